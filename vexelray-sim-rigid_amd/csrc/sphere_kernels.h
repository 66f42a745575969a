// sphere_kernels.h — device-side layout of a sphere world and the launchers of its kernels.
//
// State is structure-of-arrays in HBM: one f32 array per per-sphere field (x, y, z, u, v, w, r, im, cx..cz, dx..dz,
// ax..az, tx..tz, qx..qw), so every per-sphere pass reads and writes whole cache lines with coalesced loads.
// The contact list is compressed rows by lower sphere: sphere a's contacts occupy slots off[a] .. off[a]+cnt[a]-1,
// in the order the 27-cell walk meets them. Two such lists are kept, the substep's and the one before it, and swap.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace vr {

enum { H = 0, GX, GY, GZ, SX, SY, SZ, OMEGA, AVERAGED, RESTITUTION, MU_STATIC, MU_KINETIC, MU_ROLLING, PARAM_COUNT };
enum { LISTED = 0, MISSED = 1, LONGEST = 2 };
constexpr int IMPULSE_WORDS = 7;
constexpr int WALL_IMPULSE_WORDS = 21;
constexpr int SHOWN_STRIDE = 40;
constexpr int MOST_KEPT = 32;

// One contact list: compressed rows by lower sphere a.
struct ContactList {
    int32_t *a, *b, *done;  // cap
    float *impulse;         // 7 * cap
    int32_t *off, *cnt;     // n + 1, n: sphere a's first slot and how many it found (fitting or not)
};

struct SphereDev {
    int n, cap, nx, ny, nz, cells;
    float cell;
    int wide;
    const float *params;
    float *x, *y, *z, *u, *v, *w, *r, *im, *cx, *cy, *cz, *dx, *dy, *dz, *ax, *ay, *az, *tx, *ty, *tz, *qx, *qy,
        *qz, *qw;
    float *shown;
    int32_t *keys, *order, *starts;
    long long *claims[3];
    ContactList cur, kept;
    int32_t *count;  // contactCount: LISTED, MISSED, LONGEST
    int32_t *open;   // two words
    float *wallImpulse;
};

// A round's constants (Spheres.java:1069-1075) and the claim array it claims into (checked = slot+2, cleared =
// slot+1, all mod 3).
struct RoundConst {
    int checks, claims, pass, dispatch, mark, checksDispatch, rubs, slot;
};

// Device control of the Gauss-Seidel passes run until done (see k_round_wide / k_round_tail). Each pass of a step
// has a slot of its own, cleared once at the start of the step, so no pass needs a kernel to set it up.
constexpr int MAX_WIDE = 64;  // most wide rounds a pass launches
constexpr int SEG = 256;      // open-list places a wide round's workgroup owns (its thread count)

struct PassSlot {
    int tailCount;          // contacts the pass's last wide round left open, compacted for the tail
    int opened[MAX_WIDE];   // opened[j] != 0: round j had an open contact to take
    int openAfter[MAX_WIDE];  // contacts left open by wide round j (every wide round but the last)
};

struct PassCtl {
    int k;  // rounds the last parting pass ran, as the reference's stepper counts them
    int error, errRounds, errListed;
    int waits, rounds, most, unlisted;  // SphereStepper.Report, added up over the step
    int submissions;                    // the reference's submissions for the step's opening and more lists
    long long listedSum;                // contacts listed, added up over every pass (parting and rubbing)
    int wideRan[MAX_WIDE];              // per pass of a substep: the most rounds run wide this step
    int localFrom[MAX_WIDE];            // per pass of a substep: the latest first round with at most LOCAL_THREADS
                                        // contacts open, over the step's substeps (the local plan's wide rounds)
};

// Round j's open contacts, as {c, a, o, 0}: in segment g of buf[j & 1], seg[j & 1][g] of them, for the wide rounds
// (round 0 takes the listed contacts in place); from place 0 of buf[j & 1] for the tail.
struct OpenLists {
    int4 *buf[2];     // segments * SEG entries each
    int32_t *seg[2];  // segments each
    int segments;
};

struct SortScratch {
    int32_t *keysA, *valsA, *keysB, *hist, *scanTmp;
    int maxBlocks;
};

// Launchers (all asynchronous on `s`).
void launch_predict(const SphereDev &D, hipStream_t s);
void launch_bin(const SphereDev &D, hipStream_t s);
void launch_sort(const SphereDev &D, const SortScratch &S, hipStream_t s);
void launch_warm_walls(const SphereDev &D, hipStream_t s);
void launch_clear_contacts(const SphereDev &D, hipStream_t s);
void launch_list(const SphereDev &D, const SortScratch &S, hipStream_t s);
void launch_round(const SphereDev &D, const RoundConst &rc, hipStream_t s);
// A pass run until done (parting) or for the parting pass's rounds (rubbing), with no host wait: `wide` rounds of
// the whole chip while many contacts are open, then one workgroup finishes the pass.
// How a parting pass's lists count as the reference's submissions (SphereSimulation.submit, SphereSimulation.java:
// 231-248, of the lists SphereStep.opening and more make, SphereStep.java:348-381): `prefix` passes open its opening
// list before its rounds, after the rubbing rounds of the pass before when `carry`; pieces of at most `slice`
// dispatches (0 for no bound).
struct PassSubmit {
    int prefix, carry, slice;
};

// `local`: the pass's tail in one workgroup of LOCAL_THREADS with the claims in LDS, for at most LOCAL_MOST spheres;
// with `wide` 0, the whole pass. A round in one workgroup pays while its open contacts fit in one go of the
// workgroup's threads: each further go waits out the memory latency again, where the wide rounds take them side by
// side. So the local plan launches wide rounds until a pass has at most LOCAL_THREADS contacts open (as the step
// before found, PassCtl::localFrom) and finishes in the workgroup.
constexpr int LOCAL_MOST = 6144, LOCAL_THREADS = 1024;
void launch_pass(const SphereDev &D, PassCtl *ctl, PassSlot *slot, const OpenLists &L, int32_t *opening,
                 int32_t *enough, int it, int iterations, bool rub, int wide, bool local, const PassSubmit &submit,
                 hipStream_t s);
void launch_step_begin(PassCtl *ctl, PassSlot *slots, int nslots, hipStream_t s);
void launch_walls(const SphereDev &D, hipStream_t s);
void launch_velocity(const SphereDev &D, hipStream_t s);
void launch_bounce(const SphereDev &D, bool grid, hipStream_t s);
void launch_bounced(const SphereDev &D, hipStream_t s);
void launch_jacobi_solve(const SphereDev &D, bool grid, hipStream_t s);
void launch_jacobi_apply(const SphereDev &D, hipStream_t s);
void launch_show(const SphereDev &D, bool grid, hipStream_t s);
void launch_fill(float *p, float value, int64_t count, hipStream_t s);
void launch_fill_i32(int32_t *p, int32_t value, int64_t count, hipStream_t s);

// Exclusive scan of `n` ints from `in` to `out` (may alias); the total goes to *total if not null.
void scan_exclusive(const int32_t *in, int32_t *out, int n, int32_t *total, int32_t *tmp, hipStream_t s);
int scan_tmp_words(int n);

}  // namespace vr
