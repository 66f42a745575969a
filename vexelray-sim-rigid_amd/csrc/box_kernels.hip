// box_kernels.hip — the box world's kernels for gfx950: one per pass of the reference's box step
// (vexelray-sim-rigid-core/.../box/BoxStep.java:138-153, 199-231), over structure-of-arrays body state and the
// reference's per-slot pair layout (32 slots a body, BoxBroad.java:46).
#include "box_kernels.h"

#include <algorithm>
#include <stdexcept>
#include <string>

#include "box_math.h"

namespace vrbox {

static_assert(IMPULSE_WORDS % 4 == 0, "a slot's impulse words are written as 16-byte stores");

namespace {

constexpr int THREADS = 256;

inline int blocks_for(int64_t n) { return (int)std::max<int64_t>(1, (n + THREADS - 1) / THREADS); }

inline void check_launch(const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

__device__ __forceinline__ Bufs bufs(const BoxDev &D) {
    Bufs B;
    for (int a = 0; a < 3; a++) {
        B.x[a] = D.x[a];
        B.vel[a] = D.vel[a];
        B.l[a] = D.l[a];
        B.half[a] = D.half[a];
    }
    for (int k = 0; k < 4; k++) B.q[k] = D.q[k];
    B.im = D.im;
    B.shape = D.shape;
    B.params = D.params;
    B.lambda = D.lambda;
    B.rub = D.rub;
    B.roll = D.roll;
    return B;
}

// Boxes.velocities, Boxes.java:276-289, for body s.
__device__ __forceinline__ void velocity_of(const BoxDev &D, int s) {
    if (!(D.im[s] > 0)) return;
    const float *P = D.params;
    float h = P[H];
    float g[3] = {P[GX], P[GY], P[GZ]};
    for (int a = 0; a < 3; a++) D.vel[a][s] = D.vel[a][s] + h * g[a];
}

// With `motion`, the body's motion record is made from the state as well (the sweeps' copy of velocity and momentum).
__global__ void k_velocities(BoxDev D, int motion) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= D.n) return;
    velocity_of(D, s);
    if (motion) {
        float4 *v = D.motion + (size_t)s * MOTION_VECTORS;
        v[0] = make_float4(D.vel[0][s], D.vel[1][s], D.vel[2][s], 0.f);
        v[1] = make_float4(D.l[0][s], D.l[1][s], D.l[2][s], 0.f);
    }
}

// Boxes.positions, Boxes.java:299-342, for body s: the midpoint turn, with the angular momentum as the state.
__device__ __forceinline__ void position_of(const BoxDev &D, int s) {
    if (!(D.im[s] > 0)) return;
    Bufs B = bufs(D);
    float h = D.params[H];
    for (int a = 0; a < 3; a++) B.x[a][s] = B.x[a][s] + h * B.vel[a][s];
    Moving m = load_moving(B, s);
    float cq[4], own[3], omega[3];
    conjugate(m.q, cq);
    rotate_scaled(cq, m.twice, m.mom, own);
    for (int a = 0; a < 3; a++) own[a] = m.inverse[a] * own[a];
    rotate_scaled(m.q, m.twice, own, omega);
    float halfH = 0.5f * h;
    float mid[4], omegaMid[3], end[4], cm[4];
    turned(m.q, omega, halfH, mid);
    conjugate(mid, cm);
    float twiceMid = turn_scale(mid);
    rotate_scaled(cm, twiceMid, m.mom, own);
    for (int a = 0; a < 3; a++) own[a] = m.inverse[a] * own[a];
    rotate_scaled(mid, twiceMid, own, omegaMid);
    turned(m.q, omegaMid, h, end);
    for (int k = 0; k < 4; k++) B.q[k][s] = end[k];
}

// With `frame`, a moving body's frame record takes its new position and turn (the rest of the record is fixed over a
// step).
__global__ void k_positions(BoxDev D, int frame) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= D.n) return;
    position_of(D, s);
    if (frame && D.im[s] > 0) {
        float4 *f = D.frame + (size_t)s * FRAME_VECTORS;
        f[0] = make_float4(D.x[0][s], D.x[1][s], D.x[2][s], D.im[s]);
        float q[4] = {D.q[0][s], D.q[1][s], D.q[2][s], D.q[3][s]};
        f[1] = make_float4(q[0], q[1], q[2], q[3]);
        f[3].w = turn_scale(q);
    }
}

// The body this thread takes among the workgroup's THREADS consecutive ones, boxes first and then balls, each in
// body order (a stable partition by shape within the workgroup); n or more for a place past the last body.
__device__ __forceinline__ int shape_first(const BoxDev &D) {
    __shared__ int sBody[THREADS];
    __shared__ int sBoxes[THREADS / 64];
    int mine = blockIdx.x * THREADS + threadIdx.x;
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    bool box = mine < D.n && !(D.shape[mine] > 0.5f);  // Moving::sphere
    unsigned long long boxes = __ballot(box);
    if (lane == 0) sBoxes[wave] = __popcll(boxes);
    __syncthreads();
    int before = 0, total = 0;
    for (int w = 0; w < THREADS / 64; w++) {
        before += w < wave ? sBoxes[w] : 0;
        total += sBoxes[w];
    }
    int boxesBefore = before + __popcll(boxes & ((1ull << lane) - 1ull));
    sBody[box ? boxesBefore : total + (threadIdx.x - boxesBefore)] = mine;
    __syncthreads();
    return sBody[threadIdx.x];
}

// v[a] for a run-time a, as a register select (an indexed small array would go to scratch memory).
__device__ __forceinline__ float pick(const float *v, int a) { return a == 0 ? v[0] : (a == 1 ? v[1] : v[2]); }

// Boxes.walls, Boxes.java:357-416: one sweep of a body's corners (a sphere's nearest point) against the six walls.
// A body's sweep reads and writes only that body's state and wall impulses, so which thread takes which body changes
// nothing in the result. The wall impulses are slot-major: word w of the body at place `place` is at w * n + place,
// so a wave's loads of one word coalesce.
//
// A box's corner does something only if it holds an impulse or, in a solve or relax sweep, is within the margin of
// the wall: any other returns at once from Boxes.contact, or is passed over before it when its wall holds nothing.
// Lanes of a wave differ in which of their 48 corners do something, and a wave that ran every corner some lane needs
// would run most of them. So each box first lists its corners that may do something, in the order the reference
// takes them (a bit a corner: the holding ones exactly, from their impulses; the ones within the margin from a bound
// that never leaves one out), and the wave then takes each lane's list one corner at a time, each lane its own next
// corner, with the reference's code for a corner. The corners left out are ones the reference does nothing with, and
// each body's corners run in its order, so every body ends the sweep as the reference leaves it.

// A ball's sweep: its point nearest each wall, Boxes.contact and then Boxes.roll. `held` is the ball's cornerHeld
// word: bit `wall` clear only when the wall's words are quiet, its normal impulse not above zero, its tangent impulse
// of zero length and its rolling impulse three words of +0. At a quiet wall Boxes.contact returns at once in a warm
// sweep, and in a solve or relax sweep when the point is outside the margin; Boxes.roll then turns the ball by a
// zero change (in a warm sweep by its rolling impulse; otherwise by the new impulse less the old, which with no
// rolling friction is the old impulse less itself) and writes back what the words held. So such a wall is the turn by
// +0 alone, which does not need its words. Returns the word as the sweep leaves it.
template <int SWEEP, bool REVERSE>
__device__ unsigned long long ball_walls(const Bufs &B, Moving &box, const float *extent, unsigned long long held,
                                         size_t n, int place, const Soft &soft, const float *P, float *res) {
    const Sweep sweep = (Sweep)SWEEP;
    bool still = !(P[MU_ROLLING] > 0);  // Boxes.roll's turning stop is taken only with rolling friction
    // Moving::turn by a zero change adds the zero change to the momentum and its inverse-inertia product to the spin;
    // that product depends only on the turn and the inertia, which the sweep leaves as they are, so it is worked out
    // once for all of the ball's quiet walls.
    const float none[3] = {0.f, 0.f, 0.f};
    float spinOfNone[3];
    box.inverseInertiaTimes(none, spinOfNone);
    for (int order = 0; order < WALLS; order++) {
        int wall = REVERSE ? WALLS - 1 - order : order;
        int axis = wall / 2;
        bool high = wall % 2 == 1;
        int firstSlot = wall * CORNERS;
        float nrm[3], r[3];
        for (int a = 0; a < 3; a++) nrm[a] = a == axis ? (high ? -1.0f : 1.0f) : 0.0f;
        for (int a = 0; a < 3; a++) r[a] = -(box.half[0] * nrm[a]);
        float reach = pick(box.at, axis) + pick(r, axis);
        float gap = high ? pick(extent, axis) - reach : reach;
        bool quiet = !((held >> wall) & 1ull);
        if (quiet && (sweep == WARM || (still && !(gap < soft.margin)))) {
            for (int a = 0; a < 3; a++) {  // box.turn(none)
                box.mom[a] = box.mom[a] + none[a];
                box.omega[a] = box.omega[a] + spinOfNone[a];
            }
            continue;
        }
        size_t at = (size_t)firstSlot * n + place;
        size_t tangentAt = (size_t)firstSlot * 3 * n + place;
        float normal = B.lambda[at];
        float tangent[3] = {B.rub[tangentAt], B.rub[tangentAt + n], B.rub[tangentAt + 2 * n]};
        float was[4] = {normal, tangent[0], tangent[1], tangent[2]};
        bool wrote;
        contact_held(sweep, gap, nrm, nullptr, nullptr, box, r, normal, tangent, wrote, soft, res);
        if (wrote) {  // a word written with the bits it holds is left as it is
            if (bits(normal) != bits(was[0])) B.lambda[at] = normal;
            for (int a = 0; a < 3; a++)
                if (bits(tangent[a]) != bits(was[1 + a])) B.rub[tangentAt + a * n] = tangent[a];
        }
        float rolled[3] = {B.roll[tangentAt], B.roll[tangentAt + n], B.roll[tangentAt + 2 * n]};
        float rollWas[3] = {rolled[0], rolled[1], rolled[2]};
        roll_held(sweep, nullptr, box, box.half[0], normal, rolled, wrote, P);
        if (wrote)
            for (int a = 0; a < 3; a++)
                if (bits(rolled[a]) != bits(rollWas[a])) B.roll[tangentAt + a * n] = rolled[a];
        if (sweep != WARM) {
            bool nowQuiet = !(normal > 0) && !(dot(tangent, tangent) > 0) &&
                            (bits(rolled[0]) | bits(rolled[1]) | bits(rolled[2])) == 0u;
            held = nowQuiet ? held & ~(1ull << wall) : held | (1ull << wall);
        }
    }
    return held;
}

// The corners of a box (centre `at`, turn q, half sizes `half`) that may do something in this sweep: bit
// order * CORNERS + k for the k-th corner taken at the order-th wall taken. `held` is the box's cornerHeld word: bit
// wall * CORNERS + c is set for every corner c that holds an impulse at wall `wall` (and maybe for others).
template <int SWEEP, bool REVERSE>
__device__ unsigned long long corner_list(const float *at, const float *q, const float *half, const float *extent,
                                          unsigned long long held, const Soft &soft) {
    const Sweep sweep = (Sweep)SWEEP;
    float diagonal = sqrtf((half[0] * half[0] + half[1] * half[1]) + half[2] * half[2]);
    float twice = turn_scale(q);  // the conjugate's too
    unsigned long long todo = 0;
    for (int order = 0; order < WALLS; order++) {
        int wall = REVERSE ? WALLS - 1 - order : order;
        int axis = wall / 2;
        bool high = wall % 2 == 1;
        unsigned holding = (unsigned)(held >> (wall * CORNERS)) & 0xffu;
        if (sweep == WARM) {  // a warm sweep does something only at the corners holding an impulse
            for (int k = 0; k < CORNERS; k++)
                if ((holding >> (REVERSE ? CORNERS - 1 - k : k)) & 1u) todo |= 1ull << (order * CORNERS + k);
            continue;
        }
        // rotate turns by q / |q|, so a corner's offset along the wall's axis is row . offset for the axis's row of
        // the turn, which is the axis turned back; `slack` (1e-4 of the sizes involved) covers the rounding of both
        // ways of working it out and of the gap's two additions, which stays below 1e-6 of them. So a corner whose
        // gap so bounded is at least the margin has its own gap at least the margin, and no corner within the margin
        // is left out; NaN or infinite sizes leave every corner in.
        float centre = pick(at, axis), side = pick(extent, axis);
        float slack = 1e-4f * ((diagonal + fabsf(centre)) + fabsf(side));
        // No corner is further than the half-diagonal from the centre along any axis: a wall none of whose corners
        // holds anything and whose nearest corner could not be within the margin is passed over whole.
        float clear = high ? side - centre : centre;
        if (holding == 0 && (clear - diagonal) - slack >= soft.margin) continue;
        float along[3] = {0.f, 0.f, 0.f}, row[3], cq[4];
        along[0] = axis == 0 ? 1.f : 0.f;
        along[1] = axis == 1 ? 1.f : 0.f;
        along[2] = axis == 2 ? 1.f : 0.f;
        conjugate(q, cq);
        rotate_scaled(cq, twice, along, row);
        for (int k = 0; k < CORNERS; k++) {
            int c = REVERSE ? CORNERS - 1 - k : k;
            float offset[3];
            for (int a = 0; a < 3; a++) {
                int bit = (c / (1 << a)) % 2;
                offset[a] = half[a] * (2.0f * (float)bit - 1.0f);
            }
            float rough = dot(row, offset);
            float bound = high ? side - (centre + rough) : centre + rough;
            bool near = !((bound - slack) >= soft.margin);
            if (((holding >> c) & 1u) || near) todo |= 1ull << (order * CORNERS + k);
        }
    }
    return todo;
}

// The listed corners of a box, in list order, each through Boxes.contact; returns the corners that hold an impulse
// after it, as cornerHeld has them.
template <int SWEEP, bool REVERSE>
__device__ unsigned long long corner_sweep(const Bufs &B, Moving &box, const float *extent, unsigned long long held,
                                           unsigned long long todo, size_t n, int place, const Soft &soft,
                                           float *res) {
    const Sweep sweep = (Sweep)SWEEP;
    unsigned long long nowHeld = 0;
    while (__any(todo != 0)) {
        if (todo == 0) continue;
        int bit = __ffsll((long long)todo) - 1;
        todo &= todo - 1;
        int order = bit / CORNERS, k = bit % CORNERS;
        int wall = REVERSE ? WALLS - 1 - order : order;
        int c = REVERSE ? CORNERS - 1 - k : k;
        int axis = wall / 2;
        bool high = wall % 2 == 1;
        bool holding = (held >> (wall * CORNERS + c)) & 1;
        float nrm[3], offset[3], r[3];
        for (int a = 0; a < 3; a++) nrm[a] = a == axis ? (high ? -1.0f : 1.0f) : 0.0f;
        for (int a = 0; a < 3; a++) {
            int b = (c / (1 << a)) % 2;
            offset[a] = box.half[a] * (2.0f * (float)b - 1.0f);
        }
        box.turned_by(offset, r);
        float reach = pick(box.at, axis) + pick(r, axis);
        float gap = high ? pick(extent, axis) - reach : reach;
        // A corner holding nothing and outside the margin returns at once from Boxes.contact.
        if (!holding && !(gap < soft.margin)) continue;
        size_t at = (size_t)(wall * CORNERS + c) * n + place, tangentAt = (size_t)(wall * CORNERS + c) * 3 * n + place;
        if (contact(sweep, gap, nrm, nullptr, nullptr, box, r, B.lambda, at, B.rub, tangentAt, n, soft, res))
            nowHeld |= 1ull << (wall * CORNERS + c);
    }
    return nowHeld;
}

// Body s, its wall impulse words at place `place`. Its cornerHeld word is rewritten by a solve or relax sweep (a
// warm sweep changes no impulse). With MOTION its velocity and momentum are its motion record's, found through
// `motion` (the world's records when null); else the state's.
template <int SWEEP, bool REVERSE, bool MOTION = false, class M = GlobalMotion>
__device__ void walls_of(const BoxDev &D, int s, int place, float *res = nullptr, const M *motion = nullptr) {
    if (!(D.im[s] > 0)) return;
    Bufs B = bufs(D);
    const float *P = D.params;
    const SceneDev sc = D.sceneInfo[D.scene[s]];
    float extent[3] = {sc.sx, sc.sy, sc.sz};
    float4 *record = nullptr;
    if (MOTION) record = motion ? motion->at(s) : D.motion + (size_t)s * MOTION_VECTORS;
    Moving box = MOTION ? load_moved_at(B, record, s) : load_moving(B, s);
    const Moving loaded = box;
    Soft soft(P, true);
    size_t n = (size_t)D.n;
    unsigned long long held = D.cornerHeld[s], nowHeld;
    if (box.sphere()) {
        nowHeld = ball_walls<SWEEP, REVERSE>(B, box, extent, held, n, place, soft, P, res);
    } else {
        unsigned long long todo = corner_list<SWEEP, REVERSE>(box.at, box.q, box.half, extent, held, soft);
        nowHeld = corner_sweep<SWEEP, REVERSE>(B, box, extent, held, todo, n, place, soft, res);
    }
    if (SWEEP != WARM && nowHeld != held) D.cornerHeld[s] = nowHeld;
    if (MOTION)
        store_motion_changed_at(record, box, loaded);
    else
        store_moving_changed(B, box, loaded);
}

// The sweep over all bodies. A workgroup takes its THREADS consecutive bodies with the boxes first and then the balls
// (shape_first), so its waves run the box path or the ball path, not one after the other, while the workgroup's
// loads still fall in the same cache lines as in body order.
// Whether scene `scene` takes part in sweep a.g (the first of its phase, or the scene was open in the sweep before).
__device__ __forceinline__ bool sweeping(const BoxDev &D, const SweepArgs &a, int scene) {
    return a.first || D.sw.sceneOpen[scene] >= a.g - 1;
}

// One scene's largest residual in a sweep, recorded: past the tolerance, the scene is open in this sweep; measured,
// the scene's last sweep and residual, and the sweep's largest. Residuals are at least zero (or NaN), so their bits
// order as they do, NaN above every number.
template <int MODE>
__device__ __forceinline__ void record_one(const BoxDev &D, const SweepArgs &a, int scene, unsigned largest) {
    if ((MODE & SWEEP_STOP) && !(__uint_as_float(largest) <= a.tol)) D.sw.sceneOpen[scene] = a.g;
    if (MODE & SWEEP_MEASURE) {
        unsigned long long key = ((unsigned long long)(unsigned)a.g << 32) | (unsigned long long)largest;
        atomicMax(D.sw.sceneLast + scene, key);
        if (a.curveAt >= 0) atomicMax(D.sw.curve + a.curveAt, largest);
    }
}

// What a wave of a stopped or measured sweep leaves of its lanes' residuals, every lane of the wave calling it: a
// wave whose live lanes all hold one scene (the usual case: bodies and rounds run in scene order) records the wave's
// largest once, through one lane; a wave across scenes records lane by lane. So a sweep stores a word a wave, not a
// lane, and no word is stored to by every lane of the launch.
template <int MODE>
__device__ __forceinline__ void record(const BoxDev &D, const SweepArgs &a, bool live, int scene, float res) {
    unsigned long long lives = __ballot(live);
    if (lives == 0ull) return;
    int lane = threadIdx.x & 63;
    int leader = __ffsll((long long)lives) - 1;
    int first = __shfl(scene, leader);
    unsigned mine = live ? __float_as_uint(res) : 0u;
    if (__ballot(live && scene != first) == 0ull) {
        unsigned largest = mine;
        for (int step = 32; step > 0; step >>= 1) largest = max(largest, (unsigned)__shfl_xor((int)largest, step));
        if (lane == leader) record_one<MODE>(D, a, first, largest);
    } else if (live) {
        record_one<MODE>(D, a, scene, mine);
    }
}

// Sets lastOpen to g if any scene is open in sweep g: what the host reads to see whether a phase is done.
__global__ void k_any_open(BoxDev D, int g) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    bool open = j < D.scenes && D.sw.sceneOpen[j] == g;
    if (__ballot(open) != 0ull && (threadIdx.x & 63) == 0) *D.sw.lastOpen = g;
}

template <int SWEEP, bool REVERSE, int MODE>
__global__ void __launch_bounds__(THREADS) k_walls(BoxDev D, SweepArgs a) {
    int s = shape_first(D);
    int place = blockIdx.x * THREADS + threadIdx.x;
    if (MODE == 0) {
        if (s < D.n) walls_of<SWEEP, REVERSE, true>(D, s, place);
        return;
    }
    int scene = s < D.n ? D.scene[s] : -1;
    bool live = s < D.n && (!(MODE & SWEEP_STOP) || sweeping(D, a, scene));
    float res = 0.f;
    if (live) walls_of<SWEEP, REVERSE, true>(D, s, place, &res);
    record<MODE>(D, a, live, scene, res);
}

__global__ void __launch_bounds__(THREADS) k_wall_place(BoxDev D, int32_t *place) {
    int s = shape_first(D);
    if (s < D.n) place[s] = blockIdx.x * THREADS + threadIdx.x;
}

// One word row (w) of the three wall impulse buffers for a workgroup's bodies, moved through LDS.
__global__ void __launch_bounds__(THREADS) k_wall_move(BoxDev D, const int32_t *from, const int32_t *to) {
    __shared__ float row[THREADS];
    int s = blockIdx.x * THREADS + threadIdx.x;
    int w = blockIdx.y;  // 0 .. LAMBDA_WORDS + 2 * RUB_WORDS - 1
    float *buf = w < LAMBDA_WORDS ? D.lambda : (w < LAMBDA_WORDS + RUB_WORDS ? D.rub : D.roll);
    int k = w < LAMBDA_WORDS ? w : (w < LAMBDA_WORDS + RUB_WORDS ? w - LAMBDA_WORDS : w - LAMBDA_WORDS - RUB_WORDS);
    size_t rowAt = (size_t)k * D.n;
    bool body = s < D.n;
    if (body) row[threadIdx.x] = buf[rowAt + from[s]];
    __syncthreads();
    if (body) buf[rowAt + to[s]] = row[threadIdx.x];
}

// BoxPairs.solve, BoxPairs.java:64-151: one pair of a round (a slot, or -1 for none), its manifold's points in turn.
// Only this pair's thread touches the pair's manifold and impulse words in a round, so every load the pair needs
// (its manifold's count, normal and four points' anchors, the points' impulse words, its bodies) is issued first and
// the points then run from registers; the words a point writes are stored at the end. The arithmetic and its order
// are those of the reference, point by point.
// One manifold point of a pair as solve_pair holds it: its anchors and its impulse words (normal, tangent 3,
// rolling 3), and whether the sweep wrote them.
struct HeldPoint {
    float anchorA[3], anchorB[3], imp[POINT_IMPULSE];
    bool wroteContact, wroteRoll;
};

template <int SWEEP, bool REVERSE>
__device__ void solve_pair(const BoxDev &D, int pair, float *res = nullptr) {
    if (pair < 0) return;
    size_t out = (size_t)pair * MANIFOLD_WORDS;
    size_t held = (size_t)pair * IMPULSE_WORDS;
    int count = (int)D.manifold[out + COUNT];
    float nrm[3] = {D.manifold[out], D.manifold[out + 1], D.manifold[out + 2]};
    // Four points, each a variable of its own (not an array indexed by the point), so they stay in registers.
    HeldPoint p0, p1, p2, p3;
    auto load = [&](HeldPoint &p, int slot) {
        size_t where = out + POINTS + slot * POINT_WORDS;
        for (int a = 0; a < 3; a++) {
            p.anchorA[a] = D.manifold[where + ANCHOR_FIRST + a];
            p.anchorB[a] = D.manifold[where + ANCHOR_SECOND + a];
        }
        for (int w = 0; w < POINT_IMPULSE; w++) p.imp[w] = D.impulses[held + POINT_IMPULSE * slot + w];
        p.wroteContact = p.wroteRoll = false;
    };
    load(p0, 0);
    load(p1, 1);
    load(p2, 2);
    load(p3, 3);
    int ia = D.pairs[2 * (size_t)pair], ib = D.pairs[2 * (size_t)pair + 1];
    if (!(count > 0)) return;
    Bufs B = bufs(D);
    const float *P = D.params;
    const Sweep sweep = (Sweep)SWEEP;
    float imA = B.im[ia], imB = B.im[ib];
    Moving first = load_moving(B, ia), second = load_moving(B, ib);
    Soft soft(P, false);
    if (imA == 0 || imB == 0) soft.stiffen(P);
    auto point = [&](HeldPoint &p, int slot) {
        if (!(slot < count)) return;
        float rA[3], rB[3];
        first.turned_by(p.anchorA, rA);
        second.turned_by(p.anchorB, rB);
        if (first.sphere())
            for (int a = 0; a < 3; a++) rA[a] = first.half[0] * nrm[a];
        if (second.sphere())
            for (int a = 0; a < 3; a++) rB[a] = -(second.half[0] * nrm[a]);
        float gap = ((second.at[0] + rB[0]) - (first.at[0] + rA[0])) * nrm[0] +
                    ((second.at[1] + rB[1]) - (first.at[1] + rA[1])) * nrm[1];
        gap = gap + ((second.at[2] + rB[2]) - (first.at[2] + rA[2])) * nrm[2];
        contact_held(sweep, gap, nrm, &first, rA, second, rB, p.imp[0], &p.imp[1], p.wroteContact, soft, res);
        if (first.sphere() || second.sphere()) {
            float radius = second.half[0];
            if (first.sphere()) radius = first.half[0];
            if (first.sphere() && second.sphere())
                radius = (first.half[0] * second.half[0]) / (first.half[0] + second.half[0]);
            roll_held(sweep, &first, second, radius, p.imp[0], &p.imp[4], p.wroteRoll, P);
        }
    };
    if (REVERSE) {
        point(p3, 3);
        point(p2, 2);
        point(p1, 1);
        point(p0, 0);
    } else {
        point(p0, 0);
        point(p1, 1);
        point(p2, 2);
        point(p3, 3);
    }
    auto store = [&](const HeldPoint &p, int slot) {
        size_t at = held + POINT_IMPULSE * slot;
        if (p.wroteContact)
            for (int w = 0; w < 4; w++) D.impulses[at + w] = p.imp[w];
        if (p.wroteRoll)
            for (int w = 4; w < POINT_IMPULSE; w++) D.impulses[at + w] = p.imp[w];
    };
    store(p0, 0);
    store(p1, 1);
    store(p2, 2);
    store(p3, 3);
    store_moving(B, first);
    store_moving(B, second);
}

// ---- a pair solved by two lanes: the even lane holds the pair's first body, its neighbour the second ----
//
// About half of a point's work is per body (the anchor turned, inertia products, the point's velocity, an impulse
// applied), and the reference does it for one body and then the other. Two neighbouring lanes do it at once, each for
// its own body, and pass each other the few values the other needs (a mass term, a velocity, a position, a spin).
// Every value both bodies feed is formed in the reference's order from the two, so each lane's arithmetic, and every
// bit of the result, is the reference's; the two lanes take the same branches, since the values that decide them
// are the same in both.
struct Side {
    Moving me;
    bool second;  // this lane holds the pair's second body
};

// The neighbour's value: lanes 2k and 2k + 1 swap, a DPP quad permutation [1, 0, 3, 2] (one VALU move, where a
// general shuffle goes through the LDS unit). Both lanes of a pair are always active together.
__device__ __forceinline__ float neighbour(float mine) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(mine), 0xB1, 0xF, 0xF, false));
}

// The first body's and the second's value of a per-body quantity, from this lane's and its neighbour's.
__device__ __forceinline__ void both(const Side &S, float mine, float &ofFirst, float &ofSecond) {
    float theirs = neighbour(mine);
    ofFirst = S.second ? theirs : mine;
    ofSecond = S.second ? mine : theirs;
}

// apply_pair: p on the second at rB, -p on the first at rA, each lane on its own body.
__device__ __forceinline__ void apply_side(Side &S, const float *r, const float *p) {
    if (S.second) {
        S.me.apply(r, p);
    } else {
        float back[3] = {-p[0], -p[1], -p[2]};
        S.me.apply(r, back);
    }
}

// mass_along: the second's inverse mass along n, then the first's added.
__device__ __forceinline__ float mass_along_side(const Side &S, const float *r, const float *n) {
    float f, s;
    both(S, S.me.inverseMassAlong(r, n), f, s);
    return s + f;
}

// relative: the second's point velocity less the first's.
__device__ __forceinline__ void relative_side(const Side &S, const float *r, float *out) {
    float v[3];
    S.me.pointVelocity(r, v);
    for (int a = 0; a < 3; a++) {
        float f, s;
        both(S, v[a], f, s);
        out[a] = s - f;
    }
}

// apply_angular: j turns the second, -j the first.
__device__ __forceinline__ void apply_angular_side(Side &S, const float *j) {
    if (S.second) {
        S.me.turn(j);
    } else {
        float back[3] = {-j[0], -j[1], -j[2]};
        S.me.turn(back);
    }
}

// contact_held (Boxes.contact, Boxes.java:462-542) for the pair's point, with r this lane's body's arm.
__device__ bool contact_side(Sweep sweep, float gap, const float *n, Side &S, const float *r, float &normalHeld,
                             float *tangentHeld, bool &wrote, const Soft &soft, float *res) {
    wrote = false;
    float normal = normalHeld;
    float tangent[3] = {tangentHeld[0], tangentHeld[1], tangentHeld[2]};
    bool holding = normal > 0 || dot(tangent, tangent) > 0;
    if (sweep == WARM) {
        if (holding) {
            float p[3];
            for (int a = 0; a < 3; a++) p[a] = normal * n[a] + tangent[a];
            apply_side(S, r, p);
        }
        return holding;
    }
    if (!(gap < soft.margin || holding)) return false;
    float k = mass_along_side(S, r, n);
    float rel[3];
    relative_side(S, r, rel);
    float closing = dot(rel, n);
    float bias = 0, massScale = 1, impulseScale = 0;
    if (gap > 0) bias = gap / soft.h;
    if (sweep == SOLVE && !(gap > 0)) {
        bias = mx(soft.bias * gap, -soft.push);
        massScale = soft.massScale;
        impulseScale = soft.impulseScale;
    }
    float change = -((massScale / k) * (closing + bias)) - impulseScale * normal;
    float total = mx(0.f, normal + change);
    float p[3];
    for (int a = 0; a < 3; a++) p[a] = (total - normal) * n[a];
    apply_side(S, r, p);

    float sliding[3];
    relative_side(S, r, sliding);
    float along = dot(sliding, n);
    for (int a = 0; a < 3; a++) sliding[a] = sliding[a] - along * n[a];
    float whole[3] = {tangent[0], tangent[1], tangent[2]};
    float speed = sqrtf(dot(sliding, sliding));
    if (speed > 1e-12f) {
        float t[3];
        for (int a = 0; a < 3; a++) t[a] = sliding[a] / speed;
        float kt = mass_along_side(S, r, t);
        float stop = speed / kt;
        for (int a = 0; a < 3; a++) whole[a] = whole[a] - stop * t[a];
    }
    float size = sqrtf(dot(whole, whole));
    if (size > soft.muStatic * total) {
        float cut = (soft.muKinetic * total) / size;
        for (int a = 0; a < 3; a++) whole[a] = whole[a] * cut;
    }
    float rubbed[3];
    for (int a = 0; a < 3; a++) rubbed[a] = whole[a] - tangent[a];
    apply_side(S, r, rubbed);
    if (res) keep_largest(*res, point_residual(k, total - normal, rubbed));
    normalHeld = total;
    for (int a = 0; a < 3; a++) tangentHeld[a] = whole[a];
    wrote = true;
    return total > 0 || dot(whole, whole) > 0;
}

// roll_held (Boxes.roll, Boxes.java:553-600) for the pair's point.
__device__ void roll_side(Sweep sweep, Side &S, float radius, float normal, float *held, bool &wrote,
                          const float *P) {
    wrote = false;
    float old[3] = {held[0], held[1], held[2]};
    if (sweep == WARM) {
        apply_angular_side(S, old);
        return;
    }
    float muRolling = P[MU_ROLLING];
    float whole[3], turning[3];
    for (int a = 0; a < 3; a++) {
        whole[a] = old[a];
        float f, s;
        both(S, S.me.omega[a], f, s);
        turning[a] = s - f;
    }
    float speed = sqrtf(dot(turning, turning));
    if (speed > 1e-12f && muRolling > 0) {
        float t[3], it[3];
        for (int a = 0; a < 3; a++) t[a] = turning[a] / speed;
        S.me.inverseInertiaTimes(t, it);
        float f, s;
        both(S, dot(t, it), f, s);
        float k = s;
        k = k + f;
        float stop = speed / k;
        for (int a = 0; a < 3; a++) whole[a] = whole[a] - stop * t[a];
    }
    float most = (muRolling * radius) * mx(0.f, normal);
    float size = sqrtf(dot(whole, whole));
    if (size > most) {
        float scale = most / size;
        for (int a = 0; a < 3; a++) whole[a] = whole[a] * scale;
    }
    float change[3];
    for (int a = 0; a < 3; a++) {
        held[a] = whole[a];
        change[a] = whole[a] - old[a];
    }
    wrote = true;
    apply_angular_side(S, change);
}

// solve_pair on two lanes, for packed entry j: `second` tells which of the pair's bodies this lane holds. Both lanes
// of a pair call it together with the same entry. Each lane loads its own body's anchors and the impulses of the
// points the pair has.
// M: where the bodies' motion records are (GlobalMotion, LocalMotion).
template <int SWEEP, bool REVERSE, class M>
__device__ void solve_pair_side(const BoxDev &D, int j, bool second, float *res, const M &motion) {
    size_t T = (size_t)D.packCap;
    int side = second ? 1 : 0;
    int count = D.packCount[j];
    float nrm[3] = {D.packNormal[j], D.packNormal[T + j], D.packNormal[2 * T + j]};
    Side S;
    S.second = second;
    int body = D.packBody[side * T + j];
    S.me = load_framed_at(D.frame, motion.at(body), body);
    if (!(count > 0)) return;
    HeldPoint p0, p1, p2, p3;
    auto load = [&](HeldPoint &p, int slot) {
        p.wroteContact = p.wroteRoll = false;
        if (!(slot < count)) {
            for (int a = 0; a < 3; a++) p.anchorA[a] = p.anchorB[a] = 0.f;
            for (int w = 0; w < POINT_IMPULSE; w++) p.imp[w] = 0.f;
            return;
        }
        const float *anchor = D.packAnchor + (size_t)((side * MOST_POINTS + slot) * 3) * T + j;
        for (int a = 0; a < 3; a++) p.anchorA[a] = p.anchorB[a] = anchor[a * T];
        const float *imp = D.packImpulse + (size_t)(slot * POINT_IMPULSE) * T + j;
        for (int w = 0; w < POINT_IMPULSE; w++) p.imp[w] = imp[w * T];
    };
    load(p0, 0);
    load(p1, 1);
    load(p2, 2);
    load(p3, 3);
    const float *P = D.params;
    const Sweep sweep = (Sweep)SWEEP;
    float imA, imB;
    both(S, S.me.im, imA, imB);
    Soft soft(P, false);
    if (imA == 0 || imB == 0) soft.stiffen(P);
    float halfFirst, halfSecond, sphereFirst, sphereSecond;
    both(S, S.me.half[0], halfFirst, halfSecond);
    both(S, S.me.sphere() ? 1.f : 0.f, sphereFirst, sphereSecond);
    bool firstSphere = sphereFirst > 0.5f, secondSphere = sphereSecond > 0.5f;
    auto point = [&](HeldPoint &p, int slot) {
        if (!(slot < count)) return;
        float r[3];
        S.me.turned_by(second ? p.anchorB : p.anchorA, r);
        if (S.me.sphere())
            for (int a = 0; a < 3; a++) r[a] = second ? -(S.me.half[0] * nrm[a]) : S.me.half[0] * nrm[a];
        float at[3];
        for (int a = 0; a < 3; a++) at[a] = S.me.at[a] + r[a];  // first.at + rA, or second.at + rB
        float f0, s0, f1, s1, f2, s2;
        both(S, at[0], f0, s0);
        both(S, at[1], f1, s1);
        both(S, at[2], f2, s2);
        float gap = (s0 - f0) * nrm[0] + (s1 - f1) * nrm[1];
        gap = gap + (s2 - f2) * nrm[2];
        contact_side(sweep, gap, nrm, S, r, p.imp[0], &p.imp[1], p.wroteContact, soft, res);
        if (firstSphere || secondSphere) {
            float radius = halfSecond;
            if (firstSphere) radius = halfFirst;
            if (firstSphere && secondSphere) radius = (halfFirst * halfSecond) / (halfFirst + halfSecond);
            roll_side(sweep, S, radius, p.imp[0], &p.imp[4], p.wroteRoll, P);
        }
    };
    if (REVERSE) {
        point(p3, 3);
        point(p2, 2);
        point(p1, 1);
        point(p0, 0);
    } else {
        point(p0, 0);
        point(p1, 1);
        point(p2, 2);
        point(p3, 3);
    }
    if (!second) {  // both lanes hold the same impulse words; the first lane stores them
        auto store = [&](const HeldPoint &p, int slot) {
            float *imp = D.packImpulse + (size_t)(slot * POINT_IMPULSE) * T + j;
            if (p.wroteContact)
                for (int w = 0; w < 4; w++) imp[w * T] = p.imp[w];
            if (p.wroteRoll)
                for (int w = 4; w < POINT_IMPULSE; w++) imp[w * T] = p.imp[w];
        };
        store(p0, 0);
        store(p1, 1);
        store(p2, 2);
        store(p3, 3);
    }
    store_motion_at(motion.at(body), S.me);
}

template <int SWEEP, bool REVERSE>
__device__ void solve_pair_side(const BoxDev &D, int j, bool second, float *res = nullptr) {
    solve_pair_side<SWEEP, REVERSE>(D, j, second, res, GlobalMotion{D.motion});
}

// A round's pairs, packed entries base .. base + pairs - 1, two neighbouring lanes an entry (solve_pair_side).
// Stopped or measured (MODE), a pair of a scene that has stopped is passed over by both its lanes, and the first
// lane records the pair's residual (both lanes hold the same).
template <int SWEEP, bool REVERSE, int MODE>
__global__ void __launch_bounds__(THREADS) k_solve_round(BoxDev D, int base, int pairs, SweepArgs a) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    int k = t >> 1;
    bool second = (t & 1) != 0;
    if (MODE == 0) {
        if (k < pairs) solve_pair_side<SWEEP, REVERSE>(D, base + k, second);
        return;
    }
    int scene = k < pairs ? D.packScene[base + k] : -1;
    bool live = k < pairs && (!(MODE & SWEEP_STOP) || sweeping(D, a, scene));
    float res = 0.f;
    if (live) solve_pair_side<SWEEP, REVERSE>(D, base + k, second, &res);
    record<MODE>(D, a, live && !second, scene, res);
}

// Body s's frame record from the state, and with `motion` its motion record.
__global__ void k_frame_in(BoxDev D, int motion) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= D.n) return;
    float im = D.im[s], shape = D.shape[s], half[3], inverse[3];
    for (int a = 0; a < 3; a++) half[a] = D.half[a][s];
    inverse_inertia(half, im, shape, inverse);
    float4 *f = D.frame + (size_t)s * FRAME_VECTORS;
    f[0] = make_float4(D.x[0][s], D.x[1][s], D.x[2][s], im);
    f[1] = make_float4(D.q[0][s], D.q[1][s], D.q[2][s], D.q[3][s]);
    f[2] = make_float4(half[0], half[1], half[2], shape);
    float q[4] = {D.q[0][s], D.q[1][s], D.q[2][s], D.q[3][s]};
    f[3] = make_float4(inverse[0], inverse[1], inverse[2], turn_scale(q));
    if (motion) {
        float4 *v = D.motion + (size_t)s * MOTION_VECTORS;
        v[0] = make_float4(D.vel[0][s], D.vel[1][s], D.vel[2][s], 0.f);
        v[1] = make_float4(D.l[0][s], D.l[1][s], D.l[2][s], 0.f);
    }
}

// Body s's velocity and momentum from its motion record back to the state.
__global__ void k_motion_out(BoxDev D) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= D.n) return;
    const float4 *v = D.motion + (size_t)s * MOTION_VECTORS;
    float4 e = v[0], g = v[1];
    D.vel[0][s] = e.x, D.vel[1][s] = e.y, D.vel[2][s] = e.z;
    D.l[0][s] = g.x, D.l[1][s] = g.y, D.l[2][s] = g.z;
}

// BoxPairs' words of the touching pairs gathered into the packed fields, entry j the j-th of the colouring's order.
__global__ void k_pack_pairs(BoxDev D, Colouring C, int touching) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= touching) return;
    size_t T = (size_t)D.packCap;
    int pair = C.touch[C.order[D.packFrom[j]]];
    const float *m = D.manifold + (size_t)pair * MANIFOLD_WORDS;
    const float *imp = D.impulses + (size_t)pair * IMPULSE_WORDS;
    int count = (int)m[COUNT];
    D.packPair[j] = pair;
    D.packCount[j] = count;
    D.packBody[j] = D.pairs[2 * (size_t)pair];
    D.packBody[T + j] = D.pairs[2 * (size_t)pair + 1];
    D.packScene[j] = D.scene[D.packBody[j]];
    for (int a = 0; a < 3; a++) D.packNormal[a * T + j] = m[a];
    for (int k = 0; k < MOST_POINTS; k++) {
        if (!(k < count)) break;
        const float *point = m + POINTS + k * POINT_WORDS;
        for (int a = 0; a < 3; a++) {
            D.packAnchor[(size_t)(k * 3 + a) * T + j] = point[ANCHOR_FIRST + a];
            D.packAnchor[(size_t)((MOST_POINTS + k) * 3 + a) * T + j] = point[ANCHOR_SECOND + a];
        }
        for (int w = 0; w < POINT_IMPULSE; w++)
            D.packImpulse[(size_t)(k * POINT_IMPULSE + w) * T + j] = imp[k * POINT_IMPULSE + w];
    }
}

// The packed impulses back to their pairs' slots: the points a pair has (the rounds touch no others).
__global__ void k_unpack_pairs(BoxDev D, int touching) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= touching) return;
    size_t T = (size_t)D.packCap;
    int pair = D.packPair[j], count = D.packCount[j];
    float *imp = D.impulses + (size_t)pair * IMPULSE_WORDS;
    for (int k = 0; k < MOST_POINTS; k++) {
        if (!(k < count)) break;
        for (int w = 0; w < POINT_IMPULSE; w++)
            imp[k * POINT_IMPULSE + w] = D.packImpulse[(size_t)(k * POINT_IMPULSE + w) * T + j];
    }
}

// ---- a whole step's substeps, a workgroup a scene ----
//
// The scenes of a batch share no body and no pair, so one scene's sweeps never wait for another's. One workgroup
// takes one scene through every substep of the step: velocities, the warm sweep, the solve sweeps, positions, the
// relax sweeps (BoxStep.java:199-231), each a pass over the scene's bodies or one of its colour rounds, with a
// workgroup barrier between passes where the launches over the whole batch had a kernel boundary. A scene's bodies
// and pairs are met in the same order, by the same code, as in those launches, so each scene ends the step as they
// leave it; what changes is that a round or a wall sweep of a few hundred bodies no longer costs a launch over the
// whole chip, and that the batch's scenes run their rounds side by side.
constexpr int SCENE_THREADS = 128;

// lower bound in round r's list (places 0 .. count) of the first pair whose scene is at least `scene`; a round lists
// its pairs in slot order, so by body and so by scene.
__device__ int first_pair_of(const BoxDev &D, const int32_t *list, int count, int scene) {
    int lo = 0, hi = count;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (D.scene[list[mid] / SLOTS] < scene) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

template <int SWEEP, bool REVERSE>
__device__ void scene_sweep(const BoxDev &D, int b0, int b1, const int *from, const int *to, int used, float *res) {
    if (REVERSE) {
        for (int s = b0 + (int)threadIdx.x; s < b1; s += SCENE_THREADS) walls_of<SWEEP, REVERSE>(D, s, s, res);
        __syncthreads();
    }
    for (int order = 0; order < used; order++) {  // BoxStep.sweep, BoxStep.java:217-231
        int r = REVERSE ? used - 1 - order : order;
        if (from[r] == to[r]) continue;  // no pair of this scene in round r: nothing to wait for
        const int32_t *list = D.rounds + (size_t)r * D.width;
        for (int k = from[r] + (int)threadIdx.x; k < to[r]; k += SCENE_THREADS)
            solve_pair<SWEEP, REVERSE>(D, list[k], res);
        __syncthreads();
    }
    if (!REVERSE) {
        for (int s = b0 + (int)threadIdx.x; s < b1; s += SCENE_THREADS) walls_of<SWEEP, REVERSE>(D, s, s, res);
        __syncthreads();
    }
}

// One phase of a substep in the scene step: up to `sweeps` sweeps of the solve or the relax. With `stop`, the sweeps
// end after the first whose residuals are all at most `tol`, as the pass-by-pass step ends them; with `measure`
// (phase index `phase`), the scene's sweeps and last residual, and each sweep's largest, go where that step puts them.
template <int SWEEP>
__device__ void scene_phase(const BoxDev &D, int b0, int b1, const int *from, const int *to, int used, int sweeps,
                            int symmetric, bool stop, bool measure, float tol, int phase, unsigned *largest) {
    bool residuals = stop || measure;
    unsigned last = 0;
    int ran = 0;
    for (int k = 0; k < sweeps; k++) {
        float res = 0.f;
        if (threadIdx.x == 0) *largest = 0u;
        __syncthreads();
        if (symmetric && k % 2 == 1) scene_sweep<SWEEP, true>(D, b0, b1, from, to, used, residuals ? &res : nullptr);
        else scene_sweep<SWEEP, false>(D, b0, b1, from, to, used, residuals ? &res : nullptr);
        ran = k + 1;
        if (!residuals) continue;
        atomicMax(largest, __float_as_uint(res));
        __syncthreads();
        last = *largest;
        if (measure && threadIdx.x == 0 && k < SOLVE_MOST_SWEEPS)
            atomicMax(D.sw.curve + (size_t)phase * SOLVE_MOST_SWEEPS + k, last);
        __syncthreads();  // every thread has read *largest before the next sweep clears it
        if (stop && __uint_as_float(last) <= tol) break;
    }
    if (measure && threadIdx.x == 0) {
        atomicAdd(D.sw.phaseSweeps + phase, ran);
        atomicMax(D.sw.phaseResidual + phase, last);
        atomicAdd(D.sw.phaseResidualSum + phase, __uint_as_float(last));
    }
}

// sceneFirst: scenes + 1 words, scene j's bodies sceneFirst[j] .. sceneFirst[j + 1]. starts: the colouring's round
// starts (round r has starts[r + 1] - starts[r] pairs). used: the rounds the step has.
__global__ void __launch_bounds__(SCENE_THREADS) k_scene_step(BoxDev D, const int32_t *sceneFirst,
                                                              const int32_t *starts, int used, int substeps,
                                                              int sweeps, int symmetric, int mode, float tol) {
    __shared__ int from[MOST_ROUNDS], to[MOST_ROUNDS];
    __shared__ unsigned largest;
    int j = blockIdx.x;
    int b0 = sceneFirst[j], b1 = sceneFirst[j + 1];
    for (int r = threadIdx.x; r < used; r += SCENE_THREADS) {
        const int32_t *list = D.rounds + (size_t)r * D.width;
        int count = starts[r + 1] - starts[r];
        from[r] = first_pair_of(D, list, count, j);
        to[r] = first_pair_of(D, list, count, j + 1);
    }
    __syncthreads();
    bool stop = (mode & SWEEP_STOP) != 0, measure = (mode & SWEEP_MEASURE) != 0;
    for (int sub = 0; sub < substeps; sub++) {
        bool kept = measure && sub < SOLVE_MOST_SUBSTEPS;
        for (int s = b0 + (int)threadIdx.x; s < b1; s += SCENE_THREADS) velocity_of(D, s);
        __syncthreads();
        scene_sweep<WARM, false>(D, b0, b1, from, to, used, nullptr);
        scene_phase<SOLVE>(D, b0, b1, from, to, used, sweeps, symmetric, stop, kept, tol, 2 * sub, &largest);
        for (int s = b0 + (int)threadIdx.x; s < b1; s += SCENE_THREADS) position_of(D, s);
        __syncthreads();
        scene_phase<RELAX>(D, b0, b1, from, to, used, sweeps, symmetric, stop, kept, tol, 2 * sub + 1, &largest);
    }
}

// The end of a measured phase of the pass-by-pass step that began with sweep g0: each scene's sweeps (from the last
// sweep it recorded; a scene with no point at all ran the phase's first sweep only) and the residual of its last
// sweep, added into the phase's words.
__global__ void k_phase_measure(BoxDev D, int g0, int phase) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= D.scenes) return;
    unsigned long long key = D.sw.sceneLast[j];
    int g = (int)(key >> 32);
    unsigned res = (unsigned)(key & 0xffffffffull);
    if (g < g0) {
        g = g0;
        res = 0u;
    }
    atomicAdd(D.sw.phaseSweeps + phase, g - g0 + 1);
    atomicMax(D.sw.phaseResidual + phase, res);
    atomicAdd(D.sw.phaseResidualSum + phase, __uint_as_float(res));
}

// ---- A sweep's rounds a workgroup a scene group ----
//
// A sweep's rounds launched one by one over the whole batch (k_solve_round) pay a launch and a chain of dependent
// loads and arithmetic for every round, however few pairs it holds, and each round's loads of the bodies' records
// land anywhere in the batch. Scenes share no body and no pair, so a scene needs no other scene's round to be done
// before its own next one: a workgroup takes a group of consecutive scenes through all of the sweep's rounds, in
// order (or the reverse), with a workgroup barrier where the launches have a kernel boundary. Each pair is solved by
// the same code as in those launches (solve_pair_side), and each scene meets its rounds in the same order, so every
// scene ends the rounds as they leave it. A group's bodies' records stay in its XCD's L2 from round to round. The
// walls stay one launch over the batch (k_walls), before or after the rounds as in BoxStep.sweep.

// The sort key of colouring place i: (round, group, class), the round the most significant, so that each round's
// entries stay where the colouring put the round.
__global__ void k_pack_keys(BoxDev D, Colouring C, int touching, int groups) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= touching) return;
    int pair = C.touch[C.order[i]];
    int r = min(max(C.colour[pair], 0), MOST_ROUNDS);
    int count = min(max((int)D.manifold[(size_t)pair * MANIFOLD_WORDS + COUNT], 1), MOST_POINTS);
    int a = D.pairs[2 * (size_t)pair], b = D.pairs[2 * (size_t)pair + 1];
    bool ball = D.shape[a] > 0.5f || D.shape[b] > 0.5f;  // Moving::sphere
    int g = groups > 1 ? D.sceneGroup[D.scene[a]] : 0;
    C.keys[i] = (r * groups + g) * PACK_CLASSES + (count - 1) * 2 + (ball ? 1 : 0);
}

// Group g's entries in round r: its classes' run of the packed order.
__global__ void k_group_ranges(BoxDev D, Groups G, const int32_t *starts, int used) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= G.count * MOST_ROUNDS) return;
    int g = t / MOST_ROUNDS, r = t % MOST_ROUNDS;
    int from = 0, to = 0;
    if (r < used) {
        size_t at = ((size_t)r * G.count + g) * PACK_CLASSES;
        from = starts[at];
        to = starts[at + PACK_CLASSES];
    }
    G.ranges[2 * t] = from;
    G.ranges[2 * t + 1] = to;
}

// Round r of a group's sweep: k_solve_round's work for the group's packed entries from .. to - 1, two neighbouring
// lanes a pair. Every thread runs the same number of iterations, so the waves' residual records (record) see whole
// waves.
//
// The packed fields' stride is made opaque to the compiler at each iteration (an empty asm that may change it), so
// the fifty-odd field addresses a pair's loads and stores use are worked out inside the iteration, as in a launch of
// k_solve_round, rather than kept live across the loop in scalar registers, which would spill into vector ones.
// A group's scenes' stop words in LDS over a sweep (k_group_rounds_local, stopped and not measured): whether each of
// scenes s0 ... takes part in the sweep (sweeping), and whether a point of it has gone past the tolerance in it; the
// latter reach the world's sceneOpen once, at the sweep's end, in place of a record a wave.
struct LocalStop {
    int s0;
    const int *live;
    int *open;
};

template <int SWEEP, bool REVERSE, int MODE, class M = GlobalMotion>
__device__ void group_round(const BoxDev &D0, int from, int to, const SweepArgs &a, const M &motion,
                            const LocalStop *stop = nullptr) {
    int half = (int)blockDim.x >> 1;
    bool second = (threadIdx.x & 1) != 0;
    BoxDev D = D0;
#pragma unroll 1
    for (int base = from; base < to; base += half) {
        asm volatile("" : "+s"(D.packCap));
        int k = base + ((int)threadIdx.x >> 1);
        if (MODE == 0) {
            if (k < to) solve_pair_side<SWEEP, REVERSE>(D, k, second, nullptr, motion);
            continue;
        }
        if ((MODE & SWEEP_STOP) && stop) {
            int scene = k < to ? D.packScene[k] : -1;
            int j = k < to ? scene - stop->s0 : 0;
            bool live = k < to && stop->live[j];
            float res = 0.f;
            if (live) solve_pair_side<SWEEP, REVERSE>(D, k, second, &res, motion);
            if (live && !second && !(res <= a.tol)) stop->open[j] = 1;  // record_one's test, lane by lane
            if (MODE & SWEEP_MEASURE) record<SWEEP_MEASURE>(D, a, live && !second, scene, res);
            continue;
        }
        int scene = k < to ? D.packScene[k] : -1;
        bool live = k < to && (!(MODE & SWEEP_STOP) || sweeping(D, a, scene));
        float res = 0.f;
        if (live) solve_pair_side<SWEEP, REVERSE>(D, k, second, &res, motion);
        record<MODE>(D, a, live && !second, scene, res);
    }
}

// Rounds first .. last - 1 of a sweep (in reverse with REVERSE); blockIdx.x: the group (the groups are listed
// largest first, so the longest start first).
template <int SWEEP, bool REVERSE, int MODE>
__global__ void __launch_bounds__(GROUP_MOST_THREADS) k_group_rounds(BoxDev D, Groups G, int first, int last,
                                                                     SweepArgs a) {
    __shared__ int from[MOST_ROUNDS], to[MOST_ROUNDS];
    int g = blockIdx.x;
    for (int r = first + (int)threadIdx.x; r < last; r += blockDim.x) {
        from[r] = G.ranges[2 * (g * MOST_ROUNDS + r)];
        to[r] = G.ranges[2 * (g * MOST_ROUNDS + r) + 1];
    }
    __syncthreads();
#pragma unroll 1
    for (int order = first; order < last; order++) {  // BoxStep.sweep, BoxStep.java:217-231
        int r = REVERSE ? first + last - 1 - order : order;
        if (from[r] == to[r]) continue;  // no pair of the group in round r: nothing to wait for
        group_round<SWEEP, REVERSE, MODE>(D, from[r], to[r], a, GlobalMotion{D.motion});
        __syncthreads();
    }
}

// k_group_rounds with the group's motion records in LDS from the first round to the last: a round's body gathers of
// the records the sweep changes become LDS accesses, and the records cross HBM once each way a sweep. (The frame
// records, which no round writes, are still read from the world's.)
template <int SWEEP, bool REVERSE, int MODE>
__global__ void __launch_bounds__(GROUP_MOST_THREADS) k_group_rounds_local(BoxDev D, Groups G, int first, int last,
                                                                           SweepArgs a) {
    extern __shared__ float4 sMotion[];  // the group's motion records, then (stopped) its scenes' stop words
    __shared__ int from[MOST_ROUNDS], to[MOST_ROUNDS];
    int g = blockIdx.x;
    int b0 = G.bodies[2 * g], b1 = G.bodies[2 * g + 1];
    int s0 = G.scenes[2 * g], s1 = G.scenes[2 * g + 1];
    for (int r = first + (int)threadIdx.x; r < last; r += blockDim.x) {
        from[r] = G.ranges[2 * (g * MOST_ROUNDS + r)];
        to[r] = G.ranges[2 * (g * MOST_ROUNDS + r) + 1];
    }
    const float4 *in = D.motion + (size_t)b0 * MOTION_VECTORS;
    int words = (b1 - b0) * MOTION_VECTORS;
    for (int k = threadIdx.x; k < words; k += blockDim.x) sMotion[k] = in[k];
    constexpr bool localStop = MODE == SWEEP_STOP;
    int *live = reinterpret_cast<int *>(sMotion + (size_t)G.mostBodies * MOTION_VECTORS);
    int *open = live + G.mostScenes;
    if (localStop)
        for (int j = threadIdx.x; j < s1 - s0; j += blockDim.x) {
            live[j] = sweeping(D, a, s0 + j) ? 1 : 0;
            open[j] = 0;
        }
    __syncthreads();
    LocalMotion motion{sMotion, b0};
    LocalStop stop{s0, live, open};
#pragma unroll 1
    for (int order = first; order < last; order++) {  // BoxStep.sweep, BoxStep.java:217-231
        int r = REVERSE ? first + last - 1 - order : order;
        if (from[r] == to[r]) continue;
        group_round<SWEEP, REVERSE, MODE>(D, from[r], to[r], a, motion, localStop ? &stop : nullptr);
        __syncthreads();
    }
    __syncthreads();  // every round's records and stop words written (the last rounds may have had no pair)
    float4 *out = D.motion + (size_t)b0 * MOTION_VECTORS;
    for (int k = threadIdx.x; k < words; k += blockDim.x) out[k] = sMotion[k];
    if (localStop)
        for (int j = threadIdx.x; j < s1 - s0; j += blockDim.x)
            if (open[j]) D.sw.sceneOpen[s0 + j] = a.g;
}


// ---- A whole phase of sweeps a workgroup a scene group ----
//
// Between two sweeps of a phase, and between a sweep's rounds and its walls, a scene waits only on itself. So a
// workgroup can take its group through all of a phase's sweeps (each its walls and rounds in BoxStep.sweep's order),
// its motion records in LDS throughout, with a barrier where the launches have a kernel boundary: a group whose
// sweeps are done, or whose scenes have all stopped at the tolerance, leaves its CU to the next group instead of
// waiting for the batch's slowest group at every sweep and every walls launch. Each scene's stop is decided as the
// launches decide it (sweeping, record_one), from the scene's own words in LDS.

// The order a group's walls take its bodies in: boxes, then balls, so that a wave runs one path; within each kind,
// first the bodies that may do something at a wall in this step (a corner or wall whose bit in cornerHeld is set, or
// the body within reach of a wall, the margin and a quarter of its own reach), then the rest. A body's walls read and
// write only that body, so the order changes no result. A wave's time in the walls is that of its busiest lane, and
// most bodies of a pile are nowhere near a wall: kept together, the bodies with something to do fill a few waves
// and the others' waves pass over the corner loop. 0: active boxes, 1: other boxes, 2: active balls, 3: other balls.
__device__ __forceinline__ int walls_class(const BoxDev &D, int s) {
    bool ball = D.shape[s] > 0.5f;  // Moving::sphere
    bool active = false;
    if (D.im[s] > 0) {  // a fixed body's walls return at once
        active = (D.cornerHeld[s] & (ball ? (1ull << WALLS) - 1ull : ~0ull)) != 0ull;
        if (!active) {
            float h0 = D.half[0][s], h1 = D.half[1][s], h2 = D.half[2][s];
            float reach = ball ? h0 : sqrtf((h0 * h0 + h1 * h1) + h2 * h2);
            const SceneDev sc = D.sceneInfo[D.scene[s]];
            float extent[3] = {sc.sx, sc.sy, sc.sz};
            float near = reach + (D.params[MARGIN] + 0.25f * reach);
            for (int a = 0; a < 3; a++) {
                float c = D.x[a][s];
                active = active || !(c >= near) || !(extent[a] - c >= near);  // NaN: active
            }
        }
    }
    return (ball ? 2 : 0) + (active ? 0 : 1);
}

// Group g's bodies in the order of their walls class (walls_class), each class in body order: a stable partition
// over the group, into G.order.
__global__ void __launch_bounds__(THREADS) k_group_order(BoxDev D, Groups G) {
    constexpr int CLASSES = 4, WAVES = THREADS / 64;
    __shared__ int sWave[CLASSES][WAVES];
    int g = blockIdx.x;
    int b0 = G.bodies[2 * g], b1 = G.bodies[2 * g + 1];
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int total[CLASSES] = {0, 0, 0, 0};  // the group's bodies of each class, the same in every thread
    for (int base = b0; base < b1; base += THREADS) {
        int s = base + (int)threadIdx.x;
        int c = s < b1 ? walls_class(D, s) : -1;
        for (int k = 0; k < CLASSES; k++) {
            unsigned long long mask = __ballot(c == k);
            if (lane == 0) sWave[k][wave] = __popcll(mask);
        }
        __syncthreads();
        for (int k = 0; k < CLASSES; k++)
            for (int w = 0; w < WAVES; w++) total[k] += sWave[k][w];
        __syncthreads();
    }
    int done[CLASSES];  // where each class's next body goes, the same in every thread
    done[0] = b0;
    for (int k = 1; k < CLASSES; k++) done[k] = done[k - 1] + total[k - 1];
    for (int base = b0; base < b1; base += THREADS) {
        int s = base + (int)threadIdx.x;
        int c = s < b1 ? walls_class(D, s) : -1;
        int mine = 0;
        for (int k = 0; k < CLASSES; k++) {
            unsigned long long mask = __ballot(c == k);
            if (lane == 0) sWave[k][wave] = __popcll(mask);
            if (c == k) mine = __popcll(mask & ((1ull << lane) - 1ull));
        }
        __syncthreads();
        int chunk[CLASSES];
        for (int k = 0; k < CLASSES; k++) {
            int before = 0, all = 0;
            for (int w = 0; w < WAVES; w++) {
                before += w < wave ? sWave[k][w] : 0;
                all += sWave[k][w];
            }
            if (c == k) mine += before;
            chunk[k] = all;
        }
        __syncthreads();
        if (c >= 0) G.order[(c == 0 ? done[0] : c == 1 ? done[1] : c == 2 ? done[2] : done[3]) + mine] = s;
        for (int k = 0; k < CLASSES; k++) done[k] += chunk[k];
    }
}

// What body s's walls do over a phase of sweeps (walls_of), found at the phase's start: 0 nothing, 1 something, 2 a
// ball touching none of its walls, whose walls only add +0 to its momentum. Over a phase, positions and turns stay
// as they are (only velocities and momenta change), and a corner (or a ball's wall) comes to hold an impulse only
// where the sweep takes it: a box's listed corners, a ball's walls that hold or lie within the margin. So a box with
// no corner listed at the phase's start (corner_list: none holding, none within the margin) has none listed in any
// of its sweeps, and its walls leave it as it is; a ball whose walls are all quiet and beyond the margin, with no
// rolling friction (Boxes.roll's turning stop needs it), has each wall turn it by +0 in every sweep (ball_walls),
// which changes a momentum word of -0 to +0 and nothing else. A fixed body returns at once.
template <int SWEEP>
__device__ __forceinline__ int walls_kind(const BoxDev &D, int s) {
    if (!(D.im[s] > 0)) return 0;
    const float *P = D.params;
    unsigned long long held = D.cornerHeld[s];
    const SceneDev sc = D.sceneInfo[D.scene[s]];
    float extent[3] = {sc.sx, sc.sy, sc.sz};
    float at[3] = {D.x[0][s], D.x[1][s], D.x[2][s]};
    float half[3] = {D.half[0][s], D.half[1][s], D.half[2][s]};
    Soft soft(P, true);
    if (D.shape[s] > 0.5f) {  // Moving::sphere: a bit a wall (the word's other bits mean nothing for a ball)
        if ((held & ((1ull << WALLS) - 1ull)) != 0ull || P[MU_ROLLING] > 0) return 1;
        if (SWEEP == WARM) return 2;
        for (int wall = 0; wall < WALLS; wall++) {  // ball_walls's gap at each wall
            int axis = wall / 2;
            bool high = wall % 2 == 1;
            float nrm[3], r[3];
            for (int a = 0; a < 3; a++) nrm[a] = a == axis ? (high ? -1.0f : 1.0f) : 0.0f;
            for (int a = 0; a < 3; a++) r[a] = -(half[0] * nrm[a]);
            float reach = pick(at, axis) + pick(r, axis);
            float gap = high ? pick(extent, axis) - reach : reach;
            if (gap < soft.margin) return 1;
        }
        return 2;
    }
    float q[4] = {D.q[0][s], D.q[1][s], D.q[2][s], D.q[3][s]};
    return corner_list<SWEEP, false>(at, q, half, extent, held, soft) != 0ull ? 1 : 0;
}

// A group's bodies by what their walls do in a phase: `list` holds the bodies of kind 1 at 0 .. swept - 1 and those
// of kind 2 at size - 1 down to size - still, each as its place past the group's first body.
struct WallList {
    unsigned short *list;
    int size, swept, still;
};

// Lists the group's bodies for its walls over a phase (walls_kind), taking them in the walls' order (G.order) so
// that a wave's bodies are mostly of one shape. Called by every thread of the workgroup; ends with a barrier.
template <int SWEEP>
__device__ void list_walls(const BoxDev &D, const Groups &G, int b0, int b1, WallList &W, int *counts) {
    if (threadIdx.x == 0) counts[0] = counts[1] = 0;
    __syncthreads();
    int lane = threadIdx.x & 63;
    unsigned long long below = (1ull << lane) - 1ull;
    for (int base = b0; base < b1; base += (int)blockDim.x) {
        int p = base + (int)threadIdx.x;
        int s = p < b1 ? G.order[p] : -1;
        int kind = s >= 0 ? walls_kind<SWEEP>(D, s) : 0;
        unsigned long long swept = __ballot(kind == 1), still = __ballot(kind == 2);
        int at1 = 0, at2 = 0;
        if (lane == 0 && swept) at1 = atomicAdd(&counts[0], __popcll(swept));
        if (lane == 0 && still) at2 = atomicAdd(&counts[1], __popcll(still));
        at1 = __builtin_amdgcn_readfirstlane(at1);
        at2 = __builtin_amdgcn_readfirstlane(at2);
        if (kind == 1) W.list[at1 + __popcll(swept & below)] = (unsigned short)(s - b0);
        if (kind == 2) W.list[W.size - 1 - (at2 + __popcll(still & below))] = (unsigned short)(s - b0);
    }
    __syncthreads();
    W.swept = counts[0];
    W.still = counts[1];
}

// A group's walls in a sweep: k_walls's work for the group's bodies that have some (WallList), a thread a body, the
// motion in LDS; each still ball's momentum words +0. Every thread runs the same number of iterations, so the waves'
// residual records see whole waves. Measured, each scene taking part records a residual of 0 as well, which its
// bodies of no walls work would have recorded (a scene's record is the largest of its bodies').
template <int SWEEP, bool REVERSE, int MODE>
__device__ void group_walls(const BoxDev &D0, const Groups &G, int b0, int s1, const WallList &W, const SweepArgs &a,
                            const LocalMotion &motion, const LocalStop &stop) {
    BoxDev D = D0;
#pragma unroll 1
    for (int base = 0; base < W.swept; base += (int)blockDim.x) {
        asm volatile("" : "+s"(D.n));  // the wall words' addresses worked out in the iteration, as in group_round
        int p = base + (int)threadIdx.x;
        int s = p < W.swept ? b0 + (int)W.list[p] : -1;
        if (MODE == 0) {
            if (s >= 0) walls_of<SWEEP, REVERSE, true>(D, s, D.wallPlace[s], nullptr, &motion);
            continue;
        }
        int scene = s >= 0 ? D.scene[s] : -1;
        bool live = s >= 0 && (!(MODE & SWEEP_STOP) || stop.live[scene - stop.s0]);
        float res = 0.f;
        if (live) walls_of<SWEEP, REVERSE, true>(D, s, D.wallPlace[s], &res, &motion);
        if ((MODE & SWEEP_STOP) && live && !(res <= a.tol)) stop.open[scene - stop.s0] = 1;
        if (MODE & SWEEP_MEASURE) record<SWEEP_MEASURE>(D, a, live, scene, res);
    }
    for (int p = (int)threadIdx.x; p < W.still; p += (int)blockDim.x) {
        int s = b0 + (int)W.list[W.size - 1 - p];
        if ((MODE & SWEEP_STOP) && !stop.live[D.scene[s] - stop.s0]) continue;
        float4 *mom = motion.at(s) + 1;
        float4 m = *mom;
        float x = m.x + 0.f, y = m.y + 0.f, z = m.z + 0.f;  // Moving::turn by +0 (store_motion_changed_at)
        if ((bits(x) != bits(m.x)) | (bits(y) != bits(m.y)) | (bits(z) != bits(m.z))) *mom = make_float4(x, y, z, 0.f);
    }
    if (MODE & SWEEP_MEASURE)
        for (int j = (int)threadIdx.x; j < s1 - stop.s0; j += (int)blockDim.x)
            if (!(MODE & SWEEP_STOP) || stop.live[j]) record_one<SWEEP_MEASURE>(D, a, stop.s0 + j, 0u);
}

// One sweep of a group: walls, then rounds in reverse with REVERSE; rounds, then walls otherwise (BoxStep.sweep,
// BoxStep.java:217-231). Ends with a barrier.
template <int SWEEP, bool REVERSE, int MODE>
__device__ void group_sweep(const BoxDev &D, const Groups &G, int b0, int s1, const WallList &W, int used,
                            const int *from, const int *to, const SweepArgs &a, const LocalMotion &motion,
                            const LocalStop &stop) {
    const LocalStop *stops = (MODE & SWEEP_STOP) ? &stop : nullptr;
    if (REVERSE) {
        group_walls<SWEEP, REVERSE, MODE>(D, G, b0, s1, W, a, motion, stop);
        __syncthreads();
    }
#pragma unroll 1
    for (int order = 0; order < used; order++) {
        int r = REVERSE ? used - 1 - order : order;
        if (from[r] == to[r]) continue;  // no pair of the group in round r: nothing to wait for
        group_round<SWEEP, REVERSE, MODE>(D, from[r], to[r], a, motion, stops);
        __syncthreads();
    }
    if (!REVERSE) group_walls<SWEEP, REVERSE, MODE>(D, G, b0, s1, W, a, motion, stop);
    __syncthreads();
}

// `sweeps` sweeps of a phase, numbered g0 ..., each a group's walls and rounds; with SYMMETRIC every other sweep runs
// backwards (BoxStep.java:204-214). Stopped (MODE), a scene takes part in a sweep when it is the phase's first or the
// scene had a point past `tol` in the sweep before, as sweeping decides it, and a group whose scenes have all stopped
// is done. Measured, each sweep records as the launches do, its largest residual at curveBase + k (curveBase -1:
// not kept).
template <int SWEEP, int MODE, bool SYMMETRIC>
__global__ void __launch_bounds__(GROUP_MOST_THREADS) k_group_phase(BoxDev D, Groups G, int used, int sweeps, int g0,
                                                                    float tol, int curveBase) {
    // The group's motion records, then two words a scene of the group, then its walls list (group_phase_lds).
    extern __shared__ float4 sMotion[];
    __shared__ int from[MOST_ROUNDS], to[MOST_ROUNDS];
    __shared__ int anyLive;
    __shared__ int wallCounts[2];
    __shared__ long long began;  // the group's start on the 100 MHz clock, kept in LDS rather than a register
    if (threadIdx.x == 0) began = (long long)wall_clock64();
    constexpr int kind = SWEEP == RELAX ? 1 : 0;  // which launch order and clocks (Groups::launch)
    int g = G.launch[kind * G.count + blockIdx.x];
    int b0 = G.bodies[2 * g], b1 = G.bodies[2 * g + 1];
    int s0 = G.scenes[2 * g], s1 = G.scenes[2 * g + 1];
    for (int r = threadIdx.x; r < used; r += blockDim.x) {
        from[r] = G.ranges[2 * (g * MOST_ROUNDS + r)];
        to[r] = G.ranges[2 * (g * MOST_ROUNDS + r) + 1];
    }
    const float4 *in = D.motion + (size_t)b0 * MOTION_VECTORS;
    int words = (b1 - b0) * MOTION_VECTORS;
    for (int k = threadIdx.x; k < words; k += blockDim.x) sMotion[k] = in[k];
    int *live = reinterpret_cast<int *>(sMotion + (size_t)G.mostBodies * MOTION_VECTORS);
    int *open = live + G.mostScenes;
    WallList W{reinterpret_cast<unsigned short *>(open + G.mostScenes), b1 - b0, 0, 0};
    list_walls<SWEEP>(D, G, b0, b1, W, wallCounts);  // ends with a barrier, after which the records are in
    constexpr bool stopped = (MODE & SWEEP_STOP) != 0;
    LocalMotion motion{sMotion, b0};
    LocalStop stop{s0, live, open};
#pragma unroll 1
    for (int k = 0; k < sweeps; k++) {
        SweepArgs a{g0 + k, k == 0 ? 1 : 0, tol, curveBase >= 0 && k < SOLVE_MOST_SWEEPS ? curveBase + k : -1};
        if (stopped) {
            if (threadIdx.x == 0) anyLive = 0;
            __syncthreads();
            for (int j = threadIdx.x; j < s1 - s0; j += blockDim.x) {
                int l = (k == 0 || open[j]) ? 1 : 0;  // sweeping: the first sweep, or open in the one before
                live[j] = l;
                open[j] = 0;
                if (l) anyLive = 1;
            }
            __syncthreads();
            if (!anyLive) break;  // no scene of the group takes part in this sweep or any after it
        }
        if (SYMMETRIC && k % 2 == 1)
            group_sweep<SWEEP, true, MODE>(D, G, b0, s1, W, used, from, to, a, motion, stop);
        else
            group_sweep<SWEEP, false, MODE>(D, G, b0, s1, W, used, from, to, a, motion, stop);
        if (stopped)  // the world's stop words as the launches leave them
            for (int j = threadIdx.x; j < s1 - s0; j += blockDim.x)
                if (open[j]) D.sw.sceneOpen[s0 + j] = a.g;
    }
    __syncthreads();
    float4 *out = D.motion + (size_t)b0 * MOTION_VECTORS;
    for (int k = threadIdx.x; k < words; k += blockDim.x) out[k] = sMotion[k];
    if (SWEEP != WARM && threadIdx.x == 0)  // the group's index loaded again rather than held over the phase
        G.clocks[kind * G.count + G.launch[kind * G.count + blockIdx.x]] += (int64_t)((long long)wall_clock64() - began);
}

// BoxBroad.bin, BoxBroad.java:74-89, in the body's scene's grid.
__global__ void k_bin(BoxDev D) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= D.n) return;
    const SceneDev sc = D.sceneInfo[D.scene[s]];
    int cellsAxis[3] = {sc.nx, sc.ny, sc.nz};
    int c[3];
    for (int a = 0; a < 3; a++) {
        float fl = floorf(D.x[a][s] / sc.cell);
        c[a] = (int)mx(0.f, mn((float)(cellsAxis[a] - 1), fl));
    }
    D.keys[s] = sc.cellBase + (c[0] + sc.nx * (c[1] + sc.ny * c[2]));
}

__device__ __forceinline__ void reach_of(const BoxDev &D, int s, float *out) {  // BoxBroad.reach, :196-227
    float x = D.q[0][s], y = D.q[1][s], z = D.q[2][s], w = D.q[3][s];
    float r[3][3] = {{1.0f - 2.0f * (y * y + z * z), 2.0f * (x * y - z * w), 2.0f * (x * z + y * w)},
                     {2.0f * (x * y + z * w), 1.0f - 2.0f * (x * x + z * z), 2.0f * (y * z - x * w)},
                     {2.0f * (x * z - y * w), 2.0f * (y * z + x * w), 1.0f - 2.0f * (x * x + y * y)}};
    float h[3] = {D.half[0][s], D.half[1][s], D.half[2][s]};
    for (int a = 0; a < 3; a++)
        out[a] = (h[0] * fabsf(r[a][0]) + h[1] * fabsf(r[a][1])) + h[2] * fabsf(r[a][2]);
    if (D.shape[s] > 0.5f)
        for (int a = 0; a < 3; a++) out[a] = h[0];
}

// The broad phase's record of the body at place t of the cell order: (centre, the body's index as bits), (its reach
// along each axis, BoxBroad.reach, 0). Written once a step after the sort, so the list's walk over a cell's bodies
// reads one record a candidate, neighbouring records for neighbouring places, not 11 state fields a candidate.
__global__ void k_broad_records(BoxDev D) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= D.n) return;
    int s = D.order[t];
    float reach[3];
    reach_of(D, s, reach);
    D.broad[2 * (size_t)t] = make_float4(D.x[0][s], D.x[1][s], D.x[2][s], __int_as_float(s));
    D.broad[2 * (size_t)t + 1] = make_float4(reach[0], reach[1], reach[2], 0.f);
}

// BoxBroad.list, BoxBroad.java:103-164: a body's partners of higher index, in ascending order, into its 32 slots.
__global__ void __launch_bounds__(THREADS) k_list(BoxDev D) {
    // A thread's partners so far, sorted, in LDS: word k of thread i at k * THREADS + i, so lanes at different places
    // in their lists still take different banks (a per-lane index into a register array costs a select a word).
    __shared__ int parts[SLOTS * THREADS];
    int *part = parts + threadIdx.x;
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= D.n) return;
    // threads take bodies in cell order, so a wave walks neighbouring cells together
    const float4 own = D.broad[2 * (size_t)t], ownReach = D.broad[2 * (size_t)t + 1];
    int s = __float_as_int(own.w);
    size_t base = (size_t)s * SLOTS;
    float margin = D.params[MARGIN];
    float centre[3] = {own.x, own.y, own.z};
    float rs[3] = {ownReach.x, ownReach.y, ownReach.z};
    int found = 0;
    const SceneDev sc = D.sceneInfo[D.scene[s]];
    int key = D.keys[s] - sc.cellBase;
    int ix = key % sc.nx, iy = (key / sc.nx) % sc.ny, iz = key / (sc.nx * sc.ny);
    // The 27 cells x fastest, then y, then z (BoxBroad.around, BoxBroad.java:230): the (up to) three x cells of a
    // y, z row are consecutive cells, so their bodies are one run of the cell order. 9 runs, their 18 bounds loaded
    // together, their records four at a time, each run in the walk's order.
    int x0 = max(ix - 1, 0), x1 = min(ix + 1, sc.nx - 1);
    int lo[9], hi[9];
#pragma unroll
    for (int r = 0; r < 9; r++) {
        int jy = iy + (r % 3 - 1), jz = iz + (r / 3 - 1);
        bool inside = jy >= 0 && jy < sc.ny && jz >= 0 && jz < sc.nz;
        int row = sc.cellBase + sc.nx * (jy + sc.ny * jz);
        lo[r] = inside ? D.starts[row + x0] : 0;
        hi[r] = inside ? D.starts[row + x1 + 1] : 0;
    }
#pragma unroll
    for (int r = 0; r < 9; r++) {
        for (int k0 = lo[r]; k0 < hi[r]; k0 += 4) {
            float4 at[4], reach[4];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                bool in = k0 + q < hi[r];
                at[q] = in ? D.broad[2 * (size_t)(k0 + q)] : make_float4(0.f, 0.f, 0.f, __int_as_float(-1));
                reach[q] = in ? D.broad[2 * (size_t)(k0 + q) + 1] : make_float4(0.f, 0.f, 0.f, 0.f);
            }
#pragma unroll
            for (int q = 0; q < 4; q++) {
                int o = __float_as_int(at[q].w);
                if (!(o > s)) continue;
                float where[3] = {at[q].x, at[q].y, at[q].z}, other[3] = {reach[q].x, reach[q].y, reach[q].z};
                bool close = true;
                for (int a = 0; a < 3; a++)
                    close = close && !(fabsf(where[a] - centre[a]) > (rs[a] + other[a]) + margin);
                if (!close) continue;
                if (found < SLOTS) {
                    int place = found;
                    while (place > 0 && part[(place - 1) * THREADS] > o) {
                        part[place * THREADS] = part[(place - 1) * THREADS];
                        place--;
                    }
                    part[place * THREADS] = o;
                }
                found++;
            }
        }
    }
    for (int m = 0; m < SLOTS; m++) {
        D.pairs[(base + m) * 2] = s;
        D.pairs[(base + m) * 2 + 1] = m < found ? part[m * THREADS] : -1;
    }
    D.contacts[2 * (size_t)D.slots + s] = found;
}

// BoxContact.collide, BoxContact.java:136-246, one thread a slot.
__device__ void collide_pair(const BoxDev &D, int ia, int ib, float *m) {
    Bufs B = bufs(D);
    m[COUNT] = 0;
    Solid first = load_solid(B, ia), second = load_solid(B, ib);
    float shapeA = B.shape[ia], shapeB = B.shape[ib];
    float margin = B.params[MARGIN];
    float dd[3] = {second.centre[0] - first.centre[0], second.centre[1] - first.centre[1],
                   second.centre[2] - first.centre[2]};
    float parted = -1e30f;
    float bestFace[2] = {-1e30f, -1e30f}, bestFaceAxis[2] = {0, 0};
    for (int side = 0; side < 2; side++) {
        const Solid &own = side == 0 ? first : second;
        const Solid &other = side == 0 ? second : first;
        for (int k = 0; k < 3; k++) {
            float rch = own.half[k];
            for (int j = 0; j < 3; j++) rch = rch + other.half[j] * fabsf(dot(own.axes[k], other.axes[j]));
            float sep = fabsf(dot(dd, own.axes[k])) - rch;
            parted = mx(parted, sep);
            if (sep > bestFace[side]) {
                bestFace[side] = sep;
                bestFaceAxis[side] = (float)k;
            }
        }
    }
    float bestEdge = -1e30f;
    Edge edge;
    for (int k = 0; k < 3; k++) {
        for (int j = 0; j < 3; j++) {
            float across[3];
            cross(first.axes[k], second.axes[j], across);
            float length2 = dot(across, across);
            if (!(length2 > PARALLEL)) continue;
            float length = sqrtf(length2);
            float inv = 1.0f / length;
            float ax[3] = {across[0] * inv, across[1] * inv, across[2] * inv};
            float rch = 0;
            for (int mm = 0; mm < 3; mm++)
                rch = rch + (first.half[mm] * fabsf(dot(first.axes[mm], ax)) +
                             second.half[mm] * fabsf(dot(second.axes[mm], ax)));
            float sep = fabsf(dot(dd, ax)) - rch;
            parted = mx(parted, sep);
            if (sep > bestEdge) {
                bestEdge = sep;
                edge.record(first, second, k, j, ax, dd);
            }
        }
    }
    float smallest = first.half[0];
    for (int k = 0; k < 3; k++) smallest = mn(smallest, first.half[k]);
    for (int k = 0; k < 3; k++) smallest = mn(smallest, second.half[k]);
    float slack = ABSOLUTE * smallest;
    float fromSecond = 0;
    if (bestFace[1] > RELATIVE * bestFace[0] + slack) fromSecond = 1;
    float faceSep = bestFace[0];
    if (fromSecond > 0.5f) faceSep = bestFace[1];
    float byEdge = 0;
    if (bestEdge > RELATIVE * faceSep + slack) byEdge = 1;
    int count = 0;
    if (!(parted > margin)) {
        if (byEdge > 0.5f) {
            edge.contact(first, second, bestEdge, m);
            count = 1;
        } else {
            Face face(first, second, fromSecond, bestFaceAxis, margin);
            face.contact(m, count);
        }
    }
    m[COUNT] = (float)count;
    if (shapeA > 0.5f || shapeB > 0.5f) sphere_contact(first, second, shapeA, shapeB, margin, m);
}

// The held bits of a workgroup's slots (a body's 32 slots are 32 neighbouring lanes, half a wave) as one word a body:
// the lane of each body's slot 0 stores the word the ballot of `held` gives its half of the wave.
__device__ __forceinline__ void store_held(uint32_t *words, int p, bool held) {
    unsigned long long mask = __ballot(held);
    int lane = threadIdx.x & 63;
    if (p % SLOTS == 0) words[p / SLOTS] = (uint32_t)(mask >> (lane & 32));
}

// The narrow phase over the slots with a partner only: a body's slots 0 .. min(found, 32) - 1 (k_list fills them in
// order). Per body, a slot that may hold a count and has no partner now gets count 0 (BoxContact.collide's first
// write; a count of 0 already there is left as it is) and the held word becomes the listed slots; the listed slots'
// count, scanned, places each in a dense list, so the separating-axis test runs on full waves (a body's 32 slots in
// half a wave, most without a partner, left most lanes idle through it).
__global__ void k_collide_bodies(BoxDev D) {
    static_assert(SLOTS == 32, "a body's slots are one word of held bits");
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= D.n) return;
    int f = min(max(D.contacts[2 * (size_t)D.slots + s], 0), SLOTS);
    uint32_t listed = f >= SLOTS ? 0xffffffffu : ((1u << f) - 1u);
    uint32_t stale = D.manifoldHeld[s] & ~listed;
    while (stale) {
        int k = __ffs(stale) - 1;
        stale &= stale - 1;
        D.manifold[((size_t)s * SLOTS + k) * MANIFOLD_WORDS + COUNT] = 0;
    }
    D.manifoldHeld[s] = listed;
    D.collideFirst[s] = f;
}

// After the scan of the listed counts: body s's listed slots at collideFirst[s] onwards of the dense list.
__global__ void k_collide_place(BoxDev D) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= D.n) return;
    int f = min(max(D.contacts[2 * (size_t)D.slots + s], 0), SLOTS);
    int first = D.collideFirst[s];
    for (int k = 0; k < f; k++) D.collideList[first + k] = s * SLOTS + k;
}

// BoxContact.collide for the listed slots, each thread taking every (grid)-th of the dense list.
__global__ void __launch_bounds__(THREADS) k_collide(BoxDev D) {
    int total = *D.collideTotal;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
        int p = D.collideList[i];
        float *m = D.manifold + (size_t)p * MANIFOLD_WORDS;
        m[COUNT] = 0;
        collide_pair(D, D.pairs[2 * (size_t)p], D.pairs[2 * (size_t)p + 1], m);
    }
}

// The narrow phase alone over given pairs (what the reference's tests run as a collide-only program).
__global__ void k_collide_pairs(BoxDev D, const int32_t *pairs, int npairs, float *manifold) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npairs) return;
    float *m = manifold + (size_t)p * MANIFOLD_WORDS;
    for (int w = 0; w < MANIFOLD_WORDS; w++) m[w] = 0;
    collide_pair(D, pairs[2 * p], pairs[2 * p + 1], m);
}

// BoxPairs.carry, BoxPairs.java:191-267, and BoxBroad.counts, BoxBroad.java:174-188, one thread a slot.
__global__ void __launch_bounds__(THREADS) k_carry_counts(BoxDev D) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= D.slots) return;
    size_t out = (size_t)p * MANIFOLD_WORDS, held = (size_t)p * IMPULSE_WORDS;
    int partner = D.pairs[2 * (size_t)p + 1];
    // A slot with no partner has count 0 in its manifold (k_collide wrote it, or its held bit was clear).
    int count = partner < 0 ? 0 : (int)D.manifold[out + COUNT];
    bool wasHeld = (D.impulseHeld[p / SLOTS] >> (p % SLOTS)) & 1u;
    store_held(D.impulseHeld, p, count > 0);
    if (!(count > 0)) {
        // A slot with no point carries nothing: its impulses are zero whatever it held before (the search below
        // finds the slot's old place only to copy from it for points it has), and it lists no contact. Most slots
        // are such; they skip the search, and their 28 words go out as seven 16-byte stores, or not at all where
        // the slot's held bit says they are zero already.
        if (wasHeld) {
            float4 *zero = reinterpret_cast<float4 *>(D.impulses + held);  // 112 bytes a slot: 16-byte aligned
            for (int k = 0; k < IMPULSE_WORDS / 4; k++) zero[k] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
        D.contacts[p] = -1;
        D.contacts[(size_t)D.slots + p] = count;
        return;
    }
    int firstSlot = (p / SLOTS) * SLOTS;
    int wasSlot = -1;
    for (int look = 0; look < SLOTS; look++) {
        int then = D.pairsWas[(size_t)(firstSlot + look) * 2 + 1];
        if (then == partner && !(partner < 0)) wasSlot = firstSlot + look;
    }
    int before = 0;
    size_t outWas = 0, heldWas = 0;
    if (!(wasSlot < 0)) {
        outWas = (size_t)wasSlot * MANIFOLD_WORDS;
        heldWas = (size_t)wasSlot * IMPULSE_WORDS;
        before = (int)D.manifoldWas[outWas + COUNT];
    }
    int ia = D.pairs[2 * (size_t)p];
    float smallest = mn(D.half[0][ia], mn(D.half[1][ia], D.half[2][ia]));
    float reachSq = (SAME_POINT * smallest) * (SAME_POINT * smallest);
    for (int slot = 0; slot < MOST_POINTS; slot++) {
        float taken[POINT_IMPULSE] = {0, 0, 0, 0, 0, 0, 0};
        if (slot < count) {
            const float *anchor = D.manifold + out + POINTS + slot * POINT_WORDS + ANCHOR_FIRST;
            float nearestSq = reachSq;
            for (int old = 0; old < MOST_POINTS; old++) {
                if (!(old < before)) continue;
                const float *then = D.manifoldWas + outWas + POINTS + old * POINT_WORDS + ANCHOR_FIRST;
                float gap[3] = {anchor[0] - then[0], anchor[1] - then[1], anchor[2] - then[2]};
                float d2 = dot(gap, gap);
                if (d2 < nearestSq) {
                    nearestSq = d2;
                    for (int w = 0; w < POINT_IMPULSE; w++)
                        taken[w] = D.impulsesWas[heldWas + POINT_IMPULSE * old + w];
                }
            }
        }
        if (slot == 0 && count == 1 && before == 1)
            for (int w = 0; w < POINT_IMPULSE; w++) taken[w] = D.impulsesWas[heldWas + w];
        for (int w = 0; w < POINT_IMPULSE; w++) D.impulses[held + POINT_IMPULSE * slot + w] = taken[w];
    }
    D.contacts[p] = count > 0 ? partner : -1;
    D.contacts[(size_t)D.slots + p] = count;
}

// Boxes.show, Boxes.java:801-820. A body's 20 words are five 16-byte vectors, read and written whole (words 7 and 11,
// which the show leaves as they are, go back as they were): neighbouring lanes' records are 80 bytes apart, and a
// word at a time made each store touch a cache line a lane.
__global__ void k_show(BoxDev D) {
    static_assert(SHOWN_STRIDE == 20, "five 16-byte vectors a body");
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= D.n) return;
    float4 *b = reinterpret_cast<float4 *>(D.shown + (size_t)s * SHOWN_STRIDE);
    float4 at = b[1], size = b[2], turn = b[4];  // words 4-7, 8-11, 16-19 as they were
    b[0] = make_float4(at.x, at.y, at.z, D.shape[s]);
    b[1] = make_float4(D.x[0][s], D.x[1][s], D.x[2][s], at.w);
    b[2] = make_float4(D.half[0][s], D.half[1][s], D.half[2][s], size.w);
    b[3] = turn;
    b[4] = make_float4(D.q[0][s], D.q[1][s], D.q[2][s], D.q[3][s]);
}

}  // namespace

// ---- the learned warm start's points (no counterpart in the reference) ----

// A basis of the plane across unit normal n: t1 = n x e / |n x e| (e the axis least along n), t2 = n x t1.
__device__ __forceinline__ void plane_basis(const float *n, float *t1, float *t2) {
    float e[3] = {0, 0, 0};
    e[fabsf(n[0]) < 0.57f ? 0 : (fabsf(n[1]) < 0.57f ? 1 : 2)] = 1;
    float c[3];
    cross(n, e, c);
    float len = sqrtf(dot(c, c));
    for (int a = 0; a < 3; a++) t1[a] = c[a] / len;
    cross(n, t1, t2);
}

// Impulses (as velocities) in the network's scale, sign(v) log(1 + |v|), which keeps the rare large ones from
// ruling the squared error; and back.
__device__ __forceinline__ float squash(float v) { return copysignf(log1pf(fabsf(v)), v); }
__device__ __forceinline__ float unsquash(float v) { return copysignf(expm1f(fabsf(v)), v); }

// Where point k of touching pair t's manifold sits, and what the solve sees there (BoxPairs.solve's arms).
struct PointAt {
    int p, count;
    float n[3], rA[3], rB[3], wsum, share;
    Moving first, second;
};

__device__ __forceinline__ PointAt point_at(const BoxDev &D, const Bufs &B, int p, int k) {
    PointAt P;
    P.p = p;
    size_t out = (size_t)p * MANIFOLD_WORDS;
    P.count = (int)D.manifold[out + COUNT];
    for (int a = 0; a < 3; a++) P.n[a] = D.manifold[out + a];
    int ia = p / SLOTS, ib = D.contacts[p];
    P.first = load_moving(B, ia);
    P.second = load_moving(B, ib);
    size_t where = out + POINTS + k * POINT_WORDS;
    P.first.turned_by(D.manifold + where + ANCHOR_FIRST, P.rA);
    P.second.turned_by(D.manifold + where + ANCHOR_SECOND, P.rB);
    if (P.first.sphere())
        for (int a = 0; a < 3; a++) P.rA[a] = P.first.half[0] * P.n[a];
    if (P.second.sphere())
        for (int a = 0; a < 3; a++) P.rB[a] = -(P.second.half[0] * P.n[a]);
    float sum = P.first.im + P.second.im;
    P.wsum = sum > 0 ? sum : 1.0f;
    P.share = sum > 0 ? P.first.im / sum : 0.5f;
    return P;
}

// Whether point k of slot p holds nothing from the carry (its normal and tangent words all zero, -0 too): the
// points a prediction may start (k_pack_prediction tests the same words once they are packed).
__device__ __forceinline__ bool carried_nothing(const BoxDev &D, int p, int k) {
    const float *held = D.impulses + (size_t)p * IMPULSE_WORDS + POINT_IMPULSE * k;
    return !(held[0] != 0.f || held[1] != 0.f || held[2] != 0.f || held[3] != 0.f);
}

// The rows pair t's points take in the features: every point, or with L.fresh the ones the carry left at zero.
__device__ __forceinline__ int point_rows(const BoxDev &D, const Learned &L, int p) {
    int count = (int)D.manifold[(size_t)p * MANIFOLD_WORDS + COUNT];
    if (!L.fresh) return count;
    int rows = 0;
    for (int k = 0; k < count; k++) rows += carried_nothing(D, p, k) ? 1 : 0;
    return rows;
}

__global__ void k_point_counts(BoxDev D, Colouring C, Learned L) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= *C.touching) return;
    L.pointBase[t] = point_rows(D, L, C.touch[t]);
}

// pointBase[touching]: the rows in all, after the exclusive scan of the counts.
__global__ void k_point_total(BoxDev D, Colouring C, Learned L) {
    int last = *C.touching - 1;
    L.pointBase[last + 1] = L.pointBase[last] + point_rows(D, L, C.touch[last]);
}

// Features: normal (3), relative velocity at the point (3), depth, the carried impulse (normal, two tangent),
// the first body's share of the inverse mass, the pair's shapes (box-box, box-ball, ball-ball), closing speed along
// the normal, and the manifold's point count / 4.
__global__ void k_point_features(BoxDev D, Colouring C, Learned L) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= *C.touching) return;
    Bufs B = bufs(D);
    int p = C.touch[t];
    int count = (int)D.manifold[(size_t)p * MANIFOLD_WORDS + COUNT];
    int row = L.pointBase[t];
    for (int k = 0; k < count; k++) {
        if (L.fresh && !carried_nothing(D, p, k)) continue;  // a carried point keeps its impulse: no row
        PointAt P = point_at(D, B, p, k);
        float va[3], vb[3], rel[3], t1[3], t2[3];
        P.first.pointVelocity(P.rA, va);
        P.second.pointVelocity(P.rB, vb);
        for (int a = 0; a < 3; a++) rel[a] = vb[a] - va[a];
        plane_basis(P.n, t1, t2);
        const float *held = D.impulses + (size_t)p * IMPULSE_WORDS + POINT_IMPULSE * k;
        float tangent[3] = {held[1], held[2], held[3]};
        int balls = (P.first.sphere() ? 1 : 0) + (P.second.sphere() ? 1 : 0);
        float *f = L.x + (size_t)(row++) * LEARN_FEATURES;
        f[0] = P.n[0];
        f[1] = P.n[1];
        f[2] = P.n[2];
        f[3] = rel[0];
        f[4] = rel[1];
        f[5] = rel[2];
        f[6] = D.manifold[(size_t)p * MANIFOLD_WORDS + POINTS + k * POINT_WORDS + DEPTH];
        f[7] = squash(held[0] * P.wsum);
        f[8] = squash(dot(tangent, t1) * P.wsum);
        f[9] = squash(dot(tangent, t2) * P.wsum);
        f[10] = P.share;
        f[11] = balls == 0 ? 1.f : 0.f;
        f[12] = balls == 1 ? 1.f : 0.f;
        f[13] = balls == 2 ? 1.f : 0.f;
        f[14] = dot(rel, P.n);
        f[15] = 0.25f * (float)count;
    }
}

// Targets: the impulse each point holds when the step is done, in the features' units and basis.
__global__ void k_point_targets(BoxDev D, Colouring C, Learned L) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= *C.touching) return;
    Bufs B = bufs(D);
    int p = C.touch[t];
    int count = (int)D.manifold[(size_t)p * MANIFOLD_WORDS + COUNT];
    for (int k = 0; k < count; k++) {
        PointAt P = point_at(D, B, p, k);
        float t1[3], t2[3];
        plane_basis(P.n, t1, t2);
        const float *held = D.impulses + (size_t)p * IMPULSE_WORDS + POINT_IMPULSE * k;
        float tangent[3] = {held[1], held[2], held[3]};
        float *o = L.t + (size_t)(L.pointBase[t] + k) * LEARN_OUTPUTS;
        o[0] = squash(held[0] * P.wsum);
        o[1] = squash(dot(tangent, t1) * P.wsum);
        o[2] = squash(dot(tangent, t2) * P.wsum);
    }
}

// The packed impulse fields of the touching pairs, saved and restored (28 fields of `touching` entries each).
__global__ void k_pack_copy(BoxDev D, float *save, int touching, int back) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= touching) return;
    size_t T = (size_t)D.packCap;
    for (int f = 0; f < MOST_POINTS * POINT_IMPULSE; f++) {
        float *packed = D.packImpulse + (size_t)f * T + j, *kept = save + (size_t)f * touching + j;
        if (back) *packed = *kept;
        else *kept = *packed;
    }
}

// The prediction as the start of every point of packed entry j where it is usable (finite, no normal part below
// zero, a pair with mass to push); the carried words stay elsewhere, and the rolling words everywhere.
__global__ void k_pack_prediction(BoxDev D, Colouring C, Learned L, int touching) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= touching) return;
    Bufs B = bufs(D);
    size_t T = (size_t)D.packCap;
    int t = C.order[D.packFrom[j]];
    int p = C.touch[t];
    int count = min((int)D.manifold[(size_t)p * MANIFOLD_WORDS + COUNT], MOST_POINTS);
    int row = L.pointBase[t];
    for (int k = 0; k < count; k++) {
        float *imp = D.packImpulse + (size_t)(k * POINT_IMPULSE) * T + j;
        // A point the carry started (it found the point's last impulse) keeps it: the carried impulses are the last
        // step's converged, balanced set, and a prediction made point by point is not. The prediction starts the
        // points the carry has nothing for (new contacts, and points that moved past the carry's reach); with
        // L.fresh only those have rows, in order.
        bool carried = imp[0] != 0.f || imp[T] != 0.f || imp[2 * T] != 0.f || imp[3 * T] != 0.f;
        if (L.fresh && carried) continue;
        const float *y = L.y + (size_t)(L.fresh ? row++ : L.pointBase[t] + k) * LEARN_OUTPUTS;
        if (carried) continue;
        PointAt P = point_at(D, B, p, k);
        if (!(isfinite(y[0]) && isfinite(y[1]) && isfinite(y[2]) && y[0] >= 0 && P.first.im + P.second.im > 0))
            continue;
        float t1[3], t2[3];
        plane_basis(P.n, t1, t2);
        float along1 = unsquash(y[1]), along2 = unsquash(y[2]);
        imp[0] = unsquash(y[0]) / P.wsum;
        for (int a = 0; a < 3; a++) imp[(1 + a) * T] = (along1 * t1[a] + along2 * t2[a]) / P.wsum;
    }
}

__global__ void k_trial_residual(BoxDev D, int g, uint32_t *out) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= D.scenes) return;
    unsigned long long key = D.sw.sceneLast[j];
    out[j] = (int)(key >> 32) == g ? (uint32_t)(key & 0xffffffffull) : 0u;
}

// A scene takes the prediction only where it leaves the first sweep strictly nearer its solution (NaN never).
__global__ void k_trial_choose(BoxDev D, Learned L, int32_t *fellBack) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= D.scenes) return;
    float carry = __uint_as_float(L.fromCarry[j]), predicted = __uint_as_float(L.fromPredicted[j]);
    bool take = predicted < carry;
    L.takes[j] = take ? 1 : 0;
    if (!take && L.fromCarry[j] != 0u) atomicAdd(fellBack, 1);
}

__global__ void k_pack_choose(BoxDev D, Learned L, int touching) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= touching) return;
    size_t T = (size_t)D.packCap;
    bool take = L.takes[D.packScene[j]] != 0;
    const float *from = take ? L.startPredicted : L.startCarry;
    for (int f = 0; f < MOST_POINTS * POINT_IMPULSE; f++)
        D.packImpulse[(size_t)f * T + j] = from[(size_t)f * touching + j];
    if (!take) return;
    int started = 0;  // the points whose start is the prediction's, not the carry's
    for (int k = 0; k < min(D.packCount[j], MOST_POINTS); k++) {
        bool differs = false;
        for (int w = 0; w < 4; w++) {
            size_t f = (size_t)(k * POINT_IMPULSE + w) * touching + j;
            differs = differs || __float_as_uint(L.startPredicted[f]) != __float_as_uint(L.startCarry[f]);
        }
        started += differs ? 1 : 0;
    }
    if (started > 0) atomicAdd(L.used, started);
}

__global__ void k_scene_from_params(const float *P, SceneDev *scene) {
    scene->sx = P[SX];
    scene->sy = P[SY];
    scene->sz = P[SZ];
}

void launch_point_features(const BoxDev &D, const Colouring &C, const Learned &L, int touchingPairs, hipStream_t s) {
    if (touchingPairs <= 0) return;
    k_point_counts<<<blocks_for(touchingPairs), THREADS, 0, s>>>(D, C, L);
    check_launch("point counts");
    vr::scan_exclusive(L.pointBase, L.pointBase, touchingPairs, nullptr, C.sort.scanTmp, s);
    k_point_total<<<1, 1, 0, s>>>(D, C, L);
    check_launch("point total");
    k_point_features<<<blocks_for(touchingPairs), THREADS, 0, s>>>(D, C, L);
    check_launch("point features");
}

void launch_point_targets(const BoxDev &D, const Colouring &C, const Learned &L, int touchingPairs, hipStream_t s) {
    if (touchingPairs <= 0) return;
    k_point_targets<<<blocks_for(touchingPairs), THREADS, 0, s>>>(D, C, L);
    check_launch("point targets");
}

void launch_pack_save(const BoxDev &D, float *to, int touching, hipStream_t s) {
    if (touching <= 0) return;
    k_pack_copy<<<blocks_for(touching), THREADS, 0, s>>>(D, to, touching, 0);
    check_launch("pack save");
}

void launch_pack_restore(const BoxDev &D, const float *from, int touching, hipStream_t s) {
    if (touching <= 0) return;
    k_pack_copy<<<blocks_for(touching), THREADS, 0, s>>>(D, const_cast<float *>(from), touching, 1);
    check_launch("pack restore");
}

void launch_pack_prediction(const BoxDev &D, const Colouring &C, const Learned &L, int touching, hipStream_t s) {
    if (touching <= 0) return;
    k_pack_prediction<<<blocks_for(touching), THREADS, 0, s>>>(D, C, L, touching);
    check_launch("pack prediction");
}

void launch_trial_residual(const BoxDev &D, int g, uint32_t *out, hipStream_t s) {
    if (D.scenes <= 0) return;
    k_trial_residual<<<blocks_for(D.scenes), THREADS, 0, s>>>(D, g, out);
    check_launch("trial residual");
}

void launch_pack_choose(const BoxDev &D, const Learned &L, int touching, int32_t *fellBack, hipStream_t s) {
    if (touching <= 0 || D.scenes <= 0) return;
    k_trial_choose<<<blocks_for(D.scenes), THREADS, 0, s>>>(D, L, fellBack);
    k_pack_choose<<<blocks_for(touching), THREADS, 0, s>>>(D, L, touching);
    check_launch("pack choose");
}

void launch_scene_from_params(const BoxDev &D, SceneDev *scene, hipStream_t s) {
    k_scene_from_params<<<1, 1, 0, s>>>(D.params, scene);
    check_launch("scene from params");
}

void launch_velocities(const BoxDev &D, bool motion, hipStream_t s) {
    k_velocities<<<blocks_for(D.n), THREADS, 0, s>>>(D, motion ? 1 : 0);
    check_launch("velocities");
}

void launch_positions(const BoxDev &D, bool frame, hipStream_t s) {
    k_positions<<<blocks_for(D.n), THREADS, 0, s>>>(D, frame ? 1 : 0);
    check_launch("positions");
}

template <int SWEEP, bool REVERSE>
static void walls_then(const BoxDev &D, hipStream_t s, int mode, const SweepArgs &a) {
    switch (SWEEP == WARM ? 0 : mode) {
        case 0: k_walls<SWEEP, REVERSE, 0><<<blocks_for(D.n), THREADS, 0, s>>>(D, a); break;
        case SWEEP_STOP: k_walls<SWEEP, REVERSE, SWEEP_STOP><<<blocks_for(D.n), THREADS, 0, s>>>(D, a); break;
        case SWEEP_MEASURE: k_walls<SWEEP, REVERSE, SWEEP_MEASURE><<<blocks_for(D.n), THREADS, 0, s>>>(D, a); break;
        default:
            k_walls<SWEEP, REVERSE, SWEEP_STOP | SWEEP_MEASURE><<<blocks_for(D.n), THREADS, 0, s>>>(D, a);
            break;
    }
}

void launch_wall_place(const BoxDev &D, int32_t *place, hipStream_t s) {
    k_wall_place<<<blocks_for(D.n), THREADS, 0, s>>>(D, place);
    check_launch("wall place");
}

void launch_wall_move(const BoxDev &D, const int32_t *from, const int32_t *to, hipStream_t s) {
    dim3 grid((unsigned)blocks_for(D.n), (unsigned)(LAMBDA_WORDS + 2 * RUB_WORDS));
    k_wall_move<<<grid, THREADS, 0, s>>>(D, from, to);
    check_launch("wall move");
}

void launch_walls(const BoxDev &D, int sweep, bool reverse, hipStream_t s, int mode, const SweepArgs &a) {
    switch (sweep * 2 + (reverse ? 1 : 0)) {
        case 0: walls_then<WARM, false>(D, s, 0, a); break;
        case 1: walls_then<WARM, true>(D, s, 0, a); break;
        case 2: walls_then<SOLVE, false>(D, s, mode, a); break;
        case 3: walls_then<SOLVE, true>(D, s, mode, a); break;
        case 4: walls_then<RELAX, false>(D, s, mode, a); break;
        default: walls_then<RELAX, true>(D, s, mode, a); break;
    }
    check_launch("walls");
}

void launch_any_open(const BoxDev &D, int g, hipStream_t s) {
    if (D.scenes <= 0) return;
    k_any_open<<<blocks_for(D.scenes), THREADS, 0, s>>>(D, g);
    check_launch("any open");
}

void launch_phase_measure(const BoxDev &D, int g0, int phase, hipStream_t s) {
    if (D.scenes <= 0) return;
    k_phase_measure<<<blocks_for(D.scenes), THREADS, 0, s>>>(D, g0, phase);
    check_launch("phase measure");
}

void launch_scene_step(const BoxDev &D, const int32_t *sceneFirst, const int32_t *starts, int used, int substeps,
                       int sweeps, bool symmetric, hipStream_t s, int mode, float tol) {
    if (D.scenes <= 0) return;
    if (used < 0 || used > MOST_ROUNDS) throw std::runtime_error("scene step: rounds out of range");
    k_scene_step<<<D.scenes, SCENE_THREADS, 0, s>>>(D, sceneFirst, starts, used, substeps, sweeps, symmetric ? 1 : 0,
                                                    mode, tol);
    check_launch("scene step");
}

template <int SWEEP, bool REVERSE>
static void round_then(const BoxDev &D, int blocks, int base, int pairs, hipStream_t s, int mode,
                       const SweepArgs &a) {
    switch (SWEEP == WARM ? 0 : mode) {
        case 0: k_solve_round<SWEEP, REVERSE, 0><<<blocks, THREADS, 0, s>>>(D, base, pairs, a); break;
        case SWEEP_STOP: k_solve_round<SWEEP, REVERSE, SWEEP_STOP><<<blocks, THREADS, 0, s>>>(D, base, pairs, a); break;
        case SWEEP_MEASURE:
            k_solve_round<SWEEP, REVERSE, SWEEP_MEASURE><<<blocks, THREADS, 0, s>>>(D, base, pairs, a);
            break;
        default:
            k_solve_round<SWEEP, REVERSE, SWEEP_STOP | SWEEP_MEASURE><<<blocks, THREADS, 0, s>>>(D, base, pairs, a);
            break;
    }
}

void launch_solve_round(const BoxDev &D, int sweep, bool reverse, int base, int pairs, hipStream_t s, int mode,
                        const SweepArgs &a) {
    if (pairs <= 0) return;
    if (base < 0 || (int64_t)base + pairs > D.packCap)
        throw std::runtime_error("solve round: entries past the packed pairs");
    int blocks = blocks_for(2 * (int64_t)pairs);  // two lanes a pair
    switch (sweep * 2 + (reverse ? 1 : 0)) {
        case 0: round_then<WARM, false>(D, blocks, base, pairs, s, 0, a); break;
        case 1: round_then<WARM, true>(D, blocks, base, pairs, s, 0, a); break;
        case 2: round_then<SOLVE, false>(D, blocks, base, pairs, s, mode, a); break;
        case 3: round_then<SOLVE, true>(D, blocks, base, pairs, s, mode, a); break;
        case 4: round_then<RELAX, false>(D, blocks, base, pairs, s, mode, a); break;
        default: round_then<RELAX, true>(D, blocks, base, pairs, s, mode, a); break;
    }
    check_launch("solve round");
}

void launch_pack_order(const BoxDev &D, const Colouring &C, int touching, int groups, int32_t *starts,
                       hipStream_t s) {
    if (touching <= 0) return;
    if (touching > C.most || groups < 1) throw std::runtime_error("pack order: out of range");
    k_pack_keys<<<blocks_for(touching), THREADS, 0, s>>>(D, C, touching, groups);
    check_launch("pack keys");
    // The colouring's keys are spent (its sort is done), so its key words and sort scratch serve this sort.
    vr::SphereDev view{};
    view.n = touching;
    view.cells = (MOST_ROUNDS + 1) * groups * PACK_CLASSES;
    view.keys = C.keys;
    view.order = D.packFrom;
    view.starts = starts;
    vr::launch_sort(view, C.sort, s);
}

void launch_group_ranges(const BoxDev &D, const Groups &G, const int32_t *starts, int used, hipStream_t s) {
    if (G.count <= 0) return;
    if (used < 0 || used > MOST_ROUNDS) throw std::runtime_error("group ranges: rounds out of range");
    k_group_ranges<<<blocks_for((int64_t)G.count * MOST_ROUNDS), THREADS, 0, s>>>(D, G, starts, used);
    check_launch("group ranges");
}

template <int SWEEP, bool REVERSE>
static void group_then(const BoxDev &D, const Groups &G, int first, int last, int threads, hipStream_t s, int mode,
                       const SweepArgs &a, bool local) {
    if (local) {
        size_t lds = sizeof(float4) * MOTION_VECTORS * (size_t)G.mostBodies + 2 * sizeof(int) * (size_t)G.mostScenes;
        switch (SWEEP == WARM ? 0 : mode) {
            case 0: k_group_rounds_local<SWEEP, REVERSE, 0><<<G.count, threads, lds, s>>>(D, G, first, last, a); break;
            case SWEEP_STOP:
                k_group_rounds_local<SWEEP, REVERSE, SWEEP_STOP><<<G.count, threads, lds, s>>>(D, G, first, last, a);
                break;
            case SWEEP_MEASURE:
                k_group_rounds_local<SWEEP, REVERSE, SWEEP_MEASURE><<<G.count, threads, lds, s>>>(D, G, first, last, a);
                break;
            default:
                k_group_rounds_local<SWEEP, REVERSE, SWEEP_STOP | SWEEP_MEASURE>
                    <<<G.count, threads, lds, s>>>(D, G, first, last, a);
                break;
        }
        return;
    }
    switch (SWEEP == WARM ? 0 : mode) {
        case 0: k_group_rounds<SWEEP, REVERSE, 0><<<G.count, threads, 0, s>>>(D, G, first, last, a); break;
        case SWEEP_STOP: k_group_rounds<SWEEP, REVERSE, SWEEP_STOP><<<G.count, threads, 0, s>>>(D, G, first, last, a); break;
        case SWEEP_MEASURE:
            k_group_rounds<SWEEP, REVERSE, SWEEP_MEASURE><<<G.count, threads, 0, s>>>(D, G, first, last, a);
            break;
        default:
            k_group_rounds<SWEEP, REVERSE, SWEEP_STOP | SWEEP_MEASURE><<<G.count, threads, 0, s>>>(D, G, first, last, a);
            break;
    }
}

void launch_group_rounds(const BoxDev &D, const Groups &G, int sweep, bool reverse, int first, int last, int threads,
                         hipStream_t s, int mode, const SweepArgs &a, bool local) {
    if (G.count <= 0 || first >= last) return;
    if (first < 0 || last > MOST_ROUNDS) throw std::runtime_error("group rounds: rounds out of range");
    if (local && (G.mostBodies < 1 || G.mostBodies > GROUP_LOCAL_BODIES || G.mostScenes < 1 ||
                  G.mostBodies * sizeof(float4) * MOTION_VECTORS + 2 * sizeof(int) * G.mostScenes > GROUP_LOCAL_BYTES))
        throw std::runtime_error("group rounds: a group's motion records and stop words do not fit in LDS");
    if (threads < 64 || threads > GROUP_MOST_THREADS || threads % 64 != 0)
        throw std::runtime_error("group rounds: a workgroup of whole waves, at most GROUP_MOST_THREADS threads");
    switch (sweep * 2 + (reverse ? 1 : 0)) {
        case 0: group_then<WARM, false>(D, G, first, last, threads, s, 0, a, local); break;
        case 1: group_then<WARM, true>(D, G, first, last, threads, s, 0, a, local); break;
        case 2: group_then<SOLVE, false>(D, G, first, last, threads, s, mode, a, local); break;
        case 3: group_then<SOLVE, true>(D, G, first, last, threads, s, mode, a, local); break;
        case 4: group_then<RELAX, false>(D, G, first, last, threads, s, mode, a, local); break;
        default: group_then<RELAX, true>(D, G, first, last, threads, s, mode, a, local); break;
    }
    check_launch("group rounds");
}

void launch_group_order(const BoxDev &D, const Groups &G, hipStream_t s) {
    if (G.count <= 0) return;
    k_group_order<<<G.count, THREADS, 0, s>>>(D, G);
    check_launch("group order");
}

template <int SWEEP, bool SYMMETRIC>
static void group_phase_then(const BoxDev &D, const Groups &G, int used, int sweeps, int g0, float tol, int curveBase,
                             int threads, size_t lds, hipStream_t s, int mode) {
    switch (SWEEP == WARM ? 0 : mode) {
        case 0:
            k_group_phase<SWEEP, 0, SYMMETRIC><<<G.count, threads, lds, s>>>(D, G, used, sweeps, g0, tol, curveBase);
            break;
        case SWEEP_STOP:
            k_group_phase<SWEEP, SWEEP_STOP, SYMMETRIC>
                <<<G.count, threads, lds, s>>>(D, G, used, sweeps, g0, tol, curveBase);
            break;
        case SWEEP_MEASURE:
            k_group_phase<SWEEP, SWEEP_MEASURE, SYMMETRIC>
                <<<G.count, threads, lds, s>>>(D, G, used, sweeps, g0, tol, curveBase);
            break;
        default:
            k_group_phase<SWEEP, SWEEP_STOP | SWEEP_MEASURE, SYMMETRIC>
                <<<G.count, threads, lds, s>>>(D, G, used, sweeps, g0, tol, curveBase);
            break;
    }
}

void launch_group_phase(const BoxDev &D, const Groups &G, int sweep, int used, int sweeps, bool symmetric, int g0,
                        float tol, int curveBase, int threads, hipStream_t s, int mode) {
    if (G.count <= 0 || sweeps <= 0) return;
    if (used < 0 || used > MOST_ROUNDS) throw std::runtime_error("group phase: rounds out of range");
    size_t lds = group_phase_lds(G.mostBodies, G.mostScenes);
    if (G.mostBodies < 1 || G.mostBodies > GROUP_LOCAL_BODIES || G.mostScenes < 1 || lds > GROUP_LOCAL_BYTES)
        throw std::runtime_error("group phase: a group's motion records, stop words and walls list do not fit in LDS");
    if (threads < 64 || threads > GROUP_MOST_THREADS || threads % 64 != 0)
        throw std::runtime_error("group phase: a workgroup of whole waves, at most GROUP_MOST_THREADS threads");
    if (sweep == WARM) {
        group_phase_then<WARM, false>(D, G, used, 1, g0, tol, -1, threads, lds, s, 0);
    } else if (sweep == SOLVE) {
        if (symmetric) group_phase_then<SOLVE, true>(D, G, used, sweeps, g0, tol, curveBase, threads, lds, s, mode);
        else group_phase_then<SOLVE, false>(D, G, used, sweeps, g0, tol, curveBase, threads, lds, s, mode);
    } else {
        if (symmetric) group_phase_then<RELAX, true>(D, G, used, sweeps, g0, tol, curveBase, threads, lds, s, mode);
        else group_phase_then<RELAX, false>(D, G, used, sweeps, g0, tol, curveBase, threads, lds, s, mode);
    }
    check_launch("group phase");
}

void launch_pack(const BoxDev &D, const Colouring &C, int touching, hipStream_t s) {
    if (touching <= 0) return;
    if (touching > D.packCap) throw std::runtime_error("pack: more touching pairs than the packed fields hold");
    k_pack_pairs<<<blocks_for(touching), THREADS, 0, s>>>(D, C, touching);
    check_launch("pack");
}

void launch_frame_in(const BoxDev &D, bool motion, hipStream_t s) {
    k_frame_in<<<blocks_for(D.n), THREADS, 0, s>>>(D, motion ? 1 : 0);
    check_launch("frame in");
}

void launch_motion_out(const BoxDev &D, hipStream_t s) {
    k_motion_out<<<blocks_for(D.n), THREADS, 0, s>>>(D);
    check_launch("motion out");
}

void launch_unpack(const BoxDev &D, int touching, hipStream_t s) {
    if (touching <= 0) return;
    k_unpack_pairs<<<blocks_for(touching), THREADS, 0, s>>>(D, touching);
    check_launch("unpack");
}

void launch_bin(const BoxDev &D, hipStream_t s) {
    k_bin<<<blocks_for(D.n), THREADS, 0, s>>>(D);
    check_launch("box bin");
}

void launch_list(const BoxDev &D, hipStream_t s) {
    k_broad_records<<<blocks_for(D.n), THREADS, 0, s>>>(D);
    k_list<<<blocks_for(D.n), THREADS, 0, s>>>(D);
    check_launch("box list");
}

// The dense collide's grid: enough workgroups to fill the chip a few times over.
constexpr int COLLIDE_BLOCKS = 4096;

void launch_collide(const BoxDev &D, hipStream_t s) {
    k_collide_bodies<<<blocks_for(D.n), THREADS, 0, s>>>(D);
    vr::scan_exclusive(D.collideFirst, D.collideFirst, D.n, D.collideTotal, D.collideScan, s);
    k_collide_place<<<blocks_for(D.n), THREADS, 0, s>>>(D);
    k_collide<<<(int)std::min<int64_t>(blocks_for(D.slots), COLLIDE_BLOCKS), THREADS, 0, s>>>(D);
    check_launch("box collide");
}

void launch_collide_pairs(const BoxDev &D, const int32_t *pairs, int npairs, float *manifold, hipStream_t s) {
    k_collide_pairs<<<blocks_for(npairs), THREADS, 0, s>>>(D, pairs, npairs, manifold);
    check_launch("box collide pairs");
}

void launch_carry_counts(const BoxDev &D, hipStream_t s) {
    k_carry_counts<<<blocks_for(D.slots), THREADS, 0, s>>>(D);
    check_launch("box carry");
}

void launch_show(const BoxDev &D, hipStream_t s) {
    k_show<<<blocks_for(D.n), THREADS, 0, s>>>(D);
    check_launch("box show");
}

}  // namespace vrbox
