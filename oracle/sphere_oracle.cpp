// sphere_oracle.cpp — CPU oracle for the sphere step. TEST INFRASTRUCTURE ONLY.
//
// A sequential restatement, pass by pass and in the reference's order, of vexelray-sim-rigid's sphere step
// (vexelray-sim-rigid-core/src/main/java/dev/vexelray/sim/rigid/sphere/, cited below as Spheres.java etc.). It is the
// checker the HIP path is compared against: only tests, the smoke check and the bench's CPU-baseline leg load it,
// and nothing in the product links it.
//
// Arithmetic is f32, operation by operation as the reference's IR builds it (the expression trees are followed, so
// `a + (b + c)` stays that way), compiled with -ffp-contract=off so that no multiply-add is fused. Each invocation
// of a reference pass is one loop iteration here, in invocation order; every pass whose invocations interact only
// through atomics that commute (integer max, integer add, exchange into a chain) gives the same state in any order,
// and the comments say where order is chosen.
//
// Choices where the reference leaves something open (the external SupirVast CountingSort):
//  * the sort is a stable counting sort: within a cell, bodies are listed in ascending index. The reference's is
//    unstable (atomic ranks), so its within-cell order is not pinned by anything; sets of pairs are.
//  * the contact list is appended in invocation order a = 0, 1, ..., each sphere's partners in the order `around`
//    visits them, which is one order the reference's atomic append can produce.
//
// Extension: with wide_pairs set, contact priorities are computed in 63 bits (the claims are 64-bit), which lifts
// the reference's 46,340-sphere limit on Gauss-Seidel.

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/vexelray_rigid.h"
#include "box_oracle.h"

namespace {

thread_local std::string g_error;

struct ArgError : std::runtime_error { using std::runtime_error::runtime_error; };
struct StateError : std::runtime_error { using std::runtime_error::runtime_error; };

// Spheres.java:141-153, parameter words.
enum { H = 0, GX, GY, GZ, SX, SY, SZ, OMEGA, AVERAGED, RESTITUTION, MU_STATIC, MU_KINETIC, MU_ROLLING, PARAM_COUNT };

constexpr float COINCIDENT = 1e-12f;           // Spheres.java:156
constexpr float LIST_MARGIN = 0.02f;           // Spheres.java:795
constexpr float INERTIA = 0.4f;                // Spheres.java:246, as f((float) INERTIA)
constexpr int MOST_KEPT = 32;                  // Spheres.java:942
constexpr int IMPULSE_WORDS = 7;               // Spheres.java:495
constexpr int WALL_IMPULSE_WORDS = 21;         // Spheres.java:950
constexpr int LISTED = 0, MISSED = 1, LONGEST = 2;  // Spheres.java:797-804
constexpr int SHOWN_STRIDE = 40, PRESSED_BEFORE = 8, PRESSED_AFTER = 20, TURNED_BEFORE = 32, TURNED_AFTER = 36;
constexpr int PRESSING_WORDS = 12, TRACE = 9, LEAN = 6, LEAN_LENGTH = 10;  // Spheres.java:1656-1676
constexpr int MOST_LISTED_SPHERES = 46340;     // Spheres.java:787
constexpr int BATCH = 8, CYCLE = 48;           // SphereStep.java:48-59
constexpr int SHRINK_AFTER = 120;              // SphereStepper.java:49

inline float one_plus_margin() { return 1.0f + LIST_MARGIN; }  // f(1 + LIST_MARGIN), rounded to f32

struct World;

// A sphere's position, moves and turns, loaded once and stored once (Spheres.java:532-569, the Held rule).
struct Held {
    int s;
    float at[3], displaced[3], turned[3];
    bool turns;
};

struct Impulse {  // Spheres.java:490-522
    float *base;
    float normal, p[3], q[3];
    void load(float *b) {
        base = b;
        normal = b[0];
        for (int k = 0; k < 3; k++) { p[k] = b[1 + k]; q[k] = b[4 + k]; }
    }
    void clear() {
        normal = 0;
        for (int k = 0; k < 3; k++) { p[k] = 0; q[k] = 0; }
    }
    void store() {
        base[0] = normal;
        for (int k = 0; k < 3; k++) { base[1 + k] = p[k]; base[4 + k] = q[k]; }
    }
};

inline float dot3(const float *x, const float *y) { return x[0] * y[0] + (x[1] * y[1] + x[2] * y[2]); }

inline void cross3(const float *x, const float *y, float *z) {
    for (int k = 0; k < 3; k++) {
        int k1 = (k + 1) % 3, k2 = (k + 2) % 3;
        z[k] = x[k1] * y[k2] - x[k2] * y[k1];
    }
}

// The round constants {checks, claims, pass, dispatch, mark, checksDispatch, rubs} (Spheres.java:1069-1075).
struct Round {
    int checks, claims, pass, dispatch, mark, checksDispatch, rubs;
    int slot;  // which claim array this dispatch claims into: checked = (slot+2)%3, cleared = (slot+1)%3
};

struct World {
    vr_step_desc d{};
    int n = 0, cap = 0, cells = 0;
    bool gs = false, grid = false, wide = false;
    bool friction = true;
    int slice = 0;  // SphereSimulation.slice (SphereSimulation.java:136-148)

    // SphereSimulation.submit (SphereSimulation.java:231-248) over SphereStep's lists (SphereStep.java:146-208,
    // 216-243, 348-392): predict 1; prepare = bin, CountingSort.scan (SupirVast's, not in the reference: taken as
    // three passes), order, and for Gauss-Seidel warmWalls, keepContacts, clearContacts, listContacts; velocity,
    // bounce, bounced; walls, show and a report 1 each. A list is cut into pieces of at most `slice` dispatches.
    static constexpr int SCAN_PASSES = 3, VELOCITY_PASSES = 3;
    int prepare_passes() const { return grid ? 1 + SCAN_PASSES + 1 + (gs ? 4 : 0) : 0; }
    int pieces(int passes) const { return slice == 0 ? 1 : (passes + slice - 1) / slice; }
    std::vector<int> opening, enough;  // SphereStepper.java:52-53

    std::map<std::string, std::vector<float>> F;
    std::map<std::string, std::vector<int32_t>> I;
    std::vector<int64_t> claims[3];

    // Buffer pointers, bound once the buffers are made (they never move).
    float *F_ax = nullptr, *F_ay = nullptr, *F_az = nullptr, *F_contactImpulse = nullptr, *F_cx = nullptr, *F_cy = nullptr, *F_cz = nullptr, *F_dx = nullptr, *F_dy = nullptr, *F_dz = nullptr, *F_im = nullptr, *F_keptImpulse = nullptr, *F_qw = nullptr, *F_qx = nullptr, *F_qy = nullptr, *F_qz = nullptr, *F_r = nullptr, *F_shown = nullptr, *F_tx = nullptr, *F_ty = nullptr, *F_tz = nullptr, *F_u = nullptr, *F_v = nullptr, *F_w = nullptr, *F_wallImpulse = nullptr, *F_x = nullptr, *F_y = nullptr, *F_z = nullptr, *F_params = nullptr;
    int32_t *I_contactA = nullptr, *I_contactB = nullptr, *I_contactCount = nullptr, *I_contactDone = nullptr, *I_contactHead = nullptr, *I_contactNext = nullptr, *I_counts = nullptr, *I_keptB = nullptr, *I_keptHead = nullptr, *I_keptNext = nullptr, *I_keys = nullptr, *I_open = nullptr, *I_order = nullptr, *I_ranks = nullptr, *I_readout = nullptr, *I_starts = nullptr;
    void bind() {
        F_ax = F.count("ax") ? F["ax"].data() : nullptr;
        F_ay = F.count("ay") ? F["ay"].data() : nullptr;
        F_az = F.count("az") ? F["az"].data() : nullptr;
        F_contactImpulse = F.count("contactImpulse") ? F["contactImpulse"].data() : nullptr;
        F_cx = F.count("cx") ? F["cx"].data() : nullptr;
        F_cy = F.count("cy") ? F["cy"].data() : nullptr;
        F_cz = F.count("cz") ? F["cz"].data() : nullptr;
        F_dx = F.count("dx") ? F["dx"].data() : nullptr;
        F_dy = F.count("dy") ? F["dy"].data() : nullptr;
        F_dz = F.count("dz") ? F["dz"].data() : nullptr;
        F_im = F.count("im") ? F["im"].data() : nullptr;
        F_keptImpulse = F.count("keptImpulse") ? F["keptImpulse"].data() : nullptr;
        F_qw = F.count("qw") ? F["qw"].data() : nullptr;
        F_qx = F.count("qx") ? F["qx"].data() : nullptr;
        F_qy = F.count("qy") ? F["qy"].data() : nullptr;
        F_qz = F.count("qz") ? F["qz"].data() : nullptr;
        F_r = F.count("r") ? F["r"].data() : nullptr;
        F_shown = F.count("shown") ? F["shown"].data() : nullptr;
        F_tx = F.count("tx") ? F["tx"].data() : nullptr;
        F_ty = F.count("ty") ? F["ty"].data() : nullptr;
        F_tz = F.count("tz") ? F["tz"].data() : nullptr;
        F_u = F.count("u") ? F["u"].data() : nullptr;
        F_v = F.count("v") ? F["v"].data() : nullptr;
        F_w = F.count("w") ? F["w"].data() : nullptr;
        F_wallImpulse = F.count("wallImpulse") ? F["wallImpulse"].data() : nullptr;
        F_x = F.count("x") ? F["x"].data() : nullptr;
        F_y = F.count("y") ? F["y"].data() : nullptr;
        F_z = F.count("z") ? F["z"].data() : nullptr;
        F_params = F.count("params") ? F["params"].data() : nullptr;
        I_contactA = I.count("contactA") ? I["contactA"].data() : nullptr;
        I_contactB = I.count("contactB") ? I["contactB"].data() : nullptr;
        I_contactCount = I.count("contactCount") ? I["contactCount"].data() : nullptr;
        I_contactDone = I.count("contactDone") ? I["contactDone"].data() : nullptr;
        I_contactHead = I.count("contactHead") ? I["contactHead"].data() : nullptr;
        I_contactNext = I.count("contactNext") ? I["contactNext"].data() : nullptr;
        I_counts = I.count("counts") ? I["counts"].data() : nullptr;
        I_keptB = I.count("keptB") ? I["keptB"].data() : nullptr;
        I_keptHead = I.count("keptHead") ? I["keptHead"].data() : nullptr;
        I_keptNext = I.count("keptNext") ? I["keptNext"].data() : nullptr;
        I_keys = I.count("keys") ? I["keys"].data() : nullptr;
        I_open = I.count("open") ? I["open"].data() : nullptr;
        I_order = I.count("order") ? I["order"].data() : nullptr;
        I_ranks = I.count("ranks") ? I["ranks"].data() : nullptr;
        I_readout = I.count("readout") ? I["readout"].data() : nullptr;
        I_starts = I.count("starts") ? I["starts"].data() : nullptr;
    }

    float P(int k) const { return F_params[k]; }

    // ---- SphereStep constructor: buffers (SphereStep.java:133-209, 216-262) ----
    void build(const vr_step_desc &desc) {
        d = desc;
        if (d.bodies < 1 || d.substeps < 1 || d.iterations < 1)
            throw ArgError("a step needs spheres, substeps and iterations, got " + std::to_string(d.bodies) + ", " +
                           std::to_string(d.substeps) + ", " + std::to_string(d.iterations));
        grid = d.has_grid != 0;
        gs = d.solve == VR_GAUSS_SEIDEL_BY_CONTACT;
        if (d.solve != VR_JACOBI && d.solve != VR_GAUSS_SEIDEL_BY_CONTACT) throw ArgError("unknown solve");
        if (gs && !grid) throw ArgError("Gauss–Seidel works over the grid's cells, and there is no grid");
        if (d.rounds < 1 || d.contacts_per_sphere < 1)
            throw ArgError("at least a round and a place a sphere, got " + std::to_string(d.rounds) + " and " +
                           std::to_string(d.contacts_per_sphere));
        wide = d.wide_pairs != 0;
        if (gs && !wide && d.bodies > MOST_LISTED_SPHERES)
            throw ArgError("a contact list orders at most 46340 spheres, got " + std::to_string(d.bodies));
        if (grid && (d.nx < 1 || d.ny < 1 || d.nz < 1 || !(d.cell > 0)))
            throw ArgError("a grid of at least one cell, of positive size");
        n = d.bodies;
        cap = n * d.contacts_per_sphere;
        static const char *sphere[] = {"x", "y", "z", "u", "v", "w", "r", "im", "cx", "cy", "cz", "dx",
                                       "dy", "dz", "ax", "ay", "az", "tx", "ty", "tz", "qx", "qy", "qz", "qw"};
        for (const char *s : sphere) F[s].assign(n, 0.f);
        F["params"].assign(PARAM_COUNT, 0.f);
        F["shown"].assign((size_t)SHOWN_STRIDE * n, 0.f);
        if (grid) {
            cells = d.nx * d.ny * d.nz;
            for (const char *s : {"keys", "ranks", "order"}) I[s].assign(n, 0);
            I["counts"].assign(cells + 1, 0);
            I["starts"].assign(cells + 1, 0);
        }
        if (gs) {
            for (auto &c : claims) c.assign(n, 0);
            for (const char *s : {"contactA", "contactB", "contactDone", "contactNext", "keptB", "keptNext"})
                I[s].assign(cap, 0);
            I["contactCount"].assign(3, 0);
            F["contactImpulse"].assign((size_t)cap * IMPULSE_WORDS, 0.f);
            F["keptImpulse"].assign((size_t)cap * IMPULSE_WORDS, 0.f);
            F["wallImpulse"].assign((size_t)n * WALL_IMPULSE_WORDS, 0.f);
            I["contactHead"].assign(n, 0);
            I["keptHead"].assign(n, 0);
            I["open"].assign(2, 0);
            I["readout"].assign(2, 0);
        }
        opening.assign(d.iterations, BATCH);
        enough.assign(d.iterations, 0);
        bind();
        clear();
    }

    // PassRunner.clear() and SphereStep.constants(): qw = 1, contactHead = -1 (SphereStep.java:157-160, 224-226).
    void clear() {
        for (auto &kv : F) std::fill(kv.second.begin(), kv.second.end(), 0.f);
        for (auto &kv : I) std::fill(kv.second.begin(), kv.second.end(), 0);
        for (auto &c : claims) std::fill(c.begin(), c.end(), 0);
        std::fill(F["qw"].begin(), F["qw"].end(), 1.f);
        if (gs) std::fill(I["contactHead"].begin(), I["contactHead"].end(), -1);
    }

    // ---- passes ----

    // Spheres.turn, Spheres.java:1468-1493.
    void turn(int s, const float angle[3]) {
        float *q[4] = {F_qx, F_qy, F_qz, F_qw};
        float a[3], was[4], now[4];
        for (int k = 0; k < 3; k++) a[k] = 0.5f * angle[k];
        for (int k = 0; k < 4; k++) was[k] = q[k][s];
        for (int k = 0; k < 3; k++) {
            int k1 = (k + 1) % 3, k2 = (k + 2) % 3;
            now[k] = was[k] + (a[k] * was[3] + (a[k1] * was[k2] - a[k2] * was[k1]));
        }
        now[3] = was[3] - (a[0] * was[0] + (a[1] * was[1] + a[2] * was[2]));
        float len2 = (now[0] * now[0] + now[1] * now[1]) + (now[2] * now[2] + now[3] * now[3]);
        float inverse = 1.0f / std::sqrt(std::max(len2, 1e-30f));
        for (int k = 0; k < 4; k++) q[k][s] = now[k] * inverse;
    }

    // Spheres.predict, Spheres.java:269-295.
    void predict() {
        float *at[3] = {F_x, F_y, F_z}, *vel[3] = {F_u, F_v, F_w};
        float *spin[3] = {F_ax, F_ay, F_az}, *im = F_im;
        float h = P(H), g[3] = {P(GX), P(GY), P(GZ)};
        for (int s = 0; s < n; s++) {
            if (!(im[s] > 0)) continue;
            for (int a = 0; a < 3; a++) {
                float v = vel[a][s] + h * g[a];
                vel[a][s] = v;
                at[a][s] = at[a][s] + h * v;
            }
            float angle[3] = {h * spin[0][s], h * spin[1][s], h * spin[2][s]};
            turn(s, angle);
        }
    }

    // Spheres.bin (Spheres.java:1308-1323), then a stable counting sort standing in for CountingSort.count / scan /
    // order (SphereStep.java:179-182): starts[c]..starts[c+1] is cell c's run of `order`.
    void bin_and_sort() {
        float *at[3] = {F_x, F_y, F_z};
        int32_t *keys = I_keys, *counts = I_counts, *starts = I_starts, *order = I_order,
                *ranks = I_ranks;
        int cellsAxis[3] = {d.nx, d.ny, d.nz};
        std::fill(counts, counts + cells + 1, 0);
        for (int s = 0; s < n; s++) {
            int c[3];
            for (int a = 0; a < 3; a++) {
                float fl = std::floor(at[a][s] / d.cell);
                c[a] = (int)std::max(0.f, std::min((float)(cellsAxis[a] - 1), fl));
            }
            int key = c[0] + d.nx * (c[1] + d.ny * c[2]);
            keys[s] = key;
            ranks[s] = counts[key]++;
        }
        int run = 0;
        for (int c = 0; c <= cells; c++) {
            starts[c] = run;
            run += counts[c];
        }
        for (int s = 0; s < n; s++) order[starts[keys[s]] + ranks[s]] = s;
    }

    // Spheres.around, Spheres.java:1207-1229.
    template <class Visit>
    void around(int key, Visit visit) {
        int32_t *starts = I_starts, *order = I_order;
        int ix = key % d.nx, iy = (key / d.nx) % d.ny, iz = key / (d.nx * d.ny);
        for (int m = 0; m < 27; m++) {
            int jx = ix + (m % 3 - 1), jy = iy + ((m / 3) % 3 - 1), jz = iz + (m / 9 - 1);
            if (jx < 0 || jx >= d.nx || jy < 0 || jy >= d.ny || jz < 0 || jz >= d.nz) continue;
            int cell = jx + d.nx * (jy + d.ny * jz);
            for (int k = starts[cell]; k < starts[cell + 1]; k++) visit(order[k]);
        }
    }

    Held hold(int s, bool turns) {
        Held h;
        h.s = s;
        h.turns = turns;
        float *at[3] = {F_x, F_y, F_z}, *dsp[3] = {F_dx, F_dy, F_dz},
              *tr[3] = {F_tx, F_ty, F_tz};
        for (int k = 0; k < 3; k++) {
            h.at[k] = at[k][s];
            h.displaced[k] = dsp[k][s];
            h.turned[k] = turns ? tr[k][s] : 0.f;
        }
        return h;
    }
    void store(const Held &h) {
        float *at[3] = {F_x, F_y, F_z}, *dsp[3] = {F_dx, F_dy, F_dz},
              *tr[3] = {F_tx, F_ty, F_tz};
        for (int k = 0; k < 3; k++) {
            at[k][h.s] = h.at[k];
            dsp[k][h.s] = h.displaced[k];
            if (h.turns) tr[k][h.s] = h.turned[k];
        }
    }

    static float turns_per_impulse(float w, float lever, float radius) {  // Spheres.java:743-745
        return (w * lever) / (INERTIA * (radius * radius));
    }

    // Spheres.rub, Spheres.java:662-677.
    static void rub(Held &a, float wa, float turnsA, Held *o, float wo, float turnsO, const float n[3],
                    const float p[3]) {
        float np[3];
        cross3(n, p, np);
        for (int k = 0; k < 3; k++) {
            float move = wa * p[k];
            a.at[k] = a.at[k] + move;
            a.displaced[k] = a.displaced[k] + move;
            a.turned[k] = a.turned[k] - turnsA * np[k];
            if (o) {
                float back = wo * p[k];
                o->at[k] = o->at[k] - back;
                o->displaced[k] = o->displaced[k] - back;
                o->turned[k] = o->turned[k] - turnsO * np[k];
            }
        }
    }

    // Spheres.contactTravel, Spheres.java:751-765.
    void contact_travel(const Held &s, float h, const float n[3], float lever, int side, float travel[3]) {
        float *vel[3] = {F_u, F_v, F_w}, *spin[3] = {F_ax, F_ay, F_az};
        float turn[3], turnN[3];
        for (int k = 0; k < 3; k++) turn[k] = h * spin[k][s.s] + s.turned[k];
        cross3(turn, n, turnN);
        for (int k = 0; k < 3; k++) {
            float centre = h * vel[k][s.s] + s.displaced[k];
            float spun = lever * turnN[k];
            travel[k] = side > 0 ? centre + spun : centre - spun;
        }
    }

    // Spheres.friction, Spheres.java:595-654. o == nullptr for a wall.
    void friction_at(Held &a, float wa, float leverA, Held *o, float wo, float leverO, const float n[3],
                     Impulse &taken) {
        float muS = P(MU_STATIC), muK = P(MU_KINETIC);
        float *r = F_r;
        if (!(muS + muK > 0)) return;
        float h = P(H);
        float la = leverA;
        float slip[3];
        contact_travel(a, h, n, la, -1, slip);
        float ia = wa;
        float leverOverIa = turns_per_impulse(ia, la, r[a.s]);
        float weight = ia + leverOverIa * la;
        float io = 0, lo = 0, leverOverIo = 0;
        if (o) {
            io = wo;
            lo = leverO;
            float other[3];
            contact_travel(*o, h, n, lo, 1, other);
            for (int k = 0; k < 3; k++) slip[k] = slip[k] - other[k];
            leverOverIo = turns_per_impulse(io, lo, r[o->s]);
            weight = weight + (io + leverOverIo * lo);
        }
        float along = dot3(slip, n);
        float tangent[3];
        for (int k = 0; k < 3; k++) tangent[k] = slip[k] - along * n[k];
        if (!(weight > 0)) return;
        float takenAlong = dot3(taken.p, n);
        float before[3], whole[3];
        for (int k = 0; k < 3; k++) {
            before[k] = taken.p[k] - takenAlong * n[k];
            whole[k] = before[k] - tangent[k] / weight;
        }
        float size2 = dot3(whole, whole);
        float lambdaN = taken.normal;
        float holds = muS * lambdaN;
        if (size2 > holds * holds) {
            float cut = std::min(1.0f, (muK * lambdaN) / std::sqrt(size2));
            for (int k = 0; k < 3; k++) whole[k] = whole[k] * cut;
        }
        float p[3];
        for (int k = 0; k < 3; k++) {
            p[k] = whole[k] - before[k];
            taken.p[k] = whole[k];
        }
        rub(a, ia, leverOverIa, o, io, leverOverIo, n, p);
    }

    // Spheres.roll, Spheres.java:699-741. o == nullptr for a wall.
    void roll(Held &a, float wa, Held *o, float wo, float reach, Impulse &taken) {
        float *r = F_r, *spin[3] = {F_ax, F_ay, F_az};
        float muR = P(MU_ROLLING), h = P(H);
        float ra = r[a.s];
        float ia = wa / (INERTIA * (ra * ra));
        float inverse = ia;
        float io = 0;
        float against[3];
        for (int k = 0; k < 3; k++) against[k] = h * spin[k][a.s] + a.turned[k];
        if (o) {
            float ro = r[o->s];
            io = wo / (INERTIA * (ro * ro));
            inverse = inverse + io;
            for (int k = 0; k < 3; k++) against[k] = against[k] - (h * spin[k][o->s] + o->turned[k]);
        }
        if (!(inverse > 0)) return;
        float whole[3];
        for (int k = 0; k < 3; k++) whole[k] = taken.q[k] - against[k] / inverse;
        float size2 = dot3(whole, whole);
        float holds = (muR * reach) * taken.normal;
        if (size2 > holds * holds) {
            float cut = std::min(1.0f, holds / std::sqrt(size2));
            for (int k = 0; k < 3; k++) whole[k] = whole[k] * cut;
        }
        for (int k = 0; k < 3; k++) {
            float change = whole[k] - taken.q[k];
            taken.q[k] = whole[k];
            a.turned[k] = a.turned[k] + ia * change;
            if (o) o->turned[k] = o->turned[k] - io * change;
        }
    }

    float rubbing() { return P(MU_STATIC) + (P(MU_KINETIC) + P(MU_ROLLING)); }  // Spheres.java:684-687

    // Spheres.warm, Spheres.java:448-481.
    void warm(Held &a, float wa, Held &o, float wo, const float d[3], float d2, float reach, Impulse &taken) {
        float *r = F_r;
        float dist = std::sqrt(d2);
        float n[3];
        for (int k = 0; k < 3; k++) n[k] = d[k] / dist;
        for (int k = 0; k < 3; k++) {
            float push = taken.normal * n[k];
            float ma = push * wa, mo = -(push * wo);
            a.at[k] = a.at[k] + ma;
            o.at[k] = o.at[k] + mo;
            a.displaced[k] = a.displaced[k] + ma;
            o.displaced[k] = o.displaced[k] + mo;
        }
        float along = dot3(taken.p, n);
        for (int k = 0; k < 3; k++) taken.p[k] = taken.p[k] - along * n[k];
        float share = dist / reach;
        float ra = r[a.s], ro = r[o.s];
        float turnsA = turns_per_impulse(wa, share * ra, ra);
        float turnsO = turns_per_impulse(wo, share * ro, ro);
        rub(a, wa, turnsA, &o, wo, turnsO, n, taken.p);
        float ia = wa / (INERTIA * (ra * ra));
        float io = wo / (INERTIA * (ro * ro));
        for (int k = 0; k < 3; k++) {
            a.turned[k] = a.turned[k] + ia * taken.q[k];
            o.turned[k] = o.turned[k] - io * taken.q[k];
        }
    }

    // Spheres.contact, Spheres.java:369-439.
    void contact(int c, int a, int o, float omega, int pass, int rubs) {
        float *r = F_r, *im = F_im;
        Held ha = hold(a, true), ho = hold(o, true);
        float d[3];
        for (int k = 0; k < 3; k++) d[k] = ha.at[k] - ho.at[k];
        float d2 = dot3(d, d);
        float reach = r[a] + r[o];
        float wa = im[a], wo = im[o];
        float wsum = wa + wo;
        float listed = reach * one_plus_margin();
        if (!(d2 < listed * listed) || !(d2 > COINCIDENT) || !(wsum > 0)) return;
        Impulse taken;
        taken.load(F_contactImpulse + (size_t)c * IMPULSE_WORDS);
        if (rubs == 0) {
            float rb = rubbing();
            if (pass == 1) {
                if (rb > 0) warm(ha, wa, ho, wo, d, d2, reach, taken);
                else taken.clear();
            }
            for (int k = 0; k < 3; k++) d[k] = ha.at[k] - ho.at[k];
            float dist = std::sqrt(std::max(dot3(d, d), COINCIDENT));
            float depth = std::max(-(taken.normal * wsum), omega * (reach - dist));
            taken.normal = taken.normal + depth / wsum;
            float kk = depth / (dist * wsum);
            for (int k = 0; k < 3; k++) {
                float step = kk * d[k];
                float ma = step * wa, mo = -(step * wo);
                ha.at[k] = ha.at[k] + ma;
                ho.at[k] = ho.at[k] + mo;
                ha.displaced[k] = ha.displaced[k] + ma;
                ho.displaced[k] = ho.displaced[k] + mo;
            }
        } else {
            float dist = std::sqrt(d2);
            float nn[3];
            for (int k = 0; k < 3; k++) nn[k] = d[k] / dist;
            float share = dist / reach;
            float pairRadius = (r[a] * r[o]) / reach;
            roll(ha, wa, &ho, wo, pairRadius, taken);
            friction_at(ha, wa, share * r[a], &ho, wo, share * r[o], nn, taken);
        }
        taken.store();
        store(ha);
        store(ho);
    }

    // Spheres.warmWalls, Spheres.java:890-939.
    void warm_walls() {
        float *r = F_r, *im = F_im, *wi = F_wallImpulse;
        float ext[3] = {P(SX), P(SY), P(SZ)};
        float rb = rubbing();
        for (int s = 0; s < n; s++) {
            if (!(im[s] > 0)) continue;
            float rs = r[s], w = im[s];
            float turns = turns_per_impulse(w, rs, rs);
            Held held = hold(s, true);
            for (int axis = 0; axis < 3; axis++) {
                Impulse taken;
                taken.load(wi + (size_t)s * WALL_IMPULSE_WORDS + axis * IMPULSE_WORDS);
                float reach = rs * one_plus_margin();
                float position = held.at[axis];
                float out = 0;
                if (position < reach) out = 1;
                if (position > ext[axis] - reach) out = -1;
                bool near = rb > 0 && out != 0;
                if (!near) taken.clear();
                if (near) {
                    float push = (taken.normal * w) * out;
                    held.at[axis] = held.at[axis] + push;
                    held.displaced[axis] = held.displaced[axis] + push;
                    taken.p[axis] = 0;
                    float nrm[3] = {0, 0, 0};
                    nrm[axis] = out;
                    rub(held, w, turns, nullptr, 0, 0, nrm, taken.p);
                    for (int k = 0; k < 3; k++) held.turned[k] = held.turned[k] + (turns * taken.q[k]) / rs;
                }
                taken.store();
            }
            store(held);
        }
    }

    // Spheres.keepContacts, Spheres.java:859-878.
    void keep_contacts() {
        int32_t *cnt = I_contactCount;
        int listed = std::min(cnt[LISTED], cap);
        int32_t *cB = I_contactB, *cN = I_contactNext, *cH = I_contactHead;
        int32_t *kB = I_keptB, *kN = I_keptNext, *kH = I_keptHead;
        float *ci = F_contactImpulse, *ki = F_keptImpulse;
        for (int c = 0; c < listed; c++) {
            kB[c] = cB[c];
            kN[c] = cN[c];
            for (int k = 0; k < IMPULSE_WORDS; k++) ki[(size_t)c * 7 + k] = ci[(size_t)c * 7 + k];
        }
        for (int s = 0; s < n; s++) {
            kH[s] = cH[s];
            cH[s] = -1;
        }
    }

    // Spheres.clearContacts, Spheres.java:829-844.
    void clear_contacts() {
        for (auto &c : claims) std::fill(c.begin(), c.end(), 0);
        int32_t *cnt = I_contactCount;
        cnt[LONGEST] = std::max(cnt[LONGEST], cnt[LISTED]);
        cnt[LISTED] = 0;
    }

    // Spheres.listContacts, Spheres.java:961-1022. Invocations in order a = 0, 1, ...
    void list_contacts() {
        float *at[3] = {F_x, F_y, F_z}, *r = F_r;
        int32_t *keys = I_keys, *cnt = I_contactCount;
        int32_t *cA = I_contactA, *cB = I_contactB, *cD = I_contactDone, *cN = I_contactNext,
                *cH = I_contactHead, *kB = I_keptB, *kN = I_keptNext, *kH = I_keptHead;
        float *ci = F_contactImpulse, *ki = F_keptImpulse;
        float m = one_plus_margin();
        for (int a = 0; a < n; a++) {
            float mine[3] = {at[0][a], at[1][a], at[2][a]};
            float ra = r[a];
            around(keys[a], [&](int o) {
                if (!(o > a)) return;
                float dd[3];
                for (int k = 0; k < 3; k++) dd[k] = mine[k] - at[k][o];
                float d2 = dd[0] * dd[0] + (dd[1] * dd[1] + dd[2] * dd[2]);
                float near = (ra + r[o]) * m;
                if (!(d2 < near * near)) return;
                int slot = cnt[LISTED]++;
                if (slot >= cap) return;
                cA[slot] = a;
                cB[slot] = o;
                cD[slot] = 0;
                cN[slot] = cH[a];
                cH[a] = slot;
                float warmth[IMPULSE_WORDS] = {0, 0, 0, 0, 0, 0, 0};
                int j = kH[a], walked = 0;
                while (j >= 0 && walked < MOST_KEPT) {
                    if (kB[j] == o) {
                        for (int k = 0; k < IMPULSE_WORDS; k++) warmth[k] = ki[(size_t)j * 7 + k];
                        walked = MOST_KEPT;
                    }
                    j = kN[j];
                    walked++;
                }
                for (int k = 0; k < IMPULSE_WORDS; k++) ci[(size_t)slot * 7 + k] = warmth[k];
            });
        }
    }

    // Spheres.priority, Spheres.java:1166-1168; the 63-bit form when wide.
    int64_t priority(int a, int o, int g) const {
        if (!wide) {
            uint32_t pair = (uint32_t)a * (uint32_t)n + (uint32_t)o;
            uint32_t p = pair * 0x9E3779B1u + (uint32_t)g * 0x632BE5ABu;
            return (int64_t)(p & 0x7FFFFFFFu);
        }
        uint64_t pair = (uint64_t)a * (uint64_t)n + (uint64_t)o;
        uint64_t p = pair * 0x9E3779B97F4A7C15ull + (uint64_t)(int64_t)g * 0x632BE59BD9B4E019ull;
        return (int64_t)(p & 0x7FFFFFFFFFFFFFFFull);
    }

    // Spheres.contactRound, Spheres.java:1076-1133: one dispatch, invocations in order c = 0, 1, ...
    void contact_round(const Round &rd) {
        int32_t *marks = I_open, *cnt = I_contactCount, *done = I_contactDone, *cA = I_contactA,
                *cB = I_contactB;
        std::vector<int64_t> &checked = claims[(rd.slot + 2) % 3], &claimed = claims[rd.slot % 3],
                             &cleared = claims[(rd.slot + 1) % 3];
        float omega = P(OMEGA);
        marks[1 - rd.mark] = 0;
        int listed = std::min(cnt[LISTED], cap);
        for (int c = 0; c < listed; c++) {
            if (done[c] == rd.pass) continue;
            int a = cA[c], o = cB[c];
            bool won = false;
            if (rd.checks > 0) {
                int64_t last = priority(a, o, rd.checksDispatch);
                won = checked[a] == last && checked[o] == last;
            }
            if (won) {
                contact(c, a, o, omega, rd.pass, rd.rubs);
                done[c] = rd.pass;
            } else {
                cleared[a] = 0;
                cleared[o] = 0;
                if (rd.claims > 0) {
                    marks[rd.mark] = 1;
                    int64_t mine = priority(a, o, rd.dispatch);
                    claimed[a] = std::max(claimed[a], mine);
                    claimed[o] = std::max(claimed[o], mine);
                } else {
                    cnt[MISSED]++;
                }
            }
        }
    }

    // Spheres.walls -> shift with friction, Spheres.java:1185-1198, 1351-1418.
    void walls() {
        float *r = F_r, *im = F_im, *wi = F_wallImpulse;
        float ext[3] = {P(SX), P(SY), P(SZ)};
        for (int s = 0; s < n; s++) {
            if (!(im[s] > 0)) continue;
            float rs = r[s];
            Held held = hold(s, true);
            for (int a = 0; a < 3; a++) {
                float c = 0.f;
                float moved = held.at[a] + c;
                float high = ext[a] - rs;
                float side = 1;
                float into = rs - moved;
                if (moved * 2.0f > ext[a]) {
                    side = -1;
                    into = moved - high;
                }
                float w = im[s];
                Impulse taken;
                taken.load(wi + (size_t)s * WALL_IMPULSE_WORDS + a * IMPULSE_WORDS);
                float inside = std::max(rs, std::min(high, moved));
                float push = std::max(into, 0.f);
                if (!(into > 0) && taken.normal > 0) {
                    push = std::max(into, -(taken.normal * w));
                    inside = moved + side * push;
                }
                taken.normal = taken.normal + push / w;
                held.at[a] = inside;
                held.displaced[a] = held.displaced[a] + (c + (inside - moved));
                float reach = rs * LIST_MARGIN;
                float out = 0;
                if (!(inside > rs + reach)) out = 1;
                if (!(inside < high - reach)) out = -1;
                if (out != 0) {
                    float nrm[3] = {0, 0, 0};
                    nrm[a] = out;
                    roll(held, w, nullptr, 0, rs, taken);
                    friction_at(held, w, rs, nullptr, 0, 0, nrm, taken);
                }
                taken.store();
            }
            for (int a = 0; a < 3; a++) {
                float high = ext[a] - rs;
                float inside = std::max(rs, std::min(high, held.at[a]));
                held.displaced[a] = held.displaced[a] + (inside - held.at[a]);
                held.at[a] = inside;
            }
            store(held);
        }
    }

    // Spheres.velocity, Spheres.java:1430-1460.
    void velocity() {
        float *vel[3] = {F_u, F_v, F_w}, *dsp[3] = {F_dx, F_dy, F_dz},
              *bef[3] = {F_cx, F_cy, F_cz}, *spin[3] = {F_ax, F_ay, F_az},
              *tr[3] = {F_tx, F_ty, F_tz};
        float perH = 1.0f / P(H);
        for (int s = 0; s < n; s++) {
            for (int a = 0; a < 3; a++) {
                float was = vel[a][s];
                bef[a][s] = was;
                vel[a][s] = was + dsp[a][s] * perH;
                dsp[a][s] = 0;
            }
            float angle[3];
            for (int a = 0; a < 3; a++) {
                float by = tr[a][s];
                spin[a][s] = spin[a][s] + by * perH;
                tr[a][s] = 0;
                angle[a] = by;
            }
            turn(s, angle);
        }
    }

    // Spheres.bounce, Spheres.java:1523-1627; every pair without a grid, the 27 cells with one.
    void bounce() {
        float *at[3] = {F_x, F_y, F_z}, *vel[3] = {F_u, F_v, F_w},
              *bef[3] = {F_cx, F_cy, F_cz}, *chg[3] = {F_dx, F_dy, F_dz};
        float *r = F_r, *im = F_im;
        float e = P(RESTITUTION);
        float ext[3] = {P(SX), P(SY), P(SZ)};
        float m = one_plus_margin();
        for (int s = 0; s < n; s++) {
            float wa = im[s];
            if (!(e > 0 && wa > 0)) continue;
            float h = P(H), gx = P(GX), gy = P(GY), gz = P(GZ);
            float resting = (2.0f * h) * std::sqrt(gx * gx + (gy * gy + gz * gz));
            float ra = r[s];
            float mine[3], v[3], was[3], sum[3];
            for (int a = 0; a < 3; a++) {
                mine[a] = at[a][s];
                v[a] = vel[a][s];
                was[a] = bef[a][s];
                sum[a] = 0;
            }
            auto against = [&](int o) {
                if (o == s) return;
                float dd[3];
                for (int a = 0; a < 3; a++) dd[a] = mine[a] - at[a][o];
                float d2 = dd[0] * dd[0] + (dd[1] * dd[1] + dd[2] * dd[2]);
                float near = (ra + r[o]) * m;
                float wsum = wa + im[o];
                if (!(d2 < near * near && d2 > COINCIDENT)) return;
                float dist = std::sqrt(d2);
                float closing = 0, parting = 0;
                for (int a = 0; a < 3; a++) {
                    float nn = dd[a] / dist;
                    closing = closing + nn * (was[a] - bef[a][o]);
                    parting = parting + nn * (v[a] - vel[a][o]);
                }
                if (closing < -resting) {
                    float k = ((-(e * closing)) - parting) * wa / (wsum * dist);
                    for (int a = 0; a < 3; a++) sum[a] = sum[a] + k * dd[a];
                }
            };
            if (!grid) {
                for (int o = 0; o < n; o++) against(o);
            } else {
                around(I_keys[s], against);
            }
            for (int a = 0; a < 3; a++) {
                float reach = ra * m;
                if (mine[a] < reach && was[a] < -resting) sum[a] = sum[a] + ((-(e * was[a])) - v[a]);
                if (mine[a] > ext[a] - reach && was[a] > resting) sum[a] = sum[a] + ((-(e * was[a])) - v[a]);
            }
            for (int a = 0; a < 3; a++) chg[a][s] = sum[a];
        }
    }

    // Spheres.bounced, Spheres.java:1633-1644.
    void bounced() {
        float *vel[3] = {F_u, F_v, F_w}, *dsp[3] = {F_dx, F_dy, F_dz};
        for (int s = 0; s < n; s++)
            for (int a = 0; a < 3; a++) {
                vel[a][s] = vel[a][s] + dsp[a][s];
                dsp[a][s] = 0;
            }
    }

    // Spheres.solve / solve(grid) + Solving, Spheres.java:302-347, 1245-1298: the Jacobi move into cx, cy, cz.
    void jacobi_solve() {
        float *at[3] = {F_x, F_y, F_z}, *cor[3] = {F_cx, F_cy, F_cz};
        float *r = F_r, *im = F_im;
        std::vector<float> out(3 * (size_t)n);
        for (int s = 0; s < n; s++) {
            float mine[3] = {at[0][s], at[1][s], at[2][s]};
            float rs = r[s], ws = im[s];
            float sum[3] = {0, 0, 0}, contacts = 0;
            auto against = [&](int o) {
                if (o == s) return;
                float dd[3];
                for (int a = 0; a < 3; a++) dd[a] = mine[a] - at[a][o];
                float d2 = dd[0] * dd[0] + (dd[1] * dd[1] + dd[2] * dd[2]);
                float reach = rs + r[o];
                float wsum = ws + im[o];
                if (!(d2 < reach * reach) || !(d2 > COINCIDENT) || !(wsum > 0)) return;
                float dist = std::sqrt(d2);
                float share = ((reach - dist) * (ws / wsum)) / dist;
                for (int a = 0; a < 3; a++) sum[a] = sum[a] + share * dd[a];
                contacts = contacts + 1.0f;
            };
            if (!grid) {
                for (int o = 0; o < n; o++) against(o);
            } else {
                around(I_keys[s], against);
            }
            float omega = P(OMEGA), averaged = P(AVERAGED);
            float scale = omega;
            if (averaged > 0) scale = omega / std::max(contacts, 1.0f);
            for (int a = 0; a < 3; a++) out[(size_t)a * n + s] = scale * sum[a];
        }
        for (int a = 0; a < 3; a++) std::copy(out.begin() + (size_t)a * n, out.begin() + (size_t)(a + 1) * n, cor[a]);
    }

    // Spheres.apply -> shift without friction, Spheres.java:1326-1342, 1351-1368.
    void jacobi_apply() {
        float *cor[3] = {F_cx, F_cy, F_cz};
        float *r = F_r, *im = F_im;
        float ext[3] = {P(SX), P(SY), P(SZ)};
        for (int s = 0; s < n; s++) {
            if (!(im[s] > 0)) continue;
            float rs = r[s];
            Held held = hold(s, false);
            for (int a = 0; a < 3; a++) {
                float c = cor[a][s];
                float moved = held.at[a] + c;
                float high = ext[a] - rs;
                float inside = std::max(rs, std::min(high, moved));
                held.at[a] = inside;
                float push = inside - moved;
                held.displaced[a] = held.displaced[a] + (c + push);
            }
            store(held);
        }
    }

    // Spheres.show, Spheres.java:1697-1794.
    void show() {
        float *at[3] = {F_x, F_y, F_z}, *r = F_r, *shown = F_shown;
        float *q[4] = {F_qx, F_qy, F_qz, F_qw};
        float ext[3] = {P(SX), P(SY), P(SZ)};
        for (int s = 0; s < n; s++) {
            float *b = shown + (size_t)s * SHOWN_STRIDE;
            float mine[3];
            for (int a = 0; a < 3; a++) {
                mine[a] = at[a][s];
                b[a] = b[4 + a];
                b[4 + a] = mine[a];
            }
            float rs = r[s];
            b[3] = rs;
            for (int k = 0; k < PRESSING_WORDS; k++) b[PRESSED_BEFORE + k] = b[PRESSED_AFTER + k];
            for (int k = 0; k < 4; k++) {
                b[TURNED_BEFORE + k] = b[TURNED_AFTER + k];
                b[TURNED_AFTER + k] = q[k][s];
            }
            float flat[6] = {0, 0, 0, 0, 0, 0}, lean[3] = {0, 0, 0};
            for (int a = 0; a < 3; a++) {
                float low = rs - mine[a];
                float high = (mine[a] + rs) - ext[a];
                float depths[2] = {low, high};
                for (int side = 0; side < 2; side++) {
                    float depth = depths[side];
                    if (depth > 0) {
                        float share = depth / rs;
                        flat[a] = flat[a] + share;
                        lean[a] = lean[a] + (side == 0 ? -1.0f : 1.0f) * share;
                    }
                }
            }
            static const int pairs[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {0, 2}, {1, 2}};
            auto against = [&](int o) {
                if (o == s) return;
                float dd[3];
                for (int a = 0; a < 3; a++) dd[a] = at[a][o] - mine[a];
                float d2 = dot3(dd, dd);
                float reach = rs + r[o];
                if (!(d2 < reach * reach && d2 > COINCIDENT)) return;
                float dist = std::sqrt(d2);
                float share = (reach - dist) / reach;
                float g = share / d2;
                float along = share / dist;
                for (int k = 0; k < 6; k++) flat[k] = flat[k] + g * (dd[pairs[k][0]] * dd[pairs[k][1]]);
                for (int a = 0; a < 3; a++) lean[a] = lean[a] + along * dd[a];
            };
            if (!grid) {
                for (int o = 0; o < n; o++) against(o);
            } else {
                around(I_keys[s], against);
            }
            float *after = b + PRESSED_AFTER;
            for (int k = 0; k < 6; k++) after[k] = flat[k];
            for (int a = 0; a < 3; a++) after[LEAN + a] = lean[a];
            after[TRACE] = (flat[0] + flat[1]) + flat[2];
            after[LEAN_LENGTH] = std::sqrt(dot3(lean, lean));
        }
    }

    // ---- the schedule ----

    // SphereStep.position, SphereStep.java:333-339.
    static int position(int k) { return k == 0 ? 0 : 1 + (k - 1) % CYCLE; }

    // A round at cycle position p of pass `it`: SphereStep.cycle, SphereStep.java:286-315, and roundNames :320-327.
    Round cycle_round(int it, int p, bool rubs) const {
        Round rd;
        rd.checks = p == 0 ? 0 : 1;
        rd.claims = 1;
        rd.pass = rubs ? d.iterations + it + 1 : it + 1;
        rd.dispatch = it * CYCLE + p % CYCLE;
        rd.mark = p % 2;
        rd.checksDispatch = it * CYCLE + std::max(p - 1, 0) % CYCLE;
        rd.rubs = rubs ? 1 : 0;
        rd.slot = p % 3;
        return rd;
    }

    void prepare() {  // SphereStep.java:170-182, 227-235
        if (grid) bin_and_sort();
        if (gs) {
            warm_walls();
            keep_contacts();
            clear_contacts();
            list_contacts();
        }
    }

    void velocity_list() {  // SphereStep.java:150-159
        velocity();
        bounce();
        bounced();
    }

    void rubs(int iteration, int rounds) {  // SphereStep.java:270-276
        for (int r = 0; r < rounds; r++) contact_round(cycle_round(iteration, position(r), true));
    }

    // addRounds + report: the rounds, then what the host reads (SphereStep.java:394-402; Spheres.java:1146-1159).
    int add_rounds(int it, int from, int rounds, int *listedOut) {
        for (int k = from; k < from + rounds; k++) contact_round(cycle_round(it, position(k), false));
        int mark = position(from + rounds - 1) % 2;
        int32_t *readout = I_readout;
        readout[0] = I_open[mark];
        readout[1] = I_contactCount[LISTED];
        *listedOut = readout[1];
        return readout[0];
    }

    // SphereStepper.step, SphereStepper.java:66-103, with SphereStep.opening/more/closing (:348-392).
    void step_until_done(vr_step_report *rep) {
        int waits = 0, rounds = 0, most = 0, unlisted = 0, rubbed = 0, submissions = 0;
        int limit = cap + 2 * CYCLE;
        for (int sub = 0; sub < d.substeps; sub++) {
            for (int it = 0; it < d.iterations; it++) {
                int k = opening[it];
                rubs(it == 0 ? d.iterations - 1 : it - 1, rubbed);
                if (it > 0) {
                    walls();
                } else {
                    if (sub > 0) {
                        walls();
                        velocity_list();
                    }
                    predict();
                    prepare();
                }
                int listed = 0;
                int open = add_rounds(it, 0, k, &listed);
                // SphereStep.opening: the rubbing rounds, the walls before (or the substep's close, predict and
                // prepare), the rounds and a report.
                int prefix = it > 0 ? 1 : (sub > 0 ? 1 + VELOCITY_PASSES : 0) + 1 + prepare_passes();
                submissions += pieces(rubbed + prefix + k + 1);
                waits++;
                while (open != 0) {
                    if (k > limit)
                        throw StateError("a pass is still open after " + std::to_string(k) +
                                         " rounds, on a list of " + std::to_string(listed));
                    open = add_rounds(it, k, BATCH, &listed);
                    submissions += pieces(BATCH + 1);  // SphereStep.more
                    k += BATCH;
                    waits++;
                }
                rubbed = friction ? k : 0;
                rounds += k;
                most = std::max(most, k);
                unlisted += std::max(0, listed - cap);
                adapt(it, k);
            }
        }
        rubs(d.iterations - 1, rubbed);
        walls();
        velocity_list();
        show();
        if (rep) {
            rep->waits = waits;
            rep->submissions = submissions + pieces(rubbed + 1 + VELOCITY_PASSES + 1);  // SphereStep.closing
            rep->rounds = rounds;
            rep->most_rounds = most;
            rep->unlisted = unlisted;
        }
    }

    void adapt(int it, int ran) {  // SphereStepper.java:118-126
        if (ran > opening[it]) {
            opening[it] = std::min(ran, CYCLE);
            enough[it] = 0;
        } else if (++enough[it] >= SHRINK_AFTER && opening[it] > BATCH) {
            opening[it] -= BATCH;
            enough[it] = 0;
        }
    }

    // SphereStep.step(): the one list (SphereStep.java:184-208), Jacobi, or Gauss-Seidel in fixed rounds (:245-260).
    void step_list() {
        for (int sub = 0; sub < d.substeps; sub++) {
            predict();
            prepare();
            if (!gs) {
                for (int it = 0; it < d.iterations; it++) {
                    jacobi_solve();
                    jacobi_apply();
                }
            } else {
                int g = 0;
                for (int it = 0; it < d.iterations; it++) {
                    for (int k = 0; k <= d.rounds; k++, g++) {
                        Round rd;
                        rd.checks = k == 0 ? 0 : 1;
                        rd.claims = k == d.rounds ? 0 : 1;
                        rd.pass = it + 1;
                        rd.dispatch = g;
                        rd.mark = g % 2;
                        rd.checksDispatch = g - 1;
                        rd.rubs = 0;
                        rd.slot = g % 3;
                        contact_round(rd);
                    }
                    walls();
                }
            }
            velocity_list();
        }
        show();
    }

    void step(vr_step_report *rep) {
        if (rep) std::memset(rep, 0, sizeof(*rep));
        auto t0 = std::chrono::steady_clock::now();
        if (gs && !d.fixed_rounds) {
            step_until_done(rep);
        } else {
            step_list();
            int iteration = gs ? d.rounds + 2 : 2;  // SphereStep.step(), one list
            if (rep)
                rep->submissions =
                    pieces(d.substeps * (1 + prepare_passes() + d.iterations * iteration + VELOCITY_PASSES) + 1);
        }
        auto t1 = std::chrono::steady_clock::now();
        if (rep) rep->wall_ns = std::chrono::duration_cast<std::chrono::nanoseconds>(t1 - t0).count();
    }
};

int fail(int code, const std::string &msg) {
    g_error = msg;
    return code;
}

template <class F>
int guard(F fn) {
    try {
        fn();
        return VR_OK;
    } catch (const ArgError &e) {
        return fail(VR_ERR_ARGUMENT, e.what());
    } catch (const StateError &e) {
        return fail(VR_ERR_STATE, e.what());
    } catch (const std::out_of_range &e) {
        return fail(VR_ERR_NAME, "no such buffer");
    } catch (const std::exception &e) {
        return fail(VR_ERR_STATE, e.what());
    }
}

}  // namespace

struct vro_world {
    int kind = VR_SPHERES;
    World w;
    vro::BoxWorld *box = nullptr;
    // A batch: each scene its own box world, as the reference would step it alone, walled by its own box.
    std::vector<vro::BoxWorld *> batch;
    std::vector<int64_t> batchFirst;  // first body of each scene, and the total after the last
    std::vector<float> batchBox;      // 3 a scene
    ~vro_world() {
        if (box) vro::box_destroy(box);
        for (auto *b : batch) vro::box_destroy(b);
    }
};

namespace {

bool find_buffer(vro_world *p, const std::string &s, void **ptr, int64_t *count, int32_t *dtype) {
    if (p->kind == VR_BOXES) return vro::box_buffer(p->box, s, ptr, count, dtype);
    if (p->w.F.count(s)) {
        *ptr = p->w.F[s].data();
        *count = (int64_t)p->w.F[s].size();
        *dtype = VR_F32;
        return true;
    }
    if (p->w.I.count(s)) {
        *ptr = p->w.I[s].data();
        *count = (int64_t)p->w.I[s].size();
        *dtype = VR_I32;
        return true;
    }
    return false;
}

}  // namespace

extern "C" {

int vro_world_create(int32_t kind, const vr_step_desc *desc, vro_world **out) {
    if ((kind != VR_SPHERES && kind != VR_BOXES) || !desc || !out) return fail(VR_ERR_ARGUMENT, "unknown world kind");
    auto *p = new vro_world();
    p->kind = kind;
    int rc = guard([&] {
        if (kind == VR_BOXES) {
            try {
                p->box = vro::box_create(*desc);
            } catch (const std::invalid_argument &e) {
                throw ArgError(e.what());
            }
        } else {
            p->w.build(*desc);
        }
    });
    if (rc != VR_OK) {
        delete p;
        return rc;
    }
    *out = p;
    return VR_OK;
}

int vro_batch_create(const vr_step_desc *desc, int32_t scenes, const int32_t *bodies, const float *extents,
                     const int32_t *grids, const float *cells, vro_world **out) {
    if (!desc || !out || !bodies || !extents || !grids || !cells) return fail(VR_ERR_ARGUMENT, "null argument");
    if (scenes < 1) return fail(VR_ERR_ARGUMENT, "a batch needs a scene");
    auto *p = new vro_world();
    p->kind = VR_BOXES;
    int rc = guard([&] {
        int64_t total = 0;
        p->batchFirst.push_back(0);
        for (int k = 0; k < scenes; k++) {
            vr_step_desc one = *desc;
            one.bodies = bodies[k];
            one.has_grid = 1;
            one.nx = grids[3 * k];
            one.ny = grids[3 * k + 1];
            one.nz = grids[3 * k + 2];
            one.cell = cells[k];
            try {
                p->batch.push_back(vro::box_create(one));
            } catch (const std::invalid_argument &e) {
                throw ArgError(e.what());
            }
            total += bodies[k];
            p->batchFirst.push_back(total);
            for (int a = 0; a < 3; a++) p->batchBox.push_back(extents[3 * k + a]);
        }
        if (total != desc->bodies) throw ArgError("the scenes' bodies do not add up to the step's");
    });
    if (rc != VR_OK) {
        delete p;
        return rc;
    }
    *out = p;
    return VR_OK;
}

int vro_world_clear(vro_world *p) {
    return guard([&] {
        if (!p->batch.empty())
            for (auto *b : p->batch) vro::box_clear(b);
        else if (p->kind == VR_BOXES) vro::box_clear(p->box);
        else p->w.clear();
    });
}

int vro_world_buffer_info(vro_world *p, const char *name, int64_t *count, int32_t *dtype) {
    void *ptr;
    if (!p->batch.empty()) {  // a batch's per-body buffers, over all its bodies
        int64_t size;
        int64_t first = p->batchFirst[1] - p->batchFirst[0];
        if (vro::box_buffer(p->batch[0], name, &ptr, &size, dtype) && size == first) {
            *count = p->batchFirst.back();
            return VR_OK;
        }
        return fail(VR_ERR_NAME, std::string("a batch has no per-body buffer named ") + name);
    }
    if (find_buffer(p, name, &ptr, count, dtype)) return VR_OK;
    return fail(VR_ERR_NAME, std::string("no buffer named ") + name);
}

namespace {

// A batch's per-body buffer, split over (write) or gathered from (read) its scenes' worlds.
int batch_copy(vro_world *p, const char *name, void *host, int64_t count, bool write) {
    int64_t total = p->batchFirst.back();
    if (count < 0 || count > total) return fail(VR_ERR_NAME, std::string("buffer ") + name + " is shorter");
    for (size_t k = 0; k < p->batch.size(); k++) {
        void *ptr;
        int64_t size;
        int32_t dtype;
        if (!vro::box_buffer(p->batch[k], name, &ptr, &size, &dtype) ||
            size != p->batchFirst[k + 1] - p->batchFirst[k])
            return fail(VR_ERR_NAME, std::string("a batch has no per-body buffer named ") + name);
        int64_t from = p->batchFirst[k], to = std::min(p->batchFirst[k + 1], count);
        if (to <= from) continue;
        char *h = (char *)host + from * 4;
        if (write) std::memcpy(ptr, h, (size_t)(to - from) * 4);
        else std::memcpy(h, ptr, (size_t)(to - from) * 4);
    }
    return VR_OK;
}

}  // namespace

int vro_world_write(vro_world *p, const char *name, const void *src, int64_t count) {
    if (!p->batch.empty()) return batch_copy(p, name, const_cast<void *>(src), count, true);
    void *dst;
    int64_t size;
    int32_t dtype;
    if (!find_buffer(p, name, &dst, &size, &dtype)) return fail(VR_ERR_NAME, std::string("no buffer named ") + name);
    if (count < 0 || count > size) return fail(VR_ERR_NAME, std::string("buffer ") + name + " is shorter than the write");
    std::memcpy(dst, src, (size_t)count * 4);
    return VR_OK;
}

int vro_world_read(vro_world *p, const char *name, void *dst, int64_t count) {
    if (!p->batch.empty()) return batch_copy(p, name, dst, count, false);
    void *src;
    int64_t size;
    int32_t dtype;
    if (!find_buffer(p, name, &src, &size, &dtype)) return fail(VR_ERR_NAME, std::string("no buffer named ") + name);
    if (count < 0 || count > size) return fail(VR_ERR_NAME, std::string("buffer ") + name + " is shorter than the read");
    std::memcpy(dst, src, (size_t)count * 4);
    return VR_OK;
}

int vro_world_params(vro_world *p, const int32_t *words, int32_t n) {
    if (!p->batch.empty()) {
        if (n != 18) return fail(VR_ERR_ARGUMENT, "box parameters are 18 words");
        return guard([&] {
            for (size_t k = 0; k < p->batch.size(); k++) {
                std::vector<int32_t> own(words, words + n);
                for (int a = 0; a < 3; a++) std::memcpy(&own[4 + a], &p->batchBox[3 * k + a], 4);  // SX, SY, SZ
                try {
                    vro::box_params(p->batch[k], own.data(), n);
                } catch (const std::invalid_argument &e) {
                    throw ArgError(e.what());
                }
            }
        });
    }
    if (p->kind == VR_BOXES) {
        return guard([&] {
            try {
                vro::box_params(p->box, words, n);
            } catch (const std::invalid_argument &e) {
                throw ArgError(e.what());
            }
        });
    }
    if (n != PARAM_COUNT) return fail(VR_ERR_ARGUMENT, "sphere parameters are 13 words");
    std::memcpy(p->w.F["params"].data(), words, sizeof(int32_t) * PARAM_COUNT);
    float *P = p->w.F["params"].data();
    p->w.friction = P[MU_STATIC] + P[MU_KINETIC] + P[MU_ROLLING] > 0;  // Spheres.rubs, Spheres.java:232-235
    return VR_OK;
}

int vro_world_start(vro_world *p, const float *x, const float *y, const float *z, const float *u, const float *v,
                    const float *vw, const float *r, const float *im, const int32_t *params, int32_t nparams) {
    // SphereSimulation.start, SphereSimulation.java:154-178.
    if (p->kind == VR_BOXES || !p->batch.empty())
        return fail(VR_ERR_STATE, "a box world starts through clear, writes and params");
    int64_t n = 0;
    int32_t dtype = 0;
    int rc = vro_world_buffer_info(p, "x", &n, &dtype);
    if (rc == VR_OK) rc = vro_world_clear(p);
    const char *names[8] = {"x", "y", "z", "u", "v", "w", "r", "im"};
    const float *fields[8] = {x, y, z, u, v, vw, r, im};
    for (int k = 0; k < 8 && rc == VR_OK; k++) rc = vro_world_write(p, names[k], fields[k], n);
    if (rc == VR_OK) rc = vro_world_params(p, params, nparams);
    if (rc != VR_OK) return rc;
    std::vector<float> shown((size_t)SHOWN_STRIDE * (size_t)n, 0.f);
    for (int64_t s = 0; s < n; s++) {
        float *b = shown.data() + (size_t)SHOWN_STRIDE * (size_t)s;
        b[0] = b[4] = x[s];
        b[1] = b[5] = y[s];
        b[2] = b[6] = z[s];
        b[3] = r[s];
        b[TURNED_BEFORE + 3] = b[TURNED_AFTER + 3] = 1.f;  // the turns as the identity
    }
    return vro_world_write(p, "shown", shown.data(), (int64_t)shown.size());
}

int vro_world_friction(vro_world *p, int32_t on) {
    p->w.friction = on != 0;
    return VR_OK;
}

int vro_world_slice(vro_world *p, int32_t dispatches) {  // SphereSimulation.java:136-148
    if (p->kind == VR_BOXES || !p->batch.empty())
        return fail(VR_ERR_STATE, "a box world submits each list whole: it has no slice");
    if (dispatches < 0)
        return fail(VR_ERR_ARGUMENT,
                    "a slice is some dispatches, or zero for no bound; got " + std::to_string(dispatches));
    p->w.slice = dispatches;
    return VR_OK;
}

int vro_world_step(vro_world *p, int32_t wait, vr_step_report *rep) {
    (void)wait;
    if (!p->batch.empty()) {
        return guard([&] {
            vr_step_report local{};
            vr_step_report *r = rep ? rep : &local;
            std::memset(r, 0, sizeof(*r));
            // The scenes share nothing, so they step on as many threads as OpenMP gives (vro_threads); each scene's
            // result is the same whichever thread takes it.
            std::vector<vr_step_report> ones(p->batch.size());
            std::vector<std::string> errors(p->batch.size());
#pragma omp parallel for schedule(dynamic, 1)
            for (long k = 0; k < (long)p->batch.size(); k++) {
                try {
                    std::memset(&ones[k], 0, sizeof(ones[k]));
                    vro::box_step(p->batch[k], &ones[k]);
                } catch (const std::exception &e) {
                    errors[k] = e.what();
                }
            }
            for (const std::string &e : errors)
                if (!e.empty()) throw std::runtime_error(e);
            for (const vr_step_report &one : ones) {
                r->touching += one.touching;
                r->points += one.points;
                r->unplaced += one.unplaced;
                r->unlisted += one.unlisted;
                r->rounds = std::max(r->rounds, one.rounds);
                r->most_rounds = std::max(r->most_rounds, one.most_rounds);
                r->submissions = std::max(r->submissions, one.submissions);
            }
        });
    }
    if (p->kind == VR_BOXES) {
        return guard([&] {
            vr_step_report local{};
            vr_step_report *r = rep ? rep : &local;
            std::memset(r, 0, sizeof(*r));
            auto t0 = std::chrono::steady_clock::now();
            vro::box_step(p->box, r);
            r->wall_ns = std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - t0)
                             .count();
        });
    }
    return guard([&] { p->w.step(rep); });
}

int vro_world_finish(vro_world *) { return VR_OK; }

int vro_world_destroy(vro_world *p) {
    delete p;
    return VR_OK;
}

int vro_last_error(char *buf, int32_t len) {
    if (len <= 0) return VR_OK;
    std::strncpy(buf, g_error.c_str(), (size_t)len - 1);
    buf[len - 1] = 0;
    return VR_OK;
}

int vro_box_collide(vro_world *p, const int32_t *pairs, int32_t npairs, float *manifold) {
    if (p->kind != VR_BOXES || !p->box) return fail(VR_ERR_STATE, "not a box world of one scene");
    try {
        vro::box_collide(p->box, pairs, npairs, manifold);
    } catch (const std::invalid_argument &e) {
        return fail(VR_ERR_ARGUMENT, e.what());
    }
    return VR_OK;
}

int vro_box_collide_list(vro_world *p, vr_step_report *rep) {
    if (p->kind != VR_BOXES) return fail(VR_ERR_STATE, "not a box world");
    return guard([&] {
        vr_step_report local{};
        vr_step_report *r = rep ? rep : &local;
        std::memset(r, 0, sizeof(*r));
        if (p->batch.empty()) {
            vro::box_collide_list(p->box, r);
            return;
        }
        for (auto *b : p->batch) {
            vr_step_report one{};
            vro::box_collide_list(b, &one);
            r->touching += one.touching;
            r->points += one.points;
            r->unplaced += one.unplaced;
            r->unlisted += one.unlisted;
            r->rounds = std::max(r->rounds, one.rounds);
            r->most_rounds = std::max(r->most_rounds, one.most_rounds);
            r->submissions = std::max(r->submissions, one.submissions);
        }
    });
}

// The oracle keeps no pass timing; its profile reads as zeros.
int vro_world_profile(vro_world *, int32_t) { return VR_OK; }

int vro_world_profile_read(vro_world *, vr_profile *out) {
    if (out) std::memset(out, 0, sizeof(*out));
    return VR_OK;
}

// SphereDiagnostics.of (SphereDiagnostics.java:34-83) and BoxDiagnostics.of (BoxDiagnostics.java:18-44) in f64,
// every pair scanned for the overlap; the layout of vr_world_diagnostics.
int vro_world_diagnostics(vro_world *p, double *out, int32_t n) {
    if (!p || !out || n < 14 || !p->batch.empty()) return fail(VR_ERR_ARGUMENT, "diagnostics need 14 doubles");
    auto f = [&](const char *name) -> const float * {
        void *ptr;
        int64_t count;
        int32_t dtype;
        return find_buffer(p, name, &ptr, &count, &dtype) ? (const float *)ptr : nullptr;
    };
    for (int k = 0; k < 14; k++) out[k] = 0;
    const float *P = f("params");
    if (p->kind == VR_SPHERES) {
        int m = p->w.n;
        const float *x = f("x"), *y = f("y"), *z = f("z"), *u = f("u"), *v = f("v"), *w = f("w"), *ax = f("ax"),
                    *ay = f("ay"), *az = f("az"), *r = f("r"), *im = f("im");
        for (int s = 0; s < m; s++) {
            double X = x[s], Y = y[s], Z = z[s], U = u[s], V = v[s], W = w[s], A = ax[s], B = ay[s], C = az[s];
            double R = r[s], M = im[s] > 0 ? 1.0 / (double)im[s] : 0.0;
            if (!(std::isfinite(X) && std::isfinite(Y) && std::isfinite(Z) && std::isfinite(U) && std::isfinite(V) &&
                  std::isfinite(W) && std::isfinite(A) && std::isfinite(B) && std::isfinite(C)))
                out[12] = 1;
            double at[3] = {X, Y, Z}, ext[3] = {P[SX], P[SY], P[SZ]};
            for (int a = 0; a < 3; a++)
                out[1] = std::max(out[1], std::max((R - at[a]) / R, (at[a] + R - ext[a]) / R));
            double speed2 = U * U + V * V + W * W, spin2 = A * A + B * B + C * C, inertia = 0.4 * M * R * R;
            out[11] = std::max(out[11], speed2);
            out[2] += 0.5 * M * speed2 + 0.5 * inertia * spin2;
            out[3] += 0.5 * inertia * spin2;
            out[4] += M * -(double)P[GY] * Y;
            out[5] += M * U;
            out[6] += M * V;
            out[7] += M * W;
            out[8] += M * (Y * W - Z * V) + inertia * A;
            out[9] += M * (Z * U - X * W) + inertia * B;
            out[10] += M * (X * V - Y * U) + inertia * C;
            for (int o2 = s + 1; o2 < m; o2++) {
                double dx = X - x[o2], dy = Y - y[o2], dz = Z - z[o2];
                double over = (R + r[o2] - std::sqrt(dx * dx + dy * dy + dz * dz)) / std::min(R, (double)r[o2]);
                out[0] = std::max(out[0], over);
            }
        }
    } else {
        int64_t m;
        void *ptr;
        int32_t dtype;
        find_buffer(p, "x", &ptr, &m, &dtype);
        const float *X[3] = {f("x"), f("y"), f("z")}, *V[3] = {f("u"), f("v"), f("w")},
                    *L[3] = {f("lx"), f("ly"), f("lz")}, *Q[4] = {f("qx"), f("qy"), f("qz"), f("qw")},
                    *Hh[3] = {f("hx"), f("hy"), f("hz")}, *im = f("im"), *shape = f("shape"), *man = f("manifold");
        int64_t slots = 0;
        if (man) find_buffer(p, "manifold", &ptr, &slots, &dtype), slots /= 40;
        for (int64_t s = 0; s < m; s++) {
            double at[3], vel[3], l[3], q[4], h[3];
            bool ok = true;
            for (int a = 0; a < 3; a++) {
                at[a] = X[a][s], vel[a] = V[a][s], l[a] = L[a][s], h[a] = Hh[a][s];
                ok = ok && std::isfinite(at[a]) && std::isfinite(vel[a]) && std::isfinite(l[a]);
            }
            for (int k = 0; k < 4; k++) q[k] = Q[k][s];
            if (!ok) out[12] = 1;
            out[11] = std::max(out[11], vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2]);
            if (im[s] > 0) {
                double M = 1.0 / (double)im[s], moment[3];
                if (shape[s] > 0.5f) moment[0] = moment[1] = moment[2] = 0.4 * M * h[0] * h[0];
                else moment[0] = M * (h[1] * h[1] + h[2] * h[2]) / 3, moment[1] = M * (h[0] * h[0] + h[2] * h[2]) / 3,
                     moment[2] = M * (h[0] * h[0] + h[1] * h[1]) / 3;
                auto rot = [](const double *qq, const double *pp, double *oo) {
                    double t[3] = {2 * (qq[1] * pp[2] - qq[2] * pp[1]), 2 * (qq[2] * pp[0] - qq[0] * pp[2]),
                                   2 * (qq[0] * pp[1] - qq[1] * pp[0])};
                    double c[3] = {qq[1] * t[2] - qq[2] * t[1], qq[2] * t[0] - qq[0] * t[2], qq[0] * t[1] - qq[1] * t[0]};
                    for (int a = 0; a < 3; a++) oo[a] = pp[a] + qq[3] * t[a] + c[a];
                };
                double cq[4] = {-q[0], -q[1], -q[2], q[3]}, own[3], sc[3], omega[3];
                rot(cq, l, own);
                for (int a = 0; a < 3; a++) sc[a] = moment[a] > 0 ? own[a] / moment[a] : 0;
                rot(q, sc, omega);
                double orbit[3] = {at[1] * vel[2] - at[2] * vel[1], at[2] * vel[0] - at[0] * vel[2],
                                   at[0] * vel[1] - at[1] * vel[0]};
                double spin = 0;
                for (int a = 0; a < 3; a++) {
                    out[2] += 0.5 * M * vel[a] * vel[a];
                    spin += 0.5 * l[a] * omega[a];
                    out[5 + a] += M * vel[a];
                    out[8 + a] += M * orbit[a] + l[a];
                }
                out[2] += spin;
                out[3] += spin;
                out[4] += M * -(double)P[2] * at[1];
            }
            for (int k = 0; k < 32 && s * 32 + k < slots; k++) {
                const float *mm = man + (s * 32 + k) * 40;
                for (int pt = 0; pt < (int)mm[3] && pt < 4; pt++) out[13] = std::max(out[13], (double)mm[8 + pt * 8 + 6]);
            }
        }
    }
    out[11] = std::sqrt(out[11]);
    return VR_OK;
}

// The oracle holds no device scratch: its allocation-free diagnostics are its diagnostics.
int vro_world_diagnostics_static(vro_world *p, double *out, int32_t n) { return vro_world_diagnostics(p, out, n); }

// The learned warm start runs on the device only: the oracle's box worlds keep the reference's carry.
int vro_box_tolerance(vro_world *p, float tolerance, int32_t measure) {
    if (p->kind != VR_BOXES) return fail(VR_ERR_STATE, "not a box world");
    try {
        if (p->batch.empty()) vro::box_tolerance(p->box, tolerance, measure != 0);
        for (auto *b : p->batch) vro::box_tolerance(b, tolerance, measure != 0);
    } catch (const std::invalid_argument &e) {
        return fail(VR_ERR_ARGUMENT, e.what());
    }
    return VR_OK;
}

// The oracle steps a batch one way (scene after scene on its threads): nothing to tune, its plan the pass.
int vro_box_retune(vro_world *p) {
    if (p->kind != VR_BOXES) return fail(VR_ERR_STATE, "not a box world");
    return VR_OK;
}

int vro_box_plan(vro_world *p, int32_t *plan, int32_t *tuning) {
    if (!plan || !tuning) return fail(VR_ERR_ARGUMENT, "null argument");
    if (p->kind != VR_BOXES) return fail(VR_ERR_STATE, "not a box world");
    *plan = VR_PLAN_PASS;
    *tuning = 0;
    return VR_OK;
}

// A batch's report gathers its scenes' own: the most any scene launched, the mean of their sweeps, the largest of
// their residuals.
int vro_box_solve_read(vro_world *p, vr_box_solve *out) {
    if (!out) return fail(VR_ERR_ARGUMENT, "null argument");
    if (p->kind != VR_BOXES) return fail(VR_ERR_STATE, "not a box world");
    if (p->batch.empty()) {
        *out = vro::box_solve(p->box);
        return VR_OK;
    }
    vr_box_solve r = vro::box_solve(p->batch[0]);
    r.scenes = (int32_t)p->batch.size();
    for (int k = 0; k < 2; k++)
        for (int sub = 0; sub < VR_SOLVE_SUBSTEPS; sub++) {
            r.launched[k][sub] = 0;
            r.scene_sweeps[k][sub] = 0;
            r.residual[k][sub] = 0;
            r.residual_mean[k][sub] = 0;
        }
    for (auto *b : p->batch) {
        const vr_box_solve &one = vro::box_solve(b);
        for (int k = 0; k < 2; k++)
            for (int sub = 0; sub < VR_SOLVE_SUBSTEPS; sub++) {
                r.launched[k][sub] = std::max(r.launched[k][sub], one.launched[k][sub]);
                r.scene_sweeps[k][sub] += one.scene_sweeps[k][sub] / (float)p->batch.size();
                r.residual_mean[k][sub] += one.residual_mean[k][sub] / (float)p->batch.size();
                float v = one.residual[k][sub];
                if (r.residual[k][sub] == r.residual[k][sub] && !(v <= r.residual[k][sub])) r.residual[k][sub] = v;
            }
    }
    *out = r;
    return VR_OK;
}

int vro_box_solve_curve(vro_world *p, float *out, int32_t n, int32_t *count) {
    if (!count || n < 0 || (n > 0 && !out)) return fail(VR_ERR_ARGUMENT, "null argument");
    if (p->kind != VR_BOXES) return fail(VR_ERR_STATE, "not a box world");
    std::vector<float> c(vro::box_curve(p->batch.empty() ? p->box : p->batch[0]).size(), 0.f);
    std::vector<vro::BoxWorld *> worlds = p->batch;
    if (worlds.empty()) worlds.push_back(p->box);
    for (auto *b : worlds) {
        const std::vector<float> &one = vro::box_curve(b);
        for (size_t k = 0; k < c.size(); k++)
            if (c[k] == c[k] && !(one[k] <= c[k])) c[k] = one[k];
    }
    int32_t take = std::min<int32_t>(n, (int32_t)c.size());
    std::memcpy(out, c.data(), sizeof(float) * (size_t)take);
    *count = take;
    return VR_OK;
}

int vro_box_log(vro_world *, int32_t, int64_t) { return fail(VR_ERR_STATE, "the oracle keeps no sample log"); }
int vro_box_samples(vro_world *, float *, float *, int64_t, int64_t *) {
    return fail(VR_ERR_STATE, "the oracle keeps no sample log");
}
int vro_box_learned(vro_world *, struct vro_mlp *) { return fail(VR_ERR_STATE, "the oracle has no learned start"); }

// Opening batch sizes the stepper holds, for the tests of its adaptation (SphereStepper.opening, :115-117).
int vro_world_opening(vro_world *p, int32_t iteration) { return p->w.opening.at(iteration); }

// Threads a batch's scenes step on (0: OpenMP's own choice); the oracle's CPU baseline runs with 1 and with all.
int vro_threads(int32_t n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

}  // extern "C"
