// sphere_kernels.hip — the sphere step's kernels for gfx950.
//
// One kernel per pass of the reference's sphere step (vexelray-sim-rigid-core/.../sphere/Spheres.java), over
// structure-of-arrays state. f32 arithmetic follows the reference's expression trees operation by operation and is
// compiled with -ffp-contract=off, so no multiply-add is fused; division and square root are IEEE-rounded (hipcc's
// default). Nothing here depends on the order in which threads run: contacts solved in one round share no sphere,
// claims are integer atomic maxima, and every sum over neighbours runs in one thread in a fixed order.
#include "sphere_kernels.h"

#include <algorithm>
#include <stdexcept>
#include <string>

namespace vr {

namespace {

constexpr int THREADS = 256;
static_assert(THREADS == SEG, "a wide round's workgroup takes one open-list segment, a place per thread");
constexpr float COINCIDENT = 1e-12f;   // Spheres.java:156
constexpr float LIST_MARGIN = 0.02f;   // Spheres.java:795
constexpr float INERTIA = 0.4f;        // Spheres.java:246
constexpr float MARGIN1 = 1.0f + LIST_MARGIN;  // f(1 + LIST_MARGIN)

inline int blocks_for(int64_t n) { return (int)std::max<int64_t>(1, (n + THREADS - 1) / THREADS); }

inline void check_launch(const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

__device__ __forceinline__ float fmax_(float a, float b) { return a < b ? b : a; }  // std::max
__device__ __forceinline__ float fmin_(float a, float b) { return b < a ? b : a; }  // std::min

__device__ __forceinline__ float dot3(const float *x, const float *y) {
    return x[0] * y[0] + (x[1] * y[1] + x[2] * y[2]);
}

__device__ __forceinline__ void cross3(const float *x, const float *y, float *z) {
    z[0] = x[1] * y[2] - x[2] * y[1];
    z[1] = x[2] * y[0] - x[0] * y[2];
    z[2] = x[0] * y[1] - x[1] * y[0];
}

// A sphere's position, moves and turns, loaded once and stored once (the reference's Held, Spheres.java:532-569),
// with what a pass only reads of it: radius, spin and velocity (a pass changes none of them).
struct Held {
    int s;
    float at[3], displaced[3], turned[3];
    float r, spin[3], vel[3];
};

__device__ __forceinline__ Held hold(const SphereDev &D, int s) {
    Held h;
    h.s = s;
    h.at[0] = D.x[s]; h.at[1] = D.y[s]; h.at[2] = D.z[s];
    h.displaced[0] = D.dx[s]; h.displaced[1] = D.dy[s]; h.displaced[2] = D.dz[s];
    h.turned[0] = D.tx[s]; h.turned[1] = D.ty[s]; h.turned[2] = D.tz[s];
    h.r = D.r[s];
    h.spin[0] = D.ax[s]; h.spin[1] = D.ay[s]; h.spin[2] = D.az[s];
    h.vel[0] = D.u[s]; h.vel[1] = D.v[s]; h.vel[2] = D.w[s];
    return h;
}

__device__ __forceinline__ void store(const SphereDev &D, const Held &h, bool turns) {
    D.x[h.s] = h.at[0]; D.y[h.s] = h.at[1]; D.z[h.s] = h.at[2];
    D.dx[h.s] = h.displaced[0]; D.dy[h.s] = h.displaced[1]; D.dz[h.s] = h.displaced[2];
    if (turns) { D.tx[h.s] = h.turned[0]; D.ty[h.s] = h.turned[1]; D.tz[h.s] = h.turned[2]; }
}

// The 13 parameter words (Spheres.java:143-155), loaded once a thread at the start of a kernel.
struct Params {
    float w[PARAM_COUNT];
    __device__ __forceinline__ float operator[](int i) const { return w[i]; }
};

__device__ __forceinline__ Params load_params(const float *p) {
    Params P;
    for (int k = 0; k < PARAM_COUNT; k++) P.w[k] = p[k];
    return P;
}

struct Impulse {  // Spheres.java:490-522
    float normal, p[3], q[3];
};

__device__ __forceinline__ Impulse load_impulse(const float *b) {
    Impulse t;
    t.normal = b[0];
    for (int k = 0; k < 3; k++) { t.p[k] = b[1 + k]; t.q[k] = b[4 + k]; }
    return t;
}

__device__ __forceinline__ void store_impulse(float *b, const Impulse &t) {
    b[0] = t.normal;
    for (int k = 0; k < 3; k++) { b[1 + k] = t.p[k]; b[4 + k] = t.q[k]; }
}

__device__ __forceinline__ void clear_impulse(Impulse &t) {
    t.normal = 0;
    for (int k = 0; k < 3; k++) { t.p[k] = 0; t.q[k] = 0; }
}

__device__ __forceinline__ float turns_per_impulse(float w, float lever, float radius) {  // Spheres.java:743-745
    return (w * lever) / (INERTIA * (radius * radius));
}

// Spheres.rub, Spheres.java:662-677.
__device__ __forceinline__ void rub(Held &a, float wa, float turnsA, Held *o, float wo, float turnsO,
                                    const float n[3], const float p[3]) {
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
__device__ __forceinline__ void contact_travel(const SphereDev &D, const Held &s, float h, const float n[3],
                                               float lever, int side, float travel[3]) {
    float turn[3], turnN[3];
    for (int k = 0; k < 3; k++) turn[k] = h * s.spin[k] + s.turned[k];
    cross3(turn, n, turnN);
    for (int k = 0; k < 3; k++) {
        float centre = h * s.vel[k] + s.displaced[k];
        float spun = lever * turnN[k];
        travel[k] = side > 0 ? centre + spun : centre - spun;
    }
}

// Spheres.friction, Spheres.java:595-654; o == nullptr for a wall.
__device__ void friction_at(const SphereDev &D, const Params &P, Held &a, float wa, float leverA, Held *o, float wo,
                            float leverO, const float n[3], Impulse &taken) {
    float muS = P[MU_STATIC], muK = P[MU_KINETIC];
    if (!(muS + muK > 0)) return;
    float h = P[H];
    float la = leverA;
    float slip[3];
    contact_travel(D, a, h, n, la, -1, slip);
    float ia = wa;
    float leverOverIa = turns_per_impulse(ia, la, a.r);
    float weight = ia + leverOverIa * la;
    float io = 0, lo = 0, leverOverIo = 0;
    if (o) {
        io = wo;
        lo = leverO;
        float other[3];
        contact_travel(D, *o, h, n, lo, 1, other);
        for (int k = 0; k < 3; k++) slip[k] = slip[k] - other[k];
        leverOverIo = turns_per_impulse(io, lo, o->r);
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
        float cut = fmin_(1.0f, (muK * lambdaN) / sqrtf(size2));
        for (int k = 0; k < 3; k++) whole[k] = whole[k] * cut;
    }
    float p[3];
    for (int k = 0; k < 3; k++) {
        p[k] = whole[k] - before[k];
        taken.p[k] = whole[k];
    }
    rub(a, ia, leverOverIa, o, io, leverOverIo, n, p);
}

// Spheres.roll, Spheres.java:699-741; o == nullptr for a wall.
__device__ void roll(const SphereDev &D, const Params &P, Held &a, float wa, Held *o, float wo, float reach,
                     Impulse &taken) {
    float muR = P[MU_ROLLING], h = P[H];
    float ra = a.r;
    float ia = wa / (INERTIA * (ra * ra));
    float inverse = ia;
    float io = 0;
    float against[3];
    for (int k = 0; k < 3; k++) against[k] = h * a.spin[k] + a.turned[k];
    if (o) {
        float ro = o->r;
        io = wo / (INERTIA * (ro * ro));
        inverse = inverse + io;
        for (int k = 0; k < 3; k++) against[k] = against[k] - (h * o->spin[k] + o->turned[k]);
    }
    if (!(inverse > 0)) return;
    float whole[3];
    for (int k = 0; k < 3; k++) whole[k] = taken.q[k] - against[k] / inverse;
    float size2 = dot3(whole, whole);
    float holds = (muR * reach) * taken.normal;
    if (size2 > holds * holds) {
        float cut = fmin_(1.0f, holds / sqrtf(size2));
        for (int k = 0; k < 3; k++) whole[k] = whole[k] * cut;
    }
    for (int k = 0; k < 3; k++) {
        float change = whole[k] - taken.q[k];
        taken.q[k] = whole[k];
        a.turned[k] = a.turned[k] + ia * change;
        if (o) o->turned[k] = o->turned[k] - io * change;
    }
}

__device__ __forceinline__ float rubbing(const Params &P) {  // Spheres.java:684-687
    return P[MU_STATIC] + (P[MU_KINETIC] + P[MU_ROLLING]);
}

// Spheres.warm, Spheres.java:448-481.
__device__ void warm(const SphereDev &D, Held &a, float wa, Held &o, float wo, const float d[3], float d2,
                     float reach, Impulse &taken) {
    float dist = sqrtf(d2);
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
    float ra = a.r, ro = o.r;
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

// What a contact's solve reads: both spheres and the contact's impulse.
struct PairLoad {
    Held a, o;
    float wa, wo;
    Impulse taken;
};

__device__ __forceinline__ PairLoad load_pair(const SphereDev &D, int c, int a, int o) {
    PairLoad L;
    L.a = hold(D, a);
    L.o = hold(D, o);
    L.wa = D.im[a];
    L.wo = D.im[o];
    L.taken = load_impulse(D.cur.impulse + (size_t)c * IMPULSE_WORDS);
    return L;
}

// Makes the compiler have `v` in a register here, so a load of it is issued no later than this point.
__device__ __forceinline__ void issued(float v) { asm volatile("" ::"v"(v)); }
__device__ __forceinline__ void issued(int v) { asm volatile("" ::"v"(v)); }

__device__ __forceinline__ void issued(const Held &h, bool rubs) {
    for (int k = 0; k < 3; k++) {
        issued(h.at[k]);
        issued(h.displaced[k]);
        issued(h.turned[k]);
        if (rubs) {
            issued(h.spin[k]);
            issued(h.vel[k]);
        }
    }
    issued(h.r);
}

// Issues every load of a contact's solve at once (for a rubbing pass, spins and velocities too).
template <bool RUBS>
__device__ __forceinline__ void issued(const PairLoad &L) {
    issued(L.a, RUBS);
    issued(L.o, RUBS);
    issued(L.wa);
    issued(L.wo);
    issued(L.taken.normal);
    for (int k = 0; k < 3; k++) {
        issued(L.taken.p[k]);
        issued(L.taken.q[k]);
    }
}

// Spheres.contact, Spheres.java:369-439: contact c of spheres a < o, parting (RUBS false) or rubbing, from its loads.
template <bool RUBS>
__device__ void contact(const SphereDev &D, const Params &P, int c, PairLoad &L, int pass) {
    Held &ha = L.a, &ho = L.o;
    float d[3];
    for (int k = 0; k < 3; k++) d[k] = ha.at[k] - ho.at[k];
    float d2 = dot3(d, d);
    float ra = ha.r, ro = ho.r;
    float reach = ra + ro;
    float wa = L.wa, wo = L.wo;
    float wsum = wa + wo;
    float listed = reach * MARGIN1;
    if (!(d2 < listed * listed) || !(d2 > COINCIDENT) || !(wsum > 0)) return;
    Impulse taken = L.taken;
    if (!RUBS) {
        if (pass == 1) {
            if (rubbing(P) > 0) warm(D, ha, wa, ho, wo, d, d2, reach, taken);
            else clear_impulse(taken);
        }
        for (int k = 0; k < 3; k++) d[k] = ha.at[k] - ho.at[k];
        float dist = sqrtf(fmax_(dot3(d, d), COINCIDENT));
        float depth = fmax_(-(taken.normal * wsum), P[OMEGA] * (reach - dist));
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
        float dist = sqrtf(d2);
        float nn[3];
        for (int k = 0; k < 3; k++) nn[k] = d[k] / dist;
        float share = dist / reach;
        float pairRadius = (ra * ro) / reach;
        roll(D, P, ha, wa, &ho, wo, pairRadius, taken);
        friction_at(D, P, ha, wa, share * ra, &ho, wo, share * ro, nn, taken);
    }
    store_impulse(D.cur.impulse + (size_t)c * IMPULSE_WORDS, taken);
    store(D, ha, true);
    store(D, ho, true);
}

__device__ __forceinline__ void contact(const SphereDev &D, const Params &P, int c, int a, int o, int pass, int rubs) {
    PairLoad L = load_pair(D, c, a, o);
    if (rubs) contact<true>(D, P, c, L, pass);
    else contact<false>(D, P, c, L, pass);
}

// Spheres.priority, Spheres.java:1166-1168 (31 bits); the same formula in 63 bits when wide.
__device__ __forceinline__ long long priority(const SphereDev &D, int a, int o, int g) {
    if (!D.wide) {
        uint32_t pair = (uint32_t)a * (uint32_t)D.n + (uint32_t)o;
        uint32_t p = pair * 0x9E3779B1u + (uint32_t)g * 0x632BE5ABu;
        return (long long)(p & 0x7FFFFFFFu);
    }
    uint64_t pair = (uint64_t)a * (uint64_t)D.n + (uint64_t)o;
    uint64_t p = pair * 0x9E3779B97F4A7C15ull + (uint64_t)(int64_t)g * 0x632BE59BD9B4E019ull;
    return (long long)(p & 0x7FFFFFFFFFFFFFFFull);
}

// Spheres.turn, Spheres.java:1468-1493.
__device__ __forceinline__ void turn(const SphereDev &D, int s, const float angle[3]) {
    float a[3], was[4], now[4];
    for (int k = 0; k < 3; k++) a[k] = 0.5f * angle[k];
    was[0] = D.qx[s]; was[1] = D.qy[s]; was[2] = D.qz[s]; was[3] = D.qw[s];
    now[0] = was[0] + (a[0] * was[3] + (a[1] * was[2] - a[2] * was[1]));
    now[1] = was[1] + (a[1] * was[3] + (a[2] * was[0] - a[0] * was[2]));
    now[2] = was[2] + (a[2] * was[3] + (a[0] * was[1] - a[1] * was[0]));
    now[3] = was[3] - (a[0] * was[0] + (a[1] * was[1] + a[2] * was[2]));
    float len2 = (now[0] * now[0] + now[1] * now[1]) + (now[2] * now[2] + now[3] * now[3]);
    float inverse = 1.0f / sqrtf(fmax_(len2, 1e-30f));
    D.qx[s] = now[0] * inverse; D.qy[s] = now[1] * inverse; D.qz[s] = now[2] * inverse; D.qw[s] = now[3] * inverse;
}

// Spheres.around, Spheres.java:1207-1229: the 27 cells about `key`, each cell's spheres in sorted order. The walk
// takes x fastest, then y, then z; the (up to) three cells of one y, z row are consecutive cells, so their spheres are
// one run of the sorted order, from the first cell's start to the end of the last. So the walk is 9 runs, whose 18
// bounds are loaded together (one memory latency, not one or two a cell), visited in the walk's order.
struct Runs {
    int lo[9], hi[9];
};

__device__ __forceinline__ Runs runs_around(const SphereDev &D, int key) {
    int ix = key % D.nx, iy = (key / D.nx) % D.ny, iz = key / (D.nx * D.ny);
    int x0 = max(ix - 1, 0), x1 = min(ix + 1, D.nx - 1);
    Runs R;
#pragma unroll
    for (int r = 0; r < 9; r++) {
        int jy = iy + (r % 3 - 1), jz = iz + (r / 3 - 1);
        bool inside = jy >= 0 && jy < D.ny && jz >= 0 && jz < D.nz;
        int row = D.nx * (jy + D.ny * jz);
        R.lo[r] = inside ? D.starts[row + x0] : 0;
        R.hi[r] = inside ? D.starts[row + x1 + 1] : 0;
    }
    return R;
}

template <class Visit>
__device__ __forceinline__ void around(const SphereDev &D, int key, Visit visit) {
    const Runs R = runs_around(D, key);
#pragma unroll
    for (int r = 0; r < 9; r++)
        for (int k = R.lo[r]; k < R.hi[r]; k++) visit(D.order[k]);
}

// ---------------------------------------------------------------------------------------------------------------
// Per-sphere passes.

// Spheres.predict, Spheres.java:269-295.
__global__ void k_predict(SphereDev D) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= D.n) return;
    if (!(D.im[s] > 0)) return;
    const Params P = load_params(D.params);
    float h = P[H];
    float v0 = D.u[s] + h * P[GX];
    float v1 = D.v[s] + h * P[GY];
    float v2 = D.w[s] + h * P[GZ];
    D.u[s] = v0; D.v[s] = v1; D.w[s] = v2;
    D.x[s] = D.x[s] + h * v0;
    D.y[s] = D.y[s] + h * v1;
    D.z[s] = D.z[s] + h * v2;
    float angle[3] = {h * D.ax[s], h * D.ay[s], h * D.az[s]};
    turn(D, s, angle);
}

// Spheres.bin, Spheres.java:1308-1323: the cell key of each centre, clamped into the grid.
__global__ void k_bin(SphereDev D) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= D.n) return;
    float at[3] = {D.x[s], D.y[s], D.z[s]};
    int cells[3] = {D.nx, D.ny, D.nz};
    int c[3];
    for (int a = 0; a < 3; a++) {
        float fl = floorf(at[a] / D.cell);
        c[a] = (int)fmax_(0.f, fmin_((float)(cells[a] - 1), fl));
    }
    D.keys[s] = c[0] + D.nx * (c[1] + D.ny * c[2]);
}

// Spheres.warmWalls, Spheres.java:890-939.
__global__ void k_warm_walls(SphereDev D) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= D.n) return;
    if (!(D.im[s] > 0)) return;
    const Params P = load_params(D.params);
    float ext[3] = {P[SX], P[SY], P[SZ]};
    float rb = rubbing(P);
    float rs = D.r[s], w = D.im[s];
    float turns = turns_per_impulse(w, rs, rs);
    Held held = hold(D, s);
    for (int axis = 0; axis < 3; axis++) {
        float *base = D.wallImpulse + (size_t)s * WALL_IMPULSE_WORDS + axis * IMPULSE_WORDS;
        Impulse taken = load_impulse(base);
        float reach = rs * MARGIN1;
        float position = held.at[axis];
        float out = 0;
        if (position < reach) out = 1;
        if (position > ext[axis] - reach) out = -1;
        bool near = rb > 0 && out != 0;
        if (!near) clear_impulse(taken);
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
        store_impulse(base, taken);
    }
    store(D, held, true);
}

// Spheres.clearContacts, Spheres.java:829-844: claims cleared; the list's length kept as the longest, then zeroed.
__global__ void k_clear_contacts(SphereDev D) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= D.n) return;
    D.claims[0][s] = 0;
    D.claims[1][s] = 0;
    D.claims[2][s] = 0;
    if (s == 0) {
        D.count[LONGEST] = max(D.count[LONGEST], D.count[LISTED]);
        D.count[LISTED] = 0;
    }
}

// Spheres.listContacts, Spheres.java:961-1022, in two sweeps: first each sphere counts the higher-index spheres
// within the list margin (k_list_count); the counts are scanned into rows; then each sphere writes its row in the
// order the walk meets them (k_list_write), with each pair's warm start found in the last substep's row of its
// lower sphere, walking back from the row's end as the reference walks its chain from the head.
__device__ __forceinline__ bool near_pair(const SphereDev &D, const float mine[3], float ra, int o) {
    float dd[3] = {mine[0] - D.x[o], mine[1] - D.y[o], mine[2] - D.z[o]};
    float d2 = dd[0] * dd[0] + (dd[1] * dd[1] + dd[2] * dd[2]);
    float near = (ra + D.r[o]) * MARGIN1;
    return d2 < near * near;
}

// The spheres of the walk about `key` of higher index than a and within the list margin of it, handed to `take` in
// the walk's order. The candidates of a run are loaded four at a time, their tests made, and only then handed on,
// so `take`'s stores do not hold back the next candidates' loads.
template <class Take>
__device__ __forceinline__ void near_around(const SphereDev &D, int key, int a, const float mine[3], float ra,
                                            Take take) {
    const Runs R = runs_around(D, key);
#pragma unroll
    for (int r = 0; r < 9; r++) {
        for (int k = R.lo[r]; k < R.hi[r]; k += 4) {
            int o[4];
            bool near[4];
#pragma unroll
            for (int q = 0; q < 4; q++) o[q] = k + q < R.hi[r] ? D.order[k + q] : -1;
#pragma unroll
            for (int q = 0; q < 4; q++) near[q] = o[q] > a && near_pair(D, mine, ra, o[q]);
#pragma unroll
            for (int q = 0; q < 4; q++)
                if (near[q]) take(o[q]);
        }
    }
}

__global__ void k_list_count(SphereDev D) {
    // Threads take spheres in cell order, so a wave walks neighbouring cells together.
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= D.n) return;
    int a = D.order[t];
    float mine[3] = {D.x[a], D.y[a], D.z[a]};
    float ra = D.r[a];
    int found = 0;
    near_around(D, D.keys[a], a, mine, ra, [&](int) { found++; });
    D.cur.cnt[a] = found;
}

__global__ void k_list_write(SphereDev D) {
    // Threads take spheres in cell order, so a wave walks neighbouring cells together.
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= D.n) return;
    int a = D.order[t];
    float mine[3] = {D.x[a], D.y[a], D.z[a]};
    float ra = D.r[a];
    int slot = D.cur.off[a];
    int cap = D.cap;
    int kStart = D.kept.off[a];
    int kEnd = min(kStart + D.kept.cnt[a], cap);
    int kLow = max(kStart, kEnd - MOST_KEPT);
    // The last KEPT_HELD partners of a's kept row, from its end back (-1 past its start), loaded once for the whole
    // walk: the search for each pair's old place looks through them first, then through the rest of the row in
    // memory, in the same order, so it finds the same place.
    constexpr int KEPT_HELD = 8;
    int held[KEPT_HELD];
#pragma unroll
    for (int q = 0; q < KEPT_HELD; q++) held[q] = kEnd - 1 - q >= kLow ? D.kept.b[kEnd - 1 - q] : -1;
    near_around(D, D.keys[a], a, mine, ra, [&](int o) {
        int here = slot++;
        if (here >= cap) return;
        D.cur.a[here] = a;
        D.cur.b[here] = o;
        D.cur.done[here] = 0;
        int was = -1;
#pragma unroll
        for (int q = 0; q < KEPT_HELD; q++)
            if (was < 0 && held[q] == o) was = kEnd - 1 - q;
        for (int j = kEnd - 1 - KEPT_HELD; was < 0 && j >= kLow; j--)
            if (D.kept.b[j] == o) was = j;
        float warmth[IMPULSE_WORDS] = {0, 0, 0, 0, 0, 0, 0};
        if (was >= 0) {
            const float *src = D.kept.impulse + (size_t)was * IMPULSE_WORDS;
            for (int k = 0; k < IMPULSE_WORDS; k++) warmth[k] = src[k];
        }
        float *dst = D.cur.impulse + (size_t)here * IMPULSE_WORDS;
        for (int k = 0; k < IMPULSE_WORDS; k++) dst[k] = warmth[k];
    });
}

// Spheres.contactRound, Spheres.java:1076-1133: one claim round over the list, one thread a contact.
__global__ void k_round(SphereDev D, RoundConst rc) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c == 0) D.open[1 - rc.mark] = 0;
    int listed = min(D.count[LISTED], D.cap);
    bool stillOpen = false;
    if (c < listed && D.cur.done[c] != rc.pass) {
        int a = D.cur.a[c], o = D.cur.b[c];
        const long long *checked = D.claims[(rc.slot + 2) % 3];
        long long *claimed = D.claims[rc.slot % 3];
        long long *cleared = D.claims[(rc.slot + 1) % 3];
        bool won = false;
        if (rc.checks > 0) {
            long long last = priority(D, a, o, rc.checksDispatch);
            won = checked[a] == last && checked[o] == last;
        }
        if (won) {
            contact(D, load_params(D.params), c, a, o, rc.pass, rc.rubs);
            D.cur.done[c] = rc.pass;
        } else {
            cleared[a] = 0;
            cleared[o] = 0;
            if (rc.claims > 0) {
                stillOpen = true;
                long long mine = priority(D, a, o, rc.dispatch);
                atomicMax(&claimed[a], mine);
                atomicMax(&claimed[o], mine);
            } else {
                atomicAdd(&D.count[MISSED], 1);
            }
        }
    }
    // open[mark] = 1 from one lane of each wave that has an open contact, not from every contact.
    if (__ballot(stillOpen) && (threadIdx.x & 63) == 0) D.open[rc.mark] = 1;
}

// One contact of a round run until done (the body of Spheres.contactRound, Spheres.java:1088-1131, for a contact
// still open): it solves if it holds both its spheres' claims from the round before. Returns whether it is still
// open, and so claims (claim_wave). EAGER issues the solve's loads with the claim check, not after it (in a round, a
// contact's spheres are written only by the one contact that holds both their claims, so nothing changes what was
// loaded before it is used): one memory latency fewer for the contacts that solve, but every open contact loads its
// pair, and only one in six to ten solves in a round. The wide rounds, bound by how many loads the memory system
// takes, load after the check; the tail, a few contacts in one workgroup and bound by latency, loads with it.
// The three claim arrays a round reads and writes: the world's in device memory, or a copy of them in LDS.
// Picked by selects, not by an index into an array of pointers, which would put the array in scratch memory.
struct Claims {
    long long *c0, *c1, *c2;
    __device__ __forceinline__ long long *at(int q) const { return q == 0 ? c0 : (q == 1 ? c1 : c2); }
};

__device__ __forceinline__ Claims device_claims(const SphereDev &D) {
    return Claims{D.claims[0], D.claims[1], D.claims[2]};
}

template <bool RUBS, bool EAGER>
__device__ __forceinline__ bool round_contact(const SphereDev &D, const Claims &C, const Params &P,
                                              const RoundConst &rc, int c, int a, int o) {
    if (rc.checks == 0) return true;
    const long long *checked = C.at((rc.slot + 2) % 3);
    long long last = priority(D, a, o, rc.checksDispatch);
    long long ca = checked[a], co = checked[o];
    if (EAGER) {
        PairLoad L = load_pair(D, c, a, o);
        issued<RUBS>(L);
        for (int k = 0; k < PARAM_COUNT; k++) issued(P[k]);
        if (ca != last || co != last) return true;
        contact<RUBS>(D, P, c, L, rc.pass);
    } else {
        if (ca != last || co != last) return true;
        PairLoad L = load_pair(D, c, a, o);
        contact<RUBS>(D, P, c, L, rc.pass);
    }
    D.cur.done[c] = rc.pass;
    return false;
}

// DPP lane permutations (GFX9 encodings): shifts within each row of 16 lanes, the broadcast of a row's last lane to
// the rows after it, and shifts of the whole wave by one lane.
constexpr int DPP_ROW_SHR1 = 0x111, DPP_ROW_SHR2 = 0x112, DPP_ROW_SHR4 = 0x114, DPP_ROW_SHR8 = 0x118;
constexpr int DPP_ROW_BCAST15 = 0x142, DPP_ROW_BCAST31 = 0x143, DPP_WAVE_SHL1 = 0x130, DPP_WAVE_SHR1 = 0x138;

// One step of a max scan over 64-bit values: the value CTRL brings to this lane (the lane's own where CTRL brings
// none or the lane's row is not in ROWS), taken when `take`.
template <int CTRL, int ROWS>
__device__ __forceinline__ long long scan_step(long long most, bool take) {
    int lo = (int)(unsigned)(unsigned long long)most, hi = (int)(unsigned)((unsigned long long)most >> 32);
    int l = __builtin_amdgcn_update_dpp(lo, lo, CTRL, ROWS, 0xF, false);
    int h = __builtin_amdgcn_update_dpp(hi, hi, CTRL, ROWS, 0xF, false);
    long long t = (long long)(((unsigned long long)(unsigned)h << 32) | (unsigned long long)(unsigned)l);
    return take ? max(most, t) : most;
}

// The claims of a wave's still-open contacts (`claims`): each clears both its spheres' words in the third claim
// array and raises both its spheres' claims to its priority (Spheres.java:1121-1130). Contacts are listed by lower
// sphere a, so a wave's lanes hold runs of one a side by side; clearing a word twice is clearing it once, and atomicMax
// of a run's priorities one by one leaves the word where one atomicMax of the run's largest does. So the last lane of
// each run of claiming lanes with one a clears and claims a alone, with the run's largest priority (a segmented max
// scan over the wave), and every word ends the round as it would with one clear and one atomic per contact: fewer
// device atomics, the same claims. Sphere o is claimed once per contact. Called by every lane of the wave.
__device__ __forceinline__ void claim_wave(const SphereDev &D, const Claims &C, const RoundConst &rc, bool claims, int a,
                                           int o) {
    long long *claimed = C.at(rc.slot % 3);
    long long *cleared = C.at((rc.slot + 1) % 3);
    int lane = threadIdx.x & 63;
    long long mine = claims ? priority(D, a, o, rc.dispatch) : 0;
    int key = claims ? a : -1 - lane;  // a lane that does not claim is a run of its own
    // The neighbours' keys, by whole-wave DPP shifts (lane 0 and lane 63 keep their own and are decided by the lane
    // test). Both run in every lane, before any test.
    int before = __builtin_amdgcn_update_dpp(key, key, DPP_WAVE_SHR1, 0xF, 0xF, false);
    int after = __builtin_amdgcn_update_dpp(key, key, DPP_WAVE_SHL1, 0xF, 0xF, false);
    bool head = lane == 0 || before != key;
    bool tail = lane == 63 || after != key;
    unsigned long long heads = __ballot(head) & ((2ull << lane) - 1ull);  // lane 63: every lane
    int start = 63 - __clzll((long long)heads);
    // Segmented inclusive max scan: within each row of 16 lanes by shifts of 1, 2, 4 and 8, then across rows (the
    // last lane of row 0 to row 1 and of row 2 to row 3, then the last lane of row 1 to rows 2 and 3), each step
    // taking a lane's value only from a lane of its own run. A lane the step does not feed keeps its own value, and
    // the maximum with itself changes nothing. Every lane ends with the maximum over its run up to itself.
    long long most = mine;
    most = scan_step<DPP_ROW_SHR1, 0xF>(most, lane - 1 >= start);
    most = scan_step<DPP_ROW_SHR2, 0xF>(most, lane - 2 >= start);
    most = scan_step<DPP_ROW_SHR4, 0xF>(most, lane - 4 >= start);
    most = scan_step<DPP_ROW_SHR8, 0xF>(most, lane - 8 >= start);
    most = scan_step<DPP_ROW_BCAST15, 0xA>(most, (lane & 16) != 0 && (lane & ~15) - 1 >= start);
    most = scan_step<DPP_ROW_BCAST31, 0xC>(most, lane >= 32 && 31 >= start);
    if (!claims) return;
    cleared[o] = 0;
    atomicMax(&claimed[o], mine);
    if (tail) {
        cleared[a] = 0;
        atomicMax(&claimed[a], most);
    }
}

constexpr int CYCLE = 48, BATCH = 8, SHRINK_AFTER = 120;  // SphereStep.java:48-59, SphereStepper.java:49

__device__ __forceinline__ int position(int k) { return k == 0 ? 0 : 1 + (k - 1) % CYCLE; }

// The constants of the round at cycle position p of pass `it` (SphereStep.java:286-315, roundNames :320-327).
__device__ __forceinline__ RoundConst cycle_round(int it, int p, bool rub, int iterations) {
    RoundConst rc;
    rc.checks = p == 0 ? 0 : 1;
    rc.claims = 1;
    rc.pass = rub ? iterations + it + 1 : it + 1;
    rc.dispatch = it * CYCLE + p % CYCLE;
    rc.mark = p % 2;
    rc.checksDispatch = it * CYCLE + max(p - 1, 0) % CYCLE;
    rc.rubs = rub ? 1 : 0;
    rc.slot = p % 3;
    return rc;
}

// Listed contact c as an open-list entry: {c, a, o, 0}.
__device__ __forceinline__ int4 listed_entry(const SphereDev &D, int c) {
    return make_int4(c, D.cur.a[c], D.cur.b[c], 0);
}

// Appends `e` to `out` for every lane with `open`, one atomic a wave on `count` (in LDS or in global memory).
__device__ __forceinline__ void append_open(bool open, const int4 &e, int4 *out, int *count) {
    unsigned long long mask = __ballot(open);
    if (!mask) return;
    int lane = threadIdx.x & 63;
    int leader = __ffsll((long long)mask) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(count, __popcll(mask));
    base = __builtin_amdgcn_readlane(base, leader);
    if (open) out[base + __popcll(mask & ((1ull << lane) - 1ull))] = e;
}

__global__ void k_step_begin(PassCtl *ctl, PassSlot *slots, int words) {
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < words) reinterpret_cast<int *>(slots)[q] = 0;
    if (q == 0) {
        ctl->error = 0;
        ctl->waits = 0;
        ctl->submissions = 0;
        ctl->rounds = 0;
        ctl->most = 0;
        ctl->unlisted = 0;
        ctl->listedSum = 0;
        for (int k = 0; k < MAX_WIDE; k++) ctl->wideRan[k] = 0;
        for (int k = 0; k < MAX_WIDE; k++) ctl->localFrom[k] = 0;
    }
}

// An open-list entry {c, a, o, 0} as loaded: the fourth word is always 0, so it is not taken from memory (a store
// of a loaded word would wait for that load, and with it for every memory operation issued after it).
__device__ __forceinline__ int4 open_entry(const int4 *list, size_t place) {
    int4 v = list[place];
    return make_int4(v.x, v.y, v.z, 0);
}

// Round j of a pass over the whole chip, a workgroup per segment of SEG list places. In round 0, workgroup g takes
// the listed contacts g*SEG onwards; in each later round it takes the open contacts it kept in the round before
// (segIn[g] of them from in[g*SEG]). It compacts the contacts that stay open into its own segment of `out`, counting
// in LDS, so no round has a device-wide counter; the pass's last wide round (`last`) compacts into one list
// instead, for k_round_tail. Which order open contacts land in does not matter: a round's outcome is the same in any
// order. The segment's count, the thread's entry and, in a rubbing pass, the parting pass's rounds do not depend on
// one another, so all three are loaded before any of them is tested: one memory latency, not three, before the
// contact's own loads.
// FIRST: round 0 (j == 0), an instance of its own, so neither instance joins two ways of loading.
// counted: each workgroup adds what it left open to slot->openAfter[j], which only the local plan's tail reads (to
// find the first round that fits its workgroup). Every workgroup's add lands on the same word, which costs a few
// microseconds a round when hundreds of workgroups take part, so the wide plan leaves it out.
template <bool RUB, bool FIRST>
__global__ void __launch_bounds__(THREADS) k_round_wide(SphereDev D, const PassCtl *ctl, PassSlot *slot,
                                                        const int4 *in, const int32_t *segIn, int4 *out,
                                                        int32_t *segOut, int it, int iterations, int j, int last,
                                                        int counted) {
    __shared__ int sCount;
    int g = blockIdx.x, i = threadIdx.x;
    int kRub = RUB ? ctl->k : 0;
    int nIn;
    int4 e = make_int4(0, 0, 0, 0);
    if (FIRST) {
        nIn = min(SEG, min(D.count[LISTED], D.cap) - g * SEG);
        e = listed_entry(D, min(g * SEG + i, D.cap - 1));  // launched only when cap >= 1; used only below nIn
    } else {
        nIn = segIn[g];
        e = open_entry(in, (size_t)g * SEG + i);  // in the list's segments * SEG places; used only below nIn
    }
    issued(kRub);
    issued(nIn);
    issued(e.x);
    issued(e.y);
    issued(e.z);
    if (RUB && j >= kRub) return;
    if (nIn <= 0) {
        if (!last && threadIdx.x == 0) segOut[g] = 0;
        return;
    }
    if (threadIdx.x == 0) {
        slot->opened[j] = 1;
        sCount = 0;
    }
    __syncthreads();
    RoundConst rc = cycle_round(it, position(j), RUB, iterations);
    const Params P = load_params(D.params);
    bool open = false;
    const Claims C = device_claims(D);
    if (i < nIn) open = round_contact<RUB, false>(D, C, P, rc, e.x, e.y, e.z);
    claim_wave(D, C, rc, open, e.y, e.z);
    if (last) {
        append_open(open, e, out, &slot->tailCount);
    } else {
        append_open(open, e, out + (size_t)g * SEG, &sCount);
        __syncthreads();
        if (threadIdx.x == 0) {
            segOut[g] = sCount;
            if (counted && sCount > 0) atomicAdd(&slot->openAfter[j], sCount);
        }
    }
}

// The rest of a pass in one workgroup of T threads, rounds separated by barriers rather than launches. It starts at
// the first round the wide rounds found empty, or else at round `wide` on the list the last of them compacted. A
// parting pass runs until no contact is open, then settles k, the rounds the reference's stepper would have run: its
// opening batch, then batches of 8, up to the first report that finds nothing open (SphereStepper.java:72-96); the
// opening batch then adapts as there (:118-126). A rubbing pass runs the parting pass's k rounds, or until nothing is
// open.
// LOCAL: the whole pass (wide is 0) with the three claim arrays copied into LDS at the start and back at the end, so
// a round's claim checks, clears and atomics are LDS operations; for worlds of at most LOCAL_MOST spheres. A claim
// word holds the same value at every point as in device memory, so the rounds are the same.
constexpr int TAIL = 512;

template <bool RUB, bool LOCAL, int T>
__global__ void __launch_bounds__(T) k_round_tail(SphereDev D, PassCtl *ctl, const PassSlot *slot, OpenLists L,
                                                  int32_t *opening, int32_t *enough, int it, int iterations,
                                                  int wide, int limit, PassSubmit submit) {
    __shared__ int sCount;
    extern __shared__ long long sClaims[];  // LOCAL: 3 * D.n words
    int listed = min(D.count[LISTED], D.cap);
    int j = 0, n = listed;
    if (wide > 0) {
        j = wide;
        for (int r = 0; r < wide; r++) {
            if (!slot->opened[r]) {
                j = r;
                break;
            }
        }
        n = j < wide ? 0 : slot->tailCount;
    }
    // The first round this pass entered with at most LOCAL_THREADS contacts open: round 0 takes the listed, wide
    // round j + 1 what wide round j left (its last, the tail's list); later rounds are found below. -1: not yet.
    int fits = -1;
    if (listed <= LOCAL_THREADS) fits = 0;
    for (int r = 1; fits < 0 && r < wide; r++)
        if (slot->openAfter[r - 1] <= LOCAL_THREADS) fits = r;
    if (fits < 0 && wide > 0 && j == wide && n <= LOCAL_THREADS) fits = wide;
    int open0 = RUB ? 0 : opening[it];
    int kRub = RUB ? ctl->k : 0;
    const Params P = load_params(D.params);
    Claims C = device_claims(D);
    if (LOCAL) {
        C = Claims{sClaims, sClaims + D.n, sClaims + 2 * (size_t)D.n};
        for (int s = threadIdx.x; s < D.n; s += T) {
            C.c0[s] = D.claims[0][s];
            C.c1[s] = D.claims[1][s];
            C.c2[s] = D.claims[2][s];
        }
        __syncthreads();
    }
    int k = 0;
    bool failed = false;
    while (true) {
        if (!RUB) {
            if (n == 0) {
                k = j <= open0 ? open0 : open0 + BATCH * ((j - open0 + BATCH - 1) / BATCH);
                break;
            }
            if (j >= open0 && (j - open0) % BATCH == 0 && j > limit) {
                failed = true;
                k = j;
                break;
            }
        } else if (j >= kRub || n == 0) {
            break;
        }
        RoundConst rc = cycle_round(it, position(j), RUB, iterations);
        const int4 *in = L.buf[j & 1];
        int4 *out = L.buf[(j + 1) & 1];
        if (threadIdx.x == 0) sCount = 0;
        __syncthreads();
        for (int base = 0; base < n; base += T) {
            int i = base + threadIdx.x;
            bool open = false;
            int4 e = make_int4(0, 0, 0, 0);
            if (i < n) {
                e = j == 0 ? listed_entry(D, i) : open_entry(in, i);
                open = round_contact<RUB, true>(D, C, P, rc, e.x, e.y, e.z);
            }
            claim_wave(D, C, rc, open, e.y, e.z);
            append_open(open, e, out, &sCount);
        }
        __syncthreads();
        n = sCount;
        j++;
        if (fits < 0 && n <= LOCAL_THREADS) fits = j;
        __syncthreads();
    }
    if (LOCAL) {
        for (int s = threadIdx.x; s < D.n; s += T) {
            D.claims[0][s] = C.c0[s];
            D.claims[1][s] = C.c1[s];
            D.claims[2][s] = C.c2[s];
        }
    }
    // The rounds this pass needed, for the host to launch that many wide next step.
    int used = (RUB ? iterations : 0) + it;
    if (threadIdx.x == 0 && used < MAX_WIDE) ctl->wideRan[used] = max(ctl->wideRan[used], j);
    if (LOCAL && threadIdx.x == 0 && used < MAX_WIDE)  // the wide plan counts no openAfter and reads no localFrom
        ctl->localFrom[used] = max(ctl->localFrom[used], fits < 0 ? j : fits);
    if (threadIdx.x == 0) ctl->listedSum += listed;
    if (threadIdx.x != 0 || RUB) return;
    if (failed) {
        ctl->error = 1;
        ctl->errRounds = k;
        ctl->errListed = D.count[LISTED];
    }
    // The opening list (the pass before's rubbing rounds, its prefix, its first rounds and a report) and the batches
    // of `more` after it, each cut into pieces of at most submit.slice dispatches.
    int more = k > open0 ? (k - open0) / BATCH : 0;
    int opened = (submit.carry ? ctl->k : 0) + submit.prefix + open0 + 1;
    int cut = submit.slice;
    ctl->submissions += cut == 0 ? 1 + more : (opened + cut - 1) / cut + more * ((BATCH + 1 + cut - 1) / cut);
    ctl->k = k;
    ctl->waits += 1 + more;
    ctl->rounds += k;
    ctl->most = max(ctl->most, k);
    ctl->unlisted += max(0, D.count[LISTED] - D.cap);
    if (k > opening[it]) {  // SphereStepper.adapt
        opening[it] = min(k, CYCLE);
        enough[it] = 0;
    } else if (++enough[it] >= SHRINK_AFTER && opening[it] > BATCH) {
        opening[it] -= BATCH;
        enough[it] = 0;
    }
}

// Spheres.walls -> shift with friction, Spheres.java:1185-1198, 1351-1418.
__global__ void k_walls(SphereDev D) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= D.n) return;
    if (!(D.im[s] > 0)) return;
    const Params P = load_params(D.params);
    float ext[3] = {P[SX], P[SY], P[SZ]};
    float rs = D.r[s];
    Held held = hold(D, s);
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
        float w = D.im[s];
        float *base = D.wallImpulse + (size_t)s * WALL_IMPULSE_WORDS + a * IMPULSE_WORDS;
        Impulse taken = load_impulse(base);
        float inside = fmax_(rs, fmin_(high, moved));
        float push = fmax_(into, 0.f);
        if (!(into > 0) && taken.normal > 0) {
            push = fmax_(into, -(taken.normal * w));
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
            roll(D, P, held, w, nullptr, 0, rs, taken);
            friction_at(D, P, held, w, rs, nullptr, 0, 0, nrm, taken);
        }
        store_impulse(base, taken);
    }
    for (int a = 0; a < 3; a++) {
        float high = ext[a] - rs;
        float inside = fmax_(rs, fmin_(high, held.at[a]));
        held.displaced[a] = held.displaced[a] + (inside - held.at[a]);
        held.at[a] = inside;
    }
    store(D, held, true);
}

// Spheres.velocity, Spheres.java:1430-1460.
__global__ void k_velocity(SphereDev D) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= D.n) return;
    float perH = 1.0f / D.params[H];
    float *vel[3] = {D.u, D.v, D.w}, *dsp[3] = {D.dx, D.dy, D.dz}, *bef[3] = {D.cx, D.cy, D.cz};
    float *spin[3] = {D.ax, D.ay, D.az}, *tr[3] = {D.tx, D.ty, D.tz};
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
    turn(D, s, angle);
}

// Spheres.bounce / bounce(grid), Spheres.java:1523-1627.
template <bool GRID>
__global__ void k_bounce(SphereDev D) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= D.n) return;
    int s = GRID ? D.order[t] : t;
    const Params P = load_params(D.params);
    float e = P[RESTITUTION];
    float wa = D.im[s];
    if (!(e > 0 && wa > 0)) return;
    float h = P[H], gx = P[GX], gy = P[GY], gz = P[GZ];
    float resting = (2.0f * h) * sqrtf(gx * gx + (gy * gy + gz * gz));
    float ext[3] = {P[SX], P[SY], P[SZ]};
    float ra = D.r[s];
    float mine[3] = {D.x[s], D.y[s], D.z[s]};
    float v[3] = {D.u[s], D.v[s], D.w[s]};
    float was[3] = {D.cx[s], D.cy[s], D.cz[s]};
    float sum[3] = {0, 0, 0};
    auto against = [&](int o) {
        if (o == s) return;
        float dd[3] = {mine[0] - D.x[o], mine[1] - D.y[o], mine[2] - D.z[o]};
        float d2 = dd[0] * dd[0] + (dd[1] * dd[1] + dd[2] * dd[2]);
        float near = (ra + D.r[o]) * MARGIN1;
        float wsum = wa + D.im[o];
        if (!(d2 < near * near && d2 > COINCIDENT)) return;
        float dist = sqrtf(d2);
        float bo[3] = {D.cx[o], D.cy[o], D.cz[o]};
        float vo[3] = {D.u[o], D.v[o], D.w[o]};
        float closing = 0, parting = 0;
        for (int a = 0; a < 3; a++) {
            float nn = dd[a] / dist;
            closing = closing + nn * (was[a] - bo[a]);
            parting = parting + nn * (v[a] - vo[a]);
        }
        if (closing < -resting) {
            float k = ((-(e * closing)) - parting) * wa / (wsum * dist);
            for (int a = 0; a < 3; a++) sum[a] = sum[a] + k * dd[a];
        }
    };
    if (GRID) {
        around(D, D.keys[s], against);
    } else {
        for (int o = 0; o < D.n; o++) against(o);
    }
    for (int a = 0; a < 3; a++) {
        float reach = ra * MARGIN1;
        if (mine[a] < reach && was[a] < -resting) sum[a] = sum[a] + ((-(e * was[a])) - v[a]);
        if (mine[a] > ext[a] - reach && was[a] > resting) sum[a] = sum[a] + ((-(e * was[a])) - v[a]);
    }
    D.dx[s] = sum[0];
    D.dy[s] = sum[1];
    D.dz[s] = sum[2];
}

// Spheres.bounced, Spheres.java:1633-1644.
__global__ void k_bounced(SphereDev D) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= D.n) return;
    D.u[s] = D.u[s] + D.dx[s]; D.dx[s] = 0;
    D.v[s] = D.v[s] + D.dy[s]; D.dy[s] = 0;
    D.w[s] = D.w[s] + D.dz[s]; D.dz[s] = 0;
}

// Spheres.solve / solve(grid) with Solving, Spheres.java:302-347, 1245-1298: the Jacobi move into cx, cy, cz.
template <bool GRID>
__global__ void k_jacobi_solve(SphereDev D) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= D.n) return;
    int s = GRID ? D.order[t] : t;
    float mine[3] = {D.x[s], D.y[s], D.z[s]};
    float rs = D.r[s], ws = D.im[s];
    float sum[3] = {0, 0, 0}, contacts = 0;
    auto against = [&](int o) {
        if (o == s) return;
        float dd[3] = {mine[0] - D.x[o], mine[1] - D.y[o], mine[2] - D.z[o]};
        float d2 = dd[0] * dd[0] + (dd[1] * dd[1] + dd[2] * dd[2]);
        float reach = rs + D.r[o];
        float wsum = ws + D.im[o];
        if (!(d2 < reach * reach) || !(d2 > COINCIDENT) || !(wsum > 0)) return;
        float dist = sqrtf(d2);
        float share = ((reach - dist) * (ws / wsum)) / dist;
        for (int a = 0; a < 3; a++) sum[a] = sum[a] + share * dd[a];
        contacts = contacts + 1.0f;
    };
    if (GRID) {
        around(D, D.keys[s], against);
    } else {
        for (int o = 0; o < D.n; o++) against(o);
    }
    const Params P = load_params(D.params);
    float omega = P[OMEGA];
    float scale = omega;
    if (P[AVERAGED] > 0) scale = omega / fmax_(contacts, 1.0f);
    D.cx[s] = scale * sum[0];
    D.cy[s] = scale * sum[1];
    D.cz[s] = scale * sum[2];
}

// Spheres.apply -> shift without friction, Spheres.java:1326-1342, 1351-1368.
__global__ void k_jacobi_apply(SphereDev D) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= D.n) return;
    if (!(D.im[s] > 0)) return;
    const Params P = load_params(D.params);
    float ext[3] = {P[SX], P[SY], P[SZ]};
    float rs = D.r[s];
    float cor[3] = {D.cx[s], D.cy[s], D.cz[s]};
    Held held = hold(D, s);
    for (int a = 0; a < 3; a++) {
        float c = cor[a];
        float moved = held.at[a] + c;
        float high = ext[a] - rs;
        float inside = fmax_(rs, fmin_(high, moved));
        held.at[a] = inside;
        float push = inside - moved;
        held.displaced[a] = held.displaced[a] + (c + push);
    }
    store(D, held, false);
}

// Spheres.show / show(grid), Spheres.java:1697-1794.
template <bool GRID>
__global__ void k_show(SphereDev D) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= D.n) return;
    int s = GRID ? D.order[t] : t;
    const Params P = load_params(D.params);
    float ext[3] = {P[SX], P[SY], P[SZ]};
    float *b = D.shown + (size_t)s * SHOWN_STRIDE;
    float mine[3] = {D.x[s], D.y[s], D.z[s]};
    for (int a = 0; a < 3; a++) {
        b[a] = b[4 + a];
        b[4 + a] = mine[a];
    }
    float rs = D.r[s];
    b[3] = rs;
    for (int k = 0; k < 12; k++) b[8 + k] = b[20 + k];
    float q[4] = {D.qx[s], D.qy[s], D.qz[s], D.qw[s]};
    for (int k = 0; k < 4; k++) {
        b[32 + k] = b[36 + k];
        b[36 + k] = q[k];
    }
    float flat[6] = {0, 0, 0, 0, 0, 0}, lean[3] = {0, 0, 0};
    for (int a = 0; a < 3; a++) {
        float low = rs - mine[a];
        float high = (mine[a] + rs) - ext[a];
        if (low > 0) {
            float share = low / rs;
            flat[a] = flat[a] + share;
            lean[a] = lean[a] + -1.0f * share;
        }
        if (high > 0) {
            float share = high / rs;
            flat[a] = flat[a] + share;
            lean[a] = lean[a] + 1.0f * share;
        }
    }
    auto against = [&](int o) {
        if (o == s) return;
        float dd[3] = {D.x[o] - mine[0], D.y[o] - mine[1], D.z[o] - mine[2]};
        float d2 = dot3(dd, dd);
        float reach = rs + D.r[o];
        if (!(d2 < reach * reach && d2 > COINCIDENT)) return;
        float dist = sqrtf(d2);
        float share = (reach - dist) / reach;
        float g = share / d2;
        float along = share / dist;
        flat[0] = flat[0] + g * (dd[0] * dd[0]);
        flat[1] = flat[1] + g * (dd[1] * dd[1]);
        flat[2] = flat[2] + g * (dd[2] * dd[2]);
        flat[3] = flat[3] + g * (dd[0] * dd[1]);
        flat[4] = flat[4] + g * (dd[0] * dd[2]);
        flat[5] = flat[5] + g * (dd[1] * dd[2]);
        for (int a = 0; a < 3; a++) lean[a] = lean[a] + along * dd[a];
    };
    if (GRID) {
        around(D, D.keys[s], against);
    } else {
        for (int o = 0; o < D.n; o++) against(o);
    }
    float *after = b + 20;
    for (int k = 0; k < 6; k++) after[k] = flat[k];
    for (int a = 0; a < 3; a++) after[6 + a] = lean[a];
    after[9] = (flat[0] + flat[1]) + flat[2];
    after[10] = sqrtf(dot3(lean, lean));
}

__global__ void k_fill(float *p, float value, int64_t count) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = value;
}

__global__ void k_fill_i32(int32_t *p, int32_t value, int64_t count) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = value;
}

// ---------------------------------------------------------------------------------------------------------------
// Exclusive scan: blocks of SCAN_TILE elements scanned in LDS, their sums scanned recursively, then added back.

constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = THREADS * SCAN_ITEMS;

__device__ __forceinline__ int wave_inclusive_scan(int v) {
    int lane = threadIdx.x & 63;
    for (int off = 1; off < 64; off <<= 1) {
        int t = __shfl_up(v, off, 64);
        if (lane >= off) v += t;
    }
    return v;
}

// Block-wide exclusive scan of one value per thread; returns the exclusive prefix, and the block total in *total.
__device__ __forceinline__ int block_exclusive_scan(int v, int *total) {
    __shared__ int waveSums[THREADS / 64];
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int inc = wave_inclusive_scan(v);
    if (lane == 63) waveSums[wave] = inc;
    __syncthreads();
    int base = 0, all = 0;
    for (int w = 0; w < THREADS / 64; w++) {
        int t = waveSums[w];
        if (w < wave) base += t;
        all += t;
    }
    __syncthreads();
    *total = all;
    return base + inc - v;
}

__global__ void k_scan_tile(const int32_t *in, int32_t *out, int n, int32_t *blockSums) {
    int base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    int vals[SCAN_ITEMS];
    int sum = 0;
    for (int k = 0; k < SCAN_ITEMS; k++) {
        int i = base + k;
        vals[k] = i < n ? in[i] : 0;
        sum += vals[k];
    }
    int total;
    int run = block_exclusive_scan(sum, &total);
    for (int k = 0; k < SCAN_ITEMS; k++) {
        int i = base + k;
        if (i < n) out[i] = run;
        run += vals[k];
    }
    if (threadIdx.x == 0 && blockSums) blockSums[blockIdx.x] = total;
}

__global__ void k_scan_add(int32_t *out, int n, const int32_t *blockOffsets) {
    int i = blockIdx.x * SCAN_TILE + threadIdx.x;
    int add = blockOffsets[blockIdx.x];
    for (int k = 0; k < SCAN_ITEMS; k++, i += THREADS)
        if (i < n) out[i] += add;
}

__global__ void k_scan_total(const int32_t *in, const int32_t *out, int n, int32_t *total) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *total = n > 0 ? out[n - 1] + in[n - 1] : 0;
}

// ---------------------------------------------------------------------------------------------------------------
// Stable LSD radix sort of the cell keys, 8 bits a pass. Each block owns a tile of RS_TILE keys; the scatter ranks
// each key among the tile's earlier keys of the same digit with wavefront ballots (eight ballots give the lanes that
// share a lane's digit; a popcount of the lower ones is its rank in the wave), and the waves' counts are combined in
// LDS, so equal keys keep their order and the result lists each cell's spheres by ascending index.

constexpr int RS_ITEMS = 8;
constexpr int RS_TILE = THREADS * RS_ITEMS;

__global__ void k_rs_hist(const int32_t *keys, int n, int shift, int32_t *hist, int nblocks) {
    __shared__ int h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    int base = blockIdx.x * RS_TILE;
    for (int k = 0; k < RS_ITEMS; k++) {
        int i = base + k * THREADS + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255], 1);
    }
    __syncthreads();
    hist[threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

__global__ void k_rs_scatter(const int32_t *keysIn, const int32_t *valsIn, int32_t *keysOut, int32_t *valsOut, int n,
                             int shift, const int32_t *offsets, int nblocks) {
    __shared__ int base[256];
    __shared__ int wcount[THREADS / 64][256];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    base[threadIdx.x] = offsets[threadIdx.x * nblocks + blockIdx.x];
    const unsigned long long lower = (1ull << lane) - 1ull;
    for (int k = 0; k < RS_ITEMS; k++) {
        for (int w = 0; w < THREADS / 64; w++) wcount[w][threadIdx.x] = 0;
        __syncthreads();
        int i = blockIdx.x * RS_TILE + k * THREADS + threadIdx.x;
        bool valid = i < n;
        int key = valid ? keysIn[i] : 0;
        int val = valid ? (valsIn ? valsIn[i] : i) : 0;
        int d = (key >> shift) & 255;
        unsigned long long peers = __ballot(valid);
        for (int b = 0; b < 8; b++) {
            int bit = (d >> b) & 1;
            unsigned long long bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        int rank = __popcll(peers & lower);
        if (valid && rank == 0) wcount[wave][d] = __popcll(peers);
        __syncthreads();
        {
            int dg = threadIdx.x;
            int run = base[dg];
            for (int w = 0; w < THREADS / 64; w++) {
                int t = wcount[w][dg];
                wcount[w][dg] = run;
                run += t;
            }
            base[dg] = run;
        }
        __syncthreads();
        if (valid) {
            int pos = wcount[wave][d] + rank;
            keysOut[pos] = key;
            valsOut[pos] = val;
        }
        __syncthreads();
    }
}

// starts[c] = the first sorted position whose key is at least c, for every c in 0..cells.
__global__ void k_starts(const int32_t *sortedKeys, int n, int32_t *starts, int cells) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c > cells) return;
    int lo = 0, hi = n;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (sortedKeys[mid] < c) lo = mid + 1;
        else hi = mid;
    }
    starts[c] = lo;
}

}  // namespace

int scan_tmp_words(int n) {
    int words = 0;
    while (n > SCAN_TILE) {
        n = (n + SCAN_TILE - 1) / SCAN_TILE;
        words += n + 1;
    }
    return words + 1;
}

void scan_exclusive(const int32_t *in, int32_t *out, int n, int32_t *total, int32_t *tmp, hipStream_t s) {
    if (n <= 0) {
        if (total) launch_fill_i32(total, 0, 1, s);
        return;
    }
    int blocks = (n + SCAN_TILE - 1) / SCAN_TILE;
    // The total needs in[n-1] before `out` overwrites it when the two alias.
    int32_t *last = nullptr;
    if (total && in == out) {
        last = tmp;
        tmp += 1;
        (void)hipMemcpyAsync(last, in + n - 1, sizeof(int32_t), hipMemcpyDeviceToDevice, s);
    }
    if (blocks == 1) {
        k_scan_tile<<<1, THREADS, 0, s>>>(in, out, n, nullptr);
    } else {
        int32_t *sums = tmp;
        k_scan_tile<<<blocks, THREADS, 0, s>>>(in, out, n, sums);
        scan_exclusive(sums, sums, blocks, nullptr, tmp + blocks, s);
        k_scan_add<<<blocks, THREADS, 0, s>>>(out, n, sums);
    }
    check_launch("scan");
    if (total) k_scan_total<<<1, 1, 0, s>>>(last ? last - (n - 1) : in, out, n, total);
}

void launch_fill(float *p, float value, int64_t count, hipStream_t s) {
    if (count <= 0) return;
    k_fill<<<blocks_for(count), THREADS, 0, s>>>(p, value, count);
    check_launch("fill");
}

void launch_fill_i32(int32_t *p, int32_t value, int64_t count, hipStream_t s) {
    if (count <= 0) return;
    k_fill_i32<<<blocks_for(count), THREADS, 0, s>>>(p, value, count);
    check_launch("fill");
}

void launch_predict(const SphereDev &D, hipStream_t s) {
    k_predict<<<blocks_for(D.n), THREADS, 0, s>>>(D);
    check_launch("predict");
}

void launch_bin(const SphereDev &D, hipStream_t s) {
    k_bin<<<blocks_for(D.n), THREADS, 0, s>>>(D);
    check_launch("bin");
}

void launch_sort(const SphereDev &D, const SortScratch &S, hipStream_t s) {
    int bits = 1;
    while ((1 << bits) < D.cells && bits < 31) bits++;
    int passes = (bits + 7) / 8;
    int nblocks = (D.n + RS_TILE - 1) / RS_TILE;
    const int32_t *kin = D.keys;
    const int32_t *vin = nullptr;
    for (int p = 0; p < passes; p++) {
        bool toOrder = ((passes - 1 - p) % 2) == 0;
        int32_t *kout = toOrder ? S.keysB : S.keysA;
        int32_t *vout = toOrder ? D.order : S.valsA;
        k_rs_hist<<<nblocks, THREADS, 0, s>>>(kin, D.n, 8 * p, S.hist, nblocks);
        scan_exclusive(S.hist, S.hist, 256 * nblocks, nullptr, S.scanTmp, s);
        k_rs_scatter<<<nblocks, THREADS, 0, s>>>(kin, vin, kout, vout, D.n, 8 * p, S.hist, nblocks);
        check_launch("radix sort");
        kin = kout;
        vin = vout;
    }
    k_starts<<<blocks_for((int64_t)D.cells + 1), THREADS, 0, s>>>(kin, D.n, D.starts, D.cells);
    check_launch("starts");
}

void launch_warm_walls(const SphereDev &D, hipStream_t s) {
    k_warm_walls<<<blocks_for(D.n), THREADS, 0, s>>>(D);
    check_launch("warm walls");
}

void launch_clear_contacts(const SphereDev &D, hipStream_t s) {
    k_clear_contacts<<<blocks_for(D.n), THREADS, 0, s>>>(D);
    check_launch("clear contacts");
}

void launch_list(const SphereDev &D, const SortScratch &S, hipStream_t s) {
    k_list_count<<<blocks_for(D.n), THREADS, 0, s>>>(D);
    check_launch("list count");
    scan_exclusive(D.cur.cnt, D.cur.off, D.n, D.count + LISTED, S.scanTmp, s);
    k_list_write<<<blocks_for(D.n), THREADS, 0, s>>>(D);
    check_launch("list write");
}

void launch_round(const SphereDev &D, const RoundConst &rc, hipStream_t s) {
    k_round<<<blocks_for(D.cap), THREADS, 0, s>>>(D, rc);
    check_launch("contact round");
}

void launch_step_begin(PassCtl *ctl, PassSlot *slots, int nslots, hipStream_t s) {
    int words = nslots * (int)(sizeof(PassSlot) / sizeof(int));
    k_step_begin<<<blocks_for(words), THREADS, 0, s>>>(ctl, slots, words);
    check_launch("step begin");
}

void launch_pass(const SphereDev &D, PassCtl *ctl, PassSlot *slot, const OpenLists &L, int32_t *opening,
                 int32_t *enough, int it, int iterations, bool rub, int wide, bool local, const PassSubmit &submit,
                 hipStream_t s) {
    if (wide < 0 || wide > MAX_WIDE) throw std::runtime_error("pass: wide rounds out of range");
    if ((int64_t)L.segments * SEG < D.cap) throw std::runtime_error("pass: open lists shorter than the contact list");
    int limit = D.cap + 2 * CYCLE;
    if (L.segments == 0) wide = 0;  // no contact can be listed: the tail alone finds the pass empty
    for (int j = 0; j < wide; j++) {
        int last = j == wide - 1 ? 1 : 0;
        const int4 *in = L.buf[j & 1];
        const int32_t *segIn = L.seg[j & 1];
        int4 *out = L.buf[(j + 1) & 1];
        int32_t *segOut = L.seg[(j + 1) & 1];
        auto kernel = rub ? (j == 0 ? k_round_wide<true, true> : k_round_wide<true, false>)
                          : (j == 0 ? k_round_wide<false, true> : k_round_wide<false, false>);
        kernel<<<L.segments, THREADS, 0, s>>>(D, ctl, slot, in, segIn, out, segOut, it, iterations, j, last,
                                              local ? 1 : 0);
    }
    if (local) {
        if (D.n > LOCAL_MOST) throw std::runtime_error("pass: a pass in LDS takes at most LOCAL_MOST spheres");
        // The claims follow the kernel's one static word, at the next 8-byte boundary.
        size_t lds = (size_t)3 * D.n * sizeof(long long) + sizeof(long long);
        static bool sized = false;  // both instances allowed the LDS of LOCAL_MOST spheres, once a process
        if (!sized) {
            int most = (int)((3 * LOCAL_MOST + 1) * sizeof(long long));
            if (hipFuncSetAttribute((const void *)k_round_tail<true, true, LOCAL_THREADS>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, most) != hipSuccess ||
                hipFuncSetAttribute((const void *)k_round_tail<false, true, LOCAL_THREADS>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, most) != hipSuccess)
                throw std::runtime_error("pass: the claims of LOCAL_MOST spheres do not fit in LDS");
            sized = true;
        }
        if (rub)
            k_round_tail<true, true, LOCAL_THREADS><<<1, LOCAL_THREADS, lds, s>>>(D, ctl, slot, L, opening, enough, it,
                                                                                 iterations, wide, limit, submit);
        else
            k_round_tail<false, true, LOCAL_THREADS><<<1, LOCAL_THREADS, lds, s>>>(D, ctl, slot, L, opening, enough,
                                                                                  it, iterations, wide, limit, submit);
    } else if (rub) {
        k_round_tail<true, false, TAIL><<<1, TAIL, 0, s>>>(D, ctl, slot, L, opening, enough, it, iterations, wide,
                                                           limit, submit);
    } else {
        k_round_tail<false, false, TAIL><<<1, TAIL, 0, s>>>(D, ctl, slot, L, opening, enough, it, iterations, wide,
                                                            limit, submit);
    }
    check_launch("pass");
}

void launch_walls(const SphereDev &D, hipStream_t s) {
    k_walls<<<blocks_for(D.n), THREADS, 0, s>>>(D);
    check_launch("walls");
}

void launch_velocity(const SphereDev &D, hipStream_t s) {
    k_velocity<<<blocks_for(D.n), THREADS, 0, s>>>(D);
    check_launch("velocity");
}

void launch_bounce(const SphereDev &D, bool grid, hipStream_t s) {
    if (grid) k_bounce<true><<<blocks_for(D.n), THREADS, 0, s>>>(D);
    else k_bounce<false><<<blocks_for(D.n), THREADS, 0, s>>>(D);
    check_launch("bounce");
}

void launch_bounced(const SphereDev &D, hipStream_t s) {
    k_bounced<<<blocks_for(D.n), THREADS, 0, s>>>(D);
    check_launch("bounced");
}

void launch_jacobi_solve(const SphereDev &D, bool grid, hipStream_t s) {
    if (grid) k_jacobi_solve<true><<<blocks_for(D.n), THREADS, 0, s>>>(D);
    else k_jacobi_solve<false><<<blocks_for(D.n), THREADS, 0, s>>>(D);
    check_launch("jacobi solve");
}

void launch_jacobi_apply(const SphereDev &D, hipStream_t s) {
    k_jacobi_apply<<<blocks_for(D.n), THREADS, 0, s>>>(D);
    check_launch("jacobi apply");
}

void launch_show(const SphereDev &D, bool grid, hipStream_t s) {
    if (grid) k_show<true><<<blocks_for(D.n), THREADS, 0, s>>>(D);
    else k_show<false><<<blocks_for(D.n), THREADS, 0, s>>>(D);
    check_launch("show");
}

}  // namespace vr
