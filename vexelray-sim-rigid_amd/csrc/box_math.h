// box_math.h — the box world's per-body and per-pair arithmetic, as device functions for gfx950.
//
// The rigid-body algebra of the reference's box world (vexelray-sim-rigid-core/.../box/Boxes.java, BoxContact.java,
// BoxPairs.java): a body as a solve holds it (Boxes.Moving), the soft contact and rolling resistance (Boxes.contact,
// Boxes.roll), and the separating-axis narrow phase with its clipped, reduced manifold (BoxContact). Every f32
// operation follows the reference's expression trees (a dot product is (x0 y0 + x1 y1) + x2 y2), and the library is
// built with -ffp-contract=off, so results are the reference's arithmetic operation by operation.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace vrbox {

// Boxes.java:103-121, the parameter words.
enum {
    H = 0, GX, GY, GZ, SX, SY, SZ, MU_STATIC, MU_KINETIC, MARGIN, BIAS, MASS_SCALE, IMPULSE_SCALE, STATIC_BIAS,
    STATIC_MASS_SCALE, STATIC_IMPULSE_SCALE, PUSH_SPEED, MU_ROLLING, PARAM_COUNT
};

constexpr int CORNERS = 8, WALLS = 6;
constexpr int LAMBDA_WORDS = WALLS * CORNERS;  // Boxes.java:138
constexpr int RUB_WORDS = 3 * LAMBDA_WORDS;    // Boxes.java:140
constexpr float SPHERE_INERTIA = 0.4f;          // Boxes.java:156

constexpr int SLOTS = 32;                       // BoxBroad.java:46
constexpr int MANIFOLD_WORDS = 40, COUNT = 3, AXIS = 4, POINTS = 8, POINT_WORDS = 8;  // BoxContact.java:77-88
constexpr int ANCHOR_FIRST = 0, ANCHOR_SECOND = 3, DEPTH = 6, MOST_POINTS = 4;
constexpr int CANDIDATES = 24;                  // BoxContact.java:91
constexpr float RELATIVE = 0.98f, ABSOLUTE = 0.01f;   // BoxContact.java:98-100
constexpr float PARALLEL = 1e-10f, SAME = 1e-12f, SLACK = 1e-5f, OFF_LINE = 1e-4f;  // BoxContact.java:103-117
constexpr int SPHERE_BOX = 15, SPHERE_SPHERE = 16;    // BoxContact.java:252-253
constexpr int IMPULSE_WORDS = MOST_POINTS * 7, POINT_IMPULSE = 7;  // BoxPairs.java:37-39
constexpr float SAME_POINT = 0.05f;             // BoxPairs.java:158
constexpr int MOST_ROUNDS = 32;                 // BoxStep.java:45
constexpr int SHOWN_STRIDE = 20;                // Boxes.java:792

enum Sweep { WARM, SOLVE, RELAX };

__device__ __forceinline__ float mx(float a, float b) { return a < b ? b : a; }  // Math.max / std::max
__device__ __forceinline__ unsigned bits(float v) { return __float_as_uint(v); }
__device__ __forceinline__ float mn(float a, float b) { return b < a ? b : a; }  // Math.min / std::min

// Boxes.dot / BoxContact.dot: (x0 y0 + x1 y1) + x2 y2.
__device__ __forceinline__ float dot(const float *x, const float *y) { return (x[0] * y[0] + x[1] * y[1]) + x[2] * y[2]; }

__device__ __forceinline__ void cross(const float *x, const float *y, float *z) {
    float a = x[1] * y[2] - x[2] * y[1];
    float b = x[2] * y[0] - x[0] * y[2];
    float c = x[0] * y[1] - x[1] * y[0];
    z[0] = a;
    z[1] = b;
    z[2] = c;
}

// 2 / |q|^2, the scale of Boxes.rotate's t. A turn's conjugate gives the same bits (each square is of a negated
// word, which squares to the same bits), so a body whose turn stays as it is over a sweep works it out once.
__device__ __forceinline__ float turn_scale(const float *q) {
    return 2.0f / ((q[0] * q[0] + q[1] * q[1]) + (q[2] * q[2] + q[3] * q[3]));
}

// Boxes.rotate, Boxes.java:907-921: p turned by q, p + w t + q x t with t = 2 q x p / |q|^2, given twice =
// turn_scale(q).
__device__ __forceinline__ void rotate_scaled(const float *q, float twice, const float *p, float *out) {
    float t[3], qt[3];
    cross(q, p, t);
    for (int a = 0; a < 3; a++) t[a] = twice * t[a];
    cross(q, t, qt);
    float r[3];
    for (int a = 0; a < 3; a++) r[a] = p[a] + (q[3] * t[a] + qt[a]);
    for (int a = 0; a < 3; a++) out[a] = r[a];
}

__device__ __forceinline__ void rotate(const float *q, const float *p, float *out) {
    rotate_scaled(q, turn_scale(q), p, out);
}

__device__ __forceinline__ void conjugate(const float *q, float *out) {
    out[0] = -q[0];
    out[1] = -q[1];
    out[2] = -q[2];
    out[3] = q[3];
}

// Boxes.turned, Boxes.java:877-899.
__device__ __forceinline__ void turned(const float *q, const float *omega, float dt, float *now) {
    float a[3];
    for (int k = 0; k < 3; k++) a[k] = (0.5f * dt) * omega[k];
    for (int k = 0; k < 3; k++) {
        int k1 = (k + 1) % 3, k2 = (k + 2) % 3;
        now[k] = q[k] + (a[k] * q[3] + (a[k1] * q[k2] - a[k2] * q[k1]));
    }
    now[3] = q[3] - (a[0] * q[0] + (a[1] * q[1] + a[2] * q[2]));
    float len2 = (now[0] * now[0] + now[1] * now[1]) + (now[2] * now[2] + now[3] * now[3]);
    float inverse = 1.0f / sqrtf(mx(len2, 1e-30f));
    for (int k = 0; k < 4; k++) now[k] = now[k] * inverse;
}


// ---- a body as the solve holds it (Boxes.Moving, Boxes.java:660-773) ----

struct Moving {
    int s;
    float shape, im;
    float at[3], q[4], half[3], inverse[3], vel[3], mom[3], omega[3];
    float twice;  // turn_scale(q), for every turn of a vector by q or its conjugate

    __device__ void turned_by(const float *p, float *out) const { rotate_scaled(q, twice, p, out); }

    __device__ bool sphere() const { return shape > 0.5f; }

    // R I^-1 R^T x (Boxes.java:711-713).
    __device__ void inverseInertiaTimes(const float *x, float *out) const {
        float cq[4], own[3];
        conjugate(q, cq);
        rotate_scaled(cq, twice, x, own);
        for (int a = 0; a < 3; a++) own[a] = inverse[a] * own[a];
        rotate_scaled(q, twice, own, out);
    }

    __device__ float inverseMassAlong(const float *r, const float *n) const {  // Boxes.java:716-719
        float rn[3], t[3];
        cross(r, n, rn);
        inverseInertiaTimes(rn, t);
        return im + dot(rn, t);
    }

    __device__ void pointVelocity(const float *r, float *out) const {  // Boxes.java:722-729
        float turning[3];
        cross(omega, r, turning);
        for (int a = 0; a < 3; a++) out[a] = vel[a] + turning[a];
    }

    __device__ void apply(const float *r, const float *p) {  // Boxes.java:732-742
        for (int a = 0; a < 3; a++) vel[a] = vel[a] + im * p[a];
        float angular[3], turn[3];
        cross(r, p, angular);
        inverseInertiaTimes(angular, turn);
        for (int a = 0; a < 3; a++) {
            mom[a] = mom[a] + angular[a];
            omega[a] = omega[a] + turn[a];
        }
    }

    __device__ void turn(const float *j) {  // Boxes.java:745-751
        float spin[3];
        inverseInertiaTimes(j, spin);
        for (int a = 0; a < 3; a++) {
            mom[a] = mom[a] + j[a];
            omega[a] = omega[a] + spin[a];
        }
    }
};


struct Bufs {
    float *x[3], *vel[3], *l[3], *q[4], *im, *half[3], *shape;
    const float *params;
    float *lambda, *rub, *roll;
};

// Boxes.inverseInertia, Boxes.java:854-871: the body-frame inverse inertia of a box, or of a ball (shape > 0.5).
__device__ __forceinline__ void inverse_inertia(const float *half, float im, float shape, float *inverse) {
    float h2[3];
    for (int a = 0; a < 3; a++) h2[a] = half[a] * half[a];
    for (int a = 0; a < 3; a++) inverse[a] = (3.0f * im) / (h2[(a + 1) % 3] + h2[(a + 2) % 3]);
    if (shape > 0.5f)
        for (int a = 0; a < 3; a++) inverse[a] = im / (SPHERE_INERTIA * h2[0]);
}

__device__ Moving load_moving(const Bufs &B, int s) {
    Moving m;
    m.s = s;
    m.im = B.im[s];
    m.shape = B.shape[s];
    for (int a = 0; a < 3; a++) {
        m.at[a] = B.x[a][s];
        m.half[a] = B.half[a][s];
        m.vel[a] = B.vel[a][s];
        m.mom[a] = B.l[a][s];
    }
    for (int k = 0; k < 4; k++) m.q[k] = B.q[k][s];
    m.twice = turn_scale(m.q);
    inverse_inertia(m.half, m.im, m.shape, m.inverse);
    m.inverseInertiaTimes(m.mom, m.omega);
    return m;
}

__device__ void store_moving(const Bufs &B, const Moving &m) {
    for (int a = 0; a < 3; a++) {
        B.vel[a][m.s] = m.vel[a];
        B.l[a][m.s] = m.mom[a];
    }
}

// store_moving for a body loaded as `was`: a word whose bits did not change is left as it is.
__device__ void store_moving_changed(const Bufs &B, const Moving &m, const Moving &was) {
    for (int a = 0; a < 3; a++) {
        if (bits(m.vel[a]) != bits(was.vel[a])) B.vel[a][m.s] = m.vel[a];
        if (bits(m.mom[a]) != bits(was.mom[a])) B.l[a][m.s] = m.mom[a];
    }
}

// ---- a body as the sweeps of the pass-by-pass step hold it: two records a body ----
//
// frame, 4 float4 a body: (at, im), q, (half, shape), (inverse, turn_scale(q)); fixed over a run of sweeps (positions rewrite
// it). motion, 2 float4 a body: (vel, 0), (mom, 0); what the sweeps change. A round's bodies lie anywhere in the
// world, so it gathers each from these records, two cache lines, not from the state's 18 fields. (One 128-byte record
// a body, frame and motion in one line, measured slower on the 2,048-scene batch: 116 against 111 ms of colour
// rounds a step, and 28 against 26 ms of walls sweeps, whose coalesced loads of the motion then stride.)
constexpr int FRAME_VECTORS = 4, MOTION_VECTORS = 2;

// Body s's frame record and the motion record at v.
__device__ __forceinline__ Moving load_framed_at(const float4 *frame, const float4 *v, int s) {
    const float4 *f = frame + (size_t)s * FRAME_VECTORS;
    float4 a = f[0], b = f[1], c = f[2], d = f[3], e = v[0], g = v[1];
    Moving m;
    m.s = s;
    m.at[0] = a.x, m.at[1] = a.y, m.at[2] = a.z, m.im = a.w;
    m.q[0] = b.x, m.q[1] = b.y, m.q[2] = b.z, m.q[3] = b.w;
    m.half[0] = c.x, m.half[1] = c.y, m.half[2] = c.z, m.shape = c.w;
    m.inverse[0] = d.x, m.inverse[1] = d.y, m.inverse[2] = d.z;
    m.twice = d.w;  // turn_scale(q), kept in the record beside the inverse inertia
    m.vel[0] = e.x, m.vel[1] = e.y, m.vel[2] = e.z;
    m.mom[0] = g.x, m.mom[1] = g.y, m.mom[2] = g.z;
    m.inverseInertiaTimes(m.mom, m.omega);
    return m;
}

__device__ __forceinline__ Moving load_framed(const float4 *frame, const float4 *motion, int s) {
    return load_framed_at(frame, motion + (size_t)s * MOTION_VECTORS, s);
}

// Where the sweeps find a body's motion record: the world's records, or a workgroup's copy of a run of them in LDS
// (bodies b0 ..., k_group_rounds).
struct GlobalMotion {
    float4 *m;
    __device__ __forceinline__ float4 *at(int s) const { return m + (size_t)s * MOTION_VECTORS; }
};
struct LocalMotion {
    float4 *m;
    int b0;
    __device__ __forceinline__ float4 *at(int s) const { return m + (s - b0) * MOTION_VECTORS; }
};

// The state's fields with the motion record's velocity and momentum (the walls sweep, whose neighbouring lanes hold
// neighbouring bodies, so the fields' loads coalesce).
// The state's fields of body s with the velocity and momentum of the motion record at v.
__device__ __forceinline__ Moving load_moved_at(const Bufs &B, const float4 *v, int s) {
    float4 e = v[0], g = v[1];
    Moving m;
    m.s = s;
    m.im = B.im[s];
    m.shape = B.shape[s];
    for (int a = 0; a < 3; a++) {
        m.at[a] = B.x[a][s];
        m.half[a] = B.half[a][s];
    }
    for (int k = 0; k < 4; k++) m.q[k] = B.q[k][s];
    m.twice = turn_scale(m.q);
    m.vel[0] = e.x, m.vel[1] = e.y, m.vel[2] = e.z;
    m.mom[0] = g.x, m.mom[1] = g.y, m.mom[2] = g.z;
    inverse_inertia(m.half, m.im, m.shape, m.inverse);
    m.inverseInertiaTimes(m.mom, m.omega);
    return m;
}

__device__ __forceinline__ Moving load_moved(const Bufs &B, const float4 *motion, int s) {
    return load_moved_at(B, motion + (size_t)s * MOTION_VECTORS, s);
}

__device__ __forceinline__ void store_motion_at(float4 *v, const Moving &m) {
    v[0] = make_float4(m.vel[0], m.vel[1], m.vel[2], 0.f);
    v[1] = make_float4(m.mom[0], m.mom[1], m.mom[2], 0.f);
}

__device__ __forceinline__ void store_motion(float4 *motion, const Moving &m) {
    store_motion_at(motion + (size_t)m.s * MOTION_VECTORS, m);
}

// store_motion for a body loaded as `was`: a vector none of whose words changed is left as it is.
__device__ __forceinline__ void store_motion_changed_at(float4 *v, const Moving &m, const Moving &was) {
    if ((bits(m.vel[0]) != bits(was.vel[0])) | (bits(m.vel[1]) != bits(was.vel[1])) |
        (bits(m.vel[2]) != bits(was.vel[2])))
        v[0] = make_float4(m.vel[0], m.vel[1], m.vel[2], 0.f);
    if ((bits(m.mom[0]) != bits(was.mom[0])) | (bits(m.mom[1]) != bits(was.mom[1])) |
        (bits(m.mom[2]) != bits(was.mom[2])))
        v[1] = make_float4(m.mom[0], m.mom[1], m.mom[2], 0.f);
}

__device__ __forceinline__ void store_motion_changed(float4 *motion, const Moving &m, const Moving &was) {
    store_motion_changed_at(motion + (size_t)m.s * MOTION_VECTORS, m, was);
}

// Boxes.Soft, Boxes.java:421-450.
struct Soft {
    float h, margin, bias, massScale, impulseScale, push, muStatic, muKinetic;
    __device__ Soft(const float *P, bool fixed) {
        h = P[H];
        margin = P[MARGIN];
        bias = P[fixed ? STATIC_BIAS : BIAS];
        massScale = P[fixed ? STATIC_MASS_SCALE : MASS_SCALE];
        impulseScale = P[fixed ? STATIC_IMPULSE_SCALE : IMPULSE_SCALE];
        push = P[PUSH_SPEED];
        muStatic = P[MU_STATIC];
        muKinetic = P[MU_KINETIC];
    }
    __device__ void stiffen(const float *P) {
        bias = P[STATIC_BIAS];
        massScale = P[STATIC_MASS_SCALE];
        impulseScale = P[STATIC_IMPULSE_SCALE];
    }
};

// The impulse p on the second at rB, and -p on the first at rA (Boxes.java:643-652).
__device__ void apply_pair(Moving *first, const float *rA, Moving &second, const float *rB, const float *p) {
    second.apply(rB, p);
    if (first) {
        float back[3] = {-p[0], -p[1], -p[2]};
        first->apply(rA, back);
    }
}

__device__ void relative(const Moving *first, const float *rA, const Moving &second, const float *rB, float *out) {
    float vb[3];
    second.pointVelocity(rB, vb);
    if (!first) {
        for (int a = 0; a < 3; a++) out[a] = vb[a];
        return;
    }
    float va[3];
    first->pointVelocity(rA, va);
    for (int a = 0; a < 3; a++) out[a] = vb[a] - va[a];
}

__device__ float mass_along(const Moving *first, const float *rA, const Moving &second, const float *rB, const float *n) {
    float k = second.inverseMassAlong(rB, n);
    return first ? k + first->inverseMassAlong(rA, n) : k;
}

// The residual of one point's update in a solve or relax sweep (no counterpart in the reference, whose sweeps run a
// fixed count): the size of the whole impulse change the update made, normal `dn` and tangent `rubbed` together,
// times the inverse mass along the normal `k`, so a velocity in m/s: how far the sweep still moved the point. A
// point the sweep passes over has none.
__device__ __forceinline__ float point_residual(float k, float dn, const float *rubbed) {
    return k * sqrtf(dn * dn + dot(rubbed, rubbed));
}

// res = the larger of res and r, where NaN, once met, stays (a residual that is not a number is never small).
__device__ __forceinline__ void keep_largest(float &res, float r) {
    if (res == res && !(r <= res)) res = r;
}

// Boxes.contact, Boxes.java:462-542: one point, one sweep, on the point's impulse held in registers: `normal` and
// `tangent` come in as the point holds them and leave as Boxes.contact leaves them; `wrote` says whether it writes
// them (not in a warm sweep, nor for a point it passes over). Returns whether the point holds an impulse afterwards.
// With `res`, the point's residual (point_residual) is kept in it when larger; nothing of the contact depends on it.
__device__ bool contact_held(Sweep sweep, float gap, const float *n, Moving *first, const float *rA, Moving &second,
                             const float *rB, float &normalHeld, float *tangentHeld, bool &wrote, const Soft &soft,
                             float *res = nullptr) {
    wrote = false;
    float normal = normalHeld;
    float tangent[3] = {tangentHeld[0], tangentHeld[1], tangentHeld[2]};
    bool holding = normal > 0 || dot(tangent, tangent) > 0;
    if (sweep == WARM) {
        if (holding) {
            float p[3];
            for (int a = 0; a < 3; a++) p[a] = normal * n[a] + tangent[a];
            apply_pair(first, rA, second, rB, p);
        }
        return holding;
    }
    if (!(gap < soft.margin || holding)) return false;
    float k = mass_along(first, rA, second, rB, n);
    float rel[3];
    relative(first, rA, second, rB, rel);
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
    apply_pair(first, rA, second, rB, p);

    float sliding[3];
    relative(first, rA, second, rB, sliding);
    float along = dot(sliding, n);
    for (int a = 0; a < 3; a++) sliding[a] = sliding[a] - along * n[a];
    float whole[3] = {tangent[0], tangent[1], tangent[2]};
    float speed = sqrtf(dot(sliding, sliding));
    if (speed > 1e-12f) {
        float t[3];
        for (int a = 0; a < 3; a++) t[a] = sliding[a] / speed;
        float kt = mass_along(first, rA, second, rB, t);
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
    apply_pair(first, rA, second, rB, rubbed);
    if (res) keep_largest(*res, point_residual(k, total - normal, rubbed));
    normalHeld = total;
    for (int a = 0; a < 3; a++) tangentHeld[a] = whole[a];
    wrote = true;
    return total > 0 || dot(whole, whole) > 0;
}

// Boxes.contact on the point's words in memory. The tangent impulse's three words are `stride` apart (1 for a pair's
// point; the body count for a wall's, whose impulses are laid out word by word over the bodies).
__device__ bool contact(Sweep sweep, float gap, const float *n, Moving *first, const float *rA, Moving &second, const float *rB,
             float *normals, size_t normalAt, float *tangents, size_t tangentAt, size_t stride, const Soft &soft,
             float *res = nullptr) {
    float normal = normals[normalAt];
    float tangent[3] = {tangents[tangentAt], tangents[tangentAt + stride], tangents[tangentAt + 2 * stride]};
    float was[4] = {normal, tangent[0], tangent[1], tangent[2]};
    bool wrote;
    bool holds = contact_held(sweep, gap, n, first, rA, second, rB, normal, tangent, wrote, soft, res);
    if (wrote) {  // a word written with the bits it holds is left as it is
        if (bits(normal) != bits(was[0])) normals[normalAt] = normal;
        for (int a = 0; a < 3; a++)
            if (bits(tangent[a]) != bits(was[1 + a])) tangents[tangentAt + a * stride] = tangent[a];
    }
    return holds;
}

__device__ void apply_angular(Moving *first, Moving &second, const float *j) {  // Boxes.java:603-612
    second.turn(j);
    if (first) {
        float back[3] = {-j[0], -j[1], -j[2]};
        first->turn(back);
    }
}

// Boxes.roll, Boxes.java:553-600, on the point's rolling impulse held in registers (`held`, in and out) and its
// normal impulse as the point's contact left it; `wrote` says whether Boxes.roll writes the rolling impulse.
__device__ void roll_held(Sweep sweep, Moving *first, Moving &second, float radius, float normal, float *held,
                          bool &wrote, const float *P) {
    wrote = false;
    float old[3] = {held[0], held[1], held[2]};
    if (sweep == WARM) {
        apply_angular(first, second, old);
        return;
    }
    float muRolling = P[MU_ROLLING];
    float whole[3], turning[3];
    for (int a = 0; a < 3; a++) {
        whole[a] = old[a];
        float against = first ? first->omega[a] : 0.f;
        turning[a] = second.omega[a] - against;
    }
    float speed = sqrtf(dot(turning, turning));
    if (speed > 1e-12f && muRolling > 0) {
        float t[3], it[3];
        for (int a = 0; a < 3; a++) t[a] = turning[a] / speed;
        second.inverseInertiaTimes(t, it);
        float k = dot(t, it);
        if (first) {
            float jt[3];
            first->inverseInertiaTimes(t, jt);
            k = k + dot(t, jt);
        }
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
    apply_angular(first, second, change);
}

// Boxes.roll on the point's words in memory (the rolling impulse's three words `stride` apart).
__device__ void roll(Sweep sweep, Moving *first, Moving &second, float radius, const float *normals, size_t normalAt,
          float *rolls, size_t rollAt, size_t stride, const float *P) {
    float held[3] = {rolls[rollAt], rolls[rollAt + stride], rolls[rollAt + 2 * stride]};
    float was[3] = {held[0], held[1], held[2]};
    bool wrote;
    roll_held(sweep, first, second, radius, normals[normalAt], held, wrote, P);
    if (wrote)  // a word written with the bits it holds is left as it is
        for (int a = 0; a < 3; a++)
            if (bits(held[a]) != bits(was[a])) rolls[rollAt + a * stride] = held[a];
}

// ---- the narrow phase (BoxContact.java) ----

struct Solid {  // BoxContact.Solid, :390-422
    float centre[3], axes[3][3], half[3];
    __device__ void own(const float *point, float *out) const {
        float off[3] = {point[0] - centre[0], point[1] - centre[1], point[2] - centre[2]};
        float o0 = dot(off, axes[0]), o1 = dot(off, axes[1]), o2 = dot(off, axes[2]);
        out[0] = o0;
        out[1] = o1;
        out[2] = o2;
    }
};

__device__ Solid load_solid(const Bufs &B, int s) {
    Solid S;
    for (int a = 0; a < 3; a++) {
        S.centre[a] = B.x[a][s];
        S.half[a] = B.half[a][s];
    }
    float x = B.q[0][s], y = B.q[1][s], z = B.q[2][s], w = B.q[3][s];
    S.axes[0][0] = 1.0f - 2.0f * (y * y + z * z);
    S.axes[0][1] = 2.0f * (x * y + z * w);
    S.axes[0][2] = 2.0f * (x * z - y * w);
    S.axes[1][0] = 2.0f * (x * y - z * w);
    S.axes[1][1] = 1.0f - 2.0f * (x * x + z * z);
    S.axes[1][2] = 2.0f * (y * z + x * w);
    S.axes[2][0] = 2.0f * (x * z + y * w);
    S.axes[2][1] = 2.0f * (y * z - x * w);
    S.axes[2][2] = 1.0f - 2.0f * (x * x + y * y);
    return S;
}

__device__ __forceinline__ float sign_of(float x) { return x < 0 ? -1.0f : 1.0f; }  // BoxContact.sign, :880-884

__device__ void write_normal(float *m, const float *normal, float axis) {
    for (int x = 0; x < 3; x++) m[x] = normal[x];
    m[AXIS] = axis;
}

__device__ void write_point(float *m, int at, const float *onFirst, const float *onSecond, float depth) {
    for (int x = 0; x < 3; x++) {
        m[at + ANCHOR_FIRST + x] = onFirst[x];
        m[at + ANCHOR_SECOND + x] = onSecond[x];
    }
    m[at + DEPTH] = depth;
}

struct Edge {  // BoxContact.Edge, :427-500
    float normal[3] = {0, 0, 0}, onFirst[3] = {0, 0, 0}, alongFirst[3] = {0, 0, 0};
    float halfFirst = 0, onSecond[3] = {0, 0, 0}, alongSecond[3] = {0, 0, 0}, halfSecond = 0, axis = 0;

    __device__ void record(const Solid &first, const Solid &second, int k, int j, const float *ax, const float *d) {
        float toward = 1;
        if (dot(d, ax) < 0) toward = -1;
        for (int a = 0; a < 3; a++) {
            normal[a] = ax[a] * toward;
            onFirst[a] = first.centre[a];
            onSecond[a] = second.centre[a];
            alongFirst[a] = first.axes[k][a];
            alongSecond[a] = second.axes[j][a];
        }
        for (int m = 0; m < 3; m++) {
            if (m != k) {
                float sg = sign_of(dot(first.axes[m], normal));
                for (int a = 0; a < 3; a++) onFirst[a] = onFirst[a] + (sg * first.half[m]) * first.axes[m][a];
            }
            if (m != j) {
                float sg = sign_of(dot(second.axes[m], normal));
                for (int a = 0; a < 3; a++) onSecond[a] = onSecond[a] - (sg * second.half[m]) * second.axes[m][a];
            }
        }
        halfFirst = first.half[k];
        halfSecond = second.half[j];
        axis = (float)(6 + 3 * k + j);
    }

    __device__ void contact(const Solid &first, const Solid &second, float bestEdge, float *m) const {
        float w[3] = {onFirst[0] - onSecond[0], onFirst[1] - onSecond[1], onFirst[2] - onSecond[2]};
        float c = dot(alongFirst, alongSecond);
        float dA = dot(alongFirst, w);
        float dB = dot(alongSecond, w);
        float den = 1.0f - c * c;
        float s = (c * dB - dA) / den;
        float u = (dB - c * dA) / den;
        s = mx(-halfFirst, mn(halfFirst, s));
        u = mx(-halfSecond, mn(halfSecond, u));
        float pf[3], ps[3];
        for (int a = 0; a < 3; a++) {
            pf[a] = onFirst[a] + s * alongFirst[a];
            ps[a] = onSecond[a] + u * alongSecond[a];
        }
        write_normal(m, normal, axis);
        float of[3], os[3];
        first.own(pf, of);
        second.own(ps, os);
        write_point(m, POINTS, of, os, -bestEdge);
    }
};

struct Face {  // BoxContact.Face, :505-817
    const Solid &first, &second;
    float flip, margin, refCentre[3], n[3], u[3], w[3], hn, hu, hw, axis, face[3], a[3], c[3], ha, hc;
    float cand[CANDIDATES][4];
    int count = 0;

    __device__ Face(const Solid &f, const Solid &s, float fromSecond, const float *bestFaceAxis, float mrg)
        : first(f), second(s), margin(mrg) {
        flip = 1;
        float k = bestFaceAxis[0];
        float refAxes[3][3], refHalf[3], incCentre[3], incAxes[3][3], incHalf[3];
        for (int e = 0; e < 3; e++) {
            refCentre[e] = first.centre[e];
            refHalf[e] = first.half[e];
            incCentre[e] = second.centre[e];
            incHalf[e] = second.half[e];
            for (int m = 0; m < 3; m++) {
                refAxes[m][e] = first.axes[m][e];
                incAxes[m][e] = second.axes[m][e];
            }
        }
        if (fromSecond > 0.5f) {
            flip = -1;
            k = bestFaceAxis[1];
            for (int e = 0; e < 3; e++) {
                refCentre[e] = second.centre[e];
                incCentre[e] = first.centre[e];
                refHalf[e] = second.half[e];
                incHalf[e] = first.half[e];
                for (int m = 0; m < 3; m++) {
                    refAxes[m][e] = second.axes[m][e];
                    incAxes[m][e] = first.axes[m][e];
                }
            }
        }
        axis = k + 1.5f * (1.0f - flip);
        for (int e = 0; e < 3; e++) n[e] = u[e] = w[e] = 0;
        hn = hu = hw = 0;
        for (int kk = 0; kk < 3; kk++) {
            int ku = (kk + 1) % 3, kw = (kk + 2) % 3;
            if (k > (float)kk - 0.5f && k < (float)kk + 0.5f) {
                for (int e = 0; e < 3; e++) {
                    n[e] = refAxes[kk][e];
                    u[e] = refAxes[ku][e];
                    w[e] = refAxes[kw][e];
                }
                hn = refHalf[kk];
                hu = refHalf[ku];
                hw = refHalf[kw];
            }
        }
        float apart[3] = {incCentre[0] - refCentre[0], incCentre[1] - refCentre[1], incCentre[2] - refCentre[2]};
        if (dot(apart, n) < 0)
            for (int e = 0; e < 3; e++) n[e] = -n[e];
        float most = -1, faceAxis[3] = {0, 0, 0}, hf = 0;
        for (int e = 0; e < 3; e++) a[e] = c[e] = 0;
        ha = hc = 0;
        for (int m = 0; m < 3; m++) {
            int ma = (m + 1) % 3, mc = (m + 2) % 3;
            float tn = fabsf(dot(incAxes[m], n));
            if (tn > most) {
                most = tn;
                for (int e = 0; e < 3; e++) {
                    faceAxis[e] = incAxes[m][e];
                    a[e] = incAxes[ma][e];
                    c[e] = incAxes[mc][e];
                }
                hf = incHalf[m];
                ha = incHalf[ma];
                hc = incHalf[mc];
            }
        }
        float facing = sign_of(dot(faceAxis, n));
        for (int e = 0; e < 3; e++) face[e] = incCentre[e] - (facing * hf) * faceAxis[e];
    }

    __device__ void keep(bool when, const float *point, float depth) {  // :710-722
        float deep = depth;
        if (when && deep > -margin && count < CANDIDATES) {
            for (int x = 0; x < 3; x++) cand[count][x] = point[x];
            cand[count][3] = deep;
            count++;
        }
    }

    __device__ void contact(float *m, int &outCount) {  // :621-707
        const float signs[4][2] = {{1, 1}, {-1, 1}, {-1, -1}, {1, -1}};
        const float slack1 = 1.0f + SLACK;
        float corner[4][3], cu[4], cw[4], depth[4];
        for (int e = 0; e < 4; e++) {
            for (int x = 0; x < 3; x++)
                corner[e][x] = face[x] + ((signs[e][0] * ha) * a[x] + (signs[e][1] * hc) * c[x]);
            float rel[3] = {corner[e][0] - refCentre[0], corner[e][1] - refCentre[1], corner[e][2] - refCentre[2]};
            cu[e] = dot(rel, u);
            cw[e] = dot(rel, w);
            depth[e] = hn - dot(rel, n);
            bool inside = fabsf(cu[e]) < hu * slack1 && fabsf(cw[e]) < hw * slack1;
            keep(inside, corner[e], depth[e]);
        }
        for (int e = 0; e < 4; e++) {
            int next = (e + 1) % 4;
            for (int line = 0; line < 4; line++) {
                bool acrossU = line < 2;
                float side = line % 2 == 0 ? 1.0f : -1.0f;
                float from = acrossU ? cu[e] : cw[e];
                float to = acrossU ? cu[next] : cw[next];
                float alongFrom = acrossU ? cw[e] : cu[e];
                float alongTo = acrossU ? cw[next] : cu[next];
                float limit = acrossU ? hu : hw;
                float bound = acrossU ? hw : hu;
                float span = to - from;
                if (!(fabsf(span) > SAME)) continue;
                float s = (side * limit - from) / span;
                float along = alongFrom + s * (alongTo - alongFrom);
                bool within = (s > -SLACK && s < slack1) && fabsf(along) < bound * slack1;
                float point[3];
                for (int x = 0; x < 3; x++) point[x] = corner[e][x] + s * (corner[next][x] - corner[e][x]);
                float deep = depth[e] + s * (depth[next] - depth[e]);
                keep(within, point, deep);
            }
        }
        float off[3] = {face[0] - refCentre[0], face[1] - refCentre[1], face[2] - refCentre[2]};
        float fu = dot(off, u), fw = dot(off, w);
        float au = ha * dot(a, u), aw = ha * dot(a, w), bu = hc * dot(c, u), bw = hc * dot(c, w);
        float det = au * bw - aw * bu;
        if (fabsf(det) > SAME) {
            for (int e = 0; e < 4; e++) {
                float ru = signs[e][0] * hu - fu;
                float rw = signs[e][1] * hw - fw;
                float alpha = (ru * bw - rw * bu) / det;
                float beta = (au * rw - aw * ru) / det;
                bool inside = fabsf(alpha) < slack1 && fabsf(beta) < slack1;
                float point[3];
                for (int x = 0; x < 3; x++) point[x] = face[x] + ((alpha * ha) * a[x] + (beta * hc) * c[x]);
                float rel[3] = {point[0] - refCentre[0], point[1] - refCentre[1], point[2] - refCentre[2]};
                keep(inside, point, hn - dot(rel, n));
            }
        }
        float normal[3];
        for (int x = 0; x < 3; x++) normal[x] = n[x] * flip;
        write_normal(m, normal, axis);
        int found = count;
        outCount = 0;
        if (found > 0) reduce(found, m, outCount);
    }

    __device__ void emit(int at, float *m, int &outCount) {  // :781-799
        const float *incident = cand[at];
        float deep = cand[at][3];
        float onFirstIsRef = 0.5f * (flip + 1.0f);
        float pf[3], ps[3];
        for (int x = 0; x < 3; x++) {
            float lift = deep * n[x];
            pf[x] = incident[x] + onFirstIsRef * lift;
            ps[x] = incident[x] + (1.0f - onFirstIsRef) * lift;
        }
        float of[3], os[3];
        first.own(pf, of);
        second.own(ps, os);
        write_point(m, POINTS + outCount * POINT_WORDS, of, os, deep);
        outCount++;
    }

    __device__ void reduce(int found, float *m, int &outCount) {  // :728-778
        int deepest = 0;
        float most = -1e30f;
        for (int k = 0; k < found; k++) {
            float deep = cand[k][3];
            if (deep > most) {
                most = deep;
                deepest = k;
            }
        }
        const float *pa = cand[deepest];
        int farthest = -1;
        float far = SAME;
        for (int k = 0; k < found; k++) {
            float gap[3] = {cand[k][0] - pa[0], cand[k][1] - pa[1], cand[k][2] - pa[2]};
            float d2 = dot(gap, gap);
            if (d2 > far) {
                far = d2;
                farthest = k;
            }
        }
        emit(deepest, m, outCount);
        if (farthest > -1) {
            const float *pb = cand[farthest];
            float ab[3] = {pb[0] - pa[0], pb[1] - pa[1], pb[2] - pa[2]};
            int left = -1, right = -1;
            float mostLeft = OFF_LINE * far;
            float mostRight = -mostLeft;
            for (int k = 0; k < found; k++) {
                float aq[3] = {cand[k][0] - pa[0], cand[k][1] - pa[1], cand[k][2] - pa[2]};
                float cr[3];
                cross(ab, aq, cr);
                float area = dot(cr, n);
                if (area > mostLeft) {
                    mostLeft = area;
                    left = k;
                }
                if (area < mostRight) {
                    mostRight = area;
                    right = k;
                }
            }
            emit(farthest, m, outCount);
            if (left > -1) emit(left, m, outCount);
            if (right > -1) emit(right, m, outCount);
        }
    }
};

// BoxContact.nearest, :336-385.
__device__ float nearest(const Solid &box, const float *p, float *towards, float *surface) {
    float off[3] = {p[0] - box.centre[0], p[1] - box.centre[1], p[2] - box.centre[2]};
    float local[3], held[3], outside[3];
    for (int k = 0; k < 3; k++) {
        local[k] = dot(off, box.axes[k]);
        held[k] = mx(-box.half[k], mn(box.half[k], local[k]));
    }
    for (int k = 0; k < 3; k++) outside[k] = local[k] - held[k];
    float away = sqrtf(dot(outside, outside));
    float reach = away;
    if (away > 1e-9f) {
        for (int a = 0; a < 3; a++) {
            float dir = 0, at = box.centre[a];
            for (int k = 0; k < 3; k++) {
                dir = dir + outside[k] * box.axes[k][a];
                at = at + held[k] * box.axes[k][a];
            }
            towards[a] = dir / away;
            surface[a] = at;
        }
    } else {
        float least = 1e30f;
        for (int k = 0; k < 3; k++) {
            float inward = box.half[k] - fabsf(local[k]);
            if (inward < least) {
                least = inward;
                float side = sign_of(local[k]);
                for (int a = 0; a < 3; a++) {
                    towards[a] = side * box.axes[k][a];
                    float at = box.centre[a];
                    for (int j = 0; j < 3; j++) {
                        float along = j == k ? side * box.half[k] : local[j];
                        at = at + along * box.axes[j][a];
                    }
                    surface[a] = at;
                }
            }
        }
        reach = -least;
    }
    return reach;
}

// BoxContact.sphereContact, :267-329.
__device__ void sphere_contact(const Solid &first, const Solid &second, float shapeA, float shapeB, float margin, float *m) {
    float n[3] = {0, 0, 0}, onFirst[3] = {0, 0, 0}, onSecond[3] = {0, 0, 0};
    float depth = 0, axis = (float)SPHERE_BOX;
    bool sphereA = shapeA > 0.5f, sphereB = shapeB > 0.5f;
    if (sphereA && sphereB) {
        float d[3] = {second.centre[0] - first.centre[0], second.centre[1] - first.centre[1],
                      second.centre[2] - first.centre[2]};
        float length = sqrtf(dot(d, d));
        n[1] = 1;
        if (length > 1e-9f)
            for (int a = 0; a < 3; a++) n[a] = d[a] / length;
        depth = (first.half[0] + second.half[0]) - length;
        for (int a = 0; a < 3; a++) {
            onFirst[a] = first.centre[a] + first.half[0] * n[a];
            onSecond[a] = second.centre[a] - second.half[0] * n[a];
        }
        axis = (float)SPHERE_SPHERE;
    }
    for (int which = 0; which < 2; which++) {
        bool sphereFirst = which == 0;
        const Solid &round = sphereFirst ? first : second;
        const Solid &box = sphereFirst ? second : first;
        bool only = sphereFirst ? (sphereA && !sphereB) : (sphereB && !sphereA);
        if (!only) continue;
        float towards[3] = {0, 0, 0}, surface[3] = {0, 0, 0};
        float reach = nearest(box, round.centre, towards, surface);
        float radius = round.half[0];
        depth = radius - reach;
        for (int a = 0; a < 3; a++) {
            float onSphere = round.centre[a] - radius * towards[a];
            if (sphereFirst) {
                n[a] = -towards[a];
                onFirst[a] = onSphere;
                onSecond[a] = surface[a];
            } else {
                n[a] = towards[a];
                onFirst[a] = surface[a];
                onSecond[a] = onSphere;
            }
        }
    }
    write_normal(m, n, axis);
    float kept = 0;
    if (depth > -margin) {
        float of[3], os[3];
        first.own(onFirst, of);
        second.own(onSecond, os);
        write_point(m, POINTS, of, os, depth);
        kept = 1;
    }
    m[COUNT] = kept;
}

}  // namespace vrbox
