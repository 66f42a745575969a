// box_oracle.cpp — CPU oracle for the box step. TEST INFRASTRUCTURE ONLY.
//
// A sequential restatement, pass by pass and in the reference's order, of vexelray-sim-rigid's box world
// (vexelray-sim-rigid-core/src/main/java/dev/vexelray/sim/rigid/box/, cited as Boxes.java, BoxContact.java,
// BoxBroad.java, BoxPairs.java, BoxStep.java, BoxStepper.java): the slot broad phase, the separating-axis narrow
// phase with its clipped and reduced manifold, the carry of impulses from step to step, the host's greedy colouring
// into rounds, and Box2D v3's soft step at the level of velocities. It is the checker the HIP path is compared
// against; nothing in the product links it.
//
// f32 arithmetic follows the reference's expression trees, operation by operation (a box dot product is
// (x0 y0 + x1 y1) + x2 y2, a sphere's x0 y0 + (x1 y1 + x2 y2)), compiled with -ffp-contract=off. The grid sort is
// a stable counting sort, as in sphere_oracle.cpp; the reference's within-cell order is not pinned, and only the
// order in which a box with more than 32 partners meets them depends on it.

#include "box_oracle.h"

#include <map>
#include <vector>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>

namespace vro {

namespace {

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

inline float mx(float a, float b) { return a < b ? b : a; }  // Math.max / std::max
inline float mn(float a, float b) { return b < a ? b : a; }  // Math.min / std::min

// Boxes.dot / BoxContact.dot: (x0 y0 + x1 y1) + x2 y2.
inline float dot(const float *x, const float *y) { return (x[0] * y[0] + x[1] * y[1]) + x[2] * y[2]; }

inline void cross(const float *x, const float *y, float *z) {
    float a = x[1] * y[2] - x[2] * y[1];
    float b = x[2] * y[0] - x[0] * y[2];
    float c = x[0] * y[1] - x[1] * y[0];
    z[0] = a;
    z[1] = b;
    z[2] = c;
}

// Boxes.rotate, Boxes.java:907-921: p turned by q, p + w t + q x t with t = 2 q x p / |q|^2.
inline void rotate(const float *q, const float *p, float *out) {
    float twice = 2.0f / ((q[0] * q[0] + q[1] * q[1]) + (q[2] * q[2] + q[3] * q[3]));
    float t[3], qt[3];
    cross(q, p, t);
    for (int a = 0; a < 3; a++) t[a] = twice * t[a];
    cross(q, t, qt);
    float r[3];
    for (int a = 0; a < 3; a++) r[a] = p[a] + (q[3] * t[a] + qt[a]);
    for (int a = 0; a < 3; a++) out[a] = r[a];
}

inline void conjugate(const float *q, float *out) {
    out[0] = -q[0];
    out[1] = -q[1];
    out[2] = -q[2];
    out[3] = q[3];
}

// Boxes.turned, Boxes.java:877-899.
inline void turned(const float *q, const float *omega, float dt, float *now) {
    float a[3];
    for (int k = 0; k < 3; k++) a[k] = (0.5f * dt) * omega[k];
    for (int k = 0; k < 3; k++) {
        int k1 = (k + 1) % 3, k2 = (k + 2) % 3;
        now[k] = q[k] + (a[k] * q[3] + (a[k1] * q[k2] - a[k2] * q[k1]));
    }
    now[3] = q[3] - (a[0] * q[0] + (a[1] * q[1] + a[2] * q[2]));
    float len2 = (now[0] * now[0] + now[1] * now[1]) + (now[2] * now[2] + now[3] * now[3]);
    float inverse = 1.0f / std::sqrt(mx(len2, 1e-30f));
    for (int k = 0; k < 4; k++) now[k] = now[k] * inverse;
}

}  // namespace

// ---- a body as the solve holds it (Boxes.Moving, Boxes.java:660-773) ----

struct Moving {
    int s;
    float shape, im;
    float at[3], q[4], half[3], inverse[3], vel[3], mom[3], omega[3];

    bool sphere() const { return shape > 0.5f; }

    // R I^-1 R^T x (Boxes.java:711-713).
    void inverseInertiaTimes(const float *x, float *out) const {
        float cq[4], own[3];
        conjugate(q, cq);
        rotate(cq, x, own);
        for (int a = 0; a < 3; a++) own[a] = inverse[a] * own[a];
        rotate(q, own, out);
    }

    float inverseMassAlong(const float *r, const float *n) const {  // Boxes.java:716-719
        float rn[3], t[3];
        cross(r, n, rn);
        inverseInertiaTimes(rn, t);
        return im + dot(rn, t);
    }

    void pointVelocity(const float *r, float *out) const {  // Boxes.java:722-729
        float turning[3];
        cross(omega, r, turning);
        for (int a = 0; a < 3; a++) out[a] = vel[a] + turning[a];
    }

    void apply(const float *r, const float *p) {  // Boxes.java:732-742
        for (int a = 0; a < 3; a++) vel[a] = vel[a] + im * p[a];
        float angular[3], turn[3];
        cross(r, p, angular);
        inverseInertiaTimes(angular, turn);
        for (int a = 0; a < 3; a++) {
            mom[a] = mom[a] + angular[a];
            omega[a] = omega[a] + turn[a];
        }
    }

    void turn(const float *j) {  // Boxes.java:745-751
        float spin[3];
        inverseInertiaTimes(j, spin);
        for (int a = 0; a < 3; a++) {
            mom[a] = mom[a] + j[a];
            omega[a] = omega[a] + spin[a];
        }
    }
};

namespace {

struct Bufs {
    float *x[3], *vel[3], *l[3], *q[4], *im, *half[3], *shape, *params, *lambda, *rub, *roll;
};

Moving load_moving(const Bufs &B, int s) {
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
    // Boxes.inverseInertia, Boxes.java:854-871.
    float h2[3];
    for (int a = 0; a < 3; a++) {
        float e = B.half[a][s];
        h2[a] = e * e;
    }
    for (int a = 0; a < 3; a++) m.inverse[a] = (3.0f * m.im) / (h2[(a + 1) % 3] + h2[(a + 2) % 3]);
    if (B.shape[s] > 0.5f)
        for (int a = 0; a < 3; a++) m.inverse[a] = m.im / (SPHERE_INERTIA * h2[0]);
    m.inverseInertiaTimes(m.mom, m.omega);
    return m;
}

void store_moving(const Bufs &B, const Moving &m) {
    for (int a = 0; a < 3; a++) {
        B.vel[a][m.s] = m.vel[a];
        B.l[a][m.s] = m.mom[a];
    }
}

// Boxes.Soft, Boxes.java:421-450.
struct Soft {
    float h, margin, bias, massScale, impulseScale, push, muStatic, muKinetic;
    Soft(const float *P, bool fixed) {
        h = P[H];
        margin = P[MARGIN];
        bias = P[fixed ? STATIC_BIAS : BIAS];
        massScale = P[fixed ? STATIC_MASS_SCALE : MASS_SCALE];
        impulseScale = P[fixed ? STATIC_IMPULSE_SCALE : IMPULSE_SCALE];
        push = P[PUSH_SPEED];
        muStatic = P[MU_STATIC];
        muKinetic = P[MU_KINETIC];
    }
    void stiffen(const float *P) {
        bias = P[STATIC_BIAS];
        massScale = P[STATIC_MASS_SCALE];
        impulseScale = P[STATIC_IMPULSE_SCALE];
    }
};

// The impulse p on the second at rB, and -p on the first at rA (Boxes.java:643-652).
void apply_pair(Moving *first, const float *rA, Moving &second, const float *rB, const float *p) {
    second.apply(rB, p);
    if (first) {
        float back[3] = {-p[0], -p[1], -p[2]};
        first->apply(rA, back);
    }
}

void relative(const Moving *first, const float *rA, const Moving &second, const float *rB, float *out) {
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

float mass_along(const Moving *first, const float *rA, const Moving &second, const float *rB, const float *n) {
    float k = second.inverseMassAlong(rB, n);
    return first ? k + first->inverseMassAlong(rA, n) : k;
}

// A point's residual in a solve or relax sweep (no counterpart in the reference, whose sweeps run a fixed count):
// the size of the impulse change the sweep made there, normal `dn` and tangent `rubbed` together, times the inverse
// mass along the normal, in m/s.
float point_residual(float k, float dn, const float *rubbed) { return k * std::sqrt(dn * dn + dot(rubbed, rubbed)); }

// res = the larger of res and r; NaN, once met, stays.
void keep_largest(float &res, float r) {
    if (res == res && !(r <= res)) res = r;
}

// Boxes.contact, Boxes.java:462-542: one point, one sweep. With `res`, the point's residual kept in it when larger.
void contact(Sweep sweep, float gap, const float *n, Moving *first, const float *rA, Moving &second, const float *rB,
             float *normals, size_t normalAt, float *tangents, size_t tangentAt, const Soft &soft,
             float *res = nullptr) {
    float normal = normals[normalAt];
    float tangent[3] = {tangents[tangentAt], tangents[tangentAt + 1], tangents[tangentAt + 2]};
    bool holding = normal > 0 || dot(tangent, tangent) > 0;
    if (sweep == WARM) {
        if (holding) {
            float p[3];
            for (int a = 0; a < 3; a++) p[a] = normal * n[a] + tangent[a];
            apply_pair(first, rA, second, rB, p);
        }
        return;
    }
    if (!(gap < soft.margin || holding)) return;
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
    normals[normalAt] = total;
    float p[3];
    for (int a = 0; a < 3; a++) p[a] = (total - normal) * n[a];
    apply_pair(first, rA, second, rB, p);

    float sliding[3];
    relative(first, rA, second, rB, sliding);
    float along = dot(sliding, n);
    for (int a = 0; a < 3; a++) sliding[a] = sliding[a] - along * n[a];
    float whole[3] = {tangent[0], tangent[1], tangent[2]};
    float speed = std::sqrt(dot(sliding, sliding));
    if (speed > 1e-12f) {
        float t[3];
        for (int a = 0; a < 3; a++) t[a] = sliding[a] / speed;
        float kt = mass_along(first, rA, second, rB, t);
        float stop = speed / kt;
        for (int a = 0; a < 3; a++) whole[a] = whole[a] - stop * t[a];
    }
    float size = std::sqrt(dot(whole, whole));
    if (size > soft.muStatic * total) {
        float cut = (soft.muKinetic * total) / size;
        for (int a = 0; a < 3; a++) whole[a] = whole[a] * cut;
    }
    float rubbed[3];
    for (int a = 0; a < 3; a++) {
        tangents[tangentAt + a] = whole[a];
        rubbed[a] = whole[a] - tangent[a];
    }
    apply_pair(first, rA, second, rB, rubbed);
    if (res) keep_largest(*res, point_residual(k, total - normal, rubbed));
}

void apply_angular(Moving *first, Moving &second, const float *j) {  // Boxes.java:603-612
    second.turn(j);
    if (first) {
        float back[3] = {-j[0], -j[1], -j[2]};
        first->turn(back);
    }
}

// Boxes.roll, Boxes.java:553-600.
void roll(Sweep sweep, Moving *first, Moving &second, float radius, const float *normals, size_t normalAt,
          float *rolls, size_t rollAt, const float *P) {
    float old[3] = {rolls[rollAt], rolls[rollAt + 1], rolls[rollAt + 2]};
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
    float speed = std::sqrt(dot(turning, turning));
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
    float most = (muRolling * radius) * mx(0.f, normals[normalAt]);
    float size = std::sqrt(dot(whole, whole));
    if (size > most) {
        float scale = most / size;
        for (int a = 0; a < 3; a++) whole[a] = whole[a] * scale;
    }
    float change[3];
    for (int a = 0; a < 3; a++) {
        rolls[rollAt + a] = whole[a];
        change[a] = whole[a] - old[a];
    }
    apply_angular(first, second, change);
}

// ---- the narrow phase (BoxContact.java) ----

struct Solid {  // BoxContact.Solid, :390-422
    float centre[3], axes[3][3], half[3];
    void own(const float *point, float *out) const {
        float off[3] = {point[0] - centre[0], point[1] - centre[1], point[2] - centre[2]};
        float o0 = dot(off, axes[0]), o1 = dot(off, axes[1]), o2 = dot(off, axes[2]);
        out[0] = o0;
        out[1] = o1;
        out[2] = o2;
    }
};

Solid load_solid(const Bufs &B, int s) {
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

inline float sign_of(float x) { return x < 0 ? -1.0f : 1.0f; }  // BoxContact.sign, :880-884

void write_normal(float *m, const float *normal, float axis) {
    for (int x = 0; x < 3; x++) m[x] = normal[x];
    m[AXIS] = axis;
}

void write_point(float *m, int at, const float *onFirst, const float *onSecond, float depth) {
    for (int x = 0; x < 3; x++) {
        m[at + ANCHOR_FIRST + x] = onFirst[x];
        m[at + ANCHOR_SECOND + x] = onSecond[x];
    }
    m[at + DEPTH] = depth;
}

struct Edge {  // BoxContact.Edge, :427-500
    float normal[3] = {0, 0, 0}, onFirst[3] = {0, 0, 0}, alongFirst[3] = {0, 0, 0};
    float halfFirst = 0, onSecond[3] = {0, 0, 0}, alongSecond[3] = {0, 0, 0}, halfSecond = 0, axis = 0;

    void record(const Solid &first, const Solid &second, int k, int j, const float *ax, const float *d) {
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

    void contact(const Solid &first, const Solid &second, float bestEdge, float *m) const {
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

    Face(const Solid &f, const Solid &s, float fromSecond, const float *bestFaceAxis, float mrg)
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
            float tn = std::fabs(dot(incAxes[m], n));
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

    void keep(bool when, const float *point, float depth) {  // :710-722
        float deep = depth;
        if (when && deep > -margin && count < CANDIDATES) {
            for (int x = 0; x < 3; x++) cand[count][x] = point[x];
            cand[count][3] = deep;
            count++;
        }
    }

    void contact(float *m, int &outCount) {  // :621-707
        static const float signs[4][2] = {{1, 1}, {-1, 1}, {-1, -1}, {1, -1}};
        const float slack1 = 1.0f + SLACK;
        float corner[4][3], cu[4], cw[4], depth[4];
        for (int e = 0; e < 4; e++) {
            for (int x = 0; x < 3; x++)
                corner[e][x] = face[x] + ((signs[e][0] * ha) * a[x] + (signs[e][1] * hc) * c[x]);
            float rel[3] = {corner[e][0] - refCentre[0], corner[e][1] - refCentre[1], corner[e][2] - refCentre[2]};
            cu[e] = dot(rel, u);
            cw[e] = dot(rel, w);
            depth[e] = hn - dot(rel, n);
            bool inside = std::fabs(cu[e]) < hu * slack1 && std::fabs(cw[e]) < hw * slack1;
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
                if (!(std::fabs(span) > SAME)) continue;
                float s = (side * limit - from) / span;
                float along = alongFrom + s * (alongTo - alongFrom);
                bool within = (s > -SLACK && s < slack1) && std::fabs(along) < bound * slack1;
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
        if (std::fabs(det) > SAME) {
            for (int e = 0; e < 4; e++) {
                float ru = signs[e][0] * hu - fu;
                float rw = signs[e][1] * hw - fw;
                float alpha = (ru * bw - rw * bu) / det;
                float beta = (au * rw - aw * ru) / det;
                bool inside = std::fabs(alpha) < slack1 && std::fabs(beta) < slack1;
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

    void emit(int at, float *m, int &outCount) {  // :781-799
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

    void reduce(int found, float *m, int &outCount) {  // :728-778
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
float nearest(const Solid &box, const float *p, float *towards, float *surface) {
    float off[3] = {p[0] - box.centre[0], p[1] - box.centre[1], p[2] - box.centre[2]};
    float local[3], held[3], outside[3];
    for (int k = 0; k < 3; k++) {
        local[k] = dot(off, box.axes[k]);
        held[k] = mx(-box.half[k], mn(box.half[k], local[k]));
    }
    for (int k = 0; k < 3; k++) outside[k] = local[k] - held[k];
    float away = std::sqrt(dot(outside, outside));
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
            float inward = box.half[k] - std::fabs(local[k]);
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
void sphere_contact(const Solid &first, const Solid &second, float shapeA, float shapeB, float margin, float *m) {
    float n[3] = {0, 0, 0}, onFirst[3] = {0, 0, 0}, onSecond[3] = {0, 0, 0};
    float depth = 0, axis = (float)SPHERE_BOX;
    bool sphereA = shapeA > 0.5f, sphereB = shapeB > 0.5f;
    if (sphereA && sphereB) {
        float d[3] = {second.centre[0] - first.centre[0], second.centre[1] - first.centre[1],
                      second.centre[2] - first.centre[2]};
        float length = std::sqrt(dot(d, d));
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

}  // namespace


struct BoxWorld {
    vr_step_desc d{};
    int n = 0, slots = 0, width = 1, cells = 0;
    bool grid = false;
    std::map<std::string, std::vector<float>> F;
    std::map<std::string, std::vector<int32_t>> I;

    float *f(const char *name) { return F.at(name).data(); }
    int32_t *i(const char *name) { return I.at(name).data(); }

    Bufs bufs() {
        Bufs B;
        B.x[0] = f("x"); B.x[1] = f("y"); B.x[2] = f("z");
        B.vel[0] = f("u"); B.vel[1] = f("v"); B.vel[2] = f("w");
        B.l[0] = f("lx"); B.l[1] = f("ly"); B.l[2] = f("lz");
        B.q[0] = f("qx"); B.q[1] = f("qy"); B.q[2] = f("qz"); B.q[3] = f("qw");
        B.im = f("im");
        B.half[0] = f("hx"); B.half[1] = f("hy"); B.half[2] = f("hz");
        B.shape = f("shape");
        B.params = f("params");
        B.lambda = f("lambda");
        B.rub = f("rub");
        B.roll = f("roll");
        return B;
    }

    // BoxStep's constructor (BoxStep.java:81-162).
    void build(const vr_step_desc &desc) {
        d = desc;
        if (d.bodies < 1 || d.substeps < 1 || d.sweeps < 1)
            throw std::invalid_argument("a step needs boxes, substeps and sweeps, got " + std::to_string(d.bodies) +
                                        ", " + std::to_string(d.substeps) + " and " + std::to_string(d.sweeps));
        n = d.bodies;
        grid = n > 1 && d.has_grid;
        if (grid && (d.nx < 1 || d.ny < 1 || d.nz < 1 || !(d.cell > 0)))
            throw std::invalid_argument("a grid of at least one cell, of positive size");
        slots = grid ? n * SLOTS : 0;
        width = std::max(1, n / 2);
        static const char *box[] = {"x",  "y",  "z",  "u",  "v",  "w",  "lx", "ly", "lz",
                                    "qx", "qy", "qz", "qw", "im", "hx", "hy", "hz", "shape"};
        for (const char *name : box) F[name].assign(n, 0.f);
        F["params"].assign(PARAM_COUNT, 0.f);
        F["lambda"].assign((size_t)LAMBDA_WORDS * n, 0.f);
        F["rub"].assign((size_t)RUB_WORDS * n, 0.f);
        F["roll"].assign((size_t)RUB_WORDS * n, 0.f);
        F["shown"].assign((size_t)SHOWN_STRIDE * n, 0.f);
        if (grid) {
            cells = d.nx * d.ny * d.nz;
            for (const char *s : {"keys", "ranks", "order"}) I[s].assign(n, 0);
            I["counts"].assign(cells + 1, 0);
            I["starts"].assign(cells + 1, 0);
            I["pairs"].assign((size_t)2 * slots, 0);
            I["pairsWas"].assign((size_t)2 * slots, 0);
            F["manifold"].assign((size_t)MANIFOLD_WORDS * slots, 0.f);
            F["manifoldWas"].assign((size_t)MANIFOLD_WORDS * slots, 0.f);
            F["impulses"].assign((size_t)IMPULSE_WORDS * slots, 0.f);
            F["impulsesWas"].assign((size_t)IMPULSE_WORDS * slots, 0.f);
            I["contacts"].assign((size_t)2 * slots + n, 0);
            I["rounds"].assign((size_t)MOST_ROUNDS * width, 0);
        }
        clear();
    }

    void clear() {  // PassRunner.clear(), then BoxStep.constants(): qw = 1 (BoxStep.java:103-105)
        for (auto &kv : F) std::fill(kv.second.begin(), kv.second.end(), 0.f);
        for (auto &kv : I) std::fill(kv.second.begin(), kv.second.end(), 0);
        std::fill(F["qw"].begin(), F["qw"].end(), 1.f);
    }

    // ---- moving (Boxes.java:276-342) ----

    void velocities() {
        Bufs B = bufs();
        const float *P = B.params;
        float h = P[H], g[3] = {P[GX], P[GY], P[GZ]};
        for (int s = 0; s < n; s++) {
            if (!(B.im[s] > 0)) continue;
            for (int a = 0; a < 3; a++) B.vel[a][s] = B.vel[a][s] + h * g[a];
        }
    }

    void positions() {
        Bufs B = bufs();
        float h = B.params[H];
        for (int s = 0; s < n; s++) {
            float im = B.im[s];
            if (!(im > 0)) continue;
            for (int a = 0; a < 3; a++) B.x[a][s] = B.x[a][s] + h * B.vel[a][s];
            Moving m = load_moving(B, s);  // q, L and the inverse inertia, as the pass loads them
            float cq[4], own[3], omega[3];
            conjugate(m.q, cq);
            rotate(cq, m.mom, own);
            for (int a = 0; a < 3; a++) own[a] = m.inverse[a] * own[a];
            rotate(m.q, own, omega);
            float halfH = 0.5f * h;
            float mid[4], omegaMid[3], end[4];
            turned(m.q, omega, halfH, mid);
            float cm[4];
            conjugate(mid, cm);
            rotate(cm, m.mom, own);
            for (int a = 0; a < 3; a++) own[a] = m.inverse[a] * own[a];
            rotate(mid, own, omegaMid);
            turned(m.q, omegaMid, h, end);
            for (int k = 0; k < 4; k++) B.q[k][s] = end[k];
        }
    }

    // Boxes.walls, Boxes.java:357-416: every body's corners (a sphere's nearest point) against the six walls.
    void walls(Sweep sweep, bool reverse, float *res = nullptr) {
        Bufs B = bufs();
        const float *P = B.params;
        float extent[3] = {P[SX], P[SY], P[SZ]};
        for (int s = 0; s < n; s++) {
            if (!(B.im[s] > 0)) continue;
            Moving box = load_moving(B, s);
            Soft soft(P, true);
            size_t base = (size_t)s * LAMBDA_WORDS;
            for (int order = 0; order < WALLS; order++) {
                int wall = reverse ? WALLS - 1 - order : order;
                int axis = wall / 2;
                bool high = wall % 2 == 1;
                int firstSlot = wall * CORNERS;
                float nrm[3];
                for (int a = 0; a < 3; a++) nrm[a] = a == axis ? (high ? -1.0f : 1.0f) : 0.0f;
                if (box.sphere()) {
                    float r[3];
                    for (int a = 0; a < 3; a++) r[a] = -(box.half[0] * nrm[a]);
                    size_t at = base + firstSlot;
                    size_t tangentAt = at * 3;
                    float reach = box.at[axis] + r[axis];
                    float gap = high ? extent[axis] - reach : reach;
                    contact(sweep, gap, nrm, nullptr, nullptr, box, r, B.lambda, at, B.rub, tangentAt, soft, res);
                    roll(sweep, nullptr, box, box.half[0], B.lambda, at, B.roll, tangentAt, P);
                } else {
                    for (int k = 0; k < CORNERS; k++) {
                        int c = reverse ? CORNERS - 1 - k : k;
                        float offset[3], r[3];
                        for (int a = 0; a < 3; a++) {
                            int bit = (c / (1 << a)) % 2;
                            offset[a] = box.half[a] * (2.0f * (float)bit - 1.0f);
                        }
                        rotate(box.q, offset, r);
                        size_t at = base + firstSlot + c;
                        size_t tangentAt = at * 3;
                        float reach = box.at[axis] + r[axis];
                        float gap = high ? extent[axis] - reach : reach;
                        contact(sweep, gap, nrm, nullptr, nullptr, box, r, B.lambda, at, B.rub, tangentAt, soft,
                                res);
                    }
                }
            }
            store_moving(B, box);
        }
    }

    // BoxPairs.solve, BoxPairs.java:64-151: every pair of round r, one sweep over its points.
    void solve_round(Sweep sweep, bool reverse, int r, float *res = nullptr) {
        Bufs B = bufs();
        const float *P = B.params;
        int32_t *rounds = i("rounds"), *pairs = i("pairs");
        float *manifold = f("manifold"), *impulses = f("impulses");
        for (int k = 0; k < width; k++) {
            int pair = rounds[r * width + k];
            if (pair < 0) continue;
            size_t out = (size_t)pair * MANIFOLD_WORDS;
            int count = (int)manifold[out + COUNT];
            if (!(count > 0)) continue;
            int ia = pairs[2 * (size_t)pair], ib = pairs[2 * (size_t)pair + 1];
            float imA = B.im[ia], imB = B.im[ib];
            Moving first = load_moving(B, ia), second = load_moving(B, ib);
            Soft soft(P, false);
            if (imA == 0 || imB == 0) soft.stiffen(P);
            float nrm[3] = {manifold[out], manifold[out + 1], manifold[out + 2]};
            size_t held = (size_t)pair * IMPULSE_WORDS;
            for (int order = 0; order < MOST_POINTS; order++) {
                int slot = reverse ? MOST_POINTS - 1 - order : order;
                if (!(slot < count)) continue;
                size_t where = out + POINTS + slot * POINT_WORDS;
                float rA[3], rB[3];
                rotate(first.q, manifold + where + ANCHOR_FIRST, rA);
                rotate(second.q, manifold + where + ANCHOR_SECOND, rB);
                if (first.sphere())
                    for (int a = 0; a < 3; a++) rA[a] = first.half[0] * nrm[a];
                if (second.sphere())
                    for (int a = 0; a < 3; a++) rB[a] = -(second.half[0] * nrm[a]);
                float gap = ((second.at[0] + rB[0]) - (first.at[0] + rA[0])) * nrm[0] +
                            ((second.at[1] + rB[1]) - (first.at[1] + rA[1])) * nrm[1];
                gap = gap + ((second.at[2] + rB[2]) - (first.at[2] + rA[2])) * nrm[2];
                size_t normalAt = held + POINT_IMPULSE * slot;
                size_t tangentAt = normalAt + 1;
                contact(sweep, gap, nrm, &first, rA, second, rB, impulses, normalAt, impulses, tangentAt, soft, res);
                if (first.sphere() || second.sphere()) {
                    float radius = second.half[0];
                    if (first.sphere()) radius = first.half[0];
                    if (first.sphere() && second.sphere())
                        radius = (first.half[0] * second.half[0]) / (first.half[0] + second.half[0]);
                    roll(sweep, &first, second, radius, impulses, normalAt, impulses, normalAt + 4, P);
                }
            }
            store_moving(B, first);
            store_moving(B, second);
        }
    }

    // BoxStep.sweep, BoxStep.java:217-231; with `res`, the largest residual of its points.
    void sweep(Sweep kind, bool reverse, int rounds, float *res = nullptr) {
        if (reverse) walls(kind, reverse, res);
        for (int order = 0; order < rounds; order++)
            solve_round(kind, reverse, reverse ? rounds - 1 - order : order, res);
        if (!reverse) walls(kind, reverse, res);
    }

    // Sweeps stopped on a residual (no counterpart in the reference): with tol at zero or more, a phase's sweeps end
    // after the first whose points' residuals are all at most tol; measured, each phase's sweeps, last residual and
    // every sweep's largest residual are kept.
    float tol = -1.f;
    bool measure = false;
    vr_box_solve lastSolve{};
    std::vector<float> curve = std::vector<float>((size_t)2 * VR_SOLVE_SUBSTEPS * VR_SOLVE_SWEEPS, 0.f);

    // The solve (or relax) sweeps of substep `sub`; returns how many ran.
    int phase(Sweep kind, int sub, int used) {
        bool stop = tol >= 0, kept = measure && sub < VR_SOLVE_SUBSTEPS;
        int at = 2 * sub + (kind == RELAX ? 1 : 0);
        float last = 0.f;
        int ran = 0;
        for (int k = 0; k < d.sweeps; k++) {
            float res = 0.f;
            sweep(kind, d.symmetric && k % 2 == 1, used, (stop || kept) ? &res : nullptr);
            ran = k + 1;
            last = res;
            if (kept && k < VR_SOLVE_SWEEPS) keep_largest(curve[(size_t)at * VR_SOLVE_SWEEPS + k], res);
            if (stop && res <= tol) break;
        }
        if (kept) {
            lastSolve.scene_sweeps[at % 2][sub] = (float)ran;
            lastSolve.residual[at % 2][sub] = last;
            lastSolve.residual_mean[at % 2][sub] = last;
        }
        if (sub < VR_SOLVE_SUBSTEPS) lastSolve.launched[at % 2][sub] = ran;
        return ran;
    }

    // ---- the broad phase (BoxBroad.java) and the collide ----

    void bin_and_sort() {  // BoxBroad.bin, :74-89, and a stable counting sort
        Bufs B = bufs();
        int32_t *keys = i("keys"), *counts = i("counts"), *starts = i("starts"), *order = i("order"),
                *ranks = i("ranks");
        int cellsAxis[3] = {d.nx, d.ny, d.nz};
        std::fill(counts, counts + cells + 1, 0);
        for (int s = 0; s < n; s++) {
            int c[3];
            for (int a = 0; a < 3; a++) {
                float fl = std::floor(B.x[a][s] / d.cell);
                c[a] = (int)mx(0.f, mn((float)(cellsAxis[a] - 1), fl));
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

    void reach(const Bufs &B, int s, float *out) {  // BoxBroad.reach, :196-227
        float x = B.q[0][s], y = B.q[1][s], z = B.q[2][s], w = B.q[3][s];
        float r[3][3] = {{1.0f - 2.0f * (y * y + z * z), 2.0f * (x * y - z * w), 2.0f * (x * z + y * w)},
                         {2.0f * (x * y + z * w), 1.0f - 2.0f * (x * x + z * z), 2.0f * (y * z - x * w)},
                         {2.0f * (x * z - y * w), 2.0f * (y * z + x * w), 1.0f - 2.0f * (x * x + y * y)}};
        float h[3] = {B.half[0][s], B.half[1][s], B.half[2][s]};
        for (int a = 0; a < 3; a++)
            out[a] = (h[0] * std::fabs(r[a][0]) + h[1] * std::fabs(r[a][1])) + h[2] * std::fabs(r[a][2]);
        if (B.shape[s] > 0.5f)
            for (int a = 0; a < 3; a++) out[a] = h[0];
    }

    void list() {  // BoxBroad.list, :103-164
        Bufs B = bufs();
        int32_t *pairs = i("pairs"), *contacts = i("contacts"), *keys = i("keys"), *starts = i("starts"),
                *order = i("order");
        float margin = B.params[MARGIN];
        for (int s = 0; s < n; s++) {
            size_t base = (size_t)s * SLOTS;
            for (int m = 0; m < SLOTS; m++) {
                pairs[(base + m) * 2] = s;
                pairs[(base + m) * 2 + 1] = -1;
            }
            float centre[3] = {B.x[0][s], B.x[1][s], B.x[2][s]};
            float rs[3];
            reach(B, s, rs);
            int found = 0;
            int key = keys[s];
            int ix = key % d.nx, iy = (key / d.nx) % d.ny, iz = key / (d.nx * d.ny);
            for (int m = 0; m < 27; m++) {
                int jx = ix + (m % 3 - 1), jy = iy + ((m / 3) % 3 - 1), jz = iz + (m / 9 - 1);
                if (jx < 0 || jx >= d.nx || jy < 0 || jy >= d.ny || jz < 0 || jz >= d.nz) continue;
                int cell = jx + d.nx * (jy + d.ny * jz);
                for (int k = starts[cell]; k < starts[cell + 1]; k++) {
                    int o = order[k];
                    if (!(o > s)) continue;
                    float other[3];
                    reach(B, o, other);
                    bool close = true;
                    for (int a = 0; a < 3; a++)
                        close = close && !(std::fabs(B.x[a][o] - centre[a]) > (rs[a] + other[a]) + margin);
                    if (!close) continue;
                    if (found < SLOTS) {
                        int place = found;
                        bool moving = true;
                        while (place > 0 && moving) {
                            int before = pairs[(base + place - 1) * 2 + 1];
                            if (before > o) {
                                pairs[(base + place) * 2 + 1] = before;
                                place--;
                            } else {
                                moving = false;
                            }
                        }
                        pairs[(base + place) * 2 + 1] = o;
                    }
                    found++;
                }
            }
            contacts[2 * (size_t)slots + s] = found;
        }
    }

    void collide_all() {  // BoxContact.collide over every slot
        int32_t *pairs = i("pairs");
        float *manifold = f("manifold");
        for (int p = 0; p < slots; p++) {
            float *m = manifold + (size_t)p * MANIFOLD_WORDS;
            int partner = pairs[2 * (size_t)p + 1];
            m[COUNT] = 0;
            if (partner < 0) continue;
            collide_pair(pairs[2 * (size_t)p], partner, m);
        }
    }

    // BoxContact.collide, :136-246: the manifold of one pair, written to m.
    void collide_pair(int ia, int ib, float *m) {
        Bufs B = bufs();
        m[COUNT] = 0;
        Solid first = load_solid(B, ia), second = load_solid(B, ib);
        float shapeA = B.shape[ia], shapeB = B.shape[ib];
        float margin = B.params[MARGIN];
        float dd[3] = {second.centre[0] - first.centre[0], second.centre[1] - first.centre[1],
                       second.centre[2] - first.centre[2]};
        float parted = -1e30f;
        float bestFace[2] = {-1e30f, -1e30f}, bestFaceAxis[2] = {0, 0};
        const Solid *boxes[2] = {&first, &second};
        for (int side = 0; side < 2; side++) {
            const Solid &own = *boxes[side], &other = *boxes[1 - side];
            for (int k = 0; k < 3; k++) {
                float rch = own.half[k];
                for (int j = 0; j < 3; j++) rch = rch + other.half[j] * std::fabs(dot(own.axes[k], other.axes[j]));
                float sep = std::fabs(dot(dd, own.axes[k])) - rch;
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
                float length = std::sqrt(length2);
                float inv = 1.0f / length;
                float ax[3] = {across[0] * inv, across[1] * inv, across[2] * inv};
                float rch = 0;
                for (int mm = 0; mm < 3; mm++)
                    rch = rch + (first.half[mm] * std::fabs(dot(first.axes[mm], ax)) +
                                 second.half[mm] * std::fabs(dot(second.axes[mm], ax)));
                float sep = std::fabs(dot(dd, ax)) - rch;
                parted = mx(parted, sep);
                if (sep > bestEdge) {
                    bestEdge = sep;
                    edge.record(first, second, k, j, ax, dd);
                }
            }
        }
        float smallest = first.half[0];
        for (const Solid *bx : boxes)
            for (int k = 0; k < 3; k++) smallest = mn(smallest, bx->half[k]);
        float slack = ABSOLUTE * smallest;
        float fromSecond = 0;
        if (bestFace[1] > RELATIVE * bestFace[0] + slack) fromSecond = 1;
        float faceSep = bestFace[0];
        if (fromSecond > 0.5f) faceSep = bestFace[1];
        float byEdge = 0;
        if (bestEdge > RELATIVE * faceSep + slack) byEdge = 1;
        int count = 0;
        float touching = 1;
        if (parted > margin) touching = 0;
        if (touching > 0.5f) {
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

    void keep() {  // BoxPairs.keep, :164-182
        F["manifoldWas"] = F["manifold"];
        F["impulsesWas"] = F["impulses"];
        I["pairsWas"] = I["pairs"];
    }

    void carry() {  // BoxPairs.carry, :191-267
        Bufs B = bufs();
        int32_t *pairs = i("pairs"), *pairsWas = i("pairsWas");
        float *manifold = f("manifold"), *impulses = f("impulses"), *was = f("manifoldWas"),
              *impulsesWas = f("impulsesWas");
        for (int p = 0; p < slots; p++) {
            size_t out = (size_t)p * MANIFOLD_WORDS, held = (size_t)p * IMPULSE_WORDS;
            int count = (int)manifold[out + COUNT];
            int partner = pairs[2 * (size_t)p + 1];
            int firstSlot = (p / SLOTS) * SLOTS;
            int wasSlot = -1;
            for (int look = 0; look < SLOTS; look++) {
                int then = pairsWas[(size_t)(firstSlot + look) * 2 + 1];
                if (then == partner && !(partner < 0)) wasSlot = firstSlot + look;
            }
            long outWas = (long)wasSlot * MANIFOLD_WORDS, heldWas = (long)wasSlot * IMPULSE_WORDS;
            int before = 0;
            if (!(wasSlot < 0)) before = (int)was[outWas + COUNT];
            int ia = pairs[2 * (size_t)p];
            float smallest = mn(B.half[0][ia], mn(B.half[1][ia], B.half[2][ia]));
            float reachSq = (SAME_POINT * smallest) * (SAME_POINT * smallest);
            for (int slot = 0; slot < MOST_POINTS; slot++) {
                float taken[POINT_IMPULSE] = {0, 0, 0, 0, 0, 0, 0};
                if (slot < count) {
                    const float *anchor = manifold + out + POINTS + slot * POINT_WORDS + ANCHOR_FIRST;
                    float nearestSq = reachSq;
                    for (int old = 0; old < MOST_POINTS; old++) {
                        if (!(old < before)) continue;
                        const float *then = was + outWas + POINTS + old * POINT_WORDS + ANCHOR_FIRST;
                        float gap[3] = {anchor[0] - then[0], anchor[1] - then[1], anchor[2] - then[2]};
                        float d2 = dot(gap, gap);
                        if (d2 < nearestSq) {
                            nearestSq = d2;
                            for (int w = 0; w < POINT_IMPULSE; w++)
                                taken[w] = impulsesWas[heldWas + POINT_IMPULSE * old + w];
                        }
                    }
                }
                if (slot == 0 && count == 1 && before == 1)
                    for (int w = 0; w < POINT_IMPULSE; w++) taken[w] = impulsesWas[heldWas + w];
                for (int w = 0; w < POINT_IMPULSE; w++) impulses[held + POINT_IMPULSE * slot + w] = taken[w];
            }
        }
    }

    void counts() {  // BoxBroad.counts, :174-188
        int32_t *pairs = i("pairs"), *contacts = i("contacts");
        float *manifold = f("manifold");
        for (int p = 0; p < slots; p++) {
            int partner = pairs[2 * (size_t)p + 1];
            int count = (int)manifold[(size_t)p * MANIFOLD_WORDS + COUNT];
            contacts[p] = count > 0 ? partner : -1;
            contacts[(size_t)slots + p] = count;
        }
    }

    void show() {  // Boxes.show, :801-820
        Bufs B = bufs();
        float *shown = f("shown");
        for (int s = 0; s < n; s++) {
            float *b = shown + (size_t)s * SHOWN_STRIDE;
            for (int a = 0; a < 3; a++) {
                b[a] = b[4 + a];
                b[4 + a] = B.x[a][s];
                b[8 + a] = B.half[a][s];
            }
            for (int k = 0; k < 4; k++) {
                b[12 + k] = b[16 + k];
                b[16 + k] = B.q[k][s];
            }
            b[3] = B.shape[s];
        }
    }

    // BoxStepper.step up to the rounds' write (BoxStepper.java:62-95): BoxStep.collide (BoxStep.java:138-153) and
    // the host's first-fit colouring; the counts into rep. Returns the rounds used.
    int prepare(vr_step_report *rep) {
        int used = 0, touching = 0, points = 0, unplaced = 0, unlisted = 0;
        if (slots > 0) {
            keep();
            bin_and_sort();
            list();
            collide_all();
            carry();
            counts();
            int32_t *contacts = i("contacts"), *rounds = i("rounds");
            std::vector<int> filled(MOST_ROUNDS, 0);
            std::vector<std::vector<char>> taken(MOST_ROUNDS, std::vector<char>(n, 0));
            std::fill(rounds, rounds + (size_t)MOST_ROUNDS * width, -1);
            for (int p = 0; p < slots; p++) {
                int b = contacts[p];
                if (b < 0) continue;
                int a = p / SLOTS;
                touching++;
                points += contacts[(size_t)slots + p];
                int r = 0;
                while (r < MOST_ROUNDS && (taken[r][a] || taken[r][b])) r++;
                if (r == MOST_ROUNDS) {
                    unplaced++;
                } else {
                    taken[r][a] = taken[r][b] = 1;
                    rounds[(size_t)r * width + filled[r]++] = p;
                    used = std::max(used, r + 1);
                }
            }
            for (int s = 0; s < n; s++) unlisted += std::max(0, contacts[2 * (size_t)slots + s] - SLOTS);
        }
        if (rep) {
            rep->rounds = used;
            rep->most_rounds = used;
            rep->touching = touching;
            rep->points = points;
            rep->unplaced = unplaced;
            rep->unlisted = unlisted;
            rep->submissions = slots > 0 ? 1 : 0;
        }
        return used;
    }

    // BoxStepper.step, BoxStepper.java:56-100, with BoxStep.collide and BoxStep.step (BoxStep.java:187-214).
    void step(vr_step_report *rep) {
        int used = prepare(rep);
        std::fill(curve.begin(), curve.end(), 0.f);
        vr_box_solve fresh{};
        fresh.tolerance = tol;
        fresh.measured = measure ? 1 : 0;
        fresh.substeps = d.substeps;
        fresh.sweeps = d.sweeps;
        fresh.scenes = 1;
        lastSolve = fresh;
        for (int sub = 0; sub < d.substeps; sub++) {
            velocities();
            sweep(WARM, false, used);
            phase(SOLVE, sub, used);
            positions();
            phase(RELAX, sub, used);
        }
        show();
        if (rep) rep->submissions = slots > 0 ? 2 : 1;  // BoxStepper.step: collide read back, then the step
    }
};

BoxWorld *box_create(const vr_step_desc &desc) {
    auto *w = new BoxWorld();
    try {
        w->build(desc);
    } catch (...) {
        delete w;
        throw;
    }
    return w;
}

void box_destroy(BoxWorld *w) { delete w; }
void box_clear(BoxWorld *w) { w->clear(); }

bool box_buffer(BoxWorld *w, const std::string &name, void **ptr, int64_t *count, int32_t *dtype) {
    if (w->F.count(name)) {
        *ptr = w->F[name].data();
        *count = (int64_t)w->F[name].size();
        *dtype = VR_F32;
        return true;
    }
    if (w->I.count(name)) {
        *ptr = w->I[name].data();
        *count = (int64_t)w->I[name].size();
        *dtype = VR_I32;
        return true;
    }
    return false;
}

void box_params(BoxWorld *w, const int32_t *words, int32_t n) {
    if (n != PARAM_COUNT) throw std::invalid_argument("box parameters are 18 words");
    std::memcpy(w->F["params"].data(), words, sizeof(int32_t) * PARAM_COUNT);
}

void box_step(BoxWorld *w, vr_step_report *rep) { w->step(rep); }

void box_collide_list(BoxWorld *w, vr_step_report *rep) { w->prepare(rep); }

void box_tolerance(BoxWorld *w, float tol, bool measure) {
    if (tol != tol) throw std::invalid_argument("a tolerance is a number (below zero for none)");
    w->tol = tol;
    w->measure = measure;
}

const vr_box_solve &box_solve(BoxWorld *w) { return w->lastSolve; }
const std::vector<float> &box_curve(BoxWorld *w) { return w->curve; }

void box_collide(BoxWorld *w, const int32_t *pairs, int32_t npairs, float *manifold) {
    if (npairs < 0) throw std::invalid_argument("a count of pairs is at least 0, got " + std::to_string(npairs));
    for (int64_t k = 0; k < 2 * (int64_t)npairs; k++)
        if (pairs[k] < 0 || pairs[k] >= w->n)
            throw std::invalid_argument("pair " + std::to_string(k / 2) + " names body " + std::to_string(pairs[k]) +
                                        ", outside 0 .. " + std::to_string(w->n - 1));
    for (int p = 0; p < npairs; p++) {
        float *m = manifold + (size_t)p * MANIFOLD_WORDS;
        std::fill(m, m + MANIFOLD_WORDS, 0.f);
        w->collide_pair(pairs[2 * p], pairs[2 * p + 1], m);
    }
}

}  // namespace vro
