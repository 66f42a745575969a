// diagnostics.hip — a world's diagnostics reduced on the device in f64: SphereDiagnostics.of
// (core/.../sphere/SphereDiagnostics.java:34-83) and BoxDiagnostics.of (core/.../box/BoxDiagnostics.java:18-44),
// which the reference computes on the host with an every-pair overlap scan. Each workgroup reduces its bodies into
// one row of partial results (sums and maxima in f64), and one workgroup reduces the rows.
//
// Spheres: the deepest overlap of two spheres over the smaller radius, found through the world's grid rebuilt at the
// current positions (complete when the cells are at least the widest diameter, as SphereGrid.covering makes them),
// or over every pair when the world has no grid; the deepest wall overlap over the radius; kinetic (with spin),
// rotational and potential energy; linear and angular momentum; the top speed; whether any state is not finite.
// Boxes: kinetic (with spin) and rotational energy, linear and angular momentum from the angular momentum state L
// and the body's inertia, the top speed, whether any state is not finite, and the deepest manifold point of the
// last collide (the step's start).
#include "diagnostics.h"

#include "box_math.h"

#include <algorithm>
#include <stdexcept>
#include <string>

namespace vrdiag {

namespace {

constexpr int THREADS = 256, SUMS = 9, MOSTS = 5;

struct Acc {
    double sum[SUMS];   // kinetic, rotational, potential, momentum x y z, angular x y z
    double most[MOSTS]; // overlap, wall, speed, broken, depth
};

__device__ __forceinline__ void acc_zero(Acc &a) {
    for (int k = 0; k < SUMS; k++) a.sum[k] = 0;
    for (int k = 0; k < MOSTS; k++) a.most[k] = 0;
}

// The workgroup's rows into partials[block * ROW ..]: sums added, maxima kept.
__device__ void block_reduce(Acc &a, double *partials) {
    constexpr int ROW = SUMS + MOSTS;
    __shared__ double sh[THREADS / 64][ROW];
    for (int o = 32; o > 0; o >>= 1) {
        for (int k = 0; k < SUMS; k++) a.sum[k] += __shfl_xor(a.sum[k], o, 64);
        for (int k = 0; k < MOSTS; k++) a.most[k] = fmax(a.most[k], __shfl_xor(a.most[k], o, 64));
    }
    int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        for (int k = 0; k < SUMS; k++) sh[wave][k] = a.sum[k];
        for (int k = 0; k < MOSTS; k++) sh[wave][SUMS + k] = a.most[k];
    }
    __syncthreads();
    if (threadIdx.x < ROW) {
        int k = threadIdx.x;
        double v = sh[0][k];
        for (int w = 1; w < THREADS / 64; w++) v = k < SUMS ? v + sh[w][k] : fmax(v, sh[w][k]);
        partials[(size_t)blockIdx.x * ROW + k] = v;
    }
}

__global__ void k_finish(const double *partials, int blocks, double *out) {
    constexpr int ROW = SUMS + MOSTS;
    int k = threadIdx.x;
    if (k >= ROW) return;
    double v = partials[k];
    for (int b = 1; b < blocks; b++) v = k < SUMS ? v + partials[(size_t)b * ROW + k] : fmax(v, partials[(size_t)b * ROW + k]);
    // out: overlap, wall, kinetic, rotational, potential, momentum 3, angular 3, speed, broken, depth
    const int place[ROW] = {D_KINETIC, D_ROTATIONAL, D_POTENTIAL, D_MOMENTUM, D_MOMENTUM + 1, D_MOMENTUM + 2,
                            D_ANGULAR, D_ANGULAR + 1, D_ANGULAR + 2, D_OVERLAP, D_WALL, D_SPEED, D_BROKEN, D_DEPTH};
    out[place[k]] = k == SUMS + 2 ? sqrt(v) : v;  // the speed is reduced squared
}

__device__ __forceinline__ bool finite3(double a, double b, double c) { return isfinite(a) && isfinite(b) && isfinite(c); }

constexpr double SPHERE_INERTIA = 0.4;  // Spheres.INERTIA, Spheres.java:246

template <bool GRID>
__global__ void k_spheres(vr::SphereDev D, double *partials) {
    Acc acc;
    acc_zero(acc);
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < D.n) {
        const float *P = D.params;
        double x = D.x[s], y = D.y[s], z = D.z[s], u = D.u[s], v = D.v[s], w = D.w[s];
        double ax = D.ax[s], ay = D.ay[s], az = D.az[s], r = D.r[s], im = D.im[s];
        bool ok = finite3(x, y, z) && finite3(u, v, w) && finite3(ax, ay, az);
        acc.most[3] = ok ? 0 : 1;
        double ext[3] = {P[vr::SX], P[vr::SY], P[vr::SZ]}, at[3] = {x, y, z};
        double wall = 0;
        for (int a = 0; a < 3; a++) wall = fmax(wall, fmax((r - at[a]) / r, (at[a] + r - ext[a]) / r));
        acc.most[1] = wall;
        double speed2 = u * u + v * v + w * w;
        acc.most[2] = speed2;
        double m = im > 0 ? 1.0 / im : 0.0;
        double inertia = SPHERE_INERTIA * m * r * r;
        double spin2 = ax * ax + ay * ay + az * az;
        acc.sum[0] = 0.5 * m * speed2 + 0.5 * inertia * spin2;
        acc.sum[1] = 0.5 * inertia * spin2;
        acc.sum[2] = m * -(double)P[vr::GY] * y;
        acc.sum[3] = m * u;
        acc.sum[4] = m * v;
        acc.sum[5] = m * w;
        acc.sum[6] = m * (y * w - z * v) + inertia * ax;
        acc.sum[7] = m * (z * u - x * w) + inertia * ay;
        acc.sum[8] = m * (x * v - y * u) + inertia * az;
        double deepest = 0;
        auto pair = [&](int o) {
            if (!(o > s)) return;
            double dx = x - D.x[o], dy = y - D.y[o], dz = z - D.z[o];
            double ro = D.r[o];
            double over = (r + ro - sqrt(dx * dx + dy * dy + dz * dz)) / fmin(r, ro);
            deepest = fmax(deepest, over);
        };
        if (GRID) {
            int key = D.keys[s];
            int ix = key % D.nx, iy = (key / D.nx) % D.ny, iz = key / (D.nx * D.ny);
            for (int mz = -1; mz <= 1; mz++)
                for (int my = -1; my <= 1; my++)
                    for (int mx = -1; mx <= 1; mx++) {
                        int jx = ix + mx, jy = iy + my, jz = iz + mz;
                        if (jx < 0 || jx >= D.nx || jy < 0 || jy >= D.ny || jz < 0 || jz >= D.nz) continue;
                        int cell = jx + D.nx * (jy + D.ny * jz);
                        for (int k = D.starts[cell]; k < D.starts[cell + 1]; k++) pair(D.order[k]);
                    }
        } else {
            for (int o = s + 1; o < D.n; o++) pair(o);
        }
        acc.most[0] = deepest;
    }
    block_reduce(acc, partials);
}

// R I^-1 R' L for a body of principal moments `moment` turned by q (Boxes.inverseInertia, Boxes.java:854-871).
__device__ void spin_of(const double *l, const double *q, const double *moment, double *omega) {
    auto rot = [](const double *q, const double *p, double *o) {  // p turned by q
        double t[3] = {2 * (q[1] * p[2] - q[2] * p[1]), 2 * (q[2] * p[0] - q[0] * p[2]), 2 * (q[0] * p[1] - q[1] * p[0])};
        double c[3] = {q[1] * t[2] - q[2] * t[1], q[2] * t[0] - q[0] * t[2], q[0] * t[1] - q[1] * t[0]};
        for (int a = 0; a < 3; a++) o[a] = p[a] + q[3] * t[a] + c[a];
    };
    double cq[4] = {-q[0], -q[1], -q[2], q[3]}, own[3], scaled[3];
    rot(cq, l, own);
    for (int a = 0; a < 3; a++) scaled[a] = moment[a] > 0 ? own[a] / moment[a] : 0;
    rot(q, scaled, omega);
}

__global__ void k_boxes(vrbox::BoxDev D, double *partials) {
    Acc acc;
    acc_zero(acc);
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < D.n) {
        double at[3], vel[3], l[3], q[4], h[3];
        for (int a = 0; a < 3; a++) {
            at[a] = D.x[a][s];
            vel[a] = D.vel[a][s];
            l[a] = D.l[a][s];
            h[a] = D.half[a][s];
        }
        for (int k = 0; k < 4; k++) q[k] = D.q[k][s];
        bool ok = finite3(at[0], at[1], at[2]) && finite3(vel[0], vel[1], vel[2]) && finite3(l[0], l[1], l[2]);
        acc.most[3] = ok ? 0 : 1;
        acc.most[2] = vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2];
        double im = D.im[s];
        if (im > 0) {
            double m = 1.0 / im;
            double moment[3];
            if (D.shape[s] > 0.5f) {  // Boxes.inertia, Boxes.java:257-268
                moment[0] = moment[1] = moment[2] = 0.4 * m * h[0] * h[0];
            } else {
                moment[0] = m * (h[1] * h[1] + h[2] * h[2]) / 3;
                moment[1] = m * (h[0] * h[0] + h[2] * h[2]) / 3;
                moment[2] = m * (h[0] * h[0] + h[1] * h[1]) / 3;
            }
            double omega[3];
            spin_of(l, q, moment, omega);
            double orbit[3] = {at[1] * vel[2] - at[2] * vel[1], at[2] * vel[0] - at[0] * vel[2],
                               at[0] * vel[1] - at[1] * vel[0]};
            double rot = 0;
            for (int a = 0; a < 3; a++) {
                acc.sum[0] += 0.5 * m * vel[a] * vel[a];
                rot += 0.5 * l[a] * omega[a];
                acc.sum[3 + a] = m * vel[a];
                acc.sum[6 + a] = m * orbit[a] + l[a];
            }
            acc.sum[0] += rot;
            acc.sum[1] = rot;
            acc.sum[2] = m * -(double)D.params[2] * at[1];  // Boxes' params: h, gx, gy, ...
        }
        double deepest = 0;
        if (D.slots > 0) {
            for (int k = 0; k < vrbox::SLOTS; k++) {
                size_t out = ((size_t)s * vrbox::SLOTS + k) * vrbox::MANIFOLD_WORDS;
                int count = (int)D.manifold[out + vrbox::COUNT];
                for (int p = 0; p < count && p < vrbox::MOST_POINTS; p++)
                    deepest = fmax(deepest, (double)D.manifold[out + vrbox::POINTS + p * vrbox::POINT_WORDS + vrbox::DEPTH]);
            }
        }
        acc.most[4] = deepest;
    }
    block_reduce(acc, partials);
}

inline void check_launch(const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace

int64_t partial_words(int n) { return (int64_t)((n + THREADS - 1) / THREADS + 1) * (SUMS + MOSTS); }

void launch_spheres(const vr::SphereDev &D, bool grid, double *partials, double *out, hipStream_t s) {
    int blocks = std::max(1, (D.n + THREADS - 1) / THREADS);
    if (grid) k_spheres<true><<<blocks, THREADS, 0, s>>>(D, partials);
    else k_spheres<false><<<blocks, THREADS, 0, s>>>(D, partials);
    k_finish<<<1, 64, 0, s>>>(partials, blocks, out);
    check_launch("sphere diagnostics");
}

void launch_boxes(const vrbox::BoxDev &D, double *partials, double *out, hipStream_t s) {
    int blocks = std::max(1, (D.n + THREADS - 1) / THREADS);
    k_boxes<<<blocks, THREADS, 0, s>>>(D, partials);
    k_finish<<<1, 64, 0, s>>>(partials, blocks, out);
    check_launch("box diagnostics");
}

}  // namespace vrdiag
