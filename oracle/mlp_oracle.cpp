// mlp_oracle.cpp — TEST INFRASTRUCTURE. The learned warm start's network in f64 on the CPU: the checker of the HIP
// network's inference and gradient (vexelray-sim-rigid_amd/csrc/mlp.hip), behind the oracle's C ABI (vro_mlp_*).
//
// The network has no counterpart in the reference. Same parameters and layout: W1[16][64], b1[64], W2[64][64],
// b2[64], W3[64][64], b3[64], W4[64][3], b4[3]; h_l = silu(W_l' h_(l-1) + b_l), y = W4' h3 + b4. Weights are used
// rounded to bf16, as the HIP path uses them. Inference also rounds each layer's input to bf16 (the HIP path feeds
// the matrix cores bf16); the gradient keeps activations in f64 (the HIP path trains in f32). Backpropagation is
// the textbook chain rule, pinned by the tests against central finite differences.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <algorithm>
#include <random>
#include <string>
#include <vector>

#include "../include/vexelray_rigid.h"

namespace {

constexpr int F = 16, HID = 64, OUT = 3;
constexpr int W1 = 0, B1 = W1 + F * HID, W2 = B1 + HID, B2 = W2 + HID * HID, W3 = B2 + HID, B3 = W3 + HID * HID,
              W4 = B3 + HID, B4 = W4 + HID * OUT, NPARAMS = B4 + OUT;

double bf16(double v) {  // f32, then round to nearest even bf16
    float f = (float)v;
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7F800000u) == 0x7F800000u) return v;
    u += 0x7FFFu + ((u >> 16) & 1u);
    u &= 0xFFFF0000u;
    std::memcpy(&f, &u, 4);
    return f;
}

double sig(double x) { return 1.0 / (1.0 + std::exp(-x)); }

struct Net {
    std::vector<float> p = std::vector<float>(NPARAMS, 0.f);
    bool exact = false;  // weights as they are, unrounded (for checking the gradient by finite differences)
    double w(int i) const { return exact ? (double)p[i] : bf16(p[i]); }

    // One point's forward pass; `round` feeds each layer bf16 inputs. z/h: the three hidden layers.
    void forward(const float *x, bool round, double z[3][HID], double h[3][HID], double *y) const {
        double in[HID];
        int kDim = F;
        for (int k = 0; k < F; k++) in[k] = round ? bf16(x[k]) : x[k];
        const int offW[3] = {W1, W2, W3}, offB[3] = {B1, B2, B3};
        for (int l = 0; l < 3; l++) {
            for (int o = 0; o < HID; o++) {
                double s = p[offB[l] + o];
                for (int k = 0; k < kDim; k++) s += in[k] * w(offW[l] + k * HID + o);
                z[l][o] = s;
                h[l][o] = s * sig(s);
            }
            for (int o = 0; o < HID; o++) in[o] = round ? bf16(h[l][o]) : h[l][o];
            kDim = HID;
        }
        for (int o = 0; o < OUT; o++) {
            double s = p[B4 + o];
            for (int k = 0; k < HID; k++) s += in[k] * w(W4 + k * OUT + o);
            y[o] = s;
        }
    }

    // Mean squared error over n points and its gradient (chain rule, f64 activations).
    double gradient(const float *x, const float *t, int64_t n, std::vector<double> &g) const {
        g.assign(NPARAMS, 0.0);
        double loss = 0;
        for (int64_t c = 0; c < n; c++) {
            double z[3][HID], h[3][HID], y[OUT];
            forward(x + c * F, false, z, h, y);
            double dy[OUT];
            for (int o = 0; o < OUT; o++) {
                double e = y[o] - t[c * OUT + o];
                loss += e * e / n;
                dy[o] = 2 * e / n;
            }
            double d[HID], below[HID];
            for (int o = 0; o < OUT; o++) {
                g[B4 + o] += dy[o];
                for (int k = 0; k < HID; k++) g[W4 + k * OUT + o] += h[2][k] * dy[o];
            }
            for (int k = 0; k < HID; k++) {
                double s = 0;
                for (int o = 0; o < OUT; o++) s += dy[o] * w(W4 + k * OUT + o);
                double sg = sig(z[2][k]);
                d[k] = s * sg * (1 + z[2][k] * (1 - sg));
            }
            const int offW[3] = {W1, W2, W3}, offB[3] = {B1, B2, B3};
            for (int l = 2; l >= 0; l--) {
                int kDim = l == 0 ? F : HID;
                for (int o = 0; o < HID; o++) {
                    g[offB[l] + o] += d[o];
                    for (int k = 0; k < kDim; k++)
                        g[offW[l] + k * HID + o] += (l == 0 ? (double)x[c * F + k] : h[l - 1][k]) * d[o];
                }
                if (l == 0) break;
                for (int k = 0; k < HID; k++) {
                    double s = 0;
                    for (int o = 0; o < HID; o++) s += d[o] * w(offW[l] + k * HID + o);
                    double sg = sig(z[l - 1][k]);
                    below[k] = s * sg * (1 + z[l - 1][k] * (1 - sg));
                }
                std::memcpy(d, below, sizeof(d));
            }
        }
        return loss;
    }
};

}  // namespace

struct vro_mlp {
    Net net;
    std::vector<double> m1 = std::vector<double>(NPARAMS, 0.0), m2 = std::vector<double>(NPARAMS, 0.0);
    int steps = 0;
};

extern "C" {

int vro_mlp_create(uint64_t seed, vro_mlp **out) {
    if (!out) return VR_ERR_ARGUMENT;
    auto *m = new vro_mlp();
    std::mt19937_64 rng(seed);
    auto fill = [&](int off, int fanIn, int count) {
        float bound = std::sqrt(3.0f / (float)fanIn);
        std::uniform_real_distribution<float> u(-bound, bound);
        for (int i = 0; i < count; i++) m->net.p[off + i] = u(rng);
    };
    fill(W1, F, F * HID);
    fill(W2, HID, HID * HID);
    fill(W3, HID, HID * HID);
    fill(W4, HID, HID * OUT);
    *out = m;
    return VR_OK;
}

int vro_mlp_destroy(vro_mlp *m) {
    delete m;
    return VR_OK;
}

int vro_mlp_read(vro_mlp *m, float *params) {
    if (!m || !params) return VR_ERR_ARGUMENT;
    std::memcpy(params, m->net.p.data(), sizeof(float) * NPARAMS);
    return VR_OK;
}

int vro_mlp_write(vro_mlp *m, const float *params) {
    if (!m || !params) return VR_ERR_ARGUMENT;
    std::memcpy(m->net.p.data(), params, sizeof(float) * NPARAMS);
    return VR_OK;
}

int vro_mlp_forward(vro_mlp *m, const float *x, int64_t n, float *y) {
    if (!m || n < 0) return VR_ERR_ARGUMENT;
    for (int64_t c = 0; c < n; c++) {
        double z[3][HID], h[3][HID], out[OUT];
        m->net.forward(x + c * F, true, z, h, out);
        for (int o = 0; o < OUT; o++) y[c * OUT + o] = (float)out[o];
    }
    return VR_OK;
}

// As vro_mlp_gradient, in f64 (the double-precision gradient the tests compare against).
int vro_mlp_gradient64(vro_mlp *m, const float *x, const float *t, int64_t n, double *grad, double *loss) {
    if (!m || n < 1) return VR_ERR_ARGUMENT;
    std::vector<double> g;
    *loss = m->net.gradient(x, t, n, g);
    std::memcpy(grad, g.data(), sizeof(double) * NPARAMS);
    return VR_OK;
}

int vro_mlp_gradient(vro_mlp *m, const float *x, const float *t, int64_t n, float *grad, float *loss) {
    if (!m || n < 1 || !grad || !loss) return VR_ERR_ARGUMENT;
    std::vector<double> g;
    *loss = (float)m->net.gradient(x, t, n, g);
    for (int i = 0; i < NPARAMS; i++) grad[i] = (float)g[i];
    return VR_OK;
}

int vro_mlp_train(vro_mlp *m, const float *x, const float *t, int64_t n, int32_t epochs, int32_t batch, float rate,
                  uint64_t seed, float *loss) {
    if (!m || n < 1 || epochs < 1 || batch < 1 || !(rate > 0)) return VR_ERR_ARGUMENT;
    std::mt19937_64 rng(seed);
    std::vector<int64_t> perm(n);
    std::vector<float> bx, bt;
    std::vector<double> g;
    double last = 0;
    for (int e = 0; e < epochs; e++) {
        for (int64_t i = 0; i < n; i++) perm[i] = i;
        std::shuffle(perm.begin(), perm.end(), rng);
        double sum = 0;
        int64_t batches = 0;
        for (int64_t from = 0; from < n; from += batch, batches++) {
            int64_t count = std::min<int64_t>(batch, n - from);
            bx.resize(count * F);
            bt.resize(count * OUT);
            for (int64_t c = 0; c < count; c++) {
                std::memcpy(&bx[c * F], x + perm[from + c] * F, sizeof(float) * F);
                std::memcpy(&bt[c * OUT], t + perm[from + c] * OUT, sizeof(float) * OUT);
            }
            sum += m->net.gradient(bx.data(), bt.data(), count, g);
            m->steps++;
            double c1 = 1 - std::pow(0.9, m->steps), c2 = 1 - std::pow(0.999, m->steps);
            for (int i = 0; i < NPARAMS; i++) {
                m->m1[i] = 0.9 * m->m1[i] + 0.1 * g[i];
                m->m2[i] = 0.999 * m->m2[i] + 0.001 * g[i] * g[i];
                m->net.p[i] -= (float)(rate * (m->m1[i] / c1) / (std::sqrt(m->m2[i] / c2) + 1e-8));
            }
        }
        last = sum / batches;
    }
    if (loss) *loss = (float)last;
    return VR_OK;
}

// Unrounded weights on or off: the gradient is then that of the f32 parameters themselves.
int vro_mlp_exact(vro_mlp *m, int32_t on) {
    if (!m) return VR_ERR_ARGUMENT;
    m->net.exact = on != 0;
    return VR_OK;
}

// The mean squared error alone, in f64.
double vro_mlp_loss64(vro_mlp *m, const float *x, const float *t, int64_t n) {
    double loss = 0;
    for (int64_t c = 0; c < n; c++) {
        double z[3][HID], h[3][HID], y[OUT];
        m->net.forward(x + c * F, false, z, h, y);
        for (int o = 0; o < OUT; o++) loss += (y[o] - t[c * OUT + o]) * (y[o] - t[c * OUT + o]) / n;
    }
    return loss;
}

int vro_mlp_last_error(char *buf, int32_t len) {
    if (buf && len > 0) buf[0] = 0;
    return VR_OK;
}

}  // extern "C"
