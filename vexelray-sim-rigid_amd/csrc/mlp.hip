// mlp.hip — the learned warm start's network: 16 -> 64 -> 64 -> 64 -> 3, SiLU on the hidden layers.
//
// No counterpart in the reference, whose warm starts are the previous step's impulses carried to matching points
// (core/.../box/BoxPairs.java:191-267). The network predicts a contact point's starting impulse from 16 features.
//
// Parameters (9,603 f32, the master copy training updates), in this order: W1[16][64], b1[64], W2[64][64], b2[64],
// W3[64][64], b3[64], W4[64][3], b4[3]; y = W4' h3 + b4, h_l = silu(W_l' h_(l-1) + b_l).
//
// Inference runs on the matrix cores: bf16 weights and activations, f32 accumulation. A wave takes 16 points and
// works on the transposed layers, Z' = W' X', with v_mfma_f32_16x16x16_bf16: the accumulator a lane holds after one
// layer (rows 4(lane/16)..+3 of a 16-row tile, column lane%16) is exactly the B operand the next layer's product
// needs from that lane, so activations never move between lanes; weights sit in LDS as W' tiles.
// Training runs in f32 (with the bf16-rounded weights inference uses): a workgroup takes points 32 at a time, keeps
// the layers' activations in LDS, and adds each point's gradient into its own copy of the gradient, which goes to
// the device-wide gradient once, at the end.
#include "mlp.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>

namespace vrmlp {

namespace {

typedef __attribute__((__vector_size__(4 * sizeof(short)))) short bf16x4;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

constexpr int THREADS = 256;

inline void check_launch(const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

inline int blocks_for(int64_t n) { return (int)std::max<int64_t>(1, (n + THREADS - 1) / THREADS); }

// f32 to bf16, rounding to nearest even (NaN kept a NaN).
__host__ __device__ __forceinline__ unsigned short to_bf16(float f) {
    unsigned u;
    __builtin_memcpy(&u, &f, 4);
    if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x007FFFFFu)) return (unsigned short)((u >> 16) | 0x40);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

__host__ __device__ __forceinline__ float from_bf16(unsigned short h) {
    unsigned u = (unsigned)h << 16;
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}

__device__ __forceinline__ float bf16_round(float f) { return from_bf16(to_bf16(f)); }

__device__ __forceinline__ float sigmoid(float x) { return 1.0f / (1.0f + expf(-x)); }
__device__ __forceinline__ float silu(float x) { return x * sigmoid(x); }
__device__ __forceinline__ float silu_slope(float x) {  // d/dx x s(x) = s(x) (1 + x (1 - s(x)))
    float s = sigmoid(x);
    return s * (1.0f + x * (1.0f - s));
}

// ---- inference: packed bf16 weights, W' tiles, and f32 biases (hidden ones, the output's padded to 16) ----

struct Packed {
    unsigned short w1[HIDDEN * FEATURES];  // W1' [64][16]
    unsigned short w2[HIDDEN * HIDDEN];    // W2' [64][64]
    unsigned short w3[HIDDEN * HIDDEN];    // W3' [64][64]
    unsigned short w4[16 * HIDDEN];        // W4' [16][64], rows 3.. zero
    float b[3 * HIDDEN + 16];
};

__global__ void k_pack(const float *P, Packed *out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < HIDDEN * FEATURES) {
        int o = i / FEATURES, k = i % FEATURES;
        out->w1[i] = to_bf16(P[OFF_W1 + k * HIDDEN + o]);
    }
    if (i < HIDDEN * HIDDEN) {
        int o = i / HIDDEN, k = i % HIDDEN;
        out->w2[i] = to_bf16(P[OFF_W2 + k * HIDDEN + o]);
        out->w3[i] = to_bf16(P[OFF_W3 + k * HIDDEN + o]);
    }
    if (i < 16 * HIDDEN) {
        int o = i / HIDDEN, k = i % HIDDEN;
        out->w4[i] = o < OUTPUTS ? to_bf16(P[OFF_W4 + k * OUTPUTS + o]) : 0;
    }
    if (i < HIDDEN) {
        out->b[i] = P[OFF_B1 + i];
        out->b[HIDDEN + i] = P[OFF_B2 + i];
        out->b[2 * HIDDEN + i] = P[OFF_B3 + i];
    }
    if (i < 16) out->b[3 * HIDDEN + i] = i < OUTPUTS ? P[OFF_B4 + i] : 0.f;
}

// A operand of W' tile (out rows 16t.., k columns 16kk..) for this lane: W'[16t + lane%16][16kk + 4(lane/16) + j].
__device__ __forceinline__ bf16x4 a_tile(const unsigned short *wt, int kDim, int t, int kk, int lane) {
    const unsigned short *row = wt + (16 * t + (lane & 15)) * kDim + 16 * kk + 4 * (lane >> 4);
    bf16x4 a;
    for (int j = 0; j < 4; j++) a[j] = (short)row[j];
    return a;
}

// One hidden layer over the lane's B operands (kTiles of them): 4 output tiles, bias and SiLU, back to bf16.
__device__ __forceinline__ void hidden_layer(const unsigned short *wt, int kDim, const float *bias, const bf16x4 *in,
                                             int kTiles, bf16x4 *out, int lane) {
    for (int t = 0; t < HIDDEN / 16; t++) {
        f32x4 acc = {0, 0, 0, 0};
        for (int kk = 0; kk < kTiles; kk++)
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a_tile(wt, kDim, t, kk, lane), in[kk], acc, 0, 0, 0);
        for (int j = 0; j < 4; j++) {
            float z = acc[j] + bias[16 * t + 4 * (lane >> 4) + j];
            out[t][j] = (short)to_bf16(silu(z));
        }
    }
}

__global__ void __launch_bounds__(THREADS) k_forward(const Packed *W, const float *x, int64_t n, float *y,
                                                     const int32_t *rows) {
    __shared__ Packed w;
    if (rows) n = min(n, (int64_t)*rows);  // the rows actually there, known on the device only
    if ((int64_t)blockIdx.x * (blockDim.x / 64) * 16 >= n) return;  // the whole workgroup has none
    {
        const int *src = reinterpret_cast<const int *>(W);
        int *dst = reinterpret_cast<int *>(&w);
        for (int i = threadIdx.x; i < (int)(sizeof(Packed) / 4); i += blockDim.x) dst[i] = src[i];
    }
    __syncthreads();
    int lane = threadIdx.x & 63;
    int64_t waves = (int64_t)gridDim.x * (blockDim.x / 64);
    int64_t wave = (int64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    for (int64_t base = wave * 16; base < n; base += waves * 16) {
        int64_t point = base + (lane & 15);
        bool real = point < n;
        // B operand of X': features 4(lane/16).. of this lane's point.
        bf16x4 in0;
        for (int j = 0; j < 4; j++) in0[j] = (short)to_bf16(real ? x[point * FEATURES + 4 * (lane >> 4) + j] : 0.f);
        bf16x4 h1[4], h2[4], h3[4];
        hidden_layer(w.w1, FEATURES, w.b, &in0, 1, h1, lane);
        hidden_layer(w.w2, HIDDEN, w.b + HIDDEN, h1, 4, h2, lane);
        hidden_layer(w.w3, HIDDEN, w.b + 2 * HIDDEN, h2, 4, h3, lane);
        f32x4 acc = {0, 0, 0, 0};
        for (int kk = 0; kk < 4; kk++)
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a_tile(w.w4, HIDDEN, 0, kk, lane), h3[kk], acc, 0, 0, 0);
        // Rows 0..3 of the output tile are in lanes 0..15: outputs 0..2 of point lane%16.
        if (lane < 16 && real)
            for (int j = 0; j < OUTPUTS; j++) y[point * OUTPUTS + j] = acc[j] + w.b[3 * HIDDEN + j];
    }
}

// ---- training: f32, the weights rounded to bf16 as inference uses them ----

constexpr int CHUNK = 32;

// One chunk's activations in LDS, each [CHUNK][width] row-major (z: before SiLU, h: after; d: the loss's gradient
// at z), then the workgroup's gradient, PARAMS floats.
constexpr int AT_X = 0, AT_Z1 = AT_X + CHUNK * FEATURES, AT_H1 = AT_Z1 + CHUNK * HIDDEN, AT_Z2 = AT_H1 + CHUNK * HIDDEN,
              AT_H2 = AT_Z2 + CHUNK * HIDDEN, AT_Z3 = AT_H2 + CHUNK * HIDDEN, AT_H3 = AT_Z3 + CHUNK * HIDDEN,
              AT_DY = AT_H3 + CHUNK * HIDDEN, AT_D3 = AT_DY + CHUNK * OUTPUTS, AT_D2 = AT_D3 + CHUNK * HIDDEN,
              AT_D1 = AT_D2 + CHUNK * HIDDEN, AT_GRAD = AT_D1 + CHUNK * HIDDEN, SHARED_FLOATS = AT_GRAD + PARAMS;

__device__ __forceinline__ float wq(const float *P, int i) { return bf16_round(P[i]); }

// z[c][o] = b[o] + sum_k in[c][k] W[k][o]; h = silu(z), for o < HIDDEN.
__device__ void dense_forward(const float *P, int offW, int offB, int kDim, const float *in, float *z, float *h, int m) {
    for (int i = threadIdx.x; i < m * HIDDEN; i += blockDim.x) {
        int c = i / HIDDEN, o = i % HIDDEN;
        float s = P[offB + o];
        for (int k = 0; k < kDim; k++) s += in[c * kDim + k] * wq(P, offW + k * HIDDEN + o);
        z[i] = s;
        h[i] = silu(s);
    }
}

// dIn[c][k] = silu'(zIn[c][k]) * sum_o dOut[c][o] W[k][o], for k < HIDDEN.
__device__ void dense_back(const float *P, int offW, int outDim, const float *dOut, const float *zIn, float *dIn, int m) {
    for (int i = threadIdx.x; i < m * HIDDEN; i += blockDim.x) {
        int c = i / HIDDEN, k = i % HIDDEN;
        float s = 0;
        for (int o = 0; o < outDim; o++) s += dOut[c * outDim + o] * wq(P, offW + k * outDim + o);
        dIn[i] = s * silu_slope(zIn[i]);
    }
}

// g[offW + k*outDim + o] += sum_c in[c][k] d[c][o]; g[offB + o] += sum_c d[c][o].
__device__ void dense_grad(float *g, int offW, int offB, int kDim, int outDim, const float *in, const float *d, int m) {
    for (int i = threadIdx.x; i < kDim * outDim; i += blockDim.x) {
        int k = i / outDim, o = i % outDim;
        float s = 0;
        for (int c = 0; c < m; c++) s += in[c * kDim + k] * d[c * outDim + o];
        g[offW + i] += s;
    }
    for (int o = threadIdx.x; o < outDim; o += blockDim.x) {
        float s = 0;
        for (int c = 0; c < m; c++) s += d[c * outDim + o];
        g[offB + o] += s;
    }
}

// The gradient of the loss sum over points of |y - t|^2 / n (mean squared error), added into grad; the loss
// itself added into *loss. Each workgroup takes chunks of CHUNK points.
__global__ void __launch_bounds__(THREADS) k_gradient(const float *P, const float *x, const float *t, int64_t n,
                                                      float *grad, float *loss) {
    extern __shared__ float sh[];
    __shared__ float sLoss;
    float *g = sh + AT_GRAD;
    for (int i = threadIdx.x; i < PARAMS; i += blockDim.x) g[i] = 0;
    if (threadIdx.x == 0) sLoss = 0;
    float scale = 1.0f / (float)n;
    for (int64_t base = (int64_t)blockIdx.x * CHUNK; base < n; base += (int64_t)gridDim.x * CHUNK) {
        int m = n - base < CHUNK ? (int)(n - base) : CHUNK;
        __syncthreads();
        for (int i = threadIdx.x; i < m * FEATURES; i += blockDim.x) sh[AT_X + i] = x[base * FEATURES + i];
        __syncthreads();
        dense_forward(P, OFF_W1, OFF_B1, FEATURES, sh + AT_X, sh + AT_Z1, sh + AT_H1, m);
        __syncthreads();
        dense_forward(P, OFF_W2, OFF_B2, HIDDEN, sh + AT_H1, sh + AT_Z2, sh + AT_H2, m);
        __syncthreads();
        dense_forward(P, OFF_W3, OFF_B3, HIDDEN, sh + AT_H2, sh + AT_Z3, sh + AT_H3, m);
        __syncthreads();
        for (int i = threadIdx.x; i < m * OUTPUTS; i += blockDim.x) {
            int c = i / OUTPUTS, o = i % OUTPUTS;
            float s = P[OFF_B4 + o];
            for (int k = 0; k < HIDDEN; k++) s += sh[AT_H3 + c * HIDDEN + k] * wq(P, OFF_W4 + k * OUTPUTS + o);
            float e = s - t[(base + c) * OUTPUTS + o];
            sh[AT_DY + i] = 2.0f * e * scale;
            atomicAdd(&sLoss, e * e * scale);
        }
        __syncthreads();
        dense_back(P, OFF_W4, OUTPUTS, sh + AT_DY, sh + AT_Z3, sh + AT_D3, m);
        dense_grad(g, OFF_W4, OFF_B4, HIDDEN, OUTPUTS, sh + AT_H3, sh + AT_DY, m);
        __syncthreads();
        dense_back(P, OFF_W3, HIDDEN, sh + AT_D3, sh + AT_Z2, sh + AT_D2, m);
        dense_grad(g, OFF_W3, OFF_B3, HIDDEN, HIDDEN, sh + AT_H2, sh + AT_D3, m);
        __syncthreads();
        dense_back(P, OFF_W2, HIDDEN, sh + AT_D2, sh + AT_Z1, sh + AT_D1, m);
        dense_grad(g, OFF_W2, OFF_B2, HIDDEN, HIDDEN, sh + AT_H1, sh + AT_D2, m);
        __syncthreads();
        dense_grad(g, OFF_W1, OFF_B1, FEATURES, HIDDEN, sh + AT_X, sh + AT_D1, m);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < PARAMS; i += blockDim.x)
        if (g[i] != 0) atomicAdd(&grad[i], g[i]);
    if (threadIdx.x == 0) atomicAdd(loss, sLoss);
}

// Adam (Kingma and Ba, 2015), bias-corrected, on the master weights; the gradient is cleared for the next batch.
__global__ void k_adam(float *P, float *grad, float *m1, float *m2, float rate, float beta1, float beta2, float eps,
                       float correct1, float correct2) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= PARAMS) return;
    float gi = grad[i];
    grad[i] = 0;
    m1[i] = beta1 * m1[i] + (1.0f - beta1) * gi;
    m2[i] = beta2 * m2[i] + (1.0f - beta2) * gi * gi;
    float mh = m1[i] / correct1, vh = m2[i] / correct2;
    P[i] -= rate * mh / (sqrtf(vh) + eps);
}

}  // namespace

size_t gradient_shared_bytes() { return sizeof(float) * SHARED_FLOATS; }

Net::Net() {
    if (hipMalloc(&params, sizeof(float) * PARAMS) != hipSuccess || hipMalloc(&grad, sizeof(float) * PARAMS) != hipSuccess ||
        hipMalloc(&m1, sizeof(float) * PARAMS) != hipSuccess || hipMalloc(&m2, sizeof(float) * PARAMS) != hipSuccess ||
        hipMalloc(&packed, sizeof(Packed)) != hipSuccess || hipMalloc(&loss, sizeof(float)) != hipSuccess)
        throw std::runtime_error("mlp: out of device memory");
    (void)hipMemset(grad, 0, sizeof(float) * PARAMS);
    (void)hipMemset(m1, 0, sizeof(float) * PARAMS);
    (void)hipMemset(m2, 0, sizeof(float) * PARAMS);
    if (hipFuncSetAttribute((const void *)k_gradient, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)gradient_shared_bytes()) != hipSuccess)
        throw std::runtime_error("mlp: the training kernel's LDS does not fit");
}

Net::~Net() {
    for (void *p : {(void *)params, (void *)grad, (void *)m1, (void *)m2, packed, (void *)loss})
        if (p) (void)hipFree(p);
}

void Net::pack(hipStream_t s) {
    k_pack<<<blocks_for(HIDDEN * HIDDEN), THREADS, 0, s>>>(params, (Packed *)packed);
    check_launch("mlp pack");
}

void Net::forward(const float *x, int64_t n, float *y, hipStream_t s, const int32_t *rows) {
    if (n <= 0) return;
    int64_t waves = (n + 15) / 16;
    int blocks = (int)std::min<int64_t>((waves + 3) / 4, 4096);
    k_forward<<<blocks, THREADS, 0, s>>>((const Packed *)packed, x, n, y, rows);
    check_launch("mlp forward");
}

void Net::gradient(const float *x, const float *t, int64_t n, hipStream_t s) {
    if (n <= 0) return;
    int blocks = (int)std::min<int64_t>((n + CHUNK - 1) / CHUNK, 512);
    k_gradient<<<blocks, THREADS, gradient_shared_bytes(), s>>>(params, x, t, n, grad, loss);
    check_launch("mlp gradient");
}

void Net::adam(float rate, hipStream_t s) {
    steps++;
    const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
    float c1 = 1.0f - std::pow(beta1, (float)steps), c2 = 1.0f - std::pow(beta2, (float)steps);
    k_adam<<<blocks_for(PARAMS), THREADS, 0, s>>>(params, grad, m1, m2, rate, beta1, beta2, eps, c1, c2);
    check_launch("mlp adam");
    pack(s);
}

}  // namespace vrmlp

// ---- the C ABI (include/vexelray_rigid.h): host arrays in and out ----

#include <random>
#include <vector>

#include "vexelray_rigid.h"

struct vr_mlp {
    vrmlp::Net net;
    hipStream_t stream = nullptr;
    vr_mlp() {
        if (hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess)
            throw std::runtime_error("mlp: no stream");
    }
    ~vr_mlp() {
        if (stream) {
            (void)hipStreamSynchronize(stream);
            (void)hipStreamDestroy(stream);
        }
    }
};

namespace vrmlp {

Net *net_of(vr_mlp *m) { return m ? &m->net : nullptr; }

// Weights uniform with variance 1/fan-in, biases zero, from a seeded generator (std::mt19937_64).
std::vector<float> initial_params(uint64_t seed) {
    std::vector<float> p(PARAMS, 0.f);
    std::mt19937_64 rng(seed);
    auto fill = [&](int off, int fanIn, int count) {
        float bound = std::sqrt(3.0f / (float)fanIn);
        std::uniform_real_distribution<float> u(-bound, bound);
        for (int i = 0; i < count; i++) p[off + i] = u(rng);
    };
    fill(OFF_W1, FEATURES, FEATURES * HIDDEN);
    fill(OFF_W2, HIDDEN, HIDDEN * HIDDEN);
    fill(OFF_W3, HIDDEN, HIDDEN * HIDDEN);
    fill(OFF_W4, HIDDEN, HIDDEN * OUTPUTS);
    return p;
}

// Device arrays that live for one call.
struct Scratch {
    std::vector<void *> ptrs;
    template <class T>
    T *get(int64_t count) {
        void *p = nullptr;
        if (hipMalloc(&p, sizeof(T) * (size_t)std::max<int64_t>(count, 1)) != hipSuccess)
            throw std::runtime_error("mlp: out of device memory");
        ptrs.push_back(p);
        return (T *)p;
    }
    ~Scratch() {
        for (void *p : ptrs) (void)hipFree(p);
    }
};

#define MLP_OK(expr)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess) throw std::runtime_error(std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

__global__ void k_gather(const float *src, const int32_t *order, int64_t from, int64_t count, int width, float *dst) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count * width) return;
    int64_t row = i / width;
    int col = (int)(i % width);
    dst[i] = src[(int64_t)order[from + row] * width + col];
}

}  // namespace vrmlp

namespace {

thread_local std::string g_mlpError;

template <class F>
int mlp_guard(F fn) {
    try {
        fn();
        return VR_OK;
    } catch (const std::invalid_argument &e) {
        g_mlpError = e.what();
        return VR_ERR_ARGUMENT;
    } catch (const std::exception &e) {
        g_mlpError = e.what();
        return VR_ERR_DEVICE;
    }
}

}  // namespace

extern "C" {

int vr_mlp_create(uint64_t seed, vr_mlp **out) {
    if (!out) return VR_ERR_ARGUMENT;
    vr_mlp *m = nullptr;
    int rc = mlp_guard([&] {
        int devices = 0;
        if (hipGetDeviceCount(&devices) != hipSuccess || devices < 1) throw std::runtime_error("no HIP device");
        m = new vr_mlp();
        std::vector<float> p = vrmlp::initial_params(seed);
        MLP_OK(hipMemcpy(m->net.params, p.data(), sizeof(float) * vrmlp::PARAMS, hipMemcpyHostToDevice));
        m->net.pack(m->stream);
        MLP_OK(hipStreamSynchronize(m->stream));
    });
    if (rc != VR_OK) {
        delete m;
        return rc;
    }
    *out = m;
    return VR_OK;
}

int vr_mlp_destroy(vr_mlp *m) {
    if (m && m->net.holders > 0) {
        g_mlpError = "the network is set on " + std::to_string(m->net.holders) +
                      " box world(s); unset it (vr_box_learned with null) or destroy them first";
        return VR_ERR_STATE;
    }
    delete m;
    return VR_OK;
}

int vr_mlp_read(vr_mlp *m, float *params) {
    if (!m || !params) return VR_ERR_ARGUMENT;
    return mlp_guard([&] {
        MLP_OK(hipMemcpyAsync(params, m->net.params, sizeof(float) * vrmlp::PARAMS, hipMemcpyDeviceToHost, m->stream));
        MLP_OK(hipStreamSynchronize(m->stream));
    });
}

int vr_mlp_write(vr_mlp *m, const float *params) {
    if (!m || !params) return VR_ERR_ARGUMENT;
    return mlp_guard([&] {
        MLP_OK(hipMemcpyAsync(m->net.params, params, sizeof(float) * vrmlp::PARAMS, hipMemcpyHostToDevice, m->stream));
        m->net.pack(m->stream);
        MLP_OK(hipStreamSynchronize(m->stream));
    });
}

int vr_mlp_forward(vr_mlp *m, const float *x, int64_t n, float *y) {
    if (!m || n < 0 || (n > 0 && (!x || !y))) return VR_ERR_ARGUMENT;
    return mlp_guard([&] {
        vrmlp::Scratch s;
        float *dx = s.get<float>(n * vrmlp::FEATURES), *dy = s.get<float>(n * vrmlp::OUTPUTS);
        MLP_OK(hipMemcpyAsync(dx, x, sizeof(float) * n * vrmlp::FEATURES, hipMemcpyHostToDevice, m->stream));
        m->net.forward(dx, n, dy, m->stream);
        MLP_OK(hipMemcpyAsync(y, dy, sizeof(float) * n * vrmlp::OUTPUTS, hipMemcpyDeviceToHost, m->stream));
        MLP_OK(hipStreamSynchronize(m->stream));
    });
}

int vr_mlp_gradient(vr_mlp *m, const float *x, const float *t, int64_t n, float *grad, float *loss) {
    if (!m || n < 1 || !x || !t || !grad || !loss) return VR_ERR_ARGUMENT;
    return mlp_guard([&] {
        vrmlp::Scratch s;
        float *dx = s.get<float>(n * vrmlp::FEATURES), *dt = s.get<float>(n * vrmlp::OUTPUTS);
        MLP_OK(hipMemcpyAsync(dx, x, sizeof(float) * n * vrmlp::FEATURES, hipMemcpyHostToDevice, m->stream));
        MLP_OK(hipMemcpyAsync(dt, t, sizeof(float) * n * vrmlp::OUTPUTS, hipMemcpyHostToDevice, m->stream));
        MLP_OK(hipMemsetAsync(m->net.grad, 0, sizeof(float) * vrmlp::PARAMS, m->stream));
        MLP_OK(hipMemsetAsync(m->net.loss, 0, sizeof(float), m->stream));
        m->net.gradient(dx, dt, n, m->stream);
        MLP_OK(hipMemcpyAsync(grad, m->net.grad, sizeof(float) * vrmlp::PARAMS, hipMemcpyDeviceToHost, m->stream));
        MLP_OK(hipMemcpyAsync(loss, m->net.loss, sizeof(float), hipMemcpyDeviceToHost, m->stream));
        MLP_OK(hipMemsetAsync(m->net.grad, 0, sizeof(float) * vrmlp::PARAMS, m->stream));
        MLP_OK(hipStreamSynchronize(m->stream));
    });
}

int vr_mlp_train(vr_mlp *m, const float *x, const float *t, int64_t n, int32_t epochs, int32_t batch, float rate,
                 uint64_t seed, float *loss) {
    if (!m || n < 1 || !x || !t || epochs < 1 || batch < 1 || !(rate > 0)) return VR_ERR_ARGUMENT;
    return mlp_guard([&] {
        vrmlp::Scratch s;
        float *dx = s.get<float>(n * vrmlp::FEATURES), *dt = s.get<float>(n * vrmlp::OUTPUTS);
        float *bx = s.get<float>((int64_t)batch * vrmlp::FEATURES), *bt = s.get<float>((int64_t)batch * vrmlp::OUTPUTS);
        int32_t *order = s.get<int32_t>(n);
        MLP_OK(hipMemcpyAsync(dx, x, sizeof(float) * n * vrmlp::FEATURES, hipMemcpyHostToDevice, m->stream));
        MLP_OK(hipMemcpyAsync(dt, t, sizeof(float) * n * vrmlp::OUTPUTS, hipMemcpyHostToDevice, m->stream));
        std::vector<int32_t> perm(n);
        std::mt19937_64 rng(seed);
        float last = 0;
        for (int e = 0; e < epochs; e++) {
            for (int64_t i = 0; i < n; i++) perm[i] = (int32_t)i;
            std::shuffle(perm.begin(), perm.end(), rng);
            MLP_OK(hipMemcpyAsync(order, perm.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, m->stream));
            MLP_OK(hipMemsetAsync(m->net.loss, 0, sizeof(float), m->stream));
            int64_t batches = 0;
            for (int64_t from = 0; from < n; from += batch, batches++) {
                int64_t count = std::min<int64_t>(batch, n - from);
                vrmlp::k_gather<<<vrmlp::blocks_for(count * vrmlp::FEATURES), vrmlp::THREADS, 0, m->stream>>>(
                    dx, order, from, count, vrmlp::FEATURES, bx);
                vrmlp::k_gather<<<vrmlp::blocks_for(count * vrmlp::OUTPUTS), vrmlp::THREADS, 0, m->stream>>>(
                    dt, order, from, count, vrmlp::OUTPUTS, bt);
                m->net.gradient(bx, bt, count, m->stream);
                m->net.adam(rate, m->stream);
            }
            float sum = 0;
            MLP_OK(hipMemcpyAsync(&sum, m->net.loss, sizeof(float), hipMemcpyDeviceToHost, m->stream));
            MLP_OK(hipStreamSynchronize(m->stream));
            last = sum / (float)batches;  // the mean over the epoch's batches of each batch's mean loss
        }
        if (loss) *loss = last;
    });
}

int vr_mlp_last_error(char *buf, int32_t len) {
    if (!buf || len <= 0) return VR_OK;
    std::strncpy(buf, g_mlpError.c_str(), (size_t)len - 1);
    buf[len - 1] = 0;
    return VR_OK;
}

}  // extern "C"
