// mlp.h — the learned warm start's network (mlp.hip): layout of its parameters and its device-side operations.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace vrmlp {

constexpr int FEATURES = 16, HIDDEN = 64, OUTPUTS = 3;
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + FEATURES * HIDDEN, OFF_W2 = OFF_B1 + HIDDEN,
              OFF_B2 = OFF_W2 + HIDDEN * HIDDEN, OFF_W3 = OFF_B2 + HIDDEN, OFF_B3 = OFF_W3 + HIDDEN * HIDDEN,
              OFF_W4 = OFF_B3 + HIDDEN, OFF_B4 = OFF_W4 + HIDDEN * OUTPUTS, PARAMS = OFF_B4 + OUTPUTS;  // 9,603

// The network on one device. All operations are asynchronous on the stream given; pointers are device pointers.
struct Net {
    float *params = nullptr;        // PARAMS, the f32 master copy
    float *grad = nullptr;          // PARAMS, added to by gradient(), cleared by adam()
    float *m1 = nullptr, *m2 = nullptr;  // Adam's moments
    float *loss = nullptr;          // 1, added to by gradient()
    void *packed = nullptr;         // bf16 weights for forward()
    int steps = 0;                  // Adam steps taken
    int holders = 0;                // box worlds that start their steps from this network (vr_box_learned)
    Net();
    ~Net();
    void pack(hipStream_t s);       // after params change
    // x: n x 16, y: n x 3; with `rows` (device memory), only the first *rows of the n (at most n) are taken.
    void forward(const float *x, int64_t n, float *y, hipStream_t s, const int32_t *rows = nullptr);
    void gradient(const float *x, const float *t, int64_t n, hipStream_t s);  // mean squared error over n points
    void adam(float rate, hipStream_t s);
};

}  // namespace vrmlp

struct vr_mlp;
namespace vrmlp {
Net *net_of(vr_mlp *m);  // the network behind a handle of the C ABI
}  // namespace vrmlp
