// box_colour.hip — the touching pairs coloured into rounds on the device, as the reference's host does it.
//
// BoxStepper.step (core/.../box/BoxStepper.java:71-90) reads the contacts back and colours them on the host: in slot
// order, each touching pair takes the first of the 32 rounds that neither of its bodies is in yet, or is left
// unplaced. That is sequential first fit, and its result is fixed by one rule: a pair's round is the smallest not
// taken by the earlier (lower-slot) touching pairs it shares a body with. So the pairs can be coloured in parallel
// in the order Jones and Plassmann give, with slot order as the priority: a pair takes its round once every earlier
// pair it shares a body with has one, each step of the iteration colouring the pairs whose earlier neighbours are all
// decided. Every pair ends with the round sequential first fit gives it, so the rounds, their contents (pairs in slot
// order, as the reference fills them) and the report's counts are the reference's.
#include <algorithm>
#include <climits>
#include <stdexcept>
#include <string>

#include "box_kernels.h"
#include "box_math.h"
#include "sphere_kernels.h"

namespace vrbox {

namespace {

constexpr int THREADS = 256;
constexpr int NOT_TOUCHING = MOST_ROUNDS + 1;  // a colour key after every round and "unplaced" (MOST_ROUNDS)

inline int blocks_for(int64_t n) { return (int)std::max<int64_t>(1, (n + THREADS - 1) / THREADS); }

inline void check_launch(const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

__device__ __forceinline__ bool touching(const BoxDev &D, int p) { return D.contacts[p] >= 0; }

__global__ void k_colour_begin(BoxDev D, Colouring C) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < D.n) {
        C.incDeg[s] = 0;
        C.incPut[s] = 0;
    }
    if (s < COLOUR_FLAGS) C.flags[s] = 0;
}

// Each workgroup's count of touching pairs among its slots; each touching pair counted at both its bodies.
__global__ void k_touch_count(BoxDev D, Colouring C) {
    __shared__ int sWave[THREADS / 64];
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    bool t = p < D.slots && touching(D, p);
    if (t) {
        atomicAdd(&C.incDeg[p / SLOTS], 1);
        atomicAdd(&C.incDeg[D.contacts[p]], 1);
        C.decided[p] = INT_MAX;
    }
    unsigned long long mask = __ballot(t);
    if ((threadIdx.x & 63) == 0) sWave[threadIdx.x >> 6] = __popcll(mask);
    __syncthreads();
    if (threadIdx.x == 0) {
        int total = 0;
        for (int w = 0; w < THREADS / 64; w++) total += sWave[w];
        C.touchSeg[blockIdx.x] = total;
    }
}

// The touching pairs, in slot order, at the places the scanned counts give.
__global__ void k_touch_write(BoxDev D, Colouring C) {
    __shared__ int sWave[THREADS / 64];
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    bool t = p < D.slots && touching(D, p);
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    unsigned long long mask = __ballot(t);
    if (lane == 0) sWave[wave] = __popcll(mask);
    __syncthreads();
    int before = C.touchSeg[blockIdx.x];
    for (int w = 0; w < wave; w++) before += sWave[w];
    if (t) C.touch[before + __popcll(mask & ((1ull << lane) - 1ull))] = p;
}

// Each touching pair placed in both its bodies' incidence lists (their order does not matter).
__global__ void k_inc_place(BoxDev D, Colouring C) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= *C.touching) return;
    int p = C.touch[t];
    int a = p / SLOTS, b = D.contacts[p];
    C.inc[C.incOff[a] + atomicAdd(&C.incPut[a], 1)] = p;
    C.inc[C.incOff[b] + atomicAdd(&C.incPut[b], 1)] = p;
}

// One step of the colouring. A pair still open colours itself if every earlier pair at its bodies was decided in an
// earlier step (decided[q] < iter; those deciding in this step still read as open, so nothing here races): its round
// is the first none of them took, or MOST_ROUNDS (unplaced) if they took all 32. Otherwise it marks the step as
// leaving pairs open and stays on the open list. Workgroup g owns places g * THREADS onwards of the open lists: in
// step 0 it takes the touching pairs there, in each later step the pairs it kept open in the step before (segIn[g] of
// them, in `in`), and it keeps the ones still open in its places of `out`, counting in LDS; so the later steps,
// which decide the last few pairs, read only those, with no device-wide counter. Which order a segment holds its
// pairs in does not matter.
__global__ void __launch_bounds__(THREADS) k_colour_step(BoxDev D, Colouring C, int iter, const int32_t *in,
                                                         int32_t *out, const int32_t *segIn, int32_t *segOut) {
    __shared__ int sKept;
    int g = blockIdx.x, lane = threadIdx.x & 63;
    if (g == 0 && threadIdx.x == 0) C.flags[(iter + 1) % COLOUR_FLAGS] = 0;
    int n = iter == 0 ? min(THREADS, *C.touching - g * THREADS) : segIn[g];
    if (n <= 0) {
        if (threadIdx.x == 0) segOut[g] = 0;
        return;
    }
    if (threadIdx.x == 0) sKept = 0;
    __syncthreads();
    int t = threadIdx.x;
    int p = -1;
    bool open = false;
    if (t < n) {
        p = iter == 0 ? C.touch[g * THREADS + t] : in[g * THREADS + t];
        int ends[2] = {p / SLOTS, D.contacts[p]};
        unsigned taken = 0;
        bool ready = true;
        for (int e = 0; e < 2 && ready; e++) {
            int from = C.incOff[ends[e]], to = C.incOff[ends[e] + 1];
            for (int k = from; k < to; k++) {
                int q = C.inc[k];
                if (!(q < p)) continue;
                if (C.decided[q] >= iter) {
                    ready = false;
                    break;
                }
                int r = C.colour[q];
                if (r < MOST_ROUNDS) taken |= 1u << r;
            }
        }
        if (ready) {
            C.colour[p] = taken == 0xFFFFFFFFu ? MOST_ROUNDS : __ffs(~taken) - 1;
            C.decided[p] = iter;
        } else {
            open = true;
        }
    }
    unsigned long long mask = __ballot(open);
    if (mask) {
        int leader = __ffsll((long long)mask) - 1;
        int at = 0;
        if (lane == leader) at = atomicAdd(&sKept, __popcll(mask));
        at = __builtin_amdgcn_readlane(at, leader);
        if (open) out[g * THREADS + at + __popcll(mask & ((1ull << lane) - 1ull))] = p;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        segOut[g] = sKept;
        if (sKept > 0) C.flags[iter % COLOUR_FLAGS] = 1;
    }
}

// The sort keys: each touching pair's round (MOST_ROUNDS when unplaced), in the touching list's (slot) order.
__global__ void k_colour_keys(Colouring C) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= *C.touching) return;
    C.keys[t] = C.colour[C.touch[t]];
}

// BoxStep's rounds buffer: round r's pairs in slot order, then -1 (BoxStepper.java:80-90).
__global__ void k_fill_rounds(BoxDev D, Colouring C) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)MOST_ROUNDS * D.width) return;
    int r = (int)(i / D.width), k = (int)(i % D.width);
    int from = C.starts[r], to = C.starts[r + 1];
    D.rounds[i] = k < to - from ? C.touch[C.order[from + k]] : -1;
}

// Sums for the report: partial sums a workgroup, then the partials summed the same way until one is left.
// mode 0: the values; mode 1: how far each exceeds SLOTS (a body's partners beyond its slots, unlisted).
__global__ void k_partial_sums(const int32_t *in, int n, int mode, int32_t *out) {
    __shared__ int sWave[THREADS / 64];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int v = 0;
    if (i < n) v = mode == 0 ? in[i] : max(0, in[i] - SLOTS);
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) sWave[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        int total = 0;
        for (int w = 0; w < THREADS / 64; w++) total += sWave[w];
        out[blockIdx.x] = total;
    }
}

void launch_sum(const int32_t *in, int n, int mode, int32_t *out, int32_t *tmp, hipStream_t s) {
    int blocks = blocks_for(n);
    int32_t *level = blocks == 1 ? out : tmp;
    k_partial_sums<<<blocks, THREADS, 0, s>>>(in, n, mode, level);
    check_launch("sum");
    while (blocks > 1) {
        int next = blocks_for(blocks);
        int32_t *dst = next == 1 ? out : level + blocks;
        k_partial_sums<<<next, THREADS, 0, s>>>(level, blocks, 0, dst);
        check_launch("sum");
        level = dst;
        blocks = next;
    }
}

}  // namespace

int64_t colour_tmp_words(int64_t slots) { return 2 * ((slots + THREADS - 1) / THREADS) + 64; }

void launch_colour_prepare(const BoxDev &D, const Colouring &C, const vr::SortScratch &S, hipStream_t s) {
    k_colour_begin<<<blocks_for(std::max(D.n, COLOUR_FLAGS)), THREADS, 0, s>>>(D, C);
    check_launch("colour begin");
    int blocks = blocks_for(D.slots);
    k_touch_count<<<blocks, THREADS, 0, s>>>(D, C);
    check_launch("touch count");
    vr::scan_exclusive(C.touchSeg, C.touchSeg, blocks, C.touching, S.scanTmp, s);
    k_touch_write<<<blocks, THREADS, 0, s>>>(D, C);
    check_launch("touch write");
    // incOff[n] is the total, so every body's list ends where the next begins.
    vr::scan_exclusive(C.incDeg, C.incOff, D.n, C.incOff + D.n, S.scanTmp, s);
    k_inc_place<<<blocks_for(D.slots), THREADS, 0, s>>>(D, C);
    check_launch("incidence");
}

// The open lists alternate between the colouring's key and order words, which its sort (launch_colour_finish) fills
// only after the last step; so do the segments' counts.
void launch_colour_steps(const BoxDev &D, const Colouring &C, int touchingPairs, int from, int count,
                         hipStream_t s) {
    if (touchingPairs <= 0) return;
    if (touchingPairs > C.most) throw std::runtime_error("colour: more touching pairs than slots");
    int blocks = blocks_for(touchingPairs);
    for (int k = 0; k < count; k++) {
        int iter = from + k;
        int32_t *out = (iter & 1) ? C.order : C.keys;
        const int32_t *in = (iter & 1) ? C.keys : C.order;
        k_colour_step<<<blocks, THREADS, 0, s>>>(D, C, iter, in, out, C.segOpen[(iter + 1) % 2], C.segOpen[iter % 2]);
    }
    check_launch("colour step");
}

void launch_colour_finish(const BoxDev &D, const Colouring &C, int touchingPairs, hipStream_t s) {
    if (touchingPairs > C.most) throw std::runtime_error("colour: more touching pairs than slots");
    if (touchingPairs > 0) {
        k_colour_keys<<<blocks_for(touchingPairs), THREADS, 0, s>>>(C);
        check_launch("colour keys");
        // A stable sort of the touching list by round: each round's pairs stay in slot order.
        vr::SphereDev view{};
        view.n = touchingPairs;
        view.cells = NOT_TOUCHING;
        view.keys = C.keys;
        view.order = C.order;
        view.starts = C.starts;
        vr::launch_sort(view, C.sort, s);
    } else {
        vr::launch_fill_i32(C.starts, 0, NOT_TOUCHING + 1, s);
    }
    k_fill_rounds<<<blocks_for((int64_t)MOST_ROUNDS * D.width), THREADS, 0, s>>>(D, C);
    check_launch("fill rounds");
    launch_sum(D.contacts + D.slots, D.slots, 0, C.report + REPORT_POINTS, C.sumTmp, s);
    launch_sum(D.contacts + 2 * (int64_t)D.slots, D.n, 1, C.report + REPORT_UNLISTED, C.sumTmp, s);
}

}  // namespace vrbox
