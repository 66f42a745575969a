// box_kernels.h — device layout of a box world and the launchers of its kernels.
//
// Body state is structure-of-arrays (x, y, z, u, v, w, lx, ly, lz, qx, qy, qz, qw, im, hx, hy, hz, shape), one f32
// array a field. Pairs use the reference's slot layout: 32 slots a body, slot p's pair (p / 32, partner) with a
// 40-float manifold and 28 floats of impulses (BoxBroad.java:46; BoxContact.java:77-88; BoxPairs.java:37).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace vrbox {

struct BoxDev {
    int n, slots, width, nx, ny, nz, cells;
    float cell;
    float *x[3], *vel[3], *l[3], *q[4], *im, *half[3], *shape;
    const float *params;
    float *lambda, *rub, *roll, *shown;
    int32_t *keys, *order, *starts;
    int32_t *pairs, *pairsWas, *contacts, *rounds;
    float *manifold, *manifoldWas, *impulses, *impulsesWas;
};

void launch_velocities(const BoxDev &D, hipStream_t s);
void launch_positions(const BoxDev &D, hipStream_t s);
void launch_walls(const BoxDev &D, int sweep, bool reverse, hipStream_t s);
void launch_solve_round(const BoxDev &D, int sweep, bool reverse, int r, hipStream_t s);
void launch_bin(const BoxDev &D, hipStream_t s);
void launch_list(const BoxDev &D, hipStream_t s);
void launch_collide(const BoxDev &D, hipStream_t s);
void launch_collide_pairs(const BoxDev &D, const int32_t *pairs, int npairs, float *manifold, hipStream_t s);
void launch_carry_counts(const BoxDev &D, hipStream_t s);
void launch_show(const BoxDev &D, hipStream_t s);

}  // namespace vrbox
