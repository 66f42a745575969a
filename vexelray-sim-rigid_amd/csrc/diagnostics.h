// diagnostics.h — a world's diagnostics reduced on the device (diagnostics.hip), in the order vr_world_diagnostics
// returns them.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "box_kernels.h"
#include "sphere_kernels.h"

namespace vrdiag {

enum {
    D_OVERLAP = 0, D_WALL, D_KINETIC, D_ROTATIONAL, D_POTENTIAL, D_MOMENTUM, D_ANGULAR = D_MOMENTUM + 3,
    D_SPEED = D_ANGULAR + 3, D_BROKEN, D_DEPTH, D_WORDS
};

int64_t partial_words(int n);  // doubles of scratch the launches below need for n bodies
// The world's grid (keys, order, starts) must be built at the current positions when `grid` is set.
void launch_spheres(const vr::SphereDev &D, bool grid, double *partials, double *out, hipStream_t s);
void launch_boxes(const vrbox::BoxDev &D, double *partials, double *out, hipStream_t s);

}  // namespace vrdiag
