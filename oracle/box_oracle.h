// box_oracle.h — the CPU oracle's box world, behind the oracle's C ABI (sphere_oracle.cpp). TEST INFRASTRUCTURE.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../include/vexelray_rigid.h"

namespace vro {

struct BoxWorld;

BoxWorld *box_create(const vr_step_desc &desc);  // throws std::invalid_argument for the reference's validation
void box_destroy(BoxWorld *w);
void box_clear(BoxWorld *w);
bool box_buffer(BoxWorld *w, const std::string &name, void **ptr, int64_t *count, int32_t *dtype);
void box_params(BoxWorld *w, const int32_t *words, int32_t n);
void box_step(BoxWorld *w, vr_step_report *rep);
// BoxStepper.step up to the rounds' write: the collide list and the colouring, no substep.
void box_collide_list(BoxWorld *w, vr_step_report *rep);
// Sweeps stopped on a residual, as vr_box_tolerance; what the last step's phases did, and its sweeps' residuals.
void box_tolerance(BoxWorld *w, float tol, bool measure);
const vr_box_solve &box_solve(BoxWorld *w);
const std::vector<float> &box_curve(BoxWorld *w);
// The narrow phase alone over pairs (first, second), MANIFOLD_WORDS floats a pair out.
void box_collide(BoxWorld *w, const int32_t *pairs, int32_t npairs, float *manifold);

}  // namespace vro
