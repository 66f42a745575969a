// box_world.h — the box world behind the C ABI (world.hip dispatches to it for VR_BOXES).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>
#include <vector>

#include "vexelray_rigid.h"

namespace vrmlp {
struct Net;
}

namespace vrbox {

struct BoxWorld;

struct BoxBuf {
    void *p;
    int64_t count;
    int dtype;
    // Words a body (> 0) for a buffer the device keeps slot-major (word w of body b at w * bodies + b) while the
    // buffer's name reads and writes it body by body, as the reference lays it out; 0 when the two agree.
    int bodyWords = 0;
};

BoxWorld *box_create(const vr_step_desc &desc);  // std::invalid_argument for the reference's validation
// Many independent scenes stepped together (vr_batch_create).
BoxWorld *box_create_batch(const vr_step_desc &desc, int32_t scenes, const int32_t *bodies, const float *extents,
                           const int32_t *grids, const float *cells);
void box_destroy(BoxWorld *w);
void box_clear(BoxWorld *w);
bool box_find(BoxWorld *w, const std::string &name, BoxBuf *out);
hipStream_t box_stream(BoxWorld *w);
void box_params(BoxWorld *w, const int32_t *words, int32_t n);
void box_step(BoxWorld *w, vr_step_report *rep);
// BoxStepper.step up to the rounds' write: the collide list and the colouring, no substep (vr_box_collide_list).
void box_collide_list(BoxWorld *w, vr_step_report *rep);
// Sweeps stopped on a residual (vr_box_tolerance), and what the last step's phases did.
void box_tolerance(BoxWorld *w, float tol, bool measure);
// The launch plan timed again over the next steps (vr_box_retune); the plan the next step takes (vr_box_plan).
void box_retune(BoxWorld *w);
void box_plan(BoxWorld *w, int32_t *plan, int32_t *tuning);
// Sweep timing (vr_world_profile on a box world): on zeroes the counters.
void box_profile(BoxWorld *w, bool on);
void box_profile_read(BoxWorld *w, vr_profile *out);
void box_solve_read(BoxWorld *w, vr_box_solve *out);
int32_t box_solve_curve(BoxWorld *w, float *out, int32_t n);
void box_wall_impulses_written(BoxWorld *w);  // after a write of lambda or rub through the ABI
void box_shapes_written(BoxWorld *w);         // after a write of shape: the wall impulses follow their bodies
void box_pair_buffer_written(BoxWorld *w, const std::string &name);  // after a write of a manifold or impulse buffer
// Where each body's wall impulse words sit in the slot-major buffers (word w of body s at w * n + out[s]).
void box_wall_places(BoxWorld *w, std::vector<int32_t> &out);
// The learned warm start: a log of per-point (features, targets), and the network that starts each step.
void box_log(BoxWorld *w, bool on, int64_t capacity);
int64_t box_samples(BoxWorld *w, float *x, float *t, int64_t most);
void box_learned(BoxWorld *w, vrmlp::Net *net);
void box_diagnostics(BoxWorld *w, double *out);  // vr_world_diagnostics for a box world
// The same into device scratch the caller gives: vrdiag::partial_words(bodies) partials, then D_WORDS results.
void box_diagnostics_into(BoxWorld *w, double *partials, double *devOut, double *out);
int box_bodies(BoxWorld *w);
void box_collide(BoxWorld *w, const int32_t *pairs, int32_t npairs, float *manifold);

}  // namespace vrbox
