/*
 * vexelray_rigid.h — the C ABI of the MI355X rigid-body step.
 *
 * One opaque world per simulation, owned by the caller from vr_world_create to vr_world_destroy. Every call returns
 * an int status: VR_OK (0), or a negative code whose message vr_last_error() copies out. Host arrays are copied in
 * and out; no host pointer is kept past the call. A world is not thread-safe: one thread owns it after it is made,
 * and all of its device work runs on one HIP stream of its own.
 *
 * Each entry point replaces part of the reference's Java world/body/step API (paths relative to the
 * vexelray-sim-rigid repository):
 *
 *   vr_world_create   SphereStep(spheres, substeps, iterations, grid, solve, rounds, contactsPerSphere)
 *                     (core/.../sphere/SphereStep.java:115-209) plus the runner it is registered on
 *                     (gui/.../SphereSimulation.java:70-97 -> PassRunner.gpu(...)); SphereGrid(nx, ny, nz, cell)
 *                     (core/.../sphere/SphereGrid.java:11-19). With VR_BOXES, BoxStep(boxes, substeps, sweeps,
 *                     symmetric, grid) (core/.../box/BoxStep.java:81-162) and BoxSimulation's runner
 *                     (gui/.../BoxSimulation.java:34-75).
 *   vr_world_clear    PassRunner.clear(), then SphereStep.constants() written (SphereSimulation.java:154-166,
 *                     SphereStep.java:157-160, 224-226, 419).
 *   vr_world_write    PassRunner.write(name, int[]|float[]) (ctest/.../sphere/Scene.java:127-140).
 *   vr_world_read     PassRunner.floats(name) / read(name) (Scene.java:176-204; SphereSimulation.state() :250-254).
 *   vr_world_start    SphereSimulation.start(...) (SphereSimulation.java:154-178): clear, writes, params, shown.
 *   vr_world_params   SphereSimulation.params(int[]) (SphereSimulation.java:182-185): 13 words of raw float bits
 *                     laid out as Spheres.params (core/.../sphere/Spheres.java:173-226); for boxes 18 words laid
 *                     out as Boxes.params (core/.../box/Boxes.java:212-234), BoxSimulation.params.
 *   vr_world_friction SphereStepper.friction(boolean) (core/.../sphere/SphereStepper.java:110-112).
 *   vr_world_slice    SphereSimulation.slice(int) (gui/.../SphereSimulation.java:136-148).
 *   vr_world_step     SphereStepper.step(Runner) -> Report (SphereStepper.java:66-103) as driven by
 *                     SphereSimulation.step()/stepAndWait() (SphereSimulation.java:190-229); for boxes
 *                     BoxStepper.step(Runner) -> Report (core/.../box/BoxStepper.java:56-100) as driven by
 *                     BoxSimulation.stepAndWait() (gui/.../BoxSimulation.java:116-142).
 *   vr_world_finish   SphereSimulation.finish() / PassRunner.finish() (SphereSimulation.java:243-245).
 *   vr_world_destroy  SphereSimulation.close() (SphereSimulation.java:287-295).
 *   vr_last_error     the exception message a Java caller would have seen (IllegalArgumentException for
 *                     validation, IllegalStateException for a pass that never ends or a wrong mode).
 */
#ifndef VEXELRAY_RIGID_H
#define VEXELRAY_RIGID_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vr_world vr_world;

/* Status codes. */
#define VR_OK 0
#define VR_ERR_ARGUMENT (-1)  /* IllegalArgumentException in the reference */
#define VR_ERR_STATE (-2)     /* IllegalStateException in the reference */
#define VR_ERR_DEVICE (-3)    /* a HIP call failed, or no device */
#define VR_ERR_NAME (-4)      /* no buffer of that name, or the wrong length */

/* World kinds. */
#define VR_SPHERES 1
#define VR_BOXES 2

/* SphereStep.Solve (SphereStep.java:37-43). */
#define VR_JACOBI 0
#define VR_GAUSS_SEIDEL_BY_CONTACT 1

/* Element types of the named buffers, as vr_world_buffer_info reports them. */
#define VR_F32 1
#define VR_I32 2

/*
 * What a step is made of: SphereStep's constructor arguments (SphereStep.java:115-120). has_grid = 0 is the
 * reference's null grid (every pair). fixed_rounds = 1 runs Gauss-Seidel as the one list SphereStep.step() builds,
 * `rounds` rounds a pass, instead of until done (ctest/.../sphere/Scene.java:69-75, 158-161).
 */
typedef struct vr_step_desc {
    int32_t bodies;
    int32_t substeps;
    int32_t iterations;
    int32_t has_grid;
    int32_t nx, ny, nz;
    float cell;
    int32_t solve;
    int32_t rounds;
    int32_t contacts_per_sphere;
    int32_t fixed_rounds;
    /*
     * 0: the reference's contact priority, (a*n + o)*0x9E3779B1 + g*0x632BE5AB kept to 31 bits, and its limit of
     * 46,340 spheres for Gauss-Seidel (Spheres.java:783-787, 1166-1168; SphereStep.java:128-131).
     * 1: the same formula in 63-bit arithmetic with 64-bit odd multipliers, one to one over every pair of up to
     * 2^31 spheres, so Gauss-Seidel takes any sphere count. Identical rounds wherever both apply is not claimed:
     * the priorities differ, so the colouring does.
     */
    int32_t wide_pairs;
    /* Boxes (BoxStep(boxes, substeps, sweeps, symmetric, grid), core/.../box/BoxStep.java:81-162): sweeps of the
     * solve and of the relax a substep makes, and whether every other sweep runs backwards. */
    int32_t sweeps;
    int32_t symmetric;
} vr_step_desc;

/*
 * SphereStepper.Report (SphereStepper.java:42) and SphereSimulation.StepReport's timings (SphereSimulation.java:48).
 * `waits` counts the host waits the reference's stepper makes for the same step (one per report it reads), so the
 * two reports compare; this library decides "done" on the device and waits for none of them.
 */
typedef struct vr_step_report {
    int32_t waits;
    int32_t rounds;
    int32_t most_rounds;
    int32_t unlisted;
    /* Boxes: BoxStepper.Report(touching, points, rounds, unplaced, unlisted) (core/.../box/BoxStepper.java:40). */
    int32_t touching;
    int32_t points;
    int32_t unplaced;
    int32_t reserved;
    int64_t wall_ns;
    int64_t gpu_ns;
    /*
     * SphereSimulation.StepReport.submissions (SphereSimulation.java:48, submit :231-248): the submissions the
     * reference makes for the same step, each list its stepper runs cut into pieces of at most vr_world_slice's
     * dispatches. For boxes, BoxSimulation's count (BoxSimulation.java:116-142): one a list, no slicing.
     */
    int32_t submissions;
    int32_t unused;
} vr_step_report;

/*
 * Measurement (no counterpart in the reference): HIP events on the world's stream bracket every Gauss-Seidel pass,
 * whose round kernels (k_round_wide ... k_round_tail) are launched back to back, so the time in round kernels is
 * known without a profiler. listed_sum adds up, over every pass run (parting and rubbing), the contacts listed for
 * its substep: the C of the algorithmic-byte count 72 C + 104 N a pass is credited with. enqueue_ns adds up the
 * host's time from a step's first launch to its last, before it waits for the step: when it is close to the step's
 * time, the host's launch rate, not the GPU, sets the pace.
 */
typedef struct vr_profile {
    int64_t round_ns;
    int64_t round_launches;
    int64_t passes;
    int64_t listed_sum;
    int64_t steps;
    int64_t enqueue_ns;
    /*
     * Box worlds: events bracket the colour rounds (k_solve_round, back to back) and the walls sweep (k_walls) of
     * every solve and relax sweep. round_ns / round_launches are the rounds', walls_ns / walls_launches the walls';
     * passes counts those sweeps; listed_sum and points_sum add up the touching pairs and their manifold points the
     * sweeps took (P and P k-bar of the algorithmic bytes), bodies_sum the bodies. A batch whose sweeps' rounds
     * run a workgroup a scene group (VR_PLAN_GROUP: all of a sweep's rounds in one launch) counts them in sweep_ns
     * and sweep_launches (one a sweep) in place of round_ns and round_launches, its walls as above; and one that runs
     * a whole phase of sweeps a launch (k_group_phase: each group through the phase's walls and rounds) counts each
     * phase launch in sweep_ns and sweep_launches, walls_ns and walls_launches then staying at zero.
     */
    int64_t walls_ns;
    int64_t walls_launches;
    int64_t points_sum;
    int64_t bodies_sum;
    int64_t sweep_ns;
    int64_t sweep_launches;
} vr_profile;

int vr_world_create(int32_t kind, const vr_step_desc *desc, vr_world **out);
/*
 * A batch (no counterpart in the reference, which steps one scene a world): `scenes` independent box scenes in one
 * world, stepped together. Scene s holds bodies[s] bodies, consecutive from the sum of the counts before it, in a
 * walled box extents[3s .. 3s+2] (the reference's box world's params sx, sy, sz, which a batch then ignores), with
 * a broad-phase grid of grids[3s .. 3s+2] cells of cells[s] metres (BoxBroad.grid). desc holds the step
 * (bodies = the total, substeps, sweeps, symmetric; has_grid = 1). Buffers are the box world's, over all bodies,
 * and one set of params serves every scene. Pairs never join two scenes and the rounds are coloured in slot order,
 * so each scene steps exactly as it would alone; the report adds the scenes' counts up (rounds: the most).
 */
int vr_batch_create(const vr_step_desc *desc, int32_t scenes, const int32_t *bodies, const float *extents,
                    const int32_t *grids, const float *cells, vr_world **out);
int vr_world_clear(vr_world *w);
int vr_world_write(vr_world *w, const char *name, const void *src, int64_t count);
int vr_world_read(vr_world *w, const char *name, void *dst, int64_t count);
/* Length and element type (VR_F32 / VR_I32) of a named buffer. */
int vr_world_buffer_info(vr_world *w, const char *name, int64_t *count, int32_t *dtype);
int vr_world_params(vr_world *w, const int32_t *words, int32_t n);
/*
 * SphereSimulation.start(x, y, z, u, v, w, r, im, params) (gui/.../SphereSimulation.java:154-178) for a sphere
 * world: clear and the step's constants, the eight fields (each as long as the sphere count), the 13 parameter words
 * (with SphereStepper.friction(Spheres.rubs(params))), and the picture's "before" and "after" both this state. The
 * same as vr_world_clear, vr_world_write of each field, vr_world_params and vr_world_write of "shown". A box world
 * starts through those calls (BoxSimulation.start works out the angular momentum on the host): VR_ERR_STATE.
 */
int vr_world_start(vr_world *w, const float *x, const float *y, const float *z, const float *u, const float *v,
                   const float *vw, const float *r, const float *im, const int32_t *params, int32_t nparams);
int vr_world_friction(vr_world *w, int32_t on);
/*
 * SphereSimulation.slice(int) (gui/.../SphereSimulation.java:136-148): at most `dispatches` dispatches a submission,
 * 0 for no bound; below 0 is VR_ERR_ARGUMENT. This library enqueues a step's kernels on one stream, each its own
 * dispatch, whatever the slice: the slice changes only the submissions the report counts, as it changes nothing in
 * the reference's results. A box world has no slice (VR_ERR_STATE).
 */
int vr_world_slice(vr_world *w, int32_t dispatches);
int vr_world_step(vr_world *w, int32_t wait, vr_step_report *report);
int vr_world_finish(vr_world *w);
int vr_world_destroy(vr_world *w);
int vr_last_error(char *buf, int32_t len);
/*
 * The narrow phase alone: BoxContact.collide (core/.../box/BoxContact.java:136-246) over `npairs` pairs of a box
 * world's bodies as they are now, pairs[2k] first and pairs[2k+1] second, writing 40 floats a pair to `manifold`
 * (normal, count, axis, up to 4 points of anchor on the first, anchor on the second, depth). What the reference's
 * tests run as a collide-only program (ctest/.../box/BoxContactTest.java:441-505).
 */
int vr_box_collide(vr_world *w, const int32_t *pairs, int32_t npairs, float *manifold);
/*
 * The first half of BoxStepper.step (core/.../box/BoxStepper.java:62-95) on a box world or batch: the collide list
 * BoxStep.collide() (core/.../box/BoxStep.java:138-153: keep, bin and sort, list, collide, carry, counts) run and
 * its `contacts` read back, and the touching pairs coloured into `rounds`, with no substep after it: what the
 * reference's Runner.runAndRead(step.collide(), "contacts") leaves, before runner.run(step.step(used)). The report
 * holds the step's counts (touching, points, rounds, unplaced, unlisted). A later vr_world_step collides again from
 * what this left, as the reference's next step would. Pairs out of 0 .. bodies - 1, or a negative count, given to
 * vr_box_collide are VR_ERR_ARGUMENT, checked before anything runs.
 */
int vr_box_collide_list(vr_world *w, vr_step_report *report);
/*
 * Diagnostics reduced on the device in f64 (SphereDiagnostics.of, core/.../sphere/SphereDiagnostics.java:34-83;
 * BoxDiagnostics.of, core/.../box/BoxDiagnostics.java:18-44, both host-side in the reference), into out[14]:
 * VR_DIAG_OVERLAP deepest sphere-sphere overlap over the smaller radius (spheres; through the grid, rebuilt at the
 * current positions, so the grid buffers change), VR_DIAG_WALL deepest wall overlap over the radius (spheres),
 * kinetic energy (with spin), rotational energy, potential energy (m (-gy) y), momentum x y z, angular momentum x y z
 * (about the origin), top speed, 1 if any state is not finite, and (boxes) the deepest manifold point of the last
 * collide. The box and gravity are the params' f32 words. `n` must be at least 14.
 */
#define VR_DIAG_OVERLAP 0
#define VR_DIAG_WALL 1
#define VR_DIAG_KINETIC 2
#define VR_DIAG_ROTATIONAL 3
#define VR_DIAG_POTENTIAL 4
#define VR_DIAG_MOMENTUM 5
#define VR_DIAG_ANGULAR 8
#define VR_DIAG_SPEED 11
#define VR_DIAG_BROKEN 12
#define VR_DIAG_DEPTH 13
#define VR_DIAG_WORDS 14
int vr_world_diagnostics(vr_world *w, double *out, int32_t n);
/*
 * No counterpart in the reference. vr_world_diagnostics without allocating: for callers that cannot allocate device
 * memory (a serving loop that must not call hipMalloc, a harness that forbids it). vr_world_diagnostics makes its
 * per-workgroup partials and its 14 results in device memory on its first call; this one keeps them in a fixed
 * 1 MiB array of the library's own on each GPU (held until the process ends), so it calls no hipMalloc or hipFree.
 * It takes worlds of at most 2,396,160 bodies (VR_ERR_ARGUMENT past that, before anything runs) and returns what
 * vr_world_diagnostics returns, word for word. Every call on one GPU shares that array, and nothing in it has to
 * start at zero: calls on worlds of the same GPU from two threads at once race on it, so a caller keeps them apart.
 */
int vr_world_diagnostics_static(vr_world *w, double *out, int32_t n);
/* Switches pass timing on or off and zeroes its counters. */
int vr_world_profile(vr_world *w, int32_t on);
int vr_world_profile_read(vr_world *w, vr_profile *out);

/*
 * The learned warm start's network (no counterpart in the reference): a per-contact MLP, 16 features -> 64 -> 64
 * -> 64 -> 3 outputs, SiLU on the hidden layers. 9,603 f32 parameters in this order: W1[16][64], b1[64],
 * W2[64][64], b2[64], W3[64][64], b3[64], W4[64][3], b4[3]. Inference is bf16 on the matrix cores with f32
 * accumulation; training is f32 with the bf16-rounded weights, mean squared error, Adam. All arrays are host arrays,
 * row-major (n x 16 features, n x 3 outputs). Errors: vr_mlp_last_error.
 */
typedef struct vr_mlp vr_mlp;
#define VR_MLP_FEATURES 16
#define VR_MLP_OUTPUTS 3
#define VR_MLP_PARAMS 9603
int vr_mlp_create(uint64_t seed, vr_mlp **out);
int vr_mlp_destroy(vr_mlp *m);
int vr_mlp_read(vr_mlp *m, float *params);
int vr_mlp_write(vr_mlp *m, const float *params);
int vr_mlp_forward(vr_mlp *m, const float *x, int64_t n, float *y);
/* The gradient of the mean squared error over the n points, and the loss, at the current parameters. */
int vr_mlp_gradient(vr_mlp *m, const float *x, const float *t, int64_t n, float *grad, float *loss);
/* Epochs of Adam over shuffled batches; *loss the last epoch's mean batch loss. */
int vr_mlp_train(vr_mlp *m, const float *x, const float *t, int64_t n, int32_t epochs, int32_t batch, float rate,
                 uint64_t seed, float *loss);
int vr_mlp_last_error(char *buf, int32_t len);
/*
 * The learned warm start in a box world (or batch), no counterpart in the reference, whose warm start is the carry
 * of the last step's impulses (core/.../box/BoxPairs.java:191-267). Per manifold point of each touching pair, 16
 * features are read after the carry: normal (3), relative velocity at the point (3), depth, the carried impulse as
 * a velocity (normal, and tangent in a basis of the contact plane: 3), the first body's share of the pair's
 * inverse mass, the pair's shapes one-hot (box-box, box-ball, ball-ball), the closing speed along the normal, and
 * the manifold's point count / 4. The 3 targets are the impulse the point holds after the step (what the next
 * step's carry would start it from), in those units. Logged targets are the impulse each point holds once the first
 * substep's solve sweeps are done: what a start would best begin that phase from.
 * Impulses (features 7-9, targets) are given as sign(v) log(1 + |v|) of the velocity v.
 * vr_box_log starts (or stops) logging (features, targets) of every step, up to `capacity` samples, on the device;
 * vr_box_samples copies up to `most` of them out. vr_box_learned sets the network (null: none) whose prediction
 * (where usable: finite, normal part not negative, and at a point the carry has nothing for: a new contact, or one
 * that moved past the carry's reach; a carried point keeps its impulse) is tried against the carried impulses at each
 * step's first substep:
 * the first solve sweep of each scene's pairs is run from each start and put back, and a scene takes the prediction
 * only where it leaves that sweep's largest residual smaller, the carry otherwise (vr_box_solve.fell_back counts the
 * scenes that kept it). The report's `reserved` word counts the points started from the prediction.
 * The world keeps only the network's handle: `m` must stay alive while it is set. vr_mlp_destroy refuses
 * (VR_ERR_STATE) a network that a world still holds; vr_box_learned(w, NULL) or vr_world_destroy lets it go.
 */
int vr_box_log(vr_world *w, int32_t on, int64_t capacity);
/*
 * Sweeps stopped on a residual (no counterpart in the reference, whose BoxStep runs `sweeps` sweeps of the solve and
 * of the relax every substep, BoxStep.java:199-214). A point's residual in a solve or relax sweep is the size of the
 * impulse change that sweep made there (normal and tangent together) times the inverse mass along its normal: how
 * far the sweep still moved the point's relative velocity, in m/s. With a tolerance of zero or more, each scene's
 * solve (and relax) sweeps of a substep end after the first sweep whose points all stay at or below it, or at the
 * step's `sweeps`; a scene's sweeps depend on its own residuals alone. Below zero (the default): every sweep runs, as
 * the reference runs them. With `measure`, each step keeps what its phases did (vr_box_solve_read) and each sweep's
 * largest residual (vr_box_solve_curve); a measured step computes every residual and runs slower, with the same
 * result. A NaN tolerance is VR_ERR_ARGUMENT.
 */
int vr_box_tolerance(vr_world *w, float tolerance, int32_t measure);
/*
 * The launch plan of a batch's steps (no counterpart in the reference, whose BoxStepper launches a step one way). Every
 * plan leaves every scene as the reference does; they differ in time only:
 *   VR_PLAN_PASS   kernel by kernel over the batch, a launch for each colour round;
 *   VR_PLAN_ALL    the same, every sweep of a stopped phase launched with no look at the residuals between them;
 *   VR_PLAN_GRAPH  VR_PLAN_ALL captured as a HIP graph each step and replayed;
 *   VR_PLAN_SCENE  a workgroup a scene through every substep, one launch a step;
 *   VR_PLAN_GROUP  VR_PLAN_PASS, with each sweep's colour rounds in one launch, a workgroup a group of consecutive
 *                  scenes.
 * A batch times its plans over its first steps after a start, a change of tolerance or vr_box_retune, two steps each
 * in turn, and keeps the fastest; VR_BOX_PLAN=pass|all|graph|scene|group fixes one instead. vr_box_plan reads the
 * plan the next step takes and whether the batch is still timing them (*tuning = 1); a world of one scene steps
 * VR_PLAN_PASS and never tunes.
 */
#define VR_PLAN_PASS 0
#define VR_PLAN_ALL 1
#define VR_PLAN_GRAPH 2
#define VR_PLAN_SCENE 3
#define VR_PLAN_GROUP 4
int vr_box_retune(vr_world *w);
int vr_box_plan(vr_world *w, int32_t *plan, int32_t *tuning);
#define VR_SOLVE_SUBSTEPS 64
/*
 * What the last step's phases did (the first VR_SOLVE_SUBSTEPS substeps; index 0 the solve sweeps, 1 the relax).
 * launched: sweeps the step launched, pass by pass (the most any scene needed; 0 when a workgroup steps each scene,
 * scene_step = 1). Measured steps only: scene_sweeps, the sweeps a scene ran on average; residual, the largest
 * residual a scene ended the phase with; residual_mean, their mean over the scenes. fell_back: scenes with pairs that
 * kept the carried start because the network's prediction did not leave their first solve sweep nearer its solution
 * (vr_box_learned).
 */
typedef struct vr_box_solve {
    float tolerance;
    int32_t measured;
    int32_t substeps;
    int32_t sweeps;
    int32_t scenes;
    int32_t scene_step;
    int32_t fell_back;
    int32_t unused;
    int32_t launched[2][VR_SOLVE_SUBSTEPS];
    float scene_sweeps[2][VR_SOLVE_SUBSTEPS];
    float residual[2][VR_SOLVE_SUBSTEPS];
    float residual_mean[2][VR_SOLVE_SUBSTEPS];  /* measured: the mean over the scenes of the residual each ended with */
} vr_box_solve;
int vr_box_solve_read(vr_world *w, vr_box_solve *out);
/*
 * The last measured step's largest residual over every point of each sweep, laid out [substep][phase][sweep] with
 * VR_SOLVE_SUBSTEPS substeps, 2 phases and VR_SOLVE_SWEEPS sweeps (zeros past what ran); up to n floats into out,
 * *count how many.
 */
#define VR_SOLVE_SWEEPS 64
int vr_box_solve_curve(vr_world *w, float *out, int32_t n, int32_t *count);
int vr_box_samples(vr_world *w, float *x, float *t, int64_t most, int64_t *count);
int vr_box_learned(vr_world *w, vr_mlp *m);

#ifdef __cplusplus
}
#endif

#endif
