// box_kernels.h — device layout of a box world and the launchers of its kernels.
//
// Body state is structure-of-arrays (x, y, z, u, v, w, lx, ly, lz, qx, qy, qz, qw, im, hx, hy, hz, shape), one f32
// array a field. Pairs use the reference's slot layout: 32 slots a body, slot p's pair (p / 32, partner) with a
// 40-float manifold and 28 floats of impulses (BoxBroad.java:46; BoxContact.java:77-88; BoxPairs.java:37).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "sphere_kernels.h"

namespace vrbox {

// One scene of a world: its walled box and its broad-phase grid, whose cells are numbered from cellBase in the
// world's cell order. A world of one scene is the reference's box world; a batch holds many, their bodies
// consecutive, stepped together (pairs never join bodies of two scenes).
struct SceneDev {
    int nx, ny, nz, cellBase;
    float cell, sx, sy, sz;
};

// The sweeps stopped on a residual (no counterpart in the reference, whose substeps run a fixed count of sweeps of the
// solve and of the relax). Each scene's solve (or relax) sweeps of a substep end after the first sweep whose points'
// residuals (box_math.h point_residual) are all at most the tolerance, or at the count the step allows; so a scene's
// state depends on its own residuals alone, never on how many sweeps the batch's other scenes needed. Sweeps are
// numbered over the world's life (`g`); a scene takes part in sweep g when g is the first of its phase or the scene
// had a point past the tolerance in sweep g - 1.
struct SweepCtl {
    int32_t *sceneOpen;            // scenes: the last sweep in which the scene had a point past the tolerance
    int32_t *lastOpen;             // 1: the last sweep in which any scene did
    // Measured steps: per scene (last sweep it took part in << 32 | bits of its largest residual there), the most;
    // per phase (substep * 2 + relax) the scenes' sweeps added up and their last residuals' largest, and per phase and
    // sweep the largest residual of any point.
    unsigned long long *sceneLast;
    int32_t *phaseSweeps;
    uint32_t *phaseResidual;
    float *phaseResidualSum;       // per phase, the scenes' last residuals added up (their mean, over the scenes)
    uint32_t *curve;
};

// What one sweep's launches are told: its number, whether it opens its phase, the tolerance, and where a measured
// sweep keeps its largest residual (-1: not measured).
struct SweepArgs {
    int g, first;
    float tol;
    int curveAt;
};

// Sweep modes: plain (the reference's fixed count), stopped on the residual, measured, or both.
enum { SWEEP_STOP = 1, SWEEP_MEASURE = 2 };
constexpr int SOLVE_MOST_SUBSTEPS = 64;   // phases measured: two a substep
constexpr int SOLVE_MOST_SWEEPS = 64;     // sweeps of a phase whose residuals a measured step keeps

struct BoxDev {
    int n, slots, width, cells;
    SweepCtl sw;
    int scenes;
    const int32_t *scene;       // n: each body's scene
    const SceneDev *sceneInfo;  // scenes
    float *x[3], *vel[3], *l[3], *q[4], *im, *half[3], *shape;
    const float *params;
    float *lambda, *rub, *roll, *shown;
    // n: for a box, bit wall * 8 + c set for every corner c holding an impulse at the wall (and maybe others); for a
    // ball, bit wall clear only when the wall's impulse words are quiet (nothing held, rolling words +0). All set when
    // not known.
    uint64_t *cornerHeld;
    // n: where body s's wall impulses sit: word w of body s is at w * n + wallPlace[s], wallPlace[s] the place of the
    // thread that sweeps the body (shape_first), so a wave's loads of one word take neighbouring words
    int32_t *wallPlace;
    int32_t *keys, *order, *starts;
    float4 *broad;  // 2 n: the broad phase's record of each place of the cell order (k_broad_records)
    // The slots with a partner, listed densely for the narrow phase (launch_collide): collideFirst (n) each body's
    // first place in collideList (slots), collideTotal (1) how many, collideScan the scan's scratch.
    int32_t *collideFirst, *collideList, *collideTotal, *collideScan;
    int32_t *pairs, *pairsWas, *contacts, *rounds;
    float *manifold, *manifoldWas, *impulses, *impulsesWas;
    // n words each, one a body, bit k for the body's slot k, set where the buffer beside it may hold a word other
    // than zero at that slot (its manifold's count; its impulses): a slot whose bit is clear already reads as an
    // empty slot is written, so the per-step passes over every slot leave it alone. They swap with their buffers.
    uint32_t *manifoldHeld, *manifoldHeldWas, *impulseHeld, *impulseHeldWas;
    // The touching pairs packed in round order for a step's sweeps (k_pack, then the rounds, then k_unpack), field by
    // field with stride packCap, so that a round's neighbouring lanes read neighbouring words. Entry j is the pair
    // at place j of the colouring's order: round r's pairs are entries starts[r] .. starts[r + 1] - 1.
    // n records each: a body's frame and motion over a run of sweeps of the pass-by-pass step (box_math.h), made from
    // the state by k_frame_in and the motion returned to it by k_motion_out.
    float4 *frame, *motion;
    int64_t packCap;
    // Within each round the entries are in the order of a sort by (the pair's scene group when a step's rounds run a
    // workgroup a group, then its point count and whether a ball takes part): a round's pairs share no body, so their
    // order within the round changes nothing, and a wave's lanes then mostly run the same points (PACK_CLASSES).
    int32_t *packFrom;     // the colouring's place (in C.order) whose pair the entry holds
    const int32_t *sceneGroup;  // scenes: each scene's group (Groups), for that sort
    int32_t *packPair;     // the pair's slot
    int32_t *packCount;    // its manifold's point count
    int32_t *packBody;     // 2 fields: its first body, then its second
    int32_t *packScene;    // the pair's scene
    float *packNormal;     // 3 fields
    float *packAnchor;     // 24 fields: field (side * MOST_POINTS + point) * 3 + axis, side 0 the first body's
    float *packImpulse;    // 28 fields: field point * POINT_IMPULSE + word
};

// The device's colouring of the touching pairs into rounds (box_colour.hip).
constexpr int COLOUR_FLAGS = 64;
enum { REPORT_POINTS = 0, REPORT_UNLISTED = 1, REPORT_WORDS = 2 };

struct Colouring {
    int most;                // slots: the most touching pairs there can be
    int32_t *colour;         // slots: a touching pair's round, MOST_ROUNDS if unplaced
    int32_t *decided;        // slots: the step of the colouring that decided it, INT_MAX while open
    int32_t *touch;          // slots: the touching pairs in slot order
    int32_t *touchSeg;       // slots / 256 + 1: touching pairs a workgroup of slots, scanned to where they go
    int32_t *touching;       // 1: how many touch
    int32_t *incDeg, *incPut;  // n: touching pairs at a body, and how many of them are placed
    int32_t *incOff;         // n + 1
    int32_t *inc;            // 2 * slots: each body's touching pairs
    int32_t *flags;          // COLOUR_FLAGS: flags[i % COLOUR_FLAGS] set when step i left a pair open
    int32_t *keys, *order;   // slots: the sort of the touching list by round
    int32_t *starts;         // MOST_ROUNDS + 2: where each round's pairs begin in that order
    int32_t *report;         // REPORT_WORDS: points, unlisted
    // slots / 256 + 1 each: the open pairs a colouring step left in each workgroup's segment of its open list (step
    // i reads segOpen[(i + 1) % 2], writes segOpen[i % 2]).
    int32_t *segOpen[2];
    int32_t *sumTmp;         // colour_tmp_words(slots)
    vr::SortScratch sort;    // sized for slots keys
};

// The learned warm start's per-point data (box_kernels.hip): each touching pair's manifold points, listed in the
// touching list's order (pointBase[t] the first of pair touch[t]), with 16 features read at the step's start (after
// the carry) and 3 targets, the impulses the point ends the step with. Impulses are scaled to velocities by the
// pair's inverse-mass sum and the tangent one written in a basis (t1, t2) of the plane across the normal.
constexpr int LEARN_FEATURES = 16, LEARN_OUTPUTS = 3;

struct Learned {
    // With fresh set (a network set, nothing logged), features and predictions are made only for the points the
    // carry leaves at zero, the only ones a prediction may start (launch_pack_prediction); pointBase then counts
    // those, and pointBase[touching] holds how many there are (the network's rows). Else every point has a row.
    int fresh;
    int32_t *pointBase;  // slots + 1
    float *x;            // 4 * slots * LEARN_FEATURES
    float *y;            // 4 * slots * LEARN_OUTPUTS: the network's prediction
    float *t;            // 4 * slots * LEARN_OUTPUTS: the targets
    int32_t *used;       // 1: points whose carried impulse the prediction replaced
    // The trial of the two starts in a step's first substep: the packed impulses of each start (28 fields of the
    // touching pairs), the motion records before the trial, each scene's first-sweep residual from each start, and
    // whether the scene takes the prediction.
    float *startCarry, *startPredicted;
    float4 *motionBefore;
    uint32_t *fromCarry, *fromPredicted;
    int32_t *takes;
};

// The packed impulses of the touching pairs (entries 0 .. touching - 1) into `to`, field by field, or back.
void launch_pack_save(const BoxDev &D, float *to, int touching, hipStream_t s);
void launch_pack_restore(const BoxDev &D, const float *from, int touching, hipStream_t s);
// The network's prediction (L.y) as the normal and tangent words of every packed point where it is usable.
void launch_pack_prediction(const BoxDev &D, const Colouring &C, const Learned &L, int touching, hipStream_t s);
// out[scene] = the scene's largest residual recorded in measured sweep g (0 where it recorded none).
void launch_trial_residual(const BoxDev &D, int g, uint32_t *out, hipStream_t s);
// Each scene takes the prediction where its first-sweep residual from it is smaller than from the carry (counting the
// scenes with pairs that keep the carry into fellBack), and its pairs' packed impulses are set from that start
// (counting the points started from the prediction into L.used).
void launch_pack_choose(const BoxDev &D, const Learned &L, int touching, int32_t *fellBack, hipStream_t s);

void launch_point_features(const BoxDev &D, const Colouring &C, const Learned &L, int touchingPairs, hipStream_t s);
void launch_point_targets(const BoxDev &D, const Colouring &C, const Learned &L, int touchingPairs, hipStream_t s);

int64_t colour_tmp_words(int64_t slots);
void launch_colour_prepare(const BoxDev &D, const Colouring &C, const vr::SortScratch &S, hipStream_t s);
void launch_colour_steps(const BoxDev &D, const Colouring &C, int touchingPairs, int from, int count,
                         hipStream_t s);
void launch_colour_finish(const BoxDev &D, const Colouring &C, int touchingPairs, hipStream_t s);

// With `motion`, also the motion records (box_math.h) from the state.
void launch_velocities(const BoxDev &D, bool motion, hipStream_t s);
// A world of one scene walls it with the params' box: copied into the scene before each step.
void launch_scene_from_params(const BoxDev &D, SceneDev *scene, hipStream_t s);
// With `frame`, also the moving bodies' frame records' position and turn.
void launch_positions(const BoxDev &D, bool frame, hipStream_t s);
void launch_walls(const BoxDev &D, int sweep, bool reverse, hipStream_t s, int mode = 0,
                  const SweepArgs &a = SweepArgs{0, 1, -1.f, -1});
// lastOpen = g if any scene is open in sweep g (what the host then reads).
void launch_any_open(const BoxDev &D, int g, hipStream_t s);
// At the end of a measured phase that began with sweep g0: each scene's sweeps and last residual added into `phase`.
void launch_phase_measure(const BoxDev &D, int g0, int phase, hipStream_t s);
// The wall sweep's places for the bodies' shapes into `place` (n words), as the sweep takes the bodies.
void launch_wall_place(const BoxDev &D, int32_t *place, hipStream_t s);
// Each body's wall impulse words moved from place `from[s]` to `to[s]` (both within the body's workgroup of bodies).
void launch_wall_move(const BoxDev &D, const int32_t *from, const int32_t *to, hipStream_t s);
// Every substep of a step (velocities, the warm, solve and relax sweeps, positions), a workgroup a scene, for a batch
// of scenes (k_scene_step). starts: the colouring's round starts (device); used: the step's rounds.
// mode and tol: the sweeps stopped on the residual and measured as the pass-by-pass step stops and measures them.
void launch_scene_step(const BoxDev &D, const int32_t *sceneFirst, const int32_t *starts, int used, int substeps,
                       int sweeps, bool symmetric, hipStream_t s, int mode = 0, float tol = -1.f);
// The classes of a round's pairs in the packed order: (points - 1) * 2, plus 1 when a ball takes part.
constexpr int PACK_CLASSES = 8;
// The packed order (packFrom) for the step's colouring, its sort's starts (one a round, group and class, of
// (MOST_ROUNDS + 1) * groups * PACK_CLASSES + 1) into `starts`; groups = 1 sorts by class alone within a round.
void launch_pack_order(const BoxDev &D, const Colouring &C, int touching, int groups, int32_t *starts,
                       hipStream_t s);

// Scene groups for a sweep's rounds a workgroup a group (k_group_rounds): runs of consecutive scenes, each group's
// scenes and, for the step's rounds, packed entries. Scenes share no body and no pair, so a group's rounds wait on
// nothing outside the group: they follow one another inside one workgroup, with a barrier where the launches over
// the whole batch have a kernel boundary.
constexpr int GROUP_MOST_THREADS = 1024;
struct Groups {
    int count;
    int mostBodies;   // bodies in the largest group
    int mostScenes;   // scenes in the group of most scenes
    int32_t *scenes;  // 2 a group: its first scene, and one past its last
    int32_t *bodies;  // 2 a group: its first body, and one past its last
    int32_t *order;   // n: places bodies[2g] .. bodies[2g + 1] - 1 hold group g's bodies, its boxes first
    int32_t *ranges;  // count * MOST_ROUNDS * 2: group g's packed entries in round r, [from, to)
    // The phase launches' order, one for the solve (and warm) phases and one for the relax phases: workgroup k of a
    // launch of kind f (0 solve, 1 relax) takes group launch[f * count + k]. Each solve or relax launch adds to
    // clocks[f * count + g] how long group g took (100 MHz ticks); the host orders the next step's launches of each
    // kind by the step's sums, longest first, and clears them.
    int32_t *launch;
    int64_t *clocks;
};
// Each group's packed entries in each of the step's rounds (starts: launch_pack_order's, sorted by group).
void launch_group_ranges(const BoxDev &D, const Groups &G, const int32_t *starts, int used, hipStream_t s);
// Rounds first .. last - 1 of a sweep (in the order of launch_solve_round's launches, so in reverse with `reverse`),
// a workgroup of `threads` a group.
// With `local`, each workgroup keeps its group's motion records in LDS over the rounds (MOTION_VECTORS float4 a body
// of the largest group's, at most GROUP_LOCAL_BODIES), read in at the start and written back at the end, and with
// the sweeps stopped (not measured) two words a scene of the group of most scenes, all within GROUP_LOCAL_BYTES.
constexpr int GROUP_LOCAL_BODIES = 5000;
constexpr size_t GROUP_LOCAL_BYTES = 160 * 1024 - 512;
// A group phase's dynamic LDS: the motion records (two float4 a body), two words a scene, and the walls list (a
// 16-bit place a body), for the group of most bodies and the group of most scenes.
inline size_t group_phase_lds(int mostBodies, int mostScenes) {
    return 32 * (size_t)mostBodies + 8 * (size_t)mostScenes + 2 * (((size_t)mostBodies + 1) / 2 * 2);
}
// A phase of `sweeps` sweeps numbered g0 ... (a warm sweep: one), each its walls and its `used` rounds in the order of
// launch_walls and launch_solve_round, a workgroup of `threads` a group through all of them, its motion records and
// its scenes' stop words in LDS (k_group_phase). mode and tol as for the launches; curveBase: where a measured
// phase keeps its sweeps' largest residuals (-1: not kept).
void launch_group_phase(const BoxDev &D, const Groups &G, int sweep, int used, int sweeps, bool symmetric, int g0,
                        float tol, int curveBase, int threads, hipStream_t s, int mode);
// Each group's bodies in `order`, its boxes first (each kind in body order), for the shapes as they are.
void launch_group_order(const BoxDev &D, const Groups &G, hipStream_t s);

void launch_group_rounds(const BoxDev &D, const Groups &G, int sweep, bool reverse, int first, int last, int threads,
                         hipStream_t s, int mode = 0, const SweepArgs &a = SweepArgs{0, 1, -1.f, -1},
                         bool local = false);
// Round r's pairs, packed entries base .. base + pairs - 1.
void launch_solve_round(const BoxDev &D, int sweep, bool reverse, int base, int pairs, hipStream_t s, int mode = 0,
                        const SweepArgs &a = SweepArgs{0, 1, -1.f, -1});
// The touching pairs (entries 0 .. touching - 1 of the colouring's order) into the packed fields, and the packed
// impulses back to the pairs' slots.
void launch_pack(const BoxDev &D, const Colouring &C, int touching, hipStream_t s);
// The frame records (and with `motion` the motion records) from the state; the motion records back to the state.
void launch_frame_in(const BoxDev &D, bool motion, hipStream_t s);
void launch_motion_out(const BoxDev &D, hipStream_t s);
void launch_unpack(const BoxDev &D, int touching, hipStream_t s);
void launch_bin(const BoxDev &D, hipStream_t s);
void launch_list(const BoxDev &D, hipStream_t s);
void launch_collide(const BoxDev &D, hipStream_t s);
void launch_collide_pairs(const BoxDev &D, const int32_t *pairs, int npairs, float *manifold, hipStream_t s);
void launch_carry_counts(const BoxDev &D, hipStream_t s);
void launch_show(const BoxDev &D, hipStream_t s);

}  // namespace vrbox
