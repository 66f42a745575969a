// box_world.hip — a box world behind the C ABI: its buffers by name and the step schedule.
//
// The schedule is the reference's BoxStepper.step (core/.../box/BoxStepper.java:56-100) over BoxStep's pass lists
// (core/.../box/BoxStep.java:138-153, 199-231): keep, bin and sort, list, collide, carry and counts; the touching
// pairs coloured into rounds as the reference's host colours them (first fit in slot order, at most 32; here on the
// device, box_colour.hip); then each substep velocities, the warm sweep, the solve sweeps, positions and the relax
// sweeps, each sweep every round of pairs then the walls.
#include "box_world.h"

#include <algorithm>
#include <chrono>
#include <climits>
#include <cstring>
#include <stdexcept>
#include <cstdlib>
#include <string>
#include <vector>

#include "box_kernels.h"
#include "box_math.h"
#include "diagnostics.h"
#include "mlp.h"
#include "sphere_kernels.h"

namespace vrbox {

static_assert(VR_SOLVE_SUBSTEPS == SOLVE_MOST_SUBSTEPS && VR_SOLVE_SWEEPS == SOLVE_MOST_SWEEPS,
              "the ABI's solve report and the device's phase words agree");

namespace {

#define HIP_OK(expr)                                                                                       \
    do {                                                                                                   \
        hipError_t e_ = (expr);                                                                            \
        if (e_ != hipSuccess) throw std::runtime_error(std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

}  // namespace

struct BoxWorld {
    vr_step_desc d{};
    int n = 0;
    bool grid = false;
    hipStream_t stream = nullptr;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    std::vector<void *> allocs;
    std::vector<std::pair<std::string, BoxBuf>> named;
    BoxDev D{};
    vr::SphereDev sortView{};  // the sphere kernels' grid sort reused: keys, order, starts over the same cells
    vr::SortScratch S{};
    std::vector<int32_t> sceneBodies;  // bodies a scene, consecutive
    std::vector<SceneDev> scenes;      // walls and grid a scene (cellBase worked out in build)
    bool wallsFromParams = true;       // one scene, walled by the params' box (the reference's box world)
    SceneDev *sceneDev = nullptr;
    int32_t *sceneFirst = nullptr;  // scenes + 1: each scene's first body
    int32_t *wallPlaceNext = nullptr;  // n: the wall places being worked out for newly written shapes
    bool sceneStep = false;         // substeps a workgroup a scene (k_scene_step)
    static constexpr size_t SCENE_STEP_LEAST = 64;  // scenes, for the scene step to fill the chip
    // Bodies in the largest scene, for one workgroup to keep up: a scene's sweep is a chain of dependent steps in
    // one workgroup, which beats launches over the whole chip for batches of small scenes (2,048 scenes of 16-64
    // bodies: 19 against 32 ms a step) but not once a scene holds thousands (16-4,096 bodies: 207 against 97).
    static constexpr int SCENE_STEP_MOST = 256;
    // How a step's substeps are launched. Every plan leaves every scene as the reference does; they differ in cost:
    //   scene:  a workgroup a scene through all substeps (k_scene_step);
    //   pass:   kernel by kernel over the whole batch, and with sweeps stopped on a residual the host looks at the
    //           device's last open sweep between batches of sweeps (one wait a phase at least);
    //   all:    the same launches, but every phase launches all of its sweeps and never waits (a stopped scene
    //           passes over the rest);
    //   graph:  `all`, captured as a HIP graph each step and replayed, the executable graph updated in place when
    //           the step's launches have the shape of the last;
    //   group:  `pass`, but a sweep's rounds in one launch, a workgroup a group of consecutive scenes through all of
    //           them (k_group_rounds), in place of a launch over the batch for each round.
    // A batch tunes it after each start and each change of tolerance: each candidate steps twice in turn, timed by
    // the step's own events, and the one with the smaller sum is kept. VR_BOX_PLAN=scene|pass|all|graph|group (or
    // the older VR_BOX_STEP_MODE=scene|pass) fixes it instead.
    enum Plan { PLAN_PASS, PLAN_ALL, PLAN_GRAPH, PLAN_SCENE, PLAN_GROUP };
    static_assert(PLAN_PASS == VR_PLAN_PASS && PLAN_ALL == VR_PLAN_ALL && PLAN_GRAPH == VR_PLAN_GRAPH &&
                      PLAN_SCENE == VR_PLAN_SCENE && PLAN_GROUP == VR_PLAN_GROUP,
                  "the ABI's plan numbers");
    Plan plan = PLAN_PASS;
    bool tunable = false;
    std::vector<Plan> candidates;
    std::vector<float> candidateMs;
    int tuneStep = 0;
    hipGraphExec_t graphExec = nullptr;
    int graphUpdates = 0, graphInstantiations = 0;
    // The scene groups of the group plan: consecutive scenes of at most groupBodies bodies together (a larger scene
    // alone), listed largest first (the phase launches reorder them, groupTimed), each swept by a workgroup of
    // groupThreads. With every group within
    // GROUP_LOCAL_BODIES, groups of up to 4,096 bodies on workgroups of 1,024 threads keep their motion records in LDS
    // over a sweep's rounds; else groups of up to 1,024 bodies on 512 threads read them from HBM (the faster of each
    // on the 2,048-scene batch: 61.6 and 79.1 ms of rounds a step, against 64-110 for the others tried). Each group's
    // walls in the same launch as its rounds, tried, took as long as the separate walls launch (88.3 against 61.7 +
    // 25.5 ms a step), so the walls stay a launch over the batch.
    // VR_BOX_GROUP_BODIES, VR_BOX_GROUP_THREADS and VR_BOX_GROUP_LOCAL=0|1 set them.
    // Rounds of more than groupTail pairs, and every round before them, are launched one by one over the batch
    // (VR_BOX_GROUP_TAIL); groupSplit: the first round of the step that goes to the groups.
    Groups G{};
    int groupBodies = 4096, groupThreads = 1024, groupTail = INT_MAX, groupSplit = 0;
    bool groupLocal = true;  // the groups' motion records in LDS over a sweep's rounds
    bool groupPhase = true;  // with groupLocal, a whole phase of sweeps, walls too, a launch (k_group_phase)
    // The phase launches take the groups longest first by what each took in the step before (G.clocks), not largest
    // first by bodies: a launch holds about two groups a CU, so the order decides how evenly the CUs finish.
    // VR_BOX_GROUP_TIMED=0 keeps the largest-first order.
    bool groupTimed = true;
    std::vector<int64_t> groupClock;
    std::vector<int32_t> groupLaunch;
    int32_t *packStarts = nullptr;  // the packed order's sort: where each (round, group, class) begins
    Colouring C{};
    // The learned warm start (no counterpart in the reference): per-point data, a log of (features, targets), and
    // the network whose prediction starts each step's impulses, when one is set.
    Learned L{};
    int64_t pointCap = 0;
    bool logging = false;
    float *logX = nullptr, *logT = nullptr;
    int64_t logCap = 0, logCount = 0;
    vrmlp::Net *net = nullptr;
    int32_t *learnedUsed = nullptr;
    int32_t *pinned = nullptr;  // host: the colouring's counts read back
    int colourSteps = 1;        // steps of the colouring the last step needed
    int roundPairs[MOST_ROUNDS] = {};  // pairs in each round this step
    int roundBase[MOST_ROUNDS] = {};   // its first entry in the colouring's order (the packed entries)
    // The sweeps stopped on a residual (vr_box_tolerance): off (every sweep runs) while tol is below zero.
    float tol = -1.f;
    bool measure = false;           // keep each phase's sweeps and residuals (vr_box_solve_read)
    int sweepNo = 0;                // sweeps numbered over the world's life (SweepCtl)
    std::vector<int> predicted;     // substeps x 2: sweeps each phase needed last step, launched before a look
    vr_box_solve lastSolve{};       // what the last step's phases did
    // Measurement (vr_world_profile): three events a solve or relax sweep, around its rounds and its walls.
    bool profiling = false;
    vr_profile prof{};
    std::vector<hipEvent_t> sweepEvents;
    std::vector<char> sweepReversed;
    std::vector<int> sweepWeight;  // the sweeps a timed launch holds
    int sweepsTimed = 0;
    int roundsThisStep = 0;
    int32_t *fellBack = nullptr;    // device: scenes with pairs that kept the carried start (learned_trial)

    template <class T>
    T *alloc(int64_t count) {
        void *p = nullptr;
        HIP_OK(hipMalloc(&p, (size_t)std::max<int64_t>(count, 1) * sizeof(T)));
        allocs.push_back(p);
        return (T *)p;
    }

    void reg(const char *name, void *p, int64_t count, int dtype, int bodyWords = 0) {
        named.push_back({name, BoxBuf{p, count, dtype, bodyWords}});
    }

    static int env_int(const char *name, int fallback, int least, int most) {
        const char *v = std::getenv(name);
        if (!v) return fallback;
        int k = std::atoi(v);
        return k < least || k > most ? fallback : k;
    }

    // The group plan's scene groups from the scenes' first bodies (first[k] .. first[k + 1] - 1 scene k's).
    void make_groups(const std::vector<int32_t> &first) {
        int largest = *std::max_element(sceneBodies.begin(), sceneBodies.end());
        groupLocal = env_int("VR_BOX_GROUP_LOCAL", 1, 0, 1) == 1 && largest <= GROUP_LOCAL_BODIES;
        groupBodies = groupLocal ? 4096 : 1024;
        groupThreads = groupLocal ? 1024 : 512;
        groupBodies = env_int("VR_BOX_GROUP_BODIES", groupBodies, 1, 1 << 30);
        if (groupLocal) groupBodies = std::min(groupBodies, GROUP_LOCAL_BODIES);
        groupThreads = env_int("VR_BOX_GROUP_THREADS", groupThreads, 64, GROUP_MOST_THREADS) / 64 * 64;
        groupTail = env_int("VR_BOX_GROUP_TAIL", groupTail, 0, INT_MAX);
        groupPhase = env_int("VR_BOX_GROUP_PHASE", 1, 0, 1) == 1;
        struct Group {
            int scene0, scene1, body0, body1;
        };
        std::vector<Group> groups;
        int scenesTotal = (int)sceneBodies.size();
        for (int k = 0; k < scenesTotal;) {
            Group g{k, k + 1, first[k], first[k + 1]};
            while (g.scene1 < scenesTotal && first[g.scene1 + 1] - g.body0 <= groupBodies) g.body1 = first[++g.scene1];
            groups.push_back(g);
            k = g.scene1;
        }
        std::stable_sort(groups.begin(), groups.end(),
                         [](const Group &a, const Group &b) { return a.body1 - a.body0 > b.body1 - b.body0; });
        std::vector<int32_t> groupOf((size_t)scenesTotal);
        for (size_t g = 0; g < groups.size(); g++)
            for (int k = groups[g].scene0; k < groups[g].scene1; k++) groupOf[(size_t)k] = (int32_t)g;
        G.count = (int)groups.size();
        G.ranges = alloc<int32_t>(2 * (int64_t)G.count * MOST_ROUNDS);
        std::vector<int32_t> bo, sc;
        G.mostBodies = G.mostScenes = 0;
        for (const Group &g : groups) {
            bo.push_back(g.body0), bo.push_back(g.body1);
            sc.push_back(g.scene0), sc.push_back(g.scene1);
            G.mostBodies = std::max(G.mostBodies, g.body1 - g.body0);
            G.mostScenes = std::max(G.mostScenes, g.scene1 - g.scene0);
        }
        G.bodies = alloc<int32_t>(2 * (int64_t)G.count);
        G.scenes = alloc<int32_t>(2 * (int64_t)G.count);
        // The first launches take the groups largest first; then each step's clocks order the next step's.
        groupTimed = env_int("VR_BOX_GROUP_TIMED", 1, 0, 1) == 1;
        groupLaunch.resize(2 * (size_t)G.count);  // the solve phases' order, then the relax phases'
        for (int g = 0; g < G.count; g++) groupLaunch[(size_t)g] = groupLaunch[(size_t)(G.count + g)] = g;
        groupClock.assign(2 * (size_t)G.count, 0);
        G.launch = alloc<int32_t>(2 * (int64_t)G.count);
        G.clocks = alloc<int64_t>(2 * (int64_t)G.count);
        HIP_OK(hipMemcpy(G.launch, groupLaunch.data(), sizeof(int32_t) * 2 * G.count, hipMemcpyHostToDevice));
        HIP_OK(hipMemset(G.clocks, 0, sizeof(int64_t) * 2 * G.count));
        HIP_OK(hipMemcpy(G.bodies, bo.data(), sizeof(int32_t) * bo.size(), hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(G.scenes, sc.data(), sizeof(int32_t) * sc.size(), hipMemcpyHostToDevice));
        size_t ldsBytes = sizeof(float4) * MOTION_VECTORS * (size_t)G.mostBodies + 2 * sizeof(int) * G.mostScenes;
        if (groupLocal && ldsBytes > GROUP_LOCAL_BYTES) groupLocal = false;
        groupPhase = groupPhase && groupLocal && group_phase_lds(G.mostBodies, G.mostScenes) <= GROUP_LOCAL_BYTES;
        G.order = alloc<int32_t>(n);  // filled for the shapes as they are (place_walls)
        int32_t *sceneGroup = alloc<int32_t>(scenesTotal);
        HIP_OK(hipMemcpy(sceneGroup, groupOf.data(), sizeof(int32_t) * groupOf.size(), hipMemcpyHostToDevice));
        D.sceneGroup = sceneGroup;
    }

    void build(const vr_step_desc &desc, std::vector<int32_t> bodies, std::vector<SceneDev> sceneList,
               bool fromParams) {
        d = desc;
        sceneBodies = std::move(bodies);
        scenes = std::move(sceneList);
        wallsFromParams = fromParams;
        if (d.bodies < 1 || d.substeps < 1 || d.sweeps < 1)
            throw std::invalid_argument("a step needs boxes, substeps and sweeps, got " + std::to_string(d.bodies) +
                                        ", " + std::to_string(d.substeps) + " and " + std::to_string(d.sweeps));
        n = d.bodies;
        grid = n > 1 && d.has_grid;
        for (const SceneDev &sc : scenes)
            if (grid && (sc.nx < 1 || sc.ny < 1 || sc.nz < 1 || !(sc.cell > 0)))
                throw std::invalid_argument("a grid of at least one cell, of positive size");
        int devices = 0;
        if (hipGetDeviceCount(&devices) != hipSuccess || devices < 1) throw std::runtime_error("no HIP device");
        HIP_OK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        HIP_OK(hipEventCreate(&t0));
        HIP_OK(hipEventCreate(&t1));
        D.n = n;
        if ((int64_t)n * SLOTS > INT32_MAX) throw std::invalid_argument("too many bodies for 32 slots each");
        D.slots = grid ? n * SLOTS : 0;
        // The scenes: each body's scene, and each scene's cells numbered after the scenes before it.
        int64_t cellTotal = 0;
        for (auto &sc : scenes) {
            sc.cellBase = (int)cellTotal;
            if (grid) cellTotal += (int64_t)sc.nx * sc.ny * sc.nz;
            if (cellTotal > INT32_MAX / 2) throw std::invalid_argument("too many broad-phase cells");
        }
        std::vector<int32_t> sceneOf(n);
        for (int k = 0, at = 0; k < (int)sceneBodies.size(); k++)
            for (int b = 0; b < sceneBodies[k]; b++) sceneOf[at++] = k;
        // A batch of many small scenes steps a workgroup a scene (k_scene_step); one world, or a few big scenes,
        // step pass by pass over the whole chip.
        std::vector<int32_t> first(sceneBodies.size() + 1, 0);
        int most = 0;
        for (size_t k = 0; k < sceneBodies.size(); k++) {
            first[k + 1] = first[k] + sceneBodies[k];
            most = std::max(most, sceneBodies[k]);
        }
        sceneStep = !wallsFromParams && sceneBodies.size() >= SCENE_STEP_LEAST && most <= SCENE_STEP_MOST;
        plan = sceneStep ? PLAN_SCENE : PLAN_PASS;
        // A lone scene small enough for one workgroup's LDS is tuned as a batch is, its one group and its one
        // workgroup among the candidates; the walls it reads from the params are copied into its scene record at
        // every step's start (launch_scene_from_params), where the scene and group kernels read them.
        bool loneSmall = sceneBodies.size() == 1 && n <= GROUP_LOCAL_BODIES;
        tunable = grid && (sceneBodies.size() > 1 || loneSmall);
        const char *fixed = std::getenv("VR_BOX_PLAN");
        if (!fixed) fixed = std::getenv("VR_BOX_STEP_MODE");
        if (fixed) {
            std::string m(fixed);
            if (m == "scene" && (!wallsFromParams || loneSmall)) plan = PLAN_SCENE;
            if (m == "pass") plan = PLAN_PASS;
            if (m == "all") plan = PLAN_ALL;
            if (m == "graph" && grid) plan = PLAN_GRAPH;
            if (m == "group" && grid && (!wallsFromParams || loneSmall)) plan = PLAN_GROUP;
            sceneStep = plan == PLAN_SCENE;
            tunable = false;
        }
        sceneFirst = alloc<int32_t>((int64_t)first.size());
        HIP_OK(hipMemcpy(sceneFirst, first.data(), sizeof(int32_t) * first.size(), hipMemcpyHostToDevice));
        int32_t *sceneIdx = alloc<int32_t>(n);
        sceneDev = alloc<SceneDev>((int64_t)scenes.size());
        HIP_OK(hipMemcpy(sceneIdx, sceneOf.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(sceneDev, scenes.data(), sizeof(SceneDev) * scenes.size(), hipMemcpyHostToDevice));
        D.scenes = (int)scenes.size();
        D.scene = sceneIdx;
        D.sceneInfo = sceneDev;
        if (grid && (sceneBodies.size() > 1 || loneSmall)) make_groups(first);
        if (plan == PLAN_GROUP && G.count == 0) plan = PLAN_PASS;
        D.width = std::max(1, n / 2);
        static const char *field[] = {"x",  "y",  "z",  "u",  "v",  "w",  "lx", "ly", "lz",
                                      "qx", "qy", "qz", "qw", "im", "hx", "hy", "hz", "shape"};
        float **where[] = {&D.x[0], &D.x[1], &D.x[2], &D.vel[0], &D.vel[1], &D.vel[2], &D.l[0], &D.l[1], &D.l[2],
                           &D.q[0], &D.q[1], &D.q[2], &D.q[3], &D.im, &D.half[0], &D.half[1], &D.half[2], &D.shape};
        int64_t stride = ((int64_t)n + 63) / 64 * 64;
        float *arena = alloc<float>(18 * stride);
        for (int k = 0; k < 18; k++) {
            *where[k] = arena + k * stride;
            reg(field[k], *where[k], n, VR_F32);
        }
        float *params = alloc<float>(PARAM_COUNT);
        D.params = params;
        reg("params", params, PARAM_COUNT, VR_F32);
        D.lambda = alloc<float>((int64_t)LAMBDA_WORDS * n);
        D.rub = alloc<float>((int64_t)RUB_WORDS * n);
        D.roll = alloc<float>((int64_t)RUB_WORDS * n);
        D.shown = alloc<float>((int64_t)SHOWN_STRIDE * n);
        D.sw.sceneOpen = alloc<int32_t>((int64_t)scenes.size());
        D.sw.lastOpen = alloc<int32_t>(1);
        D.sw.sceneLast = alloc<unsigned long long>((int64_t)scenes.size());
        D.sw.phaseSweeps = alloc<int32_t>(2 * SOLVE_MOST_SUBSTEPS);
        D.sw.phaseResidual = alloc<uint32_t>(2 * SOLVE_MOST_SUBSTEPS);
        D.sw.phaseResidualSum = alloc<float>(2 * SOLVE_MOST_SUBSTEPS);
        D.sw.curve = alloc<uint32_t>(2 * SOLVE_MOST_SUBSTEPS * SOLVE_MOST_SWEEPS);
        fellBack = alloc<int32_t>(1);
        predicted.assign((size_t)d.substeps * 2, d.sweeps);
        D.cornerHeld = alloc<uint64_t>(n);
        D.frame = alloc<float4>((int64_t)FRAME_VECTORS * n);
        D.motion = alloc<float4>((int64_t)MOTION_VECTORS * n);
        D.broad = alloc<float4>(2 * (int64_t)n);
        D.collideFirst = alloc<int32_t>(n);
        D.collideTotal = alloc<int32_t>(1);
        D.collideScan = alloc<int32_t>(vr::scan_tmp_words(n) + 64);
        D.wallPlace = alloc<int32_t>(n);
        wallPlaceNext = alloc<int32_t>(n);
        reg("lambda", D.lambda, (int64_t)LAMBDA_WORDS * n, VR_F32, LAMBDA_WORDS);
        reg("rub", D.rub, (int64_t)RUB_WORDS * n, VR_F32, RUB_WORDS);
        reg("roll", D.roll, (int64_t)RUB_WORDS * n, VR_F32, RUB_WORDS);
        reg("shown", D.shown, (int64_t)SHOWN_STRIDE * n, VR_F32);
        if (grid) {
            D.cells = (int)cellTotal;
            D.keys = alloc<int32_t>(n);
            D.order = alloc<int32_t>(n);
            D.starts = alloc<int32_t>((int64_t)D.cells + 1);
            reg("keys", D.keys, n, VR_I32);
            reg("order", D.order, n, VR_I32);
            reg("starts", D.starts, (int64_t)D.cells + 1, VR_I32);
            int64_t slots = D.slots;
            pack_make();
            D.pairs = alloc<int32_t>(2 * slots);
            D.collideList = alloc<int32_t>(slots);
            D.pairsWas = alloc<int32_t>(2 * slots);
            D.manifold = alloc<float>((int64_t)MANIFOLD_WORDS * slots);
            D.manifoldWas = alloc<float>((int64_t)MANIFOLD_WORDS * slots);
            D.impulses = alloc<float>((int64_t)IMPULSE_WORDS * slots);
            D.impulsesWas = alloc<float>((int64_t)IMPULSE_WORDS * slots);
            D.manifoldHeld = alloc<uint32_t>(n);
            D.manifoldHeldWas = alloc<uint32_t>(n);
            D.impulseHeld = alloc<uint32_t>(n);
            D.impulseHeldWas = alloc<uint32_t>(n);
            D.contacts = alloc<int32_t>(2 * slots + n);
            D.rounds = alloc<int32_t>((int64_t)MOST_ROUNDS * D.width);
            reg("contacts", D.contacts, 2 * slots + n, VR_I32);
            reg("rounds", D.rounds, (int64_t)MOST_ROUNDS * D.width, VR_I32);
            sortView.n = n;
            sortView.cells = D.cells;
            sortView.keys = D.keys;
            sortView.order = D.order;
            sortView.starts = D.starts;
            int nblocks = (n + 2047) / 2048;
            S.maxBlocks = nblocks;
            S.keysA = alloc<int32_t>(n);
            S.valsA = alloc<int32_t>(n);
            S.keysB = alloc<int32_t>(n);
            S.hist = alloc<int32_t>(256 * (int64_t)nblocks);
            S.scanTmp = alloc<int32_t>(std::max(vr::scan_tmp_words(256 * nblocks), vr::scan_tmp_words(n)) + 64);
            C.most = (int)slots;
            C.colour = alloc<int32_t>(slots);
            C.decided = alloc<int32_t>(slots);
            C.touch = alloc<int32_t>(slots);
            C.touchSeg = alloc<int32_t>(slots / 256 + 2);
            C.touching = alloc<int32_t>(1);
            C.incDeg = alloc<int32_t>(n);
            C.incPut = alloc<int32_t>(n);
            C.incOff = alloc<int32_t>((int64_t)n + 1);
            C.inc = alloc<int32_t>(2 * slots);
            C.flags = alloc<int32_t>(COLOUR_FLAGS);
            C.keys = alloc<int32_t>(slots);
            C.order = alloc<int32_t>(slots);
            C.starts = alloc<int32_t>(MOST_ROUNDS + 2);
            C.report = alloc<int32_t>(REPORT_WORDS);
            C.segOpen[0] = alloc<int32_t>(slots / 256 + 2);
            C.segOpen[1] = alloc<int32_t>(slots / 256 + 2);
            C.sumTmp = alloc<int32_t>(colour_tmp_words(slots));
            int sortBlocks = (int)((slots + 2047) / 2048);
            C.sort.maxBlocks = sortBlocks;
            C.sort.keysA = alloc<int32_t>(slots);
            C.sort.valsA = alloc<int32_t>(slots);
            C.sort.keysB = alloc<int32_t>(slots);
            C.sort.hist = alloc<int32_t>(256 * (int64_t)sortBlocks);
            C.sort.scanTmp = alloc<int32_t>(std::max(vr::scan_tmp_words(256 * sortBlocks), vr::scan_tmp_words((int)slots)) + 64);
        }
        HIP_OK(hipHostMalloc((void **)&pinned, 64 * sizeof(int32_t), hipHostMallocDefault));
        clear();
    }

    // Buffers whose address changes from step to step (the kept copies swap with the current ones).
    bool find(const std::string &name, BoxBuf *out) {
        for (auto &kv : named)
            if (kv.first == name) {
                *out = kv.second;
                return true;
            }
        if (grid) {
            int64_t slots = D.slots;
            if (name == "pairs") return *out = BoxBuf{D.pairs, 2 * slots, VR_I32}, true;
            if (name == "pairsWas") return *out = BoxBuf{D.pairsWas, 2 * slots, VR_I32}, true;
            if (name == "manifold") return *out = BoxBuf{D.manifold, (int64_t)MANIFOLD_WORDS * slots, VR_F32}, true;
            if (name == "manifoldWas")
                return *out = BoxBuf{D.manifoldWas, (int64_t)MANIFOLD_WORDS * slots, VR_F32}, true;
            if (name == "impulses") return *out = BoxBuf{D.impulses, (int64_t)IMPULSE_WORDS * slots, VR_F32}, true;
            if (name == "impulsesWas")
                return *out = BoxBuf{D.impulsesWas, (int64_t)IMPULSE_WORDS * slots, VR_F32}, true;
        }
        return false;
    }

    void clear() {  // PassRunner.clear(), then BoxStep.constants(): qw = 1 (BoxStep.java:103-105)
        tuneStep = 0;  // a new start: the launch plan is timed again
        sweep_control_clear();
        for (auto &kv : named) HIP_OK(hipMemsetAsync(kv.second.p, 0, (size_t)kv.second.count * 4, stream));
        if (grid) {
            size_t slots = D.slots;
            for (void *p : {(void *)D.pairs, (void *)D.pairsWas})
                HIP_OK(hipMemsetAsync(p, 0, slots * 2 * 4, stream));
            for (void *p : {(void *)D.manifold, (void *)D.manifoldWas})
                HIP_OK(hipMemsetAsync(p, 0, slots * MANIFOLD_WORDS * 4, stream));
            for (void *p : {(void *)D.impulses, (void *)D.impulsesWas})
                HIP_OK(hipMemsetAsync(p, 0, slots * IMPULSE_WORDS * 4, stream));
            for (void *p : {(void *)D.manifoldHeld, (void *)D.manifoldHeldWas, (void *)D.impulseHeld,
                            (void *)D.impulseHeldWas})
                HIP_OK(hipMemsetAsync(p, 0, (size_t)n * 4, stream));  // every slot zero
        }
        HIP_OK(hipMemsetAsync(D.cornerHeld, 0, (size_t)n * 8, stream));  // no impulse anywhere
        vr::launch_fill(D.q[3], 1.0f, n, stream);
        place_walls(false);  // every shape 0 (a box), every impulse 0
        HIP_OK(hipStreamSynchronize(stream));
    }

    // No scene open, no sweep counted, every phase predicted at the full count.
    void sweep_control_clear() {
        HIP_OK(hipMemsetAsync(D.sw.sceneOpen, 0, sizeof(int32_t) * scenes.size(), stream));
        HIP_OK(hipMemsetAsync(D.sw.lastOpen, 0, sizeof(int32_t), stream));
        HIP_OK(hipMemsetAsync(D.sw.sceneLast, 0, sizeof(unsigned long long) * scenes.size(), stream));
        HIP_OK(hipMemsetAsync(fellBack, 0, sizeof(int32_t), stream));
        sweepNo = 0;
        std::fill(predicted.begin(), predicted.end(), d.sweeps);
    }

    // vr_box_tolerance: the sweeps of each phase stop once a sweep's residuals are all at most `t` (below 0: the
    // reference's fixed count); with `m`, each step's phases are measured.
    void tolerance(float t, bool m) {
        if (t != t) throw std::invalid_argument("a tolerance is a number (below zero for none)");
        if (t != tol) tuneStep = 0;  // the plans' costs change with the stop: timed again
        tol = t;
        measure = m;
        HIP_OK(hipStreamSynchronize(stream));
        std::fill(predicted.begin(), predicted.end(), d.sweeps);
    }

    int sweep_mode() const { return (tol >= 0 ? SWEEP_STOP : 0) | (measure ? SWEEP_MEASURE : 0); }

    // Whether the host looks between a phase's sweeps (plan `pass`, and any step a plan without looks cannot run).
    bool looks() const { return plan == PLAN_PASS || plan == PLAN_SCENE || plan == PLAN_GROUP; }

    // The plan for the next step: the fixed one, or while tuning the next candidate in turn, then the fastest.
    void choose_plan() {
        if (!tunable) return;
        if (tuneStep == 0) {
            candidates.clear();
            candidates.push_back(PLAN_PASS);
            if (tol >= 0) candidates.push_back(PLAN_ALL);
            candidates.push_back(PLAN_GRAPH);
            candidates.push_back(PLAN_SCENE);
            if (G.count > 0) candidates.push_back(PLAN_GROUP);
            candidateMs.assign(candidates.size(), 0.f);
        }
        int rounds = 2, trials = rounds * (int)candidates.size();
        Plan next;
        if (tuneStep < trials) {
            next = candidates[tuneStep % candidates.size()];
        } else {
            size_t best = 0;
            for (size_t k = 1; k < candidates.size(); k++)
                if (candidateMs[k] < candidateMs[best]) best = k;
            next = candidates[best];
        }
        plan = next;
        scene_step(next == PLAN_SCENE);
    }

    // The plan the next step takes, and whether the batch is still timing its plans.
    Plan upcoming_plan() const {
        if (!tunable) return plan;
        if (tuneStep == 0) return PLAN_PASS;  // the first candidate of the timing about to begin
        int trials = 2 * (int)candidates.size();
        if (tuneStep < trials) return candidates[tuneStep % candidates.size()];
        size_t best = 0;
        for (size_t k = 1; k < candidates.size(); k++)
            if (candidateMs[k] < candidateMs[best]) best = k;
        return candidates[best];
    }
    bool tuning() const { return tunable && (tuneStep == 0 || tuneStep < 2 * (int)candidates.size()); }

    void plan_timed(float ms) {
        if (!tunable) return;
        int trials = 2 * (int)candidates.size();
        if (tuneStep < trials) candidateMs[tuneStep % candidates.size()] += ms;
        if (tuneStep <= trials) tuneStep++;
    }

    // The step's captured graph launched: the executable graph of the last step updated to it when it has the same
    // shape (the same kernels in the same order, only their arguments and grids changed), else made anew.
    void launch_graph(hipGraph_t graph) {
        if (graphExec) {
            hipGraphNode_t failed = nullptr;
            hipGraphExecUpdateResult result = hipGraphExecUpdateError;
            if (hipGraphExecUpdate(graphExec, graph, &failed, &result) == hipSuccess &&
                result == hipGraphExecUpdateSuccess) {
                graphUpdates++;
            } else {
                (void)hipGetLastError();
                HIP_OK(hipGraphExecDestroy(graphExec));
                graphExec = nullptr;
            }
        }
        if (!graphExec) {
            HIP_OK(hipGraphInstantiate(&graphExec, graph, nullptr, nullptr, 0));
            graphInstantiations++;
        }
        HIP_OK(hipGraphLaunch(graphExec, stream));
        HIP_OK(hipGraphDestroy(graph));
    }

    // A measured step's phase words start from zero.
    void measure_clear() {
        vr::launch_fill_i32(D.sw.phaseSweeps, 0, 2 * SOLVE_MOST_SUBSTEPS, stream);
        vr::launch_fill_i32(reinterpret_cast<int32_t *>(D.sw.phaseResidual), 0, 2 * SOLVE_MOST_SUBSTEPS, stream);
        vr::launch_fill(D.sw.phaseResidualSum, 0.f, 2 * SOLVE_MOST_SUBSTEPS, stream);
        vr::launch_fill_i32(reinterpret_cast<int32_t *>(D.sw.curve), 0, 2 * SOLVE_MOST_SUBSTEPS * SOLVE_MOST_SWEEPS,
                            stream);
    }

    // lastSolve from the step just run (the stream is idle).
    void solve_report() {
        vr_box_solve &r = lastSolve;
        r.tolerance = tol;
        r.measured = measure ? 1 : 0;
        r.substeps = d.substeps;
        r.sweeps = d.sweeps;
        r.scenes = (int32_t)scenes.size();
        r.scene_step = sceneStep ? 1 : 0;
        for (int k = 0; k < 2 * VR_SOLVE_SUBSTEPS; k++) {
            r.scene_sweeps[k / VR_SOLVE_SUBSTEPS][k % VR_SOLVE_SUBSTEPS] = 0.f;
            r.residual[k / VR_SOLVE_SUBSTEPS][k % VR_SOLVE_SUBSTEPS] = 0.f;
            r.residual_mean[k / VR_SOLVE_SUBSTEPS][k % VR_SOLVE_SUBSTEPS] = 0.f;
        }
        if (!measure) return;
        std::vector<int32_t> sweeps(2 * SOLVE_MOST_SUBSTEPS);
        std::vector<uint32_t> res(2 * SOLVE_MOST_SUBSTEPS);
        std::vector<float> sums(2 * SOLVE_MOST_SUBSTEPS);
        HIP_OK(hipMemcpyAsync(sums.data(), D.sw.phaseResidualSum, sizeof(float) * sums.size(), hipMemcpyDeviceToHost,
                              stream));
        HIP_OK(hipMemcpyAsync(sweeps.data(), D.sw.phaseSweeps, sizeof(int32_t) * sweeps.size(), hipMemcpyDeviceToHost,
                              stream));
        HIP_OK(hipMemcpyAsync(res.data(), D.sw.phaseResidual, sizeof(uint32_t) * res.size(), hipMemcpyDeviceToHost,
                              stream));
        HIP_OK(hipStreamSynchronize(stream));
        for (int sub = 0; sub < std::min(d.substeps, VR_SOLVE_SUBSTEPS); sub++)
            for (int k = 0; k < 2; k++) {
                r.scene_sweeps[k][sub] = (float)sweeps[2 * sub + k] / (float)std::max<size_t>(scenes.size(), 1);
                uint32_t bitsOf = res[2 * sub + k];
                float v;
                std::memcpy(&v, &bitsOf, sizeof v);
                r.residual[k][sub] = v;
                r.residual_mean[k][sub] = sums[2 * sub + k] / (float)std::max<size_t>(scenes.size(), 1);
            }
    }

    // The last measured step's largest residual of each sweep, [substep][solve, relax][sweep], into out.
    int32_t solve_curve(float *out, int32_t n) {
        int32_t words = 2 * SOLVE_MOST_SUBSTEPS * SOLVE_MOST_SWEEPS;
        std::vector<uint32_t> c((size_t)words);
        HIP_OK(hipMemcpyAsync(c.data(), D.sw.curve, sizeof(uint32_t) * words, hipMemcpyDeviceToHost, stream));
        HIP_OK(hipStreamSynchronize(stream));
        int32_t take = std::min(n, words);
        std::memcpy(out, c.data(), sizeof(float) * (size_t)take);
        return take;
    }

    // The wall sweep's places for the shapes as they are; with `move`, the wall impulse words follow their bodies.
    // A batch stepped a workgroup a scene keeps body order (the scene step sweeps bodies in order).
    // Either mode's places keep a body within its workgroup's block of places, so k_wall_move can take the words
    // from one to the other.
    void place_walls(bool move) {
        if (sceneStep) {
            wall_identity(wallPlaceNext);
        } else {
            launch_wall_place(D, wallPlaceNext, stream);
        }
        if (move) launch_wall_move(D, D.wallPlace, wallPlaceNext, stream);
        std::swap(D.wallPlace, wallPlaceNext);
        launch_group_order(D, G, stream);
        HIP_OK(hipStreamSynchronize(stream));
    }

    void wall_identity(int32_t *place) {
        std::vector<int32_t> id((size_t)n);
        for (int s = 0; s < n; s++) id[s] = s;
        HIP_OK(hipMemcpyAsync(place, id.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, stream));
        HIP_OK(hipStreamSynchronize(stream));
    }

    // Steps from now on a workgroup a scene (or pass by pass), the wall impulses moved to that mode's places.
    void scene_step(bool on) {
        if (on == sceneStep) return;
        sceneStep = on;
        place_walls(true);
    }

    // A manifold or impulse buffer written from outside: any slot may hold words other than zero.
    void pair_buffer_written(const std::string &name) {
        uint32_t *held = nullptr;
        if (name == "manifold") held = D.manifoldHeld;
        if (name == "manifoldWas") held = D.manifoldHeldWas;
        if (name == "impulses") held = D.impulseHeld;
        if (name == "impulsesWas") held = D.impulseHeldWas;
        if (!held) return;
        vr::launch_fill_i32(reinterpret_cast<int32_t *>(held), -1, n, stream);
        HIP_OK(hipStreamSynchronize(stream));
    }

    // Where each body's wall impulse words sit (host copy of D.wallPlace).
    void wall_places(std::vector<int32_t> &out) {
        out.resize((size_t)n);
        HIP_OK(hipMemcpyAsync(out.data(), D.wallPlace, sizeof(int32_t) * n, hipMemcpyDeviceToHost, stream));
        HIP_OK(hipStreamSynchronize(stream));
    }

    // The packed fields, made once at build for the most touching pairs there can be (one a slot), so no step
    // allocates.
    void pack_make() {
        int64_t cap = std::max<int64_t>(D.slots, 1);
        D.packFrom = alloc<int32_t>(cap);
        packStarts = alloc<int32_t>((int64_t)(MOST_ROUNDS + 1) * std::max(G.count, 1) * PACK_CLASSES + 1);
        D.packPair = alloc<int32_t>(cap);
        D.packScene = alloc<int32_t>(cap);
        D.packCount = alloc<int32_t>(cap);
        D.packBody = alloc<int32_t>(2 * cap);
        D.packNormal = alloc<float>(3 * cap);
        D.packAnchor = alloc<float>(2 * MOST_POINTS * 3 * cap);
        D.packImpulse = alloc<float>(MOST_POINTS * POINT_IMPULSE * cap);
        D.packCap = cap;
    }

    void pack_room(int touching) {
        if (touching > D.packCap)
            throw std::runtime_error("more touching pairs (" + std::to_string(touching) + ") than slots");
    }

    // The learned start's per-point data, made when a network is set or a log started (never inside a step), for
    // the most points there can be: four a slot.
    void points_make() {
        if (pointCap > 0 || !grid) return;
        int64_t cap = std::max<int64_t>((int64_t)MOST_POINTS * D.slots, 1);
        L.pointBase = alloc<int32_t>((int64_t)D.slots + 1);
        L.x = alloc<float>(cap * LEARN_FEATURES);
        L.y = alloc<float>(cap * LEARN_OUTPUTS);
        L.t = alloc<float>(cap * LEARN_OUTPUTS);
        L.used = alloc<int32_t>(1);
        int64_t packed = (int64_t)MOST_POINTS * POINT_IMPULSE * std::max<int64_t>(D.slots, 1);
        L.startCarry = alloc<float>(packed);
        L.startPredicted = alloc<float>(packed);
        L.motionBefore = alloc<float4>((int64_t)MOTION_VECTORS * n);
        L.fromCarry = alloc<uint32_t>((int64_t)scenes.size());
        L.fromPredicted = alloc<uint32_t>((int64_t)scenes.size());
        L.takes = alloc<int32_t>((int64_t)scenes.size());
        HIP_OK(hipMemsetAsync(L.used, 0, sizeof(int32_t), stream));
        HIP_OK(hipStreamSynchronize(stream));
        pointCap = cap;
    }

    void log(bool on, int64_t capacity) {
        if (on) points_make();
        logging = on;
        logCount = 0;
        if (on && capacity > logCap) {
            if (logX) HIP_OK(hipFree(logX));
            if (logT) HIP_OK(hipFree(logT));
            HIP_OK(hipMalloc((void **)&logX, sizeof(float) * capacity * LEARN_FEATURES));
            HIP_OK(hipMalloc((void **)&logT, sizeof(float) * capacity * LEARN_OUTPUTS));
            logCap = capacity;
        }
    }

    int64_t samples(float *x, float *t, int64_t most) {
        int64_t take = std::min(most, logCount);
        if (take > 0) {
            HIP_OK(hipMemcpyAsync(x, logX, (size_t)take * LEARN_FEATURES * 4, hipMemcpyDeviceToHost, stream));
            HIP_OK(hipMemcpyAsync(t, logT, (size_t)take * LEARN_OUTPUTS * 4, hipMemcpyDeviceToHost, stream));
            HIP_OK(hipStreamSynchronize(stream));
        }
        return take;
    }

    // Wall impulses written from outside: which corners hold one is not known, so every corner is looked at.
    void wall_impulses_written() {
        vr::launch_fill_i32(reinterpret_cast<int32_t *>(D.cornerHeld), -1, 2 * n, stream);
        HIP_OK(hipStreamSynchronize(stream));
    }

    double *diagPartials = nullptr, *diagOut = nullptr;

    void diagnostics(double *out) {  // BoxDiagnostics.of, reduced on the device
        if (!diagOut) {
            diagPartials = alloc<double>(vrdiag::partial_words(n));
            diagOut = alloc<double>(vrdiag::D_WORDS);
        }
        diagnostics_into(diagPartials, diagOut, out);
    }

    // The same into device scratch the caller provides (vrdiag::partial_words(n) partials, D_WORDS results).
    void diagnostics_into(double *partials, double *devOut, double *out) {
        vrdiag::launch_boxes(D, partials, devOut, stream);
        HIP_OK(hipMemcpyAsync(out, devOut, sizeof(double) * vrdiag::D_WORDS, hipMemcpyDeviceToHost, stream));
        HIP_OK(hipStreamSynchronize(stream));
    }

    int32_t read_word(const int32_t *p) {
        HIP_OK(hipMemcpyAsync(pinned, p, sizeof(int32_t), hipMemcpyDeviceToHost, stream));
        HIP_OK(hipStreamSynchronize(stream));
        return pinned[0];
    }

    // The touching pairs coloured into rounds, run until every pair has its round: the steps the last step needed,
    // then a few more at a time, reading back whether the last left a pair open.
    void colour(int *used, int *touching, int *points, int *unplaced, int *unlisted) {
        launch_colour_prepare(D, C, S, stream);
        int pairs = read_word(C.touching);
        int done = 0, batch = std::max(colourSteps, 1);
        while (pairs > 0) {
            launch_colour_steps(D, C, pairs, done, batch, stream);
            done += batch;
            if (read_word(C.flags + (done - 1) % COLOUR_FLAGS) == 0) break;
            if (done > pairs) throw std::runtime_error("colouring: pairs still open after a step for every pair");
            batch = 4;
        }
        colourSteps = std::max(done, 1);
        launch_colour_finish(D, C, pairs, stream);
        HIP_OK(hipMemcpyAsync(pinned, C.starts, (MOST_ROUNDS + 2) * sizeof(int32_t), hipMemcpyDeviceToHost, stream));
        HIP_OK(hipMemcpyAsync(pinned + 40, C.report, REPORT_WORDS * sizeof(int32_t), hipMemcpyDeviceToHost, stream));
        HIP_OK(hipStreamSynchronize(stream));
        *used = 0;
        for (int r = 0; r < MOST_ROUNDS; r++) {
            roundPairs[r] = pinned[r + 1] - pinned[r];
            roundBase[r] = pinned[r];
            if (roundPairs[r] > 0) *used = r + 1;
        }
        *touching = pairs;
        *unplaced = pinned[MOST_ROUNDS + 1] - pinned[MOST_ROUNDS];
        *points = pinned[40 + REPORT_POINTS];
        *unlisted = pinned[40 + REPORT_UNLISTED];
    }

    // Three events for the next timed sweep; layout 0: rounds, then walls; 1: walls, then rounds; 2 and 3: the same
    // with the rounds in one launch (k_group_rounds); 4: a phase of `weight` sweeps in one launch (k_group_phase).
    hipEvent_t *timed_sweep(int layout, int weight = 1) {
        size_t at = 3 * (size_t)sweepsTimed;
        while (sweepEvents.size() < at + 3) {
            hipEvent_t e;
            HIP_OK(hipEventCreate(&e));
            sweepEvents.push_back(e);
        }
        if (sweepReversed.size() <= (size_t)sweepsTimed) sweepReversed.resize(sweepsTimed + 1);
        sweepReversed[sweepsTimed] = (char)layout;
        if (sweepWeight.size() <= (size_t)sweepsTimed) sweepWeight.resize(sweepsTimed + 1);
        sweepWeight[sweepsTimed] = weight;
        sweepsTimed++;
        return &sweepEvents[at];
    }

    // BoxStep.sweep, BoxStep.java:217-231.
    void sweep(int kind, bool reverse, int rounds, int mode = 0, const SweepArgs &a = SweepArgs{0, 1, -1.f, -1}) {
        if (plan == PLAN_GROUP && groupPhase) {  // a warm sweep as a phase of one (the group plan's solve and relax
            launch_group_phase(D, G, kind, rounds, 1, false, a.g, a.tol, a.curveAt, groupThreads, stream, mode);
            return;                               // phases never come here)
        }
        int layout = (plan == PLAN_GROUP ? 2 : 0) + (reverse ? 1 : 0);
        hipEvent_t *ev = profiling && kind != WARM ? timed_sweep(layout) : nullptr;
        if (plan == PLAN_GROUP) {
            // Rounds before groupSplit launched one by one over the batch, the rest in one launch a workgroup a scene
            // group; the walls as below.
            int split = std::min(groupSplit, rounds);
            if (ev) HIP_OK(hipEventRecord(ev[0], stream));
            if (reverse) launch_walls(D, kind, reverse, stream, mode, a);
            if (ev && reverse) HIP_OK(hipEventRecord(ev[1], stream));
            for (int order = 0; order < split && !reverse; order++)
                launch_solve_round(D, kind, reverse, roundBase[order], roundPairs[order], stream, mode, a);
            launch_group_rounds(D, G, kind, reverse, split, rounds, groupThreads, stream, mode, a, groupLocal);
            for (int order = 0; order < split && reverse; order++) {
                int r = split - 1 - order;
                launch_solve_round(D, kind, reverse, roundBase[r], roundPairs[r], stream, mode, a);
            }
            if (ev && !reverse) HIP_OK(hipEventRecord(ev[1], stream));
            if (!reverse) launch_walls(D, kind, reverse, stream, mode, a);
            if (ev) HIP_OK(hipEventRecord(ev[2], stream));
            return;
        }
        if (ev) HIP_OK(hipEventRecord(ev[0], stream));
        if (reverse) launch_walls(D, kind, reverse, stream, mode, a);
        if (ev && reverse) HIP_OK(hipEventRecord(ev[1], stream));
        for (int order = 0; order < rounds; order++) {
            int r = reverse ? rounds - 1 - order : order;
            launch_solve_round(D, kind, reverse, roundBase[r], roundPairs[r], stream, mode, a);
            if (ev && roundPairs[r] > 0) roundsThisStep++;
        }
        if (ev && !reverse) HIP_OK(hipEventRecord(ev[1], stream));
        if (!reverse) launch_walls(D, kind, reverse, stream, mode, a);
        if (ev) HIP_OK(hipEventRecord(ev[2], stream));
    }

    // After a profiled step (the stream is idle): its sweeps' round and walls times into prof.
    void profile_step(int touching, int points) {
        int sweeps = 0;
        for (int k = 0; k < sweepsTimed; k++) {
            hipEvent_t *ev = &sweepEvents[3 * (size_t)k];
            float first = 0, second = 0;
            HIP_OK(hipEventElapsedTime(&first, ev[0], ev[1]));
            HIP_OK(hipEventElapsedTime(&second, ev[1], ev[2]));
            sweeps += sweepWeight[k];
            if (sweepReversed[k] == 4) {  // a phase in one launch, its walls and rounds together
                prof.sweep_ns += (int64_t)((double)first * 1e6);
                prof.sweep_launches++;
                continue;
            }
            bool reversed = (sweepReversed[k] & 1) != 0;
            float rounds = reversed ? second : first, walls = reversed ? first : second;
            prof.walls_launches++;
            if (sweepReversed[k] & 2) {
                prof.sweep_ns += (int64_t)((double)rounds * 1e6);
                prof.sweep_launches++;
            } else {
                prof.round_ns += (int64_t)((double)rounds * 1e6);
            }
            prof.walls_ns += (int64_t)((double)walls * 1e6);
        }
        prof.passes += sweeps;
        prof.round_launches += roundsThisStep;
        prof.listed_sum += (int64_t)touching * sweeps;
        prof.points_sum += (int64_t)points * sweeps;
        prof.bodies_sum += (int64_t)n * sweeps;
        prof.steps += 1;
        sweepsTimed = 0;
        roundsThisStep = 0;
    }

    // The solve (or relax) sweeps of substep `sub` (BoxStep.java:204-210): the configured count, or, stopped on the
    // residual, until every scene has stopped (each scene's own sweeps decided on the device, SweepCtl). The host
    // launches what the same phase needed last step, then looks at the device's last open sweep and launches two at
    // a time until none is open; a scene that has stopped passes over any sweep launched after it. Returns the
    // sweeps launched.
    int phase(int kind, int sub, int rounds) {
        int mode = sweep_mode();
        bool relax = kind == RELAX;
        int at = 2 * sub + (relax ? 1 : 0);
        bool kept = (mode & SWEEP_MEASURE) && sub < SOLVE_MOST_SUBSTEPS;
        int runMode = kept ? mode : (mode & SWEEP_STOP);
        int g0 = sweepNo + 1;
        auto one = [&](int k) {
            SweepArgs a{++sweepNo, k == 0 ? 1 : 0, tol,
                        kept && k < SOLVE_MOST_SWEEPS ? at * SOLVE_MOST_SWEEPS + k : -1};
            sweep(kind, d.symmetric && k % 2 == 1, rounds, runMode, a);
        };
        int launched = 0;
        if (plan == PLAN_GROUP && groupPhase) {  // every sweep of the phase in one launch, each group through them
            hipEvent_t *ev = profiling ? timed_sweep(4, d.sweeps) : nullptr;
            if (ev) HIP_OK(hipEventRecord(ev[0], stream));
            launch_group_phase(D, G, kind, rounds, d.sweeps, d.symmetric != 0, g0, tol,
                               kept ? at * SOLVE_MOST_SWEEPS : -1, groupThreads, stream, runMode);
            if (ev) HIP_OK(hipEventRecord(ev[1], stream));
            if (ev) HIP_OK(hipEventRecord(ev[2], stream));
            sweepNo += d.sweeps;
            launched = d.sweeps;
        } else if (!(mode & SWEEP_STOP) || !looks()) {
            for (; launched < d.sweeps; launched++) one(launched);
        } else {
            int want = std::min(std::max(predicted[at], 1), d.sweeps);
            int needed = d.sweeps;
            while (true) {
                for (; launched < want; launched++) one(launched);
                launch_any_open(D, sweepNo, stream);
                int open = read_word(D.sw.lastOpen);
                if (open < sweepNo) {  // no scene was open after the last sweep launched
                    needed = open < g0 ? 1 : std::min(open - g0 + 2, d.sweeps);
                    break;
                }
                if (launched >= d.sweeps) break;
                want = std::min(d.sweeps, launched + 2);
            }
            predicted[at] = needed;
        }
        if (kept) launch_phase_measure(D, g0, at, stream);
        return launched;
    }

    // The first solve sweep of the pairs alone (the warm sweep's rounds, then the solve sweep's, no walls), measured
    // as sweep g: what a start leaves each scene's pairs to do.
    void trial_sweep(int rounds, int g) {
        SweepArgs a{g, 1, -1.f, -1};
        if (plan == PLAN_GROUP) {  // the rounds of each in one launch, as the plan's sweeps run them
            launch_group_rounds(D, G, WARM, false, 0, rounds, groupThreads, stream, 0, a, groupLocal);
            launch_group_rounds(D, G, SOLVE, false, 0, rounds, groupThreads, stream, SWEEP_MEASURE, a, groupLocal);
            return;
        }
        for (int r = 0; r < rounds; r++) launch_solve_round(D, WARM, false, roundBase[r], roundPairs[r], stream);
        for (int r = 0; r < rounds; r++)
            launch_solve_round(D, SOLVE, false, roundBase[r], roundPairs[r], stream, SWEEP_MEASURE, a);
    }

    // The learned start, weighed against the carry by what each leaves the first solve sweep (no counterpart in the
    // reference, whose start is the carry, BoxPairs.java:191-267). From the first substep's motion records, the
    // sweep is tried once from the carried impulses and once from the prediction; each scene keeps the start whose
    // trial left its largest residual smaller (the carry on a tie), and the motion records and impulses are put back
    // as they were before the trial, with that start. Scenes share no pair, so a scene's choice rests on its own
    // pairs alone.
    void learned_trial(int rounds, int touching) {
        size_t motionBytes = sizeof(float4) * MOTION_VECTORS * (size_t)n;
        HIP_OK(hipMemcpyAsync(L.motionBefore, D.motion, motionBytes, hipMemcpyDeviceToDevice, stream));
        launch_pack_save(D, L.startCarry, touching, stream);
        int fromCarry = ++sweepNo;
        trial_sweep(rounds, fromCarry);
        launch_trial_residual(D, fromCarry, L.fromCarry, stream);
        HIP_OK(hipMemcpyAsync(D.motion, L.motionBefore, motionBytes, hipMemcpyDeviceToDevice, stream));
        launch_pack_restore(D, L.startCarry, touching, stream);
        launch_pack_prediction(D, C, L, touching, stream);
        launch_pack_save(D, L.startPredicted, touching, stream);
        int fromPrediction = ++sweepNo;
        trial_sweep(rounds, fromPrediction);
        launch_trial_residual(D, fromPrediction, L.fromPredicted, stream);
        HIP_OK(hipMemcpyAsync(D.motion, L.motionBefore, motionBytes, hipMemcpyDeviceToDevice, stream));
        launch_pack_choose(D, L, touching, fellBack, stream);
    }

    // The next step's solve and relax launches take the groups longest first by the time each took in this step's
    // launches of the same kind, summed (ties in the present order); the sums then start again from zero. A stopped
    // phase's groups take very different times in the solve and the relax, hence an order for each. Which workgroup
    // takes which group changes no result: groups share no body.
    void order_groups() {
        HIP_OK(hipMemcpy(groupClock.data(), G.clocks, sizeof(int64_t) * 2 * G.count, hipMemcpyDeviceToHost));
        for (int f = 0; f < 2; f++) {
            const int64_t *clock = groupClock.data() + (size_t)f * G.count;
            if (*std::max_element(clock, clock + G.count) <= 0) continue;  // no launch of this kind ran
            std::stable_sort(groupLaunch.begin() + (size_t)f * G.count, groupLaunch.begin() + (size_t)(f + 1) * G.count,
                             [&](int32_t a, int32_t b) { return clock[a] > clock[b]; });
        }
        HIP_OK(hipMemcpy(G.launch, groupLaunch.data(), sizeof(int32_t) * 2 * G.count, hipMemcpyHostToDevice));
        HIP_OK(hipMemsetAsync(G.clocks, 0, sizeof(int64_t) * 2 * G.count, stream));  // before the next launches
    }

    // The counts BoxStepper.step reports for the step's collide (BoxStepper.java:40, 62-90).
    struct Prepared {
        int used = 0, touching = 0, points = 0, unplaced = 0, unlisted = 0;
    };

    // BoxStepper.step up to the rounds' write (BoxStepper.java:62-95): BoxStep.collide (keep, bin and sort, list,
    // collide, carry, counts; BoxStep.java:138-153), the touching pairs coloured into rounds, and, with a network
    // set, its prediction in place of the carried impulses.
    Prepared prepare() {
        Prepared p;
        if (wallsFromParams) launch_scene_from_params(D, sceneDev, stream);
        if (grid) {
            // BoxPairs.keep: the step's pairs, manifolds and impulses become the kept ones, which carry reads.
            std::swap(D.pairs, D.pairsWas);
            std::swap(D.manifold, D.manifoldWas);
            std::swap(D.impulses, D.impulsesWas);
            std::swap(D.manifoldHeld, D.manifoldHeldWas);
            std::swap(D.impulseHeld, D.impulseHeldWas);
            launch_bin(D, stream);
            vr::launch_sort(sortView, S, stream);
            launch_list(D, stream);
            launch_collide(D, stream);
            launch_carry_counts(D, stream);
            // BoxStepper's colouring (first fit in slot order), on the device; the host reads back the counts.
            colour(&p.used, &p.touching, &p.points, &p.unplaced, &p.unlisted);
            if ((net || logging) && p.touching > 0) {
                // A logged step takes every point's features; a learned one only the points the carry left at zero,
                // the only ones its prediction may start, their count known on the device.
                L.fresh = net && !logging ? 1 : 0;
                launch_point_features(D, C, L, p.touching, stream);
                if (net) net->forward(L.x, p.points, L.y, stream, L.fresh ? L.pointBase + p.touching : nullptr);
            }
        }
        return p;
    }

    void report_counts(const Prepared &p, vr_step_report *rep) {
        rep->rounds = p.used;
        rep->most_rounds = p.used;
        rep->touching = p.touching;
        rep->points = p.points;
        rep->unplaced = p.unplaced;
        rep->unlisted = p.unlisted;
    }

    // The collide list alone, read back and coloured, with no substep run: what a step's impulses start from.
    void collide_list(vr_step_report *rep) {
        auto w0 = std::chrono::steady_clock::now();
        HIP_OK(hipEventRecord(t0, stream));
        Prepared p = prepare();
        HIP_OK(hipEventRecord(t1, stream));
        HIP_OK(hipStreamSynchronize(stream));
        float ms = 0;
        HIP_OK(hipEventElapsedTime(&ms, t0, t1));
        report_counts(p, rep);
        rep->submissions = grid ? 1 : 0;
        rep->gpu_ns = (int64_t)(ms * 1e6);
        rep->wall_ns =
            std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - w0).count();
    }

    void step(vr_step_report *rep) {
        choose_plan();
        auto w0 = std::chrono::steady_clock::now();
        HIP_OK(hipEventRecord(t0, stream));
        Prepared p = prepare();
        int used = p.used, touching = p.touching, points = p.points;
        if (sweep_mode() & SWEEP_MEASURE) measure_clear();
        std::fill(lastSolve.launched[0], lastSolve.launched[0] + 2 * VR_SOLVE_SUBSTEPS, 0);
        bool logged = logging && grid && touching > 0;
        bool trial = net && grid && touching > 0 && used > 0;
        // A logged step takes its targets between phases, and a learned start is tried in the first substep: pass
        // by pass.
        if (sceneStep && !logged && !trial) {
            launch_scene_step(D, sceneFirst, C.starts, used, d.substeps, d.sweeps, d.symmetric != 0, stream,
                              sweep_mode(), tol);
        } else {
            // A step with no host look and no host read can be captured: graph plan, nothing logged, tried, measured
            // or timed sweep by sweep.
            bool captured = plan == PLAN_GRAPH && !logged && !trial && !measure && !profiling;
            if (captured) HIP_OK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
            if (used > 0) {  // the sweeps' rounds read and write the touching pairs packed in round order
                pack_room(touching);
                bool grouped = plan == PLAN_GROUP;
                launch_pack_order(D, C, touching, grouped ? G.count : 1, packStarts, stream);
                launch_pack(D, C, touching, stream);
                if (grouped) {
                    // The rounds from the first of the tail (every round from it on of at most groupTail pairs) on
                    // go to the groups.
                    groupSplit = used;
                    while (groupSplit > 0 && roundPairs[groupSplit - 1] <= groupTail) groupSplit--;
                    launch_group_ranges(D, G, packStarts, used, stream);
                    // The walls' order of each group's bodies, from where the bodies start the step and what they
                    // hold (k_group_order): the ones with something to do at a wall kept together.
                    if (groupPhase) launch_group_order(D, G, stream);
                }
            }
            for (int sub = 0; sub < d.substeps; sub++) {  // the sweeps work on the frame and motion records
                // The frame records are made whole at the first substep; later only positions change them.
                launch_velocities(D, sub > 0, stream);
                if (sub == 0) launch_frame_in(D, true, stream);
                if (sub == 0 && trial) learned_trial(used, touching);
                sweep(WARM, false, used);
                int solved = phase(SOLVE, sub, used);
                launch_motion_out(D, stream);
                if (sub == 0 && logged) {
                    // The targets: what each point holds once the first substep's solve sweeps are done, the
                    // impulse a start would best begin that phase from.
                    if (used > 0) launch_unpack(D, touching, stream);
                    launch_point_targets(D, C, L, touching, stream);
                }
                launch_positions(D, true, stream);
                int relaxed = phase(RELAX, sub, used);
                launch_motion_out(D, stream);
                if (sub < VR_SOLVE_SUBSTEPS) {
                    lastSolve.launched[0][sub] = solved;
                    lastSolve.launched[1][sub] = relaxed;
                }
            }
            if (used > 0) launch_unpack(D, touching, stream);
            if (captured) {
                hipGraph_t graph = nullptr;
                HIP_OK(hipStreamEndCapture(stream, &graph));
                launch_graph(graph);
            }
        }
        if (logged) {
            int64_t take = std::min<int64_t>(points, logCap - logCount);
            if (take > 0) {
                HIP_OK(hipMemcpyAsync(logX + logCount * LEARN_FEATURES, L.x, (size_t)take * LEARN_FEATURES * 4,
                                      hipMemcpyDeviceToDevice, stream));
                HIP_OK(hipMemcpyAsync(logT + logCount * LEARN_OUTPUTS, L.t, (size_t)take * LEARN_OUTPUTS * 4,
                                      hipMemcpyDeviceToDevice, stream));
                logCount += take;
            }
        }
        launch_show(D, stream);
        HIP_OK(hipEventRecord(t1, stream));
        HIP_OK(hipStreamSynchronize(stream));
        float ms = 0;
        HIP_OK(hipEventElapsedTime(&ms, t0, t1));
        plan_timed(ms);
        if (groupPhase && groupTimed && plan == PLAN_GROUP && G.count > 1) order_groups();
        rep->reserved = 0;
        lastSolve.fell_back = 0;
        if (net && grid && touching > 0) {
            HIP_OK(hipMemcpyAsync(pinned + 48, L.used, sizeof(int32_t), hipMemcpyDeviceToHost, stream));
            HIP_OK(hipMemcpyAsync(pinned + 49, fellBack, sizeof(int32_t), hipMemcpyDeviceToHost, stream));
            HIP_OK(hipMemsetAsync(L.used, 0, sizeof(int32_t), stream));
            HIP_OK(hipMemsetAsync(fellBack, 0, sizeof(int32_t), stream));
            HIP_OK(hipStreamSynchronize(stream));
            rep->reserved = pinned[48];
            lastSolve.fell_back = pinned[49];
        }
        solve_report();
        if (profiling) profile_step(touching, points);
        report_counts(p, rep);
        rep->submissions = grid ? 2 : 1;  // BoxStepper.step: the collide list read back, then the step's list
        rep->gpu_ns = (int64_t)(ms * 1e6);
        rep->wall_ns =
            std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - w0).count();
    }

    void collide_pairs(const int32_t *pairs, int npairs, float *manifold) {
        // Checked before anything is allocated or launched: the kernel loads both bodies of every pair unguarded.
        if (npairs < 0) throw std::invalid_argument("a count of pairs is at least 0, got " + std::to_string(npairs));
        for (int64_t k = 0; k < 2 * (int64_t)npairs; k++)
            if (pairs[k] < 0 || pairs[k] >= n)
                throw std::invalid_argument("pair " + std::to_string(k / 2) + " names body " +
                                            std::to_string(pairs[k]) + ", outside 0 .. " + std::to_string(n - 1));
        int32_t *dp = nullptr;
        float *dm = nullptr;
        HIP_OK(hipMalloc(&dp, (size_t)std::max(1, 2 * npairs) * 4));
        HIP_OK(hipMalloc(&dm, (size_t)std::max(1, npairs) * MANIFOLD_WORDS * 4));
        HIP_OK(hipMemcpyAsync(dp, pairs, (size_t)2 * npairs * 4, hipMemcpyHostToDevice, stream));
        if (npairs > 0) launch_collide_pairs(D, dp, npairs, dm, stream);
        HIP_OK(hipMemcpyAsync(manifold, dm, (size_t)npairs * MANIFOLD_WORDS * 4, hipMemcpyDeviceToHost, stream));
        HIP_OK(hipStreamSynchronize(stream));
        HIP_OK(hipFree(dp));
        HIP_OK(hipFree(dm));
    }

    ~BoxWorld() {
        if (net) net->holders--;  // the network may be destroyed once no world starts from it
        if (stream) (void)hipStreamSynchronize(stream);
        if (pinned) (void)hipHostFree(pinned);
        for (void *p : {(void *)logX, (void *)logT})
            if (p) (void)hipFree(p);
        for (void *p : allocs) (void)hipFree(p);
        for (hipEvent_t e : sweepEvents) (void)hipEventDestroy(e);
        if (graphExec) (void)hipGraphExecDestroy(graphExec);
        if (t0) (void)hipEventDestroy(t0);
        if (t1) (void)hipEventDestroy(t1);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

BoxWorld *box_create(const vr_step_desc &desc) {
    SceneDev one{desc.nx, desc.ny, desc.nz, 0, desc.cell, 1.0f, 1.0f, 1.0f};
    auto *w = new BoxWorld();
    try {
        w->build(desc, {desc.bodies}, {one}, true);
    } catch (...) {
        delete w;
        throw;
    }
    return w;
}

BoxWorld *box_create_batch(const vr_step_desc &desc, int32_t scenes, const int32_t *bodies, const float *extents,
                           const int32_t *grids, const float *cells) {
    if (scenes < 1) throw std::invalid_argument("a batch needs a scene, got " + std::to_string(scenes));
    if (!desc.has_grid) throw std::invalid_argument("a batch steps with its scenes' grids");
    std::vector<int32_t> counts(bodies, bodies + scenes);
    std::vector<SceneDev> list(scenes);
    int64_t total = 0;
    for (int k = 0; k < scenes; k++) {
        if (counts[k] < 1) throw std::invalid_argument("scene " + std::to_string(k) + " has no bodies");
        const float *e = extents + 3 * k;
        const int32_t *g = grids + 3 * k;
        if (!(e[0] > 0 && e[1] > 0 && e[2] > 0))
            throw std::invalid_argument("scene " + std::to_string(k) + " needs a box of positive size");
        if (g[0] < 1 || g[1] < 1 || g[2] < 1 || !(cells[k] > 0))
            throw std::invalid_argument("scene " + std::to_string(k) + " needs a grid of at least one cell");
        list[k] = SceneDev{g[0], g[1], g[2], 0, cells[k], e[0], e[1], e[2]};
        total += counts[k];
    }
    if (total != desc.bodies)
        throw std::invalid_argument("the scenes hold " + std::to_string(total) + " bodies, the step " +
                                    std::to_string(desc.bodies));
    auto *w = new BoxWorld();
    try {
        w->build(desc, std::move(counts), std::move(list), false);
    } catch (...) {
        delete w;
        throw;
    }
    return w;
}

void box_destroy(BoxWorld *w) { delete w; }
void box_clear(BoxWorld *w) { w->clear(); }
bool box_find(BoxWorld *w, const std::string &name, BoxBuf *out) { return w->find(name, out); }
hipStream_t box_stream(BoxWorld *w) { return w->stream; }

void box_params(BoxWorld *w, const int32_t *words, int32_t n) {
    if (n != PARAM_COUNT) throw std::invalid_argument("box parameters are 18 words");
    HIP_OK(hipMemcpyAsync((void *)w->D.params, words, sizeof(int32_t) * PARAM_COUNT, hipMemcpyHostToDevice,
                          w->stream));
    HIP_OK(hipStreamSynchronize(w->stream));
}

void box_step(BoxWorld *w, vr_step_report *rep) { w->step(rep); }
void box_collide_list(BoxWorld *w, vr_step_report *rep) { w->collide_list(rep); }
void box_tolerance(BoxWorld *w, float tol, bool measure) { w->tolerance(tol, measure); }
void box_retune(BoxWorld *w) { w->tuneStep = 0; }
void box_plan(BoxWorld *w, int32_t *plan, int32_t *tuning) {
    *plan = (int32_t)w->upcoming_plan();
    *tuning = w->tuning() ? 1 : 0;
}
void box_profile(BoxWorld *w, bool on) {
    w->profiling = on;
    w->prof = vr_profile{};
}
void box_profile_read(BoxWorld *w, vr_profile *out) { *out = w->prof; }
void box_solve_read(BoxWorld *w, vr_box_solve *out) { *out = w->lastSolve; }
int32_t box_solve_curve(BoxWorld *w, float *out, int32_t n) { return w->solve_curve(out, n); }
void box_wall_impulses_written(BoxWorld *w) { w->wall_impulses_written(); }
void box_shapes_written(BoxWorld *w) {
    w->place_walls(true);
    // cornerHeld is kept for boxes only: a body that was a ball has none, so after new shapes every corner is looked at.
    w->wall_impulses_written();
}
void box_wall_places(BoxWorld *w, std::vector<int32_t> &out) { w->wall_places(out); }
void box_pair_buffer_written(BoxWorld *w, const std::string &name) { w->pair_buffer_written(name); }
void box_log(BoxWorld *w, bool on, int64_t capacity) { w->log(on, capacity); }
void box_diagnostics(BoxWorld *w, double *out) { w->diagnostics(out); }
void box_diagnostics_into(BoxWorld *w, double *partials, double *devOut, double *out) {
    w->diagnostics_into(partials, devOut, out);
}
int box_bodies(BoxWorld *w) { return w->n; }
int64_t box_samples(BoxWorld *w, float *x, float *t, int64_t most) { return w->samples(x, t, most); }
void box_learned(BoxWorld *w, vrmlp::Net *net) {
    if (net) w->points_make();
    if (net) net->holders++;
    if (w->net) w->net->holders--;
    w->net = net;
}

void box_collide(BoxWorld *w, const int32_t *pairs, int32_t npairs, float *manifold) {
    w->collide_pairs(pairs, npairs, manifold);
}

}  // namespace vrbox
