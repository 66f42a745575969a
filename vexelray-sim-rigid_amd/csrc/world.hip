// world.hip — the C ABI (include/vexelray_rigid.h) over the sphere kernels: buffers by name, and the step schedule.
//
// The schedule is the reference's: SphereStep's pass lists (core/.../sphere/SphereStep.java:184-208, 216-262,
// 348-392) run by SphereStepper.step (core/.../sphere/SphereStepper.java:66-103), every pass a kernel on the world's
// stream. A Gauss-Seidel pass runs until no contact is left open: its opening batch of rounds, then batches of
// SphereStep.BATCH, with the open word read back after each batch, and the opening batch adapted as the reference's
// stepper adapts it (SphereStepper.java:118-126), so the rounds run, and the report, are the reference's.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "box_world.h"
#include "diagnostics.h"
#include "mlp.h"
#include "sphere_kernels.h"
#include "vexelray_rigid.h"

namespace {

using namespace vr;

thread_local std::string g_error;

struct ArgError : std::runtime_error { using std::runtime_error::runtime_error; };
struct StateError : std::runtime_error { using std::runtime_error::runtime_error; };
struct NameError : std::runtime_error { using std::runtime_error::runtime_error; };
struct DeviceError : std::runtime_error { using std::runtime_error::runtime_error; };

#define HIP_OK(expr)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess) throw DeviceError(std::string(#expr) + ": " + hipGetErrorString(e_));    \
    } while (0)

constexpr int BATCH = 8, CYCLE = 48, SHRINK_AFTER = 120;  // SphereStep.java:48-59, SphereStepper.java:49
constexpr int MOST_LISTED_SPHERES = 46340;                 // Spheres.java:787

struct Buf {
    void *p;
    int64_t count;
    int dtype;
};

}  // namespace

struct vr_world {
    vrbox::BoxWorld *box = nullptr;  // set for a box world, which the calls below hand on to
    int device = 0;                  // the GPU current when the world was made, whose stream it steps on
    vr_step_desc d{};
    int n = 0, cap = 0, cells = 0;
    bool gs = false, grid = false, friction = true;
    hipStream_t stream = nullptr;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    std::vector<void *> allocs;
    std::map<std::string, Buf> fixed;  // buffers whose address never changes
    SphereDev D{};
    SortScratch S{};
    ContactList lists[2]{};
    int par = 0;  // lists[par] is this substep's list, lists[1 - par] the one kept from the substep before
    int32_t *control = nullptr;  // device: contactCount[3], open[2]
    int32_t *pinned = nullptr;   // host copy of control
    std::vector<int> opening, enough;  // the host stepper's (fixed-rounds and reference-wait modes)
    PassCtl *ctl = nullptr;            // device: the passes' shared control words and the step's report
    PassSlot *slots = nullptr;         // device: one per pass of a step
    int nslots = 0;
    PassCtl *ctlHost = nullptr;        // pinned copy read after a step
    OpenLists openLists{};
    int32_t *devOpening = nullptr, *devEnough = nullptr;  // SphereStepper's opening batches, kept on the device
    std::vector<int> wide, wideRub;    // wide rounds to launch per pass, from what the last step used
    bool localPasses = false;          // each pass finished in one workgroup, the claims in LDS (launch_pass's `local`)
    bool hostWaits = false;            // run the reference's launch model: host reads each report
    int slice = 0;                     // SphereSimulation.slice: most dispatches a submission, 0 for no bound
    double *diagPartials = nullptr, *diagOut = nullptr;  // device scratch of vr_world_diagnostics, made on first use
    bool profiling = false;
    vr_profile prof{};
    std::vector<hipEvent_t> passEvents;  // a pair a pass, recorded around its round kernels when profiling
    int launchesThisStep = 0;

    template <class T>
    T *alloc(int64_t count) {
        void *p = nullptr;
        size_t bytes = (size_t)std::max<int64_t>(count, 1) * sizeof(T);
        HIP_OK(hipMalloc(&p, bytes));
        allocs.push_back(p);
        return (T *)p;
    }

    void reg(const char *name, void *p, int64_t count, int dtype) { fixed[name] = Buf{p, count, dtype}; }

    void build(const vr_step_desc &desc) {
        d = desc;
        if (d.bodies < 1 || d.substeps < 1 || d.iterations < 1)
            throw ArgError("a step needs spheres, substeps and iterations, got " + std::to_string(d.bodies) + ", " +
                           std::to_string(d.substeps) + ", " + std::to_string(d.iterations));
        if (d.solve != VR_JACOBI && d.solve != VR_GAUSS_SEIDEL_BY_CONTACT) throw ArgError("unknown solve");
        grid = d.has_grid != 0;
        gs = d.solve == VR_GAUSS_SEIDEL_BY_CONTACT;
        if (gs && !grid) throw ArgError("Gauss–Seidel works over the grid's cells, and there is no grid");
        if (d.rounds < 1 || d.contacts_per_sphere < 1)
            throw ArgError("at least a round and a place a sphere, got " + std::to_string(d.rounds) + " and " +
                           std::to_string(d.contacts_per_sphere));
        if (gs && !d.wide_pairs && d.bodies > MOST_LISTED_SPHERES)
            throw ArgError("a contact list orders at most 46340 spheres, got " + std::to_string(d.bodies));
        if (grid && (d.nx < 1 || d.ny < 1 || d.nz < 1 || !(d.cell > 0)))
            throw ArgError("a grid of at least one cell, of positive size");
        int devices = 0;
        if (hipGetDeviceCount(&devices) != hipSuccess || devices < 1) throw DeviceError("no HIP device");
        n = d.bodies;
        cap = n * d.contacts_per_sphere;
        // VR_HOST_WAITS=1 runs a Gauss-Seidel pass as the reference's stepper does, the host reading each report.
        const char *hw = std::getenv("VR_HOST_WAITS");
        hostWaits = hw && hw[0] == '1';
        // A world of at most LOCAL_MOST spheres finishes each pass in one workgroup with its claims in LDS (a round
        // a barrier, not a launch), after wide rounds while more than LOCAL_THREADS contacts are open. Both plans
        // run the same rounds. VR_SPHERE_PLAN=wide|local fixes the choice (local only up to LOCAL_MOST spheres).
        localPasses = n <= LOCAL_MOST;
        if (const char *plan = std::getenv("VR_SPHERE_PLAN")) {
            if (std::string(plan) == "wide") localPasses = false;
            else if (std::string(plan) == "local") localPasses = n <= LOCAL_MOST;
            else throw ArgError("VR_SPHERE_PLAN is wide or local, got " + std::string(plan));
        }
        HIP_OK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        HIP_OK(hipEventCreate(&t0));
        HIP_OK(hipEventCreate(&t1));
        HIP_OK(hipHostMalloc((void **)&pinned, 8 * sizeof(int32_t), hipHostMallocDefault));

        D.n = n;
        D.cap = cap;
        D.wide = d.wide_pairs ? 1 : 0;
        static const char *sphere[] = {"x", "y", "z", "u", "v", "w", "r", "im", "cx", "cy", "cz", "dx",
                                       "dy", "dz", "ax", "ay", "az", "tx", "ty", "tz", "qx", "qy", "qz", "qw"};
        float **fields[] = {&D.x, &D.y, &D.z, &D.u, &D.v, &D.w, &D.r, &D.im, &D.cx, &D.cy, &D.cz, &D.dx,
                            &D.dy, &D.dz, &D.ax, &D.ay, &D.az, &D.tx, &D.ty, &D.tz, &D.qx, &D.qy, &D.qz, &D.qw};
        // One arena for the 24 fields, each padded to 64 floats so every field starts on a 256-byte boundary.
        int64_t stride = ((int64_t)n + 63) / 64 * 64;
        float *arena = alloc<float>(24 * stride);
        for (int k = 0; k < 24; k++) {
            *fields[k] = arena + k * stride;
            reg(sphere[k], *fields[k], n, VR_F32);
        }
        float *params = alloc<float>(PARAM_COUNT);
        D.params = params;
        reg("params", params, PARAM_COUNT, VR_F32);
        D.shown = alloc<float>((int64_t)SHOWN_STRIDE * n);
        reg("shown", D.shown, (int64_t)SHOWN_STRIDE * n, VR_F32);
        if (grid) {
            D.nx = d.nx;
            D.ny = d.ny;
            D.nz = d.nz;
            D.cell = d.cell;
            cells = d.nx * d.ny * d.nz;
            D.cells = cells;
            D.keys = alloc<int32_t>(n);
            D.order = alloc<int32_t>(n);
            D.starts = alloc<int32_t>((int64_t)cells + 1);
            reg("keys", D.keys, n, VR_I32);
            reg("order", D.order, n, VR_I32);
            reg("starts", D.starts, (int64_t)cells + 1, VR_I32);
            int nblocks = (n + 2047) / 2048;
            S.maxBlocks = nblocks;
            S.keysA = alloc<int32_t>(n);
            S.valsA = alloc<int32_t>(n);
            S.keysB = alloc<int32_t>(n);
            S.hist = alloc<int32_t>(256 * (int64_t)nblocks);
            int tmpWords = std::max(scan_tmp_words(256 * nblocks), scan_tmp_words(n)) + 64;
            S.scanTmp = alloc<int32_t>(tmpWords);
        }
        control = alloc<int32_t>(8);
        D.count = control;
        D.open = control + 3;
        if (gs) {
            for (int k = 0; k < 3; k++) D.claims[k] = alloc<long long>(n);
            for (auto &L : lists) {
                L.a = alloc<int32_t>(cap);
                L.b = alloc<int32_t>(cap);
                L.done = alloc<int32_t>(cap);
                L.impulse = alloc<float>((int64_t)IMPULSE_WORDS * cap);
                L.off = alloc<int32_t>((int64_t)n + 1);
                L.cnt = alloc<int32_t>(n);
            }
            D.wallImpulse = alloc<float>((int64_t)WALL_IMPULSE_WORDS * n);
            reg("wallImpulse", D.wallImpulse, (int64_t)WALL_IMPULSE_WORDS * n, VR_F32);
            reg("contactCount", D.count, 3, VR_I32);
            reg("open", D.open, 2, VR_I32);
        }
        opening.assign(d.iterations, BATCH);
        enough.assign(d.iterations, 0);
        if (gs) {
            ctl = alloc<PassCtl>(1);
            nslots = 2 * d.substeps * d.iterations + 2;
            slots = alloc<PassSlot>(nslots);
            HIP_OK(hipHostMalloc((void **)&ctlHost, sizeof(PassCtl), hipHostMallocDefault));
            openLists.segments = (cap + SEG - 1) / SEG;
            for (int k = 0; k < 2; k++) {
                openLists.buf[k] = alloc<int4>((int64_t)openLists.segments * SEG);
                openLists.seg[k] = alloc<int32_t>(openLists.segments);
            }
            devOpening = alloc<int32_t>(d.iterations);
            devEnough = alloc<int32_t>(d.iterations);
            HIP_OK(hipMemcpy(devOpening, opening.data(), sizeof(int32_t) * d.iterations, hipMemcpyHostToDevice));
            HIP_OK(hipMemset(devEnough, 0, sizeof(int32_t) * d.iterations));
            HIP_OK(hipMemset(ctl, 0, sizeof(PassCtl)));
            wide.assign(d.iterations, 4);
            wideRub.assign(d.iterations, 4);
        }
        set_lists();
        clear();
    }

    void set_lists() {
        D.cur = lists[par];
        D.kept = lists[1 - par];
    }

    Buf find(const std::string &name) {
        auto it = fixed.find(name);
        if (it != fixed.end()) return it->second;
        if (gs) {
            const ContactList &c = lists[par], &k = lists[1 - par];
            if (name == "contactA") return Buf{c.a, cap, VR_I32};
            if (name == "contactB") return Buf{c.b, cap, VR_I32};
            if (name == "contactDone") return Buf{c.done, cap, VR_I32};
            if (name == "contactImpulse") return Buf{c.impulse, (int64_t)IMPULSE_WORDS * cap, VR_F32};
            if (name == "keptB") return Buf{k.b, cap, VR_I32};
            if (name == "keptImpulse") return Buf{k.impulse, (int64_t)IMPULSE_WORDS * cap, VR_F32};
        }
        throw NameError("no buffer named " + name);
    }

    // PassRunner.clear() then SphereStep.constants(): every buffer zero, qw = 1 (SphereStep.java:157-160).
    void clear() {
        for (auto &kv : fixed) {
            const Buf &b = kv.second;
            HIP_OK(hipMemsetAsync(b.p, 0, (size_t)b.count * 4, stream));
        }
        HIP_OK(hipMemsetAsync(control, 0, 8 * sizeof(int32_t), stream));
        if (gs) {
            for (int k = 0; k < 3; k++) HIP_OK(hipMemsetAsync(D.claims[k], 0, (size_t)n * 8, stream));
            for (auto &L : lists) {
                HIP_OK(hipMemsetAsync(L.a, 0, (size_t)cap * 4, stream));
                HIP_OK(hipMemsetAsync(L.b, 0, (size_t)cap * 4, stream));
                HIP_OK(hipMemsetAsync(L.done, 0, (size_t)cap * 4, stream));
                HIP_OK(hipMemsetAsync(L.impulse, 0, (size_t)cap * IMPULSE_WORDS * 4, stream));
                HIP_OK(hipMemsetAsync(L.off, 0, (size_t)(n + 1) * 4, stream));
                HIP_OK(hipMemsetAsync(L.cnt, 0, (size_t)n * 4, stream));
            }
        }
        launch_fill(D.qw, 1.0f, n, stream);
        HIP_OK(hipStreamSynchronize(stream));
    }

    void write(const std::string &name, const void *src, int64_t count) {
        Buf b = find(name);
        if (count < 0 || count > b.count) throw NameError("buffer " + name + " is shorter than the write");
        HIP_OK(hipMemcpyAsync(b.p, src, (size_t)count * 4, hipMemcpyHostToDevice, stream));
        HIP_OK(hipStreamSynchronize(stream));
    }

    void read(const std::string &name, void *dst, int64_t count) {
        Buf b = find(name);
        if (count < 0 || count > b.count) throw NameError("buffer " + name + " is shorter than the read");
        HIP_OK(hipMemcpyAsync(dst, b.p, (size_t)count * 4, hipMemcpyDeviceToHost, stream));
        HIP_OK(hipStreamSynchronize(stream));
    }

    void params(const int32_t *words, int32_t count) {
        if (count != PARAM_COUNT) throw ArgError("sphere parameters are 13 words");
        HIP_OK(hipMemcpyAsync((void *)D.params, words, sizeof(int32_t) * PARAM_COUNT, hipMemcpyHostToDevice,
                              stream));
        HIP_OK(hipStreamSynchronize(stream));
        float f[PARAM_COUNT];
        std::memcpy(f, words, sizeof(f));
        friction = f[MU_STATIC] + f[MU_KINETIC] + f[MU_ROLLING] > 0;  // Spheres.rubs, Spheres.java:232-235
    }

    // ---- pass lists ----

    static int position(int k) { return k == 0 ? 0 : 1 + (k - 1) % CYCLE; }  // SphereStep.java:337-339

    RoundConst cycle_round(int it, int p, bool rubs) const {  // SphereStep.java:286-315, 320-327
        RoundConst rc;
        rc.checks = p == 0 ? 0 : 1;
        rc.claims = 1;
        rc.pass = rubs ? d.iterations + it + 1 : it + 1;
        rc.dispatch = it * CYCLE + p % CYCLE;
        rc.mark = p % 2;
        rc.checksDispatch = it * CYCLE + std::max(p - 1, 0) % CYCLE;
        rc.rubs = rubs ? 1 : 0;
        rc.slot = p % 3;
        return rc;
    }

    void predict() { launch_predict(D, stream); }

    void prepare() {  // SphereStep.java:170-182, 227-235
        if (grid) {
            launch_bin(D, stream);
            launch_sort(D, S, stream);
        }
        if (gs) {
            launch_warm_walls(D, stream);
            par ^= 1;  // keepContacts: the list becomes the kept one, and the chains start empty
            set_lists();
            launch_clear_contacts(D, stream);
            launch_list(D, S, stream);
        }
    }

    void velocity_list() {  // SphereStep.java:150-159
        launch_velocity(D, stream);
        launch_bounce(D, grid, stream);
        launch_bounced(D, stream);
    }

    void rubs(int iteration, int rounds) {  // SphereStep.java:270-276
        for (int r = 0; r < rounds; r++) launch_round(D, cycle_round(iteration, position(r), true), stream);
    }

    // SphereStep.addRounds and the report read back (SphereStep.java:394-402; Spheres.java:1146-1159).
    int add_rounds(int it, int from, int rounds, int *listed) {
        for (int k = from; k < from + rounds; k++) launch_round(D, cycle_round(it, position(k), false), stream);
        int mark = position(from + rounds - 1) % 2;
        HIP_OK(hipMemcpyAsync(pinned, control, 5 * sizeof(int32_t), hipMemcpyDeviceToHost, stream));
        HIP_OK(hipStreamSynchronize(stream));
        *listed = pinned[LISTED];
        return pinned[3 + mark];
    }

    void adapt(int it, int ran) {  // SphereStepper.java:118-126
        if (ran > opening[it]) {
            opening[it] = std::min(ran, CYCLE);
            enough[it] = 0;
        } else if (++enough[it] >= SHRINK_AFTER && opening[it] > BATCH) {
            opening[it] -= BATCH;
            enough[it] = 0;
        }
    }

    // The lengths of SphereStep's lists (SphereStep.java:146-208, 216-243, 348-392), which the reference's submissions
    // count: predict; prepare (bin, the counting sort's scan, order, and for Gauss-Seidel warmWalls, keepContacts,
    // clearContacts, listContacts); velocity, bounce, bounced; walls, show and a report one dispatch each.
    // CountingSort.scan's passes belong to the SupirVast library, which is not part of the reference; they are taken
    // as the three of a two-level scan.
    static constexpr int SCAN_PASSES = 3, VELOCITY_PASSES = 3;
    int prepare_passes() const { return grid ? 1 + SCAN_PASSES + 1 + (gs ? 4 : 0) : 0; }
    // The submissions a list of `passes` dispatches is cut into (SphereSimulation.submit).
    int pieces(int passes) const { return slice == 0 ? 1 : (passes + slice - 1) / slice; }
    // What opens pass `it` of substep `sub` before its rounds (SphereStep.opening): the walls of the pass before, or
    // for a substep's first pass the close of the substep before (walls, velocity) unless it is the first, then
    // predict and prepare.
    int opening_prefix(int sub, int it) const {
        if (it > 0) return 1;
        return (sub > 0 ? 1 + VELOCITY_PASSES : 0) + 1 + prepare_passes();
    }
    // SphereStep.closing: the last pass's rubbing rounds, the walls, velocity and show.
    int closing_passes(int rubbed) const { return rubbed + 1 + VELOCITY_PASSES + 1; }

    void pass(int q, int it, bool rub, int wideRounds, const PassSubmit &submit = PassSubmit{0, 0, 0}) {
        if (profiling) {
            while ((int)passEvents.size() < 2 * (q + 1)) {
                hipEvent_t e;
                HIP_OK(hipEventCreate(&e));
                passEvents.push_back(e);
            }
            HIP_OK(hipEventRecord(passEvents[2 * q], stream));
        }
        launch_pass(D, ctl, slots + q, openLists, devOpening, devEnough, it, d.iterations, rub, wideRounds,
                    localPasses, submit, stream);
        launchesThisStep += wideRounds + 1;
        if (profiling) HIP_OK(hipEventRecord(passEvents[2 * q + 1], stream));
    }

    // SphereStepper.step with every decision on the device: each pass is launched as a fixed sequence (a few wide
    // rounds, then the tail that finishes it), so the step runs with no host wait; the rounds run and the report
    // are the reference's.
    void step_device(vr_step_report *rep) {
        auto enqueueStart = std::chrono::steady_clock::now();
        launch_step_begin(ctl, slots, nslots, stream);
        int q = 0;
        launchesThisStep = 0;
        for (int sub = 0; sub < d.substeps; sub++) {
            for (int it = 0; it < d.iterations; it++) {
                if (friction && !(sub == 0 && it == 0)) {
                    int prev = it == 0 ? d.iterations - 1 : it - 1;
                    pass(q++, prev, true, wideRub[prev]);
                }
                if (it > 0) {
                    launch_walls(D, stream);
                } else {
                    if (sub > 0) {
                        launch_walls(D, stream);
                        velocity_list();
                    }
                    predict();
                    prepare();
                }
                PassSubmit submit{opening_prefix(sub, it), friction && !(sub == 0 && it == 0) ? 1 : 0, slice};
                pass(q++, it, false, wide[it], submit);
            }
        }
        if (friction)
            pass(q++, d.iterations - 1, true, wideRub[d.iterations - 1]);
        launch_walls(D, stream);
        velocity_list();
        launch_show(D, grid, stream);
        HIP_OK(hipMemcpyAsync(ctlHost, ctl, sizeof(PassCtl), hipMemcpyDeviceToHost, stream));
        int64_t enqueued = std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() -
                                                                                enqueueStart).count();
        HIP_OK(hipStreamSynchronize(stream));
        if (ctlHost->error)
            throw StateError("a pass is still open after " + std::to_string(ctlHost->errRounds) +
                             " rounds, on a list of " + std::to_string(ctlHost->errListed));
        if (profiling) {
            for (int k = 0; k < q; k++) {
                float ms = 0;
                HIP_OK(hipEventElapsedTime(&ms, passEvents[2 * k], passEvents[2 * k + 1]));
                prof.round_ns += (int64_t)((double)ms * 1e6);
            }
            prof.round_launches += launchesThisStep;
            prof.passes += q;
            prof.listed_sum += ctlHost->listedSum;
            prof.steps += 1;
            prof.enqueue_ns += enqueued;
        }
        rep->waits = ctlHost->waits;
        rep->submissions = ctlHost->submissions + pieces(closing_passes(friction ? ctlHost->k : 0));
        rep->rounds = ctlHost->rounds;
        rep->most_rounds = ctlHost->most;
        rep->unlisted = ctlHost->unlisted;
        // Wide rounds for the next step: as many as this one needed, and one more.
        for (int it = 0; it < d.iterations; it++) {
            int parted = it, rubbed = d.iterations + it;
            // The local plan: wide rounds until the pass fits one workgroup, as this step found; else as many as
            // the pass needed, and one more, so the tail finds it done.
            const int *from = localPasses ? ctlHost->localFrom : ctlHost->wideRan;
            int more = localPasses ? 0 : 1;
            if (parted < MAX_WIDE) wide[it] = std::min(from[parted] + more, MAX_WIDE);
            if (rubbed < MAX_WIDE) wideRub[it] = std::min(from[rubbed] + more, MAX_WIDE);
        }
    }

    void step_until_done(vr_step_report *rep) {  // SphereStepper.java:66-103
        int waits = 0, rounds = 0, most = 0, unlisted = 0, rubbed = 0, submissions = 0;
        int limit = cap + 2 * CYCLE;
        for (int sub = 0; sub < d.substeps; sub++) {
            for (int it = 0; it < d.iterations; it++) {
                int k = opening[it];
                rubs(it == 0 ? d.iterations - 1 : it - 1, rubbed);
                if (it > 0) {
                    launch_walls(D, stream);
                } else {
                    if (sub > 0) {
                        launch_walls(D, stream);
                        velocity_list();
                    }
                    predict();
                    prepare();
                }
                int listed = 0;
                int open = add_rounds(it, 0, k, &listed);
                submissions += pieces(rubbed + opening_prefix(sub, it) + k + 1);
                waits++;
                while (open != 0) {
                    if (k > limit)
                        throw StateError("a pass is still open after " + std::to_string(k) +
                                         " rounds, on a list of " + std::to_string(listed));
                    open = add_rounds(it, k, BATCH, &listed);
                    submissions += pieces(BATCH + 1);
                    k += BATCH;
                    waits++;
                }
                rubbed = friction ? k : 0;
                rounds += k;
                most = std::max(most, k);
                unlisted += std::max(0, listed - cap);
                adapt(it, k);
            }
        }
        rubs(d.iterations - 1, rubbed);
        launch_walls(D, stream);
        velocity_list();
        launch_show(D, grid, stream);
        rep->waits = waits;
        rep->submissions = submissions + pieces(closing_passes(rubbed));
        rep->rounds = rounds;
        rep->most_rounds = most;
        rep->unlisted = unlisted;
    }

    // SphereStep.step(): Jacobi, or Gauss-Seidel in fixed rounds (SphereStep.java:184-208, 245-260).
    void step_list() {
        for (int sub = 0; sub < d.substeps; sub++) {
            predict();
            prepare();
            if (!gs) {
                for (int it = 0; it < d.iterations; it++) {
                    launch_jacobi_solve(D, grid, stream);
                    launch_jacobi_apply(D, stream);
                }
            } else {
                int g = 0;
                for (int it = 0; it < d.iterations; it++) {
                    for (int k = 0; k <= d.rounds; k++, g++) {
                        RoundConst rc;
                        rc.checks = k == 0 ? 0 : 1;
                        rc.claims = k == d.rounds ? 0 : 1;
                        rc.pass = it + 1;
                        rc.dispatch = g;
                        rc.mark = g % 2;
                        rc.checksDispatch = g - 1;
                        rc.rubs = 0;
                        rc.slot = g % 3;
                        launch_round(D, rc, stream);
                    }
                    launch_walls(D, stream);
                }
            }
            velocity_list();
        }
        launch_show(D, grid, stream);
    }

    void step(int32_t wait, vr_step_report *rep) {
        vr_step_report local{};
        vr_step_report *r = rep ? rep : &local;
        std::memset(r, 0, sizeof(*r));
        auto w0 = std::chrono::steady_clock::now();
        HIP_OK(hipEventRecord(t0, stream));
        if (gs && !d.fixed_rounds) {
            if (hostWaits) step_until_done(r);
            else step_device(r);
        } else {
            step_list();
            // SphereStep.step(), one list: per substep predict, prepare, each iteration's passes (solve and apply,
            // or the fixed rounds and the walls), velocity; then show.
            int iteration = gs ? d.rounds + 2 : 2;
            r->submissions = pieces(d.substeps * (1 + prepare_passes() + d.iterations * iteration + VELOCITY_PASSES) + 1);
        }
        HIP_OK(hipEventRecord(t1, stream));
        if (wait) {
            HIP_OK(hipStreamSynchronize(stream));
            float ms = 0;
            HIP_OK(hipEventElapsedTime(&ms, t0, t1));
            r->gpu_ns = (int64_t)(ms * 1e6);
        }
        auto w1 = std::chrono::steady_clock::now();
        r->wall_ns = std::chrono::duration_cast<std::chrono::nanoseconds>(w1 - w0).count();
    }

    // SphereDiagnostics.of, reduced on the device; the grid is rebuilt at the current positions for the overlaps.
    void diagnostics(double *out) {
        if (!diagOut) {
            diagPartials = alloc<double>(vrdiag::partial_words(n));
            diagOut = alloc<double>(vrdiag::D_WORDS);
        }
        diagnostics_into(diagPartials, diagOut, out);
    }

    // The same into device scratch the caller provides: `partials` of vrdiag::partial_words(n) doubles and
    // `devOut` of D_WORDS.
    void diagnostics_into(double *partials, double *devOut, double *out) {
        if (grid) {
            launch_bin(D, stream);
            launch_sort(D, S, stream);
        }
        vrdiag::launch_spheres(D, grid, partials, devOut, stream);
        HIP_OK(hipMemcpyAsync(out, devOut, sizeof(double) * vrdiag::D_WORDS, hipMemcpyDeviceToHost, stream));
        HIP_OK(hipStreamSynchronize(stream));
    }

    ~vr_world() {
        if (box) vrbox::box_destroy(box);
        if (stream) (void)hipStreamSynchronize(stream);
        for (void *p : allocs) (void)hipFree(p);
        if (pinned) (void)hipHostFree(pinned);
        if (ctlHost) (void)hipHostFree(ctlHost);
        for (hipEvent_t e : passEvents) (void)hipEventDestroy(e);
        if (t0) (void)hipEventDestroy(t0);
        if (t1) (void)hipEventDestroy(t1);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

int fail(int code, const std::string &msg) {
    g_error = msg;
    return code;
}

template <class F>
int guard(F fn) {
    try {
        fn();
        return VR_OK;
    } catch (const ArgError &e) {
        return fail(VR_ERR_ARGUMENT, e.what());
    } catch (const StateError &e) {
        return fail(VR_ERR_STATE, e.what());
    } catch (const NameError &e) {
        return fail(VR_ERR_NAME, e.what());
    } catch (const std::exception &e) {
        return fail(VR_ERR_DEVICE, e.what());
    }
}

}  // namespace

// Device scratch of vr_world_diagnostics_static: each workgroup's row of partial sums and maxima, then the 14 results,
// in place of the buffers vr_world_diagnostics makes on its first call. 1 MiB, one copy on each GPU the library runs
// on, held until the process ends. In host code `scratch` is only a stand-in for the symbol: its device address on
// each GPU comes from hipGetSymbolAddress (static_scratch), and only that address is handed to kernels and copies.
constexpr int64_t STATIC_DIAG_DOUBLES = 131072;
// The most bodies that fit: (ceil(n / 256) + 1) rows of 14 doubles, and the 14 results.
constexpr int STATIC_DIAG_BODIES = (int)((STATIC_DIAG_DOUBLES / 14 - 2) * 256);
static __device__ __align__(256) double scratch[STATIC_DIAG_DOUBLES];

namespace {

constexpr int MOST_DEVICES = 64;
std::mutex g_scratchLock;
double *g_scratchAt[MOST_DEVICES] = {};

// The scratch array's address on `device`, looked up once per device with that device current.
double *static_scratch(int device) {
    if (device < 0 || device >= MOST_DEVICES) throw DeviceError("no scratch for device " + std::to_string(device));
    std::lock_guard<std::mutex> hold(g_scratchLock);
    if (!g_scratchAt[device]) {
        int current = 0;
        HIP_OK(hipGetDevice(&current));
        if (current != device) HIP_OK(hipSetDevice(device));
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (current != device) HIP_OK(hipSetDevice(current));
        if (e != hipSuccess) throw DeviceError(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratchAt[device] = static_cast<double *>(p);
    }
    return g_scratchAt[device];
}

}  // namespace

extern "C" {

int vr_world_create(int32_t kind, const vr_step_desc *desc, vr_world **out) {
    if (!desc || !out) return fail(VR_ERR_ARGUMENT, "null argument");
    if (kind != VR_SPHERES && kind != VR_BOXES)
        return fail(VR_ERR_ARGUMENT, "unknown world kind " + std::to_string(kind));
    auto *w = new vr_world();
    int rc = guard([&] {
        if (kind == VR_BOXES) {
            try {
                w->box = vrbox::box_create(*desc);
            } catch (const std::invalid_argument &e) {
                throw ArgError(e.what());
            }
        } else {
            w->build(*desc);
        }
        HIP_OK(hipGetDevice(&w->device));
    });
    if (rc != VR_OK) {
        delete w;
        return rc;
    }
    *out = w;
    return VR_OK;
}

int vr_batch_create(const vr_step_desc *desc, int32_t scenes, const int32_t *bodies, const float *extents,
                    const int32_t *grids, const float *cells, vr_world **out) {
    if (!desc || !out || !bodies || !extents || !grids || !cells) return fail(VR_ERR_ARGUMENT, "null argument");
    auto *w = new vr_world();
    int rc = guard([&] {
        try {
            w->box = vrbox::box_create_batch(*desc, scenes, bodies, extents, grids, cells);
        } catch (const std::invalid_argument &e) {
            throw ArgError(e.what());
        }
        HIP_OK(hipGetDevice(&w->device));
    });
    if (rc != VR_OK) {
        delete w;
        return rc;
    }
    *out = w;
    return VR_OK;
}

int vr_world_clear(vr_world *w) {
    if (!w) return fail(VR_ERR_ARGUMENT, "null world");
    if (w->box) return guard([&] { vrbox::box_clear(w->box); });
    return guard([&] { w->clear(); });
}

namespace {

Buf find_any(vr_world *w, const std::string &name, int *bodyWords = nullptr) {
    if (bodyWords) *bodyWords = 0;
    if (!w->box) return w->find(name);
    vrbox::BoxBuf b;
    if (!vrbox::box_find(w->box, name, &b)) throw NameError("no buffer named " + name);
    if (bodyWords) *bodyWords = b.bodyWords;
    return Buf{b.p, b.count, b.dtype};
}

hipStream_t stream_of(vr_world *w) { return w->box ? vrbox::box_stream(w->box) : w->stream; }

// A slot-major device buffer to the reference's body-by-body order (toBodies) or back, the whole buffer: word k of
// body b sits at k * bodies + place[b] on the device.
void reorder(const int32_t *from, int32_t *to, int64_t count, int words, bool toBodies,
             const std::vector<int32_t> &place) {
    int64_t bodies = count / words;
    for (int64_t b = 0; b < bodies; b++)
        for (int k = 0; k < words; k++) {
            int64_t bodyOrder = b * words + k, slotOrder = (int64_t)k * bodies + place[(size_t)b];
            if (toBodies) to[bodyOrder] = from[slotOrder];
            else to[slotOrder] = from[bodyOrder];
        }
}

void copy_in(vr_world *w, const std::string &name, const void *src, int64_t count) {
    int words = 0;
    Buf b = find_any(w, name, &words);
    if (count < 0 || count > b.count) throw NameError("buffer " + name + " is shorter than the write");
    hipStream_t s = stream_of(w);
    if (words > 0) {  // a part written: the rest kept, so the whole buffer goes through the host
        std::vector<int32_t> dev((size_t)b.count), host((size_t)b.count), place;
        vrbox::box_wall_places(w->box, place);
        HIP_OK(hipMemcpyAsync(dev.data(), b.p, (size_t)b.count * 4, hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
        reorder(dev.data(), host.data(), b.count, words, true, place);
        std::memcpy(host.data(), src, (size_t)count * 4);
        reorder(host.data(), dev.data(), b.count, words, false, place);
        HIP_OK(hipMemcpyAsync(b.p, dev.data(), (size_t)b.count * 4, hipMemcpyHostToDevice, s));
        HIP_OK(hipStreamSynchronize(s));
        vrbox::box_wall_impulses_written(w->box);
    } else {
        HIP_OK(hipMemcpyAsync(b.p, src, (size_t)count * 4, hipMemcpyHostToDevice, s));
        HIP_OK(hipStreamSynchronize(s));
        if (w->box && name == "shape") vrbox::box_shapes_written(w->box);
        if (w->box) vrbox::box_pair_buffer_written(w->box, name);
    }
    HIP_OK(hipStreamSynchronize(s));
}

void copy_out(vr_world *w, const std::string &name, void *dst, int64_t count) {
    int words = 0;
    Buf b = find_any(w, name, &words);
    if (count < 0 || count > b.count) throw NameError("buffer " + name + " is shorter than the read");
    hipStream_t s = stream_of(w);
    if (words > 0) {
        std::vector<int32_t> dev((size_t)b.count), host((size_t)b.count), place;
        vrbox::box_wall_places(w->box, place);
        HIP_OK(hipMemcpyAsync(dev.data(), b.p, (size_t)b.count * 4, hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
        reorder(dev.data(), host.data(), b.count, words, true, place);
        std::memcpy(dst, host.data(), (size_t)count * 4);
    } else {
        HIP_OK(hipMemcpyAsync(dst, b.p, (size_t)count * 4, hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
    }
}

}  // namespace

int vr_world_write(vr_world *w, const char *name, const void *src, int64_t count) {
    if (!w || !name || (!src && count > 0)) return fail(VR_ERR_ARGUMENT, "null argument");
    return guard([&] { copy_in(w, name, src, count); });
}

int vr_world_read(vr_world *w, const char *name, void *dst, int64_t count) {
    if (!w || !name || (!dst && count > 0)) return fail(VR_ERR_ARGUMENT, "null argument");
    return guard([&] { copy_out(w, name, dst, count); });
}

int vr_world_buffer_info(vr_world *w, const char *name, int64_t *count, int32_t *dtype) {
    if (!w || !name || !count || !dtype) return fail(VR_ERR_ARGUMENT, "null argument");
    return guard([&] {
        Buf b = find_any(w, name);
        *count = b.count;
        *dtype = b.dtype;
    });
}

int vr_world_params(vr_world *w, const int32_t *words, int32_t n) {
    if (!w || !words) return fail(VR_ERR_ARGUMENT, "null argument");
    if (w->box) {
        return guard([&] {
            try {
                vrbox::box_params(w->box, words, n);
            } catch (const std::invalid_argument &e) {
                throw ArgError(e.what());
            }
        });
    }
    return guard([&] { w->params(words, n); });
}

int vr_world_start(vr_world *w, const float *x, const float *y, const float *z, const float *u, const float *v,
                   const float *vw, const float *r, const float *im, const int32_t *params, int32_t nparams) {
    if (!w || !x || !y || !z || !u || !v || !vw || !r || !im || !params) return fail(VR_ERR_ARGUMENT, "null argument");
    if (w->box) return fail(VR_ERR_STATE, "a box world starts through clear, writes and params");
    int64_t n = 0;
    int32_t dtype = 0;
    int rc = vr_world_buffer_info(w, "x", &n, &dtype);
    if (rc == VR_OK) rc = vr_world_clear(w);
    const char *names[8] = {"x", "y", "z", "u", "v", "w", "r", "im"};
    const float *fields[8] = {x, y, z, u, v, vw, r, im};
    for (int k = 0; k < 8 && rc == VR_OK; k++) rc = vr_world_write(w, names[k], fields[k], n);
    if (rc == VR_OK) rc = vr_world_params(w, params, nparams);
    if (rc != VR_OK) return rc;
    // The picture's before and after (Spheres.SHOWN_STRIDE words a sphere): centre and radius, then the turns'
    // w words (Spheres.TURNED_BEFORE + 3, TURNED_AFTER + 3) as the identity.
    constexpr int TURNED_BEFORE = 32, TURNED_AFTER = 36;
    std::vector<float> shown((size_t)SHOWN_STRIDE * (size_t)n, 0.f);
    for (int64_t s = 0; s < n; s++) {
        float *b = shown.data() + (size_t)SHOWN_STRIDE * (size_t)s;
        b[0] = b[4] = x[s];
        b[1] = b[5] = y[s];
        b[2] = b[6] = z[s];
        b[3] = r[s];
        b[TURNED_BEFORE + 3] = b[TURNED_AFTER + 3] = 1.f;
    }
    return vr_world_write(w, "shown", shown.data(), (int64_t)shown.size());
}

int vr_world_friction(vr_world *w, int32_t on) {
    if (!w) return fail(VR_ERR_ARGUMENT, "null world");
    w->friction = on != 0;
    return VR_OK;
}

int vr_world_slice(vr_world *w, int32_t dispatches) {  // SphereSimulation.java:136-148
    if (!w) return fail(VR_ERR_ARGUMENT, "null world");
    if (w->box) return fail(VR_ERR_STATE, "a box world submits each list whole: it has no slice");
    if (dispatches < 0)
        return fail(VR_ERR_ARGUMENT,
                    "a slice is some dispatches, or zero for no bound; got " + std::to_string(dispatches));
    w->slice = dispatches;
    return VR_OK;
}

int vr_world_step(vr_world *w, int32_t wait, vr_step_report *report) {
    if (!w) return fail(VR_ERR_ARGUMENT, "null world");
    if (w->box) {
        return guard([&] {
            vr_step_report local{};
            vr_step_report *r = report ? report : &local;
            std::memset(r, 0, sizeof(*r));
            vrbox::box_step(w->box, r);
        });
    }
    return guard([&] { w->step(wait, report); });
}

int vr_world_finish(vr_world *w) {
    if (!w) return fail(VR_ERR_ARGUMENT, "null world");
    return guard([&] { HIP_OK(hipStreamSynchronize(stream_of(w))); });
}

int vr_world_destroy(vr_world *w) {
    delete w;
    return VR_OK;
}

int vr_box_collide(vr_world *w, const int32_t *pairs, int32_t npairs, float *manifold) {
    if (!w || (!pairs && npairs > 0) || (!manifold && npairs > 0)) return fail(VR_ERR_ARGUMENT, "null argument");
    if (!w->box) return fail(VR_ERR_STATE, "not a box world");
    return guard([&] {
        try {
            vrbox::box_collide(w->box, pairs, npairs, manifold);
        } catch (const std::invalid_argument &e) {
            throw ArgError(e.what());
        }
    });
}

int vr_box_collide_list(vr_world *w, vr_step_report *report) {
    if (!w) return fail(VR_ERR_ARGUMENT, "null world");
    if (!w->box) return fail(VR_ERR_STATE, "not a box world");
    return guard([&] {
        vr_step_report local{};
        vr_step_report *r = report ? report : &local;
        std::memset(r, 0, sizeof(*r));
        vrbox::box_collide_list(w->box, r);
    });
}

int vr_box_tolerance(vr_world *w, float tolerance, int32_t measure) {
    if (!w) return fail(VR_ERR_ARGUMENT, "null world");
    if (!w->box) return fail(VR_ERR_STATE, "not a box world");
    return guard([&] {
        try {
            vrbox::box_tolerance(w->box, tolerance, measure != 0);
        } catch (const std::invalid_argument &e) {
            throw ArgError(e.what());
        }
    });
}

int vr_box_retune(vr_world *w) {
    if (!w) return fail(VR_ERR_ARGUMENT, "null world");
    if (!w->box) return fail(VR_ERR_STATE, "not a box world");
    return guard([&] { vrbox::box_retune(w->box); });
}

int vr_box_plan(vr_world *w, int32_t *plan, int32_t *tuning) {
    if (!w || !plan || !tuning) return fail(VR_ERR_ARGUMENT, "null argument");
    if (!w->box) return fail(VR_ERR_STATE, "not a box world");
    return guard([&] { vrbox::box_plan(w->box, plan, tuning); });
}

int vr_box_solve_read(vr_world *w, vr_box_solve *out) {
    if (!w || !out) return fail(VR_ERR_ARGUMENT, "null argument");
    if (!w->box) return fail(VR_ERR_STATE, "not a box world");
    return guard([&] { vrbox::box_solve_read(w->box, out); });
}

int vr_box_solve_curve(vr_world *w, float *out, int32_t n, int32_t *count) {
    if (!w || !count || n < 0 || (n > 0 && !out)) return fail(VR_ERR_ARGUMENT, "null argument");
    if (!w->box) return fail(VR_ERR_STATE, "not a box world");
    return guard([&] { *count = vrbox::box_solve_curve(w->box, out, n); });
}

int vr_box_log(vr_world *w, int32_t on, int64_t capacity) {
    if (!w || (on && capacity < 1)) return fail(VR_ERR_ARGUMENT, "a log needs room for a sample");
    if (!w->box) return fail(VR_ERR_STATE, "not a box world");
    return guard([&] { vrbox::box_log(w->box, on != 0, capacity); });
}

int vr_box_samples(vr_world *w, float *x, float *t, int64_t most, int64_t *count) {
    if (!w || !count || most < 0 || (most > 0 && (!x || !t))) return fail(VR_ERR_ARGUMENT, "null argument");
    if (!w->box) return fail(VR_ERR_STATE, "not a box world");
    return guard([&] { *count = vrbox::box_samples(w->box, x, t, most); });
}

int vr_box_learned(vr_world *w, vr_mlp *m) {
    if (!w) return fail(VR_ERR_ARGUMENT, "null world");
    if (!w->box) return fail(VR_ERR_STATE, "not a box world");
    return guard([&] { vrbox::box_learned(w->box, vrmlp::net_of(m)); });
}

int vr_world_diagnostics(vr_world *w, double *out, int32_t n) {
    if (!w || !out || n < vrdiag::D_WORDS) return fail(VR_ERR_ARGUMENT, "diagnostics need 14 doubles");
    if (w->box) return guard([&] { vrbox::box_diagnostics(w->box, out); });
    return guard([&] { w->diagnostics(out); });
}

int vr_world_diagnostics_static(vr_world *w, double *out, int32_t n) {
    if (!w || !out || n < vrdiag::D_WORDS) return fail(VR_ERR_ARGUMENT, "diagnostics need 14 doubles");
    return guard([&] {
        int bodies = w->box ? vrbox::box_bodies(w->box) : w->n;
        int64_t rows = vrdiag::partial_words(bodies);
        if (rows + vrdiag::D_WORDS > STATIC_DIAG_DOUBLES)
            throw ArgError("diagnostics without allocation take at most " + std::to_string(STATIC_DIAG_BODIES) +
                           " bodies, got " + std::to_string(bodies));
        double *at = static_scratch(w->device);
        if (w->box)
            vrbox::box_diagnostics_into(w->box, at, at + rows, out);
        else
            w->diagnostics_into(at, at + rows, out);
    });
}

int vr_world_profile(vr_world *w, int32_t on) {
    if (!w) return fail(VR_ERR_ARGUMENT, "null world");
    if (w->box) return guard([&] { vrbox::box_profile(w->box, on != 0); });
    w->profiling = on != 0;
    w->prof = vr_profile{};
    return VR_OK;
}

int vr_world_profile_read(vr_world *w, vr_profile *out) {
    if (!w || !out) return fail(VR_ERR_ARGUMENT, "null argument");
    if (w->box) return guard([&] { vrbox::box_profile_read(w->box, out); });
    *out = w->prof;
    return VR_OK;
}

int vr_last_error(char *buf, int32_t len) {
    if (!buf || len <= 0) return VR_OK;
    std::strncpy(buf, g_error.c_str(), (size_t)len - 1);
    buf[len - 1] = 0;
    return VR_OK;
}

}  // extern "C"
