"""Benchmark: body-substeps per second on one MI355X, at equal residual, plain against batched + learned warm start.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--run R] [--cpu-seconds S] [--workload batched_2k|pile_64k|
                mixed_32k] [--train-batches B] [--c1-steps K1] [--copy-probe MiB]

The workload (batched_2k) is the batched configuration: 2,048 independent scenes of log-uniform 16-4,096 bodies
(1,549,388 in all, half boxes and half balls, mu 0.6), stepped together at dt 1/60 s in 8 substeps of at most 10 solve
and 10 relax sweeps, over the configuration's R = 300 steps. The timed window is the run's last K steps, after
R - K - W - 11 untimed steps, one measured step, 10 in which the batch times its launch plans afresh for the method
(vr_box_retune) and W warm-up steps: the piles as the run leaves them, not the free fall of its first steps.
All state is resident in HBM before the timed region; nothing crosses PCIe inside it but each step's small report.

Three runs reach the same state at the window (the stepping is deterministic) and differ only inside it:
  plain:    the reference's method: the carried warm start and every sweep (10 + 10 a substep);
  stopped:  the plain method at a residual: the carried warm start, each scene's sweeps of a phase stopped once a sweep
            moves no point's velocity by more than the tolerance (--tolerance, 0.01 m/s), at most 10;
  learned:  the same stop, each step started in each scene from the network's prediction where a trial of the first
            substep's first solve sweep finds it leaves the scene's residual smaller than the carry does (the carry
            elsewhere).
A point's residual is the impulse change a sweep makes there times the inverse mass along the normal (m/s). stopped and
learned end every scene's phases at the same residual, the tolerance (or at the 10-sweep cap), so learned / stopped is
the learned start's gain at equal residual; plain runs every sweep and ends below it. Each run steps once measured at
its tolerance from the state all three share (before the window; equal_residual compares these), and after the window
once more at its tolerance and once at 1e-5 m/s: each phase's sweeps a scene ran, its residual (largest and mean over
the scenes), and the deepest manifold point.

value = learned body-substeps per second = bodies x substeps x K / seconds. The line carries:
  plain, stopped, learned: each run's rate, residuals, deepest point, sweeps;
  roofline: the kernel that runs the solve and relax sweeps, between HIP events on the world's stream, in the plain
    window. On the group plan the tuner picks for the settled batch that is k_group_phase, one launch a phase of 10
    sweeps (each scene group through its walls and colour rounds): algorithmic bytes P (236 + 88 k-bar) + 96 N + 32 W a
    sweep (P touching pairs, k-bar points a pair, N bodies, W active wall contacts) over its time, against MI355X's
    8 TB/s. On the other plans the colour rounds (k_solve_round, a launch a round) are the roofline and k_walls
    (96 N + 32 W a sweep) is roofline_walls;
  cpu_baseline: the oracle (oracle/, a C++ restatement of the reference's passes) on the first scenes of the batch from
    the plain window's state, one thread and all threads (the scenes share nothing: one thread a scene at a time);
  pile_64k: the 65,536-sphere pile (the sphere configuration), the last K1 steps of its 600, with its own roofline;
  pile_1k, mixed_32k (--others 1): the 1,000-sphere pile (last 20 of its 300 steps) and the 33,408-body mixed box pile
    (last 20 of its 600), each a single scene;
  reference_scenes (--others 1): the reference's own measured sphere scenes (its column of 20, pile of 343 and pile of
    4,096 with friction, 10 substeps), timed as the reference times them (2 frames untimed, then the rest of the run,
    each step waited for), beside the ms a step it publishes for them on its own GPU.
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from __graft_entry__ import load_package  # noqa: E402

METRIC = "body-substeps/s at equal residual: plain vs batched+learned warm-start"
HBM_PEAK_GBS = 8000.0
STATE = ("x", "y", "z", "u", "v", "w", "lx", "ly", "lz", "qx", "qy", "qz", "qw")

# HBM bytes a launch from rocprofv3 counter passes (FETCH_SIZE, then WRITE_SIZE, each scaled by what it reads for a
# copy of known size), over the same window as the bench line's roofline. None where not measured for that window.
SOLVE_ROUND_TRAFFIC_BYTES = None
GROUP_ROUNDS_TRAFFIC_BYTES = None
GROUP_PHASE_TRAFFIC_BYTES = 13_728_569_637  # k_group_phase, plain window (profiles/traffic_box_kernels_batched_2k.md)
WALLS_TRAFFIC_BYTES = None
SPHERE_ROUND_TRAFFIC_BYTES = None

_last_progress = [time.perf_counter()]


def progress(what, done, total):
    """A progress line every 50 steps, or after 20 s without one."""
    now = time.perf_counter()
    if done % 50 == 0 or done == total or now - _last_progress[0] > 20:
        _last_progress[0] = now
        print(f"[bench] {what}: {done}/{total} steps", file=sys.stderr, flush=True)


def modules(pkg):
    get = lambda name: __import__(pkg.__name__ + "." + name, fromlist=[name])  # noqa: E731
    return get("scenes"), get("boxes"), get("mlp")


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def threads_available():
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    limit = os.environ.get("OMP_NUM_THREADS")
    return max(1, min(n, int(limit))) if limit and limit.isdigit() else n


# ---- the batched configuration ----

def batch_of(boxes, spec, abi=None):
    sim = boxes.BoxBatch(spec.bodies, spec.extents, spec.grids, spec.substeps, spec.sweeps, abi=abi)
    sim.start(spec.x, spec.y, spec.z, spec.u, spec.v, spec.w, spec.spin, spec.q, spec.half, spec.shape, spec.im,
              spec.params)
    return sim


def training_samples(pkg, batches, settle, logged):
    """Samples for the network from plain runs of `batches` other seeded batches of 2,048 scenes each: after `settle`
    steps, `logged` steps logged (features at each step's start; targets, what each point holds once the first
    substep's solve sweeps are done). The last batch is held out, so the held-out error is over scenes the network
    never saw."""
    scenes, boxes, _ = modules(pkg)
    parts = []
    for b in range(batches):
        spec = scenes.batched_2k(seed=100 + b)
        sim = batch_of(boxes, spec)
        for k in range(settle):
            sim.stepAndWait()
        sim.log(True, 8 << 20)
        for _ in range(logged):
            sim.stepAndWait()
        parts.append(sim.samples(8 << 20))
        sim.close()
        print(f"[bench] training batch {b + 1}/{batches}: {len(parts[-1][0])} samples", file=sys.stderr, flush=True)
    return parts, len(parts) * 2048


def measured_step(sim, tol, walls_too=True):
    """One step with the phases measured: the residual a scene ended a solve or relax phase with (the largest over the
    scenes, and their mean), the sweeps a scene ran on average and the most launched, the deepest manifold point, and
    the active wall contacts."""
    sim.tolerance(tol, True)
    rep = sim.stepAndWait()
    r = sim.solveReport()
    _, _, broken, depth = sim.diagnostics()
    walls = int(np.count_nonzero(sim.floats("lambda") > 0)) if walls_too else None
    sim.tolerance(tol, False)
    subs = r.residual.shape[1]
    return {
        "tolerance_m_per_s": tol if tol >= 0 else None,
        "residual_max_m_per_s": float(np.max(r.residual)),
        "residual_mean_m_per_s": float(np.mean(r.residual_mean)),
        "residual_solve_max_m_per_s": float(np.max(r.residual[0])),
        "residual_solve_mean_m_per_s": float(np.mean(r.residual_mean[0])),
        "deepest_point_m": float(depth),
        "broken": bool(broken),
        "sweeps_per_scene_per_substep": float(np.sum(r.scene_sweeps)) / subs,
        "sweeps_launched_per_substep": float(np.sum(r.launched)) / subs,
        "solve_sweeps_per_scene_by_substep": [round(float(v), 3) for v in r.scene_sweeps[0]],
        "touching": rep.touching, "points": rep.points, "rounds": rep.rounds,
        "active_wall_contacts": walls,
        "learned_points": int(sim.learnedPoints), "pairs_left_to_the_carry": int(r.fell_back),
    }


TUNE_STEPS = 10  # untimed steps in which a batch times its launch plans (two each) after vr_box_retune


def batch_run(pkg, label, spec, args, net=None, tol=-1.0, profile=False, keep_state=False):
    """Steps the batch R - K - W - 11 steps plain, then with `net` and `tol` one measured step (from the state every
    run reaches there, stepping being deterministic, so the runs' residuals differ only by their method), 10 steps in
    which it times its launch plans afresh (vr_box_retune: the plans' costs in the settled piles are not those of the
    free fall the batch timed them in), W warm-up and K timed steps, then two measured steps (at `tol` and at
    1e-5 m/s)."""
    _, boxes, _ = modules(pkg)
    sim = batch_of(boxes, spec)
    settle = args.run - args.steps - args.warmup - TUNE_STEPS - 1
    for k in range(settle):
        sim.stepAndWait()
        progress(f"{label} settle", k + 1, settle)
    state = {k: sim.floats(k).copy() for k in STATE + ("hx", "hy", "hz", "shape", "im")} if keep_state else None
    if net is not None:
        sim.learned(net)
    shared = measured_step(sim, tol)
    sim.tolerance(tol, False)
    sim.retune()
    for _ in range(TUNE_STEPS):
        sim.stepAndWait()
    plan, tuning = sim.plan()
    if tuning:
        raise RuntimeError(f"{label}: the batch was still timing its launch plans after {TUNE_STEPS} steps")
    for _ in range(args.warmup):
        sim.stepAndWait()
    if profile:
        sim._world.profile(True)
    reps = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        reps.append(sim.stepAndWait())
        progress(f"{label} timed", k + 1, args.steps)
    seconds = time.perf_counter() - t0
    prof = sim._world.profile_read() if profile else None
    if profile:
        sim._world.profile(False)
    after = measured_step(sim, tol)
    tight = measured_step(sim, 1e-5, walls_too=False)
    sim.learned(None)
    sim.close()
    n = int(spec.bodies.sum())
    out = {
        "value": n * spec.substeps * args.steps / seconds, "unit": "body-substeps/s",
        "ms_per_step": seconds * 1e3 / args.steps,
        "tolerance_m_per_s": tol if tol >= 0 else None,
        "launch_plan": plan,
        "touching_at_start": reps[0].touching, "touching_at_end": reps[-1].touching,
        "rounds_at_end": reps[-1].rounds, "points_at_end": reps[-1].points,
        "unplaced": int(sum(r.unplaced for r in reps)), "unlisted": int(sum(r.unlisted for r in reps)),
        "from_shared_state": shared,
        "after_window": after,
        "at_1e-5": {"sweeps_per_scene_per_substep": tight["sweeps_per_scene_per_substep"],
                    "solve_sweeps_per_scene_by_substep": tight["solve_sweeps_per_scene_by_substep"],
                    "residual_max_m_per_s": tight["residual_max_m_per_s"],
                    "residual_mean_m_per_s": tight["residual_mean_m_per_s"],
                    "note": "the next step measured at a tolerance of 1e-5 m/s: the sweeps each scene needs to reach "
                            "it (10 when it does not within the step's 10)"},
    }
    return out, prof, state


def box_roofline(prof, walls_per_sweep):
    """The solve and relax sweeps' kernels over the timed sweeps: algorithmic bytes over their event-timed durations.
    On the group plan with whole phases a launch (k_group_phase: each scene group through a phase's sweeps, its walls
    and its colour rounds), that one kernel carries both the rounds' P (236 + 88 k-bar) and the walls' 96 N + 32 W a
    sweep; otherwise the colour rounds (k_solve_round, or k_group_rounds_local a sweep) and k_walls are timed apart
    and reported apart."""
    phased = prof.sweep_launches > 0 and prof.walls_launches == 0
    grouped = prof.sweep_launches > 0
    round_bytes = 236.0 * prof.listed_sum + 88.0 * prof.points_sum
    walls_bytes = 96.0 * prof.bodies_sum + 32.0 * walls_per_sweep * prof.passes
    round_ns = prof.sweep_ns if grouped else prof.round_ns
    launches = prof.sweep_launches if grouped else prof.round_launches
    kbar = prof.points_sum / max(prof.listed_sum, 1)
    per_sweep = {"touching_pairs_per_sweep": prof.listed_sum / max(prof.passes, 1), "points_per_pair": kbar,
                 "bodies_per_sweep": prof.bodies_sum / max(prof.passes, 1), "active_wall_contacts": walls_per_sweep}
    if phased:
        total = round_bytes + walls_bytes
        seconds = round_ns * 1e-9
        achieved = total / seconds / 1e9 if seconds > 0 else 0.0
        rounds = {
            "bound": "hbm",
            "limited_by": "dependent latency: the largest scenes' chains of pair solves, round after round; each SIMD's "
                          "vector ALU issues about 67 % of the time (SQ counters, "
                          "profiles/valu_group_phase_batched_2k.md); no MFMA work",
            "kernel": "k_group_phase (a solve or relax phase of sweeps in one launch: each scene group, a workgroup, "
                      "through the phase's walls and colour rounds, its motion records in LDS), plain window",
            "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
            "traffic": GROUP_PHASE_TRAFFIC_BYTES,
            "bytes_per_launch": total / max(launches, 1),
            "avg_launch_us": round_ns / 1e3 / max(launches, 1),
            "launches": int(launches), "sweeps": int(prof.passes), **per_sweep,
            "formula": "P (236 + 88 k-bar) + 96 N + 32 W B a sweep: P touching pairs, k-bar points a pair (the colour "
                       "rounds), N bodies, W active wall contacts (the walls)",
        }
        return rounds, None
    round_s = round_ns * 1e-9
    achieved = round_bytes / round_s / 1e9 if round_s > 0 else 0.0
    walls_s = prof.walls_ns * 1e-9
    walls_achieved = walls_bytes / walls_s / 1e9 if walls_s > 0 else 0.0
    if grouped:
        kernel = ("k_group_rounds_local (a solve or relax sweep's colour rounds in one launch, a workgroup a scene "
                  "group, its motion records in LDS)")
    else:
        kernel = "k_solve_round (solve and relax sweeps' colour rounds, a launch a round)"
    rounds = {
        "bound": "hbm",
        "kernel": kernel + ", plain window",
        "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
        "traffic": (GROUP_ROUNDS_TRAFFIC_BYTES if grouped else SOLVE_ROUND_TRAFFIC_BYTES),
        "bytes_per_launch": round_bytes / max(launches, 1),
        "avg_launch_us": round_ns / 1e3 / max(launches, 1),
        "launches": int(launches), "sweeps": int(prof.passes), **per_sweep,
        "formula": "P (236 + 88 k-bar) B a sweep, P touching pairs, k-bar points a pair",
    }
    walls = {
        "bound": "hbm",
        "kernel": "k_walls (solve and relax sweeps), plain window",
        "achieved": walls_achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": walls_achieved / HBM_PEAK_GBS,
        "traffic": WALLS_TRAFFIC_BYTES,
        "bytes_per_launch": walls_bytes / max(prof.walls_launches, 1),
        "avg_launch_us": prof.walls_ns / 1e3 / max(prof.walls_launches, 1),
        "launches": int(prof.walls_launches),
        "formula": "96 N + 32 W B a sweep, N bodies, W active wall contacts (normal impulse above zero, after the "
                   "window)",
    }
    return rounds, walls


def cpu_baseline(pkg, spec, state, seconds, scenes_taken=48):
    """The oracle on the batch's first `scenes_taken` scenes, started from the plain run's state where the three runs
    part, 11 steps before the window (positions, velocities, turns and momenta; the impulses from zero), stepped on
    one thread and on all of them for about `seconds` each."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import scene as sc  # the oracle's loader lives with the tests
    import ctypes
    _, boxes, _ = modules(pkg)
    abi = sc.oracle_abi()
    threads_fn = abi.lib.vro_threads
    threads_fn.argtypes = [ctypes.c_int32]
    threads_fn.restype = ctypes.c_int32
    bodies = spec.bodies[:scenes_taken]
    n = int(bodies.sum())
    out = {}
    for threads in (1, threads_available()):
        threads_fn(threads)
        sim = boxes.BoxBatch(bodies, spec.extents[:scenes_taken], spec.grids[:scenes_taken], spec.substeps,
                             spec.sweeps, abi=abi)
        sim.clear()
        for name in STATE + ("hx", "hy", "hz", "shape", "im"):
            sim.write(name, state[name][:n])
        sim.params(spec.params)
        steps = 0
        t0 = time.perf_counter()
        while True:
            sim.stepAndWait()
            steps += 1
            if time.perf_counter() - t0 >= seconds:
                break
        dt = time.perf_counter() - t0
        sim.close()
        out[threads] = (n * spec.substeps * steps / dt, steps, dt)
    one, most = out[1], out[max(out)]
    return {
        "value": most[0], "unit": "body-substeps/s", "cores": max(out), "kind": "port",
        "one_core": {"value": one[0], "steps": one[1], "seconds": one[2]},
        "cpu": cpu_model(), "nproc": os.cpu_count(),
        "sample": f"the batch's first {scenes_taken} scenes ({n} bodies), {spec.substeps} substeps of {spec.sweeps} "
                  f"sweeps, on the CPU oracle (C++ -O3, the scenes on OpenMP threads), started from the GPU's state "
                  f"11 steps before the plain window (impulses from zero): {most[1]} steps in {most[2]:.1f} s on "
                  f"{max(out)} threads, {one[1]} steps in {one[2]:.1f} s on one",
    }


def plain_only(pkg, args):
    """The plain run alone, for the profilers: its rate and roofline inputs, one JSON line."""
    scenes, _, _ = modules(pkg)
    spec = scenes.batched_2k()
    plain, prof, _ = batch_run(pkg, "plain", spec, args, profile=True)
    rounds, walls = box_roofline(prof, plain["after_window"]["active_wall_contacts"])
    return {"workload": f"batched_2k plain, the last {args.steps} of {args.run} steps", "plain": plain,
            "roofline": rounds, "roofline_walls": walls}


# ---- the profilers' output over the timed window ----

WINDOW_KERNELS = {
    "k_solve_round": ("k_solve_round<1, false, 0>", "k_solve_round<2, false, 0>", "k_solve_round<1, true, 0>",
                      "k_solve_round<2, true, 0>"),
    "k_group_rounds": ("k_group_rounds<1, false, 0>", "k_group_rounds<2, false, 0>", "k_group_rounds<1, true, 0>",
                       "k_group_rounds<2, true, 0>", "k_group_rounds_local<1, false, 0>",
                       "k_group_rounds_local<2, false, 0>", "k_group_rounds_local<1, true, 0>",
                       "k_group_rounds_local<2, true, 0>"),
    "k_walls": ("k_walls<1, false, 0>", "k_walls<2, false, 0>", "k_walls<1, true, 0>", "k_walls<2, true, 0>"),
    "k_group_phase": ("k_group_phase<1, 0, false>", "k_group_phase<2, 0, false>", "k_group_phase<1, 0, true>",
                      "k_group_phase<2, 0, true>"),
}


def window_kind(name):
    for kind, names in WINDOW_KERNELS.items():
        if any(n in name for n in names):
            return kind
    return None


def summarize_steps(trace, steps, out):
    """Markdown of a kernel trace's last `steps` steps, kernel by kernel: dispatches, mean and total duration and share.
    A step ends at its one `k_show` dispatch (the shown words, written last in every step of a sphere or box world),
    so the window is every dispatch after the (steps + 1)-th last `k_show` up to the last. The trace may be gzipped."""
    import csv
    import gzip
    opener = gzip.open if trace.endswith(".gz") else open
    with opener(trace, "rt") as f:
        rows = sorted((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), r["Kernel_Name"])
                      for r in csv.DictReader(f))
    shows = [i for i, r in enumerate(rows) if "k_show" in r[2]]
    if len(shows) <= steps:
        raise SystemExit(f"the trace holds {len(shows)} steps, fewer than {steps} + 1")
    window = rows[shows[-steps - 1] + 1:shows[-1] + 1]

    def kernel(name):
        name = name.replace("void ", "").replace("vr::(anonymous namespace)::", "")
        return name.replace("vrbox::(anonymous namespace)::", "").split("(")[0]

    stats = {}
    for start, end, name in window:
        k = stats.setdefault(kernel(name), [0, 0])
        k[0] += 1
        k[1] += end - start
    busy = sum(v[1] for v in stats.values())
    lines = [f"The last {steps} steps: {len(window)} dispatches, {busy / 1e6 / steps:.2f} ms of kernel time a step.", "",
             "| Kernel | Dispatches | Mean us | Total ms | Share of the window's kernel time |",
             "| --- | --- | --- | --- | --- |"]
    for k, (n, t) in sorted(stats.items(), key=lambda kv: -kv[1][1]):
        lines.append(f"| `{k}` | {n} | {t / n / 1e3:.2f} | {t / 1e6:.2f} | {100 * t / busy:.2f} % |")
    rounds = [v for k, v in stats.items() if k.startswith("k_round_wide") or k.startswith("k_round_tail")]
    if rounds:
        n, t = sum(v[0] for v in rounds), sum(v[1] for v in rounds)
        lines += ["", f"Round kernels (`k_round_wide`, `k_round_tail`): {n} dispatches ({n / steps:.0f} a step), "
                      f"mean {t / n / 1e3:.3f} us, {t / 1e6 / steps:.2f} ms a step."]
    with open(out, "w") as f:
        f.write("\n".join(lines) + "\n")


def summarize(trace, counters, line, out):
    """Markdown of a `plain_only` run's timed window from rocprofv3's output: the solve and relax sweeps' round and
    walls dispatches of the plain method (their template's last argument 0) are the window's last `launches` of each
    kind (the measured steps after it run other instantiations). From a kernel trace, their mean duration; from a
    counter pass (FETCH_SIZE or WRITE_SIZE, KiB a dispatch), their mean reading."""
    import csv
    d = json.loads(line)
    name = d["roofline"]["kernel"]
    rounds_kind = next(k for k in ("k_group_phase", "k_group_rounds", "k_solve_round") if name.startswith(k))
    want = {rounds_kind: d["roofline"]["launches"]}
    if d.get("roofline_walls"):
        want["k_walls"] = d["roofline_walls"]["launches"]
    rows = {k: [] for k in want}
    if trace:
        for r in csv.DictReader(open(trace)):
            kind = window_kind(r["Kernel_Name"])
            if kind:
                rows[kind].append((int(r["Dispatch_Id"]), int(r["End_Timestamp"]) - int(r["Start_Timestamp"]),
                                   int(r["Grid_Size_X"]), r["VGPR_Count"], r["SGPR_Count"]))
    if counters:
        for r in csv.DictReader(open(counters)):
            kind = window_kind(r["Kernel_Name"])
            if kind:
                rows[kind].append((int(r["Dispatch_Id"]), float(r["Counter_Value"]), int(r["Grid_Size"]),
                                   r["VGPR_Count"], r["SGPR_Count"]))
    lines = ["| Kernel | Dispatches | Mean | Of | VGPRs, SGPRs |", "| --- | --- | --- | --- | --- |"]
    buckets = []
    for kind, n in want.items():
        got = sorted(rows[kind])[-n:] if n > 0 else []
        if not got:
            lines.append(f"| `{kind}` | 0 | - | the window's dispatches not found | - |")
            continue
        mean = sum(g[1] for g in got) / len(got)
        what = "duration, us" if trace else "counter reading, KiB a dispatch"
        value = mean / 1e3 if trace else mean
        regs = sorted({f"{g[3]}, {g[4]}" for g in got})
        lines.append(f"| `{kind}` (plain solve and relax sweeps) | {len(got)} | {value:.3f} | {what} | "
                     f"{'; '.join(regs)} |")
        if kind == "k_solve_round":  # launch by launch, each a round of its own size
            by = {}
            for g in got:
                by.setdefault(max(g[2], 1).bit_length() - 1, []).append(g)
            total = sum(g[1] for g in got)
            buckets.append("")
            buckets.append(f"`k_solve_round` by launch size (lanes, two a pair), mean {what}:")
            buckets.append("")
            buckets.append("| Lanes | Dispatches | Mean lanes | Mean | Share of the kernel's total |")
            buckets.append("| --- | --- | --- | --- | --- |")
            for b in sorted(by):
                v = by[b]
                m = sum(g[1] for g in v) / len(v)
                buckets.append(f"| 2^{b} .. 2^{b + 1} | {len(v)} | {sum(g[2] for g in v) / len(v):.0f} | "
                               f"{m / 1e3 if trace else m:.3f} | {sum(g[1] for g in v) / total:.3f} |")
    with open(out, "w") as f:
        f.write("\n".join(lines + buckets) + "\n")


def batched(pkg, args):
    scenes, boxes, mlp = modules(pkg)
    spec = scenes.batched_2k()
    n = int(spec.bodies.sum())
    # The network, trained on other batches' plain runs.
    parts, seen = training_samples(pkg, args.train_batches, 20, 2)
    samples = int(sum(len(p[0]) for p in parts))
    train_x = np.concatenate([p[0] for p in parts[:-1]])
    train_t = np.concatenate([p[1] for p in parts[:-1]])
    held_x, held_t = parts[-1]
    net = mlp.Mlp(seed=3)
    t0 = time.perf_counter()
    net.train(train_x, train_t, epochs=args.epochs, batch=4096, rate=1e-3, seed=1)
    train_s = time.perf_counter() - t0
    err_net = float(np.mean(np.sum((net.forward(held_x) - held_t) ** 2, axis=1)))
    err_carry = float(np.mean(np.sum((held_x[:, 7:10] - held_t) ** 2, axis=1)))
    del train_x, train_t, parts

    plain, prof, state = batch_run(pkg, "plain", spec, args, profile=True, keep_state=True)
    tol = args.tolerance
    stopped, _, _ = batch_run(pkg, "stopped", spec, args, tol=tol)
    learned, _, _ = batch_run(pkg, "learned", spec, args, net=net, tol=tol)
    net.close()
    rounds, walls = box_roofline(prof, plain["after_window"]["active_wall_contacts"])
    within = lambda a, b: abs(a - b) <= 0.01 * abs(b)  # noqa: E731
    # Each run's measured step from the state all three share (before the window): the start and the stop alone
    # make the difference. After the window the runs have stepped 31 steps apart, so their piles differ.
    compare = {"measured": "one step from the state the three runs share, before the window"}
    for key in ("residual_max_m_per_s", "residual_mean_m_per_s", "deepest_point_m"):
        a, b = learned["from_shared_state"][key], stopped["from_shared_state"][key]
        compare[key] = {"stopped": b, "learned": a, "within_1pct": within(a, b),
                        "plain": plain["from_shared_state"][key]}
    learned["network"] = {
        "samples": samples, "held_out_samples": int(len(held_x)), "scenes_logged": seen, "epochs": args.epochs,
        "train_s": train_s, "trained_on": f"{args.train_batches - 1} plain batches of 2,048 scenes (seeds 100..), "
                                          f"2 logged steps each after 20; held out: the last batch",
        "held_out_sq_error": {"network": err_net, "carried_impulse": err_carry},
    }
    line = {
        "metric": METRIC, "value": learned["value"], "unit": "body-substeps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": learned["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {
            "workload": f"batched_2k: {len(spec.bodies)} scenes, {n} bodies (half boxes, half balls), "
                        f"{spec.substeps} substeps of at most {spec.sweeps} solve and {spec.sweeps} relax sweeps, "
                        f"mu 0.6, seed {spec.seed}; the last {args.steps} of {args.run} steps",
            "bodies": n, "scenes": len(spec.bodies), "substeps": spec.substeps, "sweeps": spec.sweeps,
            "run_steps": args.run, "tolerance_m_per_s": tol,
            "method": "batched + learned warm start, each scene's sweeps stopped at the tolerance",
        },
        "ratio_learned_to_stopped": learned["value"] / stopped["value"],
        "ratio_learned_to_plain": learned["value"] / plain["value"],
        "ratio_note": "learned / stopped is the learned start's gain at equal residual (both stop each scene at the "
                      "tolerance); learned / plain also counts the stop itself, against a method that runs every "
                      "sweep and ends each scene at a lower residual",
        "equal_residual": compare,
        "plain": plain, "stopped": stopped, "learned": learned,
        "roofline": rounds, "roofline_walls": walls,
    }
    if args.cpu_seconds > 0:
        line["cpu_baseline"] = cpu_baseline(pkg, spec, state, args.cpu_seconds)
    return line


# ---- the sphere pile ----

def sphere_pile(pkg, steps, warmup, run=600, spec=None):
    """A sphere pile (the 65,536-sphere pile unless `spec` names another), the last `steps` of its `run` steps, with
    the round kernels' roofline."""
    scenes, _, _ = modules(pkg)
    spec = spec if spec is not None else scenes.pile_64k()
    sim = pkg.SphereSimulation(spec.step)
    sim.start(spec.x, spec.y, spec.z, spec.u, spec.v, spec.w, spec.r, spec.im, spec.params)
    settle = run - steps - warmup
    for k in range(settle):
        sim.stepAndWait()
        progress("pile_64k settle", k + 1, settle)
    contacts_start = int(sim.ints("contactCount")[0])
    for _ in range(warmup):
        sim.stepAndWait()
    sim._world.profile(True)
    sim.finish()
    t0 = time.perf_counter()
    reports = []
    for k in range(steps):
        reports.append(sim.stepAndWait().solve)
    sim.finish()
    seconds = time.perf_counter() - t0
    prof = sim._world.profile_read()
    contacts_end = int(sim.ints("contactCount")[0])
    sim.close()
    n, substeps = spec.step.spheres, spec.step.substeps
    bytes_total = 72.0 * prof.listed_sum + 104.0 * n * prof.passes
    round_s = prof.round_ns * 1e-9
    achieved = bytes_total / round_s / 1e9 if round_s > 0 else 0.0
    return {
        "workload": f"{spec.name}: {n} spheres, Gauss-Seidel until done, {substeps} substeps of "
                    f"{spec.step.iterations} passes, the last {steps} of {run} steps, seed {spec.seed}",
        "value": n * substeps * steps / seconds, "unit": "body-substeps/s", "ms_per_step": seconds * 1e3 / steps,
        "contacts_at_start": contacts_start, "contacts_at_end": contacts_end,
        "rounds_per_step": float(np.mean([r.rounds for r in reports])),
        "roofline": {
            "bound": "hbm",
            "limited_by": "dependent latency, launches and device-scope claim atomics (the world sits in the 256 MB "
                          "Infinity Cache)",
            "kernel": "Gauss-Seidel round kernels (k_round_wide, k_round_tail), per pass",
            "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
            "traffic": SPHERE_ROUND_TRAFFIC_BYTES,
            "bytes_per_launch": bytes_total / max(prof.round_launches, 1),
            "avg_launch_us": prof.round_ns / 1e3 / max(prof.round_launches, 1),
            "host_enqueue_ms_per_step": prof.enqueue_ns / 1e6 / max(prof.steps, 1),
            "formula": "72 C + 104 N B a pass, C listed contacts, N spheres",
        },
    }


# The reference's published ms a step for its measured scenes, on an RTX 5070 Ti through Vulkan: the sphere scenes
# with friction mu = 0.5, 10 substeps, run until done; the box piles, mu = 0.5, 10 substeps of 4 sweeps.
SWEEP_PUBLISHED = {
    "sweep_column": (1.83, "docs/TODO.md:321"),
    "sweep_pile_343": (2.39, "docs/TODO.md:322"),
    "sweep_pile_4096": (2.90, "docs/TODO.md:367-368"),
    "box_pile_343": (6.9, "vexelray-sim-rigid-core/src/test/java/dev/vexelray/sim/rigid/box/BoxStackTest.java:165"),
    "box_pile_1000": (7.9, "vexelray-sim-rigid-core/src/test/java/dev/vexelray/sim/rigid/box/BoxStackTest.java:178"),
}


def reference_scenes(pkg, only=""):
    """The reference's measured scenes, each timed as its test does: the sphere scenes as SphereSweepTest.run
    (SphereSweepTest.java:175-182), 2 frames untimed and then the rest of the run's frames; the box piles as
    BoxStackTest.pile (:213-216), 3 s untimed and then 1 s; each step waited for, on the wall clock. `only` names one
    scene (sweep_column, sweep_pile_343, sweep_pile_4096, box_pile_343, box_pile_1000) to time alone."""
    scenes, boxes, _ = modules(pkg)
    out = {}
    for spec in (scenes.sweep_column(), scenes.sweep_pile(7), scenes.sweep_pile(16)):
        if only and spec.name != only:
            continue
        sim = pkg.SphereSimulation(spec.step)
        sim.start(spec.x, spec.y, spec.z, spec.u, spec.v, spec.w, spec.r, spec.im, spec.params)
        for _ in range(2):
            sim.stepAndWait()
        sim.finish()
        t0 = time.perf_counter()
        rounds = 0
        for _ in range(spec.steps - 2):
            rounds += sim.stepAndWait().solve.rounds
        seconds = time.perf_counter() - t0
        sim.close()
        ms = seconds * 1e3 / (spec.steps - 2)
        published, source = SWEEP_PUBLISHED[spec.name]
        out[spec.name] = {
            "spheres": spec.step.spheres, "frames_timed": spec.steps - 2, "ms_per_step": ms,
            "value": spec.step.spheres * spec.step.substeps / (ms * 1e-3), "unit": "body-substeps/s",
            "rounds_per_step": rounds / (spec.steps - 2),
            "reference_ms_per_step": published, "reference_source": source + " (RTX 5070 Ti, Vulkan)",
            "reference_over_this": published / ms,
        }
        progress("reference scenes", len(out), 5)
    for spec in (scenes.box_pile(7), scenes.box_pile(10)):
        if only and spec.name != only:
            continue
        sim = boxes.BoxSimulation(spec.step)
        sim.start(spec.x, spec.y, spec.z, spec.u, spec.v, spec.w, spec.spin, spec.q, spec.half, spec.shape, spec.im,
                  spec.params)
        untimed = 180
        for _ in range(untimed):
            sim.stepAndWait()
        sim.finish()
        t0 = time.perf_counter()
        for _ in range(spec.steps - untimed):
            rep = sim.stepAndWait()
        seconds = time.perf_counter() - t0
        sim.close()
        ms = seconds * 1e3 / (spec.steps - untimed)
        published, source = SWEEP_PUBLISHED[spec.name]
        out[spec.name] = {
            "boxes": spec.step.boxes, "frames_timed": spec.steps - untimed, "ms_per_step": ms,
            "value": spec.step.boxes * spec.step.substeps / (ms * 1e-3), "unit": "body-substeps/s",
            "touching_at_end": rep.touching, "rounds_at_end": rep.rounds,
            "reference_ms_per_step": published, "reference_source": source + " (RTX 5070 Ti, Vulkan)",
            "reference_over_this": published / ms,
        }
        progress("reference scenes", len(out), 5)
    return out


def box_scene(pkg, args, run=0):
    """A box-path workload of one scene (mixed_32k): body-substeps per second, plain stepping, over the last
    `args.steps` of `run` steps (at least steps + warm-up)."""
    scenes, boxes, _ = modules(pkg)
    spec = scenes.mixed_32k()
    sim = boxes.BoxSimulation(spec.step)
    sim.start(spec.x, spec.y, spec.z, spec.u, spec.v, spec.w, spec.spin, spec.q, spec.half, spec.shape, spec.im,
              spec.params)
    settle = max(run - args.steps, args.warmup)
    for k in range(settle):
        sim.stepAndWait()
        progress("mixed_32k settle", k + 1, settle)
    t0 = time.perf_counter()
    reps = []
    for k in range(args.steps):
        reps.append(sim.stepAndWait())
        progress("mixed_32k", k + 1, args.steps)
    seconds = time.perf_counter() - t0
    _, _, broken, depth = sim.diagnostics()
    sim.close()
    n, substeps = spec.step.boxes, spec.step.substeps
    return {
        "metric": METRIC, "value": n * substeps * args.steps / seconds, "unit": "body-substeps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": seconds * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{spec.name}: {n} bodies (16,384 boxes, 16,384 balls, {spec.towers} towers of 20 "
                               f"cubes), box path, {substeps} substeps of {spec.step.sweeps} sweeps, seed {spec.seed}, "
                               f"the last {args.steps} of {settle + args.steps} steps",
                   "bodies": n, "method": "plain", "touching_at_end": reps[-1].touching,
                   "rounds_at_end": reps[-1].rounds, "deepest_point_at_end": depth, "broken": broken},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--run", type=int, default=300, help="the configuration's steps; the window is the last --steps")
    ap.add_argument("--workload", default="batched_2k")
    ap.add_argument("--cpu-seconds", type=float, default=8.0)
    ap.add_argument("--train-batches", type=int, default=5, help="batches of 2,048 scenes logged for the network")
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--tolerance", type=float, default=0.01, help="m/s: the residual each scene's sweeps stop at")
    ap.add_argument("--c1-steps", type=int, default=20, help="timed steps of the sphere pile (0: skip it)")
    ap.add_argument("--others", type=int, default=1,
                    help="1: also time the 1,000-sphere pile (last 20 of 300 steps), the mixed box pile (last 20 "
                         "of 600) and the reference's measured sphere scenes as sub-keys")
    ap.add_argument("--only", default="", help="plain: the plain run alone (for the profilers); scenes: the "
                    "reference's measured scenes alone (scenes:NAME, one of them); one JSON line")
    ap.add_argument("--summarize", nargs=3, metavar=("TRACE_OR_-", "COUNTERS_OR_-", "LINE_JSON"),
                    help="write the timed window's kernels from rocprofv3 CSVs (kernel trace or counter collection) "
                         "and a --only plain line to --summary-out, then exit")
    ap.add_argument("--summarize-steps", nargs=2, metavar=("TRACE", "STEPS"),
                    help="write the last STEPS steps' kernels of a rocprofv3 kernel trace (csv, or csv.gz) to "
                         "--summary-out, then exit")
    ap.add_argument("--summary-out", default="window.md")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="MiB: only copy a buffer of that size 5 times (the known-size kernel that calibrates the "
                         "FETCH_SIZE / WRITE_SIZE counters), then exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the library steps one world on one MI355X and has no multi-GPU path "
              f"(asked for --gpus {args.gpus})", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe > 0:
        import torch
        src = torch.ones(args.copy_probe << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": src.numel() * 4, "copy_probe_bytes_written": dst.numel() * 4}))
        return
    if args.summarize_steps:
        summarize_steps(args.summarize_steps[0], int(args.summarize_steps[1]), args.summary_out)
        return
    if args.summarize:
        trace, counters, line = args.summarize
        summarize(None if trace == "-" else trace, None if counters == "-" else counters,
                  open(line).read().strip().splitlines()[-1], args.summary_out)
        return
    args.run = max(args.run, args.steps + args.warmup)
    pkg = load_package()
    if args.only == "plain":
        print(json.dumps(plain_only(pkg, args)))
        return
    if args.only == "scenes" or args.only.startswith("scenes:"):
        print(json.dumps(reference_scenes(pkg, args.only[len("scenes:"):])))
        return
    if args.workload == "mixed_32k":
        print(json.dumps(box_scene(pkg, args)))
        return
    if args.workload == "pile_64k":
        line = sphere_pile(pkg, args.steps, args.warmup)
        line.update({"metric": METRIC, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
                     "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
                     "data": "synthetic", "config": {"workload": line.pop("workload")}})
        print(json.dumps(line))
        return
    line = batched(pkg, args)
    if args.c1_steps > 0:
        line["pile_64k"] = sphere_pile(pkg, args.c1_steps, 2)
    if args.others:
        scenes, _, _ = modules(pkg)
        line["pile_1k"] = sphere_pile(pkg, 20, 2, run=300, spec=scenes.pile_1k())
        mixed = box_scene(pkg, argparse.Namespace(steps=20, warmup=2), run=600)
        line["mixed_32k"] = {"workload": mixed["config"]["workload"], "value": mixed["value"],
                             "unit": mixed["unit"], "ms_per_step": mixed["ms_per_step"],
                             "touching_at_end": mixed["config"]["touching_at_end"],
                             "rounds_at_end": mixed["config"]["rounds_at_end"]}
        line["reference_scenes"] = reference_scenes(pkg)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
