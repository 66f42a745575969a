"""Benchmark: the Gauss-Seidel sphere step on the 65,536-sphere pile, body-substeps per second on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload pile_64k|mixed_32k] [--cpu-seconds S]

The workload is the pile of 65,536 spheres (radius log-uniform 0.1-0.5 m, density 1000 kg/m^3, dropped with
N(0, 1) m/s velocities into a 100 x 100 m walled floor; mu 0.4, e 0.1; dt 1/60 s in 8 substeps of 10 passes, 1 m
grid cells), generated from a seeded java.util.Random. W untimed steps run first, then K timed steps (the config's
600 by default). All state is resident in HBM before the timed region; nothing crosses PCIe inside it except the
step's small report.

One JSON line is printed: value = spheres x substeps x K / seconds. It carries:
  batched: the 2,048-scene box batch (about 1.5 M bodies), body-substeps per second stepped plain (the reference's
    carried warm start) and with the learned warm start (the network trained on the device for a few epochs on
    samples logged from the plain steps), with the network's and the carry's held-out error on those samples.
  roofline: the Gauss-Seidel round kernels (k_round_wide ... k_round_tail, launched back to back for each pass),
    timed with HIP events on the world's stream around every pass over the timed region; algorithmic bytes per
    pass 72 C + 104 N (C contacts listed that substep, N spheres), against MI355X's 8 TB/s.
  cpu_baseline: the CPU oracle (oracle/, a single-threaded C++ restatement of the reference's passes) timed on a
    bounded sample: steps started from the GPU's final state, for about --cpu-seconds seconds.
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
# HBM bytes a round-kernel launch, measured with rocprofv3 (FETCH_SIZE and WRITE_SIZE passes, each scaled by what it
# reads for a copy of known size) over the first 100 steps of pile_64k: profiles/traffic_round_kernels_pile_64k.md.
ROUND_TRAFFIC_BYTES = 5904187.0
ROUND_TRAFFIC_SAMPLE = ("rocprofv3 FETCH_SIZE x2.000 + WRITE_SIZE x1.000 (copy-calibrated) over the first 100 steps "
                        "of pile_64k, all round-kernel launches; 892,506 algorithmic B a launch over the same steps")


def cpu_baseline(pkg, spec, state, seconds):
    """The oracle on the same workload, started from `state`, for about `seconds` of single-core CPU time."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import scene as sc  # the oracle's loader lives with the tests
    sim = pkg.SphereSimulation(spec.step, abi=sc.oracle_abi())
    x, y, z, u, v, w, ax, ay, az, qx, qy, qz, qw = state
    sim.start(x, y, z, u, v, w, spec.r, spec.im, spec.params, ax=ax, ay=ay, az=az, shown=False)
    for name, arr in (("qx", qx), ("qy", qy), ("qz", qz), ("qw", qw)):
        sim.write(name, arr)
    steps = 0
    t0 = time.perf_counter()
    while True:
        sim.stepAndWait()
        steps += 1
        if time.perf_counter() - t0 >= seconds:
            break
    dt = time.perf_counter() - t0
    sim.close()
    return {
        "value": spec.step.spheres * spec.step.substeps * steps / dt,
        "unit": "body-substeps/s",
        "cores": 1,
        "kind": "port",
        "sample": f"{steps} steps of the same workload on the CPU oracle (single thread, -O3), started from the "
                  f"GPU's state after its timed steps (the settled pile), {dt:.1f} s",
    }


def batched(pkg, steps, warmup, epochs):
    """The 2,048-scene batch on the box path: plain (the reference's carry as the warm start), then with the learned
    warm start, trained on the device on samples logged from the plain steps. Body-substeps per second for both,
    timed over the same steps from the same state: two batches step alike (the stepping is deterministic) through the
    warm-up and the logged steps; one then runs plain, the other with the network."""
    scenes = __import__(pkg.__name__ + ".scenes", fromlist=["scenes"])
    boxes = __import__(pkg.__name__ + ".boxes", fromlist=["boxes"])
    mlp = __import__(pkg.__name__ + ".mlp", fromlist=["mlp"])
    spec = scenes.batched_2k()
    n = int(spec.bodies.sum())

    def started():
        sim = boxes.BoxBatch(spec.bodies, spec.extents, spec.grids, spec.substeps, spec.sweeps)
        sim.start(spec.x, spec.y, spec.z, spec.u, spec.v, spec.w, spec.spin, spec.q, spec.half, spec.shape,
                  spec.im, spec.params)
        for _ in range(warmup + 3):
            sim.stepAndWait()
        return sim

    def timed(sim):
        t0 = time.perf_counter()
        reps = [sim.stepAndWait() for _ in range(steps)]
        return time.perf_counter() - t0, reps

    sim = boxes.BoxBatch(spec.bodies, spec.extents, spec.grids, spec.substeps, spec.sweeps)
    sim.start(spec.x, spec.y, spec.z, spec.u, spec.v, spec.w, spec.spin, spec.q, spec.half, spec.shape, spec.im,
              spec.params)
    for _ in range(warmup):
        sim.stepAndWait()
    sim.log(True, 4 << 20)
    for _ in range(3):
        sim.stepAndWait()
    x, t = sim.samples(4 << 20)
    sim.log(False)
    plain_s, reps = timed(sim)
    plain_end = [sim.floats(k).copy() for k in ("x", "y", "z")]
    sim.close()
    net = mlp.Mlp(seed=3)
    held = np.random.default_rng(0).permutation(len(x))
    cut = int(0.9 * len(x))
    t0 = time.perf_counter()
    net.train(x[held[:cut]], t[held[:cut]], epochs=epochs, batch=4096, rate=1e-3, seed=1)
    train_s = time.perf_counter() - t0
    pred = net.forward(x[held[cut:]])
    err_net = float(np.mean(np.sum((pred - t[held[cut:]]) ** 2, axis=1)))
    err_carry = float(np.mean(np.sum((x[held[cut:]][:, 7:10] - t[held[cut:]]) ** 2, axis=1)))
    sim = started()
    sim.learned(net)
    learned_s, lreps = timed(sim)
    used = sim.learnedPoints
    # How far the learned start moved the batch from the plain one over the same steps.
    apart = max(float(np.max(np.abs(sim.floats(k) - plain_end[a]))) for a, k in enumerate(("x", "y", "z")))
    sim.close()
    net.close()
    rate = lambda secs: n * spec.substeps * steps / secs  # noqa: E731
    return {
        "workload": f"batched_2k: {len(spec.bodies)} scenes, {n} bodies (half boxes, half balls), {spec.substeps} "
                    f"substeps of {spec.sweeps} sweeps, seed {spec.seed}",
        "steps_timed": steps, "warmup": warmup,
        "timed_from": f"step {warmup + 3}, the same state for both",
        "plain": {"value": rate(plain_s), "ms_per_step": plain_s * 1e3 / steps,
                  "touching_at_end": reps[-1].touching, "rounds_at_end": reps[-1].rounds},
        "learned": {"value": rate(learned_s), "ms_per_step": learned_s * 1e3 / steps,
                    "touching_at_end": lreps[-1].touching, "rounds_at_end": lreps[-1].rounds,
                    "points_started_from_prediction": int(used), "samples": int(len(x)), "epochs": epochs,
                    "train_s": train_s, "held_out_sq_error": {"network": err_net, "carried_impulse": err_carry},
                    "largest_position_difference_from_plain_m": apart},
        "unit": "body-substeps/s",
    }


def residual_sweeps(pkg, scenes_count=256, settle=200, epochs=10):
    """The learned warm start by the residual it leaves: a batch of `scenes_count` scenes settles for `settle` steps,
    logs three more, and the network trains on the samples; then, from that same state, one step runs plain and one
    with the network at each sweep count, and the residual after it is measured: the penetration at every touching
    point (the step's listed pairs put through the library's narrow phase on the final state) and of every body past
    its scene's walls, over the body's size (a ball's radius, a box's smallest half-extent)."""
    scenes = __import__(pkg.__name__ + ".scenes", fromlist=["scenes"])
    boxes = __import__(pkg.__name__ + ".boxes", fromlist=["boxes"])
    mlp = __import__(pkg.__name__ + ".mlp", fromlist=["mlp"])
    spec = scenes.batched_2k(scenes=scenes_count)
    scene_of = np.repeat(np.arange(scenes_count), spec.bodies)
    ext = np.asarray(spec.extents, dtype=np.float64)[scene_of]
    half = spec.half.astype(np.float64)
    ball = spec.shape > 0.5
    size = np.where(ball, half[0], half.min(axis=0))
    kept = ("x", "y", "z", "u", "v", "w", "lx", "ly", "lz", "qx", "qy", "qz", "qw", "lambda", "rub", "roll",
            "pairs", "pairsWas", "manifold", "manifoldWas", "impulses", "impulsesWas")

    def make(sweeps):
        b = boxes.BoxBatch(spec.bodies, spec.extents, spec.grids, spec.substeps, sweeps)
        b.start(spec.x, spec.y, spec.z, spec.u, spec.v, spec.w, spec.spin, spec.q, spec.half, spec.shape, spec.im,
                spec.params)
        return b

    def residual(b):
        pos = np.stack([b.floats(k).astype(np.float64) for k in ("x", "y", "z")])
        qx, qy, qz, qw = (b.floats(k).astype(np.float64) for k in ("qx", "qy", "qz", "qw"))
        pairs = b.ints("pairs").reshape(-1, 2)
        listed = pairs[pairs[:, 1] >= 0]
        parts = []
        if len(listed):
            m = b._world.collide(listed.reshape(-1)).reshape(len(listed), 40)  # the library's narrow phase
            count = m[:, 3].astype(np.int64)
            depth = np.stack([m[:, 8 + 8 * k + 6] for k in range(4)], axis=1)
            valid = np.arange(4)[None, :] < count[:, None]
            scale = np.minimum(size[listed[:, 0]], size[listed[:, 1]])
            parts.append((depth / scale[:, None])[valid])
        rot = np.array([[1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qz * qw), 2 * (qx * qz + qy * qw)],
                        [2 * (qx * qy + qz * qw), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - qx * qw)],
                        [2 * (qx * qz - qy * qw), 2 * (qy * qz + qx * qw), 1 - 2 * (qx * qx + qy * qy)]])
        reach = np.where(ball[None, :], half[0][None, :], (np.abs(rot) * half[None, :, :]).sum(axis=1))
        parts.append(np.maximum(reach - pos, pos + reach - ext.T).max(axis=0) / size)
        pen = np.concatenate(parts)
        pen = pen[pen > 0]
        return {"max": float(pen.max()), "p99": float(np.percentile(pen, 99)), "mean": float(pen.mean())} \
            if len(pen) else {"max": 0.0, "p99": 0.0, "mean": 0.0}

    sim = make(spec.sweeps)
    for _ in range(settle):
        sim.stepAndWait()
    sim.log(True, 4 << 20)
    for _ in range(3):
        sim.stepAndWait()
    x, t = sim.samples(4 << 20)
    sim.log(False)
    snap = {k: (sim.ints(k).copy() if k in ("pairs", "pairsWas") else sim.floats(k).copy()) for k in kept}
    settled = residual(sim)
    sim.close()
    net = mlp.Mlp(seed=3)
    net.train(x, t, epochs=epochs, batch=4096, rate=1e-3, seed=1)

    def one_step(sweeps, learned):
        b = make(sweeps)
        for k in kept:
            b.write(k, snap[k])
        if learned:
            b.learned(net)
        b.stepAndWait()
        out = residual(b)
        b.close()
        return out

    rows = []
    for sweeps in (1, 2, 3, 5, 8, 10, 15, 20, 30):
        rows.append({"sweeps": sweeps, "plain": one_step(sweeps, False), "learned": one_step(sweeps, True)})
        progress(len(rows), 9)
    net.close()
    return {"workload": f"batched_2k with {scenes_count} scenes, {int(spec.bodies.sum())} bodies, settled "
                        f"{settle} steps, one step from the same state at each sweep count",
            "residual": "penetration over the body's size at touching points and walls (max, 99th percentile, "
                        "mean over penetrating points)",
            "settled": settled, "samples": int(len(x)), "epochs": epochs, "by_sweeps": rows}


_last_progress = [time.perf_counter()]


def progress(done, total):
    """A progress line every 100 steps, or after 30 s without one (steps run slower under a counter profiler)."""
    now = time.perf_counter()
    if done % 100 == 0 or now - _last_progress[0] > 30:
        _last_progress[0] = now
        print(f"[bench] {done}/{total} steps", file=sys.stderr, flush=True)


def box_scene(pkg, spec, args):
    """A box-path workload of one scene (mixed_32k): body-substeps per second, plain stepping."""
    boxes = __import__(pkg.__name__ + ".boxes", fromlist=["boxes"])
    sim = boxes.BoxSimulation(spec.step)
    sim.start(spec.x, spec.y, spec.z, spec.u, spec.v, spec.w, spec.spin, spec.q, spec.half, spec.shape, spec.im,
              spec.params)
    for _ in range(args.warmup):
        sim.stepAndWait()
    t0 = time.perf_counter()
    reps = []
    for k in range(args.steps):
        reps.append(sim.stepAndWait())
        progress(k + 1, args.steps)
    seconds = time.perf_counter() - t0
    diag, speed, broken, depth = sim.diagnostics()
    sim.close()
    n, substeps = spec.step.boxes, spec.step.substeps
    return {
        "metric": METRIC, "value": n * substeps * args.steps / seconds, "unit": "body-substeps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": seconds * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{spec.name}: {n} bodies (16,384 boxes, 16,384 balls, {spec.towers} towers of 20 "
                               f"cubes), box path, {substeps} substeps of {spec.step.sweeps} sweeps, seed {spec.seed}",
                   "bodies": n, "method": "plain", "touching_at_end": reps[-1].touching,
                   "rounds_at_end": reps[-1].rounds, "deepest_point_at_end": depth, "broken": broken},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=600)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="pile_64k")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--batched-steps", type=int, default=30, help="timed steps of the batched workload (0: skip)")
    ap.add_argument("--residual", action="store_true",
                    help="only measure the residual a step leaves, plain and with the learned warm start, at several "
                         "sweep counts from the same state (residual_sweeps), then exit")
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

    pkg = load_package()
    if args.residual:
        print(json.dumps(residual_sweeps(pkg)))
        return
    scenes = __import__(pkg.__name__ + ".scenes", fromlist=["scenes"])
    spec = scenes.SCENES[args.workload]()
    if hasattr(spec, "towers"):  # a box-path scene
        print(json.dumps(box_scene(pkg, spec, args)))
        return
    try:
        import torch
        sync = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except ImportError:
        sync = lambda: None  # noqa: E731

    sim = pkg.SphereSimulation(spec.step)
    sim.start(spec.x, spec.y, spec.z, spec.u, spec.v, spec.w, spec.r, spec.im, spec.params)
    for _ in range(args.warmup):
        sim.stepAndWait()
    sim._world.profile(True)
    sim.finish()
    sync()
    t0 = time.perf_counter()
    reports = []
    for k in range(args.steps):
        rep = sim.stepAndWait()
        reports.append(rep.solve)
        progress(k + 1, args.steps)
    sim.finish()
    sync()
    seconds = time.perf_counter() - t0
    prof = sim._world.profile_read()
    n, substeps = spec.step.spheres, spec.step.substeps
    value = n * substeps * args.steps / seconds

    passes = max(prof.passes, 1)
    bytes_total = 72.0 * prof.listed_sum + 104.0 * n * prof.passes
    round_s = prof.round_ns * 1e-9
    achieved = bytes_total / round_s / 1e9 if round_s > 0 else 0.0
    per_launch_bytes = bytes_total / max(prof.round_launches, 1)
    state = [sim.floats(k).copy() for k in ("x", "y", "z", "u", "v", "w", "ax", "ay", "az", "qx", "qy", "qz", "qw")]
    cc = sim.ints("contactCount")
    sim.close()

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "body-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": seconds * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{args.workload}: {n} spheres, plain XPBD Gauss-Seidel until done, {substeps} substeps, "
                        f"{spec.step.iterations} passes, 1 m cells, box {spec.box}, seed {spec.seed}",
            "bodies": n, "substeps": substeps, "iterations": spec.step.iterations, "steps_timed": args.steps,
            "method": "plain", "contacts_at_end": int(cc[0]),
            "rounds_per_step": float(np.mean([r.rounds for r in reports])),
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "Gauss-Seidel round kernels (k_round_wide, k_round_tail), per pass",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": ROUND_TRAFFIC_BYTES,
            "traffic_sample": ROUND_TRAFFIC_SAMPLE,
            "bytes_per_launch": per_launch_bytes,
            "avg_launch_us": prof.round_ns / 1e3 / max(prof.round_launches, 1),
            "avg_pass_us": prof.round_ns / 1e3 / passes,
            "host_enqueue_ms_per_step": prof.enqueue_ns / 1e6 / max(prof.steps, 1),
        },
    }
    if args.cpu_seconds > 0:
        line["cpu_baseline"] = cpu_baseline(pkg, spec, state, args.cpu_seconds)
    if args.batched_steps > 0:
        line["batched"] = batched(pkg, args.batched_steps, 5, 10)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
