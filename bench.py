"""Benchmark: the Gauss-Seidel sphere step on the 65,536-sphere pile, body-substeps per second on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload pile_64k] [--cpu-seconds S]

The workload is the pile of 65,536 spheres (radius log-uniform 0.1-0.5 m, density 1000 kg/m^3, dropped with
N(0, 1) m/s velocities into a 100 x 100 m walled floor; mu 0.4, e 0.1; dt 1/60 s in 8 substeps of 10 passes, 1 m
grid cells), generated from a seeded java.util.Random. W untimed steps run first, then K timed steps (the config's
600 by default). All state is resident in HBM before the timed region; nothing crosses PCIe inside it except the
step's small report.

One JSON line is printed: value = spheres x substeps x K / seconds. It carries:
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=600)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="pile_64k")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build steps one world on one MI355X and has no multi-GPU path "
              f"(asked for --gpus {args.gpus})", file=sys.stderr)
        sys.exit(2)

    pkg = load_package()
    scenes = __import__(pkg.__name__ + ".scenes", fromlist=["scenes"])
    spec = scenes.SCENES[args.workload]()
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
        if (k + 1) % 100 == 0:
            print(f"[bench] {k + 1}/{args.steps} steps", file=sys.stderr, flush=True)
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
            "traffic": None,
            "bytes_per_launch": per_launch_bytes,
            "avg_launch_us": prof.round_ns / 1e3 / max(prof.round_launches, 1),
            "avg_pass_us": prof.round_ns / 1e3 / passes,
        },
    }
    if args.cpu_seconds > 0:
        line["cpu_baseline"] = cpu_baseline(pkg, spec, state, args.cpu_seconds)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
