"""A batch of box scenes (vr_batch_create) steps each scene exactly as a world of that scene alone would.

Pairs never join two scenes and the rounds are coloured in slot order, so a scene's rounds, and with them every
body's state, are the same stepped in a batch or alone. The oracle's batch is its own worlds of one scene each; the
HIP batch is one world over all the scenes' bodies. Both are compared, bit for bit, with scenes stepped alone on the
oracle.
"""

import numpy as np
import pytest

import scene as sc

pkg = sc.pkg
boxes = __import__(pkg.__name__ + ".boxes", fromlist=["boxes"])
scenes = __import__(pkg.__name__ + ".scenes", fromlist=["scenes"])
BACKENDS = ["oracle", pytest.param("gpu", marks=pytest.mark.gpu)]
STATE = ("x", "y", "z", "u", "v", "w", "lx", "ly", "lz", "qx", "qy", "qz", "qw")


def small_batch():
    return scenes.batched_2k(seed=8, scenes=7, fewest=1, most=80, substeps=4, sweeps=3)


def start(sim, spec, part=slice(None)):
    sim.start(spec.x[part], spec.y[part], spec.z[part], spec.u[part], spec.v[part], spec.w[part],
              spec.spin[:, part], spec.q[:, part], spec.half[:, part], spec.shape[part], spec.im[part], spec.params)


def alone(spec, k, steps):
    """Scene k in a world of its own on the oracle: its own walled box in the params."""
    first = int(spec.bodies[:k].sum())
    part = slice(first, first + int(spec.bodies[k]))
    n = int(spec.bodies[k])
    grid = spec.grids[k] if n > 1 else None
    sim = boxes.BoxSimulation(boxes.BoxStep(n, spec.substeps, spec.sweeps, False, grid), abi=sc.oracle_abi())
    words = spec.params.copy()
    for a in range(3):
        words[boxes.Boxes.SX + a] = np.float32(spec.extents[k][a]).view(np.int32)
    start(sim, spec, part)
    sim.params(words)
    reps = [sim.stepAndWait() for _ in range(steps)]
    out = {name: sim.floats(name).copy() for name in STATE}
    sim.close()
    return out, reps


@pytest.mark.parametrize("backend", BACKENDS)
def test_batch_steps_each_scene_as_alone(backend):
    spec = small_batch()
    steps = 12
    sim = boxes.BoxBatch(spec.bodies, spec.extents, spec.grids, spec.substeps, spec.sweeps,
                         abi=sc.abi_for(backend))
    start(sim, spec)
    reports = [sim.stepAndWait() for _ in range(steps)]
    got = {name: sim.floats(name).copy() for name in STATE}
    sim.close()
    touching = [0] * steps
    for k in range(len(spec.bodies)):
        want, reps = alone(spec, k, steps)
        part = slice(int(spec.bodies[:k].sum()), int(spec.bodies[:k + 1].sum()))
        for name in STATE:
            assert np.array_equal(got[name][part].view(np.int32), want[name].view(np.int32)), (k, name)
        for t, r in enumerate(reps):
            touching[t] += r.touching
    assert [r.touching for r in reports] == touching
    assert max(touching) > 0, "the scenes should touch"


def test_batch_refuses_bad_scenes():
    spec = small_batch()
    with pytest.raises(pkg.IllegalArgumentException):
        boxes.BoxBatch([3, 0], spec.extents[:2], spec.grids[:2], 2, abi=sc.oracle_abi())


@pytest.mark.gpu
@pytest.mark.parametrize("count", [96, 100])
def test_many_scenes_match_the_oracle(count):
    """Batches against the oracle's, bit for bit, step by step. After a start the HIP library steps a batch once
    pass by pass over the whole chip, once a workgroup a scene through all of a step's substeps, and keeps the faster,
    so each run here takes both ways and the switches between them (which move the wall impulses to the other way's
    places): 96 small scenes, and 100 scenes of up to 4,906 bodies."""
    if count == 96:
        spec = scenes.batched_2k(seed=9, scenes=count, fewest=1, most=60, substeps=3, sweeps=3)
    else:
        spec = scenes.batched_2k(seed=17, scenes=count, fewest=1, most=5000, substeps=3, sweeps=3)
        assert int(spec.bodies.max()) > 256
    out = {}
    for backend in ("oracle", "gpu"):
        sim = boxes.BoxBatch(spec.bodies, spec.extents, spec.grids, spec.substeps, spec.sweeps,
                             abi=sc.abi_for(backend))
        start(sim, spec)
        reps, states = [], []
        for _ in range(8):
            r = sim.stepAndWait()
            reps.append((r.touching, r.points, r.rounds, r.unplaced, r.unlisted))
            states.append({name: sim.floats(name).copy() for name in STATE})
        out[backend] = (reps, states)
        sim.close()
    (ra, sa), (rb, sb) = out["oracle"], out["gpu"]
    assert ra == rb
    assert max(r[0] for r in ra) > 0, "the scenes should touch"
    for t in range(len(sa)):
        for name in STATE:
            assert np.array_equal(sa[t][name].view(np.int32), sb[t][name].view(np.int32)), (t, name)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["scene", "pass"])
def test_a_fixed_step_mode_matches_the_oracle(mode, monkeypatch):
    """VR_BOX_STEP_MODE fixes the batch's step mode (no timing): each way alone, bit for bit against the oracle."""
    monkeypatch.setenv("VR_BOX_STEP_MODE", mode)
    spec = scenes.batched_2k(seed=9, scenes=96, fewest=1, most=60, substeps=3, sweeps=3)
    out = {}
    for backend in ("oracle", "gpu"):
        sim = boxes.BoxBatch(spec.bodies, spec.extents, spec.grids, spec.substeps, spec.sweeps,
                             abi=sc.abi_for(backend))
        start(sim, spec)
        reps = [sim.stepAndWait() for _ in range(4)]
        out[backend] = ([(r.touching, r.points, r.rounds) for r in reps],
                        {name: sim.floats(name).copy() for name in STATE})
        sim.close()
    (ra, sa), (rb, sb) = out["oracle"], out["gpu"]
    assert ra == rb
    for name in sa:
        assert np.array_equal(sa[name].view(np.int32), sb[name].view(np.int32)), name


@pytest.mark.gpu
@pytest.mark.parametrize("symmetric", [False, True])
@pytest.mark.parametrize("tol", [-1.0, 2e-3])
@pytest.mark.parametrize("plan", ["all", "graph", "group", "group:200:128", "group:200:128:local", "group:1:512:local",
                                  "group:200:128:phase", "group:1:512:phase"])
def test_each_launch_plan_matches_the_oracle(plan, tol, symmetric, monkeypatch):
    """VR_BOX_PLAN fixes how a step is launched: `all` launches every sweep of a phase with no host look between them
    (a scene that has stopped passes over the rest), `graph` does the same captured as a HIP graph each step, the
    executable graph updated in place from step to step, `group` runs each sweep's rounds in one launch, a workgroup a
    group of consecutive scenes, or by default a whole phase of sweeps, walls too, in one launch (here also in groups
    of at most 200 bodies, scenes of more alone, on workgroups of 128 threads, so that a group's round takes several
    passes of its workgroup; with the groups' motion records in LDS over a sweep's rounds, or over the phase, in those
    groups and a scene a group). With every other sweep backwards or not. Six steps, each a replay of the step's graph, with eager
    reads of the state between them, and the sweeps stopped at a tolerance or not: bit for bit against the oracle."""
    plan, *group = plan.split(":")
    monkeypatch.setenv("VR_BOX_PLAN", plan)
    if group:
        monkeypatch.setenv("VR_BOX_GROUP_BODIES", group[0])
        monkeypatch.setenv("VR_BOX_GROUP_THREADS", group[1])
        monkeypatch.setenv("VR_BOX_GROUP_LOCAL", "1" if group[2:] in (["local"], ["phase"]) else "0")
        monkeypatch.setenv("VR_BOX_GROUP_PHASE", "1" if group[2:] == ["phase"] else "0")
    spec = scenes.batched_2k(seed=11, scenes=40, fewest=4, most=300, substeps=3, sweeps=5)
    out = {}
    for backend in ("oracle", "gpu"):
        sim = boxes.BoxBatch(spec.bodies, spec.extents, spec.grids, spec.substeps, spec.sweeps, symmetric,
                             abi=sc.abi_for(backend))
        start(sim, spec)
        sim.tolerance(tol, False)
        states = []
        for _ in range(6):
            sim.stepAndWait()
            states.append({name: sim.floats(name).copy() for name in STATE})
        out[backend] = states
        sim.close()
    for t, (a, b) in enumerate(zip(out["oracle"], out["gpu"])):
        for name in STATE:
            assert np.array_equal(a[name].view(np.int32), b[name].view(np.int32)), (plan, tol, t, name)


def test_the_oracle_reports_its_one_plan():
    """vr_box_plan on the oracle: it steps a batch one way, so the plan is the pass and it never times plans; a
    retune changes nothing."""
    spec = small_batch()
    sim = boxes.BoxBatch(spec.bodies, spec.extents, spec.grids, spec.substeps, spec.sweeps, abi=sc.oracle_abi())
    start(sim, spec)
    assert sim.plan() == ("pass", False)
    sim.retune()
    sim.stepAndWait()
    assert sim.plan() == ("pass", False)
    sim.close()


@pytest.mark.gpu
def test_a_batch_times_its_plans_after_a_start_and_a_retune(monkeypatch):
    """A batch times its launch plans over its first steps (two each in turn), then keeps one; vr_box_retune, and a
    change of tolerance, start the timing again; VR_BOX_PLAN fixes the plan, with nothing timed. Whatever the plan,
    the state is the oracle's, bit for bit."""
    monkeypatch.delenv("VR_BOX_PLAN", raising=False)
    spec = scenes.batched_2k(seed=12, scenes=30, fewest=4, most=200, substeps=2, sweeps=3)
    sim = boxes.BoxBatch(spec.bodies, spec.extents, spec.grids, spec.substeps, spec.sweeps)
    ref = boxes.BoxBatch(spec.bodies, spec.extents, spec.grids, spec.substeps, spec.sweeps, abi=sc.oracle_abi())
    for s in (sim, ref):
        start(s, spec)
    assert sim.plan()[1]
    seen = set()
    steps = 0
    while sim.plan()[1]:
        seen.add(sim.plan()[0])
        sim.stepAndWait()
        ref.stepAndWait()
        steps += 1
        assert steps <= 10, "still timing plans after ten steps"
    assert {"pass", "graph", "scene", "group"} <= seen and steps == 2 * len(seen)
    kept = sim.plan()[0]
    sim.stepAndWait()
    ref.stepAndWait()
    assert sim.plan() == (kept, False)
    sim.retune()
    assert sim.plan() == ("pass", True)
    sim.tolerance(1e-3)
    assert sim.plan()[1]
    ref.tolerance(1e-3)
    for _ in range(3):
        sim.stepAndWait()
        ref.stepAndWait()
    for name in STATE:
        assert np.array_equal(sim.floats(name).view(np.int32), ref.floats(name).view(np.int32)), name
    sim.close()
    ref.close()
    monkeypatch.setenv("VR_BOX_PLAN", "group")
    fixed = boxes.BoxBatch(spec.bodies, spec.extents, spec.grids, spec.substeps, spec.sweeps)
    start(fixed, spec)
    assert fixed.plan() == ("group", False)
    fixed.retune()
    assert fixed.plan() == ("group", False)
    fixed.close()
