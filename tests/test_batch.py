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
