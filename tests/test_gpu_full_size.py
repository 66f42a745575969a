"""The HIP path against the oracle on the measured workloads at their full sizes, bit for bit.

The workloads are the package's seeded scenes: the 1,000-sphere pile over its full 300 steps (the reference's own
CPU-runnable case), the 65,536-sphere pile the bench measures over its first steps (the oracle needs seconds a step
there, so not the bench's 600), and the 33,408-body mixed box pile over one step. State, step reports and, for the
sphere piles, the contact lists are compared word for word.
"""

import numpy as np
import pytest

import scene as sc

pkg = sc.pkg
scenes = __import__(pkg.__name__ + ".scenes", fromlist=["scenes"])
boxes = __import__(pkg.__name__ + ".boxes", fromlist=["boxes"])
pytestmark = pytest.mark.gpu
SPHERE_STATE = ("x", "y", "z", "u", "v", "w", "ax", "ay", "az", "qx", "qy", "qz", "qw")
BOX_STATE = ("x", "y", "z", "u", "v", "w", "lx", "ly", "lz", "qx", "qy", "qz", "qw")


def spheres_on(backend, name, steps):
    spec = scenes.SCENES[name]()
    sim = pkg.SphereSimulation(spec.step, abi=sc.abi_for(backend))
    sim.start(spec.x, spec.y, spec.z, spec.u, spec.v, spec.w, spec.r, spec.im, spec.params)
    reports = []
    for _ in range(steps):
        r = sim.stepAndWait().solve
        reports.append((r.waits, r.rounds, r.mostRounds, r.unlisted))
    out = {k: sim.floats(k).copy() for k in SPHERE_STATE}
    count = sim.ints("contactCount").copy()
    listed = int(min(count[0], len(sim.ints("contactA"))))
    out["contactA"] = sim.ints("contactA")[:listed].copy()
    out["contactB"] = sim.ints("contactB")[:listed].copy()
    sim.close()
    return out, reports, count


@pytest.mark.parametrize("name,steps", [("pile_1k", 300), ("pile_64k", 10)])
def test_sphere_piles_match_the_oracle(name, steps):
    (a, ra, ca), (b, rb, cb) = spheres_on("oracle", name, steps), spheres_on("gpu", name, steps)
    assert ra == rb
    assert np.array_equal(ca, cb)
    assert ca[0] > 0, "the pile should touch"
    for k in a:
        assert np.array_equal(a[k].view(np.int32), b[k].view(np.int32)), k


def test_mixed_box_pile_matches_the_oracle():
    out = {}
    for backend in ("oracle", "gpu"):
        spec = scenes.mixed_32k()
        sim = boxes.BoxSimulation(spec.step, abi=sc.abi_for(backend))
        sim.start(spec.x, spec.y, spec.z, spec.u, spec.v, spec.w, spec.spin, spec.q, spec.half, spec.shape, spec.im,
                  spec.params)
        r = sim.stepAndWait()
        out[backend] = ((r.touching, r.points, r.rounds, r.unplaced, r.unlisted),
                        {k: sim.floats(k).copy() for k in BOX_STATE + ("lambda", "rub", "roll")})
        sim.close()
    (ra, a), (rb, b) = out["oracle"], out["gpu"]
    assert ra == rb
    assert ra[0] > 0, "the towers should touch"
    for k in a:
        assert np.array_equal(a[k].view(np.int32), b[k].view(np.int32)), k
