"""The HIP path against the oracle over long runs and at the configurations' own step shapes, bit for bit.

* the 65,536-sphere pile over its first 50 steps, still falling and landing (state, reports, contact lists);
* the 33,408-body mixed box pile over 20 steps, taken five at a time so each test stays short (state, wall impulses,
  reports);
* the batched method's own step, 8 substeps of 10 sweeps, over 128 scenes, each scene against the oracle's world of
  that scene;
* static and kinetic friction apart (mu_s 0.6, mu_k 0.3) on the box batch in both of its step modes and on a sphere
  pile with friction and rolling resistance, where a swapped pair of coefficients would show.

Bit for bit is stronger than the long-run contract (maximum penetration and residual within 1%): every word of state
is compared after every chunk.
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


def same(a, b, what):
    for k in a:
        assert np.array_equal(np.asarray(a[k]).view(np.int32), np.asarray(b[k]).view(np.int32)), (what, k)


def test_the_64k_sphere_pile_matches_over_50_steps():
    out = {}
    for backend in ("oracle", "gpu"):
        spec = scenes.pile_64k()
        sim = pkg.SphereSimulation(spec.step, abi=sc.abi_for(backend))
        sim.start(spec.x, spec.y, spec.z, spec.u, spec.v, spec.w, spec.r, spec.im, spec.params)
        reps = []
        for _ in range(50):
            r = sim.stepAndWait().solve
            reps.append((r.waits, r.rounds, r.mostRounds, r.unlisted))
        state = {k: sim.floats(k).copy() for k in SPHERE_STATE}
        count = sim.ints("contactCount").copy()
        listed = int(min(count[0], len(sim.ints("contactA"))))
        state["contactA"] = sim.ints("contactA")[:listed].copy()
        state["contactB"] = sim.ints("contactB")[:listed].copy()
        out[backend] = (reps, count, state)
        sim.close()
    (ra, ca, a), (rb, cb, b) = out["oracle"], out["gpu"]
    assert ra == rb
    assert np.array_equal(ca, cb)
    assert ca[0] > 0
    same(a, b, "pile_64k after 50 steps")


@pytest.fixture(scope="module")
def mixed_pair():
    sims = {}
    for backend in ("oracle", "gpu"):
        spec = scenes.mixed_32k()
        sim = boxes.BoxSimulation(spec.step, abi=sc.abi_for(backend))
        sim.start(spec.x, spec.y, spec.z, spec.u, spec.v, spec.w, spec.spin, spec.q, spec.half, spec.shape, spec.im,
                  spec.params)
        sims[backend] = sim
    yield sims
    for sim in sims.values():
        sim.close()


@pytest.mark.timeout(240)
@pytest.mark.parametrize("chunk", [1, 2, 3, 4])
def test_the_mixed_box_pile_matches_over_20_steps(mixed_pair, chunk):
    """Steps 5 (chunk - 1) + 1 .. 5 chunk of the same two runs (the chunks run in order on one pair of worlds)."""
    got = {}
    for backend, sim in mixed_pair.items():
        reps = []
        for _ in range(5):
            r = sim.stepAndWait()
            reps.append((r.touching, r.points, r.rounds, r.unplaced, r.unlisted))
        got[backend] = (reps, {k: sim.floats(k).copy() for k in BOX_STATE + ("lambda", "rub", "roll", "impulses")})
    (ra, a), (rb, b) = got["oracle"], got["gpu"]
    assert ra == rb
    assert ra[-1][0] > 0
    same(a, b, f"mixed_32k after {5 * chunk} steps")


def batch_run(spec, backend, steps):
    sim = boxes.BoxBatch(spec.bodies, spec.extents, spec.grids, spec.substeps, spec.sweeps, abi=sc.abi_for(backend))
    sim.start(spec.x, spec.y, spec.z, spec.u, spec.v, spec.w, spec.spin, spec.q, spec.half, spec.shape, spec.im,
              spec.params)
    reps = []
    for _ in range(steps):
        r = sim.stepAndWait()
        reps.append((r.touching, r.points, r.rounds, r.unplaced, r.unlisted))
    state = {k: sim.floats(k).copy() for k in BOX_STATE}
    sim.close()
    return reps, state


def test_the_batched_method_at_8_substeps_of_10_sweeps():
    """The batched configuration's own step (8 substeps, 10 sweeps of the solve and of the relax) over 128 scenes of
    16-512 bodies, half boxes and half balls, 4 steps."""
    spec = scenes.batched_2k(seed=21, scenes=128, fewest=16, most=512)
    assert (spec.substeps, spec.sweeps) == (8, 10)
    ra, a = batch_run(spec, "oracle", 4)
    rb, b = batch_run(spec, "gpu", 4)
    assert ra == rb
    assert ra[-1][0] > 0
    same(a, b, "128 scenes at 8 x 10")


def with_friction_apart(spec, mu_static=0.6, mu_kinetic=0.3):
    p = spec.params.copy()
    p[boxes.Boxes.MU_STATIC] = np.float32(mu_static).view(np.int32)
    p[boxes.Boxes.MU_KINETIC] = np.float32(mu_kinetic).view(np.int32)
    spec.params = p
    return spec


@pytest.mark.parametrize("mode", ["scene", "pass"])
def test_box_friction_apart_matches_in_each_step_mode(mode, monkeypatch):
    monkeypatch.setenv("VR_BOX_STEP_MODE", mode)
    spec = with_friction_apart(scenes.batched_2k(seed=9, scenes=96, fewest=4, most=60, substeps=4, sweeps=4))
    ra, a = batch_run(spec, "oracle", 6)
    rb, b = batch_run(spec, "gpu", 6)
    assert ra == rb
    same(a, b, f"friction apart, {mode} mode")


def test_sphere_friction_apart_matches():
    out = {}
    for backend in ("oracle", "gpu"):
        s = sc.pile()
        s.muStatic, s.muKinetic = 0.6, 0.3
        s.muRolling = 0.05
        s.solve = pkg.Solve.GAUSS_SEIDEL_BY_CONTACT
        s.start(backend).advance(20)
        out[backend] = {k: s.floats(k).copy() for k in sc.STATE}
        s.close()
    same(out["oracle"], out["gpu"], "sphere pile, mu_s 0.6, mu_k 0.3")
