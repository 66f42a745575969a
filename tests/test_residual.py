"""Sweeps stopped on a residual (vr_box_tolerance), on each backend, and the residual itself pinned by closed forms.

No counterpart in the reference, whose BoxStep runs a fixed count of solve and relax sweeps every substep
(core/.../box/BoxStep.java:199-214). A point's residual in a solve or relax sweep is the size of the impulse change
the sweep made there, normal and tangent together, times the inverse mass along the normal: a velocity, m/s.

* A lone soft contact is solved in one sweep: Box2D's soft step has 1 - massScale - impulseScale = 0, so the second
  sweep changes nothing. A ball sliding on the floor therefore shows, in its first substep's first solve sweep, the
  residual massScale g h sqrt(1 + mu_k^2) (the normal impulse m massScale g h taken up at once and the kinetic cut of
  the tangent), and nothing after; stopped at any small tolerance its phases run two sweeps.
* Measuring changes no result; a tolerance below zero is the reference's fixed count.
* A stack needs more sweeps than a lone contact, its residual falls from sweep to sweep, and a stopped phase ends
  with every scene at or below the tolerance.
"""

import math

import numpy as np
import pytest

import scene as sc
from boxscene import BoxScene, Boxes

pkg = sc.pkg
boxes = __import__(pkg.__name__ + ".boxes", fromlist=["boxes"])
scenes = __import__(pkg.__name__ + ".scenes", fromlist=["scenes"])
BACKENDS = ["oracle", pytest.param("gpu", marks=pytest.mark.gpu)]
G = 9.81
STATE = ("x", "y", "z", "u", "v", "w", "lx", "ly", "lz", "qx", "qy", "qz", "qw")


def sliding_ball(mu):
    s = BoxScene(1).uniform(0.1, 0.1, 0.1, 2.0)
    s.sphere(0, 0.1)
    s.sx = 10
    s.x[0] = 1
    s.y[0] = np.float32(0.1)
    s.z[0] = np.float32(0.5)
    s.u[0] = 2.0
    s.muStatic = s.muKinetic = mu
    s.passes = 4
    return s


@pytest.mark.parametrize("backend", BACKENDS)
@pytest.mark.parametrize("mu", [0.0, 0.3])
def test_a_lone_contact_is_solved_in_one_sweep(backend, mu):
    with sliding_ball(mu) as s:
        s.start(backend)
        s.sim.tolerance(1e-7, measure=True)
        s.advance(1)
        h = s.dt / s.substeps
        hertz = min(Boxes.CONTACT_HERTZ, 0.25 / h)
        mass_scale = Boxes.softness(2 * hertz, Boxes.DAMPING_RATIO, h)[1]  # a wall's contact is the stiff one
        curve = s.sim.solveCurve()
        first, second = float(curve[0, 0, 0]), float(curve[0, 0, 1])
        want = mass_scale * G * h * math.sqrt(1 + mu * mu)
        assert abs(first - want) <= 1e-4 * want, (first, want)
        assert second <= 1e-6 * first, (first, second)
        rep = s.sim.solveReport()
        assert rep.measured and rep.tolerance == pytest.approx(1e-7)
        assert list(rep.scene_sweeps[0]) == [2.0] * s.substeps, rep.scene_sweeps
        assert float(rep.residual[0][0]) <= 1e-7


def pile(backend, tol, measure, steps=6):
    """The 64 cubes and 64 balls of BoxSphereTest's mixed pile, dropped: contacts in stacks."""
    spec = scenes.batched_2k(seed=3, scenes=1, fewest=128, most=128, substeps=4, sweeps=12)
    sim = boxes.BoxBatch(spec.bodies, spec.extents, spec.grids, spec.substeps, spec.sweeps, abi=sc.abi_for(backend))
    sim.start(spec.x, spec.y, spec.z, spec.u, spec.v, spec.w, spec.spin, spec.q, spec.half, spec.shape, spec.im,
              spec.params)
    sim.tolerance(tol, measure)
    for _ in range(steps):
        sim.stepAndWait()
    out = {k: sim.floats(k).copy() for k in STATE}
    rep, curve = sim.solveReport(), (sim.solveCurve() if measure else None)
    sim.close()
    return out, rep, curve


@pytest.mark.parametrize("backend", BACKENDS)
def test_measuring_changes_nothing_and_no_tolerance_is_the_fixed_count(backend):
    plain, rep, _ = pile(backend, -1.0, False)
    measured, mrep, curve = pile(backend, -1.0, True)
    for k in plain:
        assert np.array_equal(plain[k].view(np.int32), measured[k].view(np.int32)), k
    assert not rep.measured and mrep.measured
    assert np.all(mrep.scene_sweeps[:, :4] == 12.0)
    assert curve[0, 0, 0] > 0


@pytest.mark.parametrize("backend", BACKENDS)
def test_a_stack_needs_more_sweeps_and_a_stopped_phase_ends_below_the_tolerance(backend):
    _, rep, curve = pile(backend, -1.0, True)
    solve0 = curve[0, 0, :12]
    assert solve0[0] > solve0[-1], solve0  # the residual falls over the sweeps
    tol = float(solve0[5])
    _, srep, scurve = pile(backend, tol, True)
    for sub in range(4):
        for k in range(2):
            ran = float(srep.scene_sweeps[k][sub])
            assert 1 <= ran <= 12
            if ran < 12:
                assert float(srep.residual[k][sub]) <= tol, (sub, k, srep.residual[k][sub], tol)
    assert float(srep.scene_sweeps[0][0]) > 2.0, "a stack is more than one contact"
    assert float(srep.scene_sweeps[0].sum()) < 12 * 4, "the tolerance cut sweeps"


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["scene", "pass"])
def test_stopped_sweeps_match_the_oracle(mode, monkeypatch):
    """A batch stepped with the sweeps stopped at a tolerance, measured, against the oracle's scenes stepped alone
    with the same stop: every word of state, and each phase's sweeps (the scenes' mean) and last residual."""
    monkeypatch.setenv("VR_BOX_STEP_MODE", mode)
    spec = scenes.batched_2k(seed=12, scenes=96, fewest=4, most=200, substeps=4, sweeps=8)
    got = {}
    for backend in ("oracle", "gpu"):
        sim = boxes.BoxBatch(spec.bodies, spec.extents, spec.grids, spec.substeps, spec.sweeps,
                             abi=sc.abi_for(backend))
        sim.start(spec.x, spec.y, spec.z, spec.u, spec.v, spec.w, spec.spin, spec.q, spec.half, spec.shape, spec.im,
                  spec.params)
        sim.tolerance(2e-3, True)
        reps = []
        for _ in range(5):
            sim.stepAndWait()
            r = sim.solveReport()
            reps.append((r.scene_sweeps.copy(), r.residual.copy()))
        got[backend] = ({k: sim.floats(k).copy() for k in STATE}, reps)
        sim.close()
    (a, ra), (b, rb) = got["oracle"], got["gpu"]
    for k in a:
        assert np.array_equal(a[k].view(np.int32), b[k].view(np.int32)), k
    for t, ((sa, ea), (sb, eb)) in enumerate(zip(ra, rb)):
        assert np.allclose(sa, sb, rtol=1e-6), (t, sa, sb)
        assert np.array_equal(ea.view(np.int32), eb.view(np.int32)), (t, ea, eb)
    assert any(float(sa.sum()) < 8 * 4 * 2 for sa, _ in ra), "the tolerance should stop some phases early"
