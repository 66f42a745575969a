"""Diagnostics reduced in f64 by the world (vr_world_diagnostics) against the host's SphereDiagnostics.of and
BoxDiagnostics.of (SphereDiagnostics.java:34-83, BoxDiagnostics.java:18-44) on the same state: overlaps and maxima
equal, sums to f64 rounding (the world adds in another order)."""

import numpy as np
import pytest

import scene as sc
from boxscene import BoxDiagnostics
from test_gpu_box_parity import mixed_pile

pkg = sc.pkg
BACKENDS = ["oracle", pytest.param("gpu", marks=pytest.mark.gpu)]


def close(a, b, scale):
    return abs(a - b) <= 1e-9 * max(scale, 1e-12)


@pytest.mark.parametrize("backend", BACKENDS)
def test_sphere_diagnostics_match_the_host(backend):
    s = sc.crowd(9)
    s.start(backend).advance(30)
    d = s.sim.diagnostics()
    s.read()
    # The world holds the box and gravity as the params' f32 words; the host is given the same values.
    f = lambda v: float(np.float32(v))  # noqa: E731
    h = pkg.SphereDiagnostics.of(s.x, s.y, s.z, s.u, s.v, s.w, s.ax, s.ay, s.az, s.r, s.im, f(s.gy), f(s.sx),
                                 f(s.sy), f(s.sz))
    assert d.maxOverlap == pytest.approx(h.maxOverlap, rel=1e-12, abs=1e-15) and h.maxOverlap > 0
    assert d.maxWall == pytest.approx(h.maxWall, rel=1e-12, abs=1e-15)
    assert d.maxSpeed == pytest.approx(h.maxSpeed, rel=1e-12)
    energy = abs(h.kinetic) + abs(h.potential) + 1e-12
    for name in ("kinetic", "rotational", "potential"):
        assert close(getattr(d, name), getattr(h, name), energy), name
    scale = float(np.abs(h.momentum).max()) + float(np.abs(h.angular).max()) + 1e-12
    assert all(close(a, b, scale) for a, b in zip(d.momentum, h.momentum))
    assert all(close(a, b, scale) for a, b in zip(d.angular, h.angular))
    assert d.broken is False
    s.close()


@pytest.mark.parametrize("backend", BACKENDS)
def test_box_diagnostics_match_the_host(backend):
    s = mixed_pile(80, seed=13)
    s.start(backend)
    for _ in range(15):
        s.sim.stepAndWait()
    d, speed, broken, depth = s.sim.diagnostics()
    h = s.diagnostics()
    scale = abs(h.kinetic) + 1e-12
    assert close(d.kinetic, h.kinetic, scale) and close(d.rotational, h.rotational, scale)
    mscale = float(np.abs(h.momentum).max()) + float(np.abs(h.angular).max()) + 1e-12
    assert all(close(a, b, mscale) for a, b in zip(d.momentum, h.momentum))
    assert all(close(a, b, mscale) for a, b in zip(d.angular, h.angular))
    manifold = s.floats("manifold").reshape(-1, 40)
    counts = manifold[:, 3].astype(int)
    depths = [manifold[p, 8 + 8 * k + 6] for p in range(len(manifold)) for k in range(min(counts[p], 4))]
    assert depth == pytest.approx(max([0.0] + [float(v) for v in depths]), rel=1e-12)
    assert speed > 0 and not broken
    s.close()
