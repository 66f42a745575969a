"""The HIP box path against the oracle on the same seeded scenes, through the C ABI.

Manifolds (counts, axis ids, normals, anchors, depths), the broad phase's slots and contacts, the round counts and
the state after several steps are compared. Both sides do the same f32 operations in the same order with no fused
multiply-adds, so they are compared bit for bit; the stated contract (1e-4 relative on state) is checked too.
"""

import math

import numpy as np
import pytest

import scene as sc
from boxscene import BoxContact, BoxScene, collide
from test_boxes_known import broad_scene, random_pairs

pkg = sc.pkg
pytestmark = pytest.mark.gpu
STATE = ("x", "y", "z", "u", "v", "w", "lx", "ly", "lz", "qx", "qy", "qz", "qw")
WALL_IMPULSES = ("lambda", "rub", "roll")  # read back in the reference's body-by-body layout


def test_random_pair_manifolds_match_bitwise():
    bodies = random_pairs()
    a = collide("oracle", 0, bodies)
    b = collide("gpu", 0, bodies)
    for p in range(500):
        ma, mb = a[p * 40:(p + 1) * 40], b[p * 40:(p + 1) * 40]
        count = int(ma[BoxContact.COUNT])
        assert count == int(mb[BoxContact.COUNT]), p
        if count:
            used = BoxContact.POINTS + count * BoxContact.POINT_WORDS
            assert np.array_equal(ma[:5].view(np.int32), mb[:5].view(np.int32)), p
            assert np.array_equal(ma[8:used].view(np.int32), mb[8:used].view(np.int32)), p


def test_random_pair_manifolds_with_margin_and_spheres_match():
    bodies = random_pairs(200, seed=7)
    bodies = [(c, (h[0], h[0], h[0]), q, 1.0) if k % 3 == 0 else (c, h, q, s)
              for k, (c, h, q, s) in enumerate(bodies)]
    a = collide("oracle", 0.01, bodies)
    b = collide("gpu", 0.01, bodies)
    assert np.array_equal(a.view(np.int32)[3::40], b.view(np.int32)[3::40])
    for p in range(200):
        count = int(a[p * 40 + 3])
        if count:
            used = 8 + count * 8
            assert np.array_equal(a[p * 40:p * 40 + 5].view(np.int32), b[p * 40:p * 40 + 5].view(np.int32))
            assert np.array_equal(a[p * 40 + 8:p * 40 + used].view(np.int32), b[p * 40 + 8:p * 40 + used].view(np.int32))


def mixed_pile(n=120, seed=31, spheres=True):
    """Boxes and spheres dropped into a 1 m box with friction and rolling resistance."""
    rnd = pkg.JavaRandom(seed)
    s = BoxScene(n)
    s.sy = 2
    s.muStatic = s.muKinetic = 0.5
    s.muRolling = 0.05
    for k in range(n):
        if spheres and k % 2 == 1:
            s.sphere(k, 0.03 + 0.03 * rnd.nextDouble())
        else:
            for a in range(3):
                s.half[a][k] = np.float32(0.03 + 0.04 * rnd.nextDouble())
            ax = [rnd.nextGaussian() for _ in range(4)]
            length = math.sqrt(sum(v * v for v in ax))
            for a in range(4):
                s.q[a][k] = np.float32(ax[a] / length)
        s.im[k] = np.float32(1 / (0.5 + rnd.nextDouble()))
        s.x[k] = np.float32(0.1 + 0.8 * rnd.nextDouble())
        s.y[k] = np.float32(0.1 + 1.6 * rnd.nextDouble())
        s.z[k] = np.float32(0.1 + 0.8 * rnd.nextDouble())
    return s


def run_both(make, steps):
    out = {}
    for backend in ("oracle", "gpu"):
        s = make()
        s.start(backend)
        reps = []
        for _ in range(steps):
            r = s.sim.stepAndWait()
            reps.append((r.touching, r.points, r.rounds, r.unplaced, r.unlisted))
        out[backend] = ({k: s.floats(k).copy() for k in STATE + WALL_IMPULSES}, reps, s.ints("contacts").copy(),
                        s.ints("rounds").copy())
        s.close()
    return out["oracle"], out["gpu"]


@pytest.mark.parametrize("spheres", [False, True])
def test_mixed_pile_matches_oracle(spheres):
    (sa, ra, ca, qa), (sb, rb, cb, qb) = run_both(lambda: mixed_pile(spheres=spheres), 40)
    assert ra == rb
    assert any(r[0] > 0 for r in ra), "the pile should touch"
    assert np.array_equal(ca, cb)
    assert np.array_equal(qa, qb), "the rounds (colouring) differ"
    for k in WALL_IMPULSES:
        assert np.any(sa[k] != 0) or (k == "roll" and not spheres), k  # rolling resistance is for balls
        assert np.array_equal(sa[k].view(np.int32), sb[k].view(np.int32)), k
    for k in STATE:
        scale = np.maximum(np.abs(sa[k]), 1e-3)
        assert float(np.max(np.abs(sa[k] - sb[k]) / scale)) <= 1e-4, k
        assert np.array_equal(sa[k].view(np.int32), sb[k].view(np.int32)), k


def test_symmetric_sweeps_match_oracle():
    def make():
        s = mixed_pile(60, seed=3)
        s.symmetric = True
        s.passes = 3
        return s
    (sa, ra, _, qa), (sb, rb, _, qb) = run_both(make, 20)
    assert np.array_equal(qa, qb)
    assert ra == rb
    for k in STATE:
        assert np.array_equal(sa[k].view(np.int32), sb[k].view(np.int32)), k


def test_big_pile_colouring_matches_oracle():
    """A denser pile of 600 bodies over 10 steps: the colouring (rounds buffer) and the counts, step by step."""
    def make():
        s = mixed_pile(600, seed=11)
        s.sx = s.sz = 2.0
        s.sy = 3.0
        return s
    (sa, ra, ca, qa), (sb, rb, cb, qb) = run_both(make, 10)
    assert ra == rb
    assert np.array_equal(qa, qb)
    for k in STATE:
        assert np.array_equal(sa[k].view(np.int32), sb[k].view(np.int32)), k


def test_wall_impulses_written_through_the_abi_round_trip():
    """Wall impulses read and written back body by body (the reference's layout) leave the stepping unchanged."""
    out = {}
    for backend in ("oracle", "gpu"):
        s = mixed_pile(80, seed=5)
        s.start(backend)
        for _ in range(6):
            s.sim.stepAndWait()
        for name in WALL_IMPULSES:
            s.sim.write(name, s.floats(name).copy())
        for _ in range(6):
            s.sim.stepAndWait()
        out[backend] = {k: s.floats(k).copy() for k in STATE + WALL_IMPULSES}
        s.close()
    for k in STATE + WALL_IMPULSES:
        assert np.array_equal(out["oracle"][k].view(np.int32), out["gpu"][k].view(np.int32)), k
