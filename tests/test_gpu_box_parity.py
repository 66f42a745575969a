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


def mixed_pile(n=120, seed=31, spheres=True, rolling=0.05):
    """Boxes and spheres dropped into a 1 m box with friction and (unless `rolling` is 0) rolling resistance."""
    rnd = pkg.JavaRandom(seed)
    s = BoxScene(n)
    s.sy = 2
    s.muStatic = s.muKinetic = 0.5
    s.muRolling = rolling
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
        out[backend] = ({k: s.floats(k).copy() for k in STATE + WALL_IMPULSES + ("shown",)}, reps,
                        s.ints("contacts").copy(),
                        s.ints("rounds").copy())
        s.close()
    return out["oracle"], out["gpu"]


@pytest.mark.parametrize("spheres,rolling", [(False, 0.05), (True, 0.05), (True, 0.0)])
def test_mixed_pile_matches_oracle(spheres, rolling):
    (sa, ra, ca, qa), (sb, rb, cb, qb) = run_both(lambda: mixed_pile(spheres=spheres, rolling=rolling), 40)
    assert ra == rb
    assert any(r[0] > 0 for r in ra), "the pile should touch"
    assert np.array_equal(ca, cb)
    assert np.array_equal(qa, qb), "the rounds (colouring) differ"
    for k in WALL_IMPULSES:
        assert np.any(sa[k] != 0) or (k == "roll" and not (spheres and rolling)), k  # rolling resistance is for balls
        assert np.array_equal(sa[k].view(np.int32), sb[k].view(np.int32)), k
    assert np.any(sa["shown"] != 0)
    assert np.array_equal(sa["shown"].view(np.int32), sb["shown"].view(np.int32)), "shown"  # Boxes.show's record
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


def corner_on_walls(seed=2):
    """Boxes turned so a corner points straight at a wall (a diagonal along the wall's axis), each at a distance
    from its wall of half-diagonal + margin + eps, eps from -1e-4 to +3e-3 m with rounding-sized steps near 0, and
    moving towards the wall at 4 m/s, no gravity: corners start just inside, on and just outside the margin, and at
    that speed a corner inside the margin is pushed back in the first substep (the contact's bias, gap / h, is
    3 m/s at the margin), one outside only in a later one."""
    rng = np.random.default_rng(seed)
    eps = (-1e-4, -1e-5, -1e-6, -1e-7, 0.0, 1e-7, 1e-6, 1e-5, 1e-4, 1e-3, 3e-3)
    spots = [(0.25, 0.25), (0.5, 0.5), (0.75, 0.75), (0.25, 0.75), (0.75, 0.25), (0.5, 0.25), (0.25, 0.5),
             (0.75, 0.5), (0.5, 0.75), (0.35, 0.65), (0.65, 0.35)]
    n = 6 * len(eps)
    s = BoxScene(n)
    s.gx = s.gy = s.gz = 0.0
    s.muStatic = s.muKinetic = 0.4
    k = 0
    for wall in range(6):
        axis, high = wall // 2, wall % 2 == 1
        others = [a for a in range(3) if a != axis]
        for e, (p0, p1) in zip(eps, spots):
            h = rng.uniform(0.02, 0.06, 3)
            s.half[:, k] = h.astype(np.float32)
            u = h / np.linalg.norm(h)  # the corner (+hx, +hy, +hz) turned onto the wall's outward direction
            v = np.zeros(3)
            v[axis] = 1.0 if high else -1.0
            q = np.append(np.cross(u, v), 1.0 + float(np.dot(u, v)))
            q /= np.linalg.norm(q)
            s.q[:, k] = q.astype(np.float32)
            reach = float(np.float32(np.linalg.norm(s.half[:, k].astype(np.float64)))) + s.margin + e
            at = [0.0, 0.0, 0.0]
            at[axis] = 1.0 - reach if high else reach
            at[others[0]], at[others[1]] = p0, p1
            s.x[k], s.y[k], s.z[k] = (np.float32(c) for c in at)
            speed = [0.0, 0.0, 0.0]
            speed[axis] = 4.0 if high else -4.0
            s.u[k], s.v[k], s.w[k] = (np.float32(c) for c in speed)
            s.im[k] = np.float32(1.0)
            k += 1
    return s


def test_corners_at_the_wall_margin_match_oracle():
    """Where a body's corners stand about the margin of a wall it holds no impulse at, the wall sweep may pass the
    wall over from a bound on the corners' reach; bodies placed on either side of that bound step as the oracle's."""
    (sa, ra, _, _), (sb, rb, _, _) = run_both(corner_on_walls, 2)
    assert ra == rb
    speed = np.sqrt(sa["u"].astype(np.float64) ** 2 + sa["v"] ** 2 + sa["w"] ** 2)
    assert np.all(speed < 3.9), "every body should have reached its wall"
    for k in WALL_IMPULSES + STATE:
        assert np.array_equal(sa[k].view(np.int32), sb[k].view(np.int32)), k


def test_shapes_rewritten_between_steps_match_oracle():
    """The HIP wall sweep keeps each body's wall impulses where its thread takes the body (a workgroup's boxes before
    its balls), so writing the shape buffer moves them. A pile stepped with wall impulses held, then with every
    seventh body's shape flipped, steps as the oracle's, and its wall impulses read back the same."""
    out = {}
    for backend in ("oracle", "gpu"):
        s = mixed_pile(200, seed=13)
        s.start(backend)
        for _ in range(30):  # settled enough that about a hundred corners and balls hold wall impulses
            s.sim.stepAndWait()
        shape = s.floats("shape").copy()
        shape[::7] = np.float32(1.0) - shape[::7]
        s.sim.write("shape", shape)
        after = {k: s.floats(k).copy() for k in WALL_IMPULSES}
        for _ in range(6):
            s.sim.stepAndWait()
        out[backend] = (after, {k: s.floats(k).copy() for k in STATE + WALL_IMPULSES})
        s.close()
    (wa, sa), (wb, sb) = out["oracle"], out["gpu"]
    assert any(np.any(wa[k] != 0) for k in WALL_IMPULSES), "impulses should be held when the shapes change"
    for k in WALL_IMPULSES:
        assert np.array_equal(wa[k].view(np.int32), wb[k].view(np.int32)), ("after the write", k)
    for k in STATE + WALL_IMPULSES:
        assert np.array_equal(sa[k].view(np.int32), sb[k].view(np.int32)), k


def test_pair_buffers_written_from_outside_match_oracle():
    """The HIP step leaves a slot's manifold count and impulses alone where it knows they already read as an empty
    slot's; a buffer written through the ABI may hold anything, so every slot is written again. The kept manifold
    and impulses, filled with words other than zero everywhere, become the step's own buffers after the swap: the
    step must leave them, empty slots included, as the oracle's does."""
    out = {}
    for backend in ("oracle", "gpu"):
        s = mixed_pile(120, seed=17)
        s.start(backend)
        for _ in range(8):
            s.sim.stepAndWait()
        for name in ("manifoldWas", "impulsesWas"):
            junk = np.full(s.floats(name).shape, np.float32(0.25))
            s.sim.write(name, junk)
        r = s.sim.stepAndWait()
        out[backend] = ((r.touching, r.points, r.rounds),
                        {k: s.floats(k).copy() for k in ("manifold", "impulses") + STATE})
        s.close()
    (ra, a), (rb, b) = out["oracle"], out["gpu"]
    assert ra == rb
    for k in ("impulses",) + STATE:
        assert np.array_equal(a[k].view(np.int32), b[k].view(np.int32)), k
    # The manifold's words a step writes: every slot's count, and a touching pair's normal, axis and the 7 words of
    # each counted point (of its 8). The rest hold what the buffer held before; the HIP library swaps the kept and
    # current buffers where the oracle copies, so those words differ, and nothing reads them.
    ma, mb = a["manifold"].reshape(-1, 40).view(np.int32), b["manifold"].reshape(-1, 40).view(np.int32)
    assert np.array_equal(ma[:, 3], mb[:, 3]), "counts"
    count = a["manifold"].reshape(-1, 40)[:, 3].astype(np.int64)
    assert count.max() > 0, "the pile should touch"
    for p in np.nonzero(count > 0)[0]:
        assert np.array_equal(ma[p, :5], mb[p, :5]), p
        for k in range(int(count[p])):
            assert np.array_equal(ma[p, 8 + 8 * k:15 + 8 * k], mb[p, 8 + 8 * k:15 + 8 * k]), (p, k)
