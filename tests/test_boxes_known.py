"""The box world's known answers on each backend: the reference's own box test cases, restated.

ctest/box/... = vexelray-sim-rigid-core/src/test/java/dev/vexelray/sim/rigid/box/... The narrow-phase cases are
worked out by hand in the reference (BoxContactTest); the random-pair property test and the 150-box broad-phase
scene carry the reference's own published counts (docs/boxes.md:233-234, 367-368).
"""

import math

import numpy as np
import pytest

import scene as sc
from boxscene import BoxContact, BoxDiagnostics, BoxScene, Boxes, collide

pkg = sc.pkg
BACKENDS = ["oracle", pytest.param("gpu", marks=pytest.mark.gpu)]
SLAB = (0.3, 0.1, 0.3)
CUBE = (0.1, 0.1, 0.1)
IDENTITY = (0.0, 0.0, 0.0, 1.0)
R2 = math.sqrt(2)
CLOSE = 2e-6


def axis_angle(x, y, z, angle):
    length = math.sqrt(x * x + y * y + z * z)
    s = math.sin(angle / 2) / length
    return (x * s, y * s, z * s, math.cos(angle / 2))


def box(x, y, z, half, q):
    return ((x, y, z), half, q, 0.0)


def ball(x, y, z, r):
    return ((x, y, z), (r, r, r), IDENTITY, 1.0)


def world(body, m, at):
    own = np.array(m[at:at + 3], dtype=np.float64)
    return np.array(body[0]) + BoxDiagnostics.rotate(body[2], own)


class Contact:
    """BoxContactTest.Contact (:360-429): a manifold read back and the checks a case makes."""

    def __init__(self, m, first, second):
        self.m, self.first, self.second = m, first, second

    def face(self, axis, nx, ny, nz):
        assert self.m[BoxContact.AXIS] == axis
        self.normal(nx, ny, nz)

    def normal(self, nx, ny, nz):
        for a, v in enumerate((nx, ny, nz)):
            assert abs(v - self.m[a]) <= CLOSE, (a, self.m[:3])

    def points(self, depth, on_second):
        n = self.m[:3].astype(np.float64)
        on_first = [np.array(p) + depth * n for p in on_second]
        self.points_on(depth, on_first, on_second)

    def points_on(self, depth, on_first, on_second):
        count = int(self.m[BoxContact.COUNT])
        assert count == len(on_first), f"{count} points"
        used = [False] * count
        for e in range(len(on_first)):
            matched = False
            for k in range(count):
                at = BoxContact.POINTS + k * BoxContact.POINT_WORDS
                a = world(self.first, self.m, at + BoxContact.ANCHOR_FIRST)
                b = world(self.second, self.m, at + BoxContact.ANCHOR_SECOND)
                if not used[k] and np.all(np.abs(a - on_first[e]) < 1e-5) and np.all(np.abs(b - on_second[e]) < 1e-5):
                    used[k] = matched = True
                    assert abs(depth - self.m[at + BoxContact.DEPTH]) <= CLOSE
                    break
            assert matched, f"no point at {on_first[e]} / {on_second[e]}"


def corners(hx, hz, y):
    return [(hx, y, hz), (hx, y, -hz), (-hx, y, hz), (-hx, y, -hz)]


def pair(backend, first, second, margin=0.0):
    return collide(backend, margin, [first, second])


# ---- BoxContactTest.java ----

@pytest.mark.parametrize("backend", BACKENDS)
def test_boxes_apart_have_no_contact(backend):
    assert pair(backend, box(0, 0, 0, CUBE, IDENTITY), box(0.25, 0, 0, CUBE, IDENTITY))[BoxContact.COUNT] == 0


@pytest.mark.parametrize("backend", BACKENDS)
def test_cube_on_slab_four_corners(backend):
    a, b = box(0, 0, 0, SLAB, IDENTITY), box(0, 0.19, 0, CUBE, IDENTITY)
    c = Contact(pair(backend, a, b), a, b)
    c.face(1, 0, 1, 0)
    c.points(0.01, corners(0.1, 0.1, 0.09))


@pytest.mark.parametrize("backend", BACKENDS)
def test_cube_on_same_size_cube(backend):
    a, b = box(0, 0, 0, CUBE, IDENTITY), box(0, 0.19, 0, CUBE, IDENTITY)
    c = Contact(pair(backend, a, b), a, b)
    c.face(1, 0, 1, 0)
    c.points(0.01, corners(0.1, 0.1, 0.09))


@pytest.mark.parametrize("backend", BACKENDS)
def test_turned_cube_touches_at_corners(backend):
    a, b = box(0, 0, 0, SLAB, IDENTITY), box(0, 0.19, 0, CUBE, axis_angle(0, 1, 0, math.pi / 4))
    c = Contact(pair(backend, a, b), a, b)
    c.face(1, 0, 1, 0)
    d = 0.1 * R2
    c.points(0.01, [(d, 0.09, 0), (-d, 0.09, 0), (0, 0.09, d), (0, 0.09, -d)])


@pytest.mark.parametrize("backend", BACKENDS)
def test_cube_over_an_edge_is_clipped(backend):
    a, b = box(0, 0, 0, SLAB, IDENTITY), box(0.25, 0.19, 0, CUBE, IDENTITY)
    c = Contact(pair(backend, a, b), a, b)
    c.face(1, 0, 1, 0)
    c.points(0.01, [(0.15, 0.09, 0.1), (0.15, 0.09, -0.1), (0.3, 0.09, 0.1), (0.3, 0.09, -0.1)])


@pytest.mark.parametrize("backend", BACKENDS)
def test_turned_cube_over_edge_three_points(backend):
    a, b = box(0, 0, 0, SLAB, IDENTITY), box(0.3, 0.19, 0, CUBE, axis_angle(0, 1, 0, math.pi / 4))
    c = Contact(pair(backend, a, b), a, b)
    c.face(1, 0, 1, 0)
    d = 0.1 * R2
    c.points(0.01, [(0.3 - d, 0.09, 0), (0.3, 0.09, d), (0.3, 0.09, -d)])


@pytest.mark.parametrize("backend", BACKENDS)
def test_pair_the_other_way_round_turns_the_normal(backend):
    slab, cube = box(0, 0, 0, SLAB, IDENTITY), box(0, 0.19, 0, CUBE, IDENTITY)
    c = Contact(pair(backend, cube, slab), cube, slab)
    c.face(1, 0, -1, 0)
    c.points_on(0.01, corners(0.1, 0.1, 0.09), corners(0.1, 0.1, 0.1))


@pytest.mark.parametrize("backend", BACKENDS)
def test_crossed_edges_touch_at_one_point(backend):
    top = 0.1 * R2
    below = box(0, 0, 0, CUBE, axis_angle(0, 0, 1, math.pi / 4))
    above = box(0, 2 * top - 0.01, 0, CUBE, axis_angle(1, 0, 0, math.pi / 4))
    m = pair(backend, below, above)
    assert m[BoxContact.AXIS] == 6 + 3 * 2
    c = Contact(m, below, above)
    c.normal(0, 1, 0)
    c.points_on(0.01, [(0, top, 0)], [(0, top - 0.01, 0)])


@pytest.mark.parametrize("backend", BACKENDS)
def test_corner_into_face(backend):
    tilt = math.acos(1 / math.sqrt(3))
    reach = 0.1 * math.sqrt(3)
    slab, cube = box(0, 0, 0, SLAB, IDENTITY), box(0, 0.1 + reach - 0.01, 0, CUBE, axis_angle(-1, 0, 1, tilt))
    c = Contact(pair(backend, slab, cube), slab, cube)
    c.face(1, 0, 1, 0)
    c.points_on(0.01, [(0, 0.1, 0)], [(0, 0.09, 0)])


@pytest.mark.parametrize("backend", BACKENDS)
def test_an_axis_of_the_second_parts_them(backend):
    m = pair(backend, box(0, 0, 0, CUBE, IDENTITY), box(0.24, 0.24, 0, CUBE, axis_angle(0, 0, 1, math.pi / 4)))
    assert m[BoxContact.COUNT] == 0


@pytest.mark.parametrize("backend", BACKENDS)
def test_margin_finds_boxes_about_to_touch(backend):
    slab, cube = box(0, 0, 0, SLAB, IDENTITY), box(0, 0.205, 0, CUBE, IDENTITY)
    assert pair(backend, slab, cube)[BoxContact.COUNT] == 0
    c = Contact(pair(backend, slab, cube, 0.01), slab, cube)
    c.face(1, 0, 1, 0)
    c.points(-0.005, corners(0.1, 0.1, 0.105))


def random_pairs(pairs=500, seed=19):
    """BoxContactTest.everyPointLiesOnBothBoxesAtItsDepth's scene (:188-199)."""
    rnd = pkg.JavaRandom(seed)
    bodies = []
    for k in range(2 * pairs):
        half = (0.03 + 0.12 * rnd.nextDouble(), 0.03 + 0.12 * rnd.nextDouble(), 0.03 + 0.12 * rnd.nextDouble())
        spread = 0 if k % 2 == 0 else 0.3
        x = spread * (rnd.nextDouble() - 0.5)
        y = spread * (rnd.nextDouble() - 0.5)
        z = spread * (rnd.nextDouble() - 0.5)
        q = axis_angle(rnd.nextDouble() - 0.5, rnd.nextDouble() - 0.5, rnd.nextDouble() - 0.5, 4 * rnd.nextDouble())
        bodies.append(box(x, y, z, half, q))
    return bodies


@pytest.mark.parametrize("backend", BACKENDS)
def test_every_point_lies_on_both_boxes_at_its_depth(backend):
    """BoxContactTest.everyPointLiesOnBothBoxesAtItsDepth (:188-237), and the counts the reference published for this
    scene: 468 of 500 pairs touching, 1,196 points (docs/boxes.md:233-234)."""
    bodies = random_pairs()
    allm = collide(backend, 0, bodies)
    touching = points = 0
    for p in range(500):
        m = allm[p * 40:(p + 1) * 40]
        first, second = bodies[2 * p], bodies[2 * p + 1]
        count = int(m[BoxContact.COUNT])
        assert count <= 4
        inner = min(first[1]) + min(second[1])
        apart = math.dist(first[0], second[0])
        assert apart >= inner or count > 0
        touching += count > 0
        points += count
        n = m[:3].astype(np.float64)
        seen = []
        for k in range(count):
            at = BoxContact.POINTS + k * BoxContact.POINT_WORDS
            for off, half in ((BoxContact.ANCHOR_FIRST, first[1]), (BoxContact.ANCHOR_SECOND, second[1])):
                inside = [half[a] - abs(float(m[at + off + a])) for a in range(3)]
                assert min(inside) > -1e-5 and min(inside) < 1e-5
            a = world(first, m, at + BoxContact.ANCHOR_FIRST)
            b = world(second, m, at + BoxContact.ANCHOR_SECOND)
            assert abs(float(np.dot(a - b, n)) - m[at + BoxContact.DEPTH]) <= 1e-5
            for q in seen:
                assert np.linalg.norm(b - q) > 1e-6
            seen.append(b)
    assert touching > 250
    assert (touching, points) == (468, 1196)


@pytest.mark.parametrize("backend", BACKENDS)
def test_ball_cases(backend):
    """BoxContactTest's sphere cases (:243-316)."""
    slab = box(0, 0, 0, SLAB, IDENTITY)
    b = ball(0, 0.14, 0, 0.05)
    c = Contact(pair(backend, slab, b), slab, b)
    c.face(BoxContact.SPHERE_BOX, 0, 1, 0)
    c.points_on(0.01, [(0, 0.1, 0)], [(0, 0.09, 0)])
    c = Contact(pair(backend, b, slab), b, slab)
    c.face(BoxContact.SPHERE_BOX, 0, -1, 0)
    c.points_on(0.01, [(0, 0.09, 0)], [(0, 0.1, 0)])
    e = ball(0.33, 0.13, 0, 0.05)
    c = Contact(pair(backend, slab, e), slab, e)
    d = 1 / R2
    c.face(BoxContact.SPHERE_BOX, d, d, 0)
    c.points_on(0.05 - 0.03 * R2, [(0.3, 0.1, 0)], [(0.33 - 0.05 * d, 0.13 - 0.05 * d, 0)])
    inside = ball(0.1, 0.09, 0.05, 0.05)
    c = Contact(pair(backend, slab, inside), slab, inside)
    c.face(BoxContact.SPHERE_BOX, 0, 1, 0)
    c.points_on(0.06, [(0.1, 0.1, 0.05)], [(0.1, 0.04, 0.05)])
    f, s = ball(0, 0, 0, 0.05), ball(0.08, 0, 0, 0.04)
    c = Contact(pair(backend, f, s), f, s)
    c.face(BoxContact.SPHERE_SPHERE, 1, 0, 0)
    c.points_on(0.01, [(0.05, 0, 0)], [(0.04, 0, 0)])
    assert pair(backend, slab, ball(0, 0.2, 0, 0.05))[BoxContact.COUNT] == 0


# ---- BoxBroadTest.java ----

def broad_scene(seed=23, n=150):
    rnd = pkg.JavaRandom(seed)
    bodies = []
    widest = 0.0
    for s in range(n):
        half = (0.02 + 0.06 * rnd.nextDouble(), 0.02 + 0.06 * rnd.nextDouble(), 0.02 + 0.06 * rnd.nextDouble())
        widest = max(widest, math.sqrt(half[0] ** 2 + half[1] ** 2 + half[2] ** 2))
        ax = (rnd.nextDouble() - 0.5, rnd.nextDouble() - 0.5, rnd.nextDouble() - 0.5)
        length = math.sqrt(ax[0] ** 2 + ax[1] ** 2 + ax[2] ** 2)
        angle = 4 * rnd.nextDouble()
        sn = math.sin(angle / 2) / length
        q = (ax[0] * sn, ax[1] * sn, ax[2] * sn, math.cos(angle / 2))
        centre = tuple(0.1 + 0.8 * rnd.nextDouble() for _ in range(3))
        bodies.append((centre, half, q, 0.0))
    return bodies, widest


@pytest.mark.parametrize("backend", BACKENDS)
def test_grid_finds_what_every_pair_finds(backend):
    """BoxBroadTest.theGridFindsWhatEveryPairFinds (:37-101), and the 218 touching pairs the reference published for
    this scene (docs/boxes.md:367-368)."""
    bodies, widest = broad_scene()
    n = len(bodies)
    margin = 0.005
    boxes = __import__(pkg.__name__ + ".boxes", fromlist=["boxes"])
    step = boxes.BoxStep(n, 1, 1, False, boxes.BoxBroad.grid(1, 1, 1, widest, margin))
    sim = boxes.BoxSimulation(step, abi=sc.abi_for(backend))
    f32 = lambda f: np.array([f(b) for b in bodies], dtype=np.float64).astype(np.float32)
    for k, name in enumerate(("x", "y", "z")):
        sim.write(name, f32(lambda b, k=k: b[0][k]))
    for k, name in enumerate(("qx", "qy", "qz", "qw")):
        sim.write(name, f32(lambda b, k=k: b[2][k]))
    for k, name in enumerate(("hx", "hy", "hz")):
        sim.write(name, f32(lambda b, k=k: b[1][k]))
    sim.write("im", np.zeros(n, np.float32))  # fixed: the step only lists and collides
    sim.params(Boxes.params(1e-3, 0, 0, 0, 1, 1, 1, 0, 0, margin))
    rep = sim.stepAndWait()
    contacts = sim.ints("contacts")
    slots = step.slots
    grid = {}
    for p in range(slots):
        if contacts[p] >= 0:
            grid[(p // 32, int(contacts[p]))] = int(contacts[slots + p])
    unlisted = sum(max(0, int(contacts[2 * slots + s]) - 32) for s in range(n))
    pairs = np.array([(a, b) for a in range(n) for b in range(a + 1, n)], dtype=np.int32)
    allm = sim.collide(pairs.reshape(-1))
    sim.close()
    every = {}
    for p, (a, b) in enumerate(pairs):
        count = int(allm[p * 40 + BoxContact.COUNT])
        if count > 0:
            every[(int(a), int(b))] = count
    assert len(every) > 100
    assert unlisted == 0
    assert every == grid
    assert len(grid) == 218
    assert rep.touching == 218 and rep.points == sum(grid.values())


# ---- dynamics: BoxFreeTest, BoxWallTest, BoxStackTest, BoxSphereTest ----

HX, HY, HZ = 0.05, 0.1, 0.2


@pytest.mark.parametrize("backend", BACKENDS)
def test_free_fall_is_symplectic_euler(backend):
    """BoxFreeTest.aFreeFallIsSymplecticEulers (:31-49)."""
    with BoxScene(1).uniform(HX, HY, HZ, 1) as s:
        s.sy = 20
        s.x[0] = s.z[0] = 0.5
        s.y[0] = 8
        s.start(backend).advanceSeconds(0.5).read()
        h = s.dt / s.substeps
        y, vy = 8.0, 0.0
        for _ in range(30 * s.substeps):
            vy += h * s.gy
            y += h * vy
        assert abs(vy - s.v[0]) <= 1e-5
        assert abs(y - s.y[0]) <= 1e-5
        assert s.q[3][0] == 1


@pytest.mark.parametrize("backend", BACKENDS)
def test_spin_about_stable_axis_is_kept(backend):
    """BoxFreeTest.aSpinAboutAStableAxisIsKept (:55-75)."""
    for axis in (0, 2):
        with BoxScene(1).uniform(HX, HY, HZ, 1) as s:
            s.gy = 0
            s.x[0] = s.y[0] = s.z[0] = 0.5
            (s.ax, s.ay, s.az)[axis][0] = 5
            s.start(backend).advanceSeconds(10).read()
            own = BoxDiagnostics.rotate(BoxDiagnostics.conjugate(s.q[:, 0].astype(np.float64)),
                                        [s.ax[0], s.ay[0], s.az[0]])
            for a in range(3):
                assert abs((5 if a == axis else 0) - own[a]) <= 1e-4
            angle = 2 * math.atan2(s.q[axis][0], s.q[3][0])
            expected = math.remainder(50, 2 * math.pi)
            assert abs(math.remainder(angle - expected, 2 * math.pi)) <= 1e-3


@pytest.mark.parametrize("backend", BACKENDS)
def test_tumble_keeps_angular_momentum_and_energy(backend):
    """BoxFreeTest.aTumbleKeepsItsAngularMomentumAndEnergy (:109-126): L to the bit, energy to 1e-6."""
    with BoxScene(1).uniform(HX, HY, HZ, 2) as s:
        s.gy = 0
        s.sx = s.sy = s.sz = 100
        s.x[0] = s.y[0] = s.z[0] = 50
        s.u[0] = 1
        s.ax[0], s.ay[0], s.az[0] = 1, 2, 3
        before = s.start(backend).diagnostics()
        s.advanceSeconds(10)
        after = s.diagnostics()
        drift = np.linalg.norm(after.angular - before.angular) / np.linalg.norm(before.angular)
        energy = abs(after.kinetic - before.kinetic) / before.kinetic
        assert drift == 0
        assert energy < 1e-6
        own = BoxDiagnostics.rotate(BoxDiagnostics.conjugate(s.q[:, 0].astype(np.float64)), [s.ax[0], s.ay[0], s.az[0]])
        assert abs(own[1]) > 0.1


@pytest.mark.parametrize("backend", BACKENDS)
def test_box_dropped_flat_rests_flat(backend):
    """BoxWallTest.aBoxDroppedFlatRestsFlat (:22-40)."""
    with BoxScene(1).uniform(0.1, 0.05, 0.1, 1) as s:
        s.x[0] = s.z[0] = 0.5
        s.y[0] = 0.06
        s.muStatic = s.muKinetic = 0.5
        s.start(backend).advanceSeconds(2)
        st = s.diagnostics()
        up = BoxDiagnostics.rotate(s.q[:, 0].astype(np.float64), [0, 1, 0])
        tilt = math.acos(min(1.0, abs(up[1])))
        assert tilt < 1e-4
        assert abs(0.05 - s.y[0]) <= 1e-4
        assert abs(0.5 - s.x[0]) <= 1e-4 and abs(0.5 - s.z[0]) <= 1e-4
        assert st.kinetic < 1e-6


@pytest.mark.parametrize("backend", BACKENDS)
def test_head_on_collision_keeps_momentum(backend):
    """BoxStackTest.aHeadOnCollisionKeepsMomentum (:22-42)."""
    with BoxScene(2).uniform(0.05, 0.05, 0.05, 1) as s:
        s.gy = 0
        s.sx = s.sy = s.sz = 2
        s.x[0], s.x[1] = 0.8, 1.2
        s.y[:] = 1
        s.z[:] = 1
        s.u[0], s.u[1] = 1, -1
        s.start(backend).advanceSeconds(0.5)
        st = s.diagnostics()
        assert abs(st.momentum[0]) <= 1e-6
        assert abs(s.u[0]) <= 1e-3 and abs(s.u[1]) <= 1e-3
        assert abs(0.1 - (s.x[1] - s.x[0])) <= 1e-3


@pytest.mark.parametrize("backend", BACKENDS)
def test_column_of_ten_rests(backend):
    """BoxStackTest.aColumnOfTenRests (:44-52, column :234-261); the reference's column takes 2 rounds
    (docs/boxes.md:265-266)."""
    n = 10
    with BoxScene(n).uniform(0.05, 0.05, 0.05, 1) as s:
        s.sx = s.sz = 1
        s.sy = 0.11 * n + 0.5
        for k in range(n):
            s.x[k] = s.z[k] = 0.5
            s.y[k] = np.float32(0.05 + 0.101 * k)
        s.muStatic = s.muKinetic = 0.5
        s.start(backend).advanceSeconds(3)
        st = s.diagnostics()
        drift = math.hypot(s.x[n - 1] - 0.5, s.z[n - 1] - 0.5)
        assert drift < 1e-3
        assert st.kinetic < 1e-6
        assert s.unplaced == 0
        assert s.mostRounds == 2


@pytest.mark.parametrize("backend", BACKENDS)
def test_ball_rests_on_a_box(backend):
    """BoxSphereTest.aBallRestsOnABox (:21-40)."""
    with BoxScene(2).uniform(0.05, 0.05, 0.05, 1) as s:
        s.sphere(1, 0.05)
        s.im[1] = 2
        s.muStatic = s.muKinetic = 0.5
        s.x[:] = 0.5
        s.z[:] = 0.5
        s.y[0] = 0.05
        s.y[1] = 0.151
        s.start(backend).advanceSeconds(2)
        st = s.diagnostics()
        assert abs(0.15 - s.y[1]) <= 5e-4
        assert math.hypot(s.x[1] - 0.5, s.z[1] - 0.5) < 1e-4
        assert st.kinetic < 1e-6


def crowded_body(n=80):
    """n small balls resting on one big ball, the big ball last: each small ball lists the big one as its partner,
    so the big ball is in n touching pairs, one listed by each. First fit in slot order gives the first 32 of them
    rounds 0..31 and leaves the rest unplaced (BoxStepper.java:71-90)."""
    s = BoxScene(n + 1)
    s.gy = 0.0
    s.sx = s.sy = s.sz = 2.0
    big = n
    s.sphere(big, 0.5)
    s.x[big] = s.y[big] = s.z[big] = 1.0
    golden = math.pi * (3 - math.sqrt(5))
    for k in range(n):
        s.sphere(k, 0.06)
        yk = 1 - 2 * (k + 0.5) / n
        rk = math.sqrt(1 - yk * yk)
        dist = 0.5 + 0.06 - 0.002
        s.x[k] = np.float32(1.0 + dist * rk * math.cos(golden * k))
        s.y[k] = np.float32(1.0 + dist * yk)
        s.z[k] = np.float32(1.0 + dist * rk * math.sin(golden * k))
    return s


@pytest.mark.parametrize("backend", BACKENDS)
def test_crowded_body_fills_every_round_and_leaves_the_rest_unplaced(backend):
    with crowded_body() as s:
        s.substeps = 1
        s.start(backend).advance(1)
        r = s.last
        assert r.touching == 80
        assert r.rounds == 32
        assert r.unplaced == 80 - 32
        rounds = s.ints("rounds")
        width = len(rounds) // 32
        firsts = [int(rounds[k * width]) for k in range(32)]
        # Round k holds the k-th small ball's pair with the big one (slot k * 32), and nothing else.
        assert firsts == [k * 32 for k in range(32)]
        assert all(int(rounds[k * width + 1]) == -1 for k in range(32))


def test_the_picture_is_given_each_step_and_the_one_before():
    """BoxShowTest.thePictureIsGivenEachStepAndTheOneBefore (core/src/test/.../box/BoxShowTest.java:18-52), on the
    oracle: two boxes, one tumbling in free fall and one fixed; after three steps the "after" words hold the state
    the step ended on, the "before" words what the step before left there, and the half-extents are each box's own
    (Boxes.java:781-791: centre before 0, after 4, half 8, turn before 12, after 16). The HIP library's picture is
    the oracle's bit for bit (test_gpu_box_parity.py)."""
    centre_before, centre_after, half, turned_before, turned_after = 0, 4, 8, 12, 16
    with BoxScene(2).uniform(0.05, 0.1, 0.15, 1) as s:
        s.x[0] = s.y[0] = s.z[0] = 0.5
        s.ax[0], s.ay[0], s.az[0] = 1, 2, 3
        s.x[1], s.y[1], s.z[1] = 0.2, 0.1, 0.8
        s.im[1] = 0
        s.half[2][1] = 0.05
        s.start("oracle").advance(2)
        earlier = s.floats("shown").copy()
        s.advance(1).read()
        shown = s.floats("shown")
        at = (s.x, s.y, s.z)
        for b in range(2):
            base = b * Boxes.SHOWN_STRIDE
            for a in range(3):
                assert shown[base + centre_after + a] == at[a][b], (b, a)
                assert shown[base + centre_before + a] == earlier[base + centre_after + a], (b, a)
                assert shown[base + half + a] == s.half[a][b], (b, a)
            for k in range(4):
                assert shown[base + turned_after + k] == s.q[k][b], (b, k)
                assert shown[base + turned_before + k] == earlier[base + turned_after + k], (b, k)
        assert shown[centre_after + 1] < shown[centre_before + 1], "the first box is falling"
        assert shown[turned_after] != shown[turned_before], "and turning"


@pytest.mark.parametrize("backend", BACKENDS)
@pytest.mark.parametrize("pairs", [[0, 2], [-1, 0], [1, 0, 0, 7]])
def test_the_narrow_phase_refuses_a_body_out_of_the_world(backend, pairs):
    """vr_box_collide names each pair's bodies; one outside 0 .. bodies - 1 is an IllegalArgumentException, found
    before anything runs (the kernel loads both bodies of every pair as given)."""
    s = BoxScene(2).uniform(0.1, 0.1, 0.1, 1.0)
    s.x[:] = 0.5
    s.y[:] = [0.1, 0.25]
    s.z[:] = 0.5
    with s.start(backend):
        with pytest.raises(sc.pkg.IllegalArgumentException, match="names body"):
            s.sim.collide(np.array(pairs, dtype=np.int32))
        assert s.sim.collide(np.array([0, 1], dtype=np.int32)).shape[0] == BoxContact.MANIFOLD_WORDS
