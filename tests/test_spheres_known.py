"""The sphere world's known answers, on each backend: the reference's own test cases, restated.

Each test names the reference test it restates (ctest/sphere/... = vexelray-sim-rigid-core/src/test/java/dev/vexelray/
sim/rigid/sphere/...). The "oracle" leg runs the CPU restatement and pins it to closed forms; the "gpu" leg runs the
HIP library through the C ABI and is held to the same numbers, as the reference holds each backend on its own.
"""

import math

import numpy as np
import pytest

import scene as sc
from scene import Scene

pkg = sc.pkg
GS = pkg.Solve.GAUSS_SEIDEL_BY_CONTACT
JACOBI = pkg.Solve.JACOBI

BACKENDS = ["oracle", pytest.param("gpu", marks=pytest.mark.gpu)]
EVERY_SOLVE = [(b, s) for b in BACKENDS for s in (JACOBI, GS)]
G = 9.81


def _every_solve():
    out = []
    for b in ("oracle", "gpu"):
        for s in (JACOBI, GS):
            marks = [pytest.mark.gpu] if b == "gpu" else []
            out.append(pytest.param(b, s, marks=marks, id=f"{b}-{s.name}"))
    return out


EVERY = _every_solve()


# ---- SpheresTest.java ----

@pytest.mark.parametrize("backend,solve", EVERY)
def test_falling_sphere_follows_symplectic_euler_exactly(backend, solve):
    """SpheresTest.aFallingSphereFollowsSymplecticEulerExactly (:41-70)."""
    with Scene(1).uniform(0.1, 1) as s:
        s.sx = s.sy = s.sz = 10
        s.x[0] = s.z[0] = 5
        s.y[0] = 8
        s.solve = solve
        s.start(backend).advance(30)
        s.read()
        k = 30 * s.substeps
        h = s.dt / s.substeps
        assert abs(-9.81 * h * k - s.v[0]) <= 1e-4
        assert abs(8 - 9.81 * h * h * k * (k + 1) / 2 - s.y[0]) <= 1e-4
        assert s.x[0] == 5


@pytest.mark.parametrize("backend,solve", EVERY)
def test_dropped_sphere_rests_on_floor(backend, solve):
    """SpheresTest.aDroppedSphereComesToRestOnTheFloor (:75-88)."""
    with Scene(1).uniform(0.1, 1) as s:
        s.x[0] = s.z[0] = 0.5
        s.y[0] = 0.8
        s.solve = solve
        s.start(backend).advanceSeconds(1)
        st = s.diagnostics()
        assert abs(s.y[0] - 0.1) <= 1e-6
        assert st.maxSpeed < 1e-4
        assert st.maxWall < 1e-5


@pytest.mark.parametrize("backend,solve", EVERY)
def test_collision_keeps_momentum(backend, solve):
    """SpheresTest.aCollisionKeepsMomentumAndTheSpheresLeaveTogether (:94-115)."""
    with Scene(2).uniform(0.1, 1) as s:
        s.gy = 0
        s.sx = 4
        s.y[:] = 0.5
        s.z[:] = 0.5
        s.x[0] = 1
        s.x[1] = 2
        s.u[0] = 2
        s.im[1] = np.float32(1 / 3)
        s.solve = solve
        s.start(backend).advanceSeconds(1)
        st = s.diagnostics()
        assert abs(st.momentum[0] - 2) <= 1e-6
        assert abs(s.u[0] - 0.5) <= 1e-3
        assert abs(s.u[1] - 0.5) <= 1e-3
        assert s.x[0] < s.x[1]
        assert st.maxOverlap < 1e-3
        assert not st.broken


# ---- BounceTest.java ----

RADIUS = 0.05
DROP = 1.0


def _rebound(backend, solve, e):
    with Scene(1).uniform(RADIUS, 1) as s:
        s.sx = s.sz = 0.4
        s.sy = 2
        s.substeps = 20
        s.solve = solve
        s.restitution = e
        s.x[0] = s.z[0] = 0.2
        s.y[0] = RADIUS + DROP
        s.start(backend)
        hit = False
        highest = 0.0
        for _ in range(150):
            s.advance(1).read()
            above = float(s.y[0]) - RADIUS
            if not hit and s.v[0] > 0:
                hit = True
            if hit:
                highest = max(highest, above)
                if s.v[0] < 0 and highest > 0:
                    break
        return highest


@pytest.mark.parametrize("backend,solve", EVERY)
def test_ball_rebounds_to_square_of_restitution(backend, solve):
    """BounceTest.aBallReboundsToTheSquareOfItsRestitution (:37-49)."""
    for e in (0.5, 0.9):
        rose = _rebound(backend, solve, e)
        assert abs(e * e * DROP - rose) <= 0.01 * DROP, (e, rose)


@pytest.mark.parametrize("backend,solve", EVERY)
def test_without_restitution_no_bounce(backend, solve):
    """BounceTest.withoutRestitutionABallDoesNotBounce (:51-55)."""
    assert _rebound(backend, solve, 0) < 0.01 * DROP


@pytest.mark.parametrize("backend,solve", EVERY)
def test_elastic_collision_swaps_velocities(backend, solve):
    """BounceTest.anElasticCollisionSwapsTheVelocities (:90-104)."""
    with Scene(2).uniform(RADIUS, 1) as s:
        s.gy = 0
        s.sx = 2
        s.sy = s.sz = 0.4
        s.substeps = 20
        s.solve = solve
        s.restitution = 1
        s.x[0] = 0.5
        s.x[1] = 0.8
        s.y[:] = 0.2
        s.z[:] = 0.2
        s.u[0] = 1
        s.start(backend).advanceSeconds(0.5).read()
        assert abs(s.u[0] + s.u[1] - 1) <= 1e-5
        assert abs(s.u[0]) <= 0.03
        assert abs(s.u[1] - 1) <= 0.03


@pytest.mark.parametrize("backend,solve", EVERY)
def test_ball_at_rest_stays_at_rest(backend, solve):
    """BounceTest.aBallAtRestStaysAtRest (:111-125)."""
    with Scene(1).uniform(RADIUS, 1) as s:
        s.sx = s.sy = s.sz = 0.4
        s.substeps = 10
        s.solve = solve
        s.restitution = 0.9
        s.x[0] = s.z[0] = 0.2
        s.y[0] = RADIUS
        s.start(backend).advanceSeconds(2).read()
        assert abs(s.v[0]) < 1e-3
        assert abs(s.y[0] - RADIUS) <= 1e-4


# ---- FrictionTest.java ----

def _on_the_floor(mu):
    s = Scene(1).uniform(0.1, 1)
    s.sx = 4
    s.x[0] = 0.5
    s.y[0] = 0.1
    s.z[0] = 0.5
    s.solve = GS
    s.muStatic = s.muKinetic = mu
    return s


@pytest.mark.parametrize("backend", BACKENDS)
def test_sliding_sphere_settles_into_rolling(backend):
    """FrictionTest.aSlidingSphereSettlesIntoRollingAtFiveSevenths (:38-57)."""
    mu, v0 = 0.3, 2.0
    with _on_the_floor(mu) as s:
        s.u[0] = v0
        s.start(backend).advanceSeconds(0.1)
        s.read()
        t = 0.1
        assert abs(v0 - mu * G * t - s.u[0]) <= 1e-2
        assert abs(-2.5 * mu * G * t / 0.1 - s.az[0]) <= 0.1
        s.advanceSeconds(0.4)
        s.diagnostics()
        assert abs(5 * v0 / 7 - s.u[0]) <= 5e-3
        assert abs(-s.u[0] / 0.1 - s.az[0]) <= 1e-2
        assert abs(0.1 - s.y[0]) <= 1e-5
        assert abs(s.ax[0]) <= 1e-6
        assert abs(s.ay[0]) <= 1e-6


@pytest.mark.parametrize("backend", BACKENDS)
def test_on_a_slope_rolls_or_slides(backend):
    """FrictionTest.onASlopeASphereRollsWithEnoughFrictionAndSlidesWithout (:66-90)."""
    theta = math.radians(20)
    t = 0.5
    with _on_the_floor(0.5) as s:
        s.x[0] = 0.2
        s.gx = G * math.sin(theta)
        s.gy = -G * math.cos(theta)
        s.start(backend).advanceSeconds(t)
        s.read()
        assert abs(5.0 / 7 * G * math.sin(theta) * t - s.u[0]) <= 1e-2
        assert abs(-s.u[0] / 0.1 - s.az[0]) <= 1e-2
    mu = 0.05
    with _on_the_floor(mu) as s:
        s.x[0] = 0.2
        s.gx = G * math.sin(theta)
        s.gy = -G * math.cos(theta)
        s.start(backend).advanceSeconds(t)
        s.read()
        assert abs(G * (math.sin(theta) - mu * math.cos(theta)) * t - s.u[0]) <= 1e-2
        assert abs(-2.5 * mu * G * math.cos(theta) * t / 0.1 - s.az[0]) <= 0.05


@pytest.mark.parametrize("backend", BACKENDS)
@pytest.mark.parametrize("mu_static,mu_kinetic", [(0.6, 0.2), (0.2, 0.6)])
def test_a_slide_is_slowed_by_kinetic_friction_alone(mu_static, mu_kinetic, backend):
    """Static and kinetic friction apart, which no reference test sets apart (Spheres.friction, Spheres.java:
    595-654), on each backend: a sphere sliding at 2 m/s on the floor slows at mu_k g and spins up at 5/2 mu_k g / r
    until it rolls, whatever mu_s is. Read at 0.05 s, before it rolls at either mu_k."""
    v0, t = 2.0, 0.05
    with _on_the_floor(0) as s:
        s.muStatic, s.muKinetic = mu_static, mu_kinetic
        s.u[0] = v0
        s.start(backend).advanceSeconds(t).read()
        assert abs(v0 - mu_kinetic * G * t - s.u[0]) <= 1e-2
        assert abs(-2.5 * mu_kinetic * G * t / 0.1 - s.az[0]) <= 0.1


@pytest.mark.parametrize("mu_static,mu_kinetic,rolls", [(0.3, 0.05, True), (0.1, 0.05, False), (0.05, 0.1, False),
                                                      (0.05, 0.3, True)])
@pytest.mark.parametrize("backend", BACKENDS)
def test_on_a_slope_friction_decides_whether_it_rolls(mu_static, mu_kinetic, rolls, backend):
    """On a 30 degree slope rolling without slipping takes (2/7) tan 30 = 0.165 of the normal impulse, at 5/7 g sin t.
    Friction holds it whole within mu_s, else cuts it to mu_k, but never past what stops the slip (Spheres.friction's
    min(1, mu_k lambda_n / |whole|), Spheres.java:636-641). So the sphere rolls when either mu reaches 0.165 and
    otherwise slides at g (sin t - mu_k cos t). On each backend, to 1e-2 after 0.5 s."""
    theta, t = math.radians(30), 0.5
    with _on_the_floor(0) as s:
        s.muStatic, s.muKinetic = mu_static, mu_kinetic
        s.x[0] = 0.2
        s.gx = G * math.sin(theta)
        s.gy = -G * math.cos(theta)
        s.start(backend).advanceSeconds(t).read()
        if rolls:
            assert abs(5.0 / 7 * G * math.sin(theta) * t - s.u[0]) <= 1e-2
            assert abs(-s.u[0] / 0.1 - s.az[0]) <= 1e-2
        else:
            assert abs(G * (math.sin(theta) - mu_kinetic * math.cos(theta)) * t - s.u[0]) <= 1e-2
            assert abs(-2.5 * mu_kinetic * G * math.cos(theta) * t / 0.1 - s.az[0]) <= 0.05


@pytest.mark.parametrize("omega", [1.0, 0.5])
@pytest.mark.parametrize("averaged", [True, False])
def test_a_jacobi_move_is_the_summed_shares_relaxed(omega, averaged):
    """Spheres.Solving (Spheres.java:1259-1297), on the oracle: each overlap adds this sphere's share of it,
    (reach - dist) w_s / w_sum along the separation, and the sum is scaled by omega, or by omega over the sphere's own
    contact count when averaged (Spheres.java:75-88). Sphere A is pressed 1 cm by B along x and by C along y; B and C
    are apart. One Jacobi pass, no gravity: A moves omega (0.005, 0.005) back, halved when averaged (two contacts),
    and B and C each move omega 0.005 away (one contact each)."""
    with Scene(3).uniform(0.1, 1) as s:
        s.gy = 0
        s.substeps = 1
        s.solve = JACOBI
        s.omega, s.averaged = omega, averaged
        s.x[:], s.y[:], s.z[:] = 0.5, 0.5, 0.5
        s.x[1] = np.float32(0.69)
        s.y[2] = np.float32(0.69)
        x0, y0 = s.x.astype(np.float64), s.y.astype(np.float64)
        s.start("oracle").advance(1).read()
        a = omega * 0.005 * (0.5 if averaged else 1.0)
        b = omega * 0.005
        moved = np.stack([s.x - x0, s.y - y0])
        want = np.array([[-a, b, 0.0], [-a, 0.0, b]])
        assert np.all(np.abs(moved - want) <= 1e-6), moved
        assert np.all(s.z == np.float32(0.5))


def _glancing(mu):
    s = Scene(2).uniform(0.1, 1)
    s.gy = 0
    s.sx = 4
    s.im[1] = np.float32(1 / 3)
    s.x[0] = 1
    s.x[1] = 1.5
    s.y[0] = 0.5
    s.y[1] = 0.55
    s.z[:] = 0.5
    s.u[0] = 2
    s.solve = GS
    s.muStatic = s.muKinetic = mu
    return s


@pytest.mark.parametrize("backend", BACKENDS)
def test_a_crowd_of_600_keeps_its_momentum_through_a_step(backend):
    """docs/TODO.md:123-125: through one step of 600 crowded spheres of mixed masses, Gauss-Seidel until done, the
    momentum stays at zero to 1e-6, so no two contacts that share a sphere were solved in the same round (two such
    solved side by side on the device would each read the sphere before the other moved it, and the pair of moves
    would no longer be equal and opposite by mass). GridTest.crowd's draws (radius 0.02-0.05 m, mass 0.1-1.1 kg,
    java.util.Random(5)) packed into a 0.4 m cube at the middle of a 6 m box, far enough from the walls that nothing
    reaches them within the step, no gravity, at rest. The spheres fly apart at up to ~110 m/s, so the sum is judged
    against the momentum's own scale, sum m |v| (f32 velocities carry about 1e-7 of it each): 1e-6 of it."""
    n = 600
    rnd = pkg.JavaRandom(5)
    with Scene(n) as s:
        s.sx = s.sy = s.sz = 6.0
        s.gy = 0.0
        for k in range(n):
            s.r[k] = np.float32(0.02 + 0.03 * rnd.nextDouble())
            s.im[k] = np.float32(1 / (0.1 + rnd.nextDouble()))
            s.x[k] = np.float32(2.8 + 0.4 * rnd.nextDouble())
            s.y[k] = np.float32(2.8 + 0.4 * rnd.nextDouble())
            s.z[k] = np.float32(2.8 + 0.4 * rnd.nextDouble())
        s.solve = GS
        s.start(backend).advance(1).read()
        m = 1 / s.im.astype(np.float64)
        assert s.x.min() > 2 * 0.05 and s.x.max() < 6 - 2 * 0.05, "a sphere reached a wall"
        for name in ("u", "v", "w"):
            vel = getattr(s, name).astype(np.float64)
            scale = float(np.sum(m * np.abs(vel)))
            assert scale > 100.0, "the crowd did not part"
            assert abs(float(np.sum(m * vel))) <= 1e-6 * scale, (name, float(np.sum(m * vel)), scale)


@pytest.mark.parametrize("backend", BACKENDS)
def test_glancing_collision_keeps_momentum_and_angular_momentum(backend):
    """FrictionTest.aGlancingCollisionKeepsMomentumAndAngularMomentumToFirstOrder (:110-150)."""
    frictionless = [0.0, 0.0]
    with_friction = [0.0, 0.0]
    for run in range(2):
        for mu in (0.0, 0.5):
            with _glancing(mu) as s:
                s.substeps = 10 << run
                start = pkg.SphereDiagnostics.of(s.x, s.y, s.z, s.u, s.v, s.w, s.ax, s.ay, s.az, s.r, s.im, 0,
                                                 s.sx, s.sy, s.sz)
                s.start(backend).advanceSeconds(0.5)
                end = s.diagnostics()
                for k in range(3):
                    assert abs(start.momentum[k] - end.momentum[k]) <= 1e-5
                change = abs(end.angular[2] - start.angular[2])
                assert abs(end.angular[0]) <= 1e-6
                assert abs(start.angular[1] - end.angular[1]) <= 1e-6
                if mu == 0:
                    frictionless[run] = change
                    assert end.rotational == 0
                else:
                    with_friction[run] = change
                    assert abs(s.az[0]) > 1 and abs(s.az[1]) > 0.1
                    assert end.kinetic < start.kinetic
        assert with_friction[run] <= frictionless[run]
    assert with_friction[1] < 0.6 * with_friction[0]
    assert with_friction[0] < 1.5e-3


@pytest.mark.parametrize("backend", BACKENDS)
def test_sphere_wedged_on_a_slope_rests(backend):
    """FrictionTest.aSphereWedgedOnASlopeRests (:159-180)."""
    theta = math.radians(20)
    with Scene(2).uniform(0.03, 0.1) as s:
        s.solve = GS
        s.muStatic = s.muKinetic = 0.5
        s.sx = s.sy = s.sz = 0.5
        s.gx = G * math.sin(theta)
        s.gy = -G * math.cos(theta)
        s.x[0] = 0.2
        s.x[1] = 0.26
        s.y[:] = 0.03
        s.z[:] = 0.25
        s.im[1] = 0
        s.start(backend).advanceSeconds(1)
        s.read()
        assert abs(s.az[0]) < 1.5e-4
        assert math.hypot(s.u[0], s.v[0]) < 5e-6
        assert abs(0.2 - s.x[0]) <= 1e-5


@pytest.mark.parametrize("backend", BACKENDS)
def test_sphere_in_a_pocket_stops_spinning(backend):
    """FrictionTest.aSphereInAPocketStopsSpinning (:189-215)."""
    r = 0.03
    with Scene(4).uniform(r, 0.1) as s:
        s.solve = GS
        s.muStatic = s.muKinetic = 0.5
        s.sx = s.sy = s.sz = 0.5
        c = 0.25
        for k in range(3):
            angle = 2 * math.pi * k / 3
            s.x[k] = c + 2 * r / math.sqrt(3) * math.cos(angle)
            s.z[k] = c + 2 * r / math.sqrt(3) * math.sin(angle)
            s.y[k] = r
            s.im[k] = 0
        s.x[3] = s.z[3] = c
        s.y[3] = r + 2 * r * math.sqrt(2.0 / 3)
        s.ax[3], s.ay[3], s.az[3] = 1, 2, 3
        s.start(backend).advanceSeconds(1)
        s.read()
        spin = math.sqrt(float(s.ax[3]) ** 2 + float(s.ay[3]) ** 2 + float(s.az[3]) ** 2)
        assert spin < 2e-3


@pytest.mark.parametrize("backend", BACKENDS)
def test_column_rests_with_friction(backend):
    """FrictionTest.aColumnStillRestsWithFriction (:219-232)."""
    with Scene(10).uniform(0.05, 1) as s:
        for k in range(10):
            s.x[k] = s.z[k] = 0.5
            s.y[k] = np.float32(0.05) + np.float32(0.1) * np.float32(k)
        s.solve = GS
        s.muStatic = s.muKinetic = 0.5
        s.start(backend).advanceSeconds(2)
        st = s.diagnostics()
        assert st.kinetic < 1e-6
        assert st.maxOverlap < 0.05
        assert abs(s.x[9] - 0.5) < 1e-4


# ---- RollingTest.java ----

def _rolling(v0, mu_rolling):
    s = Scene(1).uniform(0.1, 1)
    s.sx = 6
    s.x[0] = 0.5
    s.y[0] = 0.1
    s.z[0] = 0.5
    s.u[0] = v0
    s.az[0] = -v0 / 0.1
    s.solve = GS
    s.muStatic = s.muKinetic = 0.5
    s.muRolling = mu_rolling
    return s


@pytest.mark.parametrize("backend", BACKENDS)
def test_rolling_sphere_slows_at_five_sevenths(backend):
    """RollingTest.aRollingSphereSlowsAtFiveSeventhsOfMuGAndStops (:40-58)."""
    mu = 0.1
    with _rolling(1, mu) as s:
        s.start(backend).advanceSeconds(0.5)
        s.read()
        expected = 1 - 5.0 / 7 * mu * G * 0.5
        assert abs(expected - s.u[0]) <= 1e-2
        assert abs(-s.u[0] / 0.1 - s.az[0]) <= 5e-2
        s.advanceSeconds(1.5)
        s.read()
        assert abs(s.u[0]) < 1e-4
        assert abs(s.az[0]) < 1e-3
        stopped = float(s.x[0])
        s.advanceSeconds(1)
        s.read()
        assert abs(stopped - s.x[0]) <= 1e-5


@pytest.mark.parametrize("backend", BACKENDS)
def test_without_rolling_resistance_rolls_on(backend):
    """RollingTest.withoutItASphereRollsOn (:62-70)."""
    with _rolling(1, 0) as s:
        s.start(backend).advanceSeconds(1)
        s.read()
        assert abs(1 - s.u[0]) <= 1e-3
        assert abs(-10 - s.az[0]) <= 1e-2


@pytest.mark.parametrize("backend", BACKENDS)
def test_sphere_spinning_on_floor_stops(backend):
    """RollingTest.aSphereSpinningOnTheFloorStops (:77-88)."""
    with _rolling(0, 0.1) as s:
        s.ay[0] = 5
        s.start(backend).advanceSeconds(0.1)
        s.read()
        assert abs(5 - 0.1 * G / (0.4 * 0.1) * 0.1 - s.ay[0]) <= 5e-2
        s.advanceSeconds(0.5)
        s.read()
        assert abs(s.ay[0]) < 1e-3


# ---- SpinTest.java ----

@pytest.mark.parametrize("backend,solve", EVERY)
def test_spin_is_kept_and_turns_at_its_rate(backend, solve):
    """SpinTest.aSpinIsKeptAndTurnsTheSphereAtItsRate (:29-51)."""
    with Scene(1).uniform(0.1, 1) as s:
        s.gy = 0
        s.x[0] = s.y[0] = s.z[0] = 0.5
        rate = np.float32(math.pi / 3)
        s.ax[0] = rate
        s.ay[0] = 2 * rate
        s.az[0] = 2 * rate
        s.solve = solve
        s.start(backend).advanceSeconds(1)
        s.read()
        assert s.ax[0] == rate and s.ay[0] == 2 * rate and s.az[0] == 2 * rate
        q = s.orientation(0)
        expected = [1 / 3.0, 2 / 3.0, 2 / 3.0, 0]
        for k in range(4):
            assert abs(expected[k] - q[k]) <= 2e-4
        assert s.x[0] == np.float32(0.5)


@pytest.mark.parametrize("backend,solve", EVERY)
def test_picture_gets_orientation_before_and_after(backend, solve):
    """SpinTest.thePictureIsGivenTheOrientationBeforeAndAfterTheStep (:57-78)."""
    with Scene(1).uniform(0.1, 1) as s:
        s.gy = 0
        s.x[0] = s.y[0] = s.z[0] = 0.5
        s.ay[0] = 3
        s.solve = solve
        s.start(backend).advance(2)
        earlier = s.shown().copy()
        s.advance(1)
        shown = s.shown()
        q = s.orientation(0)
        for k in range(4):
            assert q[k] == shown[pkg.Spheres.TURNED_AFTER + k]
            assert earlier[pkg.Spheres.TURNED_AFTER + k] == shown[pkg.Spheres.TURNED_BEFORE + k]
        assert abs(0.075 - q[1]) <= 1e-3


@pytest.mark.parametrize("backend,solve", EVERY)
def test_spinning_sphere_rests_still_spinning(backend, solve):
    """SpinTest.aSpinningSphereRestsOnTheFloorStillSpinning (:84-95)."""
    with Scene(1).uniform(0.1, 2) as s:
        s.x[0] = s.z[0] = 0.5
        s.y[0] = 0.3
        s.ax[0] = 10
        s.solve = solve
        s.start(backend).advanceSeconds(1)
        st = s.diagnostics()
        assert s.ax[0] == 10
        assert abs(0.1 - s.y[0]) <= 1e-6
        assert abs(0.5 * 0.4 * 2 * 0.1 * 0.1 * 100 - st.rotational) <= 1e-6
        assert st.kinetic - st.rotational < 1e-6


# ---- PressingTest.java ----

def _pressing(sim_scene, s):
    p = np.zeros(12)
    c = np.array([sim_scene.x[s], sim_scene.y[s], sim_scene.z[s]], dtype=np.float64)
    r = float(sim_scene.r[s])
    extent = (sim_scene.sx, sim_scene.sy, sim_scene.sz)
    pairs = ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))

    def add(f, n):
        for k in range(6):
            p[k] += f * n[pairs[k][0]] * n[pairs[k][1]]
        for a in range(3):
            p[6 + a] += f * n[a]

    for a in range(3):
        depths = (r - c[a], c[a] + r - extent[a])
        for side in range(2):
            if depths[side] > 0:
                n = np.zeros(3)
                n[a] = -1 if side == 0 else 1
                add(depths[side] / r, n)
    for o in range(sim_scene.n):
        d = np.array([sim_scene.x[o], sim_scene.y[o], sim_scene.z[o]], dtype=np.float64) - c
        dist = float(np.linalg.norm(d))
        reach = r + float(sim_scene.r[o])
        if o != s and dist < reach:
            add((reach - dist) / reach, d / dist)
    p[9] = p[0] + p[1] + p[2]
    p[10] = math.sqrt(p[6] ** 2 + p[7] ** 2 + p[8] ** 2)
    return p


@pytest.mark.parametrize("backend,solve", EVERY)
def test_pressing_is_the_overlaps_the_step_ended_with(backend, solve):
    """PressingTest.thePressingIsTheOverlapsTheStepEndedWith (:22-62)."""
    with Scene(4) as s:
        at = [(0.05, 0.5, 0.5), (0.2, 0.55, 0.5), (0.25, 0.72, 0.5), (0.5, 0.95, 0.5)]
        radii = [0.1, 0.1, 0.15, 0.1]
        for k in range(4):
            s.x[k], s.y[k], s.z[k] = at[k]
            s.r[k] = radii[k]
            s.im[k] = 0
        s.solve = solve
        s.start(backend).advance(1)
        earlier = s.shown().copy()
        s.advance(1).read()
        shown = s.shown()
        expected = []
        for k in range(4):
            base = k * pkg.Spheres.SHOWN_STRIDE
            e = _pressing(s, k)
            expected.append(e)
            for w in range(pkg.Spheres.LEAN_LENGTH + 1):
                assert abs(e[w] - shown[base + pkg.Spheres.PRESSED_AFTER + w]) <= 1e-6
                assert earlier[base + pkg.Spheres.PRESSED_AFTER + w] == shown[base + pkg.Spheres.PRESSED_BEFORE + w]
        assert abs(0.7094 - expected[0][pkg.Spheres.TRACE]) <= 1e-3
        assert abs(-0.3013 - expected[0][pkg.Spheres.LEAN]) <= 1e-3
        assert abs(0.2912 - expected[2][pkg.Spheres.TRACE]) <= 1e-3
        assert abs(0.5 - expected[3][pkg.Spheres.LEAN + 1]) <= 1e-3


# ---- GridTest.java and UntilDoneTest.java ----

@pytest.mark.parametrize("backend", BACKENDS)
def test_grid_finds_what_every_pair_finds(backend):
    """GridTest.theGridFindsWhatEveryPairFinds (:25-44)."""
    pairs = sc.crowd()
    grid = sc.crowd()
    grid.grid = True
    with pairs, grid:
        pairs.start(backend).advance(1).read()
        grid.start(backend).advance(1).read()
        worst = max(float(np.max(np.abs(pairs.x - grid.x))), float(np.max(np.abs(pairs.y - grid.y))),
                    float(np.max(np.abs(pairs.z - grid.z))))
        assert worst < 1e-5


def test_grid_covers_the_box():
    """GridTest.aGridCoversTheBoxWithCellsAsWideAsTheWidestSphere (:47-55)."""
    g = pkg.SphereGrid.covering(1, 0.45, 0.1, 0.1)
    assert (g.nx, g.ny, g.nz, g.cells()) == (10, 5, 1, 50)
    with pytest.raises(pkg.IllegalArgumentException):
        pkg.SphereGrid.covering(1, 1, 1, 0)


@pytest.mark.parametrize("backend", BACKENDS)
def test_until_done_is_enough_rounds_to_the_bit(backend):
    """UntilDoneTest.untilDoneIsEnoughRoundsAndLeavesNothingUnsolved (:34-52)."""
    enough = pkg.SphereStep.CYCLE - 1
    with sc.pile() as s:
        s.fixedRounds = True
        s.rounds = enough
        s.start(backend).advanceSeconds(0.25).read()
        fixed = [s.x.copy(), s.y.copy(), s.z.copy(), s.v.copy()]
        assert s.contactCount()[pkg.Spheres.MISSED] == 0
    with sc.pile() as s:
        s.start(backend).advanceSeconds(0.25).read()
        assert s.mostRounds <= enough
        assert s.unlisted == 0
        for a, b in zip(fixed, (s.x, s.y, s.z, s.v)):
            assert np.array_equal(a, b)


# ---- validation (SphereStep.java:117-131, Spheres.java:208-223) ----

def test_validation_errors():
    with pytest.raises(pkg.IllegalArgumentException):
        pkg.SphereStep(0, 1)
    with pytest.raises(pkg.IllegalArgumentException):
        pkg.SphereStep(10, 1, 1, None, GS)
    with pytest.raises(pkg.IllegalArgumentException):
        pkg.SphereStep(46_341, 1, 1, pkg.SphereGrid(1, 1, 1, 1.0), GS)
    pkg.SphereStep(46_341, 1, 1, pkg.SphereGrid(1, 1, 1, 1.0), GS, widePairs=True)
    with pytest.raises(pkg.IllegalArgumentException):
        pkg.Spheres.params(0, 0, -9.81, 0, 1, 1, 1, 1, False)
    with pytest.raises(pkg.IllegalArgumentException):
        pkg.Spheres.params(0.01, 0, -9.81, 0, 1, 1, 1, 1, False, 1.5)
    with pytest.raises(pkg.IllegalArgumentException):
        pkg.Spheres.params(0.01, 0, -9.81, 0, 1, 1, 1, 1, False, 0, -1, 0)
    p = pkg.Spheres.params(0.01, 0, -9.81, 0, 1, 1, 1, 1, False, 0, 0.5, 0.5)
    assert pkg.Spheres.rubs(p)
    assert not pkg.Spheres.rubs(pkg.Spheres.params(0.01, 0, -9.81, 0, 1, 1, 1, 1, False))
    assert p.view(np.float32)[pkg.Spheres.GY] == np.float32(-9.81)


def test_oracle_world_rejects_bad_step():
    abi = sc.oracle_abi()
    d = pkg.SphereStep(10, 1).desc()
    d.bodies = 0
    with pytest.raises(pkg.IllegalArgumentException):
        abi.create(1, d, pkg.IllegalArgumentException, pkg.IllegalStateException)


def lattice(side, seed, spacing, y0, dy):
    """PileTest's loose lattice (core/src/test/.../sphere/PileTest.java:39-50, 77-88): java.util.Random(seed)."""
    rnd = sc.pkg.JavaRandom(seed)
    s = Scene(side ** 3).uniform(0.03, 0.1)
    k = 0
    for i in range(side):
        for j in range(side):
            for l in range(side):
                s.x[k] = np.float32(0.07 + spacing * i + 0.005 * rnd.nextDouble())
                s.y[k] = np.float32(y0 + dy * j)
                s.z[k] = np.float32(0.07 + spacing * l + 0.005 * rnd.nextDouble())
                k += 1
    return s


@pytest.mark.parametrize("grid,solve,mu", [(False, JACOBI, 0.0), (True, JACOBI, 0.0), (True, GS, 0.0),
                                            (True, GS, 0.5)])
def test_a_dropped_pile_stays_inside_apart_and_loses_energy(grid, solve, mu):
    """PileTest.aDroppedPileStaysInsideApartAndLosesEnergy (PileTest.java:30-68), on the oracle: 125 spheres dropped
    into a box for three seconds. Nothing goes non-finite or leaves the box, no sphere passes into another by a fifth
    of a radius, and the pile ends with less energy than it started with."""
    with lattice(5, 11, 0.065, 0.2, 0.07) as s:
        s.sx = s.sz = 0.4
        s.sy = 1
        s.substeps = 20
        s.grid = grid
        s.solve = solve
        s.muStatic = s.muKinetic = mu
        zero = np.zeros(s.n, dtype=np.float32)
        start = sc.pkg.SphereDiagnostics.of(s.x, s.y, s.z, s.u, s.v, s.w, zero, zero, zero, s.r, s.im, s.gy, s.sx,
                                            s.sy, s.sz)
        s.start("oracle")
        worst = 0.0
        for second in range(3):
            state = s.advanceSeconds(1).diagnostics()
            assert not state.broken, f"non-finite after {second + 1} s"
            assert state.maxWall < 1e-5, f"outside the box: {state}"
            worst = max(worst, state.maxOverlap)
            assert state.kinetic + state.potential < start.kinetic + start.potential, f"energy grew: {state}"
        assert worst < 0.2, f"spheres passed into each other by {worst} of a radius"


def test_a_pile_with_friction_comes_to_rest():
    """PileTest.aPileWithFrictionComesToRest (PileTest.java:76-103), on the oracle: 343 spheres dropped with
    friction 0.5 at 10 substeps settle for four seconds, then lose under 5e-3 J a second and keep under 1e-4 J
    moving."""
    side = 7
    with lattice(side, 7, 0.06, 0.3, 0.07) as s:
        s.solve = GS
        s.muStatic = s.muKinetic = 0.5
        s.sx = s.sz = 0.08 + 0.06 * side
        s.sy = 0.5 + 0.07 * side
        s.start("oracle").advanceSeconds(4)
        settled = s.diagnostics().potential
        later = s.advanceSeconds(2).diagnostics()
        creep = (settled - later.potential) / 2
        assert creep < 5e-3, f"still creeping, {creep} J a second: {later}"
        assert later.kinetic < 1e-4, f"at rest: {later}"
