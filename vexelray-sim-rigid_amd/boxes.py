"""The box world's host API: the reference's Java box classes, restated over the C ABI.

* ``Boxes.params`` / ``Boxes.softness`` / ``Boxes.inertia`` — ``core/.../box/Boxes.java:173-268``
* ``BoxBroad.grid`` — ``core/.../box/BoxBroad.java:66-68``
* ``BoxStep`` — ``core/.../box/BoxStep.java:81-162`` (validation :82-85)
* ``BoxStepper.Report`` — ``core/.../box/BoxStepper.java:40``
* ``BoxSimulation`` — ``gui/.../BoxSimulation.java:22-189``: ``start``, ``params``, ``stepAndWait``, ``last``,
  ``state``, ``readShown``, ``close``
* ``BoxDiagnostics`` — ``core/.../box/BoxDiagnostics.java:12-74``

Bodies are boxes or spheres (``shape`` 1, radius in ``hx``); every step runs on the device through the HIP library.
"""

import ctypes
import math
import struct
from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from . import _abi
from .spheres import IllegalArgumentException, IllegalStateException, SphereGrid


def _bits(value: float) -> int:
    return struct.unpack("<i", struct.pack("<f", float(np.float32(value))))[0]


class Boxes:
    """Constants and the parameter packer of ``Boxes.java``."""

    PARAM_COUNT = 18
    (H, GX, GY, GZ, SX, SY, SZ, MU_STATIC, MU_KINETIC, MARGIN, BIAS, MASS_SCALE, IMPULSE_SCALE, STATIC_BIAS,
     STATIC_MASS_SCALE, STATIC_IMPULSE_SCALE, PUSH_SPEED, MU_ROLLING) = range(18)
    CONTACT_HERTZ = 150.0
    DAMPING_RATIO = 10.0
    PUSH = 3.0
    WALL_PASSES = 4
    CORNERS = 8
    WALLS = 6
    LAMBDA_WORDS = 48
    RUB_WORDS = 144
    ROLL_WORDS = 144
    BOX = ("x", "y", "z", "u", "v", "w", "lx", "ly", "lz", "qx", "qy", "qz", "qw", "im", "hx", "hy", "hz", "shape")
    BOX_SHAPE = 0.0
    SPHERE_SHAPE = 1.0
    SPHERE_INERTIA = 0.4
    SHOWN_STRIDE = 20

    @staticmethod
    def softness(hertz, zeta, h):
        """Box2D's b2MakeSoft: [bias rate, mass scale, impulse scale] (Boxes.java:242-251)."""
        if hertz == 0:
            return [0.0, 1.0, 0.0]
        omega = 2 * math.pi * hertz
        a1 = 2 * zeta + h * omega
        a2 = h * omega * a1
        a3 = 1 / (1 + a2)
        return [omega / a1, a2 * a3, a3]

    @staticmethod
    def params(h, gx, gy, gz, sx, sy, sz, muStatic=0.0, muKinetic=0.0, margin=0.0, muRolling=0.0) -> np.ndarray:
        """``Boxes.params`` (Boxes.java:212-234): validated, the softness worked out in double, packed as f32 bits."""
        if not (muRolling >= 0) or math.isinf(muRolling):
            raise IllegalArgumentException(f"rolling resistance is finite and at least 0, got {muRolling}")
        if not (margin >= 0) or math.isinf(margin):
            raise IllegalArgumentException(f"a margin is finite and at least 0, got {margin}")
        if not (h > 0 and sx > 0 and sy > 0 and sz > 0):
            raise IllegalArgumentException(f"h > 0 and a box of positive size, got h {h}, box {sx} × {sy} × {sz}")
        if not (muStatic >= 0 and muKinetic >= 0) or math.isinf(muStatic) or math.isinf(muKinetic):
            raise IllegalArgumentException(f"friction finite and at least 0, got {muStatic} and {muKinetic}")
        hertz = min(Boxes.CONTACT_HERTZ, 0.25 / h)
        soft = Boxes.softness(hertz, Boxes.DAMPING_RATIO, h)
        stiff = Boxes.softness(2 * hertz, Boxes.DAMPING_RATIO, h)
        values = (h, gx, gy, gz, sx, sy, sz, muStatic, muKinetic, margin, soft[0], soft[1], soft[2], stiff[0],
                  stiff[1], stiff[2], Boxes.PUSH, muRolling)
        return np.array([_bits(v) for v in values], dtype=np.int32)

    @staticmethod
    def inertia(shape, m, hx, hy, hz):
        """A body's principal moments (Boxes.java:257-268)."""
        if shape > 0.5:
            moment = Boxes.SPHERE_INERTIA * m * hx * hx
            return [moment, moment, moment]
        return [m * (hy * hy + hz * hz) / 3, m * (hx * hx + hz * hz) / 3, m * (hx * hx + hy * hy) / 3]


class BoxBroad:
    SLOTS = 32

    @staticmethod
    def grid(sx, sy, sz, widestHalfDiagonal, margin) -> SphereGrid:
        """Cells as wide as the widest bounding sphere plus the margin (BoxBroad.java:66-68)."""
        return SphereGrid.covering(sx, sy, sz, 2 * widestHalfDiagonal + margin)


class BoxContact:
    MANIFOLD_WORDS = 40
    NORMAL, COUNT, AXIS, POINTS, POINT_WORDS = 0, 3, 4, 8, 8
    ANCHOR_FIRST, ANCHOR_SECOND, DEPTH = 0, 3, 6
    MOST_POINTS = 4
    SPHERE_BOX, SPHERE_SPHERE = 15, 16


class BoxStep:
    """``BoxStep``'s constructor (BoxStep.java:81-162)."""

    MOST_ROUNDS = 32

    def __init__(self, boxes: int, substeps: int, sweeps: int = Boxes.WALL_PASSES, symmetric: bool = False,
                 grid: Optional[SphereGrid] = None):
        if boxes < 1 or substeps < 1 or sweeps < 1:
            raise IllegalArgumentException(
                f"a step needs boxes, substeps and sweeps, got {boxes}, {substeps} and {sweeps}")
        self.boxes = boxes
        self.substeps = substeps
        self.sweeps = sweeps
        self.symmetric = symmetric
        self.grid = grid if boxes > 1 else None
        self.slots = boxes * BoxBroad.SLOTS if boxes > 1 and grid is not None else 0
        self.width = max(1, boxes // 2)

    def desc(self) -> "_abi.StepDesc":
        g = self.grid
        return _abi.StepDesc(bodies=self.boxes, substeps=self.substeps, iterations=1, has_grid=1 if g else 0,
                             nx=g.nx if g else 0, ny=g.ny if g else 0, nz=g.nz if g else 0,
                             cell=g.cell if g else 0.0, solve=0, rounds=1, contacts_per_sphere=1, fixed_rounds=0,
                             wide_pairs=0, sweeps=self.sweeps, symmetric=1 if self.symmetric else 0)


@dataclass
class BoxReport:
    """``BoxStepper.Report`` (BoxStepper.java:40)."""

    touching: int = 0
    points: int = 0
    rounds: int = 0
    unplaced: int = 0
    unlisted: int = 0


@dataclass
class SolveReport:
    """vr_box_solve over the step's substeps: per substep, sweeps launched and (measured) sweeps a scene ran on average
    and the largest residual a scene ended with, for the solve and the relax; and the pairs the learned start left to
    the carry."""

    tolerance: float
    measured: bool
    scene_step: bool
    launched: np.ndarray      # 2 x substeps
    scene_sweeps: np.ndarray  # 2 x substeps
    residual: np.ndarray      # 2 x substeps
    fell_back: int
    residual_mean: np.ndarray  # 2 x substeps

    @staticmethod
    def of(raw, substeps) -> "SolveReport":
        m = min(substeps, _abi.SOLVE_SUBSTEPS)
        grab = lambda field, dtype: np.array([[getattr(raw, field)[k][s] for s in range(m)] for k in range(2)],  # noqa
                                             dtype=dtype)
        return SolveReport(float(raw.tolerance), bool(raw.measured), bool(raw.scene_step),
                           grab("launched", np.int32), grab("scene_sweeps", np.float32),
                           grab("residual", np.float32), int(raw.fell_back), grab("residual_mean", np.float32))


class BoxSimulation:
    """Boxes stepped on resident device buffers (``gui/.../BoxSimulation.java``)."""

    def __init__(self, step: BoxStep, abi: Optional["_abi.Abi"] = None):
        self._abi = abi if abi is not None else _abi.product()
        self.step_ = step
        self._world = self._abi.create(_abi.VR_BOXES, step.desc(), IllegalArgumentException, IllegalStateException)
        self._last = BoxReport()

    def bodies(self) -> int:
        return self.step_.boxes

    def shownWords(self) -> int:
        return Boxes.SHOWN_STRIDE * self.step_.boxes

    def write(self, name, values):
        self._world.write(name, values)

    def floats(self, name) -> np.ndarray:
        return self._world.read(name, np.float32)

    def ints(self, name) -> np.ndarray:
        return self._world.read(name, np.int32)

    def clear(self):
        self._world.clear()

    def start(self, x, y, z, u, v, w, spin, q, half, shape, im, params) -> None:
        """``BoxSimulation.start`` (BoxSimulation.java:77-104): spin is the angular velocity in the world's frame,
        written as the angular momentum it makes."""
        n = self.step_.boxes
        self._world.clear()
        spin = np.asarray(spin, dtype=np.float64)
        q = np.asarray(q, dtype=np.float64)
        half = np.asarray(half, dtype=np.float64)
        shape = np.asarray(shape, dtype=np.float32)
        im = np.asarray(im, dtype=np.float32)
        l = np.zeros((3, n), dtype=np.float32)
        for s in (np.flatnonzero(np.any(spin != 0, axis=0)) if n else []):  # no spin, no angular momentum
            if im[s] <= 0:
                continue
            moment = Boxes.inertia(shape[s], 1 / float(im[s]), half[0][s], half[1][s], half[2][s])
            l[:, s] = BoxDiagnostics.momentumOf(spin[:, s], q[:, s], moment)
        for name, arr in (("x", x), ("y", y), ("z", z), ("u", u), ("v", v), ("w", w), ("lx", l[0]), ("ly", l[1]),
                          ("lz", l[2]), ("qx", q[0]), ("qy", q[1]), ("qz", q[2]), ("qw", q[3]), ("hx", half[0]),
                          ("hy", half[1]), ("hz", half[2]), ("shape", shape), ("im", im)):
            self._world.write(name, np.asarray(arr, dtype=np.float32))
        self.params(params)

    def params(self, params) -> None:
        self._world.params(np.asarray(params, dtype=np.int32))

    def step(self) -> BoxReport:
        return self.stepAndWait()

    def stepAndWait(self) -> BoxReport:
        rep = self._world.step(True)
        self._last = BoxReport(rep.touching, rep.points, rep.rounds, rep.unplaced, rep.unlisted)
        self.wall_ns, self.gpu_ns = rep.wall_ns, rep.gpu_ns
        self.submissions = rep.submissions  # StepReport.submissions (BoxSimulation.java:116-142)
        self.learnedPoints = rep.reserved
        return self._last

    def last(self) -> BoxReport:
        return self._last

    def collideList(self) -> BoxReport:
        """The first half of ``BoxStepper.step`` (BoxStepper.java:62-95): the collide list (keep, broad phase,
        narrow phase, carry, counts) and the colouring, with no substep run (vr_box_collide_list)."""
        rep = self._world.collide_list()
        return BoxReport(rep.touching, rep.points, rep.rounds, rep.unplaced, rep.unlisted)

    def collide(self, pairs) -> np.ndarray:
        """The narrow phase alone over ``pairs`` (first, second, first, second, ...)."""
        return self._world.collide(pairs)

    # Sweeps stopped on a residual (no counterpart in the reference): vr_box_tolerance / vr_box_solve_read /
    # vr_box_solve_curve.
    def tolerance(self, tol: float, measure: bool = False) -> None:
        """Each scene's solve (and relax) sweeps of a substep end after the first sweep whose points' residuals (the
        impulse change times the inverse mass along the normal, m/s) are all at most `tol`; below zero, every sweep
        runs (the reference). With `measure`, each step keeps its phases' sweeps and residuals."""
        if tol != tol:
            raise IllegalArgumentException("a tolerance is a number (below zero for none)")
        self._world.tolerance(float(tol), measure)

    # The launch plan (no counterpart in the reference): vr_box_retune / vr_box_plan.
    def retune(self) -> None:
        """Times the batch's launch plans again over its next steps (two each) and keeps the fastest."""
        self._world.retune()

    def plan(self):
        """(the plan the next step takes: pass, all, graph, scene or group; whether the plans are still being timed)."""
        return self._world.plan()

    def solveReport(self) -> "SolveReport":
        """What the last step's phases did (vr_box_solve)."""
        return SolveReport.of(self._world.solve_read(), self.step_.substeps)

    def solveCurve(self) -> np.ndarray:
        """The last measured step's largest residual of each sweep: [substep][solve, relax][sweep]."""
        return self._world.solve_curve()

    # The learned warm start (no counterpart in the reference): vr_box_log / vr_box_samples / vr_box_learned.
    FEATURES, TARGETS = 16, 3

    def log(self, on: bool, capacity: int = 1 << 20) -> None:
        """Starts (clearing it) or stops the device log of per-point (features, targets) of every step."""
        self._world._check(self._abi.fn("box_log")(self._world.handle, 1 if on else 0, int(capacity)))

    def samples(self, most: int = 1 << 20):
        """Up to `most` logged samples: features (n x 16) and targets (n x 3)."""
        x = np.empty((most, self.FEATURES), dtype=np.float32)
        t = np.empty((most, self.TARGETS), dtype=np.float32)
        count = ctypes.c_int64()
        f32 = ctypes.POINTER(ctypes.c_float)
        self._world._check(self._abi.fn("box_samples")(self._world.handle, x.ctypes.data_as(f32),
                                                        t.ctypes.data_as(f32), most, ctypes.byref(count)))
        return x[:count.value], t[:count.value]

    def learned(self, net) -> None:
        """Starts each step's impulses from `net`'s prediction (an Mlp of the same library), or from the carry
        (None), as the reference does. The world holds the network until learned(None) or close(), so it is kept
        here as long: the library keeps only its handle."""
        if net is not None and getattr(net, "_h", None) is None:
            raise IllegalArgumentException("the network is closed")
        self._world._check(self._abi.fn("box_learned")(self._world.handle, net._h if net is not None else None))
        self._net = net

    def diagnostics(self):
        """BoxDiagnostics.of on the device (vr_world_diagnostics), with the top speed, whether any state is not
        finite, and the deepest manifold point of the last collide: (BoxDiagnostics, speed, broken, depth)."""
        d = self._world.diagnostics()
        return (BoxDiagnostics(float(d[2]), float(d[3]), d[5:8].copy(), d[8:11].copy()), float(d[11]),
                bool(d[12] != 0), float(d[13]))

    def state(self) -> List[np.ndarray]:
        """``BoxSimulation.state`` (BoxSimulation.java:150-160): x, y, z, u, v, w, lx, ly, lz, qx, qy, qz, qw."""
        return [self.floats(k) for k in ("x", "y", "z", "u", "v", "w", "lx", "ly", "lz", "qx", "qy", "qz", "qw")]

    def readShown(self) -> np.ndarray:
        return self.ints("shown")

    def finish(self):
        self._world.finish()

    def close(self):
        if self._world is not None:
            self._world.destroy()
            self._world = None
        self._net = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BoxBatch(BoxSimulation):
    """Many independent box scenes stepped together in one world (vr_batch_create; the reference steps one scene
    a world). Scene k holds bodies[k] consecutive bodies in a walled box extents[k] = (sx, sy, sz) with broad-phase
    grid grids[k] (a SphereGrid, as BoxBroad.grid makes for the scene). Buffers, start(), state() and the report
    are the box world's over all bodies; params() serves every scene, its box ignored for each scene's own."""

    def __init__(self, bodies, extents, grids, substeps: int, sweeps: int = Boxes.WALL_PASSES,
                 symmetric: bool = False, abi: Optional["_abi.Abi"] = None):
        self._abi = abi if abi is not None else _abi.product()
        bodies = [int(b) for b in bodies]
        if len(bodies) != len(extents) or len(bodies) != len(grids):
            raise IllegalArgumentException("a batch needs a box and a grid for every scene")
        self.scenes = len(bodies)
        self.firsts = np.concatenate([[0], np.cumsum(bodies)]).astype(np.int64)
        self.step_ = BoxStep(int(self.firsts[-1]), substeps, sweeps, symmetric, SphereGrid(1, 1, 1, 1.0))
        cells = [[g.nx, g.ny, g.nz] for g in grids]
        sizes = [g.cell for g in grids]
        self._world = self._abi.create_batch(self.step_.desc(), bodies, extents, cells, sizes,
                                             IllegalArgumentException, IllegalStateException)
        self._last = BoxReport()

    def scene(self, k: int) -> slice:
        """The bodies of scene k, as a slice of the world's buffers."""
        return slice(int(self.firsts[k]), int(self.firsts[k + 1]))

    def collide(self, pairs) -> np.ndarray:
        raise IllegalStateException("the narrow phase alone runs on a world of one scene")


@dataclass
class BoxDiagnostics:
    """``BoxDiagnostics`` (BoxDiagnostics.java:12-74), in doubles."""

    kinetic: float
    rotational: float
    momentum: np.ndarray
    angular: np.ndarray

    @staticmethod
    def of(x, y, z, u, v, w, l, q, half, shape, im) -> "BoxDiagnostics":
        kinetic = rotational = 0.0
        momentum = np.zeros(3)
        angular = np.zeros(3)
        for s in range(len(x)):
            if im[s] <= 0:
                continue
            m = 1 / float(im[s])
            at = np.array([x[s], y[s], z[s]], dtype=np.float64)
            vel = np.array([u[s], v[s], w[s]], dtype=np.float64)
            spun = np.array([l[0][s], l[1][s], l[2][s]], dtype=np.float64)
            omega = BoxDiagnostics.spin(spun, [q[0][s], q[1][s], q[2][s], q[3][s]],
                                        Boxes.inertia(shape[s], m, half[0][s], half[1][s], half[2][s]))
            orbit = np.cross(at, vel)
            for a in range(3):
                kinetic += 0.5 * m * vel[a] * vel[a]
                rotational += 0.5 * spun[a] * omega[a]
                momentum[a] += m * vel[a]
                angular[a] += m * orbit[a] + spun[a]
        return BoxDiagnostics(kinetic + rotational, rotational, momentum, angular)

    @staticmethod
    def rotate(q, p):
        q = np.asarray(q, dtype=np.float64)
        p = np.asarray(p, dtype=np.float64)
        t = 2 * np.cross(q[:3], p)
        return p + q[3] * t + np.cross(q[:3], t)

    @staticmethod
    def conjugate(q):
        return np.array([-q[0], -q[1], -q[2], q[3]], dtype=np.float64)

    @staticmethod
    def spin(l, q, moment):
        own = BoxDiagnostics.rotate(BoxDiagnostics.conjugate(q), l)
        return BoxDiagnostics.rotate(q, own / np.asarray(moment, dtype=np.float64))

    @staticmethod
    def momentumOf(omega, q, moment):
        own = BoxDiagnostics.rotate(BoxDiagnostics.conjugate(q), omega)
        return BoxDiagnostics.rotate(q, own * np.asarray(moment, dtype=np.float64))
