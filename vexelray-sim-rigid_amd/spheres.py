"""The sphere world's host API: the reference's Java world/body/step classes, restated over the C ABI.

Names, argument meanings and exceptions follow the reference (paths relative to the vexelray-sim-rigid repository):

* ``Spheres.params`` / ``Spheres.rubs`` — ``core/.../sphere/Spheres.java:173-235``
* ``SphereGrid`` / ``SphereGrid.covering`` — ``core/.../sphere/SphereGrid.java:11-36``
* ``SphereStep`` — ``core/.../sphere/SphereStep.java:115-209`` (validation :117-131)
* ``SphereStepper.Report`` — ``core/.../sphere/SphereStepper.java:42``
* ``SphereSimulation`` — ``gui/.../SphereSimulation.java:36-296``: ``start``, ``params``, ``step``,
  ``stepAndWait``, ``state``, ``readShown``, ``finish``, ``close``
* ``SphereDiagnostics.of`` — ``core/.../sphere/SphereDiagnostics.java:34-83``

Every step runs on the device, through the HIP library ``libvexelray_rigid.so`` built beside this package; there is
no CPU fallback, and a missing library is an error.
"""

import ctypes
import enum
import math
import os
import struct
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import _abi


class IllegalArgumentException(ValueError):
    """The reference's validation error."""


class IllegalStateException(RuntimeError):
    """The reference's state error: a pass that never ends, a wrong mode."""


def _f32_bits(value: float) -> int:
    return struct.unpack("<i", struct.pack("<f", float(np.float32(value))))[0]


class Spheres:
    """Constants and the parameter packer of ``Spheres.java``."""

    PARAM_COUNT = 13
    H, GX, GY, GZ, SX, SY, SZ, OMEGA, AVERAGED, RESTITUTION, MU_STATIC, MU_KINETIC, MU_ROLLING = range(13)
    INERTIA = 0.4
    LIST_MARGIN = 0.02
    MOST_LISTED_SPHERES = 46_340
    CONTACT_COUNT_WORDS = 3
    LISTED, MISSED, LONGEST = 0, 1, 2
    IMPULSE_WORDS = 7
    WALL_IMPULSE_WORDS = 21
    SHOWN_STRIDE = 40
    PRESSED_BEFORE, PRESSED_AFTER = 8, 20
    TURNED_BEFORE, TURNED_AFTER = 32, 36
    PRESSING_WORDS = 12
    FLATTENING, LEAN, TRACE, LEAN_LENGTH = 0, 6, 9, 10
    SPHERE = ("x", "y", "z", "u", "v", "w", "r", "im", "cx", "cy", "cz", "dx", "dy", "dz",
              "ax", "ay", "az", "tx", "ty", "tz", "qx", "qy", "qz", "qw")

    @staticmethod
    def params(h, gx, gy, gz, sx, sy, sz, omega, averaged, restitution=0.0, muStatic=0.0, muKinetic=0.0,
               muRolling=0.0) -> np.ndarray:
        """``Spheres.params`` (Spheres.java:208-226): validated, then packed as raw float bits."""
        finite = all(math.isfinite(m) for m in (muStatic, muKinetic, muRolling))
        if not (muStatic >= 0 and muKinetic >= 0 and muRolling >= 0) or not finite:
            raise IllegalArgumentException(
                f"friction coefficients are finite and at least 0, got {muStatic}, {muKinetic} and {muRolling}")
        if not (0 <= restitution <= 1):
            raise IllegalArgumentException(f"restitution is in [0, 1], got {restitution}")
        if not (h > 0 and sx > 0 and sy > 0 and sz > 0 and omega > 0 and omega <= 1):
            raise IllegalArgumentException(
                f"h > 0, a box of positive size and 0 < omega <= 1, got h {h}, box {sx} × {sy} × {sz}, "
                f"omega {omega}")
        values = (h, gx, gy, gz, sx, sy, sz, omega, 1 if averaged else 0, restitution, muStatic, muKinetic,
                  muRolling)
        return np.array([_f32_bits(v) for v in values], dtype=np.int32)

    @staticmethod
    def rubs(params) -> bool:
        """``Spheres.rubs`` (Spheres.java:232-235): whether there is friction or rolling resistance."""
        f = np.asarray(params, dtype=np.int32).view(np.float32)
        return float(np.float32(np.float32(f[10]) + np.float32(f[11])) + np.float32(f[12])) > 0


@dataclass(frozen=True)
class SphereGrid:
    """``SphereGrid`` (SphereGrid.java:11-36): nx × ny × nz cubic cells of side ``cell`` from the origin."""

    nx: int
    ny: int
    nz: int
    cell: float

    def __post_init__(self):
        if self.nx < 1 or self.ny < 1 or self.nz < 1 or not (self.cell > 0):
            raise IllegalArgumentException(
                f"a grid of at least one cell, of positive size, got {self.nx} × {self.ny} × {self.nz} of "
                f"{self.cell}")
        object.__setattr__(self, "cell", float(np.float32(self.cell)))

    @staticmethod
    def covering(sx, sy, sz, widest) -> "SphereGrid":
        if not (widest > 0):
            raise IllegalArgumentException(f"the widest sphere must have a positive diameter, got {widest}")

        def cells(extent):
            return max(1, int(math.ceil(extent / widest)))

        return SphereGrid(cells(sx), cells(sy), cells(sz), float(np.float32(widest)))

    def cells(self) -> int:
        return self.nx * self.ny * self.nz


class Solve(enum.IntEnum):
    """``SphereStep.Solve`` (SphereStep.java:37-43)."""

    JACOBI = 0
    GAUSS_SEIDEL_BY_CONTACT = 1


class SphereStep:
    """``SphereStep``'s constructor and constants (SphereStep.java:48-59, 115-209)."""

    ROUNDS = 32
    BATCH = 8
    CYCLE = 48
    CONTACTS_PER_SPHERE = 6

    def __init__(self, spheres: int, substeps: int, iterations: int = 1, grid: Optional[SphereGrid] = None,
                 solve: Solve = Solve.JACOBI, rounds: int = ROUNDS, contactsPerSphere: int = CONTACTS_PER_SPHERE,
                 fixedRounds: bool = False, widePairs: bool = False):
        if spheres < 1 or substeps < 1 or iterations < 1:
            raise IllegalArgumentException(
                f"a step needs spheres, substeps and iterations, got {spheres}, {substeps}, {iterations}")
        if solve != Solve.JACOBI and grid is None:
            raise IllegalArgumentException("Gauss–Seidel works over the grid's cells, and there is no grid")
        if rounds < 1 or contactsPerSphere < 1:
            raise IllegalArgumentException(
                f"at least a round and a place a sphere, got {rounds} and {contactsPerSphere}")
        if solve == Solve.GAUSS_SEIDEL_BY_CONTACT and spheres > Spheres.MOST_LISTED_SPHERES and not widePairs:
            raise IllegalArgumentException(
                f"a contact list orders at most {Spheres.MOST_LISTED_SPHERES} spheres, got {spheres}")
        self.spheres = spheres
        self.substeps = substeps
        self.iterations = iterations
        self.grid = grid
        self.solve = Solve(solve)
        self.rounds = rounds
        self.contactsPerSphere = contactsPerSphere
        self.capacity = spheres * contactsPerSphere
        self.fixedRounds = fixedRounds
        self.widePairs = widePairs

    def untilDone(self) -> bool:
        return self.solve == Solve.GAUSS_SEIDEL_BY_CONTACT and not self.fixedRounds

    @staticmethod
    def position(k: int) -> int:
        """``SphereStep.position`` (SphereStep.java:337-339)."""
        return 0 if k == 0 else 1 + (k - 1) % SphereStep.CYCLE

    def desc(self) -> "_abi.StepDesc":
        g = self.grid
        return _abi.StepDesc(
            bodies=self.spheres, substeps=self.substeps, iterations=self.iterations,
            has_grid=1 if g else 0, nx=g.nx if g else 0, ny=g.ny if g else 0, nz=g.nz if g else 0,
            cell=g.cell if g else 0.0, solve=int(self.solve), rounds=self.rounds,
            contacts_per_sphere=self.contactsPerSphere, fixed_rounds=1 if self.fixedRounds else 0,
            wide_pairs=1 if self.widePairs else 0)


@dataclass
class Report:
    """``SphereStepper.Report`` (SphereStepper.java:42)."""

    waits: int = 0
    rounds: int = 0
    mostRounds: int = 0
    unlisted: int = 0


@dataclass
class StepReport:
    """``SphereSimulation.StepReport`` (SphereSimulation.java:48)."""

    solve: Report
    submissions: int
    wallNanos: int
    gpuNanos: int


class SphereSimulation:
    """Spheres in a box, stepped on resident device buffers (``gui/.../SphereSimulation.java``).

    ``abi`` selects the library that runs the step; it defaults to the HIP library and only test code passes
    another one.
    """

    def __init__(self, step: SphereStep, abi: Optional["_abi.Abi"] = None):
        self._abi = abi if abi is not None else _abi.product()
        self.step_ = step
        self._world = self._abi.create(_abi.VR_SPHERES, step.desc(), IllegalArgumentException,
                                       IllegalStateException)

    # -- SphereSimulation's accessors --
    def spheres(self) -> int:
        return self.step_.spheres

    def bodies(self) -> int:
        return self.step_.spheres

    def substeps(self) -> int:
        return self.step_.substeps

    def iterations(self) -> int:
        return self.step_.iterations

    def solve(self) -> Solve:
        return self.step_.solve

    def shownWords(self) -> int:
        return Spheres.SHOWN_STRIDE * self.step_.spheres

    def where(self) -> str:
        return self._abi.where()

    # -- buffers --
    def write(self, name: str, values) -> None:
        self._world.write(name, values)

    def floats(self, name: str) -> np.ndarray:
        return self._world.read(name, np.float32)

    def ints(self, name: str) -> np.ndarray:
        return self._world.read(name, np.int32)

    def clear(self) -> None:
        """``PassRunner.clear()`` and the step's constants written (SphereStep.constants())."""
        self._world.clear()

    def start(self, x, y, z, u, v, w, r, im, params, ax=None, ay=None, az=None, shown=True) -> None:
        """``SphereSimulation.start`` (SphereSimulation.java:154-178), as ``vr_world_start``. The angular velocities,
        which the reference's test harness writes after the constants (ctest/.../sphere/Scene.java:136-138), may be
        given too; ``shown=False`` leaves the picture's buffer cleared, as that harness does."""
        fields = tuple(np.asarray(a, dtype=np.float32) for a in (x, y, z, u, v, w, r, im))
        if shown:
            self._world.start(fields, np.asarray(params, dtype=np.int32))
        else:
            self._world.clear()
            for name, values in zip(("x", "y", "z", "u", "v", "w", "r", "im"), fields):
                self._world.write(name, values)
            self.params(params)
        for name, values in (("ax", ax), ("ay", ay), ("az", az)):
            if values is not None:
                self._world.write(name, np.asarray(values, dtype=np.float32))

    def params(self, params) -> None:
        """``SphereSimulation.params`` (SphereSimulation.java:182-185); friction follows ``Spheres.rubs``."""
        self._world.params(np.asarray(params, dtype=np.int32))

    def friction(self, on: bool) -> None:
        """``SphereStepper.friction`` (SphereStepper.java:110-112)."""
        self._world.friction(on)

    def slice(self, dispatches: int) -> None:
        """``SphereSimulation.slice`` (SphereSimulation.java:136-148): at most ``dispatches`` dispatches a submission,
        0 for no bound; a negative count is an IllegalArgumentException. The step's kernels and results are the same
        whatever the slice; the report's ``submissions`` counts the pieces the reference's lists are cut into."""
        self._world.slice(dispatches)

    def step(self) -> StepReport:
        return self._run(False)

    def stepAndWait(self) -> StepReport:
        return self._run(True)

    def _run(self, wait: bool) -> StepReport:
        rep = self._world.step(wait)
        solve = Report(rep.waits, rep.rounds, rep.most_rounds, rep.unlisted)
        return StepReport(solve, rep.submissions, rep.wall_ns, rep.gpu_ns)

    def finish(self) -> None:
        self._world.finish()

    def state(self) -> List[np.ndarray]:
        """``SphereSimulation.state`` (SphereSimulation.java:250-254): x, y, z, u, v, w, ax, ay, az."""
        return [self.floats(k) for k in ("x", "y", "z", "u", "v", "w", "ax", "ay", "az")]

    def readShown(self) -> np.ndarray:
        return self.ints("shown")

    def diagnostics(self) -> "SphereDiagnostics":
        """SphereDiagnostics.of on the device (vr_world_diagnostics): f64 reductions, overlaps through the grid
        rebuilt at the current positions."""
        d = self._world.diagnostics()
        return SphereDiagnostics(float(d[0]), float(d[1]), float(d[2]), float(d[3]), float(d[4]), d[5:8].copy(),
                                 d[8:11].copy(), float(d[11]), bool(d[12] != 0))

    def contactCount(self) -> Optional[np.ndarray]:
        return self.ints("contactCount") if self.step_.solve == Solve.GAUSS_SEIDEL_BY_CONTACT else None

    def close(self) -> None:
        if self._world is not None:
            self._world.destroy()
            self._world = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class SphereDiagnostics:
    """``SphereDiagnostics`` (SphereDiagnostics.java:22-83), read on the host in double precision.

    The reference's overlap scan is over every pair; only overlapping pairs can raise the maximum above zero, so
    pairs are found here through a grid of the widest diameter, which gives the same value.
    """

    maxOverlap: float
    maxWall: float
    kinetic: float
    rotational: float
    potential: float
    momentum: np.ndarray
    angular: np.ndarray
    maxSpeed: float
    broken: bool

    @staticmethod
    def of(x, y, z, u, v, w, ax, ay, az, r, im, gy, sx, sy, sz) -> "SphereDiagnostics":
        f64 = lambda a: np.asarray(a, dtype=np.float32).astype(np.float64)
        x, y, z, u, v, w, ax, ay, az, r, im = map(f64, (x, y, z, u, v, w, ax, ay, az, r, im))
        n = len(x)
        broken = not all(np.all(np.isfinite(a)) for a in (x, y, z, u, v, w, ax, ay, az))
        max_overlap = _max_overlap(x, y, z, r)
        at = (x, y, z)
        extent = (sx, sy, sz)
        max_wall = 0.0
        for axis in range(3):
            max_wall = max(max_wall, float(np.max((r - at[axis]) / r)) if n else 0.0)
            max_wall = max(max_wall, float(np.max((at[axis] + r - extent[axis]) / r)) if n else 0.0)
        speed2 = u * u + v * v + w * w
        max_speed = float(np.sqrt(speed2.max())) if n else 0.0
        free = im > 0
        m = np.where(free, 1.0 / np.where(free, im, 1.0), 0.0)
        inertia = Spheres.INERTIA * m * r * r
        spin2 = ax * ax + ay * ay + az * az
        rotational = float(np.sum(0.5 * inertia * spin2))
        kinetic = float(np.sum(0.5 * m * speed2 + 0.5 * inertia * spin2))
        potential = float(np.sum(m * -gy * y))
        momentum = np.array([np.sum(m * u), np.sum(m * v), np.sum(m * w)])
        angular = np.array([np.sum(m * (y * w - z * v) + inertia * ax), np.sum(m * (z * u - x * w) + inertia * ay),
                            np.sum(m * (x * v - y * u) + inertia * az)])
        return SphereDiagnostics(max_overlap, max_wall, kinetic, rotational, potential, momentum, angular,
                                 max_speed, broken)


def _max_overlap(x, y, z, r) -> float:
    """Deepest overlap between two spheres over the smaller radius; 0 if none overlap."""
    n = len(x)
    if n < 2:
        return 0.0
    if n <= 2048:
        dx = x[:, None] - x[None, :]
        dy = y[:, None] - y[None, :]
        dz = z[:, None] - z[None, :]
        ov = (r[:, None] + r[None, :] - np.sqrt(dx * dx + dy * dy + dz * dz)) / np.minimum(r[:, None], r[None, :])
        np.fill_diagonal(ov, -np.inf)
        return max(0.0, float(ov.max()))
    cell = 2.0 * float(r.max())
    keys = np.floor(np.stack([x, y, z], axis=1) / cell).astype(np.int64)
    keys -= keys.min(axis=0)
    dims = keys.max(axis=0) + 3
    flat = (keys[:, 0] + 1) + dims[0] * ((keys[:, 1] + 1) + dims[1] * (keys[:, 2] + 1))
    order = np.argsort(flat, kind="stable")
    sorted_keys = flat[order]
    best = 0.0
    for ox in (-1, 0, 1):
        for oy in (-1, 0, 1):
            for oz in (-1, 0, 1):
                target = flat + ox + dims[0] * (oy + dims[1] * oz)
                lo = np.searchsorted(sorted_keys, target, side="left")
                hi = np.searchsorted(sorted_keys, target, side="right")
                counts = hi - lo
                if counts.max() == 0:
                    continue
                a = np.repeat(np.arange(n), counts)
                idx = np.concatenate([order[l:h] for l, h in zip(lo, hi) if h > l]) if counts.sum() else []
                b = np.asarray(idx, dtype=np.int64)
                keep = b > a
                a, b = a[keep], b[keep]
                if len(a) == 0:
                    continue
                d = np.sqrt((x[a] - x[b]) ** 2 + (y[a] - y[b]) ** 2 + (z[a] - z[b]) ** 2)
                ov = (r[a] + r[b] - d) / np.minimum(r[a], r[b])
                best = max(best, float(ov.max()))
    return best
