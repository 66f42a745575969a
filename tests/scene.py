"""The test harness: a set of spheres in a box, put on a backend and stepped.

Restates ctest/sphere/Scene.java (vexelray-sim-rigid-core/src/test/java/dev/vexelray/sim/rigid/sphere/Scene.java):
set the fields, start(backend), advance(frames), read(), diagnostics(). Backends: "gpu", the HIP library (the
product), and "oracle", the CPU restatement in oracle/ that the product is checked against.
"""

import os

import numpy as np

from __graft_entry__ import ROOT, load_package

pkg = load_package()
_abi = __import__(pkg.__name__ + "._abi", fromlist=["_abi"])

ORACLE_LIB = os.path.join(ROOT, "oracle", "liboracle_rigid.so")
_ORACLE = None


def oracle_abi():
    global _ORACLE
    if _ORACLE is None:
        _ORACLE = _abi.Abi(ORACLE_LIB, "vro_", "the CPU oracle")
    return _ORACLE


def abi_for(backend):
    if backend == "oracle":
        return oracle_abi()
    if backend == "gpu":
        return _abi.product()
    raise ValueError(backend)


class Scene:
    """Scene.java's fields and defaults (Scene.java:20-98)."""

    def __init__(self, n):
        self.n = n
        z = lambda: np.zeros(n, dtype=np.float32)
        self.x, self.y, self.z, self.u, self.v, self.w = z(), z(), z(), z(), z(), z()
        self.ax, self.ay, self.az = z(), z(), z()
        self.r = z()
        self.im = np.ones(n, dtype=np.float32)
        self.gx = 0.0
        self.gy = -9.81
        self.gz = 0.0
        self.sx = self.sy = self.sz = 1.0
        self.dt = 1.0 / 60
        self.substeps = 10
        self.iterations = 1
        self.omega = 1.0
        self.averaged = True
        self.restitution = 0.0
        self.muStatic = 0.0
        self.muKinetic = 0.0
        self.muRolling = 0.0
        self.grid = False
        self.solve = pkg.Solve.JACOBI
        self.rounds = pkg.SphereStep.ROUNDS
        self.fixedRounds = False
        self.widePairs = False
        self.contactsPerSphere = pkg.SphereStep.CONTACTS_PER_SPHERE
        self.waits = 0
        self.roundsRun = 0
        self.mostRounds = 0
        self.unlisted = 0
        self.sim = None

    def uniform(self, radius, mass):
        self.r[:] = np.float32(radius)
        self.im[:] = np.float32(1 / mass)
        return self

    def widest(self):
        return float(np.max(2 * self.r.astype(np.float64))) if self.n else 0.0

    def start(self, backend):
        gridded = self.grid or self.solve != pkg.Solve.JACOBI
        grid = pkg.SphereGrid.covering(self.sx, self.sy, self.sz, self.widest()) if gridded else None
        step = pkg.SphereStep(self.n, self.substeps, self.iterations, grid, self.solve, self.rounds,
                              self.contactsPerSphere, fixedRounds=self.fixedRounds, widePairs=self.widePairs)
        self.sim = pkg.SphereSimulation(step, abi=abi_for(backend))
        params = pkg.Spheres.params(self.dt / self.substeps, self.gx, self.gy, self.gz, self.sx, self.sy, self.sz,
                                    self.omega, self.averaged, self.restitution, self.muStatic, self.muKinetic,
                                    self.muRolling)
        # Scene.start: clear, constants, params, then the state (Scene.java:126-140); shown is left cleared.
        self.sim.clear()
        self.sim.params(params)
        self.sim.friction(self.muStatic + self.muKinetic + self.muRolling > 0)
        for name in ("x", "y", "z", "u", "v", "w", "ax", "ay", "az", "r", "im"):
            self.sim.write(name, getattr(self, name))
        return self

    def advance(self, frames):
        for _ in range(frames):
            rep = self.sim.stepAndWait()
            self.waits += rep.solve.waits
            self.roundsRun += rep.solve.rounds
            self.mostRounds = max(self.mostRounds, rep.solve.mostRounds)
            self.unlisted += rep.solve.unlisted
        return self

    def advanceSeconds(self, seconds):
        return self.advance(int(round(seconds / self.dt)))

    def read(self):
        for name in ("x", "y", "z", "u", "v", "w", "ax", "ay", "az"):
            setattr(self, name, self.sim.floats(name).copy())
        return self

    def orientation(self, s):
        return np.array([self.sim.floats(k)[s] for k in ("qx", "qy", "qz", "qw")])

    def floats(self, name):
        return self.sim.floats(name)

    def ints(self, name):
        return self.sim.ints(name)

    def shown(self):
        return self.sim.floats("shown")

    def contactCount(self):
        return self.sim.contactCount()

    def diagnostics(self):
        self.read()
        return pkg.SphereDiagnostics.of(self.x, self.y, self.z, self.u, self.v, self.w, self.ax, self.ay, self.az,
                                        self.r, self.im, self.gy, self.sx, self.sy, self.sz)

    def close(self):
        if self.sim is not None:
            self.sim.close()
            self.sim = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def pile(n=300, seed=5):
    """UntilDoneTest.pile (ctest/sphere/UntilDoneTest.java:56-68): 300 mixed spheres crowded in a 0.5 x 1 x 0.5 box."""
    rnd = pkg.JavaRandom(seed)
    s = Scene(n)
    s.sx = s.sz = 0.5
    s.sy = 1.0
    s.substeps = 10
    s.solve = pkg.Solve.GAUSS_SEIDEL_BY_CONTACT
    for k in range(n):
        s.r[k] = np.float32(0.02 + 0.015 * rnd.nextDouble())
        s.im[k] = np.float32(1 / (0.05 + 0.2 * rnd.nextDouble()))
        s.x[k] = np.float32(0.05 + 0.4 * rnd.nextDouble())
        s.y[k] = np.float32(0.05 + 0.3 * rnd.nextDouble())
        s.z[k] = np.float32(0.05 + 0.4 * rnd.nextDouble())
    return s


def crowd(seed=5):
    """GridTest.crowd (ctest/sphere/GridTest.java:52-66): 600 spheres crowded into 0.6 x 0.4 x 0.5."""
    n = 600
    rnd = pkg.JavaRandom(seed)
    s = Scene(n)
    s.sx, s.sy, s.sz = 0.6, 0.4, 0.5
    for k in range(n):
        s.r[k] = np.float32(0.02 + 0.03 * rnd.nextDouble())
        s.im[k] = np.float32(1 / (0.1 + rnd.nextDouble()))
        s.x[k] = np.float32(-0.05 + 0.7 * rnd.nextDouble())
        s.y[k] = np.float32(0.4 * rnd.nextDouble())
        s.z[k] = np.float32(0.5 * rnd.nextDouble())
    return s


STATE = ("x", "y", "z", "u", "v", "w", "ax", "ay", "az", "qx", "qy", "qz", "qw")


def pile_until_done(steps, backend, friction=0.5):
    """The 300-sphere pile with friction and rolling resistance, stepped until done: the smoke check's scene."""
    s = pile()
    s.muStatic = s.muKinetic = friction
    s.muRolling = 0.05
    s.restitution = 0.3
    s.start(backend).advance(steps)
    out = {k: s.floats(k).copy() for k in STATE}
    out["contactCount"] = s.ints("contactCount").copy()
    s.close()
    return out
