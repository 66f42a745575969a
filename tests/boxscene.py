"""The box test harness: a set of boxes (and spheres) in a walled box, put on a backend and stepped.

Restates ctest/box/BoxScene.java (vexelray-sim-rigid-core/src/test/java/dev/vexelray/sim/rigid/box/BoxScene.java):
set the fields, start(backend), advance(frames), read(), diagnostics(). Backends as in scene.py: "gpu" (the HIP
library) and "oracle" (the CPU restatement).
"""

import numpy as np

import scene as sc

pkg = sc.pkg
boxes = __import__(pkg.__name__ + ".boxes", fromlist=["boxes"])
Boxes, BoxBroad, BoxStep, BoxSimulation, BoxDiagnostics, BoxContact = (
    boxes.Boxes, boxes.BoxBroad, boxes.BoxStep, boxes.BoxSimulation, boxes.BoxDiagnostics, boxes.BoxContact)


class BoxScene:
    """BoxScene.java's fields and defaults (BoxScene.java:29-103)."""

    def __init__(self, n):
        self.n = n
        z = lambda: np.zeros(n, dtype=np.float32)
        self.x, self.y, self.z, self.u, self.v, self.w = z(), z(), z(), z(), z(), z()
        self.ax, self.ay, self.az = z(), z(), z()
        self.l = np.zeros((3, n), dtype=np.float32)
        self.q = np.zeros((4, n), dtype=np.float32)
        self.q[3] = 1
        self.half = np.zeros((3, n), dtype=np.float32)
        self.shape = z()
        self.im = np.ones(n, dtype=np.float32)
        self.gx, self.gy, self.gz = 0.0, -9.81, 0.0
        self.sx = self.sy = self.sz = 1.0
        self.dt = 1.0 / 60
        self.substeps = 10
        self.muStatic = 0.0
        self.muKinetic = 0.0
        self.muRolling = 0.0
        self.passes = Boxes.WALL_PASSES
        self.margin = 0.005
        self.symmetric = False
        self.mostRounds = 0
        self.unplaced = 0
        self.unlisted = 0
        self.last = None
        self.step = None
        self.sim = None

    def sphere(self, s, r):
        self.shape[s] = Boxes.SPHERE_SHAPE
        self.half[:, s] = np.float32(r)
        return self

    def uniform(self, hx, hy, hz, mass):
        self.half[0][:] = np.float32(hx)
        self.half[1][:] = np.float32(hy)
        self.half[2][:] = np.float32(hz)
        self.im[:] = np.float32(1 / mass)
        return self

    def start(self, backend):
        h = self.half.astype(np.float64)
        widest = float(np.max(np.sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]))) if self.n else 0.0
        step = BoxStep(self.n, self.substeps, self.passes, self.symmetric,
                       BoxBroad.grid(self.sx, self.sy, self.sz, widest, self.margin))
        self.step = step
        self.sim = BoxSimulation(step, abi=sc.abi_for(backend))
        self.sim.clear()
        self.sim.params(Boxes.params(self.dt / self.substeps, self.gx, self.gy, self.gz, self.sx, self.sy, self.sz,
                                     self.muStatic, self.muKinetic, self.margin, self.muRolling))
        for s in range(self.n):
            if self.im[s] <= 0:
                continue
            moment = Boxes.inertia(self.shape[s], 1 / float(self.im[s]), float(self.half[0][s]),
                                   float(self.half[1][s]), float(self.half[2][s]))
            spun = BoxDiagnostics.momentumOf([self.ax[s], self.ay[s], self.az[s]], self.q[:, s].astype(np.float64),
                                           moment)
            self.l[:, s] = spun.astype(np.float32)
        for name, arr in (("x", self.x), ("y", self.y), ("z", self.z), ("u", self.u), ("v", self.v), ("w", self.w),
                          ("lx", self.l[0]), ("ly", self.l[1]), ("lz", self.l[2]), ("hx", self.half[0]),
                          ("hy", self.half[1]), ("hz", self.half[2]), ("qx", self.q[0]), ("qy", self.q[1]),
                          ("qz", self.q[2]), ("qw", self.q[3]), ("im", self.im), ("shape", self.shape)):
            self.sim.write(name, arr)
        return self

    def advance(self, frames):
        for _ in range(frames):
            self.last = self.sim.stepAndWait()
            self.mostRounds = max(self.mostRounds, self.last.rounds)
            self.unplaced += self.last.unplaced
            self.unlisted += self.last.unlisted
        return self

    def advanceSeconds(self, seconds):
        return self.advance(int(round(seconds / self.dt)))

    def read(self):
        names = ("x", "y", "z", "u", "v", "w")
        for k in names:
            setattr(self, k, self.sim.floats(k).copy())
        for a, k in enumerate(("lx", "ly", "lz")):
            self.l[a] = self.sim.floats(k)
        for a, k in enumerate(("qx", "qy", "qz", "qw")):
            self.q[a] = self.sim.floats(k)
        for s in range(self.n):
            if self.im[s] > 0:
                moment = Boxes.inertia(self.shape[s], 1 / float(self.im[s]), float(self.half[0][s]),
                                       float(self.half[1][s]), float(self.half[2][s]))
                om = BoxDiagnostics.spin(self.l[:, s].astype(np.float64), self.q[:, s].astype(np.float64), moment)
            else:
                om = np.zeros(3)
            self.ax[s], self.ay[s], self.az[s] = om
        return self

    def diagnostics(self):
        self.read()
        return BoxDiagnostics.of(self.x, self.y, self.z, self.u, self.v, self.w, self.l, self.q, self.half,
                                 self.shape, self.im)

    def ownSpin(self, s):
        """Body s's spin in its own frame, as the state was last read (BoxScene.java's ownSpin)."""
        return BoxDiagnostics.rotate(BoxDiagnostics.conjugate(self.q[:, s].astype(np.float64)),
                                     [float(self.ax[s]), float(self.ay[s]), float(self.az[s])])

    def axis(self, s, k):
        """Body s's own axis k in the world's frame, as the state was last read."""
        unit = [0.0, 0.0, 0.0]
        unit[k] = 1.0
        return BoxDiagnostics.rotate(self.q[:, s].astype(np.float64), unit)

    def slots(self):
        return self.step.slots

    def deepest(self):
        """The deepest point of any manifold the last step left (the pile tests' own sweep over the slots)."""
        m = self.floats("manifold")
        most = 0.0
        for p in range(self.slots()):
            at = p * BoxContact.MANIFOLD_WORDS
            for c in range(int(m[at + BoxContact.COUNT])):
                most = max(most, float(m[at + BoxContact.POINTS + c * BoxContact.POINT_WORDS + BoxContact.DEPTH]))
        return most

    def floats(self, name):
        return self.sim.floats(name)

    def ints(self, name):
        return self.sim.ints(name)

    def close(self):
        if self.sim is not None:
            self.sim.close()
            self.sim = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def collide(backend, margin, bodies):
    """BoxContactTest.collide (ctest/box/BoxContactTest.java:441-477): the narrow phase alone over pairs (bodies[0],
    bodies[1]), (bodies[2], bodies[3]), ... Each body is (centre, half, q, shape)."""
    n = len(bodies)
    step = BoxStep(n, 1, 1, False, None)
    sim = BoxSimulation(step, abi=sc.abi_for(backend))
    arr = lambda f: np.array([f(b) for b in bodies], dtype=np.float64).astype(np.float32)
    for k, name in enumerate(("x", "y", "z")):
        sim.write(name, arr(lambda b, k=k: b[0][k]))
    for k, name in enumerate(("qx", "qy", "qz", "qw")):
        sim.write(name, arr(lambda b, k=k: b[2][k]))
    for k, name in enumerate(("hx", "hy", "hz")):
        sim.write(name, arr(lambda b, k=k: b[1][k]))
    sim.write("shape", arr(lambda b: b[3]))
    sim.params(Boxes.params(1e-3, 0, 0, 0, 1, 1, 1, 0, 0, margin))
    out = sim.collide(np.arange(n, dtype=np.int32))
    sim.close()
    return out
