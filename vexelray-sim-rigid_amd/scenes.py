"""Seeded synthetic scenes for the benchmark workloads, drawn with java.util.Random (javarandom.py).

Each builder returns a SceneSpec: the SphereStep to make, the initial state, and the parameter words.

* pile_1k: 1,000 spheres of r = 0.5 m, density 1000 kg/m^3, centres uniform in a 20 x 40 x 20 m walled box, at
  rest; mu = 0.5, e = 0; dt = 1/60 s in 4 substeps, 8 Gauss-Seidel passes, 1 m cells.
* pile_64k: 65,536 spheres, r log-uniform in [0.1, 0.5] m, density 1000 kg/m^3, centres uniform in a
  100 x 50 x 100 m region of a walled box whose ceiling is at 60 m (out of reach, so the box acts as ground plus 4
  walls), velocity N(0, 1) m/s per axis; mu = 0.4, e = 0.1; dt = 1/60 s in 8 substeps, 10 passes, 1 m cells. More
  spheres than the reference's 31-bit contact priorities order, so the step uses 63-bit ones (wide pairs).

* sweep_column, sweep_pile(side): the reference's own measured sphere scenes (SphereSweepTest.java:93-143, with
  Scene.java's defaults): a column of 20 touching spheres of r = 5 cm, 1 kg, in a 0.1 x 2.5 x 0.1 m box; and a
  lattice of side^3 spheres of r = 3 cm, 0.1 kg (x, z = 0.07 + 0.06 i + 0.005 U, y = 0.3 + 0.07 j, drawn from
  java.util.Random(7)) in a box 0.08 + 0.06 side wide and 0.5 + 0.07 side tall. Both with friction mu = 0.5 static
  and kinetic, no rolling resistance, e = 0, dt = 1/60 s in 10 substeps of one Gauss-Seidel pass run until done,
  cells as wide as a sphere (SphereSweepTest's Solver.friction(10, 0.5)). run 180 frames (column) or 240 (piles).

* box_pile(side): the reference's measured box pile (BoxStackTest.java:184-232, with BoxScene.java's defaults): side^3
  cubes of half-extent 4 cm, 0.5 kg, x, z = 0.25 + 0.1 i + 0.02 U, y = 0.3 + 0.1 j, each turned by 0.6 U rad about the
  axis (U - 0.5, U - 0.5, U - 0.5), drawn from java.util.Random(7), in a box 0.1 side + 0.4 wide and 0.1 side + 1.5
  tall; mu = 0.5, margin 0.005 m; dt = 1/60 s in 10 substeps of 4 sweeps. Side 7 is its pile of 343, 10 its 1,000.
  run 240 frames: 3 s settling, then the second it times.

* mixed_32k: the mixed-shape pile, on the box path. 16,384 oriented boxes (half-extents uniform in [0.2, 1.0] m per
  axis, orientation a normalised Gaussian quaternion), 16,384 spheres (r uniform in [0.2, 0.6] m), density
  1000 kg/m^3, dropped at rest from heights 22-70 m into a 128 x 72 x 128 m walled box, beside 32 towers of 20
  stacked unit cubes (half-extent 0.5 m) standing on the floor in an 8 x 4 grid 16 m apart; mu = 0.6, margin
  0.005 m; dt = 1/60 s in 8 substeps of 20 sweeps, broad-phase cells of the widest bounding sphere plus margin.

* batched_2k: 2,048 independent scenes stepped together as one batch. Scene i holds N_i = round(exp(U[ln 16,
  ln 4096])) bodies (about 1.5 M in all), half boxes and half spheres drawn as in mixed_32k, in a walled cube of
  side 2 N_i^(1/3) m, on a lattice of that cube with jitter; mu = 0.6, margin 0.005 m; dt = 1/60 s in 8 substeps
  of 10 sweeps. Drawn with numpy's PCG64 from the seed (the reference has no such scene, and a million-body
  java.util.Random draw is too slow to make in Python).

Draw order per sphere: radius (if drawn), x, y, z, then the three velocity components. Per body of mixed_32k:
shape (box, then sphere, alternating), extents (3 halves, or the radius), quaternion (boxes: 4 Gaussians), x, y, z.
"""

import math
from dataclasses import dataclass

import numpy as np

from .javarandom import JavaRandom
from .boxes import BoxBroad, Boxes, BoxStep
from .spheres import Solve, SphereGrid, SphereStep, Spheres


@dataclass
class SceneSpec:
    name: str
    step: SphereStep
    params: np.ndarray
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    u: np.ndarray
    v: np.ndarray
    w: np.ndarray
    r: np.ndarray
    im: np.ndarray
    box: tuple
    steps: int
    dt: float
    seed: int


def _mass(r, density=1000.0):
    return density * 4.0 / 3.0 * math.pi * r ** 3


def pile_1k(seed: int = 1) -> SceneSpec:
    n = 1000
    rnd = JavaRandom(seed)
    box = (20.0, 40.0, 20.0)
    r = np.full(n, 0.5, dtype=np.float32)
    im = np.full(n, 1.0 / _mass(0.5), dtype=np.float32)
    xyz = np.array([[rnd.nextDouble() * box[0], rnd.nextDouble() * box[1], rnd.nextDouble() * box[2]]
                    for _ in range(n)], dtype=np.float64).astype(np.float32)
    zeros = np.zeros(n, dtype=np.float32)
    substeps, dt = 4, 1.0 / 60
    grid = SphereGrid.covering(*box, 1.0)
    step = SphereStep(n, substeps, 8, grid, Solve.GAUSS_SEIDEL_BY_CONTACT)
    params = Spheres.params(dt / substeps, 0, -9.81, 0, *box, 1.0, False, 0.0, 0.5, 0.5, 0.0)
    return SceneSpec("pile_1k", step, params, xyz[:, 0].copy(), xyz[:, 1].copy(), xyz[:, 2].copy(), zeros,
                     zeros.copy(), zeros.copy(), r, im, box, 300, dt, seed)


def pile_64k(seed: int = 2, n: int = 65_536) -> SceneSpec:
    rnd = JavaRandom(seed)
    region = (100.0, 50.0, 100.0)
    box = (100.0, 60.0, 100.0)
    lo, hi = math.log(0.1), math.log(0.5)
    r = np.empty(n, dtype=np.float64)
    pos = np.empty((n, 3), dtype=np.float64)
    vel = np.empty((n, 3), dtype=np.float64)
    for k in range(n):
        r[k] = math.exp(lo + (hi - lo) * rnd.nextDouble())
        pos[k] = (rnd.nextDouble() * region[0], rnd.nextDouble() * region[1], rnd.nextDouble() * region[2])
        vel[k] = (rnd.nextGaussian(), rnd.nextGaussian(), rnd.nextGaussian())
    im = (1.0 / _mass(r)).astype(np.float32)
    r32 = r.astype(np.float32)
    p32 = pos.astype(np.float32)
    v32 = vel.astype(np.float32)
    substeps, dt = 8, 1.0 / 60
    grid = SphereGrid.covering(*box, 1.0)
    step = SphereStep(n, substeps, 10, grid, Solve.GAUSS_SEIDEL_BY_CONTACT, widePairs=True)
    params = Spheres.params(dt / substeps, 0, -9.81, 0, *box, 1.0, False, 0.1, 0.4, 0.4, 0.0)
    return SceneSpec("pile_64k", step, params, p32[:, 0].copy(), p32[:, 1].copy(), p32[:, 2].copy(),
                     v32[:, 0].copy(), v32[:, 1].copy(), v32[:, 2].copy(), r32, im, box, 600, dt, seed)


def _sweep_scene(name, n, box, x, y, z, radius, mass, steps, seed, substeps=10, mu=0.5):
    """Scene.start for a Gauss-Seidel friction solve (Scene.java:107-128): one pass a substep, omega 1, averaged,
    cells of the widest sphere."""
    r = np.full(n, radius, dtype=np.float32)
    im = np.full(n, np.float32(1 / mass), dtype=np.float32)
    zeros = np.zeros(n, dtype=np.float32)
    dt = 1.0 / 60
    step = SphereStep(n, substeps, 1, SphereGrid.covering(*box, 2 * radius), Solve.GAUSS_SEIDEL_BY_CONTACT)
    params = Spheres.params(dt / substeps, 0, -9.81, 0, *box, 1.0, True, 0.0, mu, mu, 0.0)
    return SceneSpec(name, step, params, np.asarray(x, dtype=np.float32), np.asarray(y, dtype=np.float32),
                     np.asarray(z, dtype=np.float32), zeros, zeros.copy(), zeros.copy(), r, im, box, steps, dt, seed)


def sweep_column() -> SceneSpec:
    """SphereSweepTest.column (SphereSweepTest.java:93-110): 20 spheres of r = 5 cm stacked touching."""
    n = 20
    box = (0.1, 2 * 0.05 * n + 0.5, 0.1)
    return _sweep_scene("sweep_column", n, box, [0.05] * n, [0.05 + 0.1 * k for k in range(n)], [0.05] * n, 0.05,
                        1.0, 180, 0)


def sweep_pile(side: int = 16) -> SceneSpec:
    """SphereSweepTest.pile (SphereSweepTest.java:124-143): side^3 spheres of r = 3 cm as a loose lattice; side 7 is
    its pile of 343, side 16 its large pile of 4,096."""
    n = side ** 3
    rnd = JavaRandom(7)
    x = np.empty(n, dtype=np.float32)
    y = np.empty(n, dtype=np.float32)
    z = np.empty(n, dtype=np.float32)
    k = 0
    for i in range(side):
        for j in range(side):
            for l in range(side):
                x[k] = np.float32(0.07 + 0.06 * i + 0.005 * rnd.nextDouble())
                y[k] = np.float32(0.3 + 0.07 * j)
                z[k] = np.float32(0.07 + 0.06 * l + 0.005 * rnd.nextDouble())
                k += 1
    wide = 0.08 + 0.06 * side
    return _sweep_scene(f"sweep_pile_{n}", n, (wide, 0.5 + 0.07 * side, wide), x, y, z, 0.03, 0.1, 240, 7)


@dataclass
class BoxSceneSpec:
    """A box-path workload: the BoxStep to make, the initial state (BoxSimulation.start's arguments), the params."""
    name: str
    step: BoxStep
    params: np.ndarray
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    u: np.ndarray
    v: np.ndarray
    w: np.ndarray
    spin: np.ndarray   # 3 x n angular velocity
    q: np.ndarray      # 4 x n
    half: np.ndarray   # 3 x n
    shape: np.ndarray
    im: np.ndarray
    box: tuple
    steps: int
    dt: float
    seed: int
    towers: int


def mixed_32k(seed: int = 3, boxes: int = 16_384, spheres: int = 16_384, towers: int = 32, tall: int = 20,
              box=(128.0, 72.0, 128.0), substeps: int = 8, sweeps: int = 20) -> BoxSceneSpec:
    rnd = JavaRandom(seed)
    loose = boxes + spheres
    n = loose + towers * tall
    half = np.zeros((3, n), dtype=np.float64)
    q = np.zeros((4, n), dtype=np.float64)
    q[3] = 1.0
    pos = np.zeros((3, n), dtype=np.float64)
    shape = np.zeros(n, dtype=np.float32)
    mass = np.zeros(n, dtype=np.float64)
    b = s = 0
    for k in range(loose):
        sphere = (k % 2 == 1 and s < spheres) or b >= boxes
        if sphere:
            s += 1
            r = 0.2 + 0.4 * rnd.nextDouble()
            half[:, k] = r
            shape[k] = Boxes.SPHERE_SHAPE
            mass[k] = _mass(r)
        else:
            b += 1
            for a in range(3):
                half[a, k] = 0.2 + 0.8 * rnd.nextDouble()
            g = [rnd.nextGaussian() for _ in range(4)]
            length = math.sqrt(sum(v * v for v in g))
            q[:, k] = [v / length for v in g]
            mass[k] = 1000.0 * 8.0 * half[0, k] * half[1, k] * half[2, k]
        pos[0, k] = 1.0 + (box[0] - 2.0) * rnd.nextDouble()
        pos[1, k] = 22.0 + 48.0 * rnd.nextDouble()
        pos[2, k] = 1.0 + (box[2] - 2.0) * rnd.nextDouble()
    for t in range(towers):
        cx, cz = 8.0 + 16.0 * (t % 8), 8.0 + 16.0 * (t // 8)
        for level in range(tall):
            k = loose + t * tall + level
            half[:, k] = 0.5
            pos[:, k] = (cx, 0.5 + level, cz)
            mass[k] = 1000.0
    widest = float(np.max(np.sqrt((half ** 2).sum(axis=0))))
    margin = 0.005
    step = BoxStep(n, substeps, sweeps, False, BoxBroad.grid(*box, widest, margin))
    dt = 1.0 / 60
    params = Boxes.params(dt / substeps, 0, -9.81, 0, *box, 0.6, 0.6, margin)
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)
    zeros = np.zeros(n, dtype=np.float32)
    return BoxSceneSpec("mixed_32k", step, params, f32(pos[0]), f32(pos[1]), f32(pos[2]), zeros, zeros.copy(),
                        zeros.copy(), np.zeros((3, n), dtype=np.float32), f32(q), f32(half), shape,
                        f32(1.0 / mass), box, 600, dt, seed, towers)


def box_pile(side: int = 7) -> BoxSceneSpec:
    """BoxStackTest.pile (BoxStackTest.java:184-232): a loose block of side^3 cubes, jittered and turned."""
    n = side ** 3
    rnd = JavaRandom(7)
    pos = np.empty((3, n), dtype=np.float32)
    q = np.zeros((4, n), dtype=np.float32)
    k = 0
    for i in range(side):
        for j in range(side):
            for l in range(side):
                pos[0, k] = np.float32(0.25 + 0.1 * i + 0.02 * rnd.nextDouble())
                pos[1, k] = np.float32(0.3 + 0.1 * j)
                pos[2, k] = np.float32(0.25 + 0.1 * l + 0.02 * rnd.nextDouble())
                axis = [rnd.nextDouble() - 0.5, rnd.nextDouble() - 0.5, rnd.nextDouble() - 0.5]
                length = math.sqrt(axis[0] * axis[0] + axis[1] * axis[1] + axis[2] * axis[2])
                angle = 0.6 * rnd.nextDouble()
                sin = math.sin(angle / 2) / length
                for a in range(3):
                    q[a, k] = np.float32(axis[a] * sin)
                q[3, k] = np.float32(math.cos(angle / 2))
                k += 1
    half = np.full((3, n), np.float32(0.04), dtype=np.float32)
    h = half.astype(np.float64)
    widest = float(np.max(np.sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2])))
    margin = 0.005
    box = (0.1 * side + 0.4, 0.1 * side + 1.5, 0.1 * side + 0.4)
    substeps, dt = 10, 1.0 / 60
    step = BoxStep(n, substeps, Boxes.WALL_PASSES, False, BoxBroad.grid(*box, widest, margin))
    params = Boxes.params(dt / substeps, 0, -9.81, 0, *box, 0.5, 0.5, margin, 0.0)
    zeros = np.zeros(n, dtype=np.float32)
    return BoxSceneSpec(f"box_pile_{n}", step, params, pos[0].copy(), pos[1].copy(), pos[2].copy(), zeros,
                        zeros.copy(), zeros.copy(), np.zeros((3, n), dtype=np.float32), q, half,
                        np.zeros(n, dtype=np.float32), np.full(n, np.float32(1 / 0.5), dtype=np.float32), box, 240,
                        dt, 7, 0)


@dataclass
class BatchSpec:
    """A batched box-path workload: per scene its body count, walled box and grid; the state over all bodies."""
    name: str
    bodies: np.ndarray       # per scene
    extents: np.ndarray      # scenes x 3
    grids: list              # per scene, a SphereGrid
    substeps: int
    sweeps: int
    params: np.ndarray
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    u: np.ndarray
    v: np.ndarray
    w: np.ndarray
    spin: np.ndarray
    q: np.ndarray
    half: np.ndarray
    shape: np.ndarray
    im: np.ndarray
    steps: int
    dt: float
    seed: int


def batched_2k(seed: int = 4, scenes: int = 2048, fewest: int = 16, most: int = 4096, substeps: int = 8,
               sweeps: int = 10) -> BatchSpec:
    rng = np.random.default_rng(seed)
    counts = np.rint(np.exp(rng.uniform(math.log(fewest), math.log(most), scenes))).astype(np.int64)
    n = int(counts.sum())
    sphere = rng.random(n) < 0.5
    half = np.where(sphere, rng.uniform(0.2, 0.6, n), 0.0)[None, :].repeat(3, axis=0)
    boxes = ~sphere
    half[:, boxes] = rng.uniform(0.2, 1.0, (3, int(boxes.sum())))
    q = np.zeros((4, n))
    q[3] = 1.0
    g = rng.standard_normal((4, int(boxes.sum())))
    q[:, boxes] = g / np.sqrt((g * g).sum(axis=0))
    mass = np.where(sphere, 1000.0 * 4.0 / 3.0 * math.pi * half[0] ** 3, 1000.0 * 8.0 * half[0] * half[1] * half[2])
    pos = np.zeros((3, n))
    extents = np.zeros((scenes, 3))
    grids = []
    margin = 0.005
    first = 0
    for k in range(scenes):
        m = int(counts[k])
        side = 2.0 * m ** (1.0 / 3.0)
        per = int(math.ceil(m ** (1.0 / 3.0) - 1e-9))
        spacing = side / per
        cell = rng.permutation(per ** 3)[:m]
        at = np.stack([cell % per, (cell // per) % per, cell // (per * per)]).astype(np.float64)
        jitter = rng.uniform(-0.1, 0.1, (3, m)) * spacing
        pos[:, first:first + m] = (at + 0.5) * spacing + jitter
        extents[k] = side
        widest = float(np.max(np.sqrt((half[:, first:first + m] ** 2).sum(axis=0))))
        grids.append(BoxBroad.grid(side, side, side, widest, margin))
        first += m
    dt = 1.0 / 60
    params = Boxes.params(dt / substeps, 0, -9.81, 0, 1.0, 1.0, 1.0, 0.6, 0.6, margin)
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)
    zeros = np.zeros(n, dtype=np.float32)
    return BatchSpec("batched_2k", counts.astype(np.int32), extents.astype(np.float32), grids, substeps, sweeps,
                     params, f32(pos[0]), f32(pos[1]), f32(pos[2]), zeros, zeros.copy(), zeros.copy(),
                     np.zeros((3, n), dtype=np.float32), f32(q), f32(half),
                     np.where(sphere, Boxes.SPHERE_SHAPE, Boxes.BOX_SHAPE).astype(np.float32), f32(1.0 / mass),
                     300, dt, seed)


SCENES = {"pile_1k": pile_1k, "pile_64k": pile_64k, "mixed_32k": mixed_32k, "batched_2k": batched_2k}
