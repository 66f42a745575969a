"""Seeded synthetic scenes for the benchmark workloads, drawn with java.util.Random (javarandom.py).

Each builder returns a SceneSpec: the SphereStep to make, the initial state, and the parameter words.

* pile_1k: 1,000 spheres of r = 0.5 m, density 1000 kg/m^3, centres uniform in a 20 x 40 x 20 m walled box, at
  rest; mu = 0.5, e = 0; dt = 1/60 s in 4 substeps, 8 Gauss-Seidel passes, 1 m cells.
* pile_64k: 65,536 spheres, r log-uniform in [0.1, 0.5] m, density 1000 kg/m^3, centres uniform in a
  100 x 50 x 100 m region of a walled box whose ceiling is at 60 m (out of reach, so the box acts as ground plus 4
  walls), velocity N(0, 1) m/s per axis; mu = 0.4, e = 0.1; dt = 1/60 s in 8 substeps, 10 passes, 1 m cells. More
  spheres than the reference's 31-bit contact priorities order, so the step uses 63-bit ones (wide pairs).

Draw order per sphere: radius (if drawn), x, y, z, then the three velocity components.
"""

import math
from dataclasses import dataclass

import numpy as np

from .javarandom import JavaRandom
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


SCENES = {"pile_1k": pile_1k, "pile_64k": pile_64k}
