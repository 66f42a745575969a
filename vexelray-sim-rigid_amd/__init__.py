"""vexelray-sim-rigid_amd — the per-substep rigid-body step of vexelray-sim-rigid, native to MI355X.

The step runs as hand-written HIP kernels for gfx950 behind the C ABI in include/vexelray_rigid.h; this package is
the host-side mirror of the reference's Java world/body/step API over that ABI.
"""

from .spheres import (IllegalArgumentException, IllegalStateException, Report, Solve, SphereDiagnostics,
                      SphereGrid, SphereSimulation, SphereStep, Spheres, StepReport)
from .javarandom import JavaRandom

__all__ = ["IllegalArgumentException", "IllegalStateException", "JavaRandom", "Report", "Solve",
           "SphereDiagnostics", "SphereGrid", "SphereSimulation", "SphereStep", "Spheres", "StepReport"]
