"""The C ABI: both libraries load and export every entry point include/vexelray_rigid.h declares (no GPU needed)."""

import ctypes
import os
import re

import scene as sc

HEADER = os.path.join(sc.ROOT, "include", "vexelray_rigid.h")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\bint\s+(vr_\w+)\s*\(", text)))


def test_header_declares_the_abi():
    names = declared()
    assert {"vr_" + e for e in sc._abi.ENTRY_POINTS} <= set(names)


def test_product_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(sc._abi.PRODUCT_LIB)
    for name in declared():
        assert hasattr(lib, name), name


def test_oracle_exports_the_same_abi():
    lib = ctypes.CDLL(sc.ORACLE_LIB)
    for name in declared():
        assert hasattr(lib, "vro_" + name[3:]), name


def test_product_refuses_bad_arguments_without_a_device():
    abi = sc._abi.product()
    d = sc.pkg.SphereStep(10, 1).desc()
    d.bodies = 0
    handle = ctypes.c_void_p()
    rc = abi.fn("world_create")(1, ctypes.byref(d), ctypes.byref(handle))
    assert rc == sc._abi.VR_ERR_ARGUMENT
    assert "spheres, substeps and iterations" in abi.last_error()


def start_both_ways(backend):
    """A world started by vr_world_start, and one by the calls it stands for (SphereSimulation.java:154-178)."""
    import numpy as np
    s = sc.pile()
    s.muStatic = s.muKinetic = 0.3
    params = sc.pkg.Spheres.params(s.dt / s.substeps, s.gx, s.gy, s.gz, s.sx, s.sy, s.sz, s.omega, s.averaged,
                                   s.restitution, s.muStatic, s.muKinetic, s.muRolling)
    out = []
    for standard in (True, False):
        grid = sc.pkg.SphereGrid.covering(s.sx, s.sy, s.sz, s.widest())
        sim = sc.pkg.SphereSimulation(sc.pkg.SphereStep(s.n, s.substeps, s.iterations, grid, s.solve),
                                      abi=sc.abi_for(backend))
        fields = (s.x, s.y, s.z, s.u, s.v, s.w, s.r, s.im)
        if standard:
            sim.start(*fields, params)
        else:
            sim.clear()
            for name, values in zip(("x", "y", "z", "u", "v", "w", "r", "im"), fields):
                sim.write(name, values)
            sim.params(params)
            shown = np.zeros((s.n, sc.pkg.Spheres.SHOWN_STRIDE), dtype=np.float32)
            for k, arr in enumerate((s.x, s.y, s.z)):
                shown[:, k] = shown[:, 4 + k] = arr
            shown[:, 3] = s.r
            shown[:, sc.pkg.Spheres.TURNED_BEFORE + 3] = shown[:, sc.pkg.Spheres.TURNED_AFTER + 3] = 1
            sim.write("shown", shown.reshape(-1))
        sim.stepAndWait()
        out.append({k: sim.floats(k).copy() for k in sc.STATE + ("shown",)})
        sim.close()
    return out


def test_start_is_the_calls_it_stands_for():
    import numpy as np
    a, b = start_both_ways("oracle")
    for k in a:
        assert np.array_equal(a[k].view(np.int32), b[k].view(np.int32)), k


def test_start_refuses_short_fields_and_box_worlds():
    import numpy as np
    import pytest
    from boxscene import BoxScene
    s = sc.pile().start("oracle")
    with pytest.raises(sc.pkg.IllegalArgumentException):
        s.sim._world.start([np.zeros(3, np.float32)] * 8, np.zeros(13, np.int32))
    s.close()
    b = BoxScene(4)
    b.half[:, :] = 0.05
    b.start("oracle")
    with pytest.raises(sc.pkg.IllegalStateException):
        b.sim._world.start([np.zeros(4, np.float32)] * 8, np.zeros(18, np.int32))
    b.close()


import pytest  # noqa: E402


@pytest.mark.gpu
def test_hip_start_is_the_calls_it_stands_for():
    import numpy as np
    a, b = start_both_ways("gpu")
    for k in a:
        assert np.array_equal(a[k].view(np.int32), b[k].view(np.int32)), k
