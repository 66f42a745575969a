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
