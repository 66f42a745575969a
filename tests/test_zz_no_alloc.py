"""vr_world_diagnostics_static, the allocation-free diagnostics, against vr_world_diagnostics through the C ABI.

No counterpart in the reference. Two calls in a row on one GPU, the second on a smaller world, each compared word for
word with the allocating entry point on the same world; a world past the fixed scratch's limit is refused before
anything runs; and the free device memory is the same before and after each call but the first on the device
(whose look-up of the array's address has the runtime load the library's device code, the array with it).
"""

import ctypes

import numpy as np
import pytest

import scene as sc

pkg = sc.pkg
scenes = __import__(pkg.__name__ + ".scenes", fromlist=["scenes"])
boxes = __import__(pkg.__name__ + ".boxes", fromlist=["boxes"])
pytestmark = pytest.mark.gpu
LIMIT = 2_396_160  # (1 MiB / 8 / 14 - 2) * 256 bodies


def words(world, name):
    out = np.zeros(14, dtype=np.float64)
    rc = world.abi.fn(name)(world.handle, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 14)
    return rc, out


def free_bytes():
    hip = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def static_then_allocating(world, set_up):
    abi = world.abi
    before = free_bytes()
    rc, got = words(world, "world_diagnostics_static")
    assert rc == 0, abi.last_error()
    if not set_up:
        assert free_bytes() == before, "the call took device memory"
        rc, again = words(world, "world_diagnostics_static")  # the scratch keeps the last call's words: harmless
        assert rc == 0 and np.array_equal(again.view(np.int64), got.view(np.int64))
    rc, want = words(world, "world_diagnostics")
    assert rc == 0, abi.last_error()
    return got, want


def test_static_diagnostics_match_on_a_batch_then_a_smaller_pile():
    spec = scenes.batched_2k(seed=4, scenes=64, fewest=16, most=2048, substeps=2, sweeps=2)
    big = boxes.BoxBatch(spec.bodies, spec.extents, spec.grids, spec.substeps, spec.sweeps)
    big.start(spec.x, spec.y, spec.z, spec.u, spec.v, spec.w, spec.spin, spec.q, spec.half, spec.shape, spec.im,
              spec.params)
    for _ in range(2):
        big.stepAndWait()
    got, want = static_then_allocating(big._world, set_up=True)
    assert np.array_equal(got.view(np.int64), want.view(np.int64)), (got, want)
    got, want = static_then_allocating(big._world, set_up=False)
    assert np.array_equal(got.view(np.int64), want.view(np.int64)), (got, want)
    assert got[2] > 0, "the batch is moving"
    n_big = int(spec.bodies.sum())
    big.close()

    small = scenes.pile_1k()
    assert small.step.spheres < n_big
    sim = pkg.SphereSimulation(small.step)
    sim.start(small.x, small.y, small.z, small.u, small.v, small.w, small.r, small.im, small.params)
    for _ in range(3):
        sim.stepAndWait()
    got, want = static_then_allocating(sim._world, set_up=False)
    assert np.array_equal(got.view(np.int64), want.view(np.int64)), (got, want)
    assert got[2] > 0
    sim.close()


def test_static_diagnostics_refuse_a_world_past_the_limit():
    step = pkg.SphereStep(LIMIT + 1, 1, 1, None, pkg.Solve.JACOBI)
    sim = pkg.SphereSimulation(step)
    rc, _ = words(sim._world, "world_diagnostics_static")
    assert rc == sc._abi.VR_ERR_ARGUMENT
    assert "at most 2396160 bodies" in sim._world.abi.last_error()
    sim.close()
