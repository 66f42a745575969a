"""ctypes binding of the C ABI declared in include/vexelray_rigid.h.

``product()`` loads the HIP library built beside this package (``libvexelray_rigid.so``). There is no fallback: if
the library is missing or cannot load, this raises.
"""

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PRODUCT_LIB = os.path.join(HERE, "libvexelray_rigid.so")

VR_OK = 0
VR_ERR_ARGUMENT = -1
VR_ERR_STATE = -2
VR_ERR_DEVICE = -3
VR_ERR_NAME = -4
VR_SPHERES = 1
VR_BOXES = 2
VR_F32 = 1
VR_I32 = 2

# Every entry point the header declares, for the export check.
ENTRY_POINTS = ("world_create", "world_clear", "world_write", "world_read", "world_buffer_info", "world_params",
                "world_start", "world_friction", "world_slice", "world_step", "world_finish", "world_destroy", "last_error", "world_profile",
                "world_profile_read", "box_collide", "batch_create", "mlp_create", "mlp_destroy", "mlp_read",
                "mlp_write", "mlp_forward", "mlp_gradient", "mlp_train", "mlp_last_error", "box_log",
                "box_samples", "box_learned", "world_diagnostics", "box_collide_list",
                "world_diagnostics_static", "box_tolerance", "box_solve_read", "box_solve_curve", "box_retune",
                "box_plan")
# vr_box_plan's plans.
PLANS = ("pass", "all", "graph", "scene", "group")


class StepDesc(ctypes.Structure):
    _fields_ = [("bodies", ctypes.c_int32), ("substeps", ctypes.c_int32), ("iterations", ctypes.c_int32),
                ("has_grid", ctypes.c_int32), ("nx", ctypes.c_int32), ("ny", ctypes.c_int32),
                ("nz", ctypes.c_int32), ("cell", ctypes.c_float), ("solve", ctypes.c_int32),
                ("rounds", ctypes.c_int32), ("contacts_per_sphere", ctypes.c_int32),
                ("fixed_rounds", ctypes.c_int32), ("wide_pairs", ctypes.c_int32), ("sweeps", ctypes.c_int32),
                ("symmetric", ctypes.c_int32)]


class StepReport(ctypes.Structure):
    _fields_ = [("waits", ctypes.c_int32), ("rounds", ctypes.c_int32), ("most_rounds", ctypes.c_int32),
                ("unlisted", ctypes.c_int32), ("touching", ctypes.c_int32), ("points", ctypes.c_int32),
                ("unplaced", ctypes.c_int32), ("reserved", ctypes.c_int32), ("wall_ns", ctypes.c_int64),
                ("gpu_ns", ctypes.c_int64), ("submissions", ctypes.c_int32), ("unused", ctypes.c_int32)]


SOLVE_SUBSTEPS = 64
SOLVE_SWEEPS = 64


class SolveReport(ctypes.Structure):
    """vr_box_solve: what the last step's phases did (index 0 the solve sweeps, 1 the relax)."""
    _fields_ = [("tolerance", ctypes.c_float), ("measured", ctypes.c_int32), ("substeps", ctypes.c_int32),
                ("sweeps", ctypes.c_int32), ("scenes", ctypes.c_int32), ("scene_step", ctypes.c_int32),
                ("fell_back", ctypes.c_int32), ("unused", ctypes.c_int32),
                ("launched", (ctypes.c_int32 * SOLVE_SUBSTEPS) * 2),
                ("scene_sweeps", (ctypes.c_float * SOLVE_SUBSTEPS) * 2),
                ("residual", (ctypes.c_float * SOLVE_SUBSTEPS) * 2),
                ("residual_mean", (ctypes.c_float * SOLVE_SUBSTEPS) * 2)]


class Profile(ctypes.Structure):
    _fields_ = [("round_ns", ctypes.c_int64), ("round_launches", ctypes.c_int64), ("passes", ctypes.c_int64),
                ("listed_sum", ctypes.c_int64), ("steps", ctypes.c_int64), ("enqueue_ns", ctypes.c_int64),
                ("walls_ns", ctypes.c_int64), ("walls_launches", ctypes.c_int64), ("points_sum", ctypes.c_int64),
                ("bodies_sum", ctypes.c_int64), ("sweep_ns", ctypes.c_int64), ("sweep_launches", ctypes.c_int64)]


class Abi:
    """One library exposing the ABI, its symbols named ``prefix + entry``."""

    def __init__(self, path: str, prefix: str, where: str):
        if not os.path.exists(path):
            raise OSError(f"{path} is not built; run __graft_entry__.build() first")
        self.path = path
        self.lib = ctypes.CDLL(path)
        self.prefix = prefix
        self._where = where
        p = ctypes.c_void_p
        sig = {
            "world_create": ([ctypes.c_int32, ctypes.POINTER(StepDesc), ctypes.POINTER(p)]),
            "world_clear": ([p]),
            "world_write": ([p, ctypes.c_char_p, p, ctypes.c_int64]),
            "world_read": ([p, ctypes.c_char_p, p, ctypes.c_int64]),
            "world_buffer_info": ([p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int64),
                                   ctypes.POINTER(ctypes.c_int32)]),
            "world_params": ([p, ctypes.POINTER(ctypes.c_int32), ctypes.c_int32]),
            "world_friction": ([p, ctypes.c_int32]),
            "world_slice": ([p, ctypes.c_int32]),
            "world_start": ([p] + [p] * 8 + [p, ctypes.c_int32]),
            "world_step": ([p, ctypes.c_int32, ctypes.POINTER(StepReport)]),
            "world_finish": ([p]),
            "world_destroy": ([p]),
            "last_error": ([ctypes.c_char_p, ctypes.c_int32]),
            "world_profile": ([p, ctypes.c_int32]),
            "world_profile_read": ([p, ctypes.POINTER(Profile)]),
            "box_collide": ([p, ctypes.POINTER(ctypes.c_int32), ctypes.c_int32, ctypes.POINTER(ctypes.c_float)]),
            "box_collide_list": ([p, ctypes.POINTER(StepReport)]),
            "box_tolerance": ([p, ctypes.c_float, ctypes.c_int32]),
            "box_retune": ([p]),
            "box_plan": ([p, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]),
            "box_solve_read": ([p, ctypes.POINTER(SolveReport)]),
            "box_solve_curve": ([p, ctypes.POINTER(ctypes.c_float), ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]),
            "mlp_create": ([ctypes.c_uint64, ctypes.POINTER(p)]),
            "mlp_destroy": ([p]),
            "mlp_read": ([p, ctypes.POINTER(ctypes.c_float)]),
            "mlp_write": ([p, ctypes.POINTER(ctypes.c_float)]),
            "mlp_forward": ([p, ctypes.POINTER(ctypes.c_float), ctypes.c_int64, ctypes.POINTER(ctypes.c_float)]),
            "mlp_gradient": ([p, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float), ctypes.c_int64,
                              ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]),
            "mlp_train": ([p, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float), ctypes.c_int64,
                           ctypes.c_int32, ctypes.c_int32, ctypes.c_float, ctypes.c_uint64,
                           ctypes.POINTER(ctypes.c_float)]),
            "mlp_last_error": ([ctypes.c_char_p, ctypes.c_int32]),
            "box_log": ([p, ctypes.c_int32, ctypes.c_int64]),
            "box_samples": ([p, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float), ctypes.c_int64,
                             ctypes.POINTER(ctypes.c_int64)]),
            "box_learned": ([p, p]),
            "world_diagnostics": ([p, ctypes.POINTER(ctypes.c_double), ctypes.c_int32]),
            "world_diagnostics_static": ([p, ctypes.POINTER(ctypes.c_double), ctypes.c_int32]),
            "batch_create": ([ctypes.POINTER(StepDesc), ctypes.c_int32, ctypes.POINTER(ctypes.c_int32),
                              ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_int32),
                              ctypes.POINTER(ctypes.c_float), ctypes.POINTER(p)]),
        }
        for name, args in sig.items():
            fn = getattr(self.lib, prefix + name)
            fn.argtypes = args
            fn.restype = ctypes.c_int

    def fn(self, name):
        return getattr(self.lib, self.prefix + name)

    def last_error(self) -> str:
        buf = ctypes.create_string_buffer(1024)
        self.fn("last_error")(buf, 1024)
        return buf.value.decode("utf-8", "replace")

    def where(self) -> str:
        return self._where

    def create(self, kind, desc, arg_error, state_error) -> "World":
        handle = ctypes.c_void_p()
        rc = self.fn("world_create")(kind, ctypes.byref(desc), ctypes.byref(handle))
        World.check(self, rc, arg_error, state_error)
        return World(self, handle, arg_error, state_error)


    def create_batch(self, desc, bodies, extents, grids, cells, arg_error, state_error) -> "World":
        bodies = np.ascontiguousarray(bodies, dtype=np.int32)
        extents = np.ascontiguousarray(extents, dtype=np.float32).reshape(-1)
        grids = np.ascontiguousarray(grids, dtype=np.int32).reshape(-1)
        cells = np.ascontiguousarray(cells, dtype=np.float32)
        i32, f32 = ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_float)
        handle = ctypes.c_void_p()
        rc = self.fn("batch_create")(ctypes.byref(desc), len(bodies), bodies.ctypes.data_as(i32),
                                     extents.ctypes.data_as(f32), grids.ctypes.data_as(i32),
                                     cells.ctypes.data_as(f32), ctypes.byref(handle))
        World.check(self, rc, arg_error, state_error)
        return World(self, handle, arg_error, state_error)


class World:
    def __init__(self, abi: Abi, handle, arg_error, state_error):
        self.abi = abi
        self.handle = handle
        self.arg_error = arg_error
        self.state_error = state_error

    @staticmethod
    def check(abi, rc, arg_error=ValueError, state_error=RuntimeError):
        if rc == VR_OK:
            return
        msg = abi.last_error()
        if rc == VR_ERR_ARGUMENT:
            raise arg_error(msg)
        if rc == VR_ERR_STATE:
            raise state_error(msg)
        if rc == VR_ERR_NAME:
            raise KeyError(msg)
        raise RuntimeError(f"device error: {msg}")

    def _check(self, rc):
        World.check(self.abi, rc, self.arg_error, self.state_error)

    def info(self, name: str):
        count = ctypes.c_int64()
        dtype = ctypes.c_int32()
        self._check(self.abi.fn("world_buffer_info")(self.handle, name.encode(), ctypes.byref(count),
                                                     ctypes.byref(dtype)))
        return count.value, dtype.value

    def write(self, name: str, values) -> None:
        count, dtype = self.info(name)
        arr = np.ascontiguousarray(values, dtype=np.float32 if dtype == VR_F32 else np.int32)
        if arr.dtype == np.float32 and dtype == VR_I32:
            arr = arr.view(np.int32)
        arr = np.ascontiguousarray(arr.reshape(-1))
        if arr.size > count:
            raise KeyError(f"buffer {name} holds {count} words, not {arr.size}")
        self._check(self.abi.fn("world_write")(self.handle, name.encode(), arr.ctypes.data_as(ctypes.c_void_p),
                                               arr.size))

    def read(self, name: str, as_dtype) -> np.ndarray:
        count, dtype = self.info(name)
        arr = np.empty(count, dtype=np.float32 if dtype == VR_F32 else np.int32)
        self._check(self.abi.fn("world_read")(self.handle, name.encode(), arr.ctypes.data_as(ctypes.c_void_p),
                                              count))
        return arr.view(as_dtype) if arr.dtype != as_dtype else arr

    def clear(self) -> None:
        self._check(self.abi.fn("world_clear")(self.handle))

    def params(self, words: np.ndarray) -> None:
        words = np.ascontiguousarray(words, dtype=np.int32)
        self._check(self.abi.fn("world_params")(self.handle, words.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                                words.size))

    def friction(self, on: bool) -> None:
        self._check(self.abi.fn("world_friction")(self.handle, 1 if on else 0))

    def start(self, fields, params) -> None:
        """vr_world_start: the eight fields x, y, z, u, v, w, r, im (each as long as the world's bodies) and the
        parameter words."""
        count, _ = self.info("x")
        arrs = [np.ascontiguousarray(f, dtype=np.float32) for f in fields]
        if len(arrs) != 8 or any(a.shape != (count,) for a in arrs):
            raise self.arg_error(f"a start takes 8 fields of {count} floats")
        words = np.ascontiguousarray(params, dtype=np.int32)
        self._check(self.abi.fn("world_start")(self.handle, *[a.ctypes.data_as(ctypes.c_void_p) for a in arrs],
                                               words.ctypes.data_as(ctypes.c_void_p), int(words.size)))

    def slice(self, dispatches: int) -> None:
        self._check(self.abi.fn("world_slice")(self.handle, int(dispatches)))

    def step(self, wait: bool) -> StepReport:
        rep = StepReport()
        self._check(self.abi.fn("world_step")(self.handle, 1 if wait else 0, ctypes.byref(rep)))
        return rep

    def finish(self) -> None:
        self._check(self.abi.fn("world_finish")(self.handle))

    def diagnostics(self) -> np.ndarray:
        out = np.zeros(14, dtype=np.float64)
        self._check(self.abi.fn("world_diagnostics")(self.handle, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                                     14))
        return out

    def profile(self, on: bool) -> None:
        self._check(self.abi.fn("world_profile")(self.handle, 1 if on else 0))

    def profile_read(self) -> Profile:
        out = Profile()
        self._check(self.abi.fn("world_profile_read")(self.handle, ctypes.byref(out)))
        return out

    def collide(self, pairs) -> np.ndarray:
        pairs = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1)
        out = np.zeros((pairs.size // 2) * 40, dtype=np.float32)
        self._check(self.abi.fn("box_collide")(self.handle, pairs.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                               pairs.size // 2, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return out

    def collide_list(self) -> StepReport:
        rep = StepReport()
        self._check(self.abi.fn("box_collide_list")(self.handle, ctypes.byref(rep)))
        return rep

    def tolerance(self, tol: float, measure: bool) -> None:
        self._check(self.abi.fn("box_tolerance")(self.handle, ctypes.c_float(tol), 1 if measure else 0))

    def retune(self) -> None:
        self._check(self.abi.fn("box_retune")(self.handle))

    def plan(self):
        """(the plan the next step takes, by name; whether the batch is still timing its plans)."""
        plan, tuning = ctypes.c_int32(), ctypes.c_int32()
        self._check(self.abi.fn("box_plan")(self.handle, ctypes.byref(plan), ctypes.byref(tuning)))
        return PLANS[plan.value], bool(tuning.value)

    def solve_read(self) -> SolveReport:
        out = SolveReport()
        self._check(self.abi.fn("box_solve_read")(self.handle, ctypes.byref(out)))
        return out

    def solve_curve(self) -> np.ndarray:
        out = np.zeros(2 * SOLVE_SUBSTEPS * SOLVE_SWEEPS, dtype=np.float32)
        count = ctypes.c_int32()
        self._check(self.abi.fn("box_solve_curve")(self.handle, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                                   out.size, ctypes.byref(count)))
        return out[:count.value].reshape(-1, 2, SOLVE_SWEEPS) if count.value == out.size else out[:count.value]

    def destroy(self) -> None:
        if self.handle:
            self.abi.fn("world_destroy")(self.handle)
            self.handle = None


_PRODUCT = None


def product() -> Abi:
    """The HIP library. Raises if it is not built."""
    global _PRODUCT
    if _PRODUCT is None:
        _PRODUCT = Abi(PRODUCT_LIB, "vr_", "MI355X (HIP, gfx950)")
    return _PRODUCT
