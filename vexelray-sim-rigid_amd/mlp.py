"""The learned warm start's network over the C ABI (vr_mlp_*): 16 features -> 64 -> 64 -> 64 -> 3, SiLU.

No counterpart in the reference. Inference is bf16 on the matrix cores with f32 accumulation; training is f32
(with the bf16-rounded weights), mean squared error, Adam. Arrays are numpy, row-major: n x 16 in, n x 3 out.
"""

import ctypes
from typing import Optional

import numpy as np

from . import _abi

FEATURES, OUTPUTS, PARAMS = 16, 3, 9603
_f32 = ctypes.POINTER(ctypes.c_float)


def _ptr(a):
    return a.ctypes.data_as(_f32)


class Mlp:
    def __init__(self, seed: int = 1, abi: Optional["_abi.Abi"] = None):
        self._abi = abi if abi is not None else _abi.product()
        handle = ctypes.c_void_p()
        self._check(self._abi.fn("mlp_create")(ctypes.c_uint64(seed), ctypes.byref(handle)))
        self._h = handle

    def _check(self, rc):
        if rc != _abi.VR_OK:
            buf = ctypes.create_string_buffer(1024)
            self._abi.fn("mlp_last_error")(buf, 1024)
            raise (ValueError if rc == _abi.VR_ERR_ARGUMENT else RuntimeError)(
                f"mlp: {buf.value.decode('utf-8', 'replace')} ({rc})")

    @property
    def params(self) -> np.ndarray:
        out = np.empty(PARAMS, dtype=np.float32)
        self._check(self._abi.fn("mlp_read")(self._h, _ptr(out)))
        return out

    @params.setter
    def params(self, values):
        v = np.ascontiguousarray(values, dtype=np.float32)
        if v.shape != (PARAMS,):
            raise ValueError(f"{PARAMS} parameters, got {v.shape}")
        self._check(self._abi.fn("mlp_write")(self._h, _ptr(v)))

    def forward(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, FEATURES)
        y = np.empty((len(x), OUTPUTS), dtype=np.float32)
        self._check(self._abi.fn("mlp_forward")(self._h, _ptr(x), len(x), _ptr(y)))
        return y

    def gradient(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, FEATURES)
        t = np.ascontiguousarray(t, dtype=np.float32).reshape(-1, OUTPUTS)
        g = np.empty(PARAMS, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        self._check(self._abi.fn("mlp_gradient")(self._h, _ptr(x), _ptr(t), len(x), _ptr(g), _ptr(loss)))
        return g, float(loss[0])

    def train(self, x, t, epochs: int = 1, batch: int = 256, rate: float = 1e-3, seed: int = 1) -> float:
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, FEATURES)
        t = np.ascontiguousarray(t, dtype=np.float32).reshape(-1, OUTPUTS)
        loss = np.zeros(1, dtype=np.float32)
        self._check(self._abi.fn("mlp_train")(self._h, _ptr(x), _ptr(t), len(x), epochs, batch,
                                              ctypes.c_float(rate), ctypes.c_uint64(seed), _ptr(loss)))
        return float(loss[0])

    def close(self):
        if self._h:
            self._abi.fn("mlp_destroy")(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
