"""The learned warm start's network: the f64 CPU network pinned by finite differences, then the HIP network's
inference (bf16 matrix cores) and gradient (f32) against it, and training that lowers the loss.

The network has no counterpart in the reference, so the f64 network (oracle/mlp_oracle.cpp) is its checker; its
backpropagation is checked here against central differences of its own loss, with the weights unrounded.
"""

import ctypes

import numpy as np
import pytest

import scene as sc

pkg = sc.pkg
mlp = __import__(pkg.__name__ + ".mlp", fromlist=["mlp"])
F, O, PARAMS = mlp.FEATURES, mlp.OUTPUTS, mlp.PARAMS


def data(n, seed=3):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((n, F)).astype(np.float32)
    t = np.stack([x[:, 0] * x[:, 1], np.sin(x[:, 2]), x[:, 3] - 0.5 * x[:, 4]], axis=1).astype(np.float32)
    return x, t


def oracle_lib():
    lib = sc.oracle_abi().lib
    lib.vro_mlp_gradient64.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float),
                                       ctypes.POINTER(ctypes.c_float), ctypes.c_int64,
                                       ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    lib.vro_mlp_loss64.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float),
                                   ctypes.POINTER(ctypes.c_float), ctypes.c_int64]
    lib.vro_mlp_loss64.restype = ctypes.c_double
    lib.vro_mlp_exact.argtypes = [ctypes.c_void_p, ctypes.c_int32]
    return lib


def f32p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def gradient64(net, x, t):
    lib = oracle_lib()
    g = np.empty(PARAMS, dtype=np.float64)
    loss = ctypes.c_double()
    assert lib.vro_mlp_gradient64(net._h, f32p(x), f32p(t), len(x), g.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                  ctypes.byref(loss)) == 0
    return g, loss.value


def test_f64_backpropagation_matches_central_differences():
    x, t = data(24)
    net = mlp.Mlp(seed=5, abi=sc.oracle_abi())
    lib = oracle_lib()
    lib.vro_mlp_exact(net._h, 1)
    g, _ = gradient64(net, x, t)
    p = net.params
    rng = np.random.default_rng(1)
    # Parameters from every layer's weights and biases.
    picks = np.concatenate([rng.integers(a, b, 4) for a, b in ((0, 1024), (1024, 1088), (1088, 5184), (5184, 5248),
                                                                (5248, 9344), (9344, 9408), (9408, 9600),
                                                                (9600, 9603))])
    for i in picks:
        eps = 1e-3 * max(1.0, abs(float(p[i])))
        up, down = p.copy(), p.copy()
        up[i] = np.float32(p[i] + eps)
        down[i] = np.float32(p[i] - eps)
        step = float(up[i]) - float(down[i])
        net.params = up
        lu = lib.vro_mlp_loss64(net._h, f32p(x), f32p(t), len(x))
        net.params = down
        ld = lib.vro_mlp_loss64(net._h, f32p(x), f32p(t), len(x))
        fd = (lu - ld) / step
        assert abs(fd - g[i]) <= 1e-4 * max(1e-3, abs(g[i])) + 1e-7, (i, fd, g[i])
    net.params = p
    net.close()


def test_f64_training_lowers_the_loss():
    x, t = data(256)
    net = mlp.Mlp(seed=2, abi=sc.oracle_abi())
    _, before = gradient64(net, x, t)
    net.train(x, t, epochs=20, batch=64, rate=3e-3, seed=1)
    _, after = gradient64(net, x, t)
    assert after < 0.5 * before
    net.close()


@pytest.mark.gpu
def test_inference_matches_the_f64_network():
    x, t = data(1000)
    gpu = mlp.Mlp(seed=7)
    ref = mlp.Mlp(seed=7, abi=sc.oracle_abi())
    assert np.array_equal(gpu.params, ref.params), "the same seed draws the same weights"
    yg, yr = gpu.forward(x), ref.forward(x)
    # bf16 operands (8 significant bits): a rounding flip of one activation moves an output by well under 2 %.
    assert np.all(np.abs(yg - yr) <= 2e-2 * np.maximum(1.0, np.abs(yr))), float(np.max(np.abs(yg - yr)))
    gpu.close()
    ref.close()


@pytest.mark.gpu
@pytest.mark.parametrize("n", [1, 31, 1000])
def test_gradient_matches_the_f64_gradient(n):
    x, t = data(n, seed=n)
    gpu = mlp.Mlp(seed=11)
    ref = mlp.Mlp(seed=11, abi=sc.oracle_abi())
    g, loss = gpu.gradient(x, t)
    want, want_loss = gradient64(ref, x, t)
    assert abs(loss - want_loss) <= 1e-4 * want_loss
    assert np.linalg.norm(g - want) <= 1e-4 * np.linalg.norm(want)
    gpu.close()
    ref.close()


@pytest.mark.gpu
def test_training_on_the_device_lowers_the_loss():
    x, t = data(4096)
    net = mlp.Mlp(seed=4)
    _, before = net.gradient(x, t)
    net.train(x, t, epochs=30, batch=256, rate=3e-3, seed=2)
    _, after = net.gradient(x, t)
    assert after < 0.2 * before, (before, after)
    net.close()
