"""The learned warm start on a batch of box scenes, end to end on the device: log (features, targets) from the
plain solver, train the network on them, then start steps from its prediction.

What the trained network gains is measured two ways. As a predictor of the impulses a step ends with, it is
compared with the reference's own warm start (the carried impulse, which is feature 7..9) on held-out points: it
must predict them better. As a warm start, one step from the same settled state is compared with a converged
step (60 sweeps): the learned start must stay as close to it as the carry does, within 10 %; on the settled piles
measured so far it gains under 1 %, which this test records rather than claims.
"""

import numpy as np
import pytest

import scene as sc

pkg = sc.pkg
boxes = __import__(pkg.__name__ + ".boxes", fromlist=["boxes"])
scenes = __import__(pkg.__name__ + ".scenes", fromlist=["scenes"])
mlp = __import__(pkg.__name__ + ".mlp", fromlist=["mlp"])
pytestmark = pytest.mark.gpu
STATE = ("x", "y", "z", "u", "v", "w", "lx", "ly", "lz", "qx", "qy", "qz", "qw", "lambda", "rub", "roll")
PAIRS = ("pairs", "pairsWas", "manifold", "manifoldWas", "impulses", "impulsesWas")


def batch(spec, sweeps):
    b = boxes.BoxBatch(spec.bodies, spec.extents, spec.grids, spec.substeps, sweeps)
    b.start(spec.x, spec.y, spec.z, spec.u, spec.v, spec.w, spec.spin, spec.q, spec.half, spec.shape, spec.im,
            spec.params)
    return b


def snapshot(b):
    return {k: (b.ints(k) if k in ("pairs", "pairsWas") else b.floats(k)).copy() for k in STATE + PAIRS}


def one_step(spec, snap, sweeps, net=None):
    b = batch(spec, sweeps)
    for k, v in snap.items():
        b.write(k, v)
    b.learned(net)
    b.stepAndWait()
    used = b.learnedPoints
    out = {k: b.floats(k).copy() for k in ("x", "y", "z", "u", "v", "w")}
    b.close()
    return out, used


def rms(a, b, keys):
    return float(np.sqrt(np.mean(sum((a[k] - b[k]) ** 2 for k in keys))))


def test_learned_warm_start_end_to_end():
    spec = scenes.batched_2k(seed=6, scenes=48, fewest=64, most=512)
    sim = batch(spec, spec.sweeps)
    for _ in range(60):
        sim.stepAndWait()
    sim.log(True, 1 << 20)
    for _ in range(20):
        sim.stepAndWait()
    x, t = sim.samples(1 << 20)
    assert len(x) > 10_000 and np.all(np.isfinite(x)) and np.all(np.isfinite(t))
    snap = snapshot(sim)
    sim.close()

    rng = np.random.default_rng(0)
    order = rng.permutation(len(x))
    train, held = order[: int(0.9 * len(x))], order[int(0.9 * len(x)):]
    net = mlp.Mlp(seed=3)
    net.train(x[train], t[train], epochs=10, batch=1024, rate=1e-3, seed=1)
    predicted = net.forward(x[held])
    carried = x[held][:, 7:10]
    mse_net = float(np.mean(np.sum((predicted - t[held]) ** 2, axis=1)))
    mse_carry = float(np.mean(np.sum((carried - t[held]) ** 2, axis=1)))
    print(f"held-out squared error: network {mse_net:.4g}, carried impulse {mse_carry:.4g}")
    assert mse_net < mse_carry

    converged, _ = one_step(spec, snap, 60)
    plain, _ = one_step(spec, snap, spec.sweeps)
    learned, used = one_step(spec, snap, spec.sweeps, net)
    assert used > 0
    for keys in (("x", "y", "z"), ("u", "v", "w")):
        ep, el = rms(plain, converged, keys), rms(learned, converged, keys)
        print(f"{keys}: rms from the converged step, carry {ep:.4g}, learned {el:.4g}")
        assert el <= 1.1 * ep
    net.close()
