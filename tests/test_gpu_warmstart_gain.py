"""What the trained warm-start network gains: a network trained on the GPU on contacts logged
from the plain solver (16 seeded 1k-body scenes, 240 steps, 60 sweeps) against the plain
previous-step warm start, on a held-out scene, in sweeps to reach residual <= 1e-3.

Everything here is deterministic (fixed-order gradient reduction, deterministic step), so the
measured numbers are reproducible; the bar is that the learned start needs fewer sweeps."""
import numpy as np
import pytest

from helpers import seed

pytestmark = pytest.mark.gpu


def test_trained_network_needs_fewer_sweeps(axm, gpu):
    import warmstart
    xs, ys = [], []
    for sd in range(1, 17):
        b, d = axm.scene(0, seed(0, 10 + sd), 1000)
        X, Y = warmstart.collect(b, d, 240, 60)
        xs.append(X)
        ys.append(Y)
    X, Y = np.concatenate(xs), np.concatenate(ys)
    assert len(X) > 500000
    net, losses = warmstart.train(X, Y, epochs=10, batch=2048, lr=1e-3)
    assert losses[-1] < losses[0]
    b, d = axm.scene(0, seed(0, 1), 1000)
    ip, rp = warmstart.iterations_to_tolerance(b, d, 120, 1e-3, 300, axm.WARM_CACHE, warmup=60)
    il, rl = warmstart.iterations_to_tolerance(b, d, 120, 1e-3, 300, axm.WARM_MLP, net=net, warmup=60)
    print("sweeps to 1e-3: plain mean %.2f (reached %.3f), learned mean %.2f (reached %.3f)" %
          (ip.mean(), rp.mean(), il.mean(), rl.mean()))
    assert il.mean() < ip.mean()
    assert rl.mean() >= rp.mean()
