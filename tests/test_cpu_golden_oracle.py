"""The oracle reproduces its committed golden fixture (1k-body mixed scene, steps 1 and 10)
bit for bit: pair lists, manifolds, colouring and body state."""
import os

import numpy as np

from helpers import seed

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "config0_oracle.npz")
DT = 1.0 / 120.0


def test_oracle_reproduces_golden_fixture(orc, axm):
    g = np.load(GOLD)
    b, d = axm.scene(0, seed(0), 1000)
    ow = orc.OracleWorld(orc.params_from_desc(d))
    ow.set_bodies(b)
    for step in range(1, 11):
        ow.step(DT, 10)
        if step not in (1, 10):
            continue
        s = "s%d_" % step
        pa, pb = ow.pairs()
        m = ow.manifolds()
        col, order, K = ow.colours()
        pos, quat, v, w = ow.get_state()
        assert np.array_equal(pa, g[s + "pair_a"]) and np.array_equal(pb, g[s + "pair_b"])
        assert np.array_equal(m["a"], g[s + "man_a"]) and np.array_equal(m["b"], g[s + "man_b"])
        assert np.array_equal(m["npts"], g[s + "man_np"])
        assert K == int(g[s + "colours"][0])
        assert np.array_equal(col, g[s + "colour"]) and np.array_equal(order, g[s + "order"])
        for k, x in (("pos", pos), ("quat", quat), ("lin_vel", v), ("ang_vel", w)):
            assert np.array_equal(x, g[s + k]), k
    assert len(g["s10_man_a"]) > 50
