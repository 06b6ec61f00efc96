"""Islands and sleeping on the GPU against the CPU oracle, through the C ABI: island labels,
asleep flags and rest times bit-exact every step, the island and sleeping counts of the step
statistics equal, and the trajectory within the usual bars while islands fall asleep and wake."""
import numpy as np
import pytest

from helpers import seed, rel_err
from test_gpu_parity import DT, _compare_step

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("config,n,steps,iters", [(0, 1000, 240, 10), (1, 6000, 200, 20)])
def test_islands_and_sleep_match_oracle(axm, orc, gpu, config, n, steps, iters):
    b, d = axm.scene(config, seed(config), n)
    d.allow_sleep = 1
    d.sleep_time = 0.25
    w = axm.World(d)
    w.set_bodies(b)
    ow = orc.OracleWorld(orc.params_from_desc(d))
    ow.set_bodies(b)
    max_sleep = 0
    for k in range(steps):
        st, ost = _compare_step(axm, w, ow, iters, check_state=False)
        gl, ga, gr = w.islands()
        ol, oa, orr = ow.islands()
        assert np.array_equal(gl, ol), k
        assert np.array_equal(ga, oa), k
        assert np.array_equal(gr, orr), k
        assert st.islands == ost.islands and st.sleeping == ost.sleeping, k
        max_sleep = max(max_sleep, st.sleeping)
        g = w.get_bodies()
        pos, quat, v, av = ow.get_state()
        assert rel_err(g.lin_vel, v) <= 1e-4, k
        if not (np.array_equal(g.pos, pos) and np.array_equal(g.lin_vel, v)):
            # continue both from the GPU state (keeps later steps single-step comparisons)
            nb = b.copy()
            nb.pos, nb.quat, nb.lin_vel, nb.ang_vel = g.pos, g.quat, g.lin_vel, g.ang_vel
            raise AssertionError("state differs at step %d" % k)
    assert max_sleep > 0
