"""GPU parity against the CPU oracle at the full sizes of the benchmark configurations, through
the C ABI: the settled 100k-body pile, the 1M-body sparse scene (sparse broadphase build, residual-
stopped sweeps) and the 4096-world batch.  Same bars as test_gpu_parity: pair lists, manifold
pairs, feature ids and colourings bit-exact; contact geometry 1e-5 relative; velocities 1e-4
relative at the same sweep count and colour order.  The oracle runs on all host cores (its
results do not depend on the thread count)."""
import os

import numpy as np
import pytest

from helpers import seed, rel_err
from test_gpu_parity import DT, _compare_step

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 1))


def _oracle(orc, d, bodies):
    ow = orc.OracleWorld(orc.params_from_desc(d))
    ow.set_bodies(bodies)
    ow.set_threads(THREADS)
    return ow


def _state(w, bodies):
    g = w.get_bodies()
    s = bodies.copy()
    s.pos, s.quat, s.lin_vel, s.ang_vel = g.pos, g.quat, g.lin_vel, g.ang_vel
    return s


def test_pile_100k_settled_steps_match_oracle(axm, orc, gpu):
    """The dense pile at its full 100k bodies after 300 GPU steps: a fresh GPU world and the oracle
    take two steps from that state (the first colours every manifold, the second keeps colours and
    starts from the cached impulses)."""
    b, d = axm.scene(1, seed(1), 0)
    assert b.n == 100000
    w = axm.World(d)
    w.set_bodies(b)
    w.run(DT, 20, 300, graph=True)
    settled = _state(w, b)
    w.close()
    w2 = axm.World(d)
    w2.set_bodies(settled)
    ow = _oracle(orc, d, settled)
    for _ in range(2):
        st, ost = _compare_step(axm, w2, ow, 20)
    print("pile after 300 steps: %d pairs, %d manifolds, %d contacts, %d colours" % (
        st.pairs, st.manifolds, st.contacts, st.colours))
    assert st.contacts > 20000


def test_scene_1m_steps_match_oracle(axm, orc, gpu):
    """The 1M-body sparse scene at full size: two steps at its 10 sweeps from the initial state."""
    b, d = axm.scene(2, seed(2), 0)
    assert b.n == 1000000
    w = axm.World(d)
    w.set_bodies(b)
    ow = _oracle(orc, d, b)
    for _ in range(2):
        st, ost = _compare_step(axm, w, ow, 10)
    assert st.pairs > 10000 and st.manifolds > 1000


def test_scene_1m_residual_stopped_matches_oracle(axm, orc, gpu):
    """The 1M-body scene with every step's sweeps stopped at residual <= 1e-4 (checked after every
    sweep inside the GPU's sweep launch; the oracle checks the same residual the same way): each step
    runs the same number of sweeps on both, reaches the tolerance, and ends in the same state."""
    b, d = axm.scene(2, seed(2), 0)
    d.residual_tol = 1e-4
    d.max_iterations = 300
    d.residual_every = 1
    w = axm.World(d)
    w.set_bodies(b)
    w.run(DT, 300, 20, graph=True)
    s0 = _state(w, b)
    w.close()
    w = axm.World(d)
    w.set_bodies(s0)
    ow = _oracle(orc, d, s0)
    ow.set_tolerance(1e-4, 300, 1)
    for k in range(3):
        st = w.step(DT, 300, axm.STEP_RESIDUAL)
        ost = ow.step(DT, 300)
        assert st.iterations == ost.iterations, (k, st.iterations, ost.iterations)
        assert st.residual <= 1e-4 and ost.residual <= 1e-4
        g = w.get_bodies()
        pos, quat, v, av = ow.get_state()
        assert rel_err(g.lin_vel, v) <= 1e-4 and rel_err(g.ang_vel, av) <= 1e-4
        assert rel_err(g.pos, pos) <= 1e-5


def test_batch_4096_worlds_match_oracle(axm, orc, gpu):
    """4096 independent worlds of 16-4096 bodies (about 3M bodies) stepped together: two steps match
    the oracle stepping the same concatenated worlds."""
    b, d = axm.scene(3, seed(3), 0, 4096)
    assert b.n > 2500000 and int(b.world.max()) == 4095
    w = axm.World(d)
    w.set_bodies(b)
    ow = _oracle(orc, d, b)
    for _ in range(2):
        st, ost = _compare_step(axm, w, ow, 10)
    assert st.manifolds > 1000
