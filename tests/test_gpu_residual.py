"""Residual checks inside the persistent sweep launch: the per-sweep residual history and
the tolerance stop, against the per-colour launch path (which checks the residual with a
separate kernel and stops on the host) and against the oracle.

Bars: the two GPU paths agree bit for bit (same arithmetic on the same state); the
history agrees with the oracle's to 1e-4 relative (floor 1, velocity units), the bar
of the post-solve velocities it is computed from."""
import numpy as np
import pytest

from helpers import seed, rel_err

pytestmark = pytest.mark.gpu

DT = 1.0 / 120.0


def _settled(axm, n, steps):
    """A pile after `steps` steps (so the first step of a fresh world has many contacts)."""
    b, d = axm.scene(1, seed(1), n)
    w = axm.World(d)
    w.set_bodies(b)
    for _ in range(steps):
        w.step(DT, 20)
    g = w.get_bodies()
    w.close()
    nb = b.copy()
    nb.pos, nb.quat, nb.lin_vel, nb.ang_vel = g.pos, g.quat, g.lin_vel, g.ang_vel
    return nb, d


def _state(w):
    g = w.get_bodies()
    return g.pos.copy(), g.quat.copy(), g.lin_vel.copy(), g.ang_vel.copy()


def test_residual_history_matches_colour_launches_and_oracle(axm, orc, gpu):
    nb, d = _settled(axm, 3000, 60)
    d.record_residual = 1
    hist, states = [], []
    for mode in (0, 1):
        d.solver_mode = mode
        w = axm.World(d)
        w.set_bodies(nb)
        for _ in range(3):
            st = w.step(DT, 12)
        hist.append(np.asarray(w.residuals(), np.float32))
        states.append(_state(w))
        assert st.iterations == 12
        w.close()
    assert len(hist[0]) == 12 and np.array_equal(hist[0], hist[1])
    for x, y in zip(states[0], states[1]):
        assert np.array_equal(x, y)
    d.solver_mode = 0
    ow = orc.OracleWorld(orc.params_from_desc(d))
    ow.set_bodies(nb)
    for _ in range(3):
        ow.step(DT, 12)
    oh = np.asarray(ow.residuals(), np.float64)
    assert len(oh) == 12 and oh[0] > 1e-3
    assert rel_err(hist[0], oh) <= 1e-4


@pytest.mark.parametrize("every", [1, 4])
def test_tolerance_stop_matches_fixed_sweeps(axm, orc, gpu, every):
    """A step with a residual tolerance stops at the first check (every `every` sweeps) whose
    residual is <= tol, reports that residual, and ends where a step with that fixed sweep
    count ends; the per-colour launch path, which checks on the host, stops at the same sweep."""
    nb, d = _settled(axm, 3000, 60)
    d.record_residual = 1
    w = axm.World(d)
    w.set_bodies(nb)
    w.step(DT, 60)
    h = np.asarray(w.residuals(), np.float32)
    w.close()
    assert len(h) == 60
    tol = float(h[22])
    checks = [k for k in range(every, 61, every)] + ([60] if 60 % every else [])
    want = next((k for k in checks if h[k - 1] <= np.float32(tol)), 60)
    d.record_residual = 0
    d.residual_tol = tol
    d.residual_every = every
    d.max_iterations = 60
    d.island_solve = 0  # one residual over all manifolds, as the recorded history has it
    got = []
    for mode in (0, 1):
        d.solver_mode = mode
        w = axm.World(d)
        w.set_bodies(nb)
        st = w.step(DT, 1, axm.STEP_RESIDUAL)
        assert st.iterations == want, (mode, st.iterations, want)
        assert np.float32(st.residual) == h[want - 1]
        got.append(_state(w))
        w.close()
    d.residual_tol = 0.0
    d.solver_mode = 0
    w = axm.World(d)
    w.set_bodies(nb)
    w.step(DT, want)
    ref = _state(w)
    w.close()
    for s in got:
        for x, y in zip(s, ref):
            assert np.array_equal(x, y)


def test_tolerance_steps_replay_from_graphs(axm, orc, gpu):
    """Residual-stopped steps need no host round trip, so they replay from captured graphs: the
    replayed steps (plain and profiled) end bit for bit where eagerly launched ones end, and the
    sweeps they ran add up to the eager steps' counts."""
    nb, d = _settled(axm, 3000, 60)
    d.island_solve = 0  # every step's count is a multiple of the check interval
    d.residual_tol = 2e-3
    d.residual_every = 2
    d.max_iterations = 80
    w = axm.World(d)
    w.set_bodies(nb)
    counts = [w.step(DT, 1).iterations for _ in range(8)]
    ref = _state(w)
    w.close()
    assert 0 < min(counts) and max(counts) <= 80 and all(c % 2 == 0 for c in counts)
    for prof in (False, True):
        w = axm.World(d)
        w.set_bodies(nb)
        if prof:
            w.profile(True)
        st = w.run(DT, 1, 8, graph=True)
        assert st.iterations == counts[-1]
        if prof:
            p = w.profile(False)
            assert p.sweeps == sum(counts), (p.sweeps, counts)
        for x, y in zip(_state(w), ref):
            assert np.array_equal(x, y)
        w.close()
