"""Islands solved on their own (axm_world_desc::island_solve) on the GPU against the CPU oracle,
through the C ABI.  An island of at most 4096 manifolds is swept by itself (one lane up to 8
manifolds, else one workgroup) and, with a residual tolerance, stops after the first sweep that
leaves its own residual <= tol; the large islands sweep together as before.

Bars: with a tolerance, every step runs the oracle's sweep count (the most any island ran) and
ends within 1e-4 relative of the oracle's velocities (the usual step bar); without a tolerance
the split changes nothing, so the GPU's results with and without it are bit-identical; the
per-colour launch path, graph replay and eager steps agree bit for bit."""
import os

import numpy as np
import pytest

from helpers import seed, rel_err

pytestmark = pytest.mark.gpu

DT = 1.0 / 120.0
THREADS = max(1, min(16, os.cpu_count() or 1))


def _state(w, bodies):
    g = w.get_bodies()
    s = bodies.copy()
    s.pos, s.quat, s.lin_vel, s.ang_vel = g.pos, g.quat, g.lin_vel, g.ang_vel
    return s


def _arrays(w):
    g = w.get_bodies()
    return g.pos.copy(), g.quat.copy(), g.lin_vel.copy(), g.ang_vel.copy()


def _settled(axm, config, n, steps, worlds=0):
    b, d = axm.scene(config, seed(config), n, worlds) if worlds else axm.scene(config, seed(config), n)
    w = axm.World(d)
    w.set_bodies(b)
    w.run(DT, 20, steps, graph=True)
    s = _state(w, b)
    w.close()
    return s, d


def _compare_tol(axm, orc, s, d, steps, tol, cap=300):
    d.residual_tol = tol
    d.max_iterations = cap
    d.residual_every = 1
    w = axm.World(d)
    w.set_bodies(s)
    ow = orc.OracleWorld(orc.params_from_desc(d))
    ow.set_bodies(s)
    ow.set_threads(THREADS)
    ow.set_tolerance(tol, cap, 1)
    sts = []
    for k in range(steps):
        st = w.step(DT, cap, axm.STEP_RESIDUAL)
        ost = ow.step(DT, cap)
        assert st.manifolds == ost.manifolds
        assert st.iterations == ost.iterations, (k, st.iterations, ost.iterations)
        assert st.residual <= max(tol, 1.0001 * ost.residual) or st.iterations == cap
        g = w.get_bodies()
        pos, quat, v, av = ow.get_state()
        assert rel_err(g.lin_vel, v) <= 1e-4 and rel_err(g.ang_vel, av) <= 1e-4, k
        assert rel_err(g.pos, pos) <= 1e-5
        sts.append(st)
    w.close()
    return sts


@pytest.mark.parametrize("config,n,warm", [(0, 1000, 120), (1, 6000, 150)])
def test_island_stops_match_oracle(axm, orc, gpu, config, n, warm):
    """A settling scene (small islands beside large ones) stepped at residual 1e-4."""
    s, d = _settled(axm, config, n, warm)
    d.island_solve = 1
    sts = _compare_tol(axm, orc, s, d, 4, 1e-4)
    assert max(st.manifolds for st in sts) > 50


def test_island_stops_batch_matches_oracle(axm, orc, gpu):
    """256 worlds of the batched configuration: each world's pile is an island solved by one
    workgroup, stopping at its own residual."""
    s, d = _settled(axm, 3, 0, 60, worlds=256)
    d.island_solve = 1
    sts = _compare_tol(axm, orc, s, d, 3, 1e-4)
    assert sts[-1].manifolds > 1000


def test_fixed_sweeps_are_unchanged_by_the_split(axm, gpu):
    """Without a tolerance every island runs every sweep, and since islands share no dynamic body
    the visits of each body come in the same order: bit-identical results with and without it,
    in the persistent and the per-colour launch paths."""
    s, d = _settled(axm, 0, 1000, 120)
    out = []
    for split in (0, 1):
        for mode in (0, 1):
            d.island_solve = split
            d.solver_mode = mode
            w = axm.World(d)
            w.set_bodies(s)
            for _ in range(5):
                st = w.step(DT, 12, axm.STEP_RESIDUAL)
            assert st.iterations == 12
            out.append((_arrays(w), st.residual))
            w.close()
    for arrs, res in out[1:]:
        assert res == out[0][1]
        for x, y in zip(arrs, out[0][0]):
            assert np.array_equal(x, y)


def test_island_stops_colour_launches_and_graphs_agree(axm, gpu):
    """With a tolerance: the per-colour launch path (host checks), the persistent launch and
    graph replay end bit for bit in the same state with the same sweep counts."""
    s, d = _settled(axm, 0, 1000, 120)
    d.island_solve = 1
    d.residual_tol = 1e-4
    d.max_iterations = 300
    d.residual_every = 1
    runs = []
    for mode in (0, 1):
        d.solver_mode = mode
        w = axm.World(d)
        w.set_bodies(s)
        its = [w.step(DT, 300).iterations for _ in range(6)]
        runs.append((its, _arrays(w)))
        w.close()
    assert runs[0][0] == runs[1][0]
    for x, y in zip(runs[0][1], runs[1][1]):
        assert np.array_equal(x, y)
    d.solver_mode = 0
    w = axm.World(d)
    w.set_bodies(s)
    st = w.run(DT, 300, 6, graph=True)
    assert st.iterations == runs[0][0][-1]
    for x, y in zip(_arrays(w), runs[0][1]):
        assert np.array_equal(x, y)
    t = w.totals()
    assert t.steps == 6 and t.steps_at_tol == 6 and t.max_sweeps == max(runs[0][0])
    w.close()
