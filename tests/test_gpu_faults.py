"""Failure handling through the C ABI (the reference's ErrorCode values,
include/axiom/core/error_code.hpp:35-42): GJK / EPA iteration caps reached -> 301 / 302 per pair,
exactly where the oracle with the same caps fails; bodies with NaN / Inf / negative mass or a
broken quaternion refused with InvalidRigidBody (400), leaving the world's bodies as they were;
and strict mode turning a step whose sweeps stop at max_iterations above the tolerance into
ConstraintSolverFailure (401) while keeping its results.  Every condition is forced from the
host (iteration caps, tolerances); no kernel is made to fault."""
import numpy as np
import pytest

from helpers import seed

pytestmark = pytest.mark.gpu

DT = 1.0 / 120.0


@pytest.mark.parametrize("gi,ei,want", [(1, 48, {301}), (2, 48, {301}), (32, 1, {302}), (32, 3, {302}),
                                        (6, 6, {301, 302})])
def test_capped_gjk_epa_status_matches_oracle(axm, orc, gpu, gi, ei, want):
    rng = np.random.default_rng(71 + gi * 7 + ei)
    n = 256
    q = rng.normal(size=(2, n, 4))
    q = (q / np.linalg.norm(q, axis=2, keepdims=True)).astype(np.float32)
    pa = rng.uniform(0.25, 0.75, (n, 3)).astype(np.float32)
    pb = rng.uniform(0.25, 0.75, (n, 3)).astype(np.float32)
    xa = np.zeros((n, 3), np.float32)
    xb = rng.uniform(-1.2, 1.2, (n, 3)).astype(np.float32)
    ta = np.ones(n, np.uint32)
    out, st = axm.gjk_epa_pairs(ta, pa, xa, q[0], ta, pb, xb, q[1], gjk_iters=gi, epa_iters=ei)
    seen = set()
    for i in range(n):
        o, s = orc.gjk_epa(1, pa[i], xa[i], q[0, i], 1, pb[i], xb[i], q[1, i], gjk_iters=gi, epa_iters=ei)
        assert st[i] == s, (i, st[i], s)
        seen.add(int(s))
    assert seen & want, seen
    # the defaults succeed on the same pairs
    _, st0 = axm.gjk_epa_pairs(ta, pa, xa, q[0], ta, pb, xb, q[1])
    assert set(np.unique(st0)) <= {0, 1}
    with pytest.raises(axm.AxmError) as e:
        axm.gjk_epa_pairs(ta[:4], pa[:4], xa[:4], q[0, :4], ta[:4], pb[:4], xb[:4], q[1, :4], gjk_iters=5000)
    assert e.value.code == axm.INVALID_PARAMETER


@pytest.mark.parametrize("field,index,value", [
    ("pos", (3, 1), np.nan), ("lin_vel", (5, 0), np.inf), ("ang_vel", (7, 2), -np.inf),
    ("quat", (2, 3), np.nan), ("quat", (4, slice(None)), 0.0), ("inv_mass", 9, -1.0),
    ("inv_inertia", (11, 0), np.nan)])
def test_invalid_body_is_refused(axm, gpu, field, index, value):
    b, d = axm.scene(0, seed(0), 200)
    w = axm.World(d)
    w.set_bodies(b)
    w.run(DT, 10, 5, graph=False)
    before = w.get_bodies()
    bad = b.copy()
    arr = getattr(bad, field).copy()
    arr[index] = value
    setattr(bad, field, arr)
    with pytest.raises(axm.AxmError) as e:
        w.set_bodies(bad)
    assert e.value.code == 400  # InvalidRigidBody
    after = w.get_bodies()
    assert np.array_equal(before.pos, after.pos) and np.array_equal(before.lin_vel, after.lin_vel)
    w.step(DT, 10)  # the world still steps
    w.close()


@pytest.mark.parametrize("graph", [False, True])
def test_strict_mode_reports_unconverged_steps(axm, gpu, graph):
    b, d = axm.scene(0, seed(0), 1000)
    w = axm.World(d)
    w.set_bodies(b)
    w.run(DT, 10, 60, graph=False)
    state = w.get_bodies()
    s = b.copy()
    s.pos, s.quat, s.lin_vel, s.ang_vel = state.pos, state.quat, state.lin_vel, state.ang_vel
    w.close()
    # a tolerance the solver meets: strict steps succeed
    d.strict = 1
    d.residual_tol = 1e-4
    d.max_iterations = 300
    w = axm.World(d)
    w.set_bodies(s)
    st = w.run(DT, 300, 4, graph=graph)
    assert st.residual <= 1e-4
    # one it cannot meet in 3 sweeps: ConstraintSolverFailure, results kept
    d.residual_tol = 1e-9
    d.max_iterations = 3
    w.configure(d)
    before = w.get_bodies()
    with pytest.raises(axm.AxmError) as e:
        w.run(DT, 3, 2, graph=graph)
    assert e.value.code == 401
    after = w.get_bodies()
    assert not np.array_equal(before.pos, after.pos)
    # the same without strict: success, reported in the statistics only
    d.strict = 0
    w.configure(d)
    st = w.step(DT, 3)
    assert st.iterations == 3 and st.residual > 1e-9
    w.close()
