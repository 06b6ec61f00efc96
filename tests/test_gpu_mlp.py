"""The learned warm start on the GPU against the CPU oracle: BF16 MFMA forward and backward,
Adam, the per-point network inputs of a step, and a step that starts from the network's
prediction.  The bf16 emulation of the oracle rounds where the kernels round; the kernels sum
in a different order (MFMA), so the bars below are stated per quantity."""
import numpy as np
import pytest

from helpers import rel_err, seed

pytestmark = pytest.mark.gpu

DT = 1.0 / 120.0


def rand_params(rng, scale=1.0):
    from test_cpu_oracle_mlp import init
    p = init(rng)
    return (p * scale).astype(np.float32)


def test_forward_matches_oracle(axm, orc, gpu):
    rng = np.random.default_rng(10)
    net = axm.Mlp(0)
    prm = rand_params(rng)
    net.params = prm
    assert np.array_equal(net.params, prm)
    X = rng.normal(size=(1000, 24)).astype(np.float32)  # not a multiple of the 64-row block
    y = net.forward(X)
    yo = orc.mlp_forward(prm, X, emu=True)
    scale = max(1.0, float(np.abs(yo).max()))
    # same rounding points; only the fp32 summation order differs (a rare bf16 tie flips a hidden unit)
    assert np.abs(y - yo).max() <= 1e-2 * scale
    assert np.abs(y - yo).mean() <= 1e-4 * scale
    # the bf16 path is not the fp32 one: the emulation matters
    y32 = orc.mlp_forward(prm, X, emu=False)
    assert np.abs(y - yo).mean() < np.abs(y32 - yo).mean() or np.abs(y32 - yo).mean() == 0


def test_gradient_matches_oracle(axm, orc, gpu):
    rng = np.random.default_rng(11)
    net = axm.Mlp(0)
    prm = rand_params(rng)
    net.params = prm
    X = rng.normal(size=(777, 24)).astype(np.float32)
    Y = rng.normal(size=(777, 3)).astype(np.float32)
    loss, g = net.loss_grad(X, Y)
    lo, go = orc.mlp_grad(prm, X, Y, emu=True, fp64=False)
    assert abs(loss - lo) <= 1e-4 * max(1.0, lo)
    go = go.astype(np.float64)
    err = np.linalg.norm(g - go) / np.linalg.norm(go)
    cos = g @ go / (np.linalg.norm(g) * np.linalg.norm(go))
    assert err <= 5e-3 and cos > 0.99999, (err, cos)
    # every parameter block is exercised: W1, b1, W2, b2, W3, b3
    offs = [0, 3072, 3200, 19584, 19712, 20096, 20099]
    for a, b in zip(offs[:-1], offs[1:]):
        blk = np.linalg.norm(g[a:b] - go[a:b]) / max(np.linalg.norm(go[a:b]), 1e-12)
        assert blk <= 2e-2, (a, blk)
    # and against the exact FP64 gradient of the unrounded network
    _, g64 = orc.mlp_grad(prm, X, Y, emu=False, fp64=True)
    assert g @ g64 / (np.linalg.norm(g) * np.linalg.norm(g64)) > 0.999


def test_adam_matches_oracle(axm, orc, gpu):
    rng = np.random.default_rng(12)
    net = axm.Mlp(0)
    prm = rand_params(rng)
    net.params = prm
    p = prm.copy()
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    for t in (1, 2, 3):
        g = rng.normal(size=p.size).astype(np.float32)
        net.adam_step(g, lr=1e-3)
        orc.adam(p, m, v, g, t, lr=1e-3)
    assert np.allclose(net.params, p, rtol=0, atol=1e-6)


def test_training_reduces_loss(axm, gpu):
    """A teacher network's outputs are learnable: minibatch Adam on the GPU cuts the loss."""
    rng = np.random.default_rng(13)
    teacher = axm.Mlp(0)
    teacher.params = rand_params(rng, 0.5)
    X = rng.normal(size=(20000, 24)).astype(np.float32)
    Y = teacher.forward(X)
    net = axm.Mlp(0, seed=5)
    losses = net.train(X, Y, epochs=8, batch=1024, lr=2e-3)
    assert np.all(np.isfinite(losses))
    assert losses[-1] < 0.2 * losses[0], losses


def test_step_features_match_oracle(axm, orc, gpu):
    """The preparation kernel's per-point network inputs equal the oracle's."""
    b, d = axm.scene(0, seed(0), 1000)
    w = axm.World(d)
    w.set_bodies(b)
    w.set_feature_logging(True)
    ow = orc.OracleWorld(orc.params_from_desc(d))
    ow.set_bodies(b)
    for _ in range(90):
        w.step(DT, 10)
        ow.step(DT, 10)
    X, Y = w.features()
    Xo = ow.features()
    m = w.manifolds()
    assert np.array_equal(m["a"], ow.manifolds()["a"])
    valid = np.arange(4)[None, :] < m["npts"][:, None]
    assert valid.sum() > 40
    assert rel_err(X[valid], Xo[valid]) <= 1e-5
    assert X[valid][:, 21].sum() > 0  # previous-impulse hits are present
    # targets: (final - previous impulse) / u with u = mN g dt; previous / u is input 18..20
    lam = w.impulses()[1]
    final_u = Y[valid][:, 0] + X[valid][:, 18]
    assert np.all(np.isfinite(Y[valid])) and np.all(final_u >= -1e-4)
    # final / u: zero exactly where the final normal impulse is zero (the two terms cancel), positive
    # wherever it is not negligibly small
    ln = lam[valid][:, 0]
    assert np.all(final_u[ln == 0] == 0)
    assert np.all(final_u[ln > 1e-6 * np.abs(ln).max()] > 0)


def test_step_from_network_prediction(axm, orc, gpu):
    """With the network attached, the solver starts from its (projected) prediction and the step
    matches the oracle solving the same manifolds from the same warm start."""
    rng = np.random.default_rng(14)
    b, d = axm.scene(0, seed(0), 1000)
    d.warm_start = axm.WARM_MLP
    net = axm.Mlp(0)
    net.params = rand_params(rng, 0.3)
    w = axm.World(d)
    w.set_bodies(b)
    w.set_mlp(net)
    for _ in range(90):
        w.step(DT, 10)
    g0 = w.get_bodies()
    st = w.step(DT, 10, axm.STEP_KEEP_STAGES)
    X, Yt = w.features()
    m = w.manifolds()
    lam0, lam1 = w.impulses()
    valid = np.arange(4)[None, :] < m["npts"][:, None]
    # the warm start is previous + network output on those inputs, times u = mN g dt, projected onto
    # lambda_n >= 0 and the pyramid; u is recovered from the logged targets ((final - previous) / u)
    y = orc.mlp_forward(net.params, X[valid], emu=True)
    l0, l1, yt, xv = lam0[valid], lam1[valid], Yt[valid], X[valid]
    prev_u = xv[:, 18]  # previous normal impulse / u
    have_u = np.abs(yt[:, 0] + prev_u) > 1e-3
    assert have_u.sum() > 20
    u = l1[have_u, 0] / (yt[have_u, 0] + prev_u[have_u])
    assert np.all(u > 0)
    ln = np.maximum((prev_u[have_u] + y[have_u, 0]) * u, 0)
    assert np.allclose(l0[have_u, 0], ln, rtol=2e-2, atol=2e-2 * u)
    fb = w.friction_bounds()
    lim = fb[valid][have_u] * (1 + 1e-6) + 1e-7
    assert np.all(np.abs(l0[have_u, 1]) <= lim) and np.all(np.abs(l0[have_u, 2]) <= lim)
    # the same step on the oracle from the GPU's warm start
    ob = b.copy()
    ob.pos, ob.quat, ob.lin_vel, ob.ang_vel = g0.pos, g0.quat, g0.lin_vel, g0.ang_vel
    p = orc.params_from_desc(d)
    p.warm_start = 1
    ow = orc.OracleWorld(p)
    ow.set_bodies(ob)
    ow.step_given(DT, 10, m, warm=lam0, colours=w.colours()[0], bounds=fb)
    g1 = w.get_bodies()
    pos1, quat1, v1, w1 = ow.get_state()
    assert rel_err(g1.lin_vel, v1) <= 1e-4
    assert rel_err(g1.ang_vel, w1) <= 1e-4
    assert st.manifolds == len(m["a"])


def test_pile_network_step_matches_oracle(axm, orc, gpu):
    """A network trained on the GPU on contacts of the dense-pile configuration (6000 bodies,
    converged impulses as targets) and applied on a held-out pile of that configuration: the step
    starts from the projected prediction and matches the oracle solving the same manifolds from
    the same warm start."""
    import warmstart
    xs, ys = [], []
    for k in range(2):
        b, d = axm.scene(1, seed(1, 20 + k), 6000)
        X, Y = warmstart.collect(b, d, 160, 0, tol=1e-5)
        xs.append(X[len(X) // 3:])
        ys.append(Y[len(Y) // 3:])
    net, losses = warmstart.train(np.concatenate(xs), np.concatenate(ys), epochs=3, batch=2048, lr=1e-3, clip=64.0)
    assert losses[-1] < losses[0]
    b, d = axm.scene(1, seed(1, 30), 6000)
    d.warm_start = axm.WARM_MLP
    w = axm.World(d)
    w.set_bodies(b)
    w.set_mlp(net)
    for _ in range(150):
        w.step(DT, 20)
    g0 = w.get_bodies()
    st = w.step(DT, 20)
    m = w.manifolds()
    lam0, lam1 = w.impulses()
    fb = w.friction_bounds()
    assert st.manifolds > 2000
    ob = b.copy()
    ob.pos, ob.quat, ob.lin_vel, ob.ang_vel = g0.pos, g0.quat, g0.lin_vel, g0.ang_vel
    p = orc.params_from_desc(d)
    p.warm_start = 1
    ow = orc.OracleWorld(p)
    ow.set_bodies(ob)
    ow.step_given(DT, 20, m, warm=lam0, colours=w.colours()[0], bounds=fb)
    g1 = w.get_bodies()
    pos1, quat1, v1, w1 = ow.get_state()
    assert rel_err(g1.lin_vel, v1) <= 1e-4
    assert rel_err(g1.ang_vel, w1) <= 1e-4
    w.close()
    net.close()


def _pile_state(axm, steps=60):
    b, d = axm.scene(0, seed(0), 1000)
    w = axm.World(d)
    w.set_bodies(b)
    for _ in range(steps):
        w.step(DT, 10)
    g = w.get_bodies()
    s = b.copy()
    s.pos, s.quat, s.lin_vel, s.ang_vel = g.pos, g.quat, g.lin_vel, g.ang_vel
    w.close()
    return s, d


def _run(axm, state, d, warm, net, steps, tol, with_first=False):
    dd = axm.WorldDesc.from_buffer_copy(d)
    dd.warm_start = warm
    dd.residual_tol = tol
    dd.max_iterations = 1000
    w = axm.World(dd)
    w.set_bodies(state)
    if net is not None:
        w.set_mlp(net)
    its, first = [], []
    for _ in range(steps):
        st = w.step(DT, 1000, axm.STEP_RESIDUAL)
        its.append(st.iterations)
        first.append(st.residual_first)
    g = w.get_bodies()
    w.close()
    if with_first:
        return g, its, first
    return g, its


def test_nan_network_falls_back_to_the_cached_start(axm, gpu):
    """A network whose outputs are NaN everywhere leaves every point at its cached impulse: the steps
    are bit-identical to the plain previous-step warm start.  (Parameters must be finite; these
    overflow: hidden layer 1 at 1e38, layer 2 at +inf, and output weights of alternating sign
    sum +inf and -inf.)"""
    state, d = _pile_state(axm)
    net = axm.Mlp(0)
    p = np.zeros(axm.MLP_PARAMS, np.float32)
    p[3072:3200] = 1e38                    # b1
    p[3200:19584] = 1e10                   # W2
    w3 = np.ones((3, 128), np.float32)
    w3[:, 1::2] = -1.0
    p[19712:20096] = w3.ravel()            # W3
    net.params = p
    y = net.forward(np.zeros((4, axm.MLP_INPUTS), np.float32))
    assert np.all(np.isnan(y))
    g_nan, it_nan = _run(axm, state, d, axm.WARM_MLP, net, 20, 1e-4)
    g_ref, it_ref = _run(axm, state, d, axm.WARM_CACHE, None, 20, 1e-4)
    assert it_nan == it_ref
    assert np.array_equal(g_nan.lin_vel, g_ref.lin_vel) and np.array_equal(g_nan.pos, g_ref.pos)


def test_adversarial_network_ends_at_the_same_answer(axm, gpu):
    """A network whose outputs are huge (capped at the largest prediction used) starts the sweeps far
    from the solution; iterated to a tight residual, the step ends at the plain start's answer."""
    state, d = _pile_state(axm)
    net = axm.Mlp(0)
    rng = np.random.default_rng(3)
    net.params = (rng.normal(size=axm.MLP_PARAMS) * 50.0).astype(np.float32)
    tol = 2e-5
    dd = axm.WorldDesc.from_buffer_copy(d)
    dd.warm_start = axm.WARM_MLP
    dd.residual_tol = tol
    dd.max_iterations = 1000
    w = axm.World(dd)
    w.set_bodies(state)
    w.set_mlp(net)
    w.set_feature_logging(True)
    st_adv = w.step(DT, 1000, axm.STEP_RESIDUAL)
    X, _ = w.features()
    m = w.manifolds()
    g_adv = w.get_bodies()
    w.close()
    valid = np.arange(4)[None, :] < m["npts"][:, None]
    y = net.forward(np.ascontiguousarray(X[valid]))
    print("adversarial outputs: finite %.3f, max |y| %.3g" % (np.isfinite(y).mean(), np.abs(y[np.isfinite(y)]).max()))
    assert np.isfinite(y).mean() > 0.5 and np.abs(y[np.isfinite(y)]).max() > 10.0  # the premise
    g_ref, it_ref, r_ref = _run(axm, state, d, axm.WARM_CACHE, None, 1, tol, with_first=True)
    print("sweeps: adversarial %d plain %d; first residual %.3g vs %.3g"
          % (st_adv.iterations, it_ref[0], st_adv.residual_first, r_ref[0]))
    assert st_adv.residual_first > r_ref[0]  # it started further from the solution
    assert st_adv.residual <= tol
    assert rel_err(g_adv.lin_vel, g_ref.lin_vel) <= 1e-4
    assert rel_err(g_adv.ang_vel, g_ref.ang_vel) <= 1e-4
