"""The warm-start MLP oracle (24 -> 128 -> 128 -> 3, ReLU, bf16 operands / fp32 accumulation)
pinned by an independent numpy forward, FP64 finite differences, bf16 rounding known answers
and the closed form of Adam (Kingma & Ba 2015)."""
import numpy as np
import pytest

IN, H, OUT = 24, 128, 3


def unpack(prm):
    p = np.asarray(prm, np.float64)
    o = 0
    W1 = p[o:o + H * IN].reshape(H, IN); o += H * IN
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    W3 = p[o:o + OUT * H].reshape(OUT, H); o += OUT * H
    b3 = p[o:o + OUT]; o += OUT
    assert o == p.size
    return W1, b1, W2, b2, W3, b3


def np_forward(prm, X):
    W1, b1, W2, b2, W3, b3 = unpack(prm)
    h1 = np.maximum(X @ W1.T + b1, 0)
    h2 = np.maximum(h1 @ W2.T + b2, 0)
    return h2 @ W3.T + b3


def init(rng):
    prm = np.zeros(IN * H + H + H * H + H + OUT * H + OUT, np.float32)
    W1, b1, W2, b2, W3, b3 = (rng.normal(0, s, n) for s, n in
                              ((1 / np.sqrt(IN), IN * H), (0.1, H), (1 / np.sqrt(H), H * H), (0.1, H),
                               (1 / np.sqrt(H), OUT * H), (0.1, OUT)))
    prm[:] = np.concatenate([W1, b1, W2, b2, W3, b3])
    return prm


def test_param_count(orc):
    assert orc.lib().orc_mlp_nparam() == 20099


def test_bf16_round_to_nearest_even(orc):
    u = np.array([0x3F800000, 0x3F808000, 0x3F818000, 0x3F80C000, 0x3F807FFF, 0xBF808001, 0x7F7FFFFF,
                  0x00000001, 0x7FC00000], np.uint32)
    want = np.array([0x3F800000, 0x3F800000, 0x3F820000, 0x3F810000, 0x3F800000, 0xBF810000, 0x7F800000,
                     0x00000000], np.uint32)
    got = orc.bf16_round(u.view(np.float32)).view(np.uint32)
    assert got[:8].tolist() == want.tolist()
    assert np.isnan(got.view(np.float32)[8])


def test_forward_matches_independent_numpy(orc):
    rng = np.random.default_rng(0)
    prm = init(rng)
    X = rng.normal(size=(64, IN)).astype(np.float32)
    y = orc.mlp_forward(prm, X, emu=False)
    assert np.allclose(y, np_forward(prm, X.astype(np.float64)), rtol=1e-5, atol=1e-5)
    # bf16 emulation stays within bf16 precision of the exact network
    ye = orc.mlp_forward(prm, X, emu=True)
    ref = np_forward(prm, X.astype(np.float64))
    assert np.abs(ye - ref).max() <= 0.03 * max(1.0, np.abs(ref).max())
    assert not np.array_equal(ye, y)


def test_gradient_matches_fp64_finite_differences(orc):
    rng = np.random.default_rng(1)
    prm = init(rng)
    X = rng.normal(size=(16, IN)).astype(np.float32)
    Y = rng.normal(size=(16, OUT)).astype(np.float32)
    loss, g = orc.mlp_grad(prm, X, Y, emu=False, fp64=True)
    assert abs(loss - orc.mlp_loss(prm, X, Y)) <= 1e-12 * max(1.0, loss)
    idx = np.concatenate([rng.choice(IN * H, 40, replace=False),                      # W1
                          IN * H + rng.choice(H, 10, replace=False),                  # b1
                          IN * H + H + rng.choice(H * H, 40, replace=False),          # W2
                          IN * H + H + H * H + rng.choice(H, 10, replace=False),      # b2
                          IN * H + 2 * H + H * H + rng.choice(OUT * H, 20, replace=False),  # W3
                          IN * H + 2 * H + H * H + OUT * H + np.arange(OUT)])         # b3
    # the loss is evaluated in double on float parameters: perturb by float-representable steps
    for i in idx:
        h = np.float32(2e-6) * max(np.float32(1.0), abs(prm[i]))
        p1, p2 = prm.copy(), prm.copy()
        p1[i] += h
        p2[i] -= h
        fd = (orc.mlp_loss(p1, X, Y) - orc.mlp_loss(p2, X, Y)) / (float(p1[i]) - float(p2[i]))
        assert abs(fd - g[i]) <= 2e-5 * max(1e-3, abs(g[i])) + 1e-7, (i, fd, g[i])


def test_bf16_gradient_tracks_fp64(orc):
    rng = np.random.default_rng(2)
    prm = init(rng)
    X = rng.normal(size=(256, IN)).astype(np.float32)
    Y = rng.normal(size=(256, OUT)).astype(np.float32)
    _, g64 = orc.mlp_grad(prm, X, Y, emu=False, fp64=True)
    _, g16 = orc.mlp_grad(prm, X, Y, emu=True, fp64=False)
    cos = g64 @ g16 / (np.linalg.norm(g64) * np.linalg.norm(g16))
    assert cos > 0.999
    _, g32 = orc.mlp_grad(prm, X, Y, emu=False, fp64=False)
    assert np.allclose(g32, g64, rtol=1e-3, atol=1e-6 * np.abs(g64).max())


def test_adam_closed_form(orc):
    rng = np.random.default_rng(3)
    n = 1000
    prm = rng.normal(size=n).astype(np.float32)
    m = np.zeros(n, np.float32)
    v = np.zeros(n, np.float32)
    p64, m64, v64 = prm.astype(np.float64), np.zeros(n), np.zeros(n)
    lr, b1, b2, eps = 1e-3, 0.9, 0.999, 1e-8
    g1 = rng.normal(size=n).astype(np.float32)
    orc.adam(prm, m, v, g1, 1, lr, b1, b2, eps)
    # step 1: m_hat = g, v_hat = g^2  =>  update = lr * g / (|g| + eps)
    assert np.allclose(prm, p64 - lr * g1 / (np.abs(g1.astype(np.float64)) + eps), rtol=0, atol=2e-7)
    p64 = p64 - lr * g1 / (np.abs(g1.astype(np.float64)) + eps)
    m64 = (1 - b1) * g1
    v64 = (1 - b2) * g1.astype(np.float64) ** 2
    g2 = rng.normal(size=n).astype(np.float32)
    orc.adam(prm, m, v, g2, 2, lr, b1, b2, eps)
    m64 = b1 * m64 + (1 - b1) * g2
    v64 = b2 * v64 + (1 - b2) * g2.astype(np.float64) ** 2
    p64 = p64 - lr * (m64 / (1 - b1 ** 2)) / (np.sqrt(v64 / (1 - b2 ** 2)) + eps)
    assert np.allclose(prm, p64, rtol=0, atol=5e-7)
    # float32 b2 = 0.999f makes (1 - b2) carry a 5e-5 relative offset
    assert np.allclose(m, m64, rtol=1e-5) and np.allclose(v, v64, rtol=1e-4)
