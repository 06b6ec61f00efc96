"""World-level oracle checks: closed forms of the integrator and the solver, invariants of
the contact impulses, the broadphase against brute force and the colouring's validity.
All run on the CPU oracle (no GPU)."""
import math

import numpy as np
import pytest

from helpers import seed

DT = 1.0 / 120.0
G = np.float32(-9.81)


def world(orc, axm, bodies, **kw):
    d = axm.default_desc()
    for k, v in kw.items():
        if k == "planes":
            d.num_planes = len(v)
            for i, p in enumerate(v):
                for c in range(4):
                    d.planes[i][c] = p[c]
        elif k == "gravity":
            for c in range(3):
                d.gravity[c] = v[c]
        else:
            setattr(d, k, v)
    d.cell_size = 4.0
    for c in range(3):
        d.origin[c] = -2048.0
    ow = orc.OracleWorld(orc.params_from_desc(d))
    ow.set_bodies(bodies)
    return ow, d


def make(axm, specs):
    """specs: list of (shape_type, params, pos, vel, angvel, static)."""
    b = axm.Bodies(len(specs))
    for i, (t, prm, pos, vel, w, static) in enumerate(specs):
        b.shape_type[i] = t
        b.shape_params[i, :len(prm)] = prm
        b.pos[i] = pos
        b.lin_vel[i] = vel
        b.ang_vel[i] = w
        if not static:
            im, ii = axm.mass_properties(t, prm)
            b.inv_mass[i] = im
            b.inv_inertia[i] = ii
    return b


def test_free_fall_closed_form(orc, axm):
    """Symplectic Euler: v_n = v0 + n g dt, x_n = x0 + n v0 dt + g dt^2 n (n + 1) / 2."""
    v0 = np.array([1.5, 4.0, -2.0])
    b = make(axm, [(0, [0.5], [0, 100, 0], v0, [0, 0, 0], False)])
    ow, _ = world(orc, axm, b, planes=[])
    n = 60
    for _ in range(n):
        ow.step(DT, 10)
    pos, quat, v, w = ow.get_state()
    g = np.array([0, -9.81, 0])
    dt = float(np.float32(DT))
    assert np.allclose(v[0], v0 + n * g * dt, rtol=1e-5, atol=1e-5)
    assert np.allclose(pos[0], np.array([0, 100, 0]) + n * v0 * dt + g * dt * dt * n * (n + 1) / 2, rtol=1e-5,
                       atol=1e-4)


def test_torque_free_rotation_closed_form(orc, axm):
    """q <- normalise(q + dt/2 (w,0) q) is a rotation by 2 atan(|w| dt / 2) about w per step."""
    w0 = np.array([0.3, -1.2, 2.0])
    b = make(axm, [(0, [0.5], [0, 50, 0], [0, 0, 0], w0, False)])
    ow, _ = world(orc, axm, b, planes=[], gravity=[0, 0, 0])
    n = 120
    for _ in range(n):
        ow.step(DT, 10)
    pos, quat, v, w = ow.get_state()
    ang = 2 * n * math.atan(np.linalg.norm(w0) * float(np.float32(DT)) / 2)
    ax = w0 / np.linalg.norm(w0)
    want = np.concatenate([ax * math.sin(ang / 2), [math.cos(ang / 2)]])
    assert abs(abs(np.dot(quat[0], want)) - 1) < 1e-5
    assert np.allclose(w[0], w0, rtol=1e-6)
    assert abs(np.linalg.norm(quat[0]) - 1) < 1e-6


@pytest.mark.parametrize("soft", [0.0, 0.1])
def test_resting_sphere_normal_impulse_is_weight(orc, axm, soft):
    """A sphere resting on the ground: once settled, lambda_n = m g dt and it stays put.  With rigid
    rows any depth inside the slop is a rest state; a soft row (compliance eps k) only holds still
    where the Baumgarte target equals eps k lambda_n = eps g dt, i.e. at depth slop + eps g dt^2 / beta."""
    r = 0.5
    b = make(axm, [(0, [r], [0, r, 0], [0, 0, 0], [0, 0, 0], False)])
    ow, d = world(orc, axm, b, softness=soft)
    for _ in range(300):
        ow.step(DT, 10)
    lam = ow.impulses()
    m = 1.0 / b.inv_mass[0]
    assert lam.shape[0] == 1
    assert abs(lam[0, 0, 0] - m * 9.81 * DT) / (m * 9.81 * DT) < 1e-4
    pos, quat, v, w = ow.get_state()
    assert np.abs(v[0]).max() < 1e-4
    depth = r - pos[0, 1]
    if soft == 0.0:
        assert -1e-6 <= depth <= d.slop + 1e-6
    else:
        want = d.slop + soft * 9.81 * DT * DT / d.baumgarte
        assert abs(depth - want) < 2e-6, (depth, want)


@pytest.mark.parametrize("soft", [0.0, 0.1])
def test_sliding_box_decelerates_at_mu_g(orc, axm, soft):
    """Friction saturates while sliding: each point's |lambda_t| reaches its bound mu B (B: the point's
    normal impulse of the previous step), so v drops by mu g dt per step.  The box starts at its resting
    depth (slop + eps g dt^2 / beta for soft rows), where the normal impulses sum to m g dt."""
    mu, v0 = 0.5, 4.0
    d0 = 0.005 + soft * 9.81 * DT * DT / 0.2
    b = make(axm, [(1, [0.5, 0.25, 0.5], [0, 0.25 - d0, 0], [v0, 0, 0], [0, 0, 0], False)])
    ow, _ = world(orc, axm, b, friction=mu, restitution=0.0, softness=soft)
    for _ in range(5):
        ow.step(DT, 20)
    _, _, v_a, _ = ow.get_state()
    prev_n = ow.impulses()[0][:, 0].sum()
    n = 30
    for _ in range(n):
        ow.step(DT, 20)
        lam = ow.impulses()[0]
        man = ow.manifolds()
        # friction impulse along the sliding direction (-x on the box), from the pyramid rows t1, t2
        fx = 0.0
        for k in range(man["npts"][0]):
            t1, t2 = orc.tangents(man["normal"][0, k])
            fx += lam[k, 1] * t1[0] + lam[k, 2] * t2[0]
        # the impulse on B (the plane) is +P, on the box -P; every point's friction is saturated
        # at its bound mu B, B = that point's normal impulse of the previous step
        fb = ow.friction_bounds()[0, :man["npts"][0]]
        assert abs(abs(fx) - fb.sum()) <= 1e-4 * fb.sum()
        assert abs(fb.sum() - mu * prev_n) <= 1e-5 * fb.sum()
        prev_n = lam[:, 0].sum()
    _, _, v_b, _ = ow.get_state()
    dv = v_a[0, 0] - v_b[0, 0]
    assert abs(dv - n * mu * 9.81 * DT) / (n * mu * 9.81 * DT) < 2e-3
    assert abs(v_b[0, 1]) < 1e-3


@pytest.mark.parametrize("first", ["box", "sphere"])
def test_momentum_conserved_without_static_bodies(orc, axm, first):
    """Two bodies collide with no gravity, planes or damping: total linear momentum holds.  With two
    spheres (constant world inertia: the integrator has no gyroscopic term) angular momentum about the
    origin holds too, up to the lever-arm offset of the two witness points (depth x friction)."""
    s0 = (1, [0.5, 0.4, 0.3]) if first == "box" else (0, [0.55])
    b = make(axm, [(s0[0], s0[1], [0, 0, 0], [2.0, 0.1, 0], [0.3, 0.2, -0.1], False),
                   (0, [0.45], [1.2, 0.3, 0.1], [-1.0, 0, 0.2], [0, 0, 0], False)])
    b.quat[0] = np.array([0.1, 0.2, 0.3, 0.9]) / np.linalg.norm([0.1, 0.2, 0.3, 0.9])
    ow, _ = world(orc, axm, b, planes=[], gravity=[0, 0, 0], restitution=0.3)

    def momenta():
        pos, quat, v, w = ow.get_state()
        m = 1.0 / b.inv_mass.astype(np.float64)
        p = (m[:, None] * v).sum(0)
        L = np.zeros(3)
        for i in range(2):
            x, y, z, ww = quat[i].astype(np.float64)
            R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - ww * z), 2 * (x * z + ww * y)],
                          [2 * (x * y + ww * z), 1 - 2 * (x * x + z * z), 2 * (y * z - ww * x)],
                          [2 * (x * z - ww * y), 2 * (y * z + ww * x), 1 - 2 * (x * x + y * y)]])
            Iw = R @ np.diag(1.0 / b.inv_inertia[i].astype(np.float64)) @ R.T
            L += m[i] * np.cross(pos[i], v[i]) + Iw @ w[i]
        return p, L

    p0, L0 = momenta()
    hit = 0
    for _ in range(40):
        st = ow.step(DT, 10)
        hit += st.manifolds
    p1, L1 = momenta()
    assert hit > 0
    scale = np.abs(p0).max()
    assert np.abs(p1 - p0).max() <= 1e-4 * scale
    if first == "sphere":
        assert np.abs(L1 - L0).max() <= 1e-3 * max(1.0, np.abs(L0).max())


@pytest.mark.parametrize("config,n,warm,sweeps,tol", [(1, 2000, 40, 20, None), (0, 1000, 150, 1000, 1e-2)])
def test_impulses_stay_in_the_cone(orc, axm, config, n, warm, sweeps, tol):
    """lambda_n >= 0 always, and each friction impulse lies inside its point's box |lambda_t| <= mu B
    (the bound fixed at the step's start)."""
    b, d = axm.scene(config, seed(config), n)
    ow = orc.OracleWorld(orc.params_from_desc(d))
    ow.set_bodies(b)
    for _ in range(warm):
        ow.step(DT, 20)
    ow.step(DT, sweeps)
    m = ow.manifolds()
    lam = ow.impulses()
    fb = ow.friction_bounds()
    assert lam.shape[0] > 100 and fb.shape[0] == lam.shape[0]
    for k in range(lam.shape[0]):
        np_ = m["npts"][k]
        ln = lam[k, :np_, 0]
        assert np.all(ln >= 0) and np.all(fb[k, :np_] >= 0)
        lim = fb[k, :np_]
        assert np.all(np.abs(lam[k, :np_, 1]) <= lim) and np.all(np.abs(lam[k, :np_, 2]) <= lim)


def test_quaternions_stay_unit(orc, axm):
    b, d = axm.scene(0, seed(0))
    ow = orc.OracleWorld(orc.params_from_desc(d))
    ow.set_bodies(b)
    for _ in range(60):
        ow.step(DT, 10)
    _, quat, _, _ = ow.get_state()
    assert np.abs(np.linalg.norm(quat.astype(np.float64), axis=1) - 1).max() < 1e-6


@pytest.mark.parametrize("config,n", [(0, 1000), (2, 5000), (1, 3000)])
def test_broadphase_equals_brute_force(orc, axm, config, n):
    b, d = axm.scene(config, seed(config), n)
    ow = orc.OracleWorld(orc.params_from_desc(d))
    ow.set_bodies(b)
    ow.broadphase()
    a1, b1 = ow.pairs()
    a2, b2 = ow.brute_pairs()
    got = set(zip(a1.tolist(), b1.tolist()))
    assert len(got) == len(a1)  # no duplicates
    assert np.all(a1 < 0x80000000)  # a plane is always the second member (PLANE_BIT | plane index)
    norm = lambda ps: {p if p[1] & 0x80000000 else (min(p), max(p)) for p in ps}
    assert norm(got) == norm(zip(a2.tolist(), b2.tolist()))
    assert len(got) > 100


def test_touching_aabbs_count_as_overlap(orc, axm):
    """AABB overlap is inclusive (include/axiom/math/aabb.hpp:132-135; tests/math/aabb_test.cpp:187-191)."""
    h = 0.5 - 0.02  # the margin widens each AABB by 0.02, so these boxes' AABBs touch exactly
    b = make(axm, [(1, [h, h, h], [0, 5, 0], [0, 0, 0], [0, 0, 0], False),
                   (1, [h, h, h], [1.0, 5, 0], [0, 0, 0], [0, 0, 0], False),
                   (1, [h, h, h], [2.0001, 5, 0], [0, 0, 0], [0, 0, 0], False)])
    ow, _ = world(orc, axm, b, planes=[])
    ow.broadphase()
    a, bb = ow.pairs()
    assert set(zip(a.tolist(), bb.tolist())) == {(0, 1)}


def test_keys_sorted_stably(orc, axm):
    b, d = axm.scene(0, seed(0))
    ow = orc.OracleWorld(orc.params_from_desc(d))
    ow.set_bodies(b)
    ow.broadphase()
    keys, order = ow.keys()
    sk = keys[order]
    assert np.all(np.diff(sk.astype(np.float64)) >= 0) or np.all(sk[:-1] <= sk[1:])
    assert np.all(sk[:-1] <= sk[1:])
    # stable: equal keys keep ascending body index
    same = sk[:-1] == sk[1:]
    assert np.all(order[:-1][same] < order[1:][same])
    assert sorted(order.tolist()) == list(range(b.n))


@pytest.mark.parametrize("config,n,steps", [(0, 1000, 120), (1, 3000, 40), (3, 0, 30)])
def test_colouring_is_valid(orc, axm, config, n, steps):
    """No colour holds two manifolds sharing a dynamic body; static partners (planes, static bodies)
    do not constrain.  The solve order: the manifolds of islands of more than 4096 manifolds (and
    of none) by colour, then manifold index; after them the smaller islands, in the order of their
    smallest body index, each by colour, then manifold index (islands restated here by a plain
    union-find over the manifolds between two dynamic bodies)."""
    b, d = axm.scene(config, seed(config), n, 8 if config == 3 else 0)
    ow = orc.OracleWorld(orc.params_from_desc(d))
    ow.set_bodies(b)
    for _ in range(steps):
        ow.step(DT, 10)
    m = ow.manifolds()
    col, order, K = ow.colours()
    M = len(col)
    assert M > 50
    assert K == col.max() + 1
    dyn = b.inv_mass > 0
    seen = {}
    for k in range(M):
        for body in (int(m["a"][k]), int(m["b"][k])):
            if body & 0x80000000 or not dyn[body]:
                continue
            key = (int(col[k]), body)
            assert key not in seen, (k, seen.get(key), body)
            seen[key] = k
    parent = list(range(len(dyn)))

    def find(x):
        while parent[x] != x:
            x = parent[x]
        return x

    def is_dyn(body):
        return not (body & 0x80000000) and dyn[body]

    for k in range(M):
        a_, b_ = int(m["a"][k]), int(m["b"][k])
        if is_dyn(a_) and is_dyn(b_):
            ra, rb = find(a_), find(b_)
            parent[max(ra, rb)] = min(ra, rb)
    root = []
    for k in range(M):
        a_, b_ = int(m["a"][k]), int(m["b"][k])
        root.append(find(a_) if is_dyn(a_) else (find(b_) if is_dyn(b_) else -1))
    size = {}
    for r in root:
        size[r] = size.get(r, 0) + 1
    local = [r >= 0 and size[r] <= 4096 for r in root]
    glob = sorted((int(col[k]), k) for k in range(M) if not local[k])
    loc = sorted((root[k], int(col[k]), k) for k in range(M) if local[k])
    want = [k for _, k in glob] + [k for _, _, k in loc]
    assert np.array_equal(order, np.array(want, order.dtype))
    # deterministic: a second oracle run gives the same colouring
    ow2 = orc.OracleWorld(orc.params_from_desc(d))
    ow2.set_bodies(b)
    for _ in range(steps):
        ow2.step(DT, 10)
    col2, order2, K2 = ow2.colours()
    assert K2 == K and np.array_equal(col, col2) and np.array_equal(order, order2)


def test_single_contact_residual_vanishes_after_one_sweep(orc, axm):
    """One sphere on the ground is one normal row (plus friction at rest): one Gauss-Seidel update
    solves it exactly, so the fixed-point residual after the first sweep is ~0; a pile's is not."""
    b = make(axm, [(0, [0.5], [0, 0.49, 0], [0.0, -1.0, 0], [0, 0, 0], False)])
    ow, d = world(orc, axm, b, record_residual=1)
    ow.step(DT, 3)
    r = ow.residuals()
    assert len(r) == 3 and r[0] < 1e-5
    b2, d2 = axm.scene(1, seed(1), 1000)
    d2.record_residual = 1
    ow2 = orc.OracleWorld(orc.params_from_desc(d2))
    ow2.set_bodies(b2)
    for _ in range(30):
        ow2.step(DT, 5)
    assert ow2.residuals()[0] > 1e-3


def test_warm_start_matches_features_across_frames(orc, axm):
    """A box resting on the ground keeps its four corner features: every point of the second frame
    finds its previous impulse (warm hits = 4) and starts from it."""
    b = make(axm, [(1, [0.5, 0.25, 0.5], [0, 0.245, 0], [0, 0, 0], [0, 0, 0], False)])
    ow, _ = world(orc, axm, b)
    st1 = ow.step(DT, 10)
    lam1 = ow.impulses()
    st2 = ow.step(DT, 10)
    assert st1.warm_hits == 0 and st2.warm_hits == 4
    assert lam1[0, :4, 0].min() > 0


@pytest.mark.parametrize("soft", [0.0, 0.1])
def test_restitution_closed_form(orc, axm, soft):
    """A sphere hitting the ground at v > rest_threshold leaves at e * v (single frictionless point).
    A soft row converges to v' = v + (target - v) / (1 + eps) with target = -e v."""
    e, v = 0.4, -3.0
    b = make(axm, [(0, [0.5], [0, 0.5, 0], [0, v, 0], [0, 0, 0], False)])
    ow, d = world(orc, axm, b, restitution=e, friction=0.0, gravity=[0, 0, 0], softness=soft)
    ow.step(DT, 10)
    _, _, vel, _ = ow.get_state()
    want = v + (-e * v - v) / (1.0 + soft)
    assert abs(vel[0, 1] - want) < 1e-5, (vel[0, 1], want)


@pytest.mark.parametrize("soft", [0.0, 0.1])
def test_baumgarte_bias_closed_form(orc, axm, soft):
    """A sphere resting at depth d > slop is pushed out at beta / dt * (d - slop), capped at max_bias
    (divided by 1 + eps for a soft row starting at rest)."""
    for depth in (0.02, 0.5):
        b = make(axm, [(0, [0.5], [0, 0.5 - depth, 0], [0, 0, 0], [0, 0, 0], False)])
        ow, d = world(orc, axm, b, gravity=[0, 0, 0], friction=0.0, softness=soft)
        ow.step(DT, 10)
        _, _, vel, _ = ow.get_state()
        want = min(d.baumgarte / np.float32(DT) * (depth - d.slop), d.max_bias) / (1.0 + soft)
        assert abs(vel[0, 1] - want) <= 1e-4 * max(1.0, want), (depth, vel[0, 1], want)
