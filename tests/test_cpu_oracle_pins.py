"""Closed-form pins of the oracle functions the metric and the learned warm start rest on: the
contact residual, the 24 per-point network inputs and the energy.  Each expected value is derived
here from the geometry and mechanics of a hand-built scene, not from the oracle's own formulas."""
import numpy as np
import pytest

from test_cpu_oracle_world import DT, make, world

G = 9.81


def sphere_at_rest(orc, axm, soft, mu=0.5):
    """A sphere resting on the ground after settling; returns (oracle world, desc, bodies, manifolds, lambda)."""
    r = 0.5
    b = make(axm, [(0, [r], [0, r, 0], [0, 0, 0], [0, 0, 0], False)])
    ow, d = world(orc, axm, b, softness=soft, friction=mu)
    for _ in range(300):
        ow.step(DT, 10)
    return ow, d, b


@pytest.mark.parametrize("soft", [0.0, 0.1])
@pytest.mark.parametrize("dn,dt1,dt2", [(0.3, 0.0, 0.0), (0.0, 0.2, 0.0), (0.0, 0.0, -0.25), (0.1, -0.3, 0.2)])
def test_residual_of_perturbed_impulse(orc, axm, soft, dn, dt1, dt2):
    """A sphere of mass m resting on a plane, solved exactly: adding (dn, dt1, dt2) to its impulse gives
    a residual of max(|dn| kN, |dt1| kT, |dt2| kT) with kN = (1 + eps) / m for the normal row (the lever
    arm is parallel to n) and kT = (1 + eps) (1/m + R^2 / I) = 3.5 (1 + eps) / m for either friction row
    (I = 2/5 m R^2); the rows do not couple (r x t1 and r x t2 are orthogonal, I is isotropic)."""
    ow, d, b = sphere_at_rest(orc, axm, soft)
    pos, quat, v, w = ow.get_state()
    man = ow.manifolds()
    lam = ow.impulses()
    assert man["npts"][0] == 1
    m = 1.0 / b.inv_mass[0]
    # the exact fixed point of the settled contact: the solve's own residual is ~0
    st = ow.step(DT, 10)
    assert st.residual < 1e-5
    state = b.copy()
    state.pos, state.quat, state.lin_vel, state.ang_vel = ow.get_state()
    man = ow.manifolds()
    lam = ow.impulses().copy()
    col, _, _ = ow.colours()
    # from the state before the next step: solve that step exactly, then perturb and re-measure
    ow2 = orc.OracleWorld(orc.params_from_desc(d))
    ow2.set_bodies(state)
    ow2.step_given(DT, 200, man, warm=lam.reshape(-1, 12), colours=col)
    lam_star = ow2.impulses().copy()
    ow3 = orc.OracleWorld(orc.params_from_desc(d))
    ow3.set_bodies(state)
    warm = lam_star.copy()
    warm[0, 0] += [dn, dt1, dt2]
    st3 = ow3.step_given(DT, 0, man, warm=warm.reshape(-1, 12), colours=col)
    kn = (1.0 + soft) / m
    kt = 3.5 * (1.0 + soft) / m
    want = max(abs(dn) * kn, abs(dt1) * kt, abs(dt2) * kt)
    assert abs(st3.residual - want) <= 1e-3 * want + 1e-7, (st3.residual, want)


def test_network_inputs_of_a_box_on_the_ground(orc, axm):
    """The 24 inputs of the four corners of a box lying flat on the ground (no warm start yet), moving
    with v = (vx, vy, vz) and spinning about the vertical at wy, derived from the box's geometry: the
    normal points from the box (A) into the ground (B), n = (0,-1,0), with tangents t1 = (0,0,1) and
    t2 = n x t1 = (-1,0,0); depth; the ground-relative velocity of the corner r = (+-hx, -hy, +-hz) is
    -(v + w x r) in that frame; lever arms; inverse-mass ratios from I = m/3 (b^2 + c^2)."""
    hx, hy, hz = 0.5, 0.25, 0.4
    depth0 = 0.003
    vx, vy, vz, wy = 0.7, -0.2, 0.3, 0.9
    soft = 0.1
    b = make(axm, [(1, [hx, hy, hz], [0, hy - depth0, 0], [vx, vy, vz], [0, wy, 0], False)])
    ow, d = world(orc, axm, b, softness=soft, restitution=0.0)
    ow.step(DT, 1)
    X = ow.features()
    man = ow.manifolds()
    assert man["npts"][0] == 4
    m = 1.0 / b.inv_mass[0]
    Ixx, Iyy, Izz = m / 3 * (hy * hy + hz * hz), m / 3 * (hx * hx + hz * hz), m / 3 * (hx * hx + hy * hy)
    vy_pre = vy - G * DT
    gdt = G * DT
    for k in range(4):
        p = man["pos"][0, k].astype(np.float64)
        r = p - np.array([0.0, hy - depth0, 0.0])
        assert abs(abs(r[0]) - hx) < 1e-5 and abs(r[1] + hy) < 1e-5 and abs(abs(r[2]) - hz) < 1e-5
        f = X[0, k].astype(np.float64)
        n, t1, t2 = np.array([0, -1, 0.0]), np.array([0, 0, 1.0]), np.array([-1.0, 0, 0])
        vc = -(np.array([vx, vy_pre, vz]) + np.cross([0, wy, 0], r))
        # effective masses (soft): k = (1/m + sum (r x d)_i^2 / I_i) (1 + eps)
        Iinv = np.array([1 / Ixx, 1 / Iyy, 1 / Izz])
        kd = [(1 / m + (np.cross(r, dd) ** 2 * Iinv).sum()) * (1 + soft) for dd in (n, t1, t2)]
        depth = depth0
        bias = 0.0 if depth <= d.slop else min(d.baumgarte / DT * (depth - d.slop), d.max_bias)
        want = [0, -1, 0, depth * 20, vc @ n / gdt, vc @ t1 / gdt, vc @ t2 / gdt, r @ n, r @ t1, r @ t2, 0, 0, 0,
                (1 / m) / kd[0], 0, kd[0] / kd[1], kd[0] / kd[2], bias / gdt, 0, 0, 0, 0, 1.0, 1.0]
        assert np.allclose(f, want, rtol=2e-5, atol=2e-5), (k, np.round(f, 5), np.round(want, 5))


def test_energy_of_a_free_body(orc, axm):
    """A free box (no planes, gravity -9.81 y) spinning about its body y axis, which is a principal
    axis that stays vertical: after one step KE = m |v + g dt|^2 / 2 + I_yy wy^2 / 2 and
    PE = m g y1 with y1 = y0 + (vy + g_y dt) dt (semi-implicit Euler)."""
    hx, hy, hz = 0.6, 0.3, 0.45
    v0 = np.array([1.0, 2.0, -3.0])
    wy, y0 = 0.8, 5.0
    b = make(axm, [(1, [hx, hy, hz], [0.0, y0, 0.0], list(v0), [0, wy, 0], False)])
    ow, d = world(orc, axm, b, planes=[])
    st = ow.step(DT, 10)
    m = 1.0 / float(b.inv_mass[0])
    Iyy = m / 3 * (hx * hx + hz * hz)
    v1 = v0 + np.array([0.0, -G * DT, 0.0])
    ke = 0.5 * m * (v1 @ v1) + 0.5 * Iyy * wy * wy
    pe = m * G * (y0 + v1[1] * DT)
    assert abs(st.kinetic - ke) <= 1e-5 * ke
    assert abs(st.potential - pe) <= 1e-5 * pe
