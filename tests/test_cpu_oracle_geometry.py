"""Oracle narrowphase, GJK and EPA pinned by closed forms, brute force and SAT.

Conventions (the reference's debug contact, include/axiom/debug/physics_debug_draw.hpp:128-132):
normal points from A to B, depth > 0 means penetration, the contact position is
the witness point on A and the witness on B is position - normal * depth."""
import math

import numpy as np
import pytest
from scipy import optimize

SPHERE, BOX, CAPSULE, PLANE = 0, 1, 2, 3
I4 = np.array([0, 0, 0, 1], np.float32)


def rand_quat(rng):
    q = rng.normal(size=4)
    return (q / np.linalg.norm(q)).astype(np.float32)


def rot(q):
    x, y, z, w = np.asarray(q, np.float64)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def sat_depth(ha, xa, qa, hb, xb, qb):
    """Minimum overlap of two oriented boxes over the 15 separating axes, and its axis (A -> B)."""
    Ra, Rb = rot(qa), rot(qb)
    axes = [Ra[:, i] for i in range(3)] + [Rb[:, i] for i in range(3)]
    for i in range(3):
        for j in range(3):
            c = np.cross(Ra[:, i], Rb[:, j])
            if np.linalg.norm(c) > 1e-6:
                axes.append(c / np.linalg.norm(c))
    d = np.asarray(xb, np.float64) - np.asarray(xa, np.float64)
    best, best_axis = np.inf, None
    for L in axes:
        ra = sum(ha[i] * abs(Ra[:, i] @ L) for i in range(3))
        rb = sum(hb[i] * abs(Rb[:, i] @ L) for i in range(3))
        ov = ra + rb - abs(d @ L)
        if ov < best:
            best, best_axis = ov, (L if d @ L >= 0 else -L)
    return best, best_axis


def seg_dist(p0, p1, q0, q1):
    """Brute-force closest distance of two segments (convex in (s, t): grid, then bounded refinement)."""
    p0, p1, q0, q1 = (np.asarray(v, np.float64) for v in (p0, p1, q0, q1))
    f = lambda st: np.linalg.norm(p0 + st[0] * (p1 - p0) - q0 - st[1] * (q1 - q0))
    g = np.linspace(0, 1, 41)
    s0 = min(((s, t) for s in g for t in g), key=f)
    r = optimize.minimize(f, s0, bounds=[(0, 1), (0, 1)], method="L-BFGS-B", options={"ftol": 1e-15, "gtol": 1e-12})
    return min(f(r.x), f(s0))


def capsule_ends(x, q, h):
    ax = rot(q)[:, 1] * h
    return np.asarray(x, np.float64) - ax, np.asarray(x, np.float64) + ax


def check_conventions(c):
    for k in range(c["npts"]):
        assert abs(np.linalg.norm(c["normal"][k]) - 1) < 1e-5


def test_sphere_sphere_closed_form(orc):
    rng = np.random.default_rng(11)
    for _ in range(200):
        ra, rb = rng.uniform(0.25, 0.75, 2)
        xa = rng.uniform(-5, 5, 3).astype(np.float32)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        dist = rng.uniform(0.2, ra + rb)
        xb = (xa + d * dist).astype(np.float32)
        c = orc.collide(SPHERE, [ra, 0, 0], xa, I4, SPHERE, [rb, 0, 0], xb, I4)
        dd = np.asarray(xb, np.float64) - xa
        L = np.linalg.norm(dd)
        n = dd / L
        assert c["npts"] == 1 and c["status"] == 0
        assert abs(c["depth"][0] - (ra + rb - L)) <= 1e-5 * max(1.0, ra + rb)
        assert np.allclose(c["normal"][0], n, atol=1e-5)
        assert np.allclose(c["pos"][0], xa + n * ra, atol=1e-5 * max(1, np.abs(xa).max()))


def test_sphere_plane_and_margin(orc):
    c = orc.collide(SPHERE, [0.5, 0, 0], [1, 0.4, 2], I4, PLANE, [0, 1, 0], [0, 0, 0], None)
    assert c["npts"] == 1 and abs(c["depth"][0] - 0.1) < 1e-6
    assert np.allclose(c["normal"][0], [0, -1, 0]) and np.allclose(c["pos"][0], [1, -0.1, 2], atol=1e-6)
    # speculative contact inside the margin: negative depth; beyond the margin: none
    c = orc.collide(SPHERE, [0.5, 0, 0], [0, 0.51, 0], I4, PLANE, [0, 1, 0], [0, 0, 0], None, margin=0.02)
    assert c["npts"] == 1 and abs(c["depth"][0] + 0.01) < 1e-6
    c = orc.collide(SPHERE, [0.5, 0, 0], [0, 0.53, 0], I4, PLANE, [0, 1, 0], [0, 0, 0], None, margin=0.02)
    assert c["npts"] == 0


def test_box_resting_on_plane_four_corners(orc):
    h = np.array([1.0, 0.5, 0.75], np.float32)
    c = orc.collide(BOX, h, [3, 0.49, -2], I4, PLANE, [0, 1, 0], [0, 0, 0], None)
    assert c["npts"] == 4
    assert np.allclose(c["depth"], 0.01, atol=1e-6)
    corners = {(round(float(p[0]), 4), round(float(p[2]), 4)) for p in c["pos"]}
    assert corners == {(3 + sx * 1.0, -2 + sz * 0.75) for sx in (-1, 1) for sz in (-1, 1)}
    assert np.allclose(c["pos"][:, 1], -0.01, atol=1e-6)
    assert len(set(c["feature"].tolist())) == 4


def test_capsule_plane_two_points(orc):
    # lying capsule (axis along X after a 90 degree turn about Z)
    q = np.array([0, 0, math.sin(math.pi / 4), math.cos(math.pi / 4)], np.float32)
    c = orc.collide(CAPSULE, [0.3, 0.8, 0], [0, 0.29, 0], q, PLANE, [0, 1, 0], [0, 0, 0], None)
    assert c["npts"] == 2
    assert np.allclose(sorted(c["pos"][:, 0]), [-0.8, 0.8], atol=1e-5)
    assert np.allclose(c["depth"], 0.01, atol=1e-5)


def test_stacked_axis_aligned_boxes_exact_overlap(orc):
    """Depth equals the exact overlap; four points at the smaller box's bottom corners on A's top face."""
    c = orc.collide(BOX, [1, 0.5, 1], [0, 0, 0], I4, BOX, [0.4, 0.4, 0.4], [0.1, 0.88, 0], I4)
    assert c["npts"] == 4
    assert np.allclose(c["depth"], 0.02, atol=1e-6)
    assert np.allclose(c["normal"], [0, 1, 0], atol=1e-6)
    assert np.allclose(c["pos"][:, 1], 0.5, atol=1e-6)
    assert {(round(float(p[0]), 4), round(float(p[2]), 4)) for p in c["pos"]} == \
        {(x, z) for x in (-0.3, 0.5) for z in (-0.4, 0.4)}


def test_swap_flips_normal(orc):
    """Mixed-type pairs are computed in one canonical type order and flipped: swapping A and B
    gives the same contacts seen from the other body.  (Same-type pairs are not symmetric: the
    box-box reference face prefers A's faces on ties.)"""
    rng = np.random.default_rng(5)
    shapes = [(SPHERE, [0.5, 0, 0]), (BOX, [0.5, 0.3, 0.4]), (CAPSULE, [0.3, 0.5, 0])]
    n_checked = 0
    for _ in range(400):
        ia = int(rng.integers(3))
        ib = (ia + 1 + int(rng.integers(2))) % 3
        (ta, pa), (tb, pb) = shapes[ia], shapes[ib]
        xa = np.zeros(3, np.float32)
        xb = rng.uniform(-1.0, 1.0, 3).astype(np.float32)
        qa, qb = rand_quat(rng), rand_quat(rng)
        c1 = orc.collide(ta, pa, xa, qa, tb, pb, xb, qb)
        c2 = orc.collide(tb, pb, xb, qb, ta, pa, xa, qa)
        assert c1["npts"] == c2["npts"]
        if c1["npts"] == 0:
            continue
        n_checked += 1
        check_conventions(c1)
        # same contact set, seen from the other body: normals flip, positions move to the other witness
        k1 = sorted(range(c1["npts"]), key=lambda k: tuple(np.round(c1["pos"][k] - c1["normal"][k] * c1["depth"][k], 3)))
        k2 = sorted(range(c2["npts"]), key=lambda k: tuple(np.round(c2["pos"][k], 3)))
        for a, b in zip(k1, k2):
            assert np.allclose(c1["normal"][a], -c2["normal"][b], atol=1e-4)
            assert abs(c1["depth"][a] - c2["depth"][b]) < 1e-4
            assert np.allclose(c1["pos"][a] - c1["normal"][a] * c1["depth"][a], c2["pos"][b], atol=1e-4)
    assert n_checked > 100


def test_sphere_box_closest_point(orc):
    rng = np.random.default_rng(9)
    n = 0
    for _ in range(300):
        h = rng.uniform(0.25, 0.75, 3).astype(np.float32)
        q = rand_quat(rng)
        r = float(rng.uniform(0.25, 0.75))
        xs = rng.uniform(-1.5, 1.5, 3).astype(np.float32)
        R = rot(q)
        local = R.T @ xs.astype(np.float64)
        inside = np.all(np.abs(local) < h)
        if inside:
            continue
        cl = np.clip(local, -h, h)
        dist = np.linalg.norm(local - cl)
        c = orc.collide(SPHERE, [r, 0, 0], xs, I4, BOX, h, [0, 0, 0], q)
        if r - dist < -0.02:
            assert c["npts"] == 0
            continue
        n += 1
        assert c["npts"] == 1
        assert abs(c["depth"][0] - (r - dist)) <= 1e-5
        nrm = R @ (cl - local) / dist
        assert np.allclose(c["normal"][0], nrm, atol=1e-4)
    assert n > 50


def test_capsule_capsule_brute_force(orc):
    rng = np.random.default_rng(21)
    n = 0
    for _ in range(150):
        ra, rb = rng.uniform(0.2, 0.5, 2)
        ha, hb = rng.uniform(0.2, 1.0, 2)
        qa, qb = rand_quat(rng), rand_quat(rng)
        xa = np.zeros(3, np.float32)
        xb = rng.uniform(-1.2, 1.2, 3).astype(np.float32)
        d = seg_dist(*capsule_ends(xa, qa, ha), *capsule_ends(xb, qb, hb))
        c = orc.collide(CAPSULE, [ra, ha, 0], xa, qa, CAPSULE, [rb, hb, 0], xb, qb)
        if ra + rb - d < -0.02 - 1e-5:
            assert c["npts"] == 0
            continue
        if c["npts"] == 0:
            assert ra + rb - d < -0.02 + 1e-5
            continue
        n += 1
        assert abs(c["depth"].max() - (ra + rb - d)) <= 1e-5 * max(1.0, ra + rb)
    assert n > 40


def test_parallel_capsules_two_points(orc):
    c = orc.collide(CAPSULE, [0.3, 1.0, 0], [0, 0, 0], I4, CAPSULE, [0.3, 1.0, 0], [0.55, 0.5, 0], I4)
    assert c["npts"] == 2
    assert np.allclose(c["depth"], 0.05, atol=1e-5)
    assert np.allclose(c["normal"], [1, 0, 0], atol=1e-5)
    assert np.allclose(sorted(c["pos"][:, 1]), [-0.5, 1.0], atol=1e-5)


def test_box_box_depth_matches_sat(orc):
    rng = np.random.default_rng(17)
    n = 0
    for _ in range(300):
        ha, hb = rng.uniform(0.25, 0.75, 3), rng.uniform(0.25, 0.75, 3)
        qa, qb = rand_quat(rng), rand_quat(rng)
        xa = np.zeros(3)
        xb = rng.uniform(-1.3, 1.3, 3)
        depth, axis = sat_depth(ha, xa, qa, hb, xb, qb)
        c = orc.collide(BOX, ha, xa, qa, BOX, hb, xb, qb)
        if depth < -0.02 - 1e-5:
            assert c["npts"] == 0
            continue
        if depth < 0.005 or c["npts"] == 0:
            continue
        n += 1
        check_conventions(c)
        # the normal is a separating-axis candidate whose overlap bounds every point's depth; face
        # axes are preferred over an edge axis within 5 % + 1 mm, so its overlap is near the minimum
        nrm = c["normal"][0].astype(np.float64)
        Ra, Rb = rot(qa), rot(qb)
        ov = sum(ha[i] * abs(Ra[:, i] @ nrm) for i in range(3)) + sum(hb[i] * abs(Rb[:, i] @ nrm) for i in range(3)) \
            - abs((xb - xa) @ nrm)
        assert c["depth"].max() <= ov + 1e-5
        assert depth - 1e-5 <= ov <= depth / 0.95 + 0.0011
        # GJK/EPA on the same pair gives the SAT depth (penetration depth of two polytopes)
        out, st = orc.gjk_epa(BOX, ha, xa, qa, BOX, hb, xb, qb)
        assert st == 1
        assert abs(out[0] - depth) <= 1e-5 * max(1.0, depth), (out[0], depth)
        assert abs(abs(out[7:10] @ axis) - 1) < 1e-4
    assert n > 80


def test_gjk_distance_separated_boxes(orc):
    rng = np.random.default_rng(23)
    n = 0
    for _ in range(60):
        ha, hb = rng.uniform(0.25, 0.75, 3), rng.uniform(0.25, 0.75, 3)
        qa, qb = rand_quat(rng), rand_quat(rng)
        xa = np.zeros(3)
        xb = rng.normal(size=3)
        xb = xb / np.linalg.norm(xb) * rng.uniform(2.0, 3.0)
        Ra, Rb = rot(qa), rot(qb)
        f = lambda z: np.sum((Ra @ z[:3] - (xb + Rb @ z[3:])) ** 2)
        bounds = [(-ha[i], ha[i]) for i in range(3)] + [(-hb[i], hb[i]) for i in range(3)]
        best = min((optimize.minimize(f, rng.uniform(-0.2, 0.2, 6), bounds=bounds, method="L-BFGS-B",
                                      options={"ftol": 1e-16, "gtol": 1e-12}) for _ in range(4)),
                   key=lambda r: r.fun)
        want = math.sqrt(best.fun)
        if want < 0.05:
            continue
        out, st = orc.gjk_epa(BOX, ha, xa, qa, BOX, hb, xb, qb)
        assert st == 0
        n += 1
        assert abs(out[0] - want) <= 1e-5 * max(1.0, want), (out[0], want)
        # the witness points are on the boxes and realise the distance
        assert abs(np.linalg.norm(out[4:7] - out[1:4]) - out[0]) < 1e-5
    assert n > 40


def test_epa_sphere_centre_inside_box(orc):
    """Core of a sphere is its centre: EPA depth = distance from the centre to the nearest face."""
    rng = np.random.default_rng(29)
    for _ in range(50):
        h = rng.uniform(0.3, 0.75, 3)
        q = rand_quat(rng)
        local = rng.uniform(-0.9, 0.9, 3) * h
        x = rot(q) @ local
        out, st = orc.gjk_epa(BOX, h, [0, 0, 0], q, SPHERE, [0.3, 0, 0], x, I4)
        assert st == 1
        assert abs(out[0] - np.min(h - np.abs(local))) <= 1e-5


def test_sphere_capsule_point_segment_distance(orc):
    rng = np.random.default_rng(31)
    n = 0
    for _ in range(300):
        r = float(rng.uniform(0.25, 0.75))
        rc, hc = float(rng.uniform(0.2, 0.5)), float(rng.uniform(0.2, 1.0))
        qc = rand_quat(rng)
        xs = rng.uniform(-1.5, 1.5, 3).astype(np.float32)
        p0, p1 = capsule_ends([0, 0, 0], qc, hc)
        t = np.clip((xs - p0) @ (p1 - p0) / ((p1 - p0) @ (p1 - p0)), 0, 1)
        dist = np.linalg.norm(xs - (p0 + t * (p1 - p0)))
        c = orc.collide(SPHERE, [r, 0, 0], xs, I4, CAPSULE, [rc, hc, 0], [0, 0, 0], qc)
        if r + rc - dist < -0.02 - 1e-5:
            assert c["npts"] == 0
            continue
        if dist < 1e-3:
            continue
        n += 1
        assert c["npts"] == 1
        assert abs(c["depth"][0] - (r + rc - dist)) <= 1e-5 * max(1.0, r + rc)
    assert n > 50


def test_capsule_box_depth_brute_force(orc):
    """Shallow capsule-box contacts.  One point: depth = r - distance(segment, box), minimised over
    the segment by bounded search (point-box distance is a clamp in the box frame).  A capsule lying
    along a face (|axis . n| < 0.2) gets one point per segment end instead, each with depth
    r - distance(end, box)."""
    rng = np.random.default_rng(37)
    n = 0
    for _ in range(300):
        rc, hc = float(rng.uniform(0.2, 0.5)), float(rng.uniform(0.2, 1.0))
        h = rng.uniform(0.25, 0.75, 3)
        qc, qb = rand_quat(rng), rand_quat(rng)
        xc = rng.uniform(-1.8, 1.8, 3).astype(np.float32)
        Rb = rot(qb)
        p0, p1 = capsule_ends(xc, qc, hc)

        def d(t):
            p = Rb.T @ (p0 + t * (p1 - p0))
            return np.linalg.norm(p - np.clip(p, -h, h))

        ts = np.linspace(0, 1, 201)
        t0 = ts[np.argmin([d(t) for t in ts])]
        res = optimize.minimize_scalar(d, bounds=(max(0, t0 - 0.01), min(1, t0 + 0.01)), method="bounded",
                                       options={"xatol": 1e-10})
        dist = min(res.fun, d(t0))
        if dist < 0.05:  # deep or touching cores: EPA territory, checked elsewhere
            continue
        c = orc.collide(CAPSULE, [rc, hc, 0], xc, qc, BOX, h, [0, 0, 0], qb)
        if rc - dist < -0.02 - 1e-4:
            assert c["npts"] == 0
            continue
        if c["npts"] == 0:
            assert rc - dist < -0.02 + 1e-4
            continue
        n += 1
        if c["npts"] == 1:
            assert abs(c["depth"][0] - (rc - dist)) <= 2e-5, (c["depth"], rc - dist)
        else:
            assert c["npts"] == 2
            ends = sorted(rc - d(t) for t in (0.0, 1.0))
            assert np.allclose(sorted(c["depth"]), ends, atol=2e-5)
            assert c["depth"].max() <= rc - dist + 2e-5
    assert n > 30


def test_epa_box_segment_core_matches_sat(orc):
    """A capsule's core is its segment, a box with half extents (0, h, 0) along the capsule axis:
    the EPA depth of box vs segment equals the SAT overlap of the two boxes."""
    rng = np.random.default_rng(41)
    n = 0
    for _ in range(200):
        h = rng.uniform(0.3, 0.75, 3)
        hc = float(rng.uniform(0.2, 1.0))
        qa, qc = rand_quat(rng), rand_quat(rng)
        xc = rng.uniform(-0.6, 0.6, 3)
        depth, axis = sat_depth(h, np.zeros(3), qa, np.array([0.0, hc, 0.0]), xc, qc)
        if depth < 0.01:
            continue
        out, st = orc.gjk_epa(BOX, h, [0, 0, 0], qa, CAPSULE, [0.3, hc, 0], xc, qc)
        assert st == 1
        n += 1
        assert abs(out[0] - depth) <= 1e-5 * max(1.0, depth), (out[0], depth)
    assert n > 50


def test_gjk_epa_iteration_caps(orc):
    """The iteration caps: a cap below what a pair needs reports GJKFailedToConverge (301) or
    EPAFailedToConverge (302), the reference's codes (include/axiom/core/error_code.hpp:37-38);
    a cap at or above the iterations used changes nothing, bit for bit."""
    rng = np.random.default_rng(41)
    fails = {301: 0, 302: 0}
    for _ in range(80):
        ha, hb = rng.uniform(0.25, 0.75, 3), rng.uniform(0.25, 0.75, 3)
        qa, qb = rand_quat(rng), rand_quat(rng)
        xa = np.zeros(3)
        xb = rng.uniform(-1.2, 1.2, 3)
        out, st = orc.gjk_epa(BOX, ha, xa, qa, BOX, hb, xb, qb)
        assert st in (0, 1)
        big, stb = orc.gjk_epa(BOX, ha, xa, qa, BOX, hb, xb, qb, gjk_iters=200, epa_iters=200)
        assert stb == st and np.array_equal(big, out)
        _, s1 = orc.gjk_epa(BOX, ha, xa, qa, BOX, hb, xb, qb, gjk_iters=1)
        assert s1 == 301  # one GJK iteration never decides a box pair
        fails[301] += 1
        if st == 1:
            # one EPA expansion settles a pair only when the first polytope already holds the answer
            _, se = orc.gjk_epa(BOX, ha, xa, qa, BOX, hb, xb, qb, epa_iters=1)
            assert se in (1, 302)
            fails[302] += se == 302
    assert fails[302] > 10
