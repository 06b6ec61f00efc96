"""Math primitives of the oracle (and, identically, of the device code) pinned by the
reference's own known answers (tests/math/quat_test.cpp) and by closed forms."""
import math

import numpy as np
import pytest


def axis_angle(axis, angle):
    a = np.asarray(axis, np.float64)
    a = a / np.linalg.norm(a)
    s = math.sin(angle / 2)
    return np.array([a[0] * s, a[1] * s, a[2] * s, math.cos(angle / 2)], np.float32)


PI = float(np.float32(math.pi))


def test_quat_axis_angle_rotations(orc):
    # quat_test.cpp:71-84 and :86-99
    assert np.allclose(orc.quat_rotate(axis_angle([0, 0, 1], PI / 2), [1, 0, 0]), [0, 1, 0], atol=1e-5)
    assert np.allclose(orc.quat_rotate(axis_angle([1, 0, 0], PI / 2), [0, 1, 0]), [0, 0, 1], atol=1e-5)
    # quat_test.cpp:213-222: 180 degrees about Z
    assert np.allclose(orc.quat_rotate(axis_angle([0, 0, 1], PI), [1, 0, 0]), [-1, 0, 0], atol=1e-5)
    # quat_test.cpp:394-403: 360 degrees
    assert np.allclose(orc.quat_rotate(axis_angle([0, 0, 1], 2 * PI), [1, 0, 0]), [1, 0, 0], atol=1e-4)


def test_quat_product_order(orc):
    """qx * qz applies qz first (quat_test.cpp:196-211): (1,0,0) -> (0,1,0) -> (0,0,1)."""
    qz = axis_angle([0, 0, 1], PI / 2)
    qx = axis_angle([1, 0, 0], PI / 2)
    assert np.allclose(orc.quat_rotate(orc.quat_mul(qx, qz), [1, 0, 0]), [0, 0, 1], atol=1e-5)
    # the reversed product lands elsewhere: a transposed operand fails this
    assert not np.allclose(orc.quat_rotate(orc.quat_mul(qz, qx), [1, 0, 0]), [0, 0, 1], atol=1e-3)


def test_quat_continuous_rotation(orc):
    """90 x 1 degree about Z composes to 90 degrees (quat_test.cpp:338-353, 1e-4)."""
    q = axis_angle([0, 0, 1], PI / 180)
    acc = np.array([0, 0, 0, 1], np.float32)
    for _ in range(90):
        acc = orc.quat_mul(q, acc)
    assert np.allclose(orc.quat_rotate(acc, [1, 0, 0]), [0, 1, 0], atol=1e-4)


def test_quat_product_against_matrix_composition(orc):
    """Hamilton product = composition of rotation matrices, on random unit quaternions."""
    rng = np.random.default_rng(7)
    for _ in range(50):
        a = rng.normal(size=4)
        b = rng.normal(size=4)
        a /= np.linalg.norm(a)
        b /= np.linalg.norm(b)
        v = rng.normal(size=3)
        ab = orc.quat_mul(a, b)
        want = orc.quat_rotate(a, orc.quat_rotate(b, v))
        assert np.allclose(orc.quat_rotate(ab, v), want, atol=1e-5)


def test_world_inverse_inertia(orc):
    rng = np.random.default_rng(3)
    for _ in range(20):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        d = rng.uniform(0.1, 2.0, 3)
        x, y, z, w = q
        R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                      [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                      [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
        I = R @ np.diag(d) @ R.T
        got = orc.world_inv_inertia(q, d)
        assert np.allclose(got, [I[0, 0], I[1, 1], I[2, 2], I[0, 1], I[0, 2], I[1, 2]], atol=1e-5)


@pytest.mark.parametrize("n", [[0, 1, 0], [1, 0, 0], [0, 0, -1], [0.6, 0.8, 0], [0.57735, 0.57735, 0.57735],
                               [-0.2, 0.3, 0.932738]])
def test_tangent_basis_is_orthonormal_and_right_handed(orc, n):
    n = np.asarray(n, np.float32)
    n /= np.linalg.norm(n)
    t1, t2 = orc.tangents(n)
    for a, b in ((n, t1), (n, t2), (t1, t2)):
        assert abs(np.dot(a, b)) < 1e-6
    assert abs(np.linalg.norm(t1) - 1) < 1e-6 and abs(np.linalg.norm(t2) - 1) < 1e-6
    assert np.allclose(np.cross(n, t1), t2, atol=1e-6)


def test_morton_interleave_brute_force(orc):
    rng = np.random.default_rng(1)
    for _ in range(200):
        x, y, z = (int(v) for v in rng.integers(0, 1024, 3))
        want = 0
        for bit in range(10):
            want |= ((x >> bit) & 1) << (3 * bit)
            want |= ((y >> bit) & 1) << (3 * bit + 1)
            want |= ((z >> bit) & 1) << (3 * bit + 2)
        assert orc.morton3(x, y, z) == want
    assert orc.morton3(1023, 1023, 1023) == (1 << 30) - 1


def test_pair_hash_spreads(orc):
    hs = {orc.pair_hash24(a, b) for a in range(64) for b in range(64)}
    assert len(hs) > 4000 and max(hs) < (1 << 24)
    assert orc.pair_hash24(3, 5) != orc.pair_hash24(5, 3)
