// CPU oracle — TEST INFRASTRUCTURE ONLY.
//
// Scalar vector / quaternion / matrix helpers for the oracle, templated on the
// scalar type (float for parity, double for the FP64 re-solve).  Every formula
// fixes its operation order (no FMA contraction: built with -ffp-contract=off)
// because the HIP kernels evaluate the same sequence and parity is bitwise where
// the sequence is identical.
//
// Conventions restated from the reference math layer:
//   * Vec3 dot / cross / zero-length normalise -> zero     include/axiom/math/vec3.hpp:15-255, :209-215
//   * Quaternion storage x,y,z,w (w scalar)                 include/axiom/math/quat.hpp:29-34
//   * Hamilton product a*b = "apply b then a"              tests/math/quat_test.cpp:196-211
//   * rotation v' = q v q^-1                               src/math/quat.cpp:33-38
//   * AABB overlap is inclusive (touching intersects)      include/axiom/math/aabb.hpp:132-135,
//                                                          tests/math/aabb_test.cpp:187-191
#pragma once
#include <cmath>
#include <cstdint>

namespace orc {

template <class T> struct V3 {
    T x, y, z;
};
template <class T> inline V3<T> v3(T x, T y, T z) { return V3<T>{x, y, z}; }
template <class T> inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> inline V3<T> operator-(V3<T> a) { return {-a.x, -a.y, -a.z}; }
template <class T> inline V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> inline V3<T> cross(V3<T> a, V3<T> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T> inline T clampv(T v, T lo, T hi) { return v < lo ? lo : (v > hi ? hi : v); }
template <class T> inline T sgn1(T v) { return v >= T(0) ? T(1) : T(-1); }  // sign with sign(0)=+1

template <class T> struct Q4 {
    T x, y, z, w;
};

// Row-major 3x3 rotation matrix.
template <class T> struct M3 {
    V3<T> r0, r1, r2;
    V3<T> col(int k) const {
        return k == 0 ? V3<T>{r0.x, r1.x, r2.x} : (k == 1 ? V3<T>{r0.y, r1.y, r2.y} : V3<T>{r0.z, r1.z, r2.z});
    }
};

// Rotation matrix of a unit quaternion (x,y,z,w).
template <class T> inline M3<T> quat_to_mat(Q4<T> q) {
    const T xx = q.x * q.x, yy = q.y * q.y, zz = q.z * q.z;
    const T xy = q.x * q.y, xz = q.x * q.z, yz = q.y * q.z;
    const T wx = q.w * q.x, wy = q.w * q.y, wz = q.w * q.z;
    M3<T> m;
    m.r0 = {T(1) - T(2) * (yy + zz), T(2) * (xy - wz), T(2) * (xz + wy)};
    m.r1 = {T(2) * (xy + wz), T(1) - T(2) * (xx + zz), T(2) * (yz - wx)};
    m.r2 = {T(2) * (xz - wy), T(2) * (yz + wx), T(1) - T(2) * (xx + yy)};
    return m;
}
template <class T> inline V3<T> mul(const M3<T>& m, V3<T> v) { return {dot(m.r0, v), dot(m.r1, v), dot(m.r2, v)}; }
// R^T v : column dot products
template <class T> inline V3<T> mulT(const M3<T>& m, V3<T> v) {
    return {m.r0.x * v.x + m.r1.x * v.y + m.r2.x * v.z, m.r0.y * v.x + m.r1.y * v.y + m.r2.y * v.z,
            m.r0.z * v.x + m.r1.z * v.y + m.r2.z * v.z};
}

// Hamilton product a (x) b.
template <class T> inline Q4<T> qmul(Q4<T> a, Q4<T> b) {
    return {a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y, a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
            a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w, a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
}

// Symmetric 3x3 (world inverse inertia).
template <class T> struct S3 {
    T xx, yy, zz, xy, xz, yz;
};
template <class T> inline V3<T> smul(const S3<T>& s, V3<T> v) {
    return {s.xx * v.x + s.xy * v.y + s.xz * v.z, s.xy * v.x + s.yy * v.y + s.yz * v.z,
            s.xz * v.x + s.yz * v.y + s.zz * v.z};
}
// R diag(d) R^T
template <class T> inline S3<T> world_inv_inertia(const M3<T>& R, V3<T> d) {
    S3<T> s;
    s.xx = R.r0.x * d.x * R.r0.x + R.r0.y * d.y * R.r0.y + R.r0.z * d.z * R.r0.z;
    s.yy = R.r1.x * d.x * R.r1.x + R.r1.y * d.y * R.r1.y + R.r1.z * d.z * R.r1.z;
    s.zz = R.r2.x * d.x * R.r2.x + R.r2.y * d.y * R.r2.y + R.r2.z * d.z * R.r2.z;
    s.xy = R.r0.x * d.x * R.r1.x + R.r0.y * d.y * R.r1.y + R.r0.z * d.z * R.r1.z;
    s.xz = R.r0.x * d.x * R.r2.x + R.r0.y * d.y * R.r2.y + R.r0.z * d.z * R.r2.z;
    s.yz = R.r1.x * d.x * R.r2.x + R.r1.y * d.y * R.r2.y + R.r1.z * d.z * R.r2.z;
    return s;
}

// Deterministic tangent basis of a unit normal.
template <class T> inline void tangents(V3<T> n, V3<T>& t1, V3<T>& t2) {
    V3<T> t;
    if (std::fabs(n.x) >= T(0.57735)) {
        t = {n.y, -n.x, T(0)};
    } else {
        t = {T(0), n.z, -n.y};
    }
    const T inv = T(1) / std::sqrt(dot(t, t));
    t1 = t * inv;
    t2 = cross(n, t1);
}

// 10-bit-per-axis Morton interleave (x in bit 0).
inline uint32_t part1by2(uint32_t v) {
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}
inline uint32_t morton3(uint32_t x, uint32_t y, uint32_t z) {
    return part1by2(x) | (part1by2(y) << 1) | (part1by2(z) << 2);
}

// 24-bit hash of a world-local body pair (colouring priority).
inline uint32_t pair_hash24(uint32_t a, uint32_t b) {
    uint32_t h = a * 0x9E3779B1u;
    h ^= b + 0x85EBCA77u + (h << 6) + (h >> 2);
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    h *= 0xC2B2AE35u;
    h ^= h >> 16;
    return h & 0xFFFFFFu;
}

}  // namespace orc
