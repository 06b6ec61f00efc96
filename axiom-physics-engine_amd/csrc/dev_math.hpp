// Device-side vector / quaternion helpers for the gfx950 kernels.
//
// Every formula fixes its operation order and the library is built with
// -ffp-contract=off, so a kernel evaluates exactly the IEEE-754 sequence the
// project's algorithm specification states (host and device agree bit for bit
// where the sequence is the same).  Conventions follow the reference math layer:
// quaternions are stored x,y,z,w (include/axiom/math/quat.hpp:29-34), the
// Hamilton product a*b applies b first (tests/math/quat_test.cpp:196-211), AABB
// overlap is inclusive (include/axiom/math/aabb.hpp:132-135).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define AXM_D __device__ __forceinline__

namespace axm {

struct f3 {
    float x, y, z;
};
AXM_D f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
AXM_D f3 xyz(float4 v) { return f3{v.x, v.y, v.z}; }
AXM_D f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
AXM_D f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
AXM_D f3 operator-(f3 a) { return {-a.x, -a.y, -a.z}; }
AXM_D f3 operator*(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
AXM_D float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
AXM_D f3 cross(f3 a, f3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
AXM_D float clampf(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }
AXM_D float sgn1(float v) { return v >= 0.0f ? 1.0f : -1.0f; }

struct q4 {
    float x, y, z, w;
};

// Column k (run-time) of the matrix with rows r0, r1, r2, by per-component selects on
// values: all nine entries are read first, so that the choice never becomes a load at a
// run-time offset (which would keep the matrix, and whatever holds it, in scratch memory).
AXM_D f3 m3_col(f3 r0, f3 r1, f3 r2, int k) {
    const bool k0 = k == 0, k1 = k == 1;
    return f3{k0 ? r0.x : (k1 ? r0.y : r0.z), k0 ? r1.x : (k1 ? r1.y : r1.z), k0 ? r2.x : (k1 ? r2.y : r2.z)};
}
struct m3 {
    f3 r0, r1, r2;
    AXM_D f3 col(int k) const { return m3_col(r0, r1, r2, k); }
};

AXM_D m3 quat_to_mat(q4 q) {
    const float xx = q.x * q.x, yy = q.y * q.y, zz = q.z * q.z;
    const float xy = q.x * q.y, xz = q.x * q.z, yz = q.y * q.z;
    const float wx = q.w * q.x, wy = q.w * q.y, wz = q.w * q.z;
    m3 m;
    m.r0 = {1.0f - 2.0f * (yy + zz), 2.0f * (xy - wz), 2.0f * (xz + wy)};
    m.r1 = {2.0f * (xy + wz), 1.0f - 2.0f * (xx + zz), 2.0f * (yz - wx)};
    m.r2 = {2.0f * (xz - wy), 2.0f * (yz + wx), 1.0f - 2.0f * (xx + yy)};
    return m;
}
AXM_D m3 quat_to_mat(float4 q) { return quat_to_mat(q4{q.x, q.y, q.z, q.w}); }
AXM_D f3 mul(const m3& m, f3 v) { return {dot(m.r0, v), dot(m.r1, v), dot(m.r2, v)}; }
AXM_D f3 mulT(const m3& m, f3 v) {
    return {m.r0.x * v.x + m.r1.x * v.y + m.r2.x * v.z, m.r0.y * v.x + m.r1.y * v.y + m.r2.y * v.z,
            m.r0.z * v.x + m.r1.z * v.y + m.r2.z * v.z};
}
AXM_D q4 qmul(q4 a, q4 b) {
    return {a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y, a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
            a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w, a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
}

// Symmetric world inverse inertia packed as (xx, yy, zz, xy) + (xz, yz, -, -).
struct s3 {
    float xx, yy, zz, xy, xz, yz;
};
AXM_D f3 smul(const s3& s, f3 v) {
    return {s.xx * v.x + s.xy * v.y + s.xz * v.z, s.xy * v.x + s.yy * v.y + s.yz * v.z,
            s.xz * v.x + s.yz * v.y + s.zz * v.z};
}
AXM_D s3 world_inv_inertia(const m3& R, f3 d) {
    s3 s;
    s.xx = R.r0.x * d.x * R.r0.x + R.r0.y * d.y * R.r0.y + R.r0.z * d.z * R.r0.z;
    s.yy = R.r1.x * d.x * R.r1.x + R.r1.y * d.y * R.r1.y + R.r1.z * d.z * R.r1.z;
    s.zz = R.r2.x * d.x * R.r2.x + R.r2.y * d.y * R.r2.y + R.r2.z * d.z * R.r2.z;
    s.xy = R.r0.x * d.x * R.r1.x + R.r0.y * d.y * R.r1.y + R.r0.z * d.z * R.r1.z;
    s.xz = R.r0.x * d.x * R.r2.x + R.r0.y * d.y * R.r2.y + R.r0.z * d.z * R.r2.z;
    s.yz = R.r1.x * d.x * R.r2.x + R.r1.y * d.y * R.r2.y + R.r1.z * d.z * R.r2.z;
    return s;
}

AXM_D void tangents(f3 n, f3& t1, f3& t2) {
    f3 t;
    if (fabsf(n.x) >= 0.57735f) {
        t = {n.y, -n.x, 0.0f};
    } else {
        t = {0.0f, n.z, -n.y};
    }
    const float inv = 1.0f / sqrtf(dot(t, t));
    t1 = t * inv;
    t2 = cross(n, t1);
}

AXM_D uint32_t part1by2(uint32_t v) {
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}
AXM_D uint32_t compact1by2(uint32_t v) {
    v &= 0x09249249u;
    v = (v ^ (v >> 2)) & 0x030C30C3u;
    v = (v ^ (v >> 4)) & 0x0300F00Fu;
    v = (v ^ (v >> 8)) & 0x030000FFu;
    v = (v ^ (v >> 16)) & 0x000003FFu;
    return v;
}
AXM_D uint32_t morton3(uint32_t x, uint32_t y, uint32_t z) {
    return part1by2(x) | (part1by2(y) << 1) | (part1by2(z) << 2);
}

AXM_D uint32_t pair_hash24(uint32_t a, uint32_t b) {
    uint32_t h = a * 0x9E3779B1u;
    h ^= b + 0x85EBCA77u + (h << 6) + (h >> 2);
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    h *= 0xC2B2AE35u;
    h ^= h >> 16;
    return h & 0xFFFFFFu;
}

AXM_D uint64_t mix64(uint64_t k) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ull;
    k ^= k >> 33;
    return k;
}

}  // namespace axm
