// Host-side launchers and the device buffer layout of one world (or one batch of
// worlds: bodies of every world concatenated, world index per body).
//
// HBM layout (structure of arrays, 16-byte vectors so every lane issues one
// dwordx4 load and a wave touches contiguous 1 KiB):
//   bodies   x      float4 (position, -)          q     float4 (x,y,z,w)
//            vel    float4 (v, inverse mass)      angv  float4 (w, -)
//            ii     float4 (body inverse inertia) shape float4 (params, type bits)
//   per step sv/sw  float4 solver velocities,     I0/I1 float4 world inverse inertia
//            amin/amax float4 AABB, key u64 + val u32 (cell key, body) and ping-pong copies
//   pairs    pa/pb u32, per-pair contact staging (4 slots x 2 float4)
//   manifolds (pair order)  m_ab uint2, m_np u32, c_pt/c_nf float4 [m*4+k]
//   solver rows (solve order, slot-major [k*cap + s]) R0..R3, lambda float4,
//            header uint4 (a, b, npts, manifold), feature u32
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace axm {

constexpr uint32_t PLANE_BIT = 0x80000000u;
constexpr int MAX_PLANES = 8;
constexpr uint32_t UNSET = 0xFFFFFFFFu;
constexpr uint64_t EMPTY_KEY = ~0ull;

struct StepParams {
    float gx, gy, gz;
    float dt, inv_dt;
    float friction, restitution, rest_threshold;
    float baumgarte, slop, max_bias, margin;
    float lin_keep, ang_keep;
    float inv_cell;
    float ox, oy, oz;
    uint32_t num_planes;
    float4 planes[MAX_PLANES];
};

struct Bodies {
    float4* x;
    float4* q;
    float4* vel;
    float4* angv;
    float4* ii;
    float4* shape;
    uint32_t* world;
    uint32_t* world_first;
    float4* sv;
    float4* sw;
    float4* I0;
    float4* I1;
    float4* amin;
    float4* amax;
};

struct CellTable {
    uint64_t* key;
    uint2* val;  // (start, count) in sorted order
    uint32_t mask;
};

struct Counters {
    uint32_t pairs;        // total candidate pairs
    uint32_t manifolds;
    uint32_t contacts;
    uint32_t gjk_fail;
    uint32_t epa_fail;
    uint32_t warm_hits;
    uint32_t remaining;    // colouring: uncoloured manifolds
    uint32_t max_colour;   // colouring: max colour + 1
    uint32_t last_round;   // colouring: last round that coloured something + 1
    uint32_t epa_count;    // deep box-capsule pairs queued for EPA
    uint32_t residual_bits;
    uint32_t pad[1];
    unsigned long long pgs_bytes;  // algorithmic bytes of one PGS sweep (all colours)
    unsigned long long pad2;
};

// k_sort.hip
size_t scan_tmp_words(uint32_t n);
void scan_exclusive(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* tmp, hipStream_t s);
size_t radix_tmp_words(uint32_t n);
// Returns 1 when the sorted result lives in the *_alt buffers.
int radix_sort_u64(uint64_t* keys, uint32_t* vals, uint64_t* keys_alt, uint32_t* vals_alt, uint32_t n, int end_bit,
                   uint32_t* tmp, hipStream_t s);
int radix_sort_u32(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt, uint32_t n, int end_bit,
                   uint32_t* tmp, hipStream_t s);

}  // namespace axm
