// Host-side launchers and the device buffer layout of one world (or one batch of
// worlds: bodies of every world concatenated, world index per body).
//
// HBM layout (structure of arrays, 16-byte vectors so every lane issues one
// dwordx4 load and a wave touches contiguous 1 KiB):
//   bodies   x      float4 (position, -)          q     float4 (x,y,z,w)
//            vel    float4 (v, inverse mass)      angv  float4 (w, -)
//            ii     float4 (body inverse inertia) shape float4 (params, type bits)
//   per step sv/sw  float4 solver velocities
//            aabb 2 x float4 (min; max, dynamic), key u64 + val u32 (cell key, body) and ping-pong copies
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
    float soft1, keep;  // 1 + softness and its inverse (regularised rows)
    uint32_t allow_sleep;
    float sleep_lin2, sleep_ang2, sleep_time;  // squared thresholds (m/s, rad/s), s
    uint32_t use_filter;  // some body has a collision filter
    uint32_t island_split;  // islands of at most ISLAND_LOCAL_MAX manifolds are solved on their own
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
    float4* aabb;  // per body [2i] (min, 0) and [2i+1] (max, 1 if dynamic this step)
    float4* smin;  // AABBs in sorted (cell) order, w = body index bits
    float4* smax;  // w = 1 for a dynamic body
    uint32_t* vslot;  // persistent sweep: per body and colour c < 64, the slot visiting it (slot << 1 | side)
    // islands and sleeping: per body (rest time bits, asleep bit 31 | island label); per step the
    // union-find parents and each island's smallest rest time; per label the step it wakes at
    float4* mat;   // friction, restitution, linear keep (1 - damping), angular keep
    uint4* bfl;    // flags (bit 0 kinematic), filter category, mask, group
    uint2* slp;
    uint32_t* upar;
    uint32_t* imin;
    uint32_t* wake;
    uint32_t* icnt;  // per step, per island root: the island's manifolds
};
// Islands of at most ISLAND_LOCAL_MAX manifolds are solved on their own (one lane each up
// to ISLAND_LANE_MAX manifolds, else one workgroup), each to its own residual; their
// solve slots follow those of the large islands, keyed LOCAL_KEY | root << 8 | colour.
constexpr uint32_t ISLAND_LOCAL_MAX = 4096u, ISLAND_LANE_MAX = 8u;
constexpr uint32_t LOCAL_KEY = 0x80000000u, ISLAND_ROOT_LIMIT = 1u << 23;
constexpr uint32_t ASLEEP_BIT = 0x80000000u, LABEL_MASK = 0x7FFFFFFFu, LABEL_NONE = 0x7FFFFFFFu;
constexpr uint32_t BODY_KINEMATIC = 1u;
// The broadphase counts and emits each body's pairs in three parts, one per z
// slice of its 27 neighbour cells (the planes go with the first), one thread each.
constexpr uint32_t BP_PARTS = 3;
constexpr uint32_t PAIR_STASH = 16;  // partners kept per part by the counting pass

// Cell hash table (linear probing): entry = (key lo, key hi, first sorted index,
// count), one 16-byte access per probe; plus an occupancy bitmap that sparse
// scenes test before probing: one bit per hashed cell, or, for one large world,
// one bit per cell at its 30-bit Morton code (neighbouring cells share words, and
// bodies in sorted order probe the same lines).  The Morton bits a step sets are
// listed per sorted body (the word, or UNSET) so the next step clears only those.
struct CellTable {
    uint4* ent;
    uint32_t mask;
    uint32_t* bits;
    uint32_t bmask;  // bit index mask (power of two minus one)
    uint32_t* mbits;  // Morton bitmap (2^25 words), or null
    uint32_t* mlist;  // per sorted body of the last step: the Morton word it set, or UNSET
    uint32_t mlist_n; // bodies of the last step (entries of mlist to clear)
    // With the Morton bitmap, the cells are found by rank instead of by hash: the occupied
    // cells in Morton order are the distinct keys of the sorted bodies, so cell c is number
    // wrank[c >> 5] + (set bits of its word below c) and its bodies are sorted positions
    // cstart[k] .. cstart[k + 1] - 1.
    uint32_t* wrank;  // per bitmap word with a set bit: the number of occupied cells before it
    uint32_t* cstart; // per occupied cell in Morton order: its first sorted position (+ n at the end)
    uint32_t* crank;  // per sorted position: 1 at a cell's first body, then its exclusive scan
};
constexpr uint32_t MORTON_WORDS = 1u << 25;        // 2^30 cells / 32
constexpr uint32_t MORTON_MIN_BODIES = 1u << 16;   // one world at least this large uses the Morton bitmap

// Sums that thousands of workgroups add to are spread over SPREAD rows of one
// 128-byte line each (workgroup b adds to row b mod SPREAD) and summed where they
// are read: device-scope atomics on one line queue one after another.
constexpr uint32_t SPREAD = 64;
struct SpreadRow {
    uint32_t contacts;   // contact points
    uint32_t warm_hits;  // points that found last step's impulse
    uint32_t cells;      // broadphase: occupied cells
    uint32_t pad;
    unsigned long long pgs_bytes;  // algorithmic bytes of one sweep (all colours)
    uint32_t islands;    // islands of dynamic bodies after the step
    uint32_t sleeping;   // bodies asleep after the step
    uint32_t fill[24];
};
static_assert(sizeof(SpreadRow) == 128, "one line per row");

// Device-resident step control block.  Every count a later kernel needs is read
// here on the device, so a step launches without host synchronisation (and can
// be captured into a HIP graph).  The first block and the spread sums are
// cleared at every step start.
struct Counters {
    uint32_t pairs;        // candidate pairs found (may exceed the capacity)
    uint32_t manifolds;    // manifolds found (may exceed the capacity)
    uint32_t bp_sparse;    // broadphase mode of the step (1: fewer than four bodies per occupied cell)
    uint32_t gjk_fail;
    uint32_t epa_fail;
    uint32_t m_glob;       // solve slots of the large islands, [0, m_glob); the rest belong to islands
                           // solved on their own
    uint32_t remaining;    // colouring: uncoloured manifolds
    uint32_t max_colour;   // colouring: max colour + 1
    uint32_t last_round;   // colouring: last round that coloured something + 1
    uint32_t epa_count;    // deep box-capsule pairs queued for EPA
    uint32_t residual_bits;
    uint32_t p_eff;        // min(pairs, pair capacity)
    uint32_t m_eff;        // min(manifolds, manifold capacity)
    uint32_t loc_sweeps;   // the most sweeps an island solved on its own ran
    uint32_t sweeps_run;   // sweeps of a launch with residual checks (set when it ends)
    uint32_t tol_hit;      // 1: the step's sweeps stopped because the residual met the tolerance
    uint32_t loc_res_bits; // the largest residual an island solved on its own ended with
    uint32_t loc_miss;     // 1: an island solved on its own stopped at the sweep cap above the tolerance
    uint32_t isl_lane_n;   // islands solved by one lane each
    uint32_t isl_group_n;  // islands solved by one workgroup each
    uint32_t loc_points;   // contact points of the islands solved on their own
    uint32_t pad5;
    unsigned long long loc_point_sweeps;     // their contact points x the sweeps each island ran
    unsigned long long bundle_point_sweeps;  // the same for the islands of the bundle kernel
    // ---- kept across steps
    uint32_t sticky;       // STICKY_* bits; while set the integration is skipped (state untouched)
    uint32_t steps_done;   // steps whose integration ran
    unsigned long long pgs_bytes_total;  // sweep bytes of all completed steps (pgs_bytes x sweeps)
    unsigned long long sweeps_total;     // sweeps of all completed steps
    unsigned long long bundle_point_sweeps_total;  // bundle_point_sweeps of all completed steps
    uint32_t cache_m[2];  // per row set: entries its (a, b) -> slot table holds (their positions: pc_pos)
    // ---- run totals since the last axm_world_get_totals reset (steps whose integration ran)
    uint32_t acc_steps;
    uint32_t acc_reached;  // steps that stopped at the residual tolerance
    uint32_t acc_max_sweeps;
    uint32_t acc_pad;
    unsigned long long acc_sweeps;
    unsigned long long acc_contact_sweeps;  // sum of contact points x sweeps
    unsigned long long acc_local_contact_sweeps;  // the part in islands solved on their own
    uint32_t acc_gjk_fail, acc_epa_fail;          // GJK / EPA failures
    // ---- spread sums, cleared per step
    alignas(128) SpreadRow sp[SPREAD];
};
template <class T> inline T spread_total(const Counters& c, T SpreadRow::*f) {
    T s = 0;
    for (uint32_t k = 0; k < SPREAD; ++k) s += c.sp[k].*f;
    return s;
}
constexpr uint32_t COUNTERS_RESET_WORDS = 26;  // words of the first block (cleared per step)
static_assert(offsetof(Counters, sticky) == 4 * COUNTERS_RESET_WORDS, "per-step block ends at sticky");
enum : uint32_t {
    STICKY_PAIRS = 1u,        // pair capacity exceeded
    STICKY_MANIFOLDS = 2u,    // manifold capacity exceeded
    STICKY_COLOURS = 4u,      // more colours than the colour sort supports (256)
    STICKY_TIMEOUT = 8u,      // a bounded wait expired (grid barrier or sweep dependency)
    STICKY_FLOW_COLOURS = 16u, // persistent sweep needs <= 64 colours: rerun with one launch per colour
    STICKY_INTERNAL = 32u      // an internal consistency check failed (colour masks vs manifold counts)
};

// Grid barrier state (generation counter; `count` returns to 0 after each barrier).
struct GridBar {
    uint32_t count;
    uint32_t gen;
    uint32_t pad[2];
};

// k_sort.hip
size_t scan_tmp_words(uint32_t n);
// out[0..n] = exclusive prefix sums, out[n] = total (also written to *total_out when given).
// With n_dev the count is min(n, *n_dev), read on the device (the grid is sized for n);
// with flags the input is read as (in[i] != 0).  With err, a look-back that exceeds its
// bounded wait (only a broken invariant can cause it) ORs STICKY_TIMEOUT into *err, so
// the step that used the partial prefix fails instead of running on it.
void scan_exclusive(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* tmp, hipStream_t s,
                    uint32_t* total_out = nullptr, const uint32_t* n_dev = nullptr, bool flags = false,
                    uint32_t* err = nullptr);
size_t radix_tmp_words(uint32_t n);
// Returns 1 when the sorted result lives in the *_alt buffers.  `n` is the
// capacity the grid is sized for; with n_dev the element count is read on the
// device (min(*n_dev, n)), so the launch needs no host synchronisation.
int radix_sort_u64(uint64_t* keys, uint32_t* vals, uint64_t* keys_alt, uint32_t* vals_alt, uint32_t n, int end_bit,
                   uint32_t* tmp, hipStream_t s, const uint32_t* n_dev = nullptr, uint32_t* err = nullptr);
int radix_sort_u32(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt, uint32_t n, int end_bit,
                   uint32_t* tmp, hipStream_t s, const uint32_t* n_dev = nullptr, uint32_t* err = nullptr);

}  // namespace axm
