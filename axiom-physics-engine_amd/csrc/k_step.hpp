// Launch wrappers of the step kernels (k_step.hip) and the solver-row layout.
#pragma once
#include "launch.hpp"

namespace axm {

// Solver rows in solve order, slot-major.  Per contact point k of slot s, JROWS
// float4 at J[(k * JROWS + j) * cap + s] hold the three constraint directions
// d = n, t1, t2 with their angular Jacobians a = rA x d, b = rB x d and the
// inverse-inertia-weighted terms ia = IA a, ib = IB b (precomputed once per
// step, so a sweep's row update is dot products and fused multiply-adds):
//   j0 (n, mN)     j1 (aN, kN)   j2 (bN, -)  j3 (iaN, target)  j4 (ibN, -)
//   j5 (t1, mT1)   j6 (a1, kT1)  j7 (b1, -)  j8 (ia1, -)       j9 (ib1, -)
//   j10 (t2, mT2)  j11 (a2, kT2) j12 (b2, -) j13 (ia2, -)      j14 (ib2, -)
// (m = effective mass, k = 1 / m).  lam = (lambda_n, lambda_t1, lambda_t2, -) at
// k * cap + s, lam0 = its warm-start value, fid = feature id.
constexpr uint32_t JROWS = 15;
constexpr uint32_t INBOX_WORDS = 8;  // 64-bit inbox words per slot of the persistent sweep
struct SolverRows {
    uint4* hdr;  // (a, b, npts, manifold) per slot
    float4* J;
    float4* lam;
    float4* lam0;
    uint32_t* fid;
    uint2* dep;  // per slot: (deg << 8 | rank) of body a and body b in the colour order, UNSET if not dynamic
    uint32_t* colour;  // per slot: the manifold's colour (kept by the next step's incremental colouring)
    uint64_t* inbox;   // persistent sweep: per slot INBOX_WORDS words, velocities of body a (0..3) and b (4..7)
    uint2* next;       // persistent sweep: per slot, the next visit of body a and of b (slot << 1 | side)
};
// Steps between full recolourings (in between, persisting contacts keep their colour).
constexpr uint32_t RECOLOUR_PERIOD = 64;

struct PrevCacheArgs {
    const uint64_t* key;
    const uint32_t* slot;
    uint32_t mask;
    const float4* lam;
    const uint32_t* fid;
    const uint32_t* np;
    const uint32_t* colour;
    uint32_t cap;
    uint32_t valid;
};

// Clears the per-step counters and, when given, the 32 pair-kind counters.
void launch_counters_begin(Counters* cnt, uint32_t* kinds, hipStream_t s);
// Start of a step: velocities, world inertia, AABBs, cell keys; clears the
// colouring's per-body words.
// The sorted cell keys of the broadphase: 32-bit in the Morton rank mode (one world, where
// a key is the cell's Morton code), else 64-bit (world << 30 | Morton code); one is set.
struct SortedKeys {
    const uint64_t* k64 = nullptr;
    const uint32_t* k32 = nullptr;
    __host__ __device__ uint64_t operator[](uint32_t s) const { return k32 ? (uint64_t)k32[s] : k64[s]; }
};
// Also resets the step's counters (as launch_counters_begin).  The keys go to key32 (the
// Morton code) with the Morton bitmap, else to key.
void launch_begin(const Bodies& B, const StepParams& sp, uint32_t n, uint64_t* key, uint32_t* key32, uint32_t* val,
                  uint64_t* bmax0,
                  uint64_t* bmax1, uint64_t* bmask, uint32_t* bdeg, Counters* cnt, uint32_t* kinds,
                  const CellTable& ct, hipStream_t s);
// Islands of the step, after the narrowphase: union-find over the manifolds between awake
// dynamic bodies (root = smallest body index), and wake requests for sleeping islands an
// awake body touched.
void launch_isl_union(const Bodies& B, uint32_t n, const uint2* m_ab, uint32_t cap_man, Counters* cnt,
                      hipStream_t s);
// End of a step (after the integration): island labels, rest times, and islands falling asleep.
void launch_islands(const Bodies& B, const StepParams& sp, uint32_t n, Counters* cnt, hipStream_t s);
void launch_fill_u64(uint64_t* p, uint64_t v, uint32_t n, hipStream_t s);
void launch_fill_u32(uint32_t* p, uint32_t v, uint32_t n, hipStream_t s);
void launch_iota(uint32_t* p, uint32_t n, hipStream_t s);
// Empties the cell hash table and its occupancy bits (nothing with the Morton bitmap,
// whose words k_begin clears).
void launch_cell_clear(const CellTable& t, hipStream_t s);
void launch_cell_table(const Bodies& B, const SortedKeys& skey, const uint32_t* sval, uint32_t n, const CellTable& t,
                       Counters* cnt, uint32_t* scan_tmp, hipStream_t s);
// Pairs per (body, z slice) part, BP_PARTS * n counts in canonical order (stashing
// the first PAIR_STASH partners of each part), then the emit pass into pa / pb at
// the scanned offsets.
// Pair counts per (body, z slice) part in cnt, and per 256-part workgroup in bsum (the
// emit pass takes the exclusive scan of bsum, with the total at [workgroups], as boff).
void launch_pair_count(const Bodies& B, const StepParams& sp, const SortedKeys& skey, const uint32_t* sval,
                       const CellTable& ct, uint32_t n, uint32_t* cnt, uint32_t* bsum, uint32_t* stash, Counters* ctr,
                       hipStream_t s);
void launch_pair_emit(const Bodies& B, const StepParams& sp, const SortedKeys& skey, const uint32_t* sval,
                      const CellTable& ct, uint32_t n, const uint32_t* pcnt, const uint32_t* boff, uint32_t cap,
                      uint32_t* pa, uint32_t* pb, Counters* cnt, const uint32_t* stash, hipStream_t s);
// Narrowphase of every pair: the pairs are first ordered by kind of test (p_kind,
// the 32 counters `kinds` cleared at the step start, `order`), then tested; deep
// box-capsule pairs queue for EPA, run by `epa_groups` workgroups of 64.
void launch_narrow(const Bodies& B, const StepParams& sp, const uint32_t* pa, const uint32_t* pb, uint32_t cap_pairs,
                   uint32_t* p_np, float4* p_pt, float4* p_nf, uint32_t* epa_list, uint32_t* p_kind, uint32_t* kinds,
                   uint32_t* order, uint32_t epa_groups, Counters* cnt, hipStream_t s);
void launch_compact(const uint32_t* pa, const uint32_t* pb, uint32_t cap_pairs, const uint32_t* p_np,
                    const float4* p_pt, const float4* p_nf, const uint32_t* moff, uint32_t cap, uint2* m_ab,
                    uint32_t* m_np, float4* c_pt, float4* c_nf, Counters* cnt, hipStream_t s);
int colour_blocks_per_cu();
// Jones-Plassmann colouring; `prev` (previous step's (a, b) -> slot table and colours,
// valid = a previous step exists) enables the incremental form between full recolourings.
// Colouring: the keep pass, then Jones-Plassmann rounds over the manifolds not
// kept (`list`), in one workgroup for a short list or on the whole grid.
// With split, the first launch also counts the manifolds per island (icnt, after the union
// kernel), which the colour sort input of the island split needs.
void launch_colour(const Bodies& B, const uint2* m_ab, uint32_t cap_man, uint32_t* col, uint64_t* prio,
                   uint64_t* bmax0, uint64_t* bmax1, uint64_t* bmask, uint32_t* bdeg, uint32_t* list, GridBar* bar,
                   Counters* cnt, uint32_t blocks, const PrevCacheArgs& prev, uint32_t split, hipStream_t s);
void launch_col_sort_input(const Bodies& B, const uint2* m_ab, const uint32_t* col, uint32_t cap, uint32_t* key,
                           uint32_t* val, const Counters* cnt, uint32_t split,
                           hipStream_t s);
// The islands solved on their own (after the row preparation, which built their work
// list): one lane or one workgroup per island; tol > 0 stops each at its own residual.
// With flow_bmask (the colouring's per-body colour masks), its first launch also sets up
// the large islands' sweep (as launch_pgs_flow does unless told it is inited).
void launch_pgs_local(const Bodies& B, const StepParams& sp, const SolverRows& R, Counters* cnt, uint32_t cap,
                      const uint32_t* skey, uint2* list, float tol, uint32_t max_sweeps, uint32_t groups,
                      hipStream_t s, hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr,
                      const uint64_t* flow_bmask = nullptr);
// Work the row preparation launch also does for the launches after it, each in a loop of
// its own over the grid: the colour offsets of the sorted keys skey (col_off[c] = first
// position whose key has colour >= c, c = 0..K, and m_glob = col_off[K], when col_off is
// set), the small islands' work list from the same keys (when list is set;
// launch_pgs_local then builds none), clearing zero_n words (the sweep's residual words),
// and emptying the (a, b) -> slot table of row set cset (the entries its last build
// recorded, at the positions cpos holds; when ckey is set, the table's build is then made
// without clearing it).  More than max_k colours fail the step (STICKY_COLOURS).
struct PrepExtra {
    const uint32_t* skey = nullptr;
    uint32_t* col_off = nullptr;
    uint32_t max_k = 0xFFFFFFFFu;
    uint2* list = nullptr;
    uint32_t* zero = nullptr;
    uint32_t zero_n = 0;
    uint64_t* ckey = nullptr;
    const uint32_t* cpos = nullptr;
    uint32_t cset = 0;
};
void launch_prep(const Bodies& B, const StepParams& sp, const uint32_t* order, uint32_t cap, const uint2* m_ab,
                 const uint32_t* m_np, const float4* c_pt, const float4* c_nf, const SolverRows& R,
                 const PrevCacheArgs& pca, const float4* warm_in, const uint32_t* col, const uint64_t* bmask,
                 const uint32_t* bdeg, Counters* cnt, float* feat, const PrepExtra& X, hipStream_t s);
// Writes the learned warm start (network outputs Y, rows 4 * slot + point) into the impulse rows.
void launch_mlp_apply(const SolverRows& R, const Counters* cnt, uint32_t cap, const float* Y, const StepParams& sp,
                      hipStream_t s);
// Residual checks inside one persistent sweep launch: after every `every` sweeps and
// after the last one, the residual (the k_residual value) is written to
// bits[sweeps - 1]; the launch stops at the first check whose residual is <= tol
// (never when tol < 0) and stores the sweeps it ran in Counters::sweeps_run.
// The arrivals of the check after `sweeps` sweeps are counted in CHECK_ROWS words one
// line apart, arrive[((sweeps - 1) * CHECK_ROWS + r) * CHECK_ROW_WORDS] (workgroup g
// adds to row g mod CHECK_ROWS: atomics on one line queue one after another); both
// arrays start the step at zero.  bits == nullptr means no checks.
constexpr uint32_t CHECK_ROWS = 8, CHECK_ROW_WORDS = 32;
// Words of a residual buffer for `cap` sweeps: the residual of each sweep, then the
// arrival rows of a check after each sweep.
inline uint32_t res_words(uint32_t cap) { return cap * (1u + CHECK_ROWS * CHECK_ROW_WORDS); }
struct SweepStop {
    uint32_t* bits = nullptr;
    uint32_t* arrive = nullptr;
    float tol = -1.0f;
    uint32_t every = 1;
};
// Persistent sweep: `groups` workgroups of one wave, each with `lds_bytes` of LDS for
// its staged rows; all of them must be resident at once (see pgs_own_setup).  Events
// ev0 / ev1, when given, are recorded around the sweep kernel itself.  With inited, the
// inboxes and next visits were already set up by this step's launch_pgs_local (given
// bmask as flow_bmask).
void launch_pgs_flow(const Bodies& B, const StepParams& sp, const SolverRows& R, uint32_t cap, const uint64_t* bmask,
                     uint32_t sweep0, uint32_t nsweeps, Counters* cnt,
                     uint32_t spin_limit, uint32_t groups, uint32_t lds_bytes, uint32_t tune, hipStream_t s,
                     hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr, const SweepStop& stop = SweepStop{},
                     bool inited = false);
// Raises the sweep kernel's dynamic-LDS limit to max_lds_bytes and returns how many of
// its workgroups with lds_bytes each fit on one CU (0 on error).
int pgs_own_setup(uint32_t max_lds_bytes, uint32_t lds_bytes);
// Allows the bundle kernel of the small islands its LDS (once per process); 0 on failure.
int pgs_bundle_setup();
void launch_pgs(const Bodies& B, const StepParams& sp, const SolverRows& R, uint32_t s0, uint32_t s1, uint32_t cap,
                uint32_t first, hipStream_t s);
void launch_residual(const Bodies& B, const StepParams& sp, const SolverRows& R, const Counters* cnt, uint32_t cap,
                     uint32_t* out_bits, hipStream_t s);
// An (a, b) -> slot table of one row set, for its build.
struct CacheTable {
    uint64_t* key = nullptr;
    uint32_t* slot = nullptr;
    uint32_t mask = 0;
    uint32_t* np_out = nullptr;
    uint32_t* pos = nullptr;
    uint32_t set = 0;
};
// Empties the (a, b) -> slot table of row set `set` (the entries its last build made, at
// the positions pos holds; skipped with clear = false, when the row preparation did it)
// and fills it with the step's slots.
void launch_cache_build(const SolverRows& R, Counters* cnt, uint32_t cap, uint64_t* key, uint32_t* slot,
                        uint32_t mask, uint32_t* np_out, uint32_t* pos, uint32_t set, hipStream_t s,
                        bool clear = true);
// Rehashes the `on`-slot table (okey, oslot) into the empty table (key, slot, mask), with the
// new positions in pos.
void launch_cache_rehash(const uint64_t* okey, const uint32_t* oslot, uint32_t on, uint64_t* key, uint32_t* slot,
                         uint32_t mask, Counters* cnt, uint32_t* pos, hipStream_t s);
// sweeps == SWEEPS_ON_DEVICE: the count is read from Counters::sweeps_run.
constexpr uint32_t SWEEPS_ON_DEVICE = 0xFFFFFFFFu;
// With R and C, the same launch also builds the (a, b) -> slot table C of the step's rows
// (as launch_cache_build with clear = false).
void launch_integrate(const Bodies& B, const StepParams& sp, uint32_t n, Counters* cnt, uint32_t sweeps, hipStream_t s,
                      const SolverRows* R = nullptr, const CacheTable* C = nullptr);

}  // namespace axm
