// gfx950 kernels of one simulation step (single world or a batch of worlds).
//
// Stage order per step (the reference sketches the same four stages,
// examples/profiling_example.cpp:132-165, with semi-implicit Euler from
// examples/body_inspector_example.cpp:48-60):
//   k_begin            gravity/damping -> solver velocities, world inverse inertia, AABB, cell key
//   radix sort         (key, body) stable, 8-bit digits
//   k_cell_table       hash of cell -> (first sorted index, count)
//   k_pair_count/emit  27-neighbour grid scan, inclusive AABB test, j > i; exclusive scan between
//   k_narrow(+_epa)    primitive narrowphase -> <= 4 points per pair (staged per pair)
//   k_compact          pairs with points -> manifolds (pair order, exclusive scan of the flags)
//   k_col_*            Jones-Plassmann colouring with first-fit colours
//   k_prep             per-point rows + warm start (previous-step impulses keyed by (a, b, feature))
//   k_pgs              one colour of one sequential-impulse sweep
//   k_residual         velocity-space fixed-point residual
//   k_integrate        x += v dt, q += dt/2 (w,0) q, normalise
// Every body-parallel kernel is one thread per body; every manifold-parallel
// kernel one thread per manifold.  Colour batches are contiguous in the
// solve-order arrays, so a wave of 64 manifolds of one colour reads 64
// consecutive 16-byte rows.
#include "dev_geom.hpp"
#include "k_mlp.hpp"
#include "k_step.hpp"

namespace axm {

__device__ __forceinline__ uint32_t fbits(float f) { return __float_as_uint(f); }
__device__ __forceinline__ float4 f4(f3 v, float w) { return make_float4(v.x, v.y, v.z, w); }

// Adds the sum of v over the wave to field f of the workgroup's spread row with one
// atomic (every lane of the wave must call it: the reduction reads all 64 lanes).
template <class T> __device__ __forceinline__ void spread_add(Counters* cnt, T SpreadRow::*f, T v) {
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63u) == 0u && v != T(0)) atomicAdd(&(cnt->sp[blockIdx.x % SPREAD].*f), v);
}
template <class T> __device__ __forceinline__ T spread_sum(const Counters* cnt, T SpreadRow::*f) {
    T s = 0;
#pragma unroll 8
    for (uint32_t k = 0; k < SPREAD; ++k) s += cnt->sp[k].*f;
    return s;
}

constexpr uint32_t PAIR_KINDS = 16;  // narrowphase test kinds (pair_kind)
// The per-step counters back to zero (the threads of one workgroup).
__device__ __forceinline__ void counters_reset(Counters* cnt, uint32_t* __restrict__ kinds) {
    uint32_t* w = reinterpret_cast<uint32_t*>(cnt);
    if (threadIdx.x < COUNTERS_RESET_WORDS) w[threadIdx.x] = 0u;
    for (uint32_t k = threadIdx.x; k < SPREAD; k += blockDim.x) {
        cnt->sp[k].islands = 0u;
        cnt->sp[k].sleeping = 0u;
        cnt->sp[k].contacts = 0u;
        cnt->sp[k].warm_hits = 0u;
        cnt->sp[k].cells = 0u;
        cnt->sp[k].pgs_bytes = 0ull;
    }
    if (kinds && threadIdx.x < 2u * PAIR_KINDS) kinds[threadIdx.x] = 0u;
}
// ---------------------------------------------------------------- k_begin
__global__ void __launch_bounds__(256) k_begin(Bodies B, StepParams sp, uint32_t n, uint64_t* __restrict__ key,
                                               uint32_t* __restrict__ key32, uint32_t* __restrict__ val, uint64_t* __restrict__ bmax0,
                                               uint64_t* __restrict__ bmax1, uint64_t* __restrict__ bmask,
                                               uint32_t* __restrict__ bdeg, Counters* __restrict__ cnt,
                                               uint32_t* __restrict__ kinds, CellTable ct) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    // the step's counters start from zero (the kernels that count run after this launch;
    // steps_done, read below, is kept across steps)
    if (blockIdx.x == 0) counters_reset(cnt, kinds);
    // with the Morton bitmap: the words the last step set go back to zero (several bodies
    // may name one word: all write zero); the cell table sets this step's after the sort
    if (ct.mbits && i < ct.mlist_n) {
        const uint32_t w = ct.mlist[i];
        if (w != UNSET) ct.mbits[w] = 0u;
    }
    if (i >= n) return;
    // the colouring's per-body words start from zero every step
    bmax0[i] = 0ull;
    bmax1[i] = 0ull;
    bmask[i] = 0ull;
    bdeg[i] = 0u;
    // islands start as singletons; a sleeping body whose island an awake body touched in
    // the previous step wakes up (the request carries the number of this step)
    B.upar[i] = i;
    B.imin[i] = 0x7F800000u;  // +inf
    B.icnt[i] = 0u;
    bool asleep = false;
    {
        const uint2 sl = B.slp[i];
        if (sl.y & ASLEEP_BIT) {
            const uint32_t lab = sl.y & LABEL_MASK;
            if (lab < n && B.wake[lab] == cnt->steps_done) B.slp[i] = make_uint2(0u, lab);
            else asleep = true;
        }
    }
    const float4 x4 = B.x[i], q4v = B.q[i], v4 = B.vel[i], w4 = B.angv[i], sh = B.shape[i];
    const uint32_t type = fbits(sh.w);
    // a sleeping body acts as a static one for the whole step
    const bool dyn = v4.w > 0.0f && !asleep;
    f3 v = xyz(v4), w = xyz(w4);
    if (dyn) {
        const f3 g = {sp.gx, sp.gy, sp.gz};
        const float4 mt = B.mat[i];
        v = (v + g * sp.dt) * mt.z;
        w = w * mt.w;
    }
    B.sv[i] = f4(v, dyn ? v4.w : 0.0f);
    B.sw[i] = f4(w, 0.0f);
    const m3 R = quat_to_mat(q4v);
    const float mg = sp.margin;
    f3 e;
    if (type == SHAPE_SPHERE) {
        e = {sh.x + mg, sh.x + mg, sh.x + mg};
    } else if (type == SHAPE_BOX) {
        e.x = fabsf(R.r0.x) * sh.x + fabsf(R.r0.y) * sh.y + fabsf(R.r0.z) * sh.z + mg;
        e.y = fabsf(R.r1.x) * sh.x + fabsf(R.r1.y) * sh.y + fabsf(R.r1.z) * sh.z + mg;
        e.z = fabsf(R.r2.x) * sh.x + fabsf(R.r2.y) * sh.y + fabsf(R.r2.z) * sh.z + mg;
    } else {
        e.x = fabsf(R.r0.y) * sh.y + sh.x + mg;
        e.y = fabsf(R.r1.y) * sh.y + sh.x + mg;
        e.z = fabsf(R.r2.y) * sh.y + sh.x + mg;
    }
    const f3 c = xyz(x4);
    // the AABB as one 32-byte record, with whether the body is dynamic this step, so that
    // the gather in cell order reads one place per body
    B.aabb[2u * i] = f4(c - e, 0.0f);
    B.aabb[2u * i + 1u] = f4(c + e, dyn ? 1.0f : 0.0f);
    auto cc = [&](float p, float o) -> uint32_t {
        float f = (p - o) * sp.inv_cell;
        f = f > 0.0f ? f : 0.0f;
        f = f < 1023.0f ? f : 1023.0f;
        return (uint32_t)floorf(f);
    };
    const uint32_t cx = cc(c.x, sp.ox), cy = cc(c.y, sp.oy), cz = cc(c.z, sp.oz);
    if (key32) key32[i] = morton3(cx, cy, cz);  // one world: the Morton code alone
    else key[i] = ((uint64_t)B.world[i] << 30) | morton3(cx, cy, cz);
    val[i] = i;
}

// ---------------------------------------------------------------- cell table
__global__ void k_fill_u64(uint64_t* __restrict__ p, uint64_t v, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t v, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Linear-probing tables.  Every probe loop visits at most mask + 1 slots, so a
// full table ends the loop (insert returns false, lookup misses) instead of
// spinning; tables are sized >= 2x their maximum entry count.
__device__ __forceinline__ uint64_t ent_key(uint4 e) { return (uint64_t)e.x | ((uint64_t)e.y << 32); }
__device__ __forceinline__ bool ht_insert(CellTable t, uint64_t k, uint2 v, uint32_t s) {
    for (uint32_t probe = 0; probe <= t.mask; ++probe) {
        const unsigned long long prev =
            atomicCAS((unsigned long long*)&t.ent[s], (unsigned long long)EMPTY_KEY, (unsigned long long)k);
        if (prev == EMPTY_KEY || prev == k) {
            reinterpret_cast<uint2*>(&t.ent[s])[1] = v;
            return true;
        }
        s = (s + 1) & t.mask;
    }
    return false;
}
// Continues a lookup of key k at slot s (linear probing).
__device__ __forceinline__ bool ht_find_from(const CellTable& t, uint64_t k, uint32_t s, uint2& v) {
    for (uint32_t probe = 0; probe <= t.mask; ++probe) {
        const uint4 e = t.ent[s];
        const uint64_t kk = ent_key(e);
        if (kk == k) {
            v = make_uint2(e.z, e.w);
            return true;
        }
        if (kk == EMPTY_KEY) return false;
        s = (s + 1) & t.mask;
    }
    return false;
}
// Occupancy bit of a cell hash (the high half of the hash; the slot takes the low bits).
__device__ __forceinline__ uint32_t cell_bit(uint64_t h, uint32_t bmask) { return (uint32_t)(h >> 32) & bmask; }

// The hash table and its occupancy bits (the Morton bitmap is cleared by k_begin).
__global__ void __launch_bounds__(256) k_cell_clear(CellTable t) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= t.mask) t.ent[i] = make_uint4((uint32_t)EMPTY_KEY, (uint32_t)(EMPTY_KEY >> 32), 0u, 0u);
    if (i <= (t.bmask >> 5)) t.bits[i] = 0u;
}

// The cell table is built in two passes, so that most cells need no atomic.  Pass 1
// (k_cell_table): the first body of each cell stores the cell's key into its home
// slot (an untorn 8-byte store; of heads whose homes collide, one key stays), sets
// its occupancy bit and counts it.  Pass 2 (k_cell_place): a head whose key is in
// its home slot writes its range there; the others insert by compare-and-swap from
// the next slot on.  Slots claimed in pass 1 never change, so every key still lies
// after its home with no empty slot between (the linear-probing invariant).
// Each thread also gathers the AABB of sorted body s into smin / smax (body index in
// smin.w), so that the candidates of a cell are contiguous and carry what the test needs.
__global__ void __launch_bounds__(256) k_cell_table(Bodies B, SortedKeys skey,
                                                    const uint32_t* __restrict__ sval, uint32_t n, CellTable t,
                                                    Counters* __restrict__ cnt) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t lane = __lane_id();
    if (s < n) {
        const uint32_t i = sval[s];
        const float4 mn = B.aabb[2u * i], mx = B.aabb[2u * i + 1u];
        B.smin[s] = make_float4(mn.x, mn.y, mn.z, __uint_as_float(i));
        B.smax[s] = mx;
    }
    // the first body of each cell inserts it; one count per wave
    const uint64_t k = s < n ? skey[s] : EMPTY_KEY;
    const bool head = s < n && (s == 0 || skey[s - 1] != k);
    spread_add(cnt, &SpreadRow::cells, head ? 1u : 0u);
    const uint64_t heads = __ballot(head);
    const uint32_t code = (uint32_t)(k & 0x3FFFFFFFull);
    if (t.mbits) {
        if (s < n) {
            t.mlist[s] = head ? code >> 5 : UNSET;
            t.crank[s] = head ? 1u : 0u;
        }
        // the wave's cells are consecutive in Morton order and share bitmap words: one
        // atomic per distinct word, with the bits of every head in the wave that names it
        uint64_t todo = heads;
        while (todo) {
            const uint32_t lead = (uint32_t)__ffsll((unsigned long long)todo) - 1u;
            const uint32_t word = __shfl(code >> 5, (int)lead, 64);
            const bool mine = head && (code >> 5) == word;
            uint32_t b = mine ? 1u << (code & 31u) : 0u;
#pragma unroll
            for (int o = 32; o >= 1; o >>= 1) b |= __shfl_xor(b, o, 64);
            if (lane == lead) atomicOr(&t.mbits[word], b);
            todo &= ~__ballot(mine);
        }
        return;  // cells are found by rank (k_cell_rank), no hash table
    }
    if (!head) return;
    __hip_atomic_store(reinterpret_cast<uint64_t*>(&t.ent[(uint32_t)mix64(k) & t.mask]), k, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
    if (!t.mbits) {
        const uint32_t bit = cell_bit(mix64(k), t.bmask);
        atomicOr(&t.bits[bit >> 5], 1u << (bit & 31u));
    }
}
// With the Morton bitmap (crank holds the exclusive scan of the cell heads): each cell's
// first sorted position by rank, and each bitmap word's rank base (its first cell's rank).
// Also decides the broadphase mode of the step (bp_sparse: fewer than four bodies per
// occupied cell), as k_cell_place does in the hash mode.
__device__ __forceinline__ void decide_mode(Counters* cnt, uint32_t n) {
    cnt->bp_sparse = (uint64_t)spread_sum(cnt, &SpreadRow::cells) * 4u > n ? 1u : 0u;
}
__global__ void __launch_bounds__(256) k_cell_rank(SortedKeys skey, uint32_t n, CellTable t,
                                                   Counters* __restrict__ cnt) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s == 0u) decide_mode(cnt, n);
    if (s >= n) return;
    const uint64_t k = skey[s];
    const uint64_t kp = s ? skey[s - 1u] : EMPTY_KEY;
    const uint32_t r = t.crank[s];
    const bool head = s == 0u || kp != k;
    if (head) {
        t.cstart[r] = s;
        const uint32_t word = (uint32_t)(k & 0x3FFFFFFFull) >> 5;
        if (s == 0u || ((uint32_t)(kp & 0x3FFFFFFFull) >> 5) != word) t.wrank[word] = r;
    }
    if (s == n - 1u) t.cstart[r + (head ? 1u : 0u)] = n;
}
__global__ void __launch_bounds__(256) k_cell_place(SortedKeys skey, uint32_t n, CellTable t,
                                                    Counters* __restrict__ cnt) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t lane = __lane_id();
    if (s == 0u) decide_mode(cnt, n);
    const uint64_t k = s < n ? skey[s] : EMPTY_KEY;
    const bool head = s < n && (s == 0 || skey[s - 1] != k);
    const uint64_t heads = __ballot(head);
    if (!head) return;
    // the cell's count: up to the next head in the wave, else read on past the wave
    const uint64_t above = lane == 63u ? 0ull : heads & (~0ull << (lane + 1u));
    uint32_t c;
    if (above) {
        c = (uint32_t)__ffsll((unsigned long long)above) - 1u - lane;
    } else {
        c = 64u - lane;
        while (s + c < n && skey[s + c] == k) ++c;
        if (s + c > n) c = n - s;
    }
    const uint32_t h = (uint32_t)mix64(k) & t.mask;
    if (*reinterpret_cast<const uint64_t*>(&t.ent[h]) == k) {
        reinterpret_cast<uint2*>(&t.ent[h])[1] = make_uint2(s, c);
    } else if (!ht_insert(t, k, make_uint2(s, c), (h + 1u) & t.mask)) {
        atomicOr(&cnt->sticky, STICKY_INTERNAL);
    }
}

// ---------------------------------------------------------------- XCD-aware block order
// Workgroups are dispatched round-robin over the 8 XCDs (workgroup b runs on XCD b % 8),
// and each XCD has its own L2.  Kernels whose neighbouring items share gathered data
// and cost about the same (the pair kernels: neighbour cells of bodies in Morton order)
// take their block index from here instead: each XCD gets one contiguous run of blocks,
// so the shared lines stay in one L2.  Not for lists sorted by cost (pairs by kind,
// manifolds by island), where contiguous runs would load some XCDs far more than
// others.  A bijection of 0..gridDim.x-1, so a kernel that writes by item index is
// unchanged.
__device__ __forceinline__ uint32_t xcd_block() {
    const uint32_t G = gridDim.x, b = blockIdx.x;
    const uint32_t per = G >> 3, rem = G & 7u, x = b & 7u;
    return x * per + min(x, rem) + (b >> 3);
}

// ---------------------------------------------------------------- pairs
// Visits candidate partners of the body at sorted position t in the canonical
// order: planes, then the 27 neighbour cells (dz, dy, dx in -1..1), then
// ascending sorted position.  f(partner) is called for every emitted pair.
// AABBs in sorted order (k_cell_table): smin = (min, body index bits), smax = (max,
// 1 if dynamic).  The broadphase mode of the step (k_cell_rank / k_cell_place) is
// sparse when there are fewer than four bodies per occupied cell (the pair kernels
// come in a sparse and a dense build, each with only the registers its mode needs,
// and the one not chosen returns at once).

// Collision filter of two bodies (flags, category, mask, group): a shared nonzero group
// decides (positive: always, negative: never), else each category must meet the other's mask.
__device__ __forceinline__ bool should_collide(uint4 a, uint4 b) {
    const int ga = (int)a.w, gb = (int)b.w;
    if (ga == gb && ga != 0) return ga > 0;
    return (a.z & b.y) != 0u && (a.y & b.z) != 0u;
}

// Broadphase modes: cells found by rank in the occupancy bitmap (one large world), or
// through the hash table with few bodies per occupied cell (occupancy bits first) or
// many.  Each mode is its own kernel build, with only the registers its path needs.
constexpr int BP_RANK = 0, BP_SPARSE = 1, BP_DENSE = 2;

// The candidates of sorted body t in z slice dz (-1, 0, 1) of its neighbour cells,
// in canonical order; the planes come with slice -1.
template <int MODE, class F>
__device__ __forceinline__ void visit_pairs(const Bodies& B, const StepParams& sp, const SortedKeys& skey,
                                            const uint32_t* __restrict__ sval, const CellTable& ct, uint32_t t,
                                            int dz, F&& f) {
    const uint64_t key = MODE == BP_RANK ? (uint64_t)skey.k32[t] : skey.k64[t];
    const float4 mn = B.smin[t], mx = B.smax[t];
    const uint32_t i = __float_as_uint(mn.w);
    const bool dyn_i = mx.w > 0.0f;
    if (dyn_i && dz < 0) {
        const f3 c = xyz(B.x[i]);
        const f3 e = xyz(mx) - c;
        for (uint32_t k = 0; k < sp.num_planes; ++k) {
            const float4 pl = sp.planes[k];
            const float s = (pl.x * c.x + pl.y * c.y + pl.z * c.z) -
                            (fabsf(pl.x) * e.x + fabsf(pl.y) * e.y + fabsf(pl.z) * e.z);
            if (s <= pl.w) f(PLANE_BIT | k);
        }
    }
    const uint32_t mort = (uint32_t)(key & 0x3FFFFFFFull);
    const uint64_t wbits = key & ~0x3FFFFFFFull;
    const int cx = (int)compact1by2(mort), cy = (int)compact1by2(mort >> 1), cz = (int)compact1by2(mort >> 2);
    const int Z = cz + dz;
    if (Z < 0 || Z > 1023) return;
    // the cell's bodies in sorted order, four loaded at a time (a dense cell holds
    // dozens, and each load of the loop would otherwise wait on its own)
    auto test = [&](const float4 mnj, const float4 mxj) {
        const uint32_t j = __float_as_uint(mnj.w);
        if (j <= i) return;
        if (!dyn_i && !(mxj.w > 0.0f)) return;
        if (mn.x <= mxj.x && mx.x >= mnj.x && mn.y <= mxj.y && mx.y >= mnj.y && mn.z <= mxj.z && mx.z >= mnj.z) {
            if (sp.use_filter && !should_collide(B.bfl[i], B.bfl[j])) return;
            f(j);
        }
    };
    auto visit_range = [&](uint2 r) {
        const uint32_t se = r.x + r.y;
        uint32_t s = r.x;
        for (; s + 4u <= se; s += 4u) {
            float4 mn4[4], mx4[4];
#pragma unroll
            for (uint32_t u = 0; u < 4u; ++u) {
                mn4[u] = B.smin[s + u];
                mx4[u] = B.smax[s + u];
            }
#pragma unroll
            for (uint32_t u = 0; u < 4u; ++u) test(mn4[u], mx4[u]);
        }
        for (; s < se; ++s) test(B.smin[s], B.smax[s]);
    };
    // the nine cells of this z slice in canonical (dy, dx) order; Morton codes from the
    // dilated coordinates (three per axis instead of nine full interleavings)
    uint32_t px[3], py[3];
    bool okx[3], oky[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int X = cx + a - 1, Y = cy + a - 1;
        okx[a] = X >= 0 && X <= 1023;
        oky[a] = Y >= 0 && Y <= 1023;
        px[a] = part1by2((uint32_t)(okx[a] ? X : 0));
        py[a] = part1by2((uint32_t)(oky[a] ? Y : 0)) << 1;
    }
    const uint32_t pz = part1by2((uint32_t)Z) << 2;
    if constexpr (MODE == BP_RANK) {
        // one large world: the nine occupancy words are read once (they stay in the cache);
        // an occupied cell's range comes by rank from two arrays in Morton order
        // per cell: occupied, and the occupied cells of its word below it (5 bits each, six
        // to a register, so the nine words need not stay in registers)
        uint32_t occ = 0, below0 = 0, below1 = 0;
#pragma unroll
        for (int q = 0; q < 9; ++q) {
            const uint32_t code = px[q % 3] | py[q / 3] | pz;
            const uint32_t wv = okx[q % 3] && oky[q / 3] ? ct.mbits[code >> 5] : 0u;
            occ |= ((wv >> (code & 31u)) & 1u) << q;
            const uint32_t b = (uint32_t)__popc(wv & ((1u << (code & 31u)) - 1u));
            if (q < 6) below0 |= b << (5 * q);
            else below1 |= b << (5 * (q - 6));
        }
        while (occ) {
            const int q = __ffs(occ) - 1;
            occ &= occ - 1u;
            const uint32_t code = px[q % 3] | py[q / 3] | pz;
            const uint32_t below = (q < 6 ? below0 >> (5 * q) : below1 >> (5 * (q - 6))) & 31u;
            const uint32_t k = ct.wrank[code >> 5] + below;
            const uint32_t s0 = ct.cstart[k];
            visit_range(make_uint2(s0, ct.cstart[k + 1u] - s0));
        }
        return;
    }
    if constexpr (MODE == BP_SPARSE) {
        // few bodies per occupied cell: the occupancy bits (which stay in the cache) are
        // read first, then only the occupied cells are hashed and probed, in order
        uint32_t occ = 0;
#pragma unroll
        for (int q = 0; q < 9; ++q) {
            const uint32_t code = px[q % 3] | py[q / 3] | pz;
            const uint32_t bit = cell_bit(mix64(wbits | code), ct.bmask);
            const bool on = okx[q % 3] && oky[q / 3] && ((ct.bits[bit >> 5] >> (bit & 31u)) & 1u) != 0u;
            occ |= on ? 1u << q : 0u;
        }
        while (occ) {
            const int q = __ffs(occ) - 1;
            occ &= occ - 1u;
            const uint64_t nk = wbits | (px[q % 3] | py[q / 3] | pz);
            const uint32_t hs = (uint32_t)mix64(nk) & ct.mask;
            const uint4 e = ct.ent[hs];
            const uint64_t kk = ent_key(e);
            if (kk == EMPTY_KEY) continue;
            uint2 r = make_uint2(e.z, e.w);
            if (kk != nk && !ht_find_from(ct, nk, (hs + 1u) & ct.mask, r)) continue;
            visit_range(r);
        }
        return;
    }
    // dense: the first probes of all nine cells in flight together
    uint64_t nk[9];
    uint4 e[9];
    uint32_t hs[9];
    bool ok[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        ok[q] = okx[q % 3] && oky[q / 3];
        nk[q] = wbits | (px[q % 3] | py[q / 3] | pz);
        hs[q] = (uint32_t)mix64(nk[q]) & ct.mask;
    }
#pragma unroll
    for (int q = 0; q < 9; ++q)
        e[q] = ok[q] ? ct.ent[hs[q]] : make_uint4((uint32_t)EMPTY_KEY, (uint32_t)(EMPTY_KEY >> 32), 0u, 0u);
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        const uint64_t kk = ent_key(e[q]);
        if (!ok[q] || kk == EMPTY_KEY) continue;
        uint2 r = make_uint2(e[q].z, e[q].w);
        if (kk != nk[q] && !ht_find_from(ct, nk[q], (hs[q] + 1u) & ct.mask, r)) continue;
        visit_range(r);
    }
}

// Exclusive scan over a 256-thread workgroup (four waves); the workgroup's total in `total`.
__device__ __forceinline__ uint32_t wg_excl_scan(uint32_t v, uint32_t& total) {
    __shared__ uint32_t s_w[4];
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    uint32_t inc = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = __shfl_up(inc, o, 64);
        if (lane >= (uint32_t)o) inc += t;
    }
    if (lane == 63u) s_w[wv] = inc;
    __syncthreads();
    uint32_t base = 0, tot = 0;
#pragma unroll
    for (uint32_t w = 0; w < 4u; ++w) {
        const uint32_t x = s_w[w];
        if (w < wv) base += x;
        tot += x;
    }
    __syncthreads();
    total = tot;
    return base + inc - v;
}

// Counts the pairs of each (body, z slice) part u = 3 t + dz + 1 and keeps the
// first PAIR_STASH partners, so that the emit pass only re-visits the grid for the
// few parts with more.  Three threads per body: the probes of a body's slices are
// in flight together.
// The workgroup's total goes to bsum[block] (block in the XCD-aware order): the emit pass
// forms each part's offset from the scanned totals and a scan within its workgroup.
template <int MODE>
__device__ __forceinline__ void pair_count_body(const Bodies& B, const StepParams& sp, const SortedKeys& skey,
                                                const uint32_t* __restrict__ sval, const CellTable& ct, uint32_t n,
                                                uint32_t* __restrict__ cnt, uint32_t* __restrict__ bsum,
                                                uint32_t* __restrict__ stash, const Counters* __restrict__ ctr) {
    if (MODE != BP_RANK && (ctr->bp_sparse != 0u) != (MODE == BP_SPARSE)) return;
    const uint32_t lb = xcd_block();
    const uint32_t u = lb * blockDim.x + threadIdx.x;
    uint32_t c = 0;
    if (u < BP_PARTS * n) {
        const uint32_t t = u / BP_PARTS;
        const int dz = (int)(u - BP_PARTS * t) - 1;
        uint32_t* st = stash + (size_t)u * PAIR_STASH;
        visit_pairs<MODE>(B, sp, skey, sval, ct, t, dz, [&](uint32_t j) {
            if (c < PAIR_STASH) st[c] = j;
            ++c;
        });
        cnt[u] = c;
    }
    uint32_t tot;
    (void)wg_excl_scan(c, tot);
    if (threadIdx.x == 0u) bsum[lb] = tot;
}
// The rank and sparse builds fit 64 registers without scratch, so eight waves per SIMD
// hide more of the probe latency; the dense build keeps its registers.
template <int MODE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8)))
k_pair_count(Bodies B, StepParams sp, SortedKeys skey, const uint32_t* __restrict__ sval,
             CellTable ct, uint32_t n, uint32_t* __restrict__ cnt, uint32_t* __restrict__ bsum,
             uint32_t* __restrict__ stash, const Counters* __restrict__ ctr) {
    pair_count_body<MODE>(B, sp, skey, sval, ct, n, cnt, bsum, stash, ctr);
}
__global__ void __launch_bounds__(256)
k_pair_count_dense(Bodies B, StepParams sp, SortedKeys skey, const uint32_t* __restrict__ sval,
                   CellTable ct, uint32_t n, uint32_t* __restrict__ cnt, uint32_t* __restrict__ bsum,
                   uint32_t* __restrict__ stash, const Counters* __restrict__ ctr) {
    pair_count_body<BP_DENSE>(B, sp, skey, sval, ct, n, cnt, bsum, stash, ctr);
}

// Part u's pairs start at boff[block] (the scanned workgroup totals; boff[blocks] is the
// pair count) plus the counts of the parts before it in its workgroup.
template <int MODE>
__global__ void __launch_bounds__(256) k_pair_emit(Bodies B, StepParams sp, SortedKeys skey,
                                                   const uint32_t* __restrict__ sval, CellTable ct, uint32_t n,
                                                   const uint32_t* __restrict__ pcnt, const uint32_t* __restrict__ boff,
                                                   uint32_t cap, uint32_t* __restrict__ pa, uint32_t* __restrict__ pb,
                                                   Counters* __restrict__ cnt, const uint32_t* __restrict__ stash) {
    if (MODE != BP_RANK && (cnt->bp_sparse != 0u) != (MODE == BP_SPARSE)) return;
    const uint32_t lb = xcd_block();
    const uint32_t u = lb * blockDim.x + threadIdx.x;
    const uint32_t parts = BP_PARTS * n;
    if (u == 0) {
        const uint32_t P = boff[(parts + 255u) / 256u];
        cnt->p_eff = P < cap ? P : cap;
        if (P > cap) atomicOr(&cnt->sticky, STICKY_PAIRS);
    }
    const uint32_t c = u < parts ? pcnt[u] : 0u;
    uint32_t tot;
    const uint32_t ex = wg_excl_scan(c, tot);
    if (u >= parts) return;
    const uint32_t t = u / BP_PARTS;
    const int dz = (int)(u - BP_PARTS * t) - 1;
    uint32_t o = boff[lb] + ex;
    const uint32_t i = sval[t];
    if (c <= PAIR_STASH) {
        const uint32_t* st = stash + (size_t)u * PAIR_STASH;
        for (uint32_t k = 0; k < c && o + k < cap; ++k) {
            pa[o + k] = i;
            pb[o + k] = st[k];
        }
        return;
    }
    visit_pairs<MODE>(B, sp, skey, sval, ct, t, dz, [&](uint32_t j) {
        if (o < cap) {
            pa[o] = i;
            pb[o] = j;
        }
        ++o;
    });
}

// ---------------------------------------------------------------- narrowphase
__device__ __forceinline__ Shp load_shape(const Bodies& B, uint32_t i) {
    Shp s;
    const float4 sh = B.shape[i];
    s.c = xyz(B.x[i]);
    s.R = quat_to_mat(B.q[i]);
    s.type = fbits(sh.w);
    s.prm = xyz(sh);
    return s;
}

__device__ __forceinline__ void store_pair(uint32_t p, const Man& m, uint32_t* __restrict__ p_np,
                                           float4* __restrict__ p_pt, float4* __restrict__ p_nf) {
    p_np[p] = m.np;
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) {
        if (k >= m.np) break;
        p_pt[(size_t)p * 4 + k] = f4(m.pt[k].p, m.pt[k].depth);
        p_nf[(size_t)p * 4 + k] = f4(m.pt[k].n, __uint_as_float(m.pt[k].fid));
    }
}

// Processing order of the narrowphase: pairs grouped by the kind of test (the two
// shape types, or the shape type against a plane), so that a wave runs one test
// rather than every test its lanes need.  Results stay indexed by pair (PAIR_KINDS
// kinds, defined with the step counters).
__device__ __forceinline__ uint32_t pair_kind(const Bodies& B, uint32_t a, uint32_t b) {
    const uint32_t ta = fbits(B.shape[a].w) & 3u;
    if (b & PLANE_BIT) return 12u + ta;
    const uint32_t tb = fbits(B.shape[b].w) & 3u;
    const uint32_t lo = ta < tb ? ta : tb, hi = ta < tb ? tb : ta;
    return lo * 4u + hi;
}
// Per pair its kind; per kind the count (kinds[0..15]; kinds[16..31] are the fill
// cursors of k_pair_order, both cleared at the step start).
__global__ void __launch_bounds__(256) k_pair_kind(Bodies B, const uint32_t* __restrict__ pa,
                                                   const uint32_t* __restrict__ pb, uint32_t* __restrict__ p_kind,
                                                   uint32_t* __restrict__ kinds, const Counters* __restrict__ cnt) {
    __shared__ uint32_t h[PAIR_KINDS];
    if (threadIdx.x < PAIR_KINDS) h[threadIdx.x] = 0u;
    __syncthreads();
    const uint32_t P = cnt->p_eff;
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < P; p += gridDim.x * blockDim.x) {
        const uint32_t k = pair_kind(B, pa[p], pb[p]);
        p_kind[p] = k;
        atomicAdd(&h[k], 1u);
    }
    __syncthreads();
    if (threadIdx.x < PAIR_KINDS && h[threadIdx.x]) atomicAdd(&kinds[threadIdx.x], h[threadIdx.x]);
}
__global__ void __launch_bounds__(256) k_pair_order(const uint32_t* __restrict__ p_kind, uint32_t* __restrict__ kinds,
                                                    uint32_t* __restrict__ order, const Counters* __restrict__ cnt) {
    __shared__ uint32_t base[PAIR_KINDS], off[PAIR_KINDS], h[PAIR_KINDS];
    const uint32_t P = cnt->p_eff;
    if (threadIdx.x < PAIR_KINDS) {
        uint32_t o = 0;
        for (uint32_t k = 0; k < threadIdx.x; ++k) o += kinds[k];
        base[threadIdx.x] = o;
    }
    // a bounded grid strides over the pairs (block-uniform trip count: the barriers below)
    for (uint32_t b0 = blockIdx.x * blockDim.x; b0 < P; b0 += gridDim.x * blockDim.x) {
        const uint32_t p = b0 + threadIdx.x;
        if (threadIdx.x < PAIR_KINDS) h[threadIdx.x] = 0u;
        __syncthreads();
        const uint32_t k = p < P ? p_kind[p] : 0u;
        const uint32_t r = p < P ? atomicAdd(&h[k], 1u) : 0u;
        __syncthreads();
        if (threadIdx.x < PAIR_KINDS)
            off[threadIdx.x] = base[threadIdx.x] +
                               (h[threadIdx.x] ? atomicAdd(&kinds[PAIR_KINDS + threadIdx.x], h[threadIdx.x]) : 0u);
        __syncthreads();
        if (p < P) {
            const uint32_t q = off[k] + r;
            if (q < P) order[q] = p;
        }
        __syncthreads();
    }
}

// The heavy tests (box-box, box-capsule) run in a kernel of their own, so that the
// other tests keep a small register allocation and no scratch memory.  In the
// kind-sorted order the heavy pairs are one contiguous range [h0, h1).
constexpr uint32_t KIND_BOX_BOX = 1u * 4u + 1u, KIND_BOX_CAPSULE = 1u * 4u + 2u;
static_assert(KIND_BOX_CAPSULE == KIND_BOX_BOX + 1u, "heavy kinds are adjacent");
template <bool HEAVY>
__global__ void __launch_bounds__(256) k_narrow(Bodies B, StepParams sp, const uint32_t* __restrict__ pa,
                                                const uint32_t* __restrict__ pb, uint32_t* __restrict__ p_np,
                                                float4* __restrict__ p_pt, float4* __restrict__ p_nf,
                                                uint32_t* __restrict__ epa_list, Counters* __restrict__ cnt,
                                                const uint32_t* __restrict__ order, const uint32_t* __restrict__ kinds) {
    const uint32_t P = cnt->p_eff;
    uint32_t h0 = 0;
    for (uint32_t k = 0; k < KIND_BOX_BOX; ++k) h0 += kinds[k];
    uint32_t h1 = h0 + kinds[KIND_BOX_BOX] + kinds[KIND_BOX_CAPSULE];
    h1 = min(h1, P);
    h0 = min(h0, h1);
    const uint32_t nh = h1 - h0, N = HEAVY ? nh : P - nh;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += gridDim.x * blockDim.x) {
    const uint32_t t = HEAVY ? h0 + v : (v < h0 ? v : v + nh);
    const uint32_t p = order[t];
    const uint32_t a = pa[p], b = pb[p];
    Man m;
    m.np = 0;
    const Shp A = load_shape(B, a);
    if (!HEAVY && (b & PLANE_BIT)) {
        const float4 pl = sp.planes[b & ~PLANE_BIT];
        collide_plane(m, A, f3{pl.x, pl.y, pl.z}, pl.w, sp.margin);
    } else {
        const Shp Bs = load_shape(B, b);
        bool need_epa = false;
        int st;
        const bool swap = A.type > Bs.type;
        constexpr int TESTS = HEAVY ? COLLIDE_HEAVY : COLLIDE_LIGHT;
        st = swap ? collide_canon<TESTS>(m, Bs, A, sp.margin, need_epa)
                  : collide_canon<TESTS>(m, A, Bs, sp.margin, need_epa);
        if (need_epa) {
            const uint32_t slot = atomicAdd(&cnt->epa_count, 1u);
            epa_list[slot] = p;
        } else if (swap) {
            flip_manifold(m);
        }
        if (st == GEOM_GJK_FAIL) atomicAdd(&cnt->gjk_fail, 1u);
    }
    store_pair(p, m, p_np, p_pt, p_nf);
    }
}

// Deep box-capsule pairs (cores intersect): EPA, one queued pair per workgroup of
// one wave at a time (a fixed grid strides over the queue).
__global__ void __launch_bounds__(64) k_narrow_epa(Bodies B, StepParams sp, const uint32_t* __restrict__ pa,
                                                   const uint32_t* __restrict__ pb,
                                                   const uint32_t* __restrict__ epa_list, uint32_t cap,
                                                   uint32_t* __restrict__ p_np, float4* __restrict__ p_pt,
                                                   float4* __restrict__ p_nf, Counters* __restrict__ cnt) {
    const uint32_t nq = min(cnt->epa_count, cap);
    // one pair per workgroup (one wave) at a time, on its first lane: each pair's
    // expansion runs its own iteration count, and its polytope lives in the
    // workgroup's LDS instead of scratch memory
    __shared__ EpaWork W;
    if (threadIdx.x != 0u) return;
    for (uint32_t q = blockIdx.x; q < nq; q += gridDim.x) {
        const uint32_t p = epa_list[q];
        const uint32_t a = pa[p], b = pb[p];
        Shp A = load_shape(B, a), Bs = load_shape(B, b);
        const bool swap = A.type > Bs.type;
        Man m;
        m.np = 0;
        const int st = swap ? box_capsule_epa(m, Bs, A, W) : box_capsule_epa(m, A, Bs, W);
        if (swap) flip_manifold(m);
        if (st == GEOM_EPA_FAIL) atomicAdd(&cnt->epa_fail, 1u);
        store_pair(p, m, p_np, p_pt, p_nf);
    }
}

__global__ void __launch_bounds__(256) k_compact(const uint32_t* __restrict__ pa, const uint32_t* __restrict__ pb,
                                                 const uint32_t* __restrict__ p_np,
                                                 const float4* __restrict__ p_pt, const float4* __restrict__ p_nf,
                                                 const uint32_t* __restrict__ moff, uint32_t cap,
                                                 uint2* __restrict__ m_ab, uint32_t* __restrict__ m_np,
                                                 float4* __restrict__ c_pt, float4* __restrict__ c_nf,
                                                 Counters* __restrict__ cnt) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const uint32_t M = cnt->manifolds;
        cnt->m_eff = M < cap ? M : cap;
        if (M > cap) atomicOr(&cnt->sticky, STICKY_MANIFOLDS);
    }
    const uint32_t P = cnt->p_eff;
    uint32_t contacts = 0;
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < P; p += gridDim.x * blockDim.x) {
        const uint32_t np = p_np[p];
        contacts += np;
        if (np == 0) continue;
        const uint32_t m = moff[p];
        if (m >= cap) continue;
        m_ab[m] = make_uint2(pa[p], pb[p]);
        m_np[m] = np;
        for (uint32_t k = 0; k < np; ++k) {
            c_pt[(size_t)m * 4 + k] = p_pt[(size_t)p * 4 + k];
            c_nf[(size_t)m * 4 + k] = p_nf[(size_t)p * 4 + k];
        }
    }
    spread_add(cnt, &SpreadRow::contacts, contacts);
}

// ---------------------------------------------------------------- colouring
__device__ __forceinline__ bool dyn_body(const Bodies& B, uint32_t b) {
    return !(b & PLANE_BIT) && B.sv[b].w > 0.0f;
}
__device__ __forceinline__ uint32_t local_id(const Bodies& B, uint32_t b) {
    return (b & PLANE_BIT) ? b : b - B.world_first[B.world[b]];
}

// Grid-wide barrier for a kernel whose workgroups are all resident (grid sized
// from the occupancy query).  Every wave first waits until its own stores are
// complete in the XCD's L2 (a workgroup barrier does not wait for them); then one
// thread per workgroup writes the L2 back and arrives (agent-scope release), waits
// for the last arrival and invalidates the stale L1/L2 lines for the whole
// workgroup (agent-scope acquire).  One writeback per workgroup instead of one per
// wave: the writeback of an XCD's L2 is the expensive part of a fence.  A bounded
// wait sets STICKY_TIMEOUT instead of hanging.
// Grid barrier over `nblocks` resident workgroups.  The last workgroup to arrive
// copies *snap (a counter the workgroups update before arriving) into the barrier
// before releasing it, and every workgroup returns that copy: all of them see the
// value as it stood when the whole grid had arrived, even if a fast workgroup
// already updates the counter again in the next round.
__device__ __forceinline__ uint32_t grid_sync(GridBar* bar, uint32_t nblocks, Counters* cnt, const uint32_t* snap) {
    __shared__ uint32_t s_snap;
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t g = __hip_atomic_load(&bar->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t arrived = __hip_atomic_fetch_add(&bar->count, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (arrived == nblocks - 1) {
            const uint32_t v = __hip_atomic_load(snap, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&bar->pad[0], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&bar->count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_add(&bar->gen, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            s_snap = v;
        } else {
            const uint64_t t0 = wall_clock64();  // 100 MHz constant clock
            bool late = false;
            while (__hip_atomic_load(&bar->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == g) {
                __builtin_amdgcn_s_sleep(2);
                if (wall_clock64() - t0 > 200000000ull) {  // 2 s
                    atomicOr(&cnt->sticky, STICKY_TIMEOUT);
                    late = true;
                    break;
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            s_snap = late ? 0u : __hip_atomic_load(&bar->pad[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    __syncthreads();
    return s_snap;
}

// Jones-Plassmann colouring in one persistent launch.  Round r: every
// uncoloured manifold offers tag(r, m) = (r+1) << 48 | hash24(a,b) << 24 | m at
// its dynamic bodies (atomic max into the round's buffer); a manifold holding
// the max at all its dynamic bodies takes the first colour free in both bodies'
// 64-bit masks (64 + r once they are full).  Winners of one round share no
// dynamic body, so their mask updates cannot collide.  Losers offer their next
// tag into the other buffer in the same pass, so each round costs one barrier;
// tags grow with the round, so the buffers never need clearing.
struct PrevCache {
    const uint64_t* key;  // (a << 32 | b) -> previous slot
    const uint32_t* slot;
    uint32_t mask;
    const float4* lam;    // [k * cap + s]
    const uint32_t* fid;  // [k * cap + s]
    const uint32_t* np;   // points per previous slot
    const uint32_t* colour;  // colour per previous slot
    uint32_t cap;
    uint32_t valid;
};

__device__ __forceinline__ bool prev_find(const PrevCache& c, uint64_t k, uint32_t& s) {
    uint32_t h = (uint32_t)mix64(k) & c.mask;
    for (uint32_t probe = 0; probe <= c.mask; ++probe) {
        const uint64_t kk = c.key[h];
        if (kk == k) {
            s = c.slot[h];
            return true;
        }
        if (kk == EMPTY_KEY) return false;
        h = (h + 1) & c.mask;
    }
    return false;
}

// Union-find over body indices.  Every link points from a larger root to a smaller one
// (compare-and-swap on the root), so the root of a finished component is its smallest
// body index whatever the order of the links (lock-free union-find with path halving:
// Jaiganesh & Burtscher 2018, ECL-CC).  The parents are shared between workgroups, so
// they move through agent-scope atomics.
__device__ __forceinline__ uint32_t uf_load(uint32_t* p, uint32_t x) {
    return __hip_atomic_load(p + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t uf_find(uint32_t* par, uint32_t x) {
    uint32_t cur = uf_load(par, x);
    if (cur != x) {
        uint32_t prev = x, next;
        while (cur > (next = uf_load(par, cur))) {
            __hip_atomic_store(par + prev, next, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            prev = cur;
            cur = next;
        }
    }
    return cur;
}
// Root of x in the finished forest (after the union kernel, which no later kernel of the
// step changes): the union's coherent loads, without its path-halving stores (in one giant
// island every walk ends on the same root word, which stores would keep contended).
__device__ __forceinline__ uint32_t uf_root(uint32_t* par, uint32_t x) {
    uint32_t p;
    while ((p = uf_load(par, x)) != x) x = p;
    return x;
}
// Island (union-find root) of a manifold: that of its dynamic body, UNSET with none.
__device__ __forceinline__ uint32_t man_root(const Bodies& B, uint2 ab) {
    const bool plane = (ab.y & PLANE_BIT) != 0u;
    const uint32_t d = B.sv[ab.x].w > 0.0f ? ab.x : (!plane && B.sv[ab.y].w > 0.0f ? ab.y : UNSET);
    return d == UNSET ? UNSET : uf_root(B.upar, d);
}
// Per-island atomics of a wave: the lanes whose island is that of the first lane taking
// part combine their values into one atomic made by that lane; the others make their own.
// In one giant island a wave makes one atomic instead of 64 on the same word.
template <class Op>
__device__ __forceinline__ void wave_by_island(bool on, uint32_t r, uint32_t v, uint32_t idle, Op op) {
    const uint64_t act = __ballot(on);
    if (act == 0ull) return;
    const uint32_t lead = (uint32_t)__ffsll((unsigned long long)act) - 1u;
    const uint32_t lr = (uint32_t)__shfl((int)r, (int)lead, 64);
    const bool same = on && r == lr;
    uint32_t c = same ? v : idle;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) c = op.combine(c, (uint32_t)__shfl_xor((int)c, o, 64));
    if (__lane_id() == lead) op.apply(lr, c);
    if (on && !same) op.apply(r, v);
}
// Manifolds per island (after the union kernel), run by the colouring's first launch
// (k_colour_keep), every thread of its grid.
__device__ __forceinline__ void isl_count_part(const Bodies& B, const uint2* __restrict__ m_ab, uint32_t cap,
                                               const Counters* __restrict__ cnt) {
    if (cnt->sticky != 0u) return;
    const uint32_t M = min(cap, cnt->m_eff);
    struct Add {
        uint32_t* icnt;
        __device__ uint32_t combine(uint32_t a, uint32_t b) const { return a + b; }
        __device__ void apply(uint32_t r, uint32_t c) const { atomicAdd(icnt + r, c); }
    };
    // every lane of a wave runs the same number of rounds (the wave-wide combine)
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t m0 = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); m0 < M; m0 += stride) {
        const uint32_t m = m0 + __lane_id();
        const uint32_t r = m < M ? man_root(B, m_ab[m]) : UNSET;
        wave_by_island(r != UNSET, r, 1u, 0u, Add{B.icnt});
    }
}
// Pass 1 (one thread per manifold, no barrier): degrees, priorities, the kept
// colours of persisting contacts (with their mask bits) and, for every other
// manifold, its round-0 tag at its dynamic bodies and an entry in the list of
// manifolds to colour (cnt->remaining counts them).  The per-body arrays were
// cleared by k_begin.  Temporal coherence: between full recolourings (every
// recolour_period steps) a manifold whose body pair was in contact in the
// previous step keeps that step's colour (the kept ones stay a valid colouring:
// a subset of one); Jones-Plassmann then colours only the others around them.
__global__ void __launch_bounds__(256) k_colour_keep(Bodies B, const uint2* __restrict__ m_ab,
                                                     uint32_t* __restrict__ col, uint64_t* __restrict__ prio,
                                                     uint64_t* __restrict__ bmax0, uint64_t* __restrict__ bmask,
                                                     uint32_t* __restrict__ bdeg, uint32_t* __restrict__ list,
                                                     Counters* cnt, PrevCache pc, uint32_t recolour_period,
                                                     uint32_t count_cap) {
    // with the island split, the manifolds per island first (count_cap: the manifold capacity)
    if (count_cap) isl_count_part(B, m_ab, count_cap, cnt);
    const uint32_t M = cnt->m_eff;
    const bool incremental = pc.valid != 0u && (cnt->steps_done % recolour_period) != 0u;
    // per workgroup, the largest kept colour and the list entries are gathered in LDS
    // and go out with one atomic each (atomics on one address queue one after another)
    __shared__ uint32_t s_cmax, s_n, s_base;
    // a bounded grid strides over the manifolds (block-uniform trip count: the barriers below)
    for (uint32_t b0 = blockIdx.x * blockDim.x; b0 < M; b0 += gridDim.x * blockDim.x) {
        const uint32_t m = b0 + threadIdx.x;
        if (threadIdx.x == 0) {
            s_cmax = 0u;
            s_n = 0u;
        }
        __syncthreads();
        bool fresh = false;
        if (m < M) {
            const uint2 ab = m_ab[m];
            const bool bd = dyn_body(B, ab.y);
            atomicAdd(&bdeg[ab.x], 1u);
            if (bd) atomicAdd(&bdeg[ab.y], 1u);
            const uint64_t pr =
                ((uint64_t)pair_hash24(local_id(B, ab.x), local_id(B, ab.y)) << 24) | (uint64_t)(m & 0xFFFFFFu);
            prio[m] = pr;
            uint32_t c = UNSET, ps;
            if (incremental && prev_find(pc, ((uint64_t)ab.x << 32) | ab.y, ps)) {
                const uint32_t pcol = pc.colour[ps];
                if (pcol < 64u) c = pcol;
            }
            col[m] = c;
            if (c != UNSET) {
                atomicOr((unsigned long long*)&bmask[ab.x], 1ull << c);
                if (bd) atomicOr((unsigned long long*)&bmask[ab.y], 1ull << c);
                atomicMax(&s_cmax, c + 1u);
            } else {
                const uint64_t t = (1ull << 48) | pr;
                atomicMax((unsigned long long*)&bmax0[ab.x], (unsigned long long)t);
                if (bd) atomicMax((unsigned long long*)&bmax0[ab.y], (unsigned long long)t);
                fresh = true;
            }
        }
        // list entries of the manifolds to colour: a slot range per wave within the workgroup's
        const unsigned long long fb = __ballot(fresh);
        const uint32_t lane = __lane_id();
        uint32_t wbase = 0;
        if (fb && lane == 0u) wbase = atomicAdd(&s_n, (uint32_t)__popcll(fb));
        wbase = __shfl(wbase, 0, 64);
        __syncthreads();
        if (threadIdx.x == 0) {
            if (s_cmax) atomicMax(&cnt->max_colour, s_cmax);
            s_base = s_n ? atomicAdd(&cnt->remaining, s_n) : 0u;
        }
        __syncthreads();
        if (fresh) list[s_base + wbase + (uint32_t)__popcll(fb & ((1ull << lane) - 1ull))] = m;
        __syncthreads();  // s_base and s_n are reset by the next iteration
    }
}

// One Jones-Plassmann visit of list entry m in round r: a manifold holding the
// round's tag at all its dynamic bodies takes the first colour free in both
// bodies' 64-bit masks (64 + r once they are full); a loser offers its next tag
// into the other buffer.  Winners of one round share no dynamic body, so their
// mask updates cannot collide; tags grow with the round, so the buffers never
// need clearing.  AT: shared words through agent-scope atomics (the one-workgroup
// form, which has no cache invalidation between rounds).
template <bool AT>
__device__ __forceinline__ void jp_visit(const Bodies& B, const uint2* __restrict__ m_ab, uint32_t* __restrict__ col,
                                         const uint64_t* __restrict__ prio, const uint64_t* cur, uint64_t* nxt,
                                         uint64_t* bmask, uint32_t m, uint32_t r, uint32_t& won, uint32_t& cmax) {
    if (col[m] != UNSET) return;
    auto ld = [](const uint64_t* p) -> uint64_t {
        if constexpr (AT) return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        else return *p;
    };
    auto st = [](uint64_t* p, uint64_t v) {
        if constexpr (AT) __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        else *p = v;
    };
    const uint2 ab = m_ab[m];
    const bool bd = dyn_body(B, ab.y);
    const uint64_t t = ((uint64_t)(r + 1) << 48) | prio[m];
    if (ld(&cur[ab.x]) == t && (!bd || ld(&cur[ab.y]) == t)) {
        const uint64_t ma = ld(&bmask[ab.x]), mb = bd ? ld(&bmask[ab.y]) : 0ull;
        const uint64_t used = ma | mb;
        uint32_t c;
        if (~used == 0ull) {
            c = 64u + r;
        } else {
            c = (uint32_t)__ffsll((unsigned long long)~used) - 1u;
            st(&bmask[ab.x], ma | (1ull << c));
            if (bd) st(&bmask[ab.y], mb | (1ull << c));
        }
        col[m] = c;
        ++won;
        cmax = c + 1u > cmax ? c + 1u : cmax;
    } else {
        const uint64_t t2 = ((uint64_t)(r + 2) << 48) | prio[m];
        atomicMax((unsigned long long*)&nxt[ab.x], (unsigned long long)t2);
        if (bd) atomicMax((unsigned long long*)&nxt[ab.y], (unsigned long long)t2);
    }
}

// The rounds for a short list: one workgroup, a workgroup barrier per round
// (every wave first waits for its own stores: they are agent-scope and must have
// landed before another wave loads them).  Runs when 0 < list length <= small.
__global__ void __launch_bounds__(1024) k_colour_small(Bodies B, const uint2* __restrict__ m_ab,
                                                       uint32_t* __restrict__ col, const uint64_t* __restrict__ prio,
                                                       uint64_t* bmax0, uint64_t* bmax1, uint64_t* bmask,
                                                       const uint32_t* __restrict__ list, Counters* cnt,
                                                       uint32_t max_rounds, uint32_t small) {
    // per-round totals in three slots: round r adds into slot r % 3 and reads it after
    // barrier r; thread 0 then clears slot (r + 2) % 3, which nobody touches before
    // round r + 2 (it starts after barrier r + 1, which thread 0 reaches after the clear)
    __shared__ uint32_t s_won[3], s_cmax[3];
    const uint32_t L = cnt->remaining;
    if (L == 0u || L > small) return;  // uniform
    if (threadIdx.x < 3u) {
        s_won[threadIdx.x] = 0u;
        s_cmax[threadIdx.x] = 0u;
    }
    __syncthreads();
    uint32_t left = L;
    for (uint32_t r = 0;; ++r) {
        const uint32_t q = r % 3u;
        const uint64_t* cur = (r & 1u) ? bmax1 : bmax0;
        uint64_t* nxt = (r & 1u) ? bmax0 : bmax1;
        uint32_t won = 0, cmax = 0;
        for (uint32_t i = threadIdx.x; i < L; i += blockDim.x)
            jp_visit<true>(B, m_ab, col, prio, cur, nxt, bmask, list[i], r, won, cmax);
        if (won) {
            atomicAdd(&s_won[q], won);
            atomicMax(&s_cmax[q], cmax);
        }
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();
        const uint32_t w = s_won[q];
        if (threadIdx.x == 0) {
            if (w) {
                atomicMax(&cnt->max_colour, s_cmax[q]);
                atomicMax(&cnt->last_round, r + 1u);
            }
            s_won[(r + 2u) % 3u] = 0u;
            s_cmax[(r + 2u) % 3u] = 0u;
        }
        left -= w;
        if (left == 0u) break;
        if (r + 1 >= max_rounds) {
            if (threadIdx.x == 0) atomicOr(&cnt->sticky, STICKY_COLOURS);
            break;
        }
    }
    if (threadIdx.x == 0) cnt->remaining = 0u;
}

// The rounds for a long list (a full recolouring): persistent, one workgroup per
// CU, a grid barrier per round.  Runs when the list is longer than `small`.
__global__ void __launch_bounds__(256) k_colour(Bodies B, const uint2* __restrict__ m_ab, uint32_t* __restrict__ col,
                                                const uint64_t* __restrict__ prio, uint64_t* __restrict__ bmax0,
                                                uint64_t* __restrict__ bmax1, uint64_t* __restrict__ bmask,
                                                const uint32_t* __restrict__ list, GridBar* bar, Counters* cnt,
                                                uint32_t max_rounds, uint32_t small) {
    const uint32_t L = cnt->remaining;
    if (L <= small) return;  // uniform across the grid: no barrier is entered
    const uint32_t nb = gridDim.x;
    const uint32_t gs = nb * blockDim.x;
    const uint32_t g0 = blockIdx.x * blockDim.x + threadIdx.x;
    for (uint32_t r = 0;; ++r) {
        const uint64_t* cur = (r & 1u) ? bmax1 : bmax0;
        uint64_t* nxt = (r & 1u) ? bmax0 : bmax1;
        uint32_t won = 0, cmax = 0;
        for (uint32_t i = g0; i < L; i += gs) jp_visit<false>(B, m_ab, col, prio, cur, nxt, bmask, list[i], r, won, cmax);
        if (won) {
            atomicSub(&cnt->remaining, won);
            atomicMax(&cnt->max_colour, cmax);
            atomicMax(&cnt->last_round, r + 1u);
        }
        // the count left after round r, as it stood when every workgroup had finished round r
        const uint32_t left = grid_sync(bar, nb, cnt, &cnt->remaining);
        const uint32_t bad = __hip_atomic_load(&cnt->sticky, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (left == 0 || (bad & STICKY_TIMEOUT)) break;
        if (r + 1 >= max_rounds) {
            if (g0 == 0) atomicOr(&cnt->sticky, STICKY_COLOURS);
            break;
        }
    }
}

// Colour sort input: values = manifold index, keys = colour, or with the island split
// LOCAL_KEY | root << 8 | colour for a manifold of an island of at most ISLAND_LOCAL_MAX
// manifolds.  A stable sort then gives the solve order: the large islands' manifolds by
// colour, then the small islands one after another, each by colour (manifold index last).
__global__ void __launch_bounds__(256) k_col_sort_input(Bodies B, const uint2* __restrict__ m_ab,
                                                        const uint32_t* __restrict__ col, uint32_t cap,
                                                        uint32_t* __restrict__ key, uint32_t* __restrict__ val,
                                                        const Counters* __restrict__ cnt, uint32_t split) {
    const uint32_t M = min(cap, cnt->m_eff);
    for (uint32_t m = blockIdx.x * blockDim.x + threadIdx.x; m < M; m += gridDim.x * blockDim.x) {
        uint32_t k = col[m];
        if (split) {
            const uint32_t r = man_root(B, m_ab[m]);
            if (r != UNSET && B.icnt[r] <= ISLAND_LOCAL_MAX) k |= LOCAL_KEY | (r << 8);
        }
        key[m] = k;
        val[m] = m;
    }
}
// The work list of the islands solved by one workgroup each, from the sorted keys: (first
// slot, manifolds) of every small island of more than ISLAND_LANE_MAX manifolds; the smaller
// ones are only counted (the bundle kernel finds them by their slots).
// One pass over the sorted keys skey (thread gid of gsz, both whole waves; every lane of a
// wave takes part in each round): with off, the colour offsets off[c] = first position whose
// key has colour >= c (c = 0..K, a small island's key counting as colour K, since those
// keys sort above every colour) and m_glob = off[K]; with list, the work list of the small
// islands solved by one workgroup each (first slot, manifolds), the one-lane islands only
// counted (per wave).  An island starts where its key differs from the one before it above
// the colour bits.
__device__ __forceinline__ void sorted_keys_part(const Bodies& B, const uint32_t* __restrict__ skey,
                                                 uint32_t* __restrict__ off, uint2* __restrict__ list, uint32_t cap,
                                                 Counters* cnt, uint32_t gid, uint32_t gsz) {
    const uint32_t M = min(cap, cnt->m_eff), K = cnt->max_colour, lane = gid & 63u;
    auto colour_of = [K](uint32_t k) { return (k & LOCAL_KEY) != 0u ? K : k; };
    for (uint32_t base = gid - lane; base <= M; base += gsz) {
        const uint32_t s = base + lane;
        bool start = false;
        uint32_t size = 0;
        if (s <= M) {
            const uint32_t k = s < M ? skey[s] : 0u, kp = s > 0u ? skey[s - 1u] : 0u;
            if (off) {
                // the colours from just above the previous key's to this key's (to K at the end)
                const uint32_t c0 = s > 0u ? colour_of(kp) + 1u : 0u, c1 = s < M ? colour_of(k) : K;
                for (uint32_t c = c0; c <= c1; ++c) {
                    off[c] = s;
                    if (c == K) cnt->m_glob = s;
                }
            }
            if (list && s < M) {
                start = (k & LOCAL_KEY) != 0u && (s == 0u || ((kp ^ k) >> 8) != 0u);
                if (start) size = B.icnt[(k & ~LOCAL_KEY) >> 8];
            }
        }
        if (list) {
            const uint64_t one = __ballot(start && size <= ISLAND_LANE_MAX);
            if (lane == 0u && one) atomicAdd(&cnt->isl_lane_n, (uint32_t)__popcll(one));
            if (start && size > ISLAND_LANE_MAX) list[atomicAdd(&cnt->isl_group_n, 1u)] = make_uint2(s, size);
        }
    }
}

__global__ void k_iota(uint32_t* __restrict__ p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = i;
}


// ---------------------------------------------------------------- solver
struct SB {
    f3 v, w, x;
    float im;
    s3 I;
    bool dyn;
};

// A body's solver state for the row preparation.  The world inverse inertia is formed
// here, for the bodies of manifolds only (sv.w > 0 exactly when the body is dynamic and
// awake this step), rather than stored for every body at the step start.
__device__ __forceinline__ void load_sb(const Bodies& B, uint32_t i, SB& s, bool want_x) {
    const float4 v = B.sv[i], w = B.sw[i];
    s.v = xyz(v);
    s.im = v.w;
    s.w = xyz(w);
    s.dyn = v.w > 0.0f;
    s.I = s.dyn ? world_inv_inertia(quat_to_mat(B.q[i]), xyz(B.ii[i])) : s3{0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    if (want_x) s.x = xyz(B.x[i]);
}
__device__ __forceinline__ void zero_sb(SB& s) {
    s.v = s.w = s.x = f3{0.0f, 0.0f, 0.0f};
    s.im = 0.0f;
    s.I = s3{0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    s.dyn = false;
}

__device__ __forceinline__ void apply_imp(SB& A, SB& Bd, f3 rA, f3 rB, f3 P) {
    if (A.dyn) {
        A.v = A.v - P * A.im;
        A.w = A.w - smul(A.I, cross(rA, P));
    }
    if (Bd.dyn) {
        Bd.v = Bd.v + P * Bd.im;
        Bd.w = Bd.w + smul(Bd.I, cross(rB, P));
    }
}
__device__ __forceinline__ f3 vrel(const SB& A, const SB& Bd, f3 rA, f3 rB) {
    return (Bd.v + cross(Bd.w, rB)) - (A.v + cross(A.w, rA));
}

// Prepares solve slot s; returns the warm-start hits and the algorithmic bytes of one
// sweep visit of the slot (the kernel sums them per wave).
__device__ __forceinline__ void prep_slot(const Bodies& B, const StepParams& sp, const uint32_t* __restrict__ order,
                                          uint32_t cap, const uint2* __restrict__ m_ab,
                                          const uint32_t* __restrict__ m_np, const float4* __restrict__ c_pt,
                                          const float4* __restrict__ c_nf, const SolverRows& R, const PrevCache& pc,
                                          const float4* __restrict__ warm_in, const uint32_t* __restrict__ col,
                                          const uint64_t* __restrict__ bmask, const uint32_t* __restrict__ bdeg,
                                          Counters* __restrict__ cnt, float* __restrict__ feat, uint32_t s,
                                          uint32_t& hits_out, unsigned long long& bytes_out) {
    const uint32_t m = order[s];
    const uint2 ab = m_ab[m];
    const uint32_t np = m_np[m];
    const bool plane = (ab.y & PLANE_BIT) != 0;
    SB A, Bd;
    load_sb(B, ab.x, A, true);
    if (plane) zero_sb(Bd); else load_sb(B, ab.y, Bd, true);
    R.hdr[s] = make_uint4(ab.x, ab.y, np, m);
    {
        // dependency words for the persistent sweep: deg << 8 | rank per dynamic body
        const uint32_t c = col[m];
        R.colour[s] = c;
        uint2 dp = make_uint2(UNSET, UNSET);
        // (more than 64 colours: no dependency words; the sweep hands the step to per-colour launches)
        if (c < 64u && cnt->max_colour <= 64u) {
            // a body's colour mask must hold one bit per manifold of the body, or the
            // persistent sweep's visit counts could never be reached
            const unsigned long long below = (1ull << c) - 1ull;
            if (A.dyn) {
                const unsigned long long mk = bmask[ab.x];
                if ((uint32_t)__popcll(mk) != bdeg[ab.x]) atomicOr(&cnt->sticky, STICKY_INTERNAL);
                dp.x = ((uint32_t)__popcll(mk) << 8) | (uint32_t)__popcll(mk & below);
            }
            if (Bd.dyn) {
                const unsigned long long mk = bmask[ab.y];
                if ((uint32_t)__popcll(mk) != bdeg[ab.y]) atomicOr(&cnt->sticky, STICKY_INTERNAL);
                dp.y = ((uint32_t)__popcll(mk) << 8) | (uint32_t)__popcll(mk & below);
            }
            // who visits which body in which colour: the sweep's visits pass velocities on by it
            if (A.dyn) B.vslot[(size_t)ab.x * 64u + c] = s << 1;
            if (Bd.dyn) B.vslot[(size_t)ab.y * 64u + c] = (s << 1) | 1u;
        }
        R.dep[s] = dp;
    }
    uint32_t ps = 0;
    const bool have_prev = pc.valid && prev_find(pc, ((uint64_t)ab.x << 32) | ab.y, ps);
    // the pair's friction and restitution: mu = sqrt(mu_A mu_B) (mu_A when equal), e = max(e_A, e_B);
    // a plane has the world's values
    float mu, rest_e;
    {
        const float4 ma = B.mat[ab.x];
        const float mub = plane ? sp.friction : B.mat[ab.y].x, eb = plane ? sp.restitution : B.mat[ab.y].y;
        mu = ma.x == mub ? ma.x : sqrtf(ma.x * mub);
        rest_e = ma.y > eb ? ma.y : eb;
    }
    uint32_t hits = 0;
    for (uint32_t k = 0; k < np; ++k) {
        const float4 pt = c_pt[(size_t)m * 4 + k], nf = c_nf[(size_t)m * 4 + k];
        const f3 p = xyz(pt), n = xyz(nf);
        const float depth = pt.w;
        const uint32_t fid = fbits(nf.w);
        const f3 rA = p - A.x;
        const f3 rB = plane ? f3{0.0f, 0.0f, 0.0f} : p - Bd.x;
        f3 t1, t2;
        tangents(n, t1, t2);
        const f3 dirs[3] = {n, t1, t2};
        float kk[3];
        f3 ja[3], jb[3], jia[3], jib[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            ja[d] = cross(rA, dirs[d]);
            jb[d] = cross(rB, dirs[d]);
            jia[d] = smul(A.I, ja[d]);
            jib[d] = smul(Bd.I, jb[d]);
            kk[d] = (A.im + Bd.im + dot(ja[d], jia[d]) + dot(jb[d], jib[d])) * sp.soft1;
        }
        const float mN = kk[0] > 0.0f ? 1.0f / kk[0] : 0.0f;
        const float mT1 = kk[1] > 0.0f ? 1.0f / kk[1] : 0.0f;
        const float mT2 = kk[2] > 0.0f ? 1.0f / kk[2] : 0.0f;
        const f3 vr = vrel(A, Bd, rA, rB);
        const float vn0 = dot(vr, n);
        float bias = 0.0f;
        if (depth > sp.slop) {
            bias = sp.baumgarte * (depth - sp.slop) * sp.inv_dt;
            if (bias > sp.max_bias) bias = sp.max_bias;
        } else if (depth < 0.0f) {
            bias = depth * sp.inv_dt;
        }
        if (vn0 < -sp.rest_threshold) {
            const float rb = -rest_e * vn0;
            if (rb > bias) bias = rb;
        }
        const size_t o = (size_t)k * cap + s;
        float4* J = R.J + (size_t)k * JROWS * cap + s;
        J[0 * (size_t)cap] = f4(n, mN);
        J[1 * (size_t)cap] = f4(ja[0], kk[0]);
        J[2 * (size_t)cap] = f4(jb[0], 0.0f);
        J[3 * (size_t)cap] = f4(jia[0], bias);
        J[4 * (size_t)cap] = f4(jib[0], 0.0f);
        J[5 * (size_t)cap] = f4(t1, mT1);
        J[6 * (size_t)cap] = f4(ja[1], kk[1]);
        J[7 * (size_t)cap] = f4(jb[1], 0.0f);
        J[9 * (size_t)cap] = f4(jib[1], 0.0f);
        J[10 * (size_t)cap] = f4(t2, mT2);
        J[11 * (size_t)cap] = f4(ja[2], kk[2]);
        J[12 * (size_t)cap] = f4(jb[2], 0.0f);
        J[13 * (size_t)cap] = f4(jia[2], 0.0f);
        J[14 * (size_t)cap] = f4(jib[2], 0.0f);
        R.fid[o] = fid;
        // previous step's impulse of the same feature (the cached warm start)
        float4 Lp = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        bool hit = false;
        if (have_prev) {
            const uint32_t pnp = pc.np[ps];
            for (uint32_t j = 0; j < pnp; ++j) {
                const size_t po = (size_t)j * pc.cap + ps;
                if (pc.fid[po] == fid) {
                    Lp = pc.lam[po];
                    hit = true;
                    ++hits;
                    break;
                }
            }
        }
        {
            // friction bound of this step: mu x the previous normal impulse of the same feature,
            // or for a new point the impulse that stops its approach on its own
            const float est = mN * (bias - vn0);
            const float nb = hit ? Lp.x : est;
            J[8 * (size_t)cap] = f4(jia[1], mu * (nb > 0.0f ? nb : 0.0f));
        }
        // the given warm start, if any, else the cached one (chosen by value: a choice of
        // address would place Lp in scratch memory)
        float4 L = Lp;
        if (warm_in) {
            const float4 wv = warm_in[(size_t)m * 4 + k];
            L.x = wv.x;
            L.y = wv.y;
            L.z = wv.z;
        }
        L.w = 0.0f;
        R.lam[o] = L;
        R.lam0[o] = L;
        if (feat) {
            // warm-start network inputs, contact frame, impulses in units of u = mN g dt
            const float gdt = MLP_G_REF * sp.dt;
            const float inv_gdt = 1.0f / gdt;
            const float u = mN * gdt;
            const float inv_u = u > 0.0f ? 1.0f / u : 0.0f;
            float4* fr = (float4*)(feat + ((size_t)4 * s + k) * MLP_IN);
            fr[0] = make_float4(n.x, n.y, n.z, depth * 20.0f);
            fr[1] = make_float4(vn0 * inv_gdt, dot(vr, t1) * inv_gdt, dot(vr, t2) * inv_gdt, dot(rA, n));
            fr[2] = make_float4(dot(rA, t1), dot(rA, t2), dot(rB, n), dot(rB, t1));
            fr[3] = make_float4(dot(rB, t2), A.im * mN, Bd.im * mN, mT1 * kk[0]);
            fr[4] = make_float4(mT2 * kk[0], bias * inv_gdt, Lp.x * inv_u, Lp.y * inv_u);
            fr[5] = make_float4(Lp.z * inv_u, hit ? 1.0f : 0.0f, (float)np * 0.25f, plane ? 1.0f : 0.0f);
        }
    }
    hits_out = hits;
    // algorithmic bytes one sweep visit of this manifold moves: header 16; per
    // point 15 Jacobian rows (9 against a plane, whose b / ib rows are not read)
    // and lambda read + write; each body's solver velocities read (32) and, when
    // dynamic, written back (32); a plane has no body
    unsigned long long bytes = 16ull + (plane ? 9ull * 16ull + 32ull : 15ull * 16ull + 32ull) * np + 32ull +
                               (A.dyn ? 32ull : 0ull);
    if (!plane) bytes += 32ull + (Bd.dyn ? 32ull : 0ull);
    bytes_out = bytes;
}
__global__ void __launch_bounds__(256) k_prep(Bodies B, StepParams sp, const uint32_t* __restrict__ order,
                                              uint32_t cap, const uint2* __restrict__ m_ab,
                                              const uint32_t* __restrict__ m_np, const float4* __restrict__ c_pt,
                                              const float4* __restrict__ c_nf, SolverRows R, PrevCache pc,
                                              const float4* __restrict__ warm_in, const uint32_t* __restrict__ col,
                                              const uint64_t* __restrict__ bmask, const uint32_t* __restrict__ bdeg,
                                              Counters* __restrict__ cnt, float* __restrict__ feat, PrepExtra X) {
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x, gsz = gridDim.x * blockDim.x;
    for (uint32_t i = gid; i < X.zero_n; i += gsz) X.zero[i] = 0u;
    // more colours than the colour offsets hold fail the step (as a sticky error, which
    // this launch sets itself, so it reads the condition rather than the bit)
    const bool colours_ok = cnt->max_colour <= X.max_k;
    // a failed step must not touch the impulse buffers
    const bool ok = cnt->sticky == 0u && colours_ok;
    if (gid == 0 && !colours_ok) atomicOr(&cnt->sticky, STICKY_COLOURS);
    if (ok) {
        if (X.ckey) {
            const uint32_t m = cnt->cache_m[X.cset];
            for (uint32_t i = gid; i < m; i += gsz) X.ckey[X.cpos[i]] = EMPTY_KEY;
        }
        if (X.col_off || X.list) sorted_keys_part(B, X.skey, X.col_off, X.list, cap, cnt, gid, gsz);
    }
    uint32_t hits = 0;
    unsigned long long bytes = 0;
    const uint32_t M = ok ? cnt->m_eff : 0u;
    for (uint32_t s = gid; s < M; s += gsz) {
        uint32_t h = 0;
        unsigned long long b = 0;
        prep_slot(B, sp, order, cap, m_ab, m_np, c_pt, c_nf, R, pc, warm_in, col, bmask, bdeg, cnt, feat, s, h, b);
        hits += h;
        bytes += b;
    }
    spread_add(cnt, &SpreadRow::warm_hits, hits);
    spread_add(cnt, &SpreadRow::pgs_bytes, bytes);
}

// Memory policy of the sweep: plain accesses when every colour is its own
// launch; agent-coherent accesses (sc1, coherent across XCD L2s) in the
// persistent kernel, where bodies and impulses written by one workgroup are read
// by another inside the same launch.
template <bool COH> struct Mem {
    static AXM_D float4 ld(const float4* p) {
        if constexpr (COH) {
            const unsigned long long* q = reinterpret_cast<const unsigned long long*>(p);
            const unsigned long long lo = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const unsigned long long hi = __hip_atomic_load(q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return make_float4(__uint_as_float((uint32_t)lo), __uint_as_float((uint32_t)(lo >> 32)),
                               __uint_as_float((uint32_t)hi), __uint_as_float((uint32_t)(hi >> 32)));
        } else {
            return *p;
        }
    }
    static AXM_D void st(float4* p, float4 v) {
        if constexpr (COH) {
            unsigned long long* q = reinterpret_cast<unsigned long long*>(p);
            const unsigned long long lo = (unsigned long long)__float_as_uint(v.x) |
                                          ((unsigned long long)__float_as_uint(v.y) << 32);
            const unsigned long long hi = (unsigned long long)__float_as_uint(v.z) |
                                          ((unsigned long long)__float_as_uint(v.w) << 32);
            __hip_atomic_store(q, lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(q + 1, hi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            *p = v;
        }
    }
};

// Solver velocities of a body (the sweep needs no inertia: I a and I b are in the rows).
template <bool COH> __device__ __forceinline__ void load_vel(const Bodies& B, uint32_t i, SB& s) {
    const float4 v = Mem<COH>::ld(B.sv + i), w = Mem<COH>::ld(B.sw + i);
    s.v = xyz(v);
    s.im = v.w;
    s.w = xyz(w);
    s.dyn = v.w > 0.0f;
}

// One constraint direction of a contact point: d, a = rA x d, b = rB x d, ia = IA a, ib = IB b.
struct Dir {
    f3 d, a, b, ia, ib;
};
// dot product with fused multiply-adds: x.z y.z + (x.y y.y + x.x y.x)
__device__ __forceinline__ float fdot(f3 x, f3 y) { return fmaf(x.z, y.z, fmaf(x.y, y.y, x.x * y.x)); }
// relative velocity along the direction: d.(vB - vA) + b.wB - a.wA
__device__ __forceinline__ float jvel(const Dir& J, const SB& A, const SB& Bd) {
    const float r = fdot(J.d, Bd.v - A.v) + fdot(J.b, Bd.w);
    return r - fdot(J.a, A.w);
}
// impulse `lam` along the direction: A gets -lam, B +lam
__device__ __forceinline__ void japply(const Dir& J, float lam, SB& A, SB& Bd) {
    if (A.dyn) {
        const float sa = -(lam * A.im);
        A.v = f3{fmaf(sa, J.d.x, A.v.x), fmaf(sa, J.d.y, A.v.y), fmaf(sa, J.d.z, A.v.z)};
        A.w = f3{fmaf(-lam, J.ia.x, A.w.x), fmaf(-lam, J.ia.y, A.w.y), fmaf(-lam, J.ia.z, A.w.z)};
    }
    if (Bd.dyn) {
        const float sb = lam * Bd.im;
        Bd.v = f3{fmaf(sb, J.d.x, Bd.v.x), fmaf(sb, J.d.y, Bd.v.y), fmaf(sb, J.d.z, Bd.v.z)};
        Bd.w = f3{fmaf(lam, J.ib.x, Bd.w.x), fmaf(lam, J.ib.y, Bd.w.y), fmaf(lam, J.ib.z, Bd.w.z)};
    }
}
// japply without a branch per side, for visits whose lanes differ in which sides are
// dynamic.  A side that is not dynamic has zero inverse mass and zero inverse-inertia
// rows (its world inverse inertia is zero), so for a finite impulse its velocity changes
// at most in the sign of a zero component; callers never store such a side back.
__device__ __forceinline__ void japply_both(const Dir& J, float lam, SB& A, SB& Bd) {
    const float sa = -(lam * A.im);
    A.v = f3{fmaf(sa, J.d.x, A.v.x), fmaf(sa, J.d.y, A.v.y), fmaf(sa, J.d.z, A.v.z)};
    A.w = f3{fmaf(-lam, J.ia.x, A.w.x), fmaf(-lam, J.ia.y, A.w.y), fmaf(-lam, J.ia.z, A.w.z)};
    const float sb = lam * Bd.im;
    Bd.v = f3{fmaf(sb, J.d.x, Bd.v.x), fmaf(sb, J.d.y, Bd.v.y), fmaf(sb, J.d.z, Bd.v.z)};
    Bd.w = f3{fmaf(lam, J.ib.x, Bd.w.x), fmaf(lam, J.ib.y, Bd.w.y), fmaf(lam, J.ib.z, Bd.w.z)};
}
// Direction `dir` (0 n, 1 t1, 2 t2) of the point whose rows start at J, consecutive
// rows `rs` float4 apart; the .w of the first row is returned in w0.  Against a plane
// b and ib are zero and not read.
__device__ __forceinline__ Dir load_dir(const float4* J, size_t rs, int dir, bool plane, float& w0, float& w1,
                                        float& w2, float& w3, float& w4) {
    const float4* q = J + (size_t)(5 * dir) * rs;
    const float4 r0 = q[0], r1 = q[rs], r3 = q[3 * rs];
    Dir D;
    D.d = xyz(r0);
    D.a = xyz(r1);
    D.ia = xyz(r3);
    w0 = r0.w;
    w1 = r1.w;
    w3 = r3.w;
    if (plane) {
        D.b = D.ib = f3{0.0f, 0.0f, 0.0f};
        w2 = w4 = 0.0f;
    } else {
        const float4 r2 = q[2 * rs], r4 = q[4 * rs];
        D.b = xyz(r2);
        D.ib = xyz(r4);
        w2 = r2.w;
        w4 = r4.w;
    }
    return D;
}

// The projections of a visit, one instruction each: x clamped to [-m, m]
// (m >= 0) and x clamped below at zero.  For finite x they give the values of
// x < -m ? -m : (x > m ? m : x) and x > 0 ? x : 0, except that a zero may come out with the
// other sign; every later use of the impulse is a sum, product or comparison, so no later
// value differs.
__device__ __forceinline__ float clamp_sym(float x, float m) { return __builtin_amdgcn_fmed3f(x, -m, m); }
// One visit of a manifold on registers: warm start (first sweep only), friction
// rows of every point, then normal rows.  Point k's rows start at J + k * ps, rows
// `rs` apart; L[k] holds the point impulses in and out.
__device__ __forceinline__ void solve_rows(const float4* J, size_t rs, size_t ps, uint32_t np, bool plane, float keep,
                                           bool first, SB& A, SB& Bd, float4 L[4]) {
    float w0, w1, w2, w3, w4;
    if (first) {
#pragma unroll
        for (uint32_t k = 0; k < 4; ++k) {
            if (k >= np) break;
            const float4* P = J + (size_t)k * ps;
            const Dir Dn = load_dir(P, rs, 0, plane, w0, w1, w2, w3, w4);
            const Dir D1 = load_dir(P, rs, 1, plane, w0, w1, w2, w3, w4);
            const Dir D2 = load_dir(P, rs, 2, plane, w0, w1, w2, w3, w4);
            japply_both(Dn, L[k].x, A, Bd);
            japply_both(D1, L[k].y, A, Bd);
            japply_both(D2, L[k].z, A, Bd);
        }
    }
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) {
        if (k >= np) break;
        const float4* P = J + (size_t)k * ps;
        float maxf;  // the step's friction bound (row 8's .w)
        {
            float mT1;
            const Dir D1 = load_dir(P, rs, 1, plane, mT1, w1, w2, maxf, w4);
            float nl = fmaf(-mT1, jvel(D1, A, Bd), keep * L[k].y);
            nl = clamp_sym(nl, maxf);
            const float d = nl - L[k].y;
            L[k].y = nl;
            japply_both(D1, d, A, Bd);
        }
        {
            float mT2;
            const Dir D2 = load_dir(P, rs, 2, plane, mT2, w1, w2, w3, w4);
            float nl = fmaf(-mT2, jvel(D2, A, Bd), keep * L[k].z);
            nl = clamp_sym(nl, maxf);
            const float d = nl - L[k].z;
            L[k].z = nl;
            japply_both(D2, d, A, Bd);
        }
    }
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) {
        if (k >= np) break;
        const float4* P = J + (size_t)k * ps;
        float mN, x1, x2, tg, x4;
        const Dir Dn = load_dir(P, rs, 0, plane, mN, x1, x2, tg, x4);
        float nl = fmaf(mN, tg - jvel(Dn, A, Bd), keep * L[k].x);
        nl = fmaxf(nl, 0.0f);
        const float d = nl - L[k].x;
        L[k].x = nl;
        japply_both(Dn, d, A, Bd);
    }
}
// The same with every row of the visit already in registers (Jr[k][r]: row r of
// point k), so that a visit whose rows are not staged waits for one batch of loads
// instead of one per row group.  Same arithmetic, same results.  The b-side rows of a
// plane contact hold zeros (load_rows), the values the row loads above substitute, so
// no lane selects between the two.
__device__ __forceinline__ Dir dir_regs(const float4 (&Jr)[JROWS], int dir, float& w0, float& w3) {
    const float4 r0 = Jr[5 * dir], r1 = Jr[5 * dir + 1], r3 = Jr[5 * dir + 3];
    Dir D;
    D.d = xyz(r0);
    D.a = xyz(r1);
    D.ia = xyz(r3);
    w0 = r0.w;
    w3 = r3.w;
    D.b = xyz(Jr[5 * dir + 2]);
    D.ib = xyz(Jr[5 * dir + 4]);
    return D;
}
__device__ __forceinline__ void load_rows(const SolverRows& R, uint32_t s, uint32_t cap, uint32_t np, bool plane,
                                          float4 (&Jr)[4][JROWS]) {
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) {
        if (k >= np) break;
        const float4* src = R.J + (size_t)k * JROWS * cap + s;
#pragma unroll
        for (uint32_t r = 0; r < JROWS; ++r) {
            const bool bside = (r % 5u) == 2u || (r % 5u) == 4u;  // b and IB b: not read against a plane
            Jr[k][r] = (bside && plane) ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : src[(size_t)r * cap];
        }
    }
}
__device__ __forceinline__ void solve_regs(const float4 (&Jr)[4][JROWS], uint32_t np, float keep,
                                           bool first, SB& A, SB& Bd, float4 L[4]) {
    float w0, w3;
    if (first) {
#pragma unroll
        for (uint32_t k = 0; k < 4; ++k) {
            if (k >= np) break;
            const Dir Dn = dir_regs(Jr[k], 0, w0, w3);
            const Dir D1 = dir_regs(Jr[k], 1, w0, w3);
            const Dir D2 = dir_regs(Jr[k], 2, w0, w3);
            japply_both(Dn, L[k].x, A, Bd);
            japply_both(D1, L[k].y, A, Bd);
            japply_both(D2, L[k].z, A, Bd);
        }
    }
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) {
        if (k >= np) break;
        float maxf;  // the step's friction bound (row 8's .w)
        {
            float mT1;
            const Dir D1 = dir_regs(Jr[k], 1, mT1, maxf);
            float nl = fmaf(-mT1, jvel(D1, A, Bd), keep * L[k].y);
            nl = clamp_sym(nl, maxf);
            const float d = nl - L[k].y;
            L[k].y = nl;
            japply_both(D1, d, A, Bd);
        }
        {
            float mT2;
            const Dir D2 = dir_regs(Jr[k], 2, mT2, w3);
            float nl = fmaf(-mT2, jvel(D2, A, Bd), keep * L[k].z);
            nl = clamp_sym(nl, maxf);
            const float d = nl - L[k].z;
            L[k].z = nl;
            japply_both(D2, d, A, Bd);
        }
    }
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) {
        if (k >= np) break;
        float mN, tg;
        const Dir Dn = dir_regs(Jr[k], 0, mN, tg);
        float nl = fmaf(mN, tg - jvel(Dn, A, Bd), keep * L[k].x);
        nl = fmaxf(nl, 0.0f);
        const float d = nl - L[k].x;
        L[k].x = nl;
        japply_both(Dn, d, A, Bd);
    }
}
// The same with the rows of solve slot s in the SoA row arrays (row r of point k at
// J[(k * JROWS + r) * cap + s]).
__device__ __forceinline__ void solve_core(const SolverRows& R, uint32_t s, uint32_t cap, uint32_t np, bool plane,
                                           float keep, bool first, SB& A, SB& Bd, float4 L[4]) {
    solve_rows(R.J + s, cap, (size_t)JROWS * cap, np, plane, keep, first, A, Bd, L);
}

// One visit of the manifold in solve slot s with plain loads and stores (one
// launch per colour: every visit of a launch touches distinct bodies).
__device__ __forceinline__ void solve_slot(const Bodies& B, const StepParams& sp, const SolverRows& R, uint4 h,
                                           uint32_t s, uint32_t cap, bool first) {
    const uint32_t a = h.x, b = h.y, np = h.z;
    const bool plane = (b & PLANE_BIT) != 0;
    SB A, Bd;
    load_vel<false>(B, a, A);
    if (plane) zero_sb(Bd); else load_vel<false>(B, b, Bd);
    float4 L[4];
    for (uint32_t k = 0; k < np; ++k) L[k] = R.lam[(size_t)k * cap + s];
    solve_core(R, s, cap, np, plane, sp.keep, first, A, Bd, L);
    for (uint32_t k = 0; k < np; ++k) R.lam[(size_t)k * cap + s] = make_float4(L[k].x, L[k].y, L[k].z, 0.0f);
    if (A.dyn) {
        B.sv[a] = f4(A.v, A.im);
        B.sw[a] = f4(A.w, 0.0f);
    }
    if (Bd.dyn) {
        B.sv[b] = f4(Bd.v, Bd.im);
        B.sw[b] = f4(Bd.w, 0.0f);
    }
}

__global__ void __launch_bounds__(256) k_pgs(Bodies B, StepParams sp, SolverRows R, uint32_t s0, uint32_t s1,
                                             uint32_t cap, uint32_t first) {
    const uint32_t s = s0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= s1) return;
    solve_slot(B, sp, R, R.hdr[s], s, cap, first != 0);
}

// Persistent sweep kernel.  An item is the visit of one manifold slot in one
// sweep.  It may run once every earlier visit of its dynamic bodies is complete:
// the visit of slot s in sweep k is visit number k * deg(x) + rank(x, s) of body
// x, with rank = colours of x's manifolds below s's colour (the per-body colour
// masks of the colouring), so each body sees exactly the update sequence of the
// colour-by-colour sweep without a grid-wide barrier.
//
// Ownership: the slots are cut into chunks of 64; workgroup g (one wave) owns
// chunks g, g + G, g + 2G, ... and lane l of it slot 64 * chunk + l, in every
// sweep.  A lane walks its own items in increasing (sweep, slot) order on its
// own, so the smallest unfinished item of the launch always has its
// dependencies complete and the walk cannot deadlock while all G waves are
// resident (the host sizes G from the occupancy of the kernel).  Because a lane
// is the only one to visit its slots, the Jacobian rows and the impulses of its
// points are staged once per launch into LDS (up to `ppl` points per lane, the
// rest is read from and written to the row arrays) and never go back to memory
// between sweeps; only body velocities move between waves.
//
// Inboxes: during the launch a body's solver velocity travels from one visit of
// the body to the next.  Each slot has an inbox of 8 64-bit words (64 bytes) holding,
// for the slot's next visit, body a's velocity (v, w: 192 bits) in words 0..3 and
// body b's in words 4..7, each word 48 value bits under a 16-bit version: the body's
// visits completed, modulo 2^16.  An inbox only ever holds the version before the
// one its reader waits for (deg visits earlier, deg <= 64) or that one, so 16 bits
// never alias.  A visit waits until its inbox words carry the visit numbers it needs
// (a 64-bit word is single-copy atomic, so matching versions mean matching values),
// solves, and writes each dynamic body's new velocity into the inbox of that body's
// next visit (the manifold of the next colour in the body's colour mask, or its first
// one for the next sweep) with plain stores (two 16-byte ones where the inboxes lie
// within 2 GiB).  The last word of each side (3 for a, 7 for b, in the last store a
// writer makes) is a hint: a poll reads the two hints and loads the rest once both are
// in; the full load checks every version itself.
// A bounded wait sets STICKY_TIMEOUT instead of hanging.
constexpr uint32_t IB_WORDS = INBOX_WORDS;  // inbox words per slot
constexpr uint32_t IB_SIDE = 4;     // words per body
constexpr uint32_t IB_HINT = 3;     // word 3 of each side (polled first), stored with word 2
constexpr uint32_t OWN_WAVE = 64;   // threads per workgroup of the sweep kernel
// residual check word: workgroups arrived in the low 16 bits, workgroups above the tolerance above them
constexpr uint32_t CHECK_ABOVE = 1u << 16;
constexpr uint32_t NO_VISIT = 0xFFFFFFFFu;  // version no visit count reaches

// A velocity (v, w) as four inbox words under version ver (low 16 bits): the six
// floats' 192 bits in 48-bit pieces, x0 | x1[0:16], x1[16:32] | x2, x3 | x4[0:16],
// x4[16:32] | x5.
__device__ __forceinline__ void vpack4(f3 v, f3 w, uint32_t ver, uint64_t (&o)[4]) {
    const uint64_t V = (uint64_t)(ver & 0xFFFFu) << 48;
    const uint32_t x0 = __float_as_uint(v.x), x1 = __float_as_uint(v.y), x2 = __float_as_uint(v.z);
    const uint32_t x3 = __float_as_uint(w.x), x4 = __float_as_uint(w.y), x5 = __float_as_uint(w.z);
    o[0] = V | ((uint64_t)(x1 & 0xFFFFu) << 32) | x0;
    o[1] = V | ((uint64_t)x2 << 16) | (x1 >> 16);
    o[2] = V | ((uint64_t)(x4 & 0xFFFFu) << 32) | x3;
    o[3] = V | ((uint64_t)x5 << 16) | (x4 >> 16);
}
// The same words as (low, high) dword pairs: the version of a word is high >> 16.
__device__ __forceinline__ uint32_t vver2(uint2 x) { return x.y >> 16; }
__device__ __forceinline__ void vunpack4(uint2 a, uint2 b, uint2 c, uint2 d, f3& v, f3& w) {
    v = f3{__uint_as_float(a.x), __uint_as_float((a.y & 0xFFFFu) | (b.x << 16)),
           __uint_as_float((b.x >> 16) | (b.y << 16))};
    w = f3{__uint_as_float(c.x), __uint_as_float((c.y & 0xFFFFu) | (d.x << 16)),
           __uint_as_float((d.x >> 16) | (d.y << 16))};
}
__device__ __forceinline__ uint2 vsplit(uint64_t x) { return make_uint2((uint32_t)x, (uint32_t)(x >> 32)); }
__device__ __forceinline__ uint64_t vload(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void vstore(uint64_t* p, uint64_t w) {
    __hip_atomic_store(p, w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Word k (0..3) of side `side` of the inbox of slot s.
__device__ __forceinline__ size_t vin_word(uint32_t s, uint32_t side, uint32_t k) {
    return (size_t)s * IB_WORDS + side * IB_SIDE + k;
}
// Writes a velocity into side `side` of slot s's inbox, the hint word last.  With
// `wide` (inboxes within 2 GiB) words 0-1 and 2-3 go out as one 16-byte store each,
// with the cache policy of a device-scope store: two writes instead of four.  Every
// word carries its own version, so a reader that sees one half of a pair before the
// other only waits.
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void vin_store(uint64_t* inbox, uint32_t s, uint32_t side, uint32_t ver, f3 v, f3 w,
                                          bool wide) {
    uint64_t o[4];
    vpack4(v, w, ver, o);
    if (wide) {
        const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(inbox, 0, 0x7FFFFFFF, 0x00020000);
        const uint32_t base = s * (IB_WORDS * 8u) + side * (IB_SIDE * 8u);  // byte offset of word 0 of this side
        const u32x4_t q0 = {(uint32_t)o[0], (uint32_t)(o[0] >> 32), (uint32_t)o[1], (uint32_t)(o[1] >> 32)};
        const u32x4_t q1 = {(uint32_t)o[2], (uint32_t)(o[2] >> 32), (uint32_t)o[3], (uint32_t)(o[3] >> 32)};
        // aux: sc1, as the device-scope atomic store below
        __builtin_amdgcn_raw_buffer_store_b128(q0, rs, (int)base, 0, 16);
        __builtin_amdgcn_raw_buffer_store_b128(q1, rs, (int)(base + 16u), 0, 16);
        return;
    }
#pragma unroll
    for (uint32_t k = 0; k < IB_SIDE; ++k) vstore(inbox + vin_word(s, side, k), o[k]);
}
// Words k and k + 1 of slot s's inbox in one coherent 16-byte access, as (value,
// version) dword pairs; every word is checked against its own version, so a load
// that sees one word before the other only delays the visit.  Inboxes within 2 GiB.
__device__ __forceinline__ uint4 vin_quad(const uint64_t* inbox, uint32_t s, uint32_t k) {
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<uint64_t*>(inbox), 0, 0x7FFFFFFF, 0x00020000);
    // aux: sc0 | sc1, the cache policy of a system-coherent load (bypasses the non-coherent caches)
    const auto r = __builtin_amdgcn_raw_buffer_load_b128(rs, s * (IB_WORDS * 8u) + k * 8u, 0, 1 | 16);
    return make_uint4(r[0], r[1], r[2], r[3]);
}
// The visit after the one of colour c of a body with colour mask mk: the next
// colour up, or the body's first colour (next sweep).
__device__ __forceinline__ uint32_t next_colour(unsigned long long mk, uint32_t c) {
    const unsigned long long above = c >= 63u ? 0ull : mk & ~((2ull << c) - 1ull);
    return (uint32_t)__ffsll(above ? above : mk) - 1u;
}

// Before a launch, per slot: the next visit of each dynamic body, and the inbox
// words: a body's first visit (rank 0) gets the body's solver velocity with
// version 0, every other inbox a version no visit waits for.
// (Thread gid of gsz; also run by the small islands' workgroup kernel, k_pgs_group, whose
// bodies and slots are disjoint from the large islands'.)
__device__ __forceinline__ void flow_init_part(const Bodies& B, const SolverRows& R, const uint64_t* __restrict__ bmask,
                                               const Counters* __restrict__ cnt, uint32_t gid, uint32_t gsz) {
    // the rows exist only when this step's preparation ran
    if (cnt->sticky != 0u) return;
    const uint32_t M = cnt->m_glob;  // the large islands' slots
    for (uint32_t s = gid; s < M; s += gsz) {
    const uint4 h = R.hdr[s];
    const uint2 dp = R.dep[s];
    const uint32_t c = R.colour[s];
    uint2 nx = make_uint2(UNSET, UNSET);
    f3 va = {0.0f, 0.0f, 0.0f}, wa = va, vb = va, wb = va;
    uint32_t ra = NO_VISIT, rb = NO_VISIT;
    if (dp.x != UNSET) {
        nx.x = B.vslot[(size_t)h.x * 64u + next_colour(bmask[h.x], c)];
        if ((dp.x & 0xFFu) == 0u) ra = 0u;
        va = xyz(B.sv[h.x]);
        wa = xyz(B.sw[h.x]);
    }
    if (dp.y != UNSET) {
        nx.y = B.vslot[(size_t)h.y * 64u + next_colour(bmask[h.y], c)];
        if ((dp.y & 0xFFu) == 0u) rb = 0u;
        vb = xyz(B.sv[h.y]);
        wb = xyz(B.sw[h.y]);
    }
    R.next[s] = nx;
    // the whole inbox (words 0-3 and 4-7 as laid out by vin_word) in four 16-byte
    // stores; the sweep kernel, a later launch, sees them after the launch boundary
    uint64_t oa[4], ob[4];
    vpack4(va, wa, ra, oa);
    vpack4(vb, wb, rb, ob);
    uint4* q = reinterpret_cast<uint4*>(R.inbox + (size_t)s * IB_WORDS);
    q[0] = make_uint4((uint32_t)oa[0], (uint32_t)(oa[0] >> 32), (uint32_t)oa[1], (uint32_t)(oa[1] >> 32));
    q[1] = make_uint4((uint32_t)oa[2], (uint32_t)(oa[2] >> 32), (uint32_t)oa[3], (uint32_t)(oa[3] >> 32));
    q[2] = make_uint4((uint32_t)ob[0], (uint32_t)(ob[0] >> 32), (uint32_t)ob[1], (uint32_t)(ob[1] >> 32));
    q[3] = make_uint4((uint32_t)ob[2], (uint32_t)(ob[2] >> 32), (uint32_t)ob[3], (uint32_t)(ob[3] >> 32));
    }
}
__global__ void __launch_bounds__(256) k_flow_init(Bodies B, SolverRows R, const uint64_t* __restrict__ bmask,
                                                   const Counters* __restrict__ cnt) {
    flow_init_part(B, R, bmask, cnt, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
}
// After a launch: each dynamic body's final velocity is in the inbox of its first
// visit (the last visit of the last sweep passed it on there).
__global__ void __launch_bounds__(256) k_flow_finish(Bodies B, SolverRows R, const Counters* __restrict__ cnt) {
    if (cnt->sticky != 0u) return;
    const uint32_t M = cnt->m_glob;
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < M; s += gridDim.x * blockDim.x) {
    const uint4 h = R.hdr[s];
    const uint2 dp = R.dep[s];
    const uint64_t* ib = R.inbox;
    if (dp.x != UNSET && (dp.x & 0xFFu) == 0u) {
        f3 v, w;
        vunpack4(vsplit(ib[vin_word(s, 0, 0)]), vsplit(ib[vin_word(s, 0, 1)]), vsplit(ib[vin_word(s, 0, 2)]),
                 vsplit(ib[vin_word(s, 0, 3)]), v, w);
        B.sv[h.x] = f4(v, B.sv[h.x].w);
        B.sw[h.x] = f4(w, 0.0f);
    }
    if (dp.y != UNSET && (dp.y & 0xFFu) == 0u) {
        f3 v, w;
        vunpack4(vsplit(ib[vin_word(s, 1, 0)]), vsplit(ib[vin_word(s, 1, 1)]), vsplit(ib[vin_word(s, 1, 2)]),
                 vsplit(ib[vin_word(s, 1, 3)]), v, w);
        B.sv[h.y] = f4(v, B.sv[h.y].w);
        B.sw[h.y] = f4(w, 0.0f);
    }
    }
}

// Residual of one contact point: the largest velocity-space change (impulse change
// times the row's effective mass inverse) that one more visit would make along the
// normal or either friction direction, from the current velocities and impulses L.
// J points at the point's first Jacobian row, rs is the row stride.
__device__ __forceinline__ float point_residual(const float4* J, size_t rs, bool plane, float keep, const SB& A,
                                                const SB& Bd, const float4 L) {
    float mN, kN, tg, mT1, kT1, mT2, kT2, x2, x3, x4;
    const Dir Dn = load_dir(J, rs, 0, plane, mN, kN, x2, tg, x4);
    float maxf;  // the step's friction bound (row 8's .w)
    const Dir D1 = load_dir(J, rs, 1, plane, mT1, kT1, x2, maxf, x4);
    const Dir D2 = load_dir(J, rs, 2, plane, mT2, kT2, x2, x3, x4);
    float nl = fmaf(mN, tg - jvel(Dn, A, Bd), keep * L.x);
    nl = fmaxf(nl, 0.0f);
    float e = fabsf(nl - L.x) * kN;
    float q1 = fmaf(-mT1, jvel(D1, A, Bd), keep * L.y);
    q1 = clamp_sym(q1, maxf);
    const float e1 = fabsf(q1 - L.y) * kT1;
    float q2 = fmaf(-mT2, jvel(D2, A, Bd), keep * L.z);
    q2 = clamp_sym(q2, maxf);
    const float e2 = fabsf(q2 - L.z) * kT2;
    if (e1 > e) e = e1;
    if (e2 > e) e = e2;
    return e;
}

// The same with the point's rows in registers (as solve_regs).
__device__ __forceinline__ float point_residual_regs(const float4 (&Jr)[JROWS], float keep, const SB& A,
                                                     const SB& Bd, const float4 L) {
    float mN, tg, mT1, maxf, mT2, x3;
    const Dir Dn = dir_regs(Jr, 0, mN, tg);
    const Dir D1 = dir_regs(Jr, 1, mT1, maxf);
    const Dir D2 = dir_regs(Jr, 2, mT2, x3);
    const float kN = Jr[1].w, kT1 = Jr[6].w, kT2 = Jr[11].w;
    float nl = fmaf(mN, tg - jvel(Dn, A, Bd), keep * L.x);
    nl = fmaxf(nl, 0.0f);
    float e = fabsf(nl - L.x) * kN;
    float q1 = fmaf(-mT1, jvel(D1, A, Bd), keep * L.y);
    q1 = clamp_sym(q1, maxf);
    const float e1 = fabsf(q1 - L.y) * kT1;
    float q2 = fmaf(-mT2, jvel(D2, A, Bd), keep * L.z);
    q2 = clamp_sym(q2, maxf);
    const float e2 = fabsf(q2 - L.z) * kT2;
    if (e1 > e) e = e1;
    if (e2 > e) e = e2;
    return e;
}

// What a visit needs of its slot besides the rows: the slot, body indices and point count,
// dependency words, the next visits of its bodies, inverse masses and the
// velocity of a static partner (constant during the sweep).
struct Item {
    uint32_t s;  // the slot
    uint4 h;
    uint2 dp, nx;
    float imA, imB;
    f3 sv, sw;  // static partner's velocity (at most one side of a manifold is static)
};
__device__ __forceinline__ Item item_global(const Bodies& B, const SolverRows& R, uint32_t s) {
    Item it;
    it.s = s;
    it.h = R.hdr[s];
    it.dp = R.dep[s];
    it.nx = R.next[s];
    const bool plane = (it.h.y & PLANE_BIT) != 0u;
    const float4 va = B.sv[it.h.x];
    const float4 vb = plane ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : B.sv[it.h.y];
    it.imA = va.w;
    it.imB = vb.w;
    const uint32_t st = it.dp.x == UNSET ? it.h.x : it.h.y;  // the static side, if any
    const bool has_static = it.dp.x == UNSET || (!plane && it.dp.y == UNSET);
    it.sv = has_static ? xyz(it.dp.x == UNSET ? va : vb) : f3{0.0f, 0.0f, 0.0f};
    it.sw = has_static ? xyz(B.sw[st]) : f3{0.0f, 0.0f, 0.0f};
    return it;
}
// staged slot record: 4 float4 at rec[r * 64], r = 0..3
__device__ __forceinline__ void item_put(float4* rec, const Item& it) {
    rec[0] = make_float4(__uint_as_float(it.h.x), __uint_as_float(it.h.y), __uint_as_float(it.h.z),
                         __uint_as_float(it.s));
    rec[OWN_WAVE] = make_float4(__uint_as_float(it.dp.x), __uint_as_float(it.dp.y), __uint_as_float(it.nx.x),
                                __uint_as_float(it.nx.y));
    rec[2 * OWN_WAVE] = make_float4(it.imA, it.imB, it.sv.x, it.sv.y);
    rec[3 * OWN_WAVE] = make_float4(it.sv.z, it.sw.x, it.sw.y, it.sw.z);
}
__device__ __forceinline__ Item item_get(const float4* rec) {
    const float4 r0 = rec[0], r1 = rec[OWN_WAVE], r2 = rec[2 * OWN_WAVE], r3 = rec[3 * OWN_WAVE];
    Item it;
    it.h = make_uint4(__float_as_uint(r0.x), __float_as_uint(r0.y), __float_as_uint(r0.z), 0u);
    it.s = __float_as_uint(r0.w);
    it.dp = make_uint2(__float_as_uint(r1.x), __float_as_uint(r1.y));
    it.nx = make_uint2(__float_as_uint(r1.z), __float_as_uint(r1.w));
    it.imA = r2.x;
    it.imB = r2.y;
    it.sv = f3{r2.z, r2.w, r3.x};
    it.sw = f3{r3.y, r3.z, r3.w};
    return it;
}

constexpr uint32_t OWN_SLOTS = 16;  // slots per lane with a staged record

// LDS of a wave: records of its first min(owned chunks, OWN_SLOTS) slots per lane
// (as many as fit; record j of a lane at lds[(j * 4 + r) * 64 + lane]), then the
// staged points of the slots chosen for staging, packed in slot order: point q of a
// lane keeps row r at lds[P0 + (r * ppl + q) * 64 + lane] (rows 0..JROWS-1 the
// Jacobian rows of R.J[(k * JROWS + r) * cap + s], row JROWS its impulses).  Slots
// and points not staged are read from the row arrays.  With stop.bits, the sweeps
// run in chunks of stop.every with a residual check after each chunk (every wave of
// the grid meets at one arrival count per check).
__global__ void __launch_bounds__(OWN_WAVE) k_pgs_own(Bodies B, StepParams sp, SolverRows R, Counters* cnt,
                                                      uint32_t cap, uint32_t sweep0, uint32_t nsweeps,
                                                      uint32_t lds_f4, uint32_t spin_limit, uint32_t tune,
                                                      const uint64_t* __restrict__ bmask, SweepStop stop) {
    extern __shared__ float4 own_lds[];
    const uint32_t lane = threadIdx.x;
    const uint32_t M = cnt->m_glob;  // the large islands' slots (the small ones are solved on their own)
    if (cnt->max_colour > 64u) {
        // dependency ranks come from 64-bit colour masks: hand the step back for per-colour launches
        if (blockIdx.x == 0 && lane == 0) atomicOr(&cnt->sticky, STICKY_FLOW_COLOURS);
        return;
    }
    if (cnt->sticky != 0u || nsweeps == 0u) return;
    if (M == 0u && cnt->m_eff > 0u) {
        // every island was solved on its own: no sweep of the large ones runs
        if (blockIdx.x == 0 && lane == 0 && stop.bits) {
            __hip_atomic_store(&cnt->sweeps_run, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&cnt->tol_hit, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        return;
    }
    uint32_t* err = &cnt->sticky;
    const uint32_t G = gridDim.x, g = blockIdx.x;
    const uint32_t nchunks = (M + OWN_WAVE - 1) / OWN_WAVE;
    const uint32_t own = g < nchunks ? (nchunks - 1 - g) / G + 1 : 0u;
    // owned slots of this lane: j < nv (only the last owned chunk can be partial)
    const uint32_t nv = own == 0u ? 0u : own - (((g + (own - 1) * G) * OWN_WAVE + lane) >= M ? 1u : 0u);
    // slot records that fit the LDS (the rest are read from the row arrays)
    const uint32_t nrec = min(min(own, OWN_SLOTS), lds_f4 / (4u * OWN_WAVE));
    const uint32_t p0 = nrec * 4u * OWN_WAVE;
    const uint32_t ppl = lds_f4 > p0 ? (lds_f4 - p0) / ((JROWS + 1u) * OWN_WAVE) : 0u;
    const size_t rs = (size_t)ppl * OWN_WAVE;  // LDS row stride of the points (float4)
    float4* recs = own_lds + lane;
    float4* mine = own_lds + p0 + lane;
    auto slot_of = [&](uint32_t jj) { return (g + jj * G) * OWN_WAVE + lane; };

    // stage slot records (only owned slots below M exist: j < nv)
    const uint32_t nr = min(nrec, nv);
    for (uint32_t j = 0; j < nr; ++j) item_put(recs + (size_t)j * 4u * OWN_WAVE, item_global(B, R, slot_of(j)));
    // Which owned slots get their rows and impulses staged (bit j of smask): as many
    // as fit, among the slots with a staged record those on the longest dependency
    // chains first (the larger visit count per sweep of their two bodies), then the
    // other slots in order.  Only a lane's first 64 slots can be staged.
    uint64_t smask = 0;
    {
        auto rec_np_crit = [&](uint32_t j, uint32_t& np, uint32_t& crit) {
            const float4 r0 = recs[(size_t)j * 4u * OWN_WAVE], r1 = recs[(size_t)(j * 4u + 1u) * OWN_WAVE];
            const uint32_t dx = __float_as_uint(r1.x), dy = __float_as_uint(r1.y);
            np = min(__float_as_uint(r0.z), 4u);
            crit = max(dx == UNSET ? 0u : dx >> 8, dy == UNSET ? 0u : dy >> 8);
        };
        uint32_t used = 0, level = 0xFFFFFFFFu;
        for (;;) {  // criticality levels, descending
            uint32_t next = 0;
            bool any = false;
            for (uint32_t j = 0; j < nr; ++j) {
                uint32_t np, c;
                rec_np_crit(j, np, c);
                if (c < level && (!any || c > next)) {
                    next = c;
                    any = true;
                }
            }
            if (!any) break;
            for (uint32_t j = 0; j < nr; ++j) {
                uint32_t np, c;
                rec_np_crit(j, np, c);
                if (c == next && used + np <= ppl) {
                    smask |= 1ull << j;
                    used += np;
                }
            }
            level = next;
        }
        for (uint32_t j = nr; j < min(nv, 64u); ++j) {
            const uint32_t np = min(R.hdr[slot_of(j)].z, 4u);
            if (used + np <= ppl) {
                smask |= 1ull << j;
                used += np;
            }
        }
    }
    auto staged_slot = [&](uint32_t j) { return j < 64u && ((smask >> j) & 1ull) != 0ull; };
    const bool all_staged = nv <= nrec && smask == (1ull << nv) - 1ull;  // nrec <= 16
    // stage the selected slots' rows and impulses, packed in slot order
    {
        uint32_t off = 0;
        for (uint32_t j = 0; j < nv; ++j) {
            if (!staged_slot(j)) continue;
            const uint32_t s = slot_of(j);
            const uint32_t np = min(j < nrec ? __float_as_uint(recs[(size_t)j * 4u * OWN_WAVE].z) : R.hdr[s].z, 4u);
            for (uint32_t k = 0; k < np; ++k) {
                const float4* src = R.J + (size_t)k * JROWS * cap + s;
                float4* dst = mine + (size_t)(off + k) * OWN_WAVE;
#pragma unroll
                for (uint32_t r = 0; r < JROWS; ++r) dst[r * rs] = src[(size_t)r * cap];
                dst[JROWS * rs] = R.lam[(size_t)k * cap + s];
            }
            off += np;
        }
    }
    // The visits of a wave whose slots and points are all in LDS read nothing but
    // the inbox words from memory; the general loop reads what is not staged from
    // the row arrays.  Two instances, so that the first never waits on a memory
    // load other than its poll (a wait for a load also waits for the stores before it).
    const float keep = sp.keep;
    // tune bits 4-6: poll back-off level; bit 1: poll the two hint words before the full load
    // (inboxes within 2 GiB); bit 2: adaptive back-off
    const uint32_t nap = (tune >> 4) & 7u;
    const bool wide = (uint64_t)cap * IB_WORDS * 8u <= 0x7FFFFFFFull;  // 16-byte inbox loads
    const bool hint = (tune & 2u) != 0u && wide, adaptive = (tune & 4u) != 0u;
    const uint64_t t_start = wall_clock64();
    const uint64_t t_limit = (uint64_t)spin_limit * 100ull;  // wall clock: 100 MHz

    // Final velocity of dynamic body `body` after `sweeps` sweeps: the inbox of its
    // first visit (lowest colour) once it holds version sweeps * deg.  No visit of a
    // later sweep runs before the check's grid barrier, so the words stay put.
    auto final_vel = [&](uint32_t body, uint32_t dpw, uint32_t sweeps, SB& X) -> bool {
        const uint32_t v0 = B.vslot[(size_t)body * 64u + (uint32_t)__ffsll((unsigned long long)bmask[body]) - 1u];
        const uint32_t want = (sweeps * (dpw >> 8)) & 0xFFFFu, vs = v0 >> 1, side = v0 & 1u;
        const uint64_t* ib = R.inbox;
        for (;;) {
            const uint2 a0 = vsplit(vload(ib + vin_word(vs, side, 0))), a1 = vsplit(vload(ib + vin_word(vs, side, 1))),
                        a2 = vsplit(vload(ib + vin_word(vs, side, 2))), a3 = vsplit(vload(ib + vin_word(vs, side, 3)));
            const uint32_t d = (vver2(a0) ^ want) | (vver2(a1) ^ want) | (vver2(a2) ^ want) | (vver2(a3) ^ want);
            if (d == 0u) {
                vunpack4(a0, a1, a2, a3, X.v, X.w);
                return true;
            }
            if (wall_clock64() - t_start > t_limit) {
                atomicOr(err, STICKY_TIMEOUT);
                return false;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    };
    // Residual check after `sweeps` sweeps: the k_residual value over this lane's
    // slots (rows and impulses where they are staged), wave max, one atomic max per
    // wave into stop.bits[sweeps - 1] (the history the host reads), and one atomic
    // add per wave into its arrival row of the check that carries both the arrival
    // (low 16 bits) and whether this wave's maximum is above the tolerance (bit 16
    // up): the decision needs no second round trip for the maximum.  Once the rows
    // sum to every workgroup, the launch stops when no wave was above the tolerance
    // (the same test as maximum <= tol).  Everything the waves exchange here and in
    // the sweep goes through device-scope atomics, so the barrier needs no cache
    // writeback or invalidation.  Returns 1 to stop, 0 to go on, -1 on an error.
    auto check = [&](uint32_t sweeps) -> int {
        float res = 0.0f;
        uint32_t off = 0;
        for (uint32_t j = 0; j < nv; ++j) {
            const Item it = j < nrec ? item_get(recs + (size_t)j * 4u * OWN_WAVE) : item_global(B, R, slot_of(j));
            const bool plane = (it.h.y & PLANE_BIT) != 0u;
            const uint32_t np = min(it.h.z, 4u);
            SB A, Bd;
            zero_sb(A);
            zero_sb(Bd);
            A.v = it.sv;
            A.w = it.sw;
            if (!plane) {
                Bd.v = it.sv;
                Bd.w = it.sw;
            }
            bool got = true;
            if (it.dp.x != UNSET) got &= final_vel(it.h.x, it.dp.x, sweeps, A);
            if (it.dp.y != UNSET) got &= final_vel(it.h.y, it.dp.y, sweeps, Bd);
            if (!got) break;
            const bool staged = staged_slot(j);
            for (uint32_t k = 0; k < np; ++k) {
                const float4* J = staged ? mine + (size_t)(off + k) * OWN_WAVE : R.J + (size_t)k * JROWS * cap + it.s;
                const size_t jr = staged ? rs : (size_t)cap;
                const float4 L = staged ? J[JROWS * jr] : R.lam[(size_t)k * cap + it.s];
                const float e = point_residual(J, jr, plane, keep, A, Bd, L);
                if (e > res) res = e;
            }
            off += staged ? np : 0u;
        }
        uint32_t bits = __float_as_uint(res);
        for (int o = 32; o >= 1; o >>= 1) {
            const uint32_t ob = __shfl_xor(bits, o, 64);
            bits = ob > bits ? ob : bits;
        }
        uint32_t* arr = stop.arrive + (size_t)(sweeps - 1u) * CHECK_ROWS * CHECK_ROW_WORDS;
        if (lane == 0u) {
            atomicMax(stop.bits + sweeps - 1u, bits);
            const bool above = __uint_as_float(bits) > stop.tol;
            __hip_atomic_fetch_add(arr + (g % CHECK_ROWS) * CHECK_ROW_WORDS, 1u + (above ? CHECK_ABOVE : 0u),
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        bool late = false;
        uint32_t seen;
        for (;;) {
            // the rows summed (arrivals stay below 2^16, so the halves add apart); uniform
            // loads, no cross-lane steps (the sweep loop's register allocation stays as it is)
            seen = 0u;
#pragma unroll
            for (uint32_t r = 0; r < CHECK_ROWS; ++r)
                seen += __hip_atomic_load(arr + r * CHECK_ROW_WORDS, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if ((seen & (CHECK_ABOVE - 1u)) >= G) break;
            if (wall_clock64() - t_start > t_limit) {
                late = true;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        if (late) {
            if (lane == 0u) atomicOr(err, STICKY_TIMEOUT);
            return -1;
        }
        if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) return -1;
        return seen < CHECK_ABOVE ? 1 : 0;
    };
    const bool checks = stop.bits != nullptr;
    auto sweep_loop = [&]<bool STAGED>() -> bool {
        uint32_t chunk_end = checks ? min(stop.every, nsweeps) : nsweeps;
        uint32_t total = nv * chunk_end;
        uint32_t done = 0, sw = 0, j = 0, off = 0;
        Item it;
        it.s = 0u;
        if (total > 0u) it = (STAGED || nrec > 0u) ? item_get(recs) : item_global(B, R, slot_of(0));
        uint32_t s = it.s;
        bool fresh = true;
        uint32_t tA = 0, tB = 0, np = 0;
        bool plane = false;
        SB A, Bd;
        zero_sb(A);
        zero_sb(Bd);
        uint32_t spins = 0, idle = 0;
        for (;;) {
            const bool pending = done < total;
            if (pending && fresh) {
                fresh = false;
                tA = it.dp.x == UNSET ? 0u : sw * (it.dp.x >> 8) + (it.dp.x & 0xFFu);
                tB = it.dp.y == UNSET ? 0u : sw * (it.dp.y >> 8) + (it.dp.y & 0xFFu);
                plane = (it.h.y & PLANE_BIT) != 0u;
                np = min(it.h.z, 4u);
                A.im = it.imA;
                A.dyn = it.dp.x != UNSET;
                Bd.im = plane ? 0.0f : it.imB;
                Bd.dyn = it.dp.y != UNSET;
                if (!A.dyn) {
                    A.v = it.sv;
                    A.w = it.sw;
                }
                if (plane) {
                    Bd.v = Bd.w = f3{0.0f, 0.0f, 0.0f};
                } else if (!Bd.dyn) {
                    Bd.v = it.sv;
                    Bd.w = it.sw;
                }
            }
            // poll = load: all words of both bodies in this slot's inbox (one line), issued together
            bool ok = false;
            bool ready = pending;
            if (pending && hint) {
                // the two hint words first; the full load below runs only for lanes
                // whose hint words are in
                const uint64_t ha = vload(R.inbox + vin_word(s, 0, IB_HINT)), hb = vload(R.inbox + vin_word(s, 1, IB_HINT));
                ready = ((vver2(vsplit(ha)) == (tA & 0xFFFFu)) | !A.dyn) & ((vver2(vsplit(hb)) == (tB & 0xFFFFu)) | !Bd.dyn);
            }
            if (ready) {
                // (low, high) dwords of the four words of each side
                uint2 a[4], b[4];
                if (wide) {
                    // 16-byte loads, two per side
                    const uint4 q0 = vin_quad(R.inbox, s, 0), q1 = vin_quad(R.inbox, s, 2),
                                q2 = vin_quad(R.inbox, s, 4), q3 = vin_quad(R.inbox, s, 6);
                    a[0] = make_uint2(q0.x, q0.y), a[1] = make_uint2(q0.z, q0.w);
                    a[2] = make_uint2(q1.x, q1.y), a[3] = make_uint2(q1.z, q1.w);
                    b[0] = make_uint2(q2.x, q2.y), b[1] = make_uint2(q2.z, q2.w);
                    b[2] = make_uint2(q3.x, q3.y), b[3] = make_uint2(q3.z, q3.w);
                } else {
                    const uint64_t* ib = R.inbox;
#pragma unroll
                    for (uint32_t k = 0; k < IB_SIDE; ++k) {
                        a[k] = vsplit(vload(ib + vin_word(s, 0, k)));
                        b[k] = vsplit(vload(ib + vin_word(s, 1, k)));
                    }
                }
                // branch-free: every loaded word is consumed here, so no later instruction
                // waits on the memory counter (which also counts the stores before it)
                const uint32_t tA16 = tA & 0xFFFFu, tB16 = tB & 0xFFFFu;
                const uint32_t da = (vver2(a[0]) ^ tA16) | (vver2(a[1]) ^ tA16) | (vver2(a[2]) ^ tA16) |
                                    (vver2(a[3]) ^ tA16);
                const uint32_t db = (vver2(b[0]) ^ tB16) | (vver2(b[1]) ^ tB16) | (vver2(b[2]) ^ tB16) |
                                    (vver2(b[3]) ^ tB16);
                ok = ((da == 0u) | !A.dyn) & ((db == 0u) | !Bd.dyn);
                const bool ta = ok & A.dyn, tb = ok & Bd.dyn;
                f3 av, aw, bv, bw;
                vunpack4(a[0], a[1], a[2], a[3], av, aw);
                vunpack4(b[0], b[1], b[2], b[3], bv, bw);
                A.v = f3{ta ? av.x : A.v.x, ta ? av.y : A.v.y, ta ? av.z : A.v.z};
                A.w = f3{ta ? aw.x : A.w.x, ta ? aw.y : A.w.y, ta ? aw.z : A.w.z};
                Bd.v = f3{tb ? bv.x : Bd.v.x, tb ? bv.y : Bd.v.y, tb ? bv.z : Bd.v.z};
                Bd.w = f3{tb ? bw.x : Bd.w.x, tb ? bw.y : Bd.w.y, tb ? bw.z : Bd.w.z};
            }
            if (ok) {
                const bool first = sweep0 + sw == 0u;
                float4 L[4];
                // the point loops run over a fixed 4 so that L stays in registers
                const bool stg = STAGED || staged_slot(j);
                if (stg) {
                    float4* P = mine + (size_t)off * OWN_WAVE;
#pragma unroll
                    for (uint32_t k = 0; k < 4; ++k)
                        if (k < np) L[k] = P[JROWS * rs + (size_t)k * OWN_WAVE];
                    solve_rows(P, rs, OWN_WAVE, np, plane, keep, first, A, Bd, L);
#pragma unroll
                    for (uint32_t k = 0; k < 4; ++k)
                        if (k < np) P[JROWS * rs + (size_t)k * OWN_WAVE] = L[k];
                } else {
                    float4 Jr[4][JROWS];
                    load_rows(R, s, cap, np, plane, Jr);
#pragma unroll
                    for (uint32_t k = 0; k < 4; ++k)
                        if (k < np) L[k] = R.lam[(size_t)k * cap + s];
                    solve_regs(Jr, np, keep, first, A, Bd, L);
#pragma unroll
                    for (uint32_t k = 0; k < 4; ++k)
                        if (k < np) R.lam[(size_t)k * cap + s] = make_float4(L[k].x, L[k].y, L[k].z, 0.0f);
                }
                // pass the velocities on to each body's next visit
                if (A.dyn) vin_store(R.inbox, it.nx.x >> 1, it.nx.x & 1u, tA + 1u, A.v, A.w, wide);
                if (Bd.dyn) vin_store(R.inbox, it.nx.y >> 1, it.nx.y & 1u, tB + 1u, Bd.v, Bd.w, wide);
                // next item of this lane: the next owned slot, or the first one of the next sweep
                ++done;
                off += stg ? np : 0u;
                if (++j == nv) {
                    j = 0;
                    ++sw;
                    off = 0;
                }
                fresh = true;
                if (done < total) {
                    it = (STAGED || j < nrec) ? item_get(recs + (size_t)j * 4u * OWN_WAVE)
                                              : item_global(B, R, slot_of(j));
                    s = it.s;
                }
            }
            if (!__ballot(done < total)) {
                if (!checks) return true;
                // every slot of this wave is through sweep chunk_end - 1
                const int r = check(chunk_end);
                if (r < 0) return false;
                if (r > 0 || chunk_end == nsweeps) {
                    if (g == 0u && lane == 0u) {
                        __hip_atomic_store(&cnt->sweeps_run, chunk_end, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        __hip_atomic_store(&cnt->tol_hit, r > 0 ? 1u : 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    }
                    return true;
                }
                chunk_end = min(chunk_end + stop.every, nsweeps);
                total = nv * chunk_end;
                if (nv > 0u) {
                    it = (STAGED || nrec > 0u) ? item_get(recs) : item_global(B, R, slot_of(0));
                    s = it.s;
                }
                fresh = true;
                continue;
            }
            if ((++spins & 255u) == 0u) {
                bool quit = false;
                if (wall_clock64() - t_start > t_limit) {
                    if (done < total) atomicOr(err, STICKY_TIMEOUT);  // waited longer than spin_limit microseconds
                    quit = true;
                } else if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) {
                    quit = true;  // another wave already gave up: drain
                }
                if (quit) return false;
            }
            // adaptive back-off: a wave none of whose lanes got through sleeps longer each time
            if (adaptive) idle = __ballot(ok) ? 0u : min(idle + 1u, 4u);
            switch (min(nap + idle, 6u)) {
            case 0: break;
            case 1: __builtin_amdgcn_s_sleep(1); break;
            case 2: __builtin_amdgcn_s_sleep(2); break;
            case 3: __builtin_amdgcn_s_sleep(4); break;
            case 4: __builtin_amdgcn_s_sleep(8); break;
            case 5: __builtin_amdgcn_s_sleep(16); break;
            default: __builtin_amdgcn_s_sleep(32); break;
            }
        }
    };
    const bool finished = __ballot(!all_staged) == 0ull ? sweep_loop.template operator()<true>()
                                                        : sweep_loop.template operator()<false>();
    if (!finished) return;

    // impulses of the staged points back to the rows
    uint32_t woff = 0;
    for (uint32_t jj = 0; jj < nv; ++jj) {
        const uint32_t ss = slot_of(jj);
        if (!staged_slot(jj)) continue;
        const uint32_t n = min(jj < nrec ? __float_as_uint(recs[(size_t)jj * 4u * OWN_WAVE].z) : R.hdr[ss].z, 4u);
        for (uint32_t k = 0; k < n; ++k) {
            const float4 L = mine[(JROWS * rs) + (size_t)(woff + k) * OWN_WAVE];
            R.lam[(size_t)k * cap + ss] = make_float4(L.x, L.y, L.z, 0.0f);
        }
        woff += n;
    }
}

__global__ void __launch_bounds__(256) k_residual(Bodies B, StepParams sp, SolverRows R, const Counters* cnt,
                                                  uint32_t cap, uint32_t* __restrict__ out_bits) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t M = cnt->m_glob;  // the large islands' slots (the small ones report their own)
    float res = 0.0f;
    if (s < M && cnt->sticky == 0u) {  // rows exist only when the preparation ran
        uint4 h = R.hdr[s];
        h.z = min(h.z, 4u);
        const bool plane = (h.y & PLANE_BIT) != 0;
        SB A, Bd;
        load_vel<false>(B, h.x, A);
        if (plane) zero_sb(Bd); else load_vel<false>(B, h.y, Bd);
        for (uint32_t k = 0; k < h.z; ++k) {
            const float e = point_residual(R.J + (size_t)k * JROWS * cap + s, cap, plane, sp.keep, A, Bd,
                                           R.lam[(size_t)k * cap + s]);
            if (e > res) res = e;
        }
    }
    // wave max, then one atomic per wave (non-negative floats order as their bits)
    uint32_t bits = __float_as_uint(res);
    for (int o = 32; o >= 1; o >>= 1) {
        const uint32_t ob = __shfl_xor(bits, o, 64);
        bits = ob > bits ? ob : bits;
    }
    if ((threadIdx.x & 63) == 0) atomicMax(out_bits, bits);
}

// ---------------------------------------------------------------- islands solved on their own
// The manifolds of an island of at most ISLAND_LOCAL_MAX manifolds share no dynamic body
// with any other island, so the island is solved by itself: sweeps over its slots in
// solve order (colour, then manifold), and with a tolerance it stops after the first
// sweep that leaves its own residual <= tol.  Without a tolerance it runs every sweep, and
// the result is bit-identical to sweeping all islands together colour by colour.

// Residual of solve slot s (the k_residual value of its points).
__device__ __forceinline__ float slot_residual(const Bodies& B, float keep, const SolverRows& R, uint32_t s,
                                               uint32_t cap) {
    const uint4 h = R.hdr[s];
    const uint32_t np = min(h.z, 4u);
    const bool plane = (h.y & PLANE_BIT) != 0;
    SB A, Bd;
    load_vel<false>(B, h.x, A);
    if (plane) zero_sb(Bd); else load_vel<false>(B, h.y, Bd);
    float res = 0.0f;
    for (uint32_t k = 0; k < np; ++k) {
        const float e = point_residual(R.J + (size_t)k * JROWS * cap + s, cap, plane, keep, A, Bd,
                                       R.lam[(size_t)k * cap + s]);
        if (e > res) res = e;
    }
    return res;
}
// One island's result: sweeps run, final residual (bits), whether it missed the tolerance,
// contact points; folded into the counters.
__device__ __forceinline__ void local_done(Counters* cnt, uint32_t sweeps, float res, bool miss, uint32_t pts,
                                           bool bundle) {
    atomicMax(&cnt->loc_sweeps, sweeps);
    atomicMax(&cnt->loc_res_bits, __float_as_uint(res));
    if (miss) atomicOr(&cnt->loc_miss, 1u);
    atomicAdd(&cnt->loc_points, pts);
    atomicAdd(&cnt->loc_point_sweeps, (unsigned long long)pts * sweeps);
    if (bundle) atomicAdd(&cnt->bundle_point_sweeps, (unsigned long long)pts * sweeps);
}
// Islands of at most ISLAND_LANE_MAX manifolds, bundled: the islands whose first slot lies
// in window w (BUNDLE_SLOTS slots from the first small-island slot on) are solved by
// workgroup w, one wave, one lane per slot (at most BUNDLE_SLOTS + ISLAND_LANE_MAX - 1 lanes).
// Each sweep runs phases: in phase p every island visits the slots of its p-th colour, all
// at once (within an island a colour's slots share no dynamic body, and islands share none
// at all), with a barrier between phases; then each island's residual is the maximum over
// its lanes, and an island that is done stops while the others go on.
//
// Everything a sweep touches stays on chip for the whole solve: each lane's Jacobian rows
// and impulses in its registers (a lane visits the same slot in every sweep; with one wave
// per SIMD the rows fit the register file, accumulation registers included), and each
// body's solver velocity in LDS, in one home slot: the first (lane, side) of its island
// that names it (dynamic bodies appear in one island only; a static partner keeps its own
// slot).  The visits are the arithmetic of solve_slot on those copies; velocities and
// impulses go back to memory once, at the end.
constexpr uint32_t BUNDLE_SLOTS = 56;
static_assert(BUNDLE_SLOTS + ISLAND_LANE_MAX - 1u <= 64u, "a bundle fits one wave");
constexpr uint32_t BUNDLE_BODY0 = 0u;  // first body slot (float4 units)
constexpr uint32_t BUNDLE_LDS_BYTES = (BUNDLE_BODY0 + 4u * 64u) * 16u;
__device__ __forceinline__ void home_load(const float4* lds, uint32_t h, SB& X) {
    const float4 v = lds[BUNDLE_BODY0 + h], w = lds[BUNDLE_BODY0 + 128u + h];
    X.v = xyz(v);
    X.im = v.w;
    X.w = xyz(w);
    X.dyn = v.w > 0.0f;
}
__device__ __forceinline__ void home_store(float4* lds, uint32_t h, const SB& X) {
    lds[BUNDLE_BODY0 + h] = f4(X.v, X.im);
    lds[BUNDLE_BODY0 + 128u + h] = f4(X.w, 0.0f);
}
__global__ void __launch_bounds__(64) k_pgs_bundle(Bodies B, StepParams sp, SolverRows R, Counters* cnt, uint32_t cap,
                                                   const uint32_t* __restrict__ skey, float tol,
                                                   uint32_t max_sweeps) {
    extern __shared__ float4 bl[];
    __shared__ uint32_t isl_res[2][64];  // per island (first slot - window start): residual bits, two sweeps
    __shared__ uint32_t isl_pts[64];
    __shared__ uint2 ab[64];
    __shared__ uint32_t colr[64];
    __shared__ uint32_t kmax_w;
    if (cnt->sticky != 0u) return;
    const uint32_t M = min(cap, cnt->m_eff), M0 = cnt->m_glob, lane = threadIdx.x;
    const uint32_t nwin = M > M0 ? (M - M0 + BUNDLE_SLOTS - 1u) / BUNDLE_SLOTS : 0u;
    const bool tol_mode = tol > 0.0f;
    const float keep = sp.keep;
    for (uint32_t win = blockIdx.x; win < nwin; win += gridDim.x) {
        const uint32_t w0 = M0 + win * BUNDLE_SLOTS, s = w0 + lane;
        // this lane's slot belongs to the bundle when its island is a one-lane island whose
        // first slot lies in the window
        bool mine = false;
        uint32_t li = 0, colour = 0;
        if (s < M) {
            const uint32_t k = skey[s];
            if (B.icnt[(k & ~LOCAL_KEY) >> 8] <= ISLAND_LANE_MAX) {
                uint32_t st = s;
                while (st > M0 && ((skey[st - 1u] ^ k) >> 8) == 0u) --st;  // at most ISLAND_LANE_MAX - 1 steps
                mine = st >= w0 && st < w0 + BUNDLE_SLOTS;
                li = mine ? st - w0 : 0u;
                colour = k & 0xFFu;
            }
        }
        const uint4 h = mine ? R.hdr[s] : make_uint4(UNSET, UNSET, 0u, 0u);
        const uint32_t np = min(h.z, 4u);
        const bool plane = (h.y & PLANE_BIT) != 0u;
        ab[lane] = make_uint2(h.x, h.y);
        colr[lane] = mine ? colour : UNSET;
        isl_res[0][lane] = 0u;
        isl_res[1][lane] = 0u;
        isl_pts[lane] = 0u;
        if (lane == 0u) kmax_w = 0u;
        // the slot's rows and impulses in registers
        float4 Jr[4][JROWS];
        float4 L[4];
#pragma unroll
        for (uint32_t k = 0; k < 4; ++k) L[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (mine) {
            // own copies of both bodies' velocities
            const float4 va = B.sv[h.x], wa = B.sw[h.x];
            bl[BUNDLE_BODY0 + lane] = va;
            bl[BUNDLE_BODY0 + 128u + lane] = make_float4(wa.x, wa.y, wa.z, 0.0f);
            // a plane's side holds zeros (velocity, inverse mass): it loads as a static body
            float4 vb = make_float4(0.0f, 0.0f, 0.0f, 0.0f), wb = vb;
            if (!plane) {
                vb = B.sv[h.y];
                wb = B.sw[h.y];
            }
            bl[BUNDLE_BODY0 + 64u + lane] = vb;
            bl[BUNDLE_BODY0 + 192u + lane] = make_float4(wb.x, wb.y, wb.z, 0.0f);
            load_rows(R, s, cap, np, plane, Jr);
#pragma unroll
            for (uint32_t k = 0; k < 4; ++k)
                if (k < np) L[k] = R.lam[(size_t)k * cap + s];
        }
        __syncthreads();
        // home slot of each side: the first (lane, side) of the island naming the body; the
        // lane's phase: the colours of its island below its own (an island's slots ascend by
        // colour), so each island visits its own colours in turn, lanes of one phase at once
        uint32_t home_a = lane, home_b = 64u + lane, phase = 0;
        if (mine) {
            for (uint32_t j = li + 1u; j <= lane; ++j) phase += colr[j] != colr[j - 1u] ? 1u : 0u;
            atomicMax(&kmax_w, phase);
            atomicAdd(&isl_pts[li], np);
            auto home_of = [&](uint32_t body, uint32_t own) {
                for (uint32_t j = li; j <= lane; ++j) {
                    const uint2 o = ab[j];
                    if (o.x == body) return j;
                    if (o.y == body) return 64u + j;
                }
                return own;
            };
            home_a = home_of(h.x, lane);
            if (!plane) home_b = home_of(h.y, 64u + lane);
        }
        __syncthreads();
        const uint32_t kmax = kmax_w;
        bool active = mine;
        uint32_t sweeps = 0;
        float res = 0.0f;
        for (uint32_t it = 0; it < max_sweeps; ++it) {
            if (!__syncthreads_or(active)) break;
            for (uint32_t c = 0; c <= kmax; ++c) {
                if (active && phase == c) {
                    SB A, Bd;
                    home_load(bl, home_a, A);
                    home_load(bl, home_b, Bd);
                    solve_regs(Jr, np, keep, it == 0u, A, Bd, L);
#pragma unroll
                    for (uint32_t k = 0; k < 4; ++k) L[k].w = 0.0f;
                    if (A.dyn) home_store(bl, home_a, A);
                    if (Bd.dyn) home_store(bl, home_b, Bd);
                }
                __syncthreads();
            }
            if (tol_mode || it + 1u == max_sweeps) {
                uint32_t* cur = isl_res[it & 1u];
                isl_res[(it + 1u) & 1u][lane] = 0u;  // the next sweep's words (not read in this one)
                if (active) {
                    SB A, Bd;
                    home_load(bl, home_a, A);
                    home_load(bl, home_b, Bd);
                    float r = 0.0f;
#pragma unroll
                    for (uint32_t k = 0; k < 4; ++k) {
                        if (k >= np) break;
                        const float e = point_residual_regs(Jr[k], keep, A, Bd, L[k]);
                        if (e > r) r = e;
                    }
                    atomicMax(&cur[li], __float_as_uint(r));
                }
                __syncthreads();
                if (active) {
                    res = __uint_as_float(cur[li]);
                    if (tol_mode && res <= tol) {
                        sweeps = it + 1u;
                        active = false;
                    }
                }
            }
            if (active && it + 1u == max_sweeps) {
                sweeps = max_sweeps;
                active = false;
            }
        }
        __syncthreads();
        if (mine) {
            // impulses and the velocities of the bodies whose home this lane holds back to memory
#pragma unroll
            for (uint32_t k = 0; k < 4; ++k)
                if (k < np) R.lam[(size_t)k * cap + s] = L[k];
            SB A;
            home_load(bl, lane, A);
            if (home_a == lane && A.dyn) {
                B.sv[h.x] = f4(A.v, A.im);
                B.sw[h.x] = f4(A.w, 0.0f);
            }
            if (!plane && home_b == 64u + lane) {
                SB Bd;
                home_load(bl, 64u + lane, Bd);
                if (Bd.dyn) {
                    B.sv[h.y] = f4(Bd.v, Bd.im);
                    B.sw[h.y] = f4(Bd.w, 0.0f);
                }
            }
            // one lane per island (its first slot) reports it
            if (s == w0 + li) local_done(cnt, sweeps, res, tol_mode && !(res <= tol), isl_pts[li], true);
        }
        __syncthreads();
    }
}
int pgs_bundle_setup() {
    return hipFuncSetAttribute((const void*)k_pgs_bundle, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)BUNDLE_LDS_BYTES) == hipSuccess ? 1 : 0;
}
// Block-wide maximum and sum of 256 threads (red: 8 words of LDS); every thread gets both.
__device__ __forceinline__ void block_max_sum(uint32_t& mx, uint32_t& sum, uint32_t* red) {
    for (int o = 32; o >= 1; o >>= 1) {
        const uint32_t om = __shfl_xor(mx, o, 64), os = __shfl_xor(sum, o, 64);
        mx = om > mx ? om : mx;
        sum += os;
    }
    if ((threadIdx.x & 63u) == 0u) {
        red[threadIdx.x >> 6] = mx;
        red[4 + (threadIdx.x >> 6)] = sum;
    }
    __syncthreads();
    mx = max(max(red[0], red[1]), max(red[2], red[3]));
    sum = red[4] + red[5] + red[6] + red[7];
    __syncthreads();
}
// Larger islands: one workgroup each.  A colour's slots share no dynamic body, so the
// workgroup visits them together and meets at a barrier before the next colour (the
// velocities move through global memory, visible within the workgroup after the barrier).
__global__ void __launch_bounds__(256) k_pgs_group(Bodies B, StepParams sp, SolverRows R, Counters* cnt, uint32_t cap,
                                                   const uint2* __restrict__ list, const uint32_t* __restrict__ skey,
                                                   float tol, uint32_t max_sweeps,
                                                   const uint64_t* __restrict__ flow_bmask) {
    __shared__ uint32_t red[8];
    // with flow_bmask, the large islands' sweep launch is prepared here (as k_flow_init)
    if (flow_bmask) flow_init_part(B, R, flow_bmask, cnt, blockIdx.x * blockDim.x + threadIdx.x,
                                   gridDim.x * blockDim.x);
    if (cnt->sticky != 0u) return;
    const uint32_t n = cnt->isl_group_n, t = threadIdx.x;
    const bool tol_mode = tol > 0.0f;
    for (uint32_t g = blockIdx.x; g < n; g += gridDim.x) {
        const uint2 e = list[g];
        const uint32_t s0 = e.x, s1 = e.x + e.y;
        const uint32_t k0 = skey[s0], k1 = skey[s1 - 1u];
        uint32_t it = 0;
        float res = 0.0f;
        for (; it < max_sweeps; ++it) {
            uint32_t a = s0;
            for (uint32_t c = k0 & 0xFFu; c <= (k1 & 0xFFu); ++c) {
                // [a, b): the island's slots of colour c (the keys ascend, equal above the colour)
                const uint32_t kc = (k0 & ~0xFFu) | c;
                uint32_t lo = a, hi = s1;
                while (lo < hi) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (skey[mid] <= kc) lo = mid + 1u; else hi = mid;
                }
                if (lo > a) {
                    for (uint32_t s = a + t; s < lo; s += blockDim.x) solve_slot(B, sp, R, R.hdr[s], s, cap, it == 0u);
                    __syncthreads();
                }
                a = lo;
            }
            if (tol_mode || it + 1u == max_sweeps) {
                float r = 0.0f;
                for (uint32_t s = s0 + t; s < s1; s += blockDim.x) {
                    const float q = slot_residual(B, sp.keep, R, s, cap);
                    if (q > r) r = q;
                }
                uint32_t mx = __float_as_uint(r), unused = 0u;
                block_max_sum(mx, unused, red);
                res = __uint_as_float(mx);
                if (tol_mode && res <= tol) {
                    ++it;
                    break;
                }
            }
        }
        uint32_t pts = 0, mx = 0u;
        for (uint32_t s = s0 + t; s < s1; s += blockDim.x) pts += min(R.hdr[s].z, 4u);
        block_max_sum(mx, pts, red);
        if (t == 0u) local_done(cnt, it, res, tol_mode && !(res <= tol), pts, false);
    }
}

// Previous-step lookup table for the warm start: (a, b) -> slot.
// Each entry's table position goes to pos[slot], so that the next build of this table
// clears only the entries it holds (cache_m of them) instead of the whole table.
__device__ __forceinline__ void cache_build_part(const SolverRows& R, Counters* cnt, const CacheTable& C,
                                                 uint32_t gid, uint32_t gsz) {
    uint64_t* __restrict__ key = C.key;
    uint32_t* __restrict__ slot = C.slot;
    uint32_t* __restrict__ np_out = C.np_out;
    uint32_t* __restrict__ pos = C.pos;
    const uint32_t mask = C.mask;
    const uint32_t M = cnt->m_eff;
    if (gid == 0) cnt->cache_m[C.set] = M;
    for (uint32_t s = gid; s < M; s += gsz) {
        np_out[s] = R.hdr[s].z;
        const uint4 h = R.hdr[s];
        const uint64_t k = ((uint64_t)h.x << 32) | h.y;
        uint32_t q = (uint32_t)mix64(k) & mask;
        bool placed = false;
        for (uint32_t probe = 0; probe <= mask && !placed; ++probe) {
            const unsigned long long prev =
                atomicCAS((unsigned long long*)&key[q], (unsigned long long)EMPTY_KEY, (unsigned long long)k);
            if (prev == EMPTY_KEY || prev == k) {
                slot[q] = s;
                pos[s] = q;
                placed = true;
            }
            q = (q + 1) & mask;
        }
        if (!placed) atomicOr(&cnt->sticky, STICKY_INTERNAL);  // table full: never expected (sized >= 2x the slots)
    }
}
__global__ void __launch_bounds__(256) k_cache_build(const SolverRows R, Counters* cnt, CacheTable C) {
    if (cnt->sticky != 0u) return;
    cache_build_part(R, cnt, C, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
}

// Moves the entries of an (a, b) -> slot table into a larger empty one.
__global__ void __launch_bounds__(256) k_cache_rehash(const uint64_t* __restrict__ okey,
                                                      const uint32_t* __restrict__ oslot, uint32_t on,
                                                      uint64_t* __restrict__ key, uint32_t* __restrict__ slot,
                                                      uint32_t mask, Counters* __restrict__ cnt,
                                                      uint32_t* __restrict__ pos) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= on) return;
    const uint64_t k = okey[i];
    if (k == EMPTY_KEY) return;
    const uint32_t v = oslot[i];
    uint32_t q = (uint32_t)mix64(k) & mask;
    for (uint32_t probe = 0; probe <= mask; ++probe) {
        const unsigned long long prev =
            atomicCAS((unsigned long long*)&key[q], (unsigned long long)EMPTY_KEY, (unsigned long long)k);
        if (prev == EMPTY_KEY || prev == k) {
            slot[q] = v;
            pos[v] = q;
            return;
        }
        q = (q + 1) & mask;
    }
    atomicOr(&cnt->sticky, STICKY_INTERNAL);
}



// Clears the per-step part of the control block.
__global__ void k_counters_begin(Counters* cnt, uint32_t* __restrict__ kinds) { counters_reset(cnt, kinds); }

// Integration is skipped while any sticky error is set, so an overflowing step
// leaves the state exactly as it was.
__global__ void __launch_bounds__(256) k_integrate(Bodies B, StepParams sp, uint32_t n, Counters* cnt,
                                                   uint32_t sweeps, SolverRows R, CacheTable C) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (cnt->sticky != 0u) return;
    // the step's (a, b) -> slot table for the next step's warm start, in the same launch
    if (C.key) cache_build_part(R, cnt, C, i, gridDim.x * blockDim.x);
    if (i == 0) {
        cnt->steps_done += 1u;
        // sweeps of the large islands, and of the step (the most any island ran)
        const uint32_t swg = sweeps == SWEEPS_ON_DEVICE ? cnt->sweeps_run : sweeps;
        const uint32_t sw = max(swg, cnt->loc_sweeps);
        cnt->pgs_bytes_total += spread_sum(cnt, &SpreadRow::pgs_bytes) * (unsigned long long)sw;
        cnt->sweeps_total += sw;
        cnt->bundle_point_sweeps_total += cnt->bundle_point_sweeps;
        cnt->acc_steps += 1u;
        cnt->acc_reached += cnt->tol_hit != 0u && cnt->loc_miss == 0u ? 1u : 0u;
        cnt->acc_max_sweeps = max(cnt->acc_max_sweeps, sw);
        cnt->acc_sweeps += sw;
        cnt->acc_contact_sweeps +=
            (unsigned long long)(spread_sum(cnt, &SpreadRow::contacts) - cnt->loc_points) * swg + cnt->loc_point_sweeps;
        cnt->acc_local_contact_sweeps += cnt->loc_point_sweeps;
        cnt->acc_gjk_fail += cnt->gjk_fail;
        cnt->acc_epa_fail += cnt->epa_fail;
    }
    if (i >= n) return;
    const float4 v4 = B.sv[i];
    if (!(v4.w > 0.0f) && !(B.bfl[i].x & BODY_KINEMATIC)) return;
    const float4 w4 = B.sw[i];
    const float4 x4 = B.x[i];
    const float4 qv = B.q[i];
    const f3 v = xyz(v4), w = xyz(w4);
    const f3 x = xyz(x4) + v * sp.dt;
    const float hdt = 0.5f * sp.dt;
    const q4 q = {qv.x, qv.y, qv.z, qv.w};
    const q4 dq = qmul(q4{w.x, w.y, w.z, 0.0f}, q);
    const q4 nq = {q.x + dq.x * hdt, q.y + dq.y * hdt, q.z + dq.z * hdt, q.w + dq.w * hdt};
    const float len = sqrtf(nq.x * nq.x + nq.y * nq.y + nq.z * nq.z + nq.w * nq.w);
    const float inv = 1.0f / len;
    B.x[i] = make_float4(x.x, x.y, x.z, x4.w);
    B.q[i] = make_float4(nq.x * inv, nq.y * inv, nq.z * inv, nq.w * inv);
    B.vel[i] = v4;
    B.angv[i] = make_float4(w.x, w.y, w.z, 0.0f);
}

// ---------------------------------------------------------------- islands and sleeping
// Union-find over the step's manifolds between two awake dynamic bodies (uf_find).
__global__ void __launch_bounds__(256) k_isl_union(Bodies B, const uint2* __restrict__ m_ab, uint32_t cap,
                                                   uint32_t n, const Counters* __restrict__ cnt) {
    if (cnt->sticky != 0u) return;
    const uint32_t M = min(cap, cnt->m_eff);
    const uint32_t stamp = cnt->steps_done + 1u;  // runs before the integration counts this step
    for (uint32_t m = blockIdx.x * blockDim.x + threadIdx.x; m < M; m += gridDim.x * blockDim.x) {
        const uint2 ab = m_ab[m];
        if (ab.y & PLANE_BIT) continue;
        const bool da = B.sv[ab.x].w > 0.0f, db = B.sv[ab.y].w > 0.0f;
        if (da && db) {
            uint32_t ra = uf_find(B.upar, ab.x), rb = uf_find(B.upar, ab.y);
            while (ra != rb) {
                if (ra < rb) {
                    const uint32_t old = atomicCAS(B.upar + rb, rb, ra);
                    if (old == rb) break;
                    rb = uf_find(B.upar, old);
                } else {
                    const uint32_t old = atomicCAS(B.upar + ra, ra, rb);
                    if (old == ra) break;
                    ra = uf_find(B.upar, old);
                }
            }
        } else if (da || db) {
            // an awake dynamic body touches a sleeping one: its island wakes at the next step
            const uint32_t o = da ? ab.y : ab.x;
            const uint32_t y = B.slp[o].y;
            if ((y & ASLEEP_BIT) && (y & LABEL_MASK) < n) B.wake[y & LABEL_MASK] = stamp;
        }
    }
}
// Labels, rest times and each island's smallest rest time; counts islands and sleeping bodies.
__global__ void __launch_bounds__(256) k_isl_label(Bodies B, StepParams sp, uint32_t n, Counters* cnt) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (cnt->sticky != 0u) return;
    uint32_t isl = 0, slp = 0;
    // each island's least rest time (rest times are >= 0: bit order = value order)
    struct Min {
        uint32_t* imin;
        __device__ uint32_t combine(uint32_t a, uint32_t b) const { return a < b ? a : b; }
        __device__ void apply(uint32_t r, uint32_t c) const { atomicMin(imin + r, c); }
    };
    const bool dyn = i < n && B.sv[i].w > 0.0f;
    uint32_t r = UNSET, tb = 0u;
    if (dyn) {
        r = uf_root(B.upar, i);
        const float4 v = B.vel[i], w = B.angv[i];
        const bool still = v.x * v.x + v.y * v.y + v.z * v.z < sp.sleep_lin2 &&
                           w.x * w.x + w.y * w.y + w.z * w.z < sp.sleep_ang2;
        const float t = still ? __uint_as_float(B.slp[i].x) + sp.dt : 0.0f;
        B.slp[i] = make_uint2(__float_as_uint(t), r);
        tb = __float_as_uint(t);
        isl = r == i ? 1u : 0u;
    }
    wave_by_island(dyn, r, tb, 0xFFFFFFFFu, Min{B.imin});
    if (i >= n) return;
    if (!dyn && B.vel[i].w > 0.0f) {
        // asleep (a static body has no inverse mass)
        const uint32_t y = B.slp[i].y;
        isl = (y & LABEL_MASK) == i ? 1u : 0u;
        slp = 1u;
    }
    const uint32_t row = (blockIdx.x * 4u + (threadIdx.x >> 6)) % SPREAD;
    const uint32_t wi = __popcll(__ballot(isl != 0u)), ws = __popcll(__ballot(slp != 0u));
    if ((threadIdx.x & 63u) == 0u) {
        if (wi) atomicAdd(&cnt->sp[row].islands, wi);
        if (ws) atomicAdd(&cnt->sp[row].sleeping, ws);
    }
}
// Islands whose bodies have all rested sleep_time seconds fall asleep.
__global__ void __launch_bounds__(256) k_isl_sleep(Bodies B, StepParams sp, uint32_t n, Counters* cnt) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (cnt->sticky != 0u || i >= n) return;
    bool now = false;
    if (sp.allow_sleep && B.sv[i].w > 0.0f) {
        const uint2 sl = B.slp[i];
        if (B.imin[sl.y] >= __float_as_uint(sp.sleep_time)) {
            B.slp[i] = make_uint2(sl.x, sl.y | ASLEEP_BIT);
            B.vel[i] = make_float4(0.0f, 0.0f, 0.0f, B.vel[i].w);
            B.angv[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            now = true;
        }
    }
    const uint32_t row = (blockIdx.x * 4u + (threadIdx.x >> 6)) % SPREAD;
    const uint32_t ws = __popcll(__ballot(now));
    if ((threadIdx.x & 63u) == 0u && ws) atomicAdd(&cnt->sp[row].sleeping, ws);
}

// ---------------------------------------------------------------- host launchers
static inline dim3 grid_for(uint32_t n, uint32_t b = 256) { return dim3((n + b - 1) / b ? (n + b - 1) / b : 1); }
// Kernels whose element count is read on the device stride over it with at most
// 1024 workgroups, so that a capacity far above the count launches few idle waves.
static inline dim3 grid_bounded(uint32_t cap, uint32_t max_groups = 1024u) {
    const uint32_t b = (cap + 255u) / 256u;
    return dim3(b == 0u ? 1u : (b > max_groups ? max_groups : b));
}


void launch_counters_begin(Counters* cnt, uint32_t* kinds, hipStream_t s) {
    hipLaunchKernelGGL(k_counters_begin, dim3(1), dim3(64), 0, s, cnt, kinds);
}
void launch_begin(const Bodies& B, const StepParams& sp, uint32_t n, uint64_t* key, uint32_t* key32, uint32_t* val,
                  uint64_t* bmax0,
                  uint64_t* bmax1, uint64_t* bmask, uint32_t* bdeg, Counters* cnt, uint32_t* kinds,
                  const CellTable& ct, hipStream_t s) {
    const uint32_t m = std::max(n, ct.mbits ? ct.mlist_n : 0u);
    if (m)
        hipLaunchKernelGGL(k_begin, grid_for(m), dim3(256), 0, s, B, sp, n, key, key32, val, bmax0, bmax1, bmask, bdeg,
                           cnt, kinds, ct);
    else
        launch_counters_begin(cnt, kinds, s);
}
void launch_fill_u64(uint64_t* p, uint64_t v, uint32_t n, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_fill_u64, grid_for(n), dim3(256), 0, s, p, v, n);
}
void launch_fill_u32(uint32_t* p, uint32_t v, uint32_t n, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_fill_u32, grid_for(n), dim3(256), 0, s, p, v, n);
}
void launch_iota(uint32_t* p, uint32_t n, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_iota, grid_for(n), dim3(256), 0, s, p, n);
}
void launch_cell_clear(const CellTable& t, hipStream_t s) {
    if (t.mbits) return;  // k_begin cleared the bitmap words
    const uint32_t n = std::max(t.mask + 1u, (t.bmask >> 5) + 1u);
    hipLaunchKernelGGL(k_cell_clear, grid_for(n), dim3(256), 0, s, t);
}
void launch_cell_table(const Bodies& B, const SortedKeys& skey, const uint32_t* sval, uint32_t n, const CellTable& t,
                       Counters* cnt, uint32_t* scan_tmp, hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(k_cell_table, grid_for(n), dim3(256), 0, s, B, skey, sval, n, t, cnt);
    if (t.mbits) {
        scan_exclusive(t.crank, t.crank, n, scan_tmp, s, nullptr, nullptr, false, &cnt->sticky);
        hipLaunchKernelGGL(k_cell_rank, grid_for(n), dim3(256), 0, s, skey, n, t, cnt);
        return;
    }
    hipLaunchKernelGGL(k_cell_place, grid_for(n), dim3(256), 0, s, skey, n, t, cnt);
}
void launch_pair_count(const Bodies& B, const StepParams& sp, const SortedKeys& skey, const uint32_t* sval,
                       const CellTable& ct, uint32_t n, uint32_t* cnt, uint32_t* bsum, uint32_t* stash, Counters* ctr,
                       hipStream_t s) {
    if (!n) return;
    if (ct.mbits) {
        hipLaunchKernelGGL(k_pair_count<BP_RANK>, grid_for(BP_PARTS * n), dim3(256), 0, s, B, sp, skey, sval, ct, n,
                           cnt, bsum, stash, ctr);
        return;
    }
    hipLaunchKernelGGL(k_pair_count<BP_SPARSE>, grid_for(BP_PARTS * n), dim3(256), 0, s, B, sp, skey, sval, ct, n, cnt,
                       bsum, stash, ctr);
    hipLaunchKernelGGL(k_pair_count_dense, grid_for(BP_PARTS * n), dim3(256), 0, s, B, sp, skey, sval, ct, n, cnt,
                       bsum, stash, ctr);
}
void launch_pair_emit(const Bodies& B, const StepParams& sp, const SortedKeys& skey, const uint32_t* sval,
                      const CellTable& ct, uint32_t n, const uint32_t* pcnt, const uint32_t* boff, uint32_t cap,
                      uint32_t* pa, uint32_t* pb, Counters* cnt, const uint32_t* stash, hipStream_t s) {
    if (!n) return;
    if (ct.mbits) {
        hipLaunchKernelGGL(k_pair_emit<BP_RANK>, grid_for(BP_PARTS * n), dim3(256), 0, s, B, sp, skey, sval, ct, n, pcnt,
                           boff, cap, pa, pb, cnt, stash);
        return;
    }
    hipLaunchKernelGGL(k_pair_emit<BP_SPARSE>, grid_for(BP_PARTS * n), dim3(256), 0, s, B, sp, skey, sval, ct, n, pcnt,
                       boff, cap, pa, pb, cnt, stash);
    hipLaunchKernelGGL(k_pair_emit<BP_DENSE>, grid_for(BP_PARTS * n), dim3(256), 0, s, B, sp, skey, sval, ct, n, pcnt,
                       boff, cap, pa, pb, cnt, stash);
}
void launch_narrow(const Bodies& B, const StepParams& sp, const uint32_t* pa, const uint32_t* pb, uint32_t cap_pairs,
                   uint32_t* p_np, float4* p_pt, float4* p_nf, uint32_t* epa_list, uint32_t* p_kind, uint32_t* kinds,
                   uint32_t* order, uint32_t epa_groups, Counters* cnt, hipStream_t s) {
    // 256 workgroups: each adds its kind counts to the 16 shared counters once per pass
    hipLaunchKernelGGL(k_pair_kind, grid_bounded(cap_pairs, 256u), dim3(256), 0, s, B, pa, pb, p_kind, kinds, cnt);
    hipLaunchKernelGGL(k_pair_order, grid_bounded(cap_pairs, 256u), dim3(256), 0, s, p_kind, kinds, order, cnt);
    hipLaunchKernelGGL(k_narrow<false>, grid_bounded(cap_pairs), dim3(256), 0, s, B, sp, pa, pb, p_np, p_pt, p_nf,
                       epa_list, cnt, order, kinds);
    hipLaunchKernelGGL(k_narrow<true>, grid_bounded(cap_pairs), dim3(256), 0, s, B, sp, pa, pb, p_np, p_pt, p_nf,
                       epa_list, cnt, order, kinds);
    // the deep pairs are few: a fixed grid strides over the queue
    hipLaunchKernelGGL(k_narrow_epa, dim3(epa_groups), dim3(64), 0, s, B, sp, pa, pb, epa_list, cap_pairs, p_np, p_pt,
                       p_nf, cnt);
}
void launch_compact(const uint32_t* pa, const uint32_t* pb, uint32_t cap_pairs, const uint32_t* p_np,
                    const float4* p_pt, const float4* p_nf, const uint32_t* moff, uint32_t cap, uint2* m_ab,
                    uint32_t* m_np, float4* c_pt, float4* c_nf, Counters* cnt, hipStream_t s) {
    hipLaunchKernelGGL(k_compact, grid_bounded(cap_pairs), dim3(256), 0, s, pa, pb, p_np, p_pt, p_nf, moff, cap, m_ab,
                       m_np, c_pt, c_nf, cnt);
}
int colour_blocks_per_cu() {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_colour, 256, 0) != hipSuccess) return 1;
    return nb < 1 ? 1 : nb;
}
static PrevCache make_prev(const PrevCacheArgs& pca) {
    PrevCache pc;
    pc.key = pca.key;
    pc.slot = pca.slot;
    pc.mask = pca.mask;
    pc.lam = pca.lam;
    pc.fid = pca.fid;
    pc.np = pca.np;
    pc.colour = pca.colour;
    pc.cap = pca.cap;
    pc.valid = pca.valid;
    return pc;
}
void launch_colour(const Bodies& B, const uint2* m_ab, uint32_t cap_man, uint32_t* col, uint64_t* prio,
                   uint64_t* bmax0, uint64_t* bmax1, uint64_t* bmask, uint32_t* bdeg, uint32_t* list, GridBar* bar,
                   Counters* cnt, uint32_t blocks, const PrevCacheArgs& prev, uint32_t split, hipStream_t s) {
    constexpr uint32_t SMALL = 16384;  // longest list coloured by one workgroup
    hipLaunchKernelGGL(k_colour_keep, grid_bounded(cap_man), dim3(256), 0, s, B, m_ab, col, prio, bmax0, bmask, bdeg, list,
                       cnt, make_prev(prev), RECOLOUR_PERIOD, split ? cap_man : 0u);
    hipLaunchKernelGGL(k_colour_small, dim3(1), dim3(1024), 0, s, B, m_ab, col, prio, bmax0, bmax1, bmask, list, cnt,
                       65000u, SMALL);
    hipLaunchKernelGGL(k_colour, dim3(blocks), dim3(256), 0, s, B, m_ab, col, prio, bmax0, bmax1, bmask, list, bar,
                       cnt, 65000u, SMALL);
}
void launch_col_sort_input(const Bodies& B, const uint2* m_ab, const uint32_t* col, uint32_t cap, uint32_t* key,
                           uint32_t* val, const Counters* cnt, uint32_t split, hipStream_t s) {
    hipLaunchKernelGGL(k_col_sort_input, grid_bounded(cap), dim3(256), 0, s, B, m_ab, col, cap, key, val, cnt, split);
}
void launch_pgs_local(const Bodies& B, const StepParams& sp, const SolverRows& R, Counters* cnt, uint32_t cap,
                      const uint32_t* skey, uint2* list, float tol, uint32_t max_sweeps, uint32_t groups,
                      hipStream_t s, hipEvent_t ev0, hipEvent_t ev1, const uint64_t* flow_bmask) {
    if (max_sweeps == 0u) return;
    hipLaunchKernelGGL(k_pgs_group, dim3(groups), dim3(256), 0, s, B, sp, R, cnt, cap, list, skey, tol, max_sweeps,
                       flow_bmask);
    const uint32_t windows = (cap + BUNDLE_SLOTS - 1u) / BUNDLE_SLOTS;
    if (ev0) (void)hipEventRecord(ev0, s);
    hipLaunchKernelGGL(k_pgs_bundle, dim3(std::max(1u, std::min(windows, 16384u))), dim3(64), BUNDLE_LDS_BYTES, s, B,
                       sp, R, cnt, cap, skey, tol, max_sweeps);
    if (ev1) (void)hipEventRecord(ev1, s);
}
void launch_prep(const Bodies& B, const StepParams& sp, const uint32_t* order, uint32_t cap, const uint2* m_ab,
                 const uint32_t* m_np, const float4* c_pt, const float4* c_nf, const SolverRows& R,
                 const PrevCacheArgs& pca, const float4* warm_in, const uint32_t* col, const uint64_t* bmask,
                 const uint32_t* bdeg, Counters* cnt, float* feat, const PrepExtra& X, hipStream_t s) {
    const PrevCache pc = make_prev(pca);
    hipLaunchKernelGGL(k_prep, grid_bounded(cap), dim3(256), 0, s, B, sp, order, cap, m_ab, m_np, c_pt, c_nf, R, pc,
                       warm_in, col, bmask, bdeg, cnt, feat, X);
}

// Learned warm start: the network predicts the correction to the previous
// step's impulse of the same feature (already in the rows, zero for a new
// point) in units of u = mN g dt; lambda = previous + output * u, projected onto
// lambda_n >= 0 and the step's friction box |lambda_t| <= mu B.
__global__ void __launch_bounds__(256) k_mlp_apply(SolverRows R, const Counters* __restrict__ cnt, uint32_t cap,
                                                   const float* __restrict__ Y, StepParams sp) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= cnt->m_eff || cnt->sticky != 0u) return;
    const uint32_t np = R.hdr[s].z;
    const float gdt = MLP_G_REF * sp.dt;
    for (uint32_t k = 0; k < np; ++k) {
        const size_t o = (size_t)k * cap + s;
        const float u = R.J[(size_t)k * JROWS * cap + s].w * gdt;  // mN
        const float* y = Y + ((size_t)4 * s + k) * MLP_OUT;
        const float4 Lp = R.lam[o];
        // a non-finite output leaves the point exactly at its cached impulse (as the plain warm
        // start has it); finite outputs are capped at MLP_Y_MAX units, so any network gives a
        // finite start the sweeps then correct
        if (!(isfinite(y[0]) && isfinite(y[1]) && isfinite(y[2]))) continue;
        auto cap_y = [](float v) { return v < -MLP_Y_MAX ? -MLP_Y_MAX : (v > MLP_Y_MAX ? MLP_Y_MAX : v); };
        const float y0 = cap_y(y[0]), y1 = cap_y(y[1]), y2 = cap_y(y[2]);
        float ln = Lp.x + y0 * u;
        ln = ln > 0.0f ? ln : 0.0f;
        const float maxf = R.J[((size_t)k * JROWS + 8) * cap + s].w;  // the step's friction bound
        float t1 = Lp.y + y1 * u, t2 = Lp.z + y2 * u;
        t1 = t1 < -maxf ? -maxf : (t1 > maxf ? maxf : t1);
        t2 = t2 < -maxf ? -maxf : (t2 > maxf ? maxf : t2);
        const float4 L = make_float4(ln, t1, t2, 0.0f);
        R.lam[o] = L;
        R.lam0[o] = L;
    }
}
void launch_mlp_apply(const SolverRows& R, const Counters* cnt, uint32_t cap, const float* Y, const StepParams& sp,
                      hipStream_t s) {
    hipLaunchKernelGGL(k_mlp_apply, grid_for(cap), dim3(256), 0, s, R, cnt, cap, Y, sp);
}
int pgs_own_setup(uint32_t max_lds_bytes, uint32_t lds_bytes) {
    if (hipFuncSetAttribute((const void*)k_pgs_own, hipFuncAttributeMaxDynamicSharedMemorySize, (int)max_lds_bytes) !=
        hipSuccess)
        return 0;
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_pgs_own, OWN_WAVE, lds_bytes) != hipSuccess) return 0;
    return nb;
}
void launch_pgs_flow(const Bodies& B, const StepParams& sp, const SolverRows& R, uint32_t cap, const uint64_t* bmask,
                     uint32_t sweep0, uint32_t nsweeps, Counters* cnt,
                     uint32_t spin_limit, uint32_t groups, uint32_t lds_bytes, uint32_t tune, hipStream_t s,
                     hipEvent_t ev0, hipEvent_t ev1, const SweepStop& stop, bool inited) {
    if (!nsweeps) return;
    if (!inited) hipLaunchKernelGGL(k_flow_init, grid_bounded(cap), dim3(256), 0, s, B, R, bmask, cnt);
    if (ev0) (void)hipEventRecord(ev0, s);
    hipLaunchKernelGGL(k_pgs_own, dim3(groups ? groups : 1u), dim3(OWN_WAVE), lds_bytes, s, B, sp, R, cnt, cap, sweep0,
                       nsweeps, lds_bytes / (uint32_t)sizeof(float4), spin_limit, tune, bmask, stop);
    if (ev1) (void)hipEventRecord(ev1, s);
    hipLaunchKernelGGL(k_flow_finish, grid_bounded(cap), dim3(256), 0, s, B, R, cnt);
}
void launch_pgs(const Bodies& B, const StepParams& sp, const SolverRows& R, uint32_t s0, uint32_t s1, uint32_t cap,
                uint32_t first, hipStream_t s) {
    if (s1 > s0) hipLaunchKernelGGL(k_pgs, grid_for(s1 - s0), dim3(256), 0, s, B, sp, R, s0, s1, cap, first);
}
void launch_residual(const Bodies& B, const StepParams& sp, const SolverRows& R, const Counters* cnt, uint32_t cap,
                     uint32_t* out_bits, hipStream_t s) {
    hipLaunchKernelGGL(k_residual, grid_for(cap), dim3(256), 0, s, B, sp, R, cnt, cap, out_bits);
}
// Empties the entries the table holds (the positions its last build recorded).
__global__ void k_cache_clear(uint64_t* __restrict__ key, const uint32_t* __restrict__ pos, const Counters* cnt,
                              uint32_t set) {
    if (cnt->sticky != 0u) return;
    const uint32_t m = cnt->cache_m[set];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < m; i += gridDim.x * blockDim.x) key[pos[i]] = EMPTY_KEY;
}

void launch_cache_build(const SolverRows& R, Counters* cnt, uint32_t cap, uint64_t* key, uint32_t* slot,
                        uint32_t mask, uint32_t* np_out, uint32_t* pos, uint32_t set, hipStream_t s, bool clear) {
    if (clear) hipLaunchKernelGGL(k_cache_clear, grid_bounded(cap), dim3(256), 0, s, key, pos, cnt, set);
    hipLaunchKernelGGL(k_cache_build, grid_bounded(cap), dim3(256), 0, s, R, cnt,
                       CacheTable{key, slot, mask, np_out, pos, set});
}
void launch_cache_rehash(const uint64_t* okey, const uint32_t* oslot, uint32_t on, uint64_t* key, uint32_t* slot,
                         uint32_t mask, Counters* cnt, uint32_t* pos, hipStream_t s) {
    if (on)
        hipLaunchKernelGGL(k_cache_rehash, grid_for(on), dim3(256), 0, s, okey, oslot, on, key, slot, mask, cnt, pos);
}
void launch_integrate(const Bodies& B, const StepParams& sp, uint32_t n, Counters* cnt, uint32_t sweeps,
                      hipStream_t s, const SolverRows* R, const CacheTable* C) {
    const uint32_t g = std::max(n, 1u);
    hipLaunchKernelGGL(k_integrate, grid_for(g), dim3(256), 0, s, B, sp, n, cnt, sweeps, R ? *R : SolverRows{},
                       C ? *C : CacheTable{});
}

void launch_isl_union(const Bodies& B, uint32_t n, const uint2* m_ab, uint32_t cap_man, Counters* cnt,
                      hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_isl_union, grid_bounded(cap_man), dim3(256), 0, s, B, m_ab, cap_man, n, cnt);
}
void launch_islands(const Bodies& B, const StepParams& sp, uint32_t n, Counters* cnt, hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(k_isl_label, grid_for(n), dim3(256), 0, s, B, sp, n, cnt);
    hipLaunchKernelGGL(k_isl_sleep, grid_for(n), dim3(256), 0, s, B, sp, n, cnt);
}
}  // namespace axm
