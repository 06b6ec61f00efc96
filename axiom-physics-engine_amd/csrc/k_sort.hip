// Device-wide exclusive scan and stable LSD radix sort for gfx950 (wave64).
//
// Radix sort: 8-bit digits, 256-thread workgroups over 4096-key tiles.
//   pass = histogram (LDS atomics) -> exclusive scan of the digit-major
//   histogram -> stable scatter.  Stability comes from ranking keys in index
//   order: each wave finds the lanes sharing its digit with 8 ballots
//   (a match-any over the digit bits), ranks itself by the population count of
//   lower peer lanes, and per-wave digit counts are prefixed in LDS.
// Scan: one pass, tiles chained by decoupled look-back (below).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "launch.hpp"

namespace axm {

constexpr int RS_THREADS = 256;
constexpr int RS_ITEMS = 4;  // keys per thread: small tiles, so that a 100k-key sort fills the GPU
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;
constexpr int SC_THREADS = 256;
constexpr int SC_ITEMS = 16;
constexpr int SC_TILE = SC_THREADS * SC_ITEMS;
// workgroups of a scan launch (each takes tiles in turn): enough to keep many tiles in
// flight; fewer when the count is read on the device and usually far below the capacity
constexpr uint32_t SC_MAX_GROUPS = 512, SC_MAX_GROUPS_DEV = 64;

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Inclusive wave scan (wave64) of a u32.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    const uint32_t lane = lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = __shfl_up(v, o, 64);
        if (lane >= (uint32_t)o) v += t;
    }
    return v;
}

// Exclusive block scan (blockDim = 256 = 4 waves); returns the block total in `total`.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t& total) {
    __shared__ uint32_t s_w[RS_THREADS / 64];
    const uint32_t lane = lane_id(), wv = threadIdx.x >> 6;
    const uint32_t inc = wave_incl_scan(v);
    if (lane == 63) s_w[wv] = inc;
    __syncthreads();
    uint32_t base = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < RS_THREADS / 64; ++w) {
        const uint32_t x = s_w[w];
        if ((uint32_t)w < wv) base += x;
        tot += x;
    }
    __syncthreads();
    total = tot;
    return base + inc - v;
}

// Single-pass exclusive scan with decoupled look-back.  Workgroup b starts with
// tile b; when there are more tiles than workgroups, the rest are taken in turn
// from a counter (so every lower tile is already held by a running workgroup).  A
// workgroup publishes its tile total, then reads the status words of the 64 tiles
// below it at once (one per lane of its first wave), moving down 64 at a time until
// it meets an inclusive prefix, and publishes its own.  Status word: bits 62-63
// state (0 none, 1 tile total, 2 inclusive prefix), bits 0-31 the value.  The
// launch is at most a few hundred workgroups, all resident together.  Atomics on
// one address (or line) queue one after another, so the finish count that finds the
// last workgroup is spread over SC_ROWS counters one line apart, each row's last
// arrival adding to one top counter; the last workgroup to finish clears the status
// words it used and the counters, so every call starts from zeros (the scratch is
// zeroed once when allocated).  Scratch layout (u32 words): [0] tile counter,
// [SC_ROW0 + 32 r] row r's finish count, [SC_TOP] rows finished, then from
// SC_STATUS the nb 64-bit status words.
constexpr uint64_t SC_AGG = 1ull << 62, SC_PRE = 2ull << 62;
constexpr uint32_t SC_ROWS = 16, SC_ROW0 = 32, SC_TOP = SC_ROW0 + 32 * SC_ROWS, SC_STATUS = SC_TOP + 32;

// With n_dev the element count is min(n, *n_dev), read on the device (nb, the
// tiles the scratch holds, is sized for n).  With flags, the input is read as
// (in[i] != 0).  `in` may equal `out` (each thread reads its elements before it
// writes them).
__global__ void __launch_bounds__(SC_THREADS) k_scan_1p(const uint32_t* in, uint32_t* out, uint32_t n_cap,
                                                         uint32_t* __restrict__ tmp, uint32_t nb,
                                                         uint32_t* __restrict__ total_out,
                                                         const uint32_t* __restrict__ n_dev, int flags,
                                                         uint32_t* __restrict__ err) {
    __shared__ uint32_t s_tile, s_prefix;
    uint64_t* st = reinterpret_cast<uint64_t*>(tmp + SC_STATUS);
    const uint32_t n = n_dev ? min(n_cap, *n_dev) : n_cap;
    const uint32_t nt = n == 0u ? 1u : (n + SC_TILE - 1u) / SC_TILE;  // tiles with elements (tile 0 always)
    // 16-byte loads and stores when both arrays allow them (whole items of a full tile)
    const bool vec = ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out)) & 15u) == 0u;
    uint32_t tile = blockIdx.x;
    while (tile < nt) {
        // each thread owns SC_ITEMS consecutive elements of the tile
        const uint32_t base = tile * SC_TILE + threadIdx.x * SC_ITEMS;
        const bool whole = vec && base + SC_ITEMS <= n;
        uint32_t v[SC_ITEMS];
        if (whole) {
            const uint4* p = reinterpret_cast<const uint4*>(in + base);
#pragma unroll
            for (int q = 0; q < SC_ITEMS / 4; ++q) {
                const uint4 x = p[q];
                v[4 * q] = x.x;
                v[4 * q + 1] = x.y;
                v[4 * q + 2] = x.z;
                v[4 * q + 3] = x.w;
            }
        } else {
#pragma unroll
            for (int k = 0; k < SC_ITEMS; ++k) v[k] = base + k < n ? in[base + k] : 0u;
        }
        uint32_t sum = 0;
#pragma unroll
        for (int k = 0; k < SC_ITEMS; ++k) {
            if (flags) v[k] = v[k] != 0u ? 1u : 0u;
            sum += v[k];
        }
        uint32_t tot;
        const uint32_t ex = block_excl_scan(sum, tot);
        if (threadIdx.x < 64u) {
            // the first wave publishes the tile and looks back over 64 lower tiles at a time
            const uint32_t lane = threadIdx.x;
            uint32_t prefix = 0;
            if (tile == 0) {
                if (lane == 0u) __hip_atomic_store(&st[0], SC_PRE | tot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            } else {
                if (lane == 0u) __hip_atomic_store(&st[tile], SC_AGG | tot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                // lower tiles are held by running workgroups, so each publishes; the wall-clock
                // bound only keeps a broken invariant from hanging the device
                const uint64_t t0 = wall_clock64();
                int j0 = (int)tile - 1;  // window: tiles j0 - lane
                for (;;) {
                    const int j = j0 - (int)lane;
                    // below tile 0 reads as an inclusive prefix of 0 (tile 0 always publishes one first)
                    const uint64_t w =
                        j >= 0 ? __hip_atomic_load(&st[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : SC_PRE;
                    const uint32_t state = (uint32_t)(w >> 62);
                    const uint64_t pre = __ballot(state == 2u), none = __ballot(state == 0u);
                    // the nearest inclusive prefix, and every tile between it and this one published
                    const uint32_t first = pre ? (uint32_t)__ffsll((unsigned long long)pre) - 1u : 64u;
                    const uint64_t upto = first >= 63u ? ~0ull : (2ull << first) - 1ull;
                    if (none & upto) {
                        if (wall_clock64() - t0 > 200000000ull) {  // 2 s: report, do not hang
                            if (lane == 0u && err) atomicOr(err, STICKY_TIMEOUT);
                            break;
                        }
                        __builtin_amdgcn_s_sleep(1);
                        continue;
                    }
                    uint32_t val = lane <= first ? (uint32_t)w : 0u;
                    val = __shfl(wave_incl_scan(val), 63, 64);
                    prefix += val;
                    if (first < 64u) break;
                    j0 -= 64;
                }
                if (lane == 0u)
                    __hip_atomic_store(&st[tile], SC_PRE | (uint32_t)(prefix + tot), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
            }
            if (lane == 0u) s_prefix = prefix;
        }
        __syncthreads();
        uint32_t run = s_prefix + ex;
        if (whole) {
            uint4* p = reinterpret_cast<uint4*>(out + base);
#pragma unroll
            for (int q = 0; q < SC_ITEMS / 4; ++q) {
                uint4 x;
                x.x = run;
                run += v[4 * q];
                x.y = run;
                run += v[4 * q + 1];
                x.z = run;
                run += v[4 * q + 2];
                x.w = run;
                run += v[4 * q + 3];
                p[q] = x;
            }
        } else {
#pragma unroll
            for (int k = 0; k < SC_ITEMS; ++k) {
                if (base + k < n) out[base + k] = run;
                run += v[k];
            }
        }
        if (tile == nt - 1 && threadIdx.x == SC_THREADS - 1) {
            out[n] = s_prefix + tot;
            if (total_out) *total_out = s_prefix + tot;
        }
        if (nt <= gridDim.x) break;  // every tile was a first tile
        if (threadIdx.x == 0) s_tile = gridDim.x + atomicAdd(&tmp[0], 1u);
        __syncthreads();  // s_prefix is rewritten by the next tile; the ticket is in
        tile = s_tile;
        __syncthreads();  // every thread has the ticket before thread 0 takes the next one
    }
    __shared__ bool s_last;
    if (threadIdx.x == 0) {
        // finish count: this workgroup's row, then the top counter once the row is complete
        const uint32_t G = gridDim.x, r = blockIdx.x % SC_ROWS;
        const uint32_t in_row = (G - r + SC_ROWS - 1u) / SC_ROWS, rows = G < SC_ROWS ? G : SC_ROWS;
        bool last = false;
        if (atomicAdd(&tmp[SC_ROW0 + 32u * r], 1u) == in_row - 1u) {
            // every member of the row has added
            __hip_atomic_store(&tmp[SC_ROW0 + 32u * r], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            last = atomicAdd(&tmp[SC_TOP], 1u) == rows - 1u;
        }
        s_last = last;
    }
    __syncthreads();
    if (s_last) {
        // every workgroup has finished its look-back: reset the scratch for the next call
        for (uint32_t k = threadIdx.x; k < nt; k += SC_THREADS)
            __hip_atomic_store(&st[k], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (threadIdx.x == 0) {
            __hip_atomic_store(&tmp[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&tmp[SC_TOP], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

size_t scan_tmp_words(uint32_t n) { return SC_STATUS + 2 * (size_t)((n + SC_TILE - 1) / SC_TILE); }

__global__ void k_zero1(uint32_t* p, uint32_t* q) {
    *p = 0;
    if (q) *q = 0;
}

void scan_exclusive(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* tmp, hipStream_t s, uint32_t* total_out,
                    const uint32_t* n_dev, bool flags, uint32_t* err) {
    if (n == 0) {
        hipLaunchKernelGGL(k_zero1, dim3(1), dim3(1), 0, s, out, total_out);
        return;
    }
    const uint32_t nb = (n + SC_TILE - 1) / SC_TILE;
    const uint32_t groups = std::min<uint32_t>(nb, n_dev ? SC_MAX_GROUPS_DEV : SC_MAX_GROUPS);
    hipLaunchKernelGGL(k_scan_1p, dim3(groups), dim3(SC_THREADS), 0, s, in, out, n, tmp, nb, total_out, n_dev,
                       flags ? 1 : 0, err);
}

// ---------------------------------------------------------------- radix sort
template <class K> __device__ __forceinline__ uint32_t digit_of(K k, int shift) { return (uint32_t)(k >> shift) & 0xFFu; }

__device__ __forceinline__ uint32_t eff_count(uint32_t cap, const uint32_t* n_dev) {
    if (!n_dev) return cap;
    const uint32_t v = *n_dev;
    return v < cap ? v : cap;
}

// Per-tile digit counts, digit-major: hist[d * nt + tile] for the nt tiles that hold
// keys (nt from the count read on the device; the grid is sized for the capacity).
// The first workgroup writes 256 nt, the length of the scan that follows, to *len.
template <class K>
__global__ void __launch_bounds__(RS_THREADS) k_rs_hist(const K* __restrict__ keys, uint32_t cap,
                                                        const uint32_t* __restrict__ n_dev, int shift,
                                                        uint32_t* __restrict__ hist, uint32_t* __restrict__ len) {
    const uint32_t n = eff_count(cap, n_dev);
    const uint32_t nt = (n + RS_TILE - 1) / RS_TILE;
    if (blockIdx.x == 0 && threadIdx.x == 0) *len = 256u * nt;
    if (blockIdx.x >= nt) return;
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t base = blockIdx.x * RS_TILE;
#pragma unroll 4
    for (int it = 0; it < RS_ITEMS; ++it) {
        const uint32_t i = base + it * RS_THREADS + threadIdx.x;
        if (i < n) atomicAdd(&h[digit_of(keys[i], shift)], 1u);
    }
    __syncthreads();
    hist[(size_t)threadIdx.x * nt + blockIdx.x] = h[threadIdx.x];
}

template <class K>
__global__ void __launch_bounds__(RS_THREADS) k_rs_scatter(const K* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                           K* __restrict__ kout, uint32_t* __restrict__ vout, uint32_t cap,
                                                           const uint32_t* __restrict__ n_dev, int shift,
                                                           const uint32_t* __restrict__ offs) {
    const uint32_t n = eff_count(cap, n_dev);
    if (blockIdx.x * RS_TILE >= n) return;
    const uint32_t nt = (n + RS_TILE - 1) / RS_TILE;
    __shared__ uint32_t s_off[256];
    __shared__ uint32_t s_wc[RS_THREADS / 64][256];
    const uint32_t tid = threadIdx.x, lane = lane_id(), wv = tid >> 6;
    s_off[tid] = offs[(size_t)tid * nt + blockIdx.x];
    const uint64_t lt = (1ull << lane) - 1ull;
    const uint32_t base = blockIdx.x * RS_TILE;
    for (int it = 0; it < RS_ITEMS; ++it) {
        const uint32_t i = base + it * RS_THREADS + tid;
        const bool valid = i < n;
        K k = 0;
        uint32_t v = 0, d = 256u;
        if (valid) {
            k = kin[i];
            v = vin[i];
            d = digit_of(k, shift);
        }
#pragma unroll
        for (int w = 0; w < RS_THREADS / 64; ++w) s_wc[w][tid] = 0;
        // lanes with the same digit (9 bits so that invalid lanes, d = 256, form their own group)
        uint64_t peers = __ballot(1);
#pragma unroll
        for (int b = 0; b < 9; ++b) {
            const uint64_t bal = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? bal : ~bal;
        }
        const uint32_t rank = (uint32_t)__popcll(peers & lt);
        __syncthreads();
        if (valid && rank == 0) s_wc[wv][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        uint32_t pos = 0;
        if (valid) {
            pos = s_off[d] + rank;
            for (uint32_t w = 0; w < wv; ++w) pos += s_wc[w][d];
        }
        __syncthreads();
        uint32_t add = 0;
#pragma unroll
        for (int w = 0; w < RS_THREADS / 64; ++w) add += s_wc[w][tid];
        s_off[tid] += add;
        if (valid) {
            kout[pos] = k;
            vout[pos] = v;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- cooperative sort
// For a count read on the device (usually far below the capacity the grid would be
// sized for), one launch of at most CO_GROUPS workgroups, all resident, runs every
// pass with one grid barrier per pass instead of three launches per pass.  Per-tile
// digit counts live in three rotating buffers (tile-major: [tile * 256 + digit]): pass
// p reads its counts from H[p % 3]; its scatter adds each key's next digit at the key's
// destination tile into H[(p + 1) % 3]; and it clears H[(p + 2) % 3], which pass p - 1
// read and pass p + 1 adds into.  A prologue writes pass 0's counts and clears the
// buffer pass 0 adds into.  Each workgroup finds its tiles' offsets from the counts
// themselves (digit totals over all tiles, then the tiles below its own).
constexpr uint32_t CO_GROUPS = 64, CO_HDR = 4;

size_t coop_tmp_words(uint32_t n) {
    const uint32_t nb = (n + RS_TILE - 1) / RS_TILE;
    return CO_HDR + 3ull * 256u * (nb ? nb : 1u);
}

// Barrier over the grid (every workgroup resident): each wave's stores complete, then one
// thread per workgroup arrives (agent-scope release) and waits for the last arrival
// (agent-scope acquire); a bounded wait ORs STICKY_TIMEOUT into *err instead of hanging.
// bar[0] counts arrivals and returns to 0; bar[1] is the generation.
__device__ __forceinline__ void co_sync(uint32_t* bar, uint32_t* err) {
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t g = __hip_atomic_load(&bar[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t arrived = __hip_atomic_fetch_add(&bar[0], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (arrived == gridDim.x - 1) {
            __hip_atomic_store(&bar[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_add(&bar[1], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            const uint64_t t0 = wall_clock64();
            while (__hip_atomic_load(&bar[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == g) {
                __builtin_amdgcn_s_sleep(2);
                if (wall_clock64() - t0 > 200000000ull) {  // 2 s: report, do not hang
                    if (err) atomicOr(err, STICKY_TIMEOUT);
                    break;
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
    }
    __syncthreads();
}

// Sum of counts[t * 256 + d] over tiles t0 <= t < t1, eight loads in flight at a time.
__device__ __forceinline__ uint32_t sum_tiles(const uint32_t* __restrict__ h, uint32_t t0, uint32_t t1, uint32_t d) {
    uint32_t s = 0;
    uint32_t t = t0;
    for (; t + 8u <= t1; t += 8u) {
        uint32_t a[8];
#pragma unroll
        for (uint32_t u = 0; u < 8u; ++u) a[u] = h[(size_t)(t + u) * 256u + d];
#pragma unroll
        for (uint32_t u = 0; u < 8u; ++u) s += a[u];
    }
    uint32_t a[8];
#pragma unroll
    for (uint32_t u = 0; u < 8u; ++u) a[u] = t + u < t1 ? h[(size_t)(t + u) * 256u + d] : 0u;
#pragma unroll
    for (uint32_t u = 0; u < 8u; ++u) s += a[u];
    return s;
}

template <class K>
__global__ void __launch_bounds__(RS_THREADS) k_rs_coop(K* __restrict__ k0, uint32_t* __restrict__ v0,
                                                        K* __restrict__ k1, uint32_t* __restrict__ v1, uint32_t cap,
                                                        const uint32_t* __restrict__ n_dev, int passes,
                                                        uint32_t* __restrict__ tmp, uint32_t* __restrict__ err) {
    __shared__ uint32_t s_off[256], s_cnt[256];
    __shared__ uint32_t s_wc[RS_THREADS / 64][256];
    const uint32_t n = eff_count(cap, n_dev);
    const uint32_t nt = (n + RS_TILE - 1) / RS_TILE;
    const uint32_t nbc = (cap + RS_TILE - 1) / RS_TILE;
    uint32_t* bar = tmp;
    uint32_t* const H0 = tmp + CO_HDR;
    const size_t hsz = (size_t)256 * nbc;
    const uint32_t tid = threadIdx.x, lane = lane_id(), wv = tid >> 6;
    const uint64_t lt = (1ull << lane) - 1ull;
    // prologue: pass 0's counts per tile, and the buffer pass 0 adds into cleared
    for (uint32_t t = blockIdx.x; t < nt; t += gridDim.x) {
        s_cnt[tid] = 0u;
        __syncthreads();
#pragma unroll
        for (int it = 0; it < RS_ITEMS; ++it) {
            const uint32_t i = t * RS_TILE + it * RS_THREADS + tid;
            if (i < n) atomicAdd(&s_cnt[digit_of(k0[i], 0)], 1u);
        }
        __syncthreads();
        H0[(size_t)t * 256u + tid] = s_cnt[tid];
        H0[hsz + (size_t)t * 256u + tid] = 0u;
        __syncthreads();
    }
    co_sync(bar, err);
    K* ki = k0;
    K* ko = k1;
    uint32_t* vi = v0;
    uint32_t* vo = v1;
    for (int p = 0; p < passes; ++p) {
        const uint32_t* Hp = H0 + hsz * (uint32_t)(p % 3);
        uint32_t* Hn = H0 + hsz * (uint32_t)((p + 1) % 3);
        uint32_t* Hz = H0 + hsz * (uint32_t)((p + 2) % 3);
        const bool more = p + 1 < passes;
        const int shift = 8 * p;
        // digit tid's place among all keys: the counts of the lower digits over all tiles
        // the first tile's keys and values are loaded before the counts, so that the two
        // rounds of loads overlap
        K kk[RS_ITEMS];
        uint32_t vv[RS_ITEMS];
        auto load_tile = [&](uint32_t t) {
#pragma unroll
            for (int it = 0; it < RS_ITEMS; ++it) {
                const uint32_t i = t * RS_TILE + it * RS_THREADS + tid;
                kk[it] = i < n ? ki[i] : (K)0;
                vv[it] = i < n ? vi[i] : 0u;
            }
        };
        if (blockIdx.x < nt) load_tile(blockIdx.x);
        // up to 32 tiles: all of digit tid's counts in registers from one round of loads,
        // which give the total and the count below each of this workgroup's tiles
        const bool few = nt <= 32u;
        uint32_t c32[32];
#pragma unroll
        for (uint32_t u = 0; u < 32u; ++u) c32[u] = few && u < nt ? Hp[(size_t)u * 256u + tid] : 0u;
        uint32_t tot = 0;
        if (few) {
#pragma unroll
            for (uint32_t u = 0; u < 32u; ++u) tot += c32[u];
        } else {
            tot = sum_tiles(Hp, 0u, nt, tid);
        }
        uint32_t all;
        const uint32_t base = block_excl_scan(tot, all);
        uint32_t below = 0;  // count of digit tid in the tiles before the current one
        uint32_t tb = 0;
        for (uint32_t t = blockIdx.x; t < nt; t += gridDim.x) {
            if (t != blockIdx.x) load_tile(t);
            if (few) {
                below = 0;
#pragma unroll
                for (uint32_t u = 0; u < 32u; ++u) below += u < t ? c32[u] : 0u;
            } else {
                below += sum_tiles(Hp, tb, t, tid);
                tb = t;
            }
            s_off[tid] = base + below;
            if (more) Hz[(size_t)t * 256u + tid] = 0u;
            // stable scatter, item by item in index order (as k_rs_scatter), adding each
            // key's next digit at its destination tile
#pragma unroll
            for (int it = 0; it < RS_ITEMS; ++it) {
                const uint32_t i = t * RS_TILE + it * RS_THREADS + tid;
                const bool valid = i < n;
                const K k = kk[it];
                const uint32_t v = vv[it];
                const uint32_t d = valid ? digit_of(k, shift) : 256u;
#pragma unroll
                for (int w = 0; w < RS_THREADS / 64; ++w) s_wc[w][tid] = 0;
                uint64_t peers = __ballot(1);
#pragma unroll
                for (int b = 0; b < 9; ++b) {
                    const uint64_t bal = __ballot((d >> b) & 1u);
                    peers &= ((d >> b) & 1u) ? bal : ~bal;
                }
                const uint32_t rank = (uint32_t)__popcll(peers & lt);
                __syncthreads();
                if (valid && rank == 0) s_wc[wv][d] = (uint32_t)__popcll(peers);
                __syncthreads();
                uint32_t pos = 0;
                if (valid) {
                    pos = s_off[d] + rank;
                    for (uint32_t w = 0; w < wv; ++w) pos += s_wc[w][d];
                }
                __syncthreads();
                uint32_t add = 0;
#pragma unroll
                for (int w = 0; w < RS_THREADS / 64; ++w) add += s_wc[w][tid];
                s_off[tid] += add;
                if (valid && pos < n) {
                    ko[pos] = k;
                    vo[pos] = v;
                    if (more) atomicAdd(&Hn[(size_t)(pos / RS_TILE) * 256u + digit_of(k, shift + 8)], 1u);
                }
            }
            __syncthreads();
        }
        if (more) co_sync(bar, err);
        K* tk = ki; ki = ko; ko = tk;
        uint32_t* tv = vi; vi = vo; vo = tv;
    }
}

size_t radix_tmp_words(uint32_t n) {
    const uint32_t nb = (n + RS_TILE - 1) / RS_TILE;
    const size_t split = (size_t)256 * nb + 2 + scan_tmp_words(256u * nb);
    const size_t coop = coop_tmp_words(n);
    return split > coop ? split : coop;
}

template <class K>
static int radix_sort_impl(K* keys, uint32_t* vals, K* keys_alt, uint32_t* vals_alt, uint32_t n, int end_bit,
                           uint32_t* tmp, hipStream_t s, const uint32_t* n_dev, uint32_t* err) {
    if (n == 0) return 0;
    const uint32_t nb = (n + RS_TILE - 1) / RS_TILE;
    if (n_dev) {
        // the count is read on the device: one cooperative launch
        const int passes = (end_bit + 7) / 8;
        const uint32_t groups = nb < CO_GROUPS ? nb : CO_GROUPS;
        hipLaunchKernelGGL(k_rs_coop<K>, dim3(groups), dim3(RS_THREADS), 0, s, keys, vals, keys_alt, vals_alt, n, n_dev,
                           passes, tmp, err);
        return passes & 1;
    }
    uint32_t* hist = tmp;
    uint32_t* offs = tmp;  // scanned in place
    // the histogram length in use (256 x tiles with keys); the word before it takes the
    // scan's total (out[len])
    uint32_t* len = tmp + (size_t)256 * nb + 1;
    uint32_t* scan_tmp = tmp + (size_t)256 * nb + 2;  // 8-byte aligned for the scan's status words
    K* ki = keys;
    K* ko = keys_alt;
    uint32_t* vi = vals;
    uint32_t* vo = vals_alt;
    int passes = 0;
    for (int shift = 0; shift < end_bit; shift += 8) {
        hipLaunchKernelGGL(k_rs_hist<K>, dim3(nb), dim3(RS_THREADS), 0, s, ki, n, n_dev, shift, hist, len);
        scan_exclusive(hist, offs, 256u * nb, scan_tmp, s, nullptr, len, false, err);
        hipLaunchKernelGGL(k_rs_scatter<K>, dim3(nb), dim3(RS_THREADS), 0, s, ki, vi, ko, vo, n, n_dev, shift, offs);
        K* tk = ki; ki = ko; ko = tk;
        uint32_t* tv = vi; vi = vo; vo = tv;
        ++passes;
    }
    return passes & 1;
}

int radix_sort_u64(uint64_t* keys, uint32_t* vals, uint64_t* keys_alt, uint32_t* vals_alt, uint32_t n, int end_bit,
                   uint32_t* tmp, hipStream_t s, const uint32_t* n_dev, uint32_t* err) {
    return radix_sort_impl<uint64_t>(keys, vals, keys_alt, vals_alt, n, end_bit, tmp, s, n_dev, err);
}

int radix_sort_u32(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt, uint32_t n, int end_bit,
                   uint32_t* tmp, hipStream_t s, const uint32_t* n_dev, uint32_t* err) {
    return radix_sort_impl<uint32_t>(keys, vals, keys_alt, vals_alt, n, end_bit, tmp, s, n_dev, err);
}

}  // namespace axm
