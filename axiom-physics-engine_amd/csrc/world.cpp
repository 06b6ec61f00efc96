// Host side of the C ABI (include/axiom/axm.h): world lifecycle, device buffer
// management and the orchestration of one step on the world's HIP stream.
//
// Error handling follows the reference's no-exception Result/ErrorCode
// convention (include/axiom/core/result.hpp:15-283,
// include/axiom/core/error_code.hpp:21-57): every entry point returns an
// axm_status; a failed step leaves the body state untouched (the canonical
// state is written only by the final integration kernel).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "axiom/axm.h"
#include "k_step.hpp"

using namespace axm;

namespace {

#define AXM_HIP(expr)                                   \
    do {                                                \
        hipError_t e_ = (expr);                         \
        if (e_ != hipSuccess) {                         \
            std::fprintf(stderr, "axm: %s failed: %s\n", #expr, hipGetErrorString(e_)); \
            return AXM_GPU_OPERATION_FAILED;            \
        }                                               \
    } while (0)

#define AXM_TRY(expr)                    \
    do {                                 \
        axm_status s_ = (expr);          \
        if (s_ != AXM_SUCCESS) return s_; \
    } while (0)

uint32_t pow2_at_least(uint64_t v) {
    uint64_t p = 1;
    while (p < v) p <<= 1;
    return (uint32_t)p;
}
int bits_for(uint32_t v) {
    int b = 0;
    while (b < 32 && (1ull << b) <= v) ++b;
    return b;
}

struct Alloc {
    std::vector<void*> ptrs;
    template <class T> axm_status get(T*& p, size_t count) {
        p = nullptr;
        if (count == 0) count = 1;
        void* q = nullptr;
        if (hipMalloc(&q, count * sizeof(T)) != hipSuccess) return AXM_BUFFER_ALLOCATION_FAILED;
        ptrs.push_back(q);
        p = (T*)q;
        return AXM_SUCCESS;
    }
    template <class T> void release(T*& p) {
        if (!p) return;
        auto it = std::find(ptrs.begin(), ptrs.end(), (void*)p);
        if (it != ptrs.end()) ptrs.erase(it);
        hipFree(p);
        p = nullptr;
    }
    void release_all() {
        for (void* p : ptrs) hipFree(p);
        ptrs.clear();
    }
};

}  // namespace

struct axm_world_s {
    axm_world_desc desc{};
    hipStream_t stream = nullptr;
    Alloc mem;
    uint32_t n = 0, cap_n = 0, cap_pairs = 0, cap_man = 0, n_worlds = 1;
    float cell = 1.0f;

    Bodies B{};
    float4 *pre_v = nullptr, *pre_w = nullptr;  // solver input velocities of the last step (kept on request)
    bool have_pre = false;
    uint64_t *key = nullptr, *key_alt = nullptr;
    uint32_t *val = nullptr, *val_alt = nullptr, *sort_tmp = nullptr;
    CellTable ct{};

    uint32_t *pair_cnt = nullptr, *pair_off = nullptr, *scan_tmp = nullptr;
    uint32_t *pa = nullptr, *pb = nullptr;
    uint32_t *p_np = nullptr, *p_flag = nullptr, *man_off = nullptr, *epa_list = nullptr;
    float4 *p_pt = nullptr, *p_nf = nullptr;
    uint32_t pscan_tmp_words = 0;
    uint32_t* pscan_tmp = nullptr;

    uint2* m_ab = nullptr;
    uint32_t* m_np = nullptr;
    float4 *c_pt = nullptr, *c_nf = nullptr;

    uint32_t* col = nullptr;
    uint64_t *prio = nullptr, *bmax = nullptr, *bmask = nullptr;
    uint32_t* bdone = nullptr;      // persistent sweep: finished visits per body
    uint32_t* flow_ctl = nullptr;   // persistent sweep: [0] work queue, [1] spin-limit error
    uint32_t *ckey = nullptr, *ckey_alt = nullptr, *cval = nullptr, *cval_alt = nullptr, *csort_tmp = nullptr;
    uint32_t* col_off = nullptr;
    uint32_t col_off_cap = 0;
    uint32_t* order = nullptr;  // points into cval / cval_alt

    SolverRows R[2]{};
    uint32_t rcap[2] = {0, 0};
    uint32_t* np_slot[2] = {nullptr, nullptr};
    uint64_t* pc_key[2] = {nullptr, nullptr};
    uint32_t* pc_slot[2] = {nullptr, nullptr};
    uint32_t pc_mask[2] = {0, 0};
    int cur = 0;
    bool prev_valid = false;

    uint32_t* res_bits = nullptr;
    uint32_t res_cap = 0;
    Counters* cnt = nullptr;
    Counters* h_cnt = nullptr;
    float4* warm_dev = nullptr;
    uint32_t warm_cap = 0;

    // kernel profiling
    bool prof_on = false;
    std::vector<hipEvent_t> prof_ev;
    axm_kernel_profile prof{};

    // last step
    uint32_t P = 0, M = 0, K = 0, rounds = 0, iters_run = 0;
    bool have_step = false;
    std::vector<uint32_t> h_col_off;
    std::vector<float> h_res;

    ~axm_world_s() {
        if (stream) hipStreamSynchronize(stream);
        for (hipEvent_t e : prof_ev) hipEventDestroy(e);
        mem.release_all();
        if (h_cnt) hipHostFree(h_cnt);
        if (stream) hipStreamDestroy(stream);
    }
};

namespace {

axm_status alloc_bodies(axm_world_s* w, uint32_t cap) {
    Bodies& B = w->B;
    AXM_TRY(w->mem.get(B.x, cap));
    AXM_TRY(w->mem.get(B.q, cap));
    AXM_TRY(w->mem.get(B.vel, cap));
    AXM_TRY(w->mem.get(B.angv, cap));
    AXM_TRY(w->mem.get(B.ii, cap));
    AXM_TRY(w->mem.get(B.shape, cap));
    AXM_TRY(w->mem.get(B.world, cap));
    AXM_TRY(w->mem.get(B.world_first, cap + 1));
    AXM_TRY(w->mem.get(B.sv, cap));
    AXM_TRY(w->mem.get(B.sw, cap));
    AXM_TRY(w->mem.get(B.I0, cap));
    AXM_TRY(w->mem.get(B.I1, cap));
    AXM_TRY(w->mem.get(B.amin, cap));
    AXM_TRY(w->mem.get(B.amax, cap));
    AXM_TRY(w->mem.get(w->pre_v, cap));
    AXM_TRY(w->mem.get(w->pre_w, cap));
    AXM_TRY(w->mem.get(w->key, cap));
    AXM_TRY(w->mem.get(w->key_alt, cap));
    AXM_TRY(w->mem.get(w->val, cap));
    AXM_TRY(w->mem.get(w->val_alt, cap));
    AXM_TRY(w->mem.get(w->sort_tmp, radix_tmp_words(cap)));
    const uint32_t T = pow2_at_least(2ull * cap + 16);
    AXM_TRY(w->mem.get(w->ct.key, T));
    AXM_TRY(w->mem.get(w->ct.val, T));
    w->ct.mask = T - 1;
    AXM_TRY(w->mem.get(w->pair_cnt, cap));
    AXM_TRY(w->mem.get(w->pair_off, cap + 1));
    AXM_TRY(w->mem.get(w->scan_tmp, scan_tmp_words(cap)));
    AXM_TRY(w->mem.get(w->bmax, cap));
    AXM_TRY(w->mem.get(w->bmask, cap));
    AXM_TRY(w->mem.get(w->bdone, cap));
    w->cap_n = cap;
    return AXM_SUCCESS;
}

void free_bodies(axm_world_s* w) {
    Bodies& B = w->B;
    float4** f4s[] = {&B.x, &B.q, &B.vel, &B.angv, &B.ii, &B.shape, &B.sv, &B.sw, &B.I0, &B.I1, &B.amin, &B.amax,
                      &w->pre_v, &w->pre_w};
    for (float4** p : f4s) w->mem.release(*p);
    w->mem.release(B.world);
    w->mem.release(B.world_first);
    w->mem.release(w->key);
    w->mem.release(w->key_alt);
    w->mem.release(w->val);
    w->mem.release(w->val_alt);
    w->mem.release(w->sort_tmp);
    w->mem.release(w->ct.key);
    w->mem.release(w->ct.val);
    w->mem.release(w->pair_cnt);
    w->mem.release(w->pair_off);
    w->mem.release(w->scan_tmp);
    w->mem.release(w->bmax);
    w->mem.release(w->bmask);
    w->mem.release(w->bdone);
    w->cap_n = 0;
}

void free_pairs(axm_world_s* w) {
    w->mem.release(w->pa);
    w->mem.release(w->pb);
    w->mem.release(w->p_np);
    w->mem.release(w->p_flag);
    w->mem.release(w->man_off);
    w->mem.release(w->epa_list);
    w->mem.release(w->p_pt);
    w->mem.release(w->p_nf);
    w->mem.release(w->pscan_tmp);
}

axm_status alloc_pairs(axm_world_s* w, uint32_t cap) {
    free_pairs(w);
    AXM_TRY(w->mem.get(w->pa, cap));
    AXM_TRY(w->mem.get(w->pb, cap));
    AXM_TRY(w->mem.get(w->p_np, cap));
    AXM_TRY(w->mem.get(w->p_flag, cap));
    AXM_TRY(w->mem.get(w->man_off, (size_t)cap + 1));
    AXM_TRY(w->mem.get(w->epa_list, cap));
    AXM_TRY(w->mem.get(w->p_pt, (size_t)cap * 4));
    AXM_TRY(w->mem.get(w->p_nf, (size_t)cap * 4));
    AXM_TRY(w->mem.get(w->pscan_tmp, scan_tmp_words(cap)));
    w->cap_pairs = cap;
    return AXM_SUCCESS;
}

axm_status alloc_rows(axm_world_s* w, int set, uint32_t cap) {
    SolverRows& R = w->R[set];
    w->mem.release(R.hdr);
    w->mem.release(R.r0);
    w->mem.release(R.r1);
    w->mem.release(R.r2);
    w->mem.release(R.r3);
    w->mem.release(R.lam);
    w->mem.release(R.lam0);
    w->mem.release(R.fid);
    w->mem.release(R.dep);
    w->mem.release(w->np_slot[set]);
    w->mem.release(w->pc_key[set]);
    w->mem.release(w->pc_slot[set]);
    AXM_TRY(w->mem.get(R.hdr, cap));
    AXM_TRY(w->mem.get(R.r0, (size_t)cap * 4));
    AXM_TRY(w->mem.get(R.r1, (size_t)cap * 4));
    AXM_TRY(w->mem.get(R.r2, (size_t)cap * 4));
    AXM_TRY(w->mem.get(R.r3, (size_t)cap * 4));
    AXM_TRY(w->mem.get(R.lam, (size_t)cap * 4));
    AXM_TRY(w->mem.get(R.lam0, (size_t)cap * 4));
    AXM_TRY(w->mem.get(R.fid, (size_t)cap * 4));
    AXM_TRY(w->mem.get(R.dep, cap));
    AXM_TRY(w->mem.get(w->np_slot[set], cap));
    const uint32_t T = pow2_at_least(2ull * cap + 16);
    AXM_TRY(w->mem.get(w->pc_key[set], T));
    AXM_TRY(w->mem.get(w->pc_slot[set], T));
    w->pc_mask[set] = T - 1;
    w->rcap[set] = cap;
    return AXM_SUCCESS;
}

axm_status alloc_manifolds(axm_world_s* w, uint32_t cap) {
    w->mem.release(w->m_ab);
    w->mem.release(w->m_np);
    w->mem.release(w->c_pt);
    w->mem.release(w->c_nf);
    w->mem.release(w->col);
    w->mem.release(w->prio);
    w->mem.release(w->ckey);
    w->mem.release(w->ckey_alt);
    w->mem.release(w->cval);
    w->mem.release(w->cval_alt);
    w->mem.release(w->csort_tmp);
    AXM_TRY(w->mem.get(w->m_ab, cap));
    AXM_TRY(w->mem.get(w->m_np, cap));
    AXM_TRY(w->mem.get(w->c_pt, (size_t)cap * 4));
    AXM_TRY(w->mem.get(w->c_nf, (size_t)cap * 4));
    AXM_TRY(w->mem.get(w->col, cap));
    AXM_TRY(w->mem.get(w->prio, cap));
    AXM_TRY(w->mem.get(w->ckey, cap));
    AXM_TRY(w->mem.get(w->ckey_alt, cap));
    AXM_TRY(w->mem.get(w->cval, cap));
    AXM_TRY(w->mem.get(w->cval_alt, cap));
    AXM_TRY(w->mem.get(w->csort_tmp, radix_tmp_words(cap)));
    if (w->rcap[w->cur] < cap) AXM_TRY(alloc_rows(w, w->cur, cap));
    w->cap_man = cap;
    return AXM_SUCCESS;
}

StepParams make_params(const axm_world_s* w, float dt) {
    const axm_world_desc& d = w->desc;
    StepParams sp{};
    sp.gx = d.gravity[0];
    sp.gy = d.gravity[1];
    sp.gz = d.gravity[2];
    sp.dt = dt;
    sp.inv_dt = 1.0f / dt;
    sp.friction = d.friction;
    sp.restitution = d.restitution;
    sp.rest_threshold = d.rest_threshold;
    sp.baumgarte = d.baumgarte;
    sp.slop = d.slop;
    sp.max_bias = d.max_bias;
    sp.margin = d.margin;
    sp.lin_keep = 1.0f - d.lin_damp;
    sp.ang_keep = 1.0f - d.ang_damp;
    sp.inv_cell = 1.0f / w->cell;
    sp.ox = d.origin[0];
    sp.oy = d.origin[1];
    sp.oz = d.origin[2];
    sp.num_planes = d.num_planes;
    for (uint32_t k = 0; k < d.num_planes && k < (uint32_t)MAX_PLANES; ++k)
        sp.planes[k] = make_float4(d.planes[k][0], d.planes[k][1], d.planes[k][2], d.planes[k][3]);
    return sp;
}

axm_status read_counters(axm_world_s* w) {
    AXM_HIP(hipMemcpyAsync(w->h_cnt, w->cnt, sizeof(Counters), hipMemcpyDeviceToHost, w->stream));
    AXM_HIP(hipStreamSynchronize(w->stream));
    return AXM_SUCCESS;
}

struct StageTimer {
    bool on = false;
    hipStream_t s = nullptr;
    hipEvent_t ev[6] = {};
    int n = 0;
    void init(bool enable, hipStream_t st) {
        on = enable;
        s = st;
        if (!on) return;
        for (auto& e : ev) hipEventCreate(&e);
    }
    void mark() {
        if (on && n < 6) hipEventRecord(ev[n++], s);
    }
    float ms(int a, int b) {
        float r = 0.0f;
        if (on && b < n) hipEventElapsedTime(&r, ev[a], ev[b]);
        return r;
    }
    ~StageTimer() {
        if (on)
            for (auto& e : ev) hipEventDestroy(e);
    }
};

// ---- phase 1: velocity integration into the solver copies, AABBs, keys,
// sort, cell table, candidate pairs.  The canonical body state is not touched.
axm_status phase_broad(axm_world_s* w, const StepParams& sp, uint32_t& P, axm_step_stats& S) {
    const uint32_t n = w->n;
    hipStream_t s = w->stream;
    launch_begin(w->B, sp, n, w->key, w->val, s);
    const int end_bit = 30 + bits_for(w->n_worlds > 1 ? w->n_worlds - 1 : 0);
    const int alt = radix_sort_u64(w->key, w->val, w->key_alt, w->val_alt, n, end_bit, w->sort_tmp, s);
    uint64_t* skey = alt ? w->key_alt : w->key;
    uint32_t* sval = alt ? w->val_alt : w->val;
    launch_fill_u64(w->ct.key, EMPTY_KEY, w->ct.mask + 1, s);
    launch_cell_table(skey, n, w->ct, s);
    launch_pair_count(w->B, sp, skey, sval, w->ct, n, w->pair_cnt, s);
    scan_exclusive(w->pair_cnt, w->pair_off, n, w->scan_tmp, s);
    AXM_HIP(hipMemcpyAsync(&w->h_cnt->pairs, w->pair_off + n, 4, hipMemcpyDeviceToHost, s));
    AXM_HIP(hipStreamSynchronize(s));
    P = w->h_cnt->pairs;
    if (P > w->cap_pairs) {
        if (!w->desc.auto_grow) {
            S.pairs = P;
            S.overflow = 1;
            return AXM_OUT_OF_RANGE;
        }
        AXM_TRY(alloc_pairs(w, (uint32_t)std::min<uint64_t>(0xFFFFFFF0ull, (uint64_t)P + P / 2 + 1024)));
    }
    launch_pair_emit(w->B, sp, skey, sval, w->ct, n, w->pair_off, w->cap_pairs, w->pa, w->pb, s);
    return AXM_SUCCESS;
}

// ---- phase 2: narrowphase into per-pair staging, compaction into manifolds.
axm_status phase_narrow(axm_world_s* w, const StepParams& sp, uint32_t P, uint32_t& M, axm_step_stats& S) {
    hipStream_t s = w->stream;
    M = 0;
    if (!P) return AXM_SUCCESS;
    launch_narrow(w->B, sp, w->pa, w->pb, P, w->p_np, w->p_pt, w->p_nf, w->epa_list, P, w->cnt, s);
    launch_flag_nonzero(w->p_np, P, w->p_flag, s);
    scan_exclusive(w->p_flag, w->man_off, P, w->pscan_tmp, s);
    AXM_HIP(hipMemcpyAsync(&w->h_cnt->manifolds, w->man_off + P, 4, hipMemcpyDeviceToHost, s));
    AXM_HIP(hipStreamSynchronize(s));
    M = w->h_cnt->manifolds;
    if (M > w->cap_man || w->rcap[w->cur] < M) {
        if (!w->desc.auto_grow && M > w->cap_man) {
            S.pairs = P;
            S.manifolds = M;
            S.overflow = 2;
            return AXM_OUT_OF_RANGE;
        }
        AXM_TRY(alloc_manifolds(w, std::max(w->cap_man, M + M / 2 + 1024)));
    }
    launch_compact(w->pa, w->pb, P, w->p_np, w->p_pt, w->p_nf, w->man_off, w->cap_man, w->m_ab, w->m_np, w->c_pt,
                   w->c_nf, w->cnt, s);
    return AXM_SUCCESS;
}

// ---- phase 3: Jones-Plassmann colouring, stable order by colour, colour offsets.
axm_status phase_colour(axm_world_s* w, uint32_t M, uint32_t& K, uint32_t& rounds) {
    hipStream_t s = w->stream;
    K = rounds = 0;
    if (!M) {
        w->h_col_off.assign(1, 0);
        return AXM_SUCCESS;
    }
    launch_col_init(w->B, w->m_ab, M, w->col, w->prio, s);
    AXM_HIP(hipMemsetAsync(w->bmax, 0, sizeof(uint64_t) * w->n, s));
    AXM_HIP(hipMemsetAsync(w->bmask, 0, sizeof(uint64_t) * w->n, s));
    w->h_cnt->remaining = M;
    AXM_HIP(hipMemcpyAsync(&w->cnt->remaining, &w->h_cnt->remaining, 4, hipMemcpyHostToDevice, s));
    uint32_t r = 0;
    const uint32_t batch = 4;
    for (;;) {
        for (uint32_t k = 0; k < batch; ++k, ++r)
            launch_col_round(w->B, w->m_ab, M, w->col, w->prio, w->bmax, w->bmask, r, w->cnt, s);
        AXM_TRY(read_counters(w));
        if (w->h_cnt->remaining == 0) break;
        if (r >= 65000u) return AXM_CONSTRAINT_SOLVER_FAILURE;
    }
    K = w->h_cnt->max_colour;
    rounds = w->h_cnt->last_round;
    AXM_HIP(hipMemcpyAsync(w->ckey, w->col, sizeof(uint32_t) * M, hipMemcpyDeviceToDevice, s));
    launch_iota(w->cval, M, s);
    const int calt =
        radix_sort_u32(w->ckey, w->cval, w->ckey_alt, w->cval_alt, M, std::max(8, bits_for(K)), w->csort_tmp, s);
    uint32_t* scol = calt ? w->ckey_alt : w->ckey;
    w->order = calt ? w->cval_alt : w->cval;
    if (w->col_off_cap < K + 1) {
        w->mem.release(w->col_off);
        AXM_TRY(w->mem.get(w->col_off, K + 1 + 256));
        w->col_off_cap = K + 1 + 256;
    }
    launch_col_offsets(scol, M, K, w->col_off, s);
    w->h_col_off.resize(K + 1);
    AXM_HIP(hipMemcpyAsync(w->h_col_off.data(), w->col_off, sizeof(uint32_t) * (K + 1), hipMemcpyDeviceToHost, s));
    AXM_HIP(hipStreamSynchronize(s));
    return AXM_SUCCESS;
}

// One full step.  warm_in (optional, device, manifold order x 4 points) overrides the warm start.
axm_status step_impl(axm_world_s* w, float dt, uint32_t iters, uint32_t flags, axm_step_stats* st,
                     const float4* warm_in) {
    if (!(dt > 0.0f) || !std::isfinite(dt)) return AXM_INVALID_PARAMETER;
    axm_step_stats S{};
    const uint32_t n = w->n;
    hipStream_t s = w->stream;
    const StepParams sp = make_params(w, dt);
    StageTimer tm;
    tm.init((flags & AXM_STEP_TIMINGS) != 0, s);
    tm.mark();
    AXM_HIP(hipMemsetAsync(w->cnt, 0, sizeof(Counters), s));
    if (n == 0) {
        w->have_step = true;
        w->P = w->M = w->K = w->rounds = 0;
        w->h_res.clear();
        if (st) *st = S;
        return AXM_SUCCESS;
    }
    uint32_t P = 0, M = 0, K = 0, rounds = 0;
    axm_status e = phase_broad(w, sp, P, S);
    if (e != AXM_SUCCESS) {
        if (st) *st = S;
        return e;
    }
    if (flags & AXM_STEP_KEEP_STAGES) {
        AXM_HIP(hipMemcpyAsync(w->pre_v, w->B.sv, 16ull * n, hipMemcpyDeviceToDevice, s));
        AXM_HIP(hipMemcpyAsync(w->pre_w, w->B.sw, 16ull * n, hipMemcpyDeviceToDevice, s));
        w->have_pre = true;
    } else {
        w->have_pre = false;
    }
    tm.mark();
    e = phase_narrow(w, sp, P, M, S);
    if (e != AXM_SUCCESS) {
        if (st) *st = S;
        return e;
    }
    tm.mark();
    AXM_TRY(phase_colour(w, M, K, rounds));
    // ---- solver
    const int c = w->cur, pv = 1 - c;
    const SolverRows& R = w->R[c];
    const uint32_t rc = w->rcap[c];
    PrevCacheArgs pca{};
    pca.valid = (w->prev_valid && w->desc.warm_start != AXM_WARM_NONE) ? 1u : 0u;
    pca.key = w->pc_key[pv];
    pca.slot = w->pc_slot[pv];
    pca.mask = w->pc_mask[pv];
    pca.lam = w->R[pv].lam;
    pca.fid = w->R[pv].fid;
    pca.np = w->np_slot[pv];
    pca.cap = w->rcap[pv];
    if (!pca.key) pca.valid = 0;
    launch_prep(w->B, sp, w->order, M, rc, w->m_ab, w->m_np, w->c_pt, w->c_nf, R, pca, warm_in, w->col, w->bmask,
                w->cnt, s);
    const bool tol_mode = w->desc.residual_tol > 0.0f;
    const uint32_t max_it = tol_mode ? std::max<uint32_t>(w->desc.max_iterations, 1u) : iters;
    if (w->res_cap < max_it + 1) {
        w->mem.release(w->res_bits);
        AXM_TRY(w->mem.get(w->res_bits, max_it + 1));
        w->res_cap = max_it + 1;
    }
    AXM_HIP(hipMemsetAsync(w->res_bits, 0, sizeof(uint32_t) * (max_it + 1), s));
    const bool want_res = (flags & AXM_STEP_RESIDUAL) || w->desc.record_residual || tol_mode;
    const uint32_t every = std::max<uint32_t>(w->desc.residual_every, 1u);
    uint32_t it = 0;
    size_t nev = 0;
    const bool flow = w->desc.solver_mode == 0 && K <= 64 && M > 0;
    auto prof_mark = [&]() -> axm_status {
        if (!w->prof_on) return AXM_SUCCESS;
        while (w->prof_ev.size() < nev + 1) {
            hipEvent_t e;
            AXM_HIP(hipEventCreate(&e));
            w->prof_ev.push_back(e);
        }
        AXM_HIP(hipEventRecord(w->prof_ev[nev++], s));
        return AXM_SUCCESS;
    };
    if (flow) {
        AXM_HIP(hipMemsetAsync(w->flow_ctl, 0, 16, s));
        while (it < max_it) {
            uint32_t seg = max_it - it;
            if (w->desc.record_residual) seg = 1;
            else if (tol_mode) seg = std::min(seg, every);
            AXM_HIP(hipMemsetAsync(w->bdone, 0, sizeof(uint32_t) * n, s));
            AXM_HIP(hipMemsetAsync(w->flow_ctl, 0, 4, s));
            AXM_TRY(prof_mark());
            launch_pgs_flow(w->B, sp, R, M, rc, it, seg, w->bdone, w->flow_ctl, w->flow_ctl + 1, 1u << 20, 4096u, s);
            AXM_TRY(prof_mark());
            it += seg;
            const bool last = it == max_it;
            const bool check = tol_mode && (it % every == 0);
            if (w->desc.record_residual || check || (last && want_res)) {
                launch_residual(w->B, sp, R, M, rc, w->res_bits + it - 1, s);
                if (check) {
                    AXM_HIP(hipMemcpyAsync(&w->h_cnt->residual_bits, w->res_bits + it - 1, 4, hipMemcpyDeviceToHost, s));
                    AXM_HIP(hipStreamSynchronize(s));
                    float res;
                    std::memcpy(&res, &w->h_cnt->residual_bits, 4);
                    if (res <= w->desc.residual_tol) break;
                }
            }
        }
    }
    for (; !flow && it < max_it; ++it) {
        for (uint32_t col = 0; col < K; ++col) {
            const uint32_t s0 = w->h_col_off[col], s1 = w->h_col_off[col + 1];
            if (s1 <= s0) continue;
            AXM_TRY(prof_mark());
            launch_pgs(w->B, sp, R, s0, s1, rc, it == 0 ? 1u : 0u, s);
            AXM_TRY(prof_mark());
        }
        const bool last = it + 1 == max_it;
        const bool check = tol_mode && ((it + 1) % every == 0);
        if (w->desc.record_residual || check || (last && want_res)) {
            launch_residual(w->B, sp, R, M, rc, w->res_bits + it, s);
            if (check) {
                AXM_HIP(hipMemcpyAsync(&w->h_cnt->residual_bits, w->res_bits + it, 4, hipMemcpyDeviceToHost, s));
                AXM_HIP(hipStreamSynchronize(s));
                float res;
                std::memcpy(&res, &w->h_cnt->residual_bits, 4);
                if (res <= w->desc.residual_tol) {
                    ++it;
                    break;
                }
            }
        }
    }
    const uint32_t iters_run = it;
    if (w->prof_on) {
        AXM_HIP(hipStreamSynchronize(s));
        double ms = 0.0;
        for (size_t k = 0; k < nev; k += 2) {
            float e = 0.0f;
            AXM_HIP(hipEventElapsedTime(&e, w->prof_ev[k], w->prof_ev[k + 1]));
            ms += e;
        }
        AXM_TRY(read_counters(w));
        w->prof.steps += 1;
        w->prof.pgs_launches += nev / 2;
        w->prof.pgs_ms += ms;
        w->prof.pgs_bytes += (double)w->h_cnt->pgs_bytes * (double)iters_run;
        w->prof.sweeps += iters_run;
    }
    tm.mark();
    // warm-start table for the next step
    launch_fill_u64(w->pc_key[c], EMPTY_KEY, w->pc_mask[c] + 1, s);
    launch_cache_build(R, M, w->pc_key[c], w->pc_slot[c], w->pc_mask[c], w->np_slot[c], s);
    launch_integrate(w->B, sp, n, s);
    tm.mark();
    // ---- bookkeeping
    w->P = P;
    w->M = M;
    w->K = K;
    w->rounds = rounds;
    w->iters_run = iters_run;
    w->have_step = true;
    w->prev_valid = true;
    w->cur = pv;
    if (w->rcap[w->cur] < w->cap_man) AXM_TRY(alloc_rows(w, w->cur, w->cap_man));
    const bool need_sync = st || (flags & AXM_STEP_SYNC) || want_res || tm.on;
    w->h_res.clear();
    if (flow) {
        uint32_t ctl[2] = {0, 0};
        AXM_HIP(hipMemcpyAsync(ctl, w->flow_ctl, 8, hipMemcpyDeviceToHost, s));
        AXM_HIP(hipStreamSynchronize(s));
        if (ctl[1]) return AXM_GPU_TIMEOUT;
    }
    if (need_sync) {
        AXM_TRY(read_counters(w));
        if (want_res && iters_run) {
            std::vector<uint32_t> bits(iters_run);
            AXM_HIP(hipMemcpy(bits.data(), w->res_bits, 4ull * iters_run, hipMemcpyDeviceToHost));
            w->h_res.assign(iters_run, 0.0f);
            for (uint32_t k = 0; k < iters_run; ++k) std::memcpy(&w->h_res[k], &bits[k], 4);
        }
        S.pairs = P;
        S.manifolds = M;
        S.contacts = w->h_cnt->contacts;
        S.colours = K;
        S.colour_rounds = rounds;
        S.iterations = iters_run;
        S.gjk_fail = w->h_cnt->gjk_fail;
        S.epa_fail = w->h_cnt->epa_fail;
        S.warm_hits = w->h_cnt->warm_hits;
        if (!w->h_res.empty()) {
            S.residual = w->h_res.back();
            S.residual_first = w->h_res.front();
        }
        if (tm.on) {
            S.broadphase_ms = tm.ms(0, 1);
            S.narrowphase_ms = tm.ms(1, 2);
            S.solver_ms = tm.ms(2, 3);
            S.integration_ms = tm.ms(3, 4);
            S.total_ms = tm.ms(0, 4);
        }
        if (st) *st = S;
    }
    return AXM_SUCCESS;
}

bool finite3(const float* p) { return std::isfinite(p[0]) && std::isfinite(p[1]) && std::isfinite(p[2]); }

}  // namespace

extern "C" {

int32_t axm_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

axm_status axm_world_create(const axm_world_desc* desc, axm_world* out) {
    if (!desc || !out) return AXM_NULL_POINTER;
    *out = nullptr;
    if (desc->num_planes > AXM_MAX_PLANES) return AXM_INVALID_PARAMETER;
    if (desc->cell_size < 0.0f || !std::isfinite(desc->cell_size)) return AXM_INVALID_PARAMETER;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return AXM_GPU_OPERATION_FAILED;
    if (desc->device < 0 || desc->device >= ndev) return AXM_INVALID_PARAMETER;
    if (hipSetDevice(desc->device) != hipSuccess) return AXM_GPU_OPERATION_FAILED;
    axm_world_s* w = new (std::nothrow) axm_world_s;
    if (!w) return AXM_OUT_OF_MEMORY;
    w->desc = *desc;
    if (hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking) != hipSuccess) {
        delete w;
        return AXM_GPU_OPERATION_FAILED;
    }
    if (hipHostMalloc((void**)&w->h_cnt, sizeof(Counters), hipHostMallocDefault) != hipSuccess) {
        delete w;
        return AXM_BUFFER_ALLOCATION_FAILED;
    }
    axm_status st = w->mem.get(w->cnt, 1);
    if (st == AXM_SUCCESS) st = w->mem.get(w->flow_ctl, 4);
    if (st == AXM_SUCCESS) st = alloc_bodies(w, std::max<uint32_t>(desc->max_bodies, 16u));
    if (st == AXM_SUCCESS) st = alloc_pairs(w, std::max<uint32_t>(desc->max_pairs, 1024u));
    if (st == AXM_SUCCESS) st = alloc_manifolds(w, std::max<uint32_t>(desc->max_manifolds, 1024u));
    if (st == AXM_SUCCESS) st = alloc_rows(w, 1 - w->cur, w->cap_man);
    if (st != AXM_SUCCESS) {
        delete w;
        return st;
    }
    *out = w;
    return AXM_SUCCESS;
}

axm_status axm_world_destroy(axm_world w) {
    if (!w) return AXM_NULL_POINTER;
    hipSetDevice(w->desc.device);
    delete w;
    return AXM_SUCCESS;
}

axm_status axm_world_body_count(axm_world w, uint32_t* n) {
    if (!w || !n) return AXM_NULL_POINTER;
    *n = w->n;
    return AXM_SUCCESS;
}

axm_status axm_world_set_bodies(axm_world w, const axm_body_soa* b, uint32_t n) {
    if (!w || !b) return AXM_NULL_POINTER;
    if (n && (!b->pos || !b->quat || !b->lin_vel || !b->ang_vel || !b->inv_mass || !b->inv_inertia || !b->shape_type ||
              !b->shape_params))
        return AXM_NULL_POINTER;
    if (n >= (1u << 29)) return AXM_OUT_OF_RANGE;
    hipSetDevice(w->desc.device);
    std::vector<float4> x(n), q(n), v(n), av(n), ii(n), sh(n);
    std::vector<uint32_t> wd(n);
    std::vector<uint32_t> first;
    float rmax = 0.0f;
    uint32_t nw = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const float* p = b->pos + 3 * i;
        const float* qq = b->quat + 4 * i;
        const float* lv = b->lin_vel + 3 * i;
        const float* aw = b->ang_vel + 3 * i;
        const float* iv = b->inv_inertia + 3 * i;
        const float* pr = b->shape_params + 3 * i;
        const float im = b->inv_mass[i];
        const uint32_t t = b->shape_type[i];
        if (!finite3(p) || !finite3(lv) || !finite3(aw) || !finite3(iv) || !std::isfinite(im) || im < 0.0f ||
            !std::isfinite(qq[0]) || !std::isfinite(qq[1]) || !std::isfinite(qq[2]) || !std::isfinite(qq[3]) ||
            iv[0] < 0.0f || iv[1] < 0.0f || iv[2] < 0.0f)
            return AXM_INVALID_RIGID_BODY;
        const float ql = qq[0] * qq[0] + qq[1] * qq[1] + qq[2] * qq[2] + qq[3] * qq[3];
        if (!(ql > 0.5f && ql < 1.5f)) return AXM_INVALID_RIGID_BODY;
        if (t > AXM_SHAPE_CAPSULE || !finite3(pr)) return AXM_INVALID_SHAPE;
        float br;
        if (t == AXM_SHAPE_SPHERE) {
            if (!(pr[0] > 0.0f)) return AXM_INVALID_SHAPE;
            br = pr[0];
        } else if (t == AXM_SHAPE_BOX) {
            if (!(pr[0] > 0.0f && pr[1] > 0.0f && pr[2] > 0.0f)) return AXM_INVALID_SHAPE;
            br = std::sqrt(pr[0] * pr[0] + pr[1] * pr[1] + pr[2] * pr[2]);
        } else {
            if (!(pr[0] > 0.0f && pr[1] >= 0.0f)) return AXM_INVALID_SHAPE;
            br = pr[0] + pr[1];
        }
        rmax = std::max(rmax, br);
        const uint32_t wi = b->world ? b->world[i] : 0u;
        if (i > 0 && wi < wd[i - 1]) return AXM_INVALID_PARAMETER;  // worlds must be contiguous
        wd[i] = wi;
        nw = std::max(nw, wi + 1);
        const bool stat = im == 0.0f;
        x[i] = make_float4(p[0], p[1], p[2], 0.0f);
        q[i] = make_float4(qq[0], qq[1], qq[2], qq[3]);
        v[i] = make_float4(lv[0], lv[1], lv[2], im);
        av[i] = make_float4(aw[0], aw[1], aw[2], 0.0f);
        ii[i] = stat ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : make_float4(iv[0], iv[1], iv[2], 0.0f);
        uint32_t tb = t;
        float tf;
        std::memcpy(&tf, &tb, 4);
        sh[i] = make_float4(pr[0], pr[1], pr[2], tf);
    }
    if (nw > 4096u * 1024u) return AXM_OUT_OF_RANGE;
    first.assign(nw + 1, 0);
    for (uint32_t i = n; i-- > 0;) first[wd[i]] = i;
    const float need = 2.0f * (rmax + w->desc.margin);
    float cell = w->desc.cell_size;
    if (cell == 0.0f) cell = need > 0.0f ? need : 1.0f;
    if (cell < need) return AXM_INVALID_PARAMETER;
    if (n > w->cap_n) {
        free_bodies(w);
        AXM_TRY(alloc_bodies(w, n));
    }
    if (w->desc.max_pairs == 0 && w->cap_pairs < 8ull * n) AXM_TRY(alloc_pairs(w, 8 * n));
    if (w->desc.max_manifolds == 0 && w->cap_man < 4ull * n) AXM_TRY(alloc_manifolds(w, 4 * n));
    hipStream_t s = w->stream;
    if (n) {
        AXM_HIP(hipMemcpyAsync(w->B.x, x.data(), 16ull * n, hipMemcpyHostToDevice, s));
        AXM_HIP(hipMemcpyAsync(w->B.q, q.data(), 16ull * n, hipMemcpyHostToDevice, s));
        AXM_HIP(hipMemcpyAsync(w->B.vel, v.data(), 16ull * n, hipMemcpyHostToDevice, s));
        AXM_HIP(hipMemcpyAsync(w->B.angv, av.data(), 16ull * n, hipMemcpyHostToDevice, s));
        AXM_HIP(hipMemcpyAsync(w->B.ii, ii.data(), 16ull * n, hipMemcpyHostToDevice, s));
        AXM_HIP(hipMemcpyAsync(w->B.shape, sh.data(), 16ull * n, hipMemcpyHostToDevice, s));
        AXM_HIP(hipMemcpyAsync(w->B.world, wd.data(), 4ull * n, hipMemcpyHostToDevice, s));
        AXM_HIP(hipMemcpyAsync(w->B.world_first, first.data(), 4ull * (nw + 1), hipMemcpyHostToDevice, s));
    }
    AXM_HIP(hipStreamSynchronize(s));
    w->n = n;
    w->n_worlds = std::max(nw, 1u);
    w->cell = cell;
    w->prev_valid = false;
    w->have_step = false;
    return AXM_SUCCESS;
}

axm_status axm_world_get_bodies(axm_world w, axm_body_soa* out) {
    if (!w || !out) return AXM_NULL_POINTER;
    hipSetDevice(w->desc.device);
    const uint32_t n = w->n;
    std::vector<float4> buf(n);
    auto pull = [&](float4* src, float* dst, int comps) -> axm_status {
        if (!dst || !n) return AXM_SUCCESS;
        AXM_HIP(hipMemcpyAsync(buf.data(), src, 16ull * n, hipMemcpyDeviceToHost, w->stream));
        AXM_HIP(hipStreamSynchronize(w->stream));
        for (uint32_t i = 0; i < n; ++i) {
            dst[comps * i] = buf[i].x;
            dst[comps * i + 1] = buf[i].y;
            dst[comps * i + 2] = buf[i].z;
            if (comps == 4) dst[comps * i + 3] = buf[i].w;
        }
        return AXM_SUCCESS;
    };
    AXM_TRY(pull(w->B.x, out->pos, 3));
    AXM_TRY(pull(w->B.q, out->quat, 4));
    AXM_TRY(pull(w->B.vel, out->lin_vel, 3));
    AXM_TRY(pull(w->B.angv, out->ang_vel, 3));
    return AXM_SUCCESS;
}

axm_status axm_world_get_pre_solve_velocities(axm_world w, float* lin_vel, float* ang_vel) {
    if (!w) return AXM_NULL_POINTER;
    if (!w->have_pre) return AXM_GPU_INVALID_OPERATION;
    hipSetDevice(w->desc.device);
    const uint32_t n = w->n;
    std::vector<float4> v(n), a(n);
    AXM_HIP(hipMemcpy(v.data(), w->pre_v, 16ull * n, hipMemcpyDeviceToHost));
    AXM_HIP(hipMemcpy(a.data(), w->pre_w, 16ull * n, hipMemcpyDeviceToHost));
    for (uint32_t i = 0; i < n; ++i) {
        if (lin_vel) { lin_vel[3 * i] = v[i].x; lin_vel[3 * i + 1] = v[i].y; lin_vel[3 * i + 2] = v[i].z; }
        if (ang_vel) { ang_vel[3 * i] = a[i].x; ang_vel[3 * i + 1] = a[i].y; ang_vel[3 * i + 2] = a[i].z; }
    }
    return AXM_SUCCESS;
}

axm_status axm_broadphase(axm_world w, uint32_t* count) {
    if (!w) return AXM_NULL_POINTER;
    hipSetDevice(w->desc.device);
    axm_step_stats S{};
    const StepParams sp = make_params(w, 1.0f / 120.0f);
    AXM_HIP(hipMemsetAsync(w->cnt, 0, sizeof(Counters), w->stream));
    uint32_t P = 0;
    if (w->n) AXM_TRY(phase_broad(w, sp, P, S));
    AXM_HIP(hipStreamSynchronize(w->stream));
    w->P = P;
    w->M = 0;
    w->K = 0;
    w->have_step = true;
    if (count) *count = P;
    return AXM_SUCCESS;
}

axm_status axm_collide(axm_world w, uint32_t* manifold_count) {
    if (!w) return AXM_NULL_POINTER;
    hipSetDevice(w->desc.device);
    axm_step_stats S{};
    const StepParams sp = make_params(w, 1.0f / 120.0f);
    AXM_HIP(hipMemsetAsync(w->cnt, 0, sizeof(Counters), w->stream));
    uint32_t P = 0, M = 0;
    if (w->n) {
        AXM_TRY(phase_broad(w, sp, P, S));
        AXM_TRY(phase_narrow(w, sp, P, M, S));
    }
    AXM_HIP(hipStreamSynchronize(w->stream));
    w->P = P;
    w->M = M;
    w->K = 0;
    w->have_step = true;
    if (manifold_count) *manifold_count = M;
    return AXM_SUCCESS;
}

axm_status axm_world_step(axm_world w, float dt, uint32_t iters, uint32_t flags, axm_step_stats* stats) {
    if (!w) return AXM_NULL_POINTER;
    hipSetDevice(w->desc.device);
    return step_impl(w, dt, iters, flags, stats, nullptr);
}

axm_status axm_world_run(axm_world w, float dt, uint32_t iters, uint32_t steps, int graph, axm_step_stats* last) {
    if (!w) return AXM_NULL_POINTER;
    hipSetDevice(w->desc.device);
    (void)graph;
    for (uint32_t k = 0; k < steps; ++k)
        AXM_TRY(step_impl(w, dt, iters, k + 1 == steps ? AXM_STEP_SYNC : 0u, k + 1 == steps ? last : nullptr, nullptr));
    return AXM_SUCCESS;
}

axm_status axm_world_profile(axm_world w, int enable, axm_kernel_profile* out) {
    if (!w) return AXM_NULL_POINTER;
    if (out) *out = w->prof;
    w->prof = axm_kernel_profile{};
    w->prof_on = enable != 0;
    return AXM_SUCCESS;
}

axm_status axm_world_sync(axm_world w) {
    if (!w) return AXM_NULL_POINTER;
    hipSetDevice(w->desc.device);
    AXM_HIP(hipStreamSynchronize(w->stream));
    return AXM_SUCCESS;
}

axm_status axm_world_get_pairs(axm_world w, uint32_t* a, uint32_t* b, uint32_t cap, uint32_t* count) {
    if (!w || !count) return AXM_NULL_POINTER;
    if (!w->have_step) return AXM_GPU_INVALID_OPERATION;
    hipSetDevice(w->desc.device);
    *count = w->P;
    if (cap < w->P) return AXM_OUT_OF_RANGE;
    if (w->P) {
        if (a) AXM_HIP(hipMemcpy(a, w->pa, 4ull * w->P, hipMemcpyDeviceToHost));
        if (b) AXM_HIP(hipMemcpy(b, w->pb, 4ull * w->P, hipMemcpyDeviceToHost));
    }
    return AXM_SUCCESS;
}

axm_status axm_world_get_manifolds(axm_world w, uint32_t* a, uint32_t* b, uint32_t* npts, float* pos, float* normal,
                                   float* depth, uint32_t* feature, uint32_t cap, uint32_t* count) {
    if (!w || !count) return AXM_NULL_POINTER;
    if (!w->have_step) return AXM_GPU_INVALID_OPERATION;
    hipSetDevice(w->desc.device);
    const uint32_t M = w->M;
    *count = M;
    if (cap < M) return AXM_OUT_OF_RANGE;
    if (!M) return AXM_SUCCESS;
    std::vector<uint2> ab(M);
    std::vector<uint32_t> np(M);
    std::vector<float4> pt((size_t)M * 4), nf((size_t)M * 4);
    AXM_HIP(hipMemcpy(ab.data(), w->m_ab, 8ull * M, hipMemcpyDeviceToHost));
    AXM_HIP(hipMemcpy(np.data(), w->m_np, 4ull * M, hipMemcpyDeviceToHost));
    AXM_HIP(hipMemcpy(pt.data(), w->c_pt, 64ull * M, hipMemcpyDeviceToHost));
    AXM_HIP(hipMemcpy(nf.data(), w->c_nf, 64ull * M, hipMemcpyDeviceToHost));
    for (uint32_t m = 0; m < M; ++m) {
        if (a) a[m] = ab[m].x;
        if (b) b[m] = ab[m].y;
        if (npts) npts[m] = np[m];
        for (uint32_t k = 0; k < 4; ++k) {
            const size_t s = (size_t)m * 4 + k;
            const bool on = k < np[m];
            if (pos) { pos[3 * s] = on ? pt[s].x : 0.f; pos[3 * s + 1] = on ? pt[s].y : 0.f; pos[3 * s + 2] = on ? pt[s].z : 0.f; }
            if (normal) { normal[3 * s] = on ? nf[s].x : 0.f; normal[3 * s + 1] = on ? nf[s].y : 0.f; normal[3 * s + 2] = on ? nf[s].z : 0.f; }
            if (depth) depth[s] = on ? pt[s].w : 0.f;
            if (feature) {
                uint32_t f;
                std::memcpy(&f, &nf[s].w, 4);
                feature[s] = on ? f : 0u;
            }
        }
    }
    return AXM_SUCCESS;
}

axm_status axm_world_get_colours(axm_world w, uint32_t* colour, uint32_t* solve_order, uint32_t cap, uint32_t* count,
                                 uint32_t* num_colours) {
    if (!w || !count) return AXM_NULL_POINTER;
    if (!w->have_step) return AXM_GPU_INVALID_OPERATION;
    hipSetDevice(w->desc.device);
    *count = w->M;
    if (num_colours) *num_colours = w->K;
    if (cap < w->M) return AXM_OUT_OF_RANGE;
    if (w->M) {
        if (colour) AXM_HIP(hipMemcpy(colour, w->col, 4ull * w->M, hipMemcpyDeviceToHost));
        if (solve_order) AXM_HIP(hipMemcpy(solve_order, w->order, 4ull * w->M, hipMemcpyDeviceToHost));
    }
    return AXM_SUCCESS;
}

axm_status axm_world_get_impulses(axm_world w, float* initial, float* final_, uint32_t cap, uint32_t* count) {
    if (!w || !count) return AXM_NULL_POINTER;
    if (!w->have_step) return AXM_GPU_INVALID_OPERATION;
    hipSetDevice(w->desc.device);
    const uint32_t M = w->M;
    *count = M;
    if (cap < M) return AXM_OUT_OF_RANGE;
    if (!M) return AXM_SUCCESS;
    const int set = 1 - w->cur;  // rows of the last step
    const SolverRows& R = w->R[set];
    const uint32_t rc = w->rcap[set];
    std::vector<uint4> hdr(M);
    AXM_HIP(hipMemcpy(hdr.data(), R.hdr, 16ull * M, hipMemcpyDeviceToHost));
    std::vector<float4> l(M), l0(M);
    for (uint32_t k = 0; k < 4; ++k) {
        AXM_HIP(hipMemcpy(l.data(), R.lam + (size_t)k * rc, 16ull * M, hipMemcpyDeviceToHost));
        AXM_HIP(hipMemcpy(l0.data(), R.lam0 + (size_t)k * rc, 16ull * M, hipMemcpyDeviceToHost));
        for (uint32_t s = 0; s < M; ++s) {
            const uint32_t m = hdr[s].w;
            const bool on = k < hdr[s].z;
            const size_t o = ((size_t)m * 4 + k) * 3;
            if (final_) { final_[o] = on ? l[s].x : 0.f; final_[o + 1] = on ? l[s].y : 0.f; final_[o + 2] = on ? l[s].z : 0.f; }
            if (initial) { initial[o] = on ? l0[s].x : 0.f; initial[o + 1] = on ? l0[s].y : 0.f; initial[o + 2] = on ? l0[s].z : 0.f; }
        }
    }
    return AXM_SUCCESS;
}

axm_status axm_world_get_residuals(axm_world w, float* out, uint32_t cap, uint32_t* count) {
    if (!w || !count) return AXM_NULL_POINTER;
    *count = (uint32_t)w->h_res.size();
    if (cap < *count) return AXM_OUT_OF_RANGE;
    if (out)
        for (size_t k = 0; k < w->h_res.size(); ++k) out[k] = w->h_res[k];
    return AXM_SUCCESS;
}

}  // extern "C"
