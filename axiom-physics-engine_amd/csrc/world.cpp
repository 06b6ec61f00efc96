// Host side of the C ABI (include/axiom/axm.h): world lifecycle, device buffer
// management and the orchestration of one step on the world's HIP stream.
//
// A step is enqueued without host synchronisation: every count a kernel needs
// lives in the device control block (Counters), grids are sized by capacity,
// and the stage kernels read the live counts.  That makes the step capturable
// into a HIP graph (axm_world_run with graph = 1) and leaves exactly one
// synchronisation per axm_world_step (to read the statistics and the sticky
// error word).  A step that overflows a capacity, needs more colours than the
// persistent sweep supports, or times out in a bounded wait sets a sticky bit;
// every kernel that writes persistent state then skips, so the body state and
// the warm-start cache stay exactly as before the step and the host can grow
// the buffers (or switch the sweep to one launch per colour) and run it again.
//
// Error handling follows the reference's no-exception Result/ErrorCode
// convention (include/axiom/core/result.hpp:15-283,
// include/axiom/core/error_code.hpp:21-57).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstddef>
#include <cstring>
#include <new>
#include <vector>

#include "axiom/axm.h"
#include "file_io.hpp"
#include "k_step.hpp"
#include "mlp_impl.hpp"

using namespace axm;

namespace {

#define AXM_HIP(expr)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess) {                                                              \
            std::fprintf(stderr, "axm: %s failed: %s\n", #expr, hipGetErrorString(e_));    \
            return AXM_GPU_OPERATION_FAILED;                                                 \
        }                                                                                    \
    } while (0)

#define AXM_TRY(expr)                     \
    do {                                  \
        axm_status s_ = (expr);           \
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

constexpr uint32_t MAX_SORT_COLOURS = 255;  // the colour sort is one 8-bit radix pass

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
        (void)hipFree(p);
        p = nullptr;
    }
    void release_all() {
        for (void* p : ptrs) (void)hipFree(p);
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
    int cu_count = 256;
    uint32_t colour_blocks = 256;

    Bodies B{};
    float4 *pre_v = nullptr, *pre_w = nullptr;  // solver input velocities of the last step (kept on request)
    bool have_pre = false;
    uint64_t *key = nullptr, *key_alt = nullptr;
    uint32_t *val = nullptr, *val_alt = nullptr, *sort_tmp = nullptr;
    CellTable ct{};

    uint32_t *pair_cnt = nullptr, *pair_off = nullptr, *scan_tmp = nullptr, *pair_stash = nullptr;
    uint32_t *pa = nullptr, *pb = nullptr;
    uint32_t *p_np = nullptr, *man_off = nullptr, *epa_list = nullptr, *pscan_tmp = nullptr;
    uint32_t *p_kind = nullptr, *p_order = nullptr;  // narrowphase: kind of test per pair, pairs ordered by kind
    uint32_t* pkinds = nullptr;                       // 16 pair-kind counts + 16 fill cursors
    float4 *p_pt = nullptr, *p_nf = nullptr;

    uint2* m_ab = nullptr;
    uint32_t* m_np = nullptr;
    float4 *c_pt = nullptr, *c_nf = nullptr;

    // learned warm start: per-point network inputs (rows 4 * slot + point, 24 floats)
    // and outputs (3 floats), kept while logging or while a network is attached
    bool want_feat = false;
    float* feat = nullptr;
    float* ymlp = nullptr;
    axm_mlp_s* mlp = nullptr;
    float last_dt = 0.0f;

    // persistent sweep launch shape: workgroups (one wave each, all resident at
    // once), LDS per workgroup for the staged rows of its slots, and the policy
    // (bit 1 poll a 16-byte hint before the full load, bit 2 adaptive back-off,
    // bits 4-6 poll back-off).
    // Set at creation to one workgroup per CU with the largest LDS a workgroup
    // may have.
    uint32_t sweep_groups = 0;
    uint32_t sweep_lds = 0;
    uint32_t sweep_tune = 0x12;
    uint32_t max_lds = 0;       // device limit of LDS per workgroup

    uint32_t* col = nullptr;
    uint32_t* clist = nullptr;  // colouring: the manifolds not kept from the previous step
    uint64_t *prio = nullptr, *bmax0 = nullptr, *bmax1 = nullptr, *bmask = nullptr;
    uint32_t *ckey = nullptr, *ckey_alt = nullptr, *cval = nullptr, *cval_alt = nullptr, *csort_tmp = nullptr;
    uint2* isl_list = nullptr;  // the small islands' work lists (first slot, manifolds)
    const uint32_t* solve_order = nullptr;  // the sort output holding the last step's solve order
    uint32_t* col_off = nullptr;
    uint32_t* bdone = nullptr;
    GridBar* bar = nullptr;

    SolverRows R[2]{};
    uint32_t rcap[2] = {0, 0};
    uint32_t* np_slot[2] = {nullptr, nullptr};
    uint64_t* pc_key[2] = {nullptr, nullptr};
    uint32_t* pc_pos[2] = {nullptr, nullptr};  // per slot: its entry's position in pc_key (see cache_m)
    uint32_t* pc_slot[2] = {nullptr, nullptr};
    uint32_t pc_mask[2] = {0, 0};
    int cur = 0;
    bool prev_valid = false;

    uint32_t* res_bits = nullptr;  // residual per sweep (res_cap words), then check arrival rows
    uint32_t res_cap = 0;
    Counters* cnt = nullptr;
    Counters* h_cnt = nullptr;

    // kernel profiling
    bool prof_on = false;
    std::vector<hipEvent_t> prof_ev;
    axm_kernel_profile prof{};

    // captured steps (one per warm-start buffer parity); with profiling on, the
    // step is replayed as two graphs around an eagerly launched, event-timed sweep
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    hipGraphExec_t gpre[2] = {nullptr, nullptr};
    hipGraphExec_t gpost[2] = {nullptr, nullptr};
    float graph_dt = 0.0f;
    uint32_t graph_iters = 0;
    hipEvent_t gprof_ev[4] = {nullptr, nullptr, nullptr, nullptr};  // sweep-kernel events per parity
    hipEvent_t bprof_ev[4] = {nullptr, nullptr, nullptr, nullptr};  // bundle-kernel events (eager / per parity)
    const uint32_t* solve_keys = nullptr;  // the sorted keys of the last step's solve order

    // last step
    bool have_step = false;
    uint32_t last_iters = 0;
    bool last_iters_dev = false;
    bool use_filter = false;  // some body of the current set has a collision filter  // the last step's sweep count is in Counters::sweeps_run
    std::vector<float> h_res;

    void drop_graphs() {
        for (hipGraphExec_t* set : {gexec, gpre, gpost})
            for (int p = 0; p < 2; ++p)
                if (set[p]) {
                    (void)hipGraphExecDestroy(set[p]);
                    set[p] = nullptr;
                }
    }
    ~axm_world_s() {
        if (stream) (void)hipStreamSynchronize(stream);
        drop_graphs();
        for (hipEvent_t e : prof_ev) (void)hipEventDestroy(e);
        for (hipEvent_t e : gprof_ev)
            if (e) (void)hipEventDestroy(e);
        for (hipEvent_t e : bprof_ev)
            if (e) (void)hipEventDestroy(e);
        mem.release_all();
        if (h_cnt) (void)hipHostFree(h_cnt);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

// ------------------------------------------------------------------ buffers
void free_bodies(axm_world_s* w) {
    Bodies& B = w->B;
    float4** f4s[] = {&B.x, &B.q, &B.vel, &B.angv, &B.ii, &B.shape, &B.sv, &B.sw,
                      &B.smin, &B.smax, &w->pre_v, &w->pre_w};
    for (float4** p : f4s) w->mem.release(*p);
    w->mem.release(B.aabb);
    w->mem.release(w->pair_stash);
    w->mem.release(B.vslot);
    w->mem.release(B.mat);
    w->mem.release(B.bfl);
    w->mem.release(B.slp);
    w->mem.release(B.upar);
    w->mem.release(B.imin);
    w->mem.release(B.wake);
    w->mem.release(B.icnt);
    w->mem.release(B.world);
    w->mem.release(B.world_first);
    w->mem.release(w->key);
    w->mem.release(w->key_alt);
    w->mem.release(w->val);
    w->mem.release(w->val_alt);
    w->mem.release(w->sort_tmp);
    w->mem.release(w->ct.ent);
    w->mem.release(w->ct.bits);
    w->mem.release(w->ct.mlist);
    w->mem.release(w->ct.cstart);
    w->mem.release(w->ct.crank);
    w->mem.release(w->pair_cnt);
    w->mem.release(w->pair_off);
    w->mem.release(w->scan_tmp);
    w->mem.release(w->bmax0);
    w->mem.release(w->bmax1);
    w->mem.release(w->bmask);
    w->mem.release(w->bdone);
    w->cap_n = 0;
}

axm_status alloc_bodies(axm_world_s* w, uint32_t cap) {
    free_bodies(w);
    Bodies& B = w->B;
    float4** f4s[] = {&B.x, &B.q, &B.vel, &B.angv, &B.ii, &B.shape, &B.sv, &B.sw,
                      &B.smin, &B.smax, &w->pre_v, &w->pre_w};
    for (float4** p : f4s) AXM_TRY(w->mem.get(*p, cap));
    AXM_TRY(w->mem.get(B.aabb, (size_t)cap * 2));
    AXM_TRY(w->mem.get(w->pair_stash, (size_t)cap * BP_PARTS * PAIR_STASH));
    AXM_TRY(w->mem.get(B.vslot, (size_t)cap * 64));
    AXM_TRY(w->mem.get(B.mat, cap));
    AXM_TRY(w->mem.get(B.bfl, cap));
    AXM_TRY(w->mem.get(B.slp, cap));
    AXM_TRY(w->mem.get(B.upar, cap));
    AXM_TRY(w->mem.get(B.imin, cap));
    AXM_TRY(w->mem.get(B.wake, cap));
    AXM_TRY(w->mem.get(B.icnt, cap));
    AXM_TRY(w->mem.get(B.world, cap));
    AXM_TRY(w->mem.get(B.world_first, cap + 1));
    AXM_TRY(w->mem.get(w->key, cap));
    AXM_TRY(w->mem.get(w->key_alt, cap));
    AXM_TRY(w->mem.get(w->val, cap));
    AXM_TRY(w->mem.get(w->val_alt, cap));
    AXM_TRY(w->mem.get(w->sort_tmp, radix_tmp_words(cap)));
    // the scans' look-back scratch must start zeroed (each scan leaves it zeroed)
    if (hipMemset(w->sort_tmp, 0, 4 * radix_tmp_words(cap)) != hipSuccess) return AXM_GPU_OPERATION_FAILED;
    const uint32_t T = pow2_at_least(2ull * cap + 16);
    AXM_TRY(w->mem.get(w->ct.ent, T));
    w->ct.mask = T - 1;
    const uint32_t TB = std::min<uint32_t>(T, 1u << 24) * 8u;  // occupancy bits: 8 per table entry, at most 128 Mi
    AXM_TRY(w->mem.get(w->ct.bits, TB / 32u));
    w->ct.bmask = TB - 1;
    AXM_TRY(w->mem.get(w->ct.mlist, cap));
    AXM_TRY(w->mem.get(w->ct.cstart, cap + 1));
    AXM_TRY(w->mem.get(w->ct.crank, cap));
    AXM_TRY(w->mem.get(w->pair_cnt, (size_t)BP_PARTS * cap));
    AXM_TRY(w->mem.get(w->pair_off, (size_t)BP_PARTS * cap + 1));
    AXM_TRY(w->mem.get(w->scan_tmp, scan_tmp_words(BP_PARTS * cap)));
    if (hipMemset(w->scan_tmp, 0, 4 * scan_tmp_words(BP_PARTS * cap)) != hipSuccess) return AXM_GPU_OPERATION_FAILED;
    AXM_TRY(w->mem.get(w->bmax0, cap));
    AXM_TRY(w->mem.get(w->bmax1, cap));
    AXM_TRY(w->mem.get(w->bmask, cap));
    AXM_TRY(w->mem.get(w->bdone, cap));
    w->cap_n = cap;
    return AXM_SUCCESS;
}

axm_status alloc_pairs(axm_world_s* w, uint32_t cap) {
    w->mem.release(w->pa);
    w->mem.release(w->pb);
    w->mem.release(w->p_np);
    w->mem.release(w->man_off);
    w->mem.release(w->epa_list);
    w->mem.release(w->p_kind);
    w->mem.release(w->p_order);
    w->mem.release(w->p_pt);
    w->mem.release(w->p_nf);
    w->mem.release(w->pscan_tmp);
    AXM_TRY(w->mem.get(w->pa, cap));
    AXM_TRY(w->mem.get(w->pb, cap));
    AXM_TRY(w->mem.get(w->p_np, cap));
    AXM_TRY(w->mem.get(w->man_off, (size_t)cap + 1));
    AXM_TRY(w->mem.get(w->epa_list, cap));
    AXM_TRY(w->mem.get(w->p_kind, cap));
    AXM_TRY(w->mem.get(w->p_order, cap));
    AXM_TRY(w->mem.get(w->p_pt, (size_t)cap * 4));
    AXM_TRY(w->mem.get(w->p_nf, (size_t)cap * 4));
    AXM_TRY(w->mem.get(w->pscan_tmp, scan_tmp_words(cap)));
    if (hipMemset(w->pscan_tmp, 0, 4 * scan_tmp_words(cap)) != hipSuccess) return AXM_GPU_OPERATION_FAILED;
    w->cap_pairs = cap;
    return AXM_SUCCESS;
}

axm_status alloc_rows(axm_world_s* w, int set, uint32_t cap) {
    SolverRows& R = w->R[set];
    w->mem.release(R.hdr);
    w->mem.release(R.J);
    w->mem.release(R.lam);
    w->mem.release(R.lam0);
    w->mem.release(R.fid);
    w->mem.release(R.dep);
    w->mem.release(R.colour);
    w->mem.release(R.inbox);
    w->mem.release(R.next);
    w->mem.release(w->np_slot[set]);
    w->mem.release(w->pc_key[set]);
    w->mem.release(w->pc_slot[set]);
    w->mem.release(w->pc_pos[set]);
    AXM_TRY(w->mem.get(R.hdr, cap));
    AXM_TRY(w->mem.get(R.J, (size_t)cap * 4 * JROWS));
    AXM_TRY(w->mem.get(R.lam, (size_t)cap * 4));
    AXM_TRY(w->mem.get(R.lam0, (size_t)cap * 4));
    AXM_TRY(w->mem.get(R.fid, (size_t)cap * 4));
    AXM_TRY(w->mem.get(R.dep, cap));
    AXM_TRY(w->mem.get(R.colour, cap));
    AXM_TRY(w->mem.get(R.inbox, (size_t)cap * INBOX_WORDS));
    AXM_TRY(w->mem.get(R.next, cap));
    AXM_TRY(w->mem.get(w->np_slot[set], cap));
    const uint32_t T = pow2_at_least(2ull * cap + 16);
    AXM_TRY(w->mem.get(w->pc_key[set], T));
    AXM_TRY(w->mem.get(w->pc_slot[set], T));
    AXM_TRY(w->mem.get(w->pc_pos[set], cap));
    // an empty table: no warm-start hit until a step has filled it
    if (hipMemset(w->pc_key[set], 0xFF, 8ull * T) != hipSuccess) return AXM_GPU_OPERATION_FAILED;
    if (w->cnt && hipMemset(&w->cnt->cache_m[set], 0, 4) != hipSuccess) return AXM_GPU_OPERATION_FAILED;
    if (hipMemset(w->np_slot[set], 0, 4ull * cap) != hipSuccess) return AXM_GPU_OPERATION_FAILED;
    w->pc_mask[set] = T - 1;
    w->rcap[set] = cap;
    return AXM_SUCCESS;
}

// Grows the row set holding the previous step's impulses while keeping them:
// impulses and feature ids are re-strided from [k * old + s] to [k * cap + s]
// and the (a, b) -> slot table is rehashed into one sized for `cap` slots.
axm_status grow_rows_keep(axm_world_s* w, int set, uint32_t cap) {
    SolverRows& R = w->R[set];
    const uint32_t old = w->rcap[set];
    if (old >= cap) return AXM_SUCCESS;
    float4 *lam = nullptr, *lam0 = nullptr, *jr = nullptr;
    uint32_t *fid = nullptr, *np = nullptr;
    uint4* hdr = nullptr;
    uint2* dep = nullptr;
    uint32_t* colour = nullptr;
    uint64_t* inbox = nullptr;
    uint2* next = nullptr;
    AXM_TRY(w->mem.get(colour, cap));
    AXM_TRY(w->mem.get(inbox, (size_t)cap * INBOX_WORDS));
    AXM_TRY(w->mem.get(next, cap));
    AXM_TRY(w->mem.get(lam, (size_t)cap * 4));
    AXM_TRY(w->mem.get(fid, (size_t)cap * 4));
    AXM_TRY(w->mem.get(np, cap));
    AXM_TRY(w->mem.get(lam0, (size_t)cap * 4));
    AXM_TRY(w->mem.get(jr, (size_t)cap * 4 * JROWS));
    AXM_TRY(w->mem.get(hdr, cap));
    AXM_TRY(w->mem.get(dep, cap));
    const uint32_t T = pow2_at_least(2ull * cap + 16);
    uint64_t* tkey = nullptr;
    uint32_t *tslot = nullptr, *tpos = nullptr;
    AXM_TRY(w->mem.get(tkey, T));
    AXM_TRY(w->mem.get(tslot, T));
    AXM_TRY(w->mem.get(tpos, cap));
    launch_fill_u64(tkey, EMPTY_KEY, T, w->stream);
    launch_cache_rehash(w->pc_key[set], w->pc_slot[set], w->pc_mask[set] + 1, tkey, tslot, T - 1, w->cnt, tpos,
                        w->stream);
    AXM_HIP(hipMemcpy2DAsync(lam, 16ull * cap, R.lam, 16ull * old, 16ull * old, 4, hipMemcpyDeviceToDevice, w->stream));
    AXM_HIP(hipMemcpy2DAsync(fid, 4ull * cap, R.fid, 4ull * old, 4ull * old, 4, hipMemcpyDeviceToDevice, w->stream));
    AXM_HIP(hipMemcpyAsync(np, w->np_slot[set], 4ull * old, hipMemcpyDeviceToDevice, w->stream));
    AXM_HIP(hipMemcpyAsync(colour, R.colour, 4ull * old, hipMemcpyDeviceToDevice, w->stream));
    AXM_HIP(hipStreamSynchronize(w->stream));
    w->mem.release(R.lam);
    w->mem.release(R.fid);
    w->mem.release(w->np_slot[set]);
    w->mem.release(R.lam0);
    w->mem.release(R.J);
    w->mem.release(R.hdr);
    w->mem.release(R.dep);
    w->mem.release(R.colour);
    w->mem.release(R.inbox);
    w->mem.release(R.next);
    w->mem.release(w->pc_key[set]);
    w->mem.release(w->pc_slot[set]);
    w->mem.release(w->pc_pos[set]);
    w->pc_key[set] = tkey;
    w->pc_slot[set] = tslot;
    w->pc_pos[set] = tpos;
    w->pc_mask[set] = T - 1;
    R.lam = lam;
    R.fid = fid;
    w->np_slot[set] = np;
    R.lam0 = lam0;
    R.J = jr;
    R.hdr = hdr;
    R.dep = dep;
    R.colour = colour;
    R.inbox = inbox;
    R.next = next;
    w->rcap[set] = cap;
    return AXM_SUCCESS;
}

// Per-point network input / output buffers sized for the manifold capacity.
axm_status alloc_feat(axm_world_s* w) {
    w->mem.release(w->feat);
    w->mem.release(w->ymlp);
    if (!w->want_feat || !w->cap_man) return AXM_SUCCESS;
    AXM_TRY(w->mem.get(w->feat, (size_t)w->cap_man * 4 * MLP_IN));
    AXM_TRY(w->mem.get(w->ymlp, (size_t)w->cap_man * 4 * MLP_OUT));
    return AXM_SUCCESS;
}

// Grows the manifold buffers.  The current row set is reallocated; the other
// set (the previous step's impulses, i.e. the next warm start) is grown in place
// so both sets always hold at least `cap` slots.
axm_status alloc_manifolds(axm_world_s* w, uint32_t cap) {
    w->mem.release(w->m_ab);
    w->mem.release(w->m_np);
    w->mem.release(w->c_pt);
    w->mem.release(w->c_nf);
    w->mem.release(w->col);
    w->mem.release(w->prio);
    w->mem.release(w->clist);
    w->mem.release(w->ckey);
    w->mem.release(w->ckey_alt);
    w->mem.release(w->cval);
    w->mem.release(w->cval_alt);
    w->mem.release(w->csort_tmp);
    w->mem.release(w->isl_list);
    AXM_TRY(w->mem.get(w->m_ab, cap));
    AXM_TRY(w->mem.get(w->m_np, cap));
    AXM_TRY(w->mem.get(w->c_pt, (size_t)cap * 4));
    AXM_TRY(w->mem.get(w->c_nf, (size_t)cap * 4));
    AXM_TRY(w->mem.get(w->col, cap));
    AXM_TRY(w->mem.get(w->prio, cap));
    AXM_TRY(w->mem.get(w->clist, cap));
    AXM_TRY(w->mem.get(w->ckey, cap));
    AXM_TRY(w->mem.get(w->ckey_alt, cap));
    AXM_TRY(w->mem.get(w->cval, cap));
    AXM_TRY(w->mem.get(w->cval_alt, cap));
    AXM_TRY(w->mem.get(w->csort_tmp, radix_tmp_words(cap)));
    AXM_TRY(w->mem.get(w->isl_list, cap));
    if (hipMemset(w->csort_tmp, 0, 4 * radix_tmp_words(cap)) != hipSuccess) return AXM_GPU_OPERATION_FAILED;
    w->cap_man = cap;
    if (w->rcap[w->cur] < cap) AXM_TRY(alloc_rows(w, w->cur, cap));
    if (w->rcap[1 - w->cur] > 0 && w->rcap[1 - w->cur] < cap) AXM_TRY(grow_rows_keep(w, 1 - w->cur, cap));
    return alloc_feat(w);
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
    sp.soft1 = 1.0f + d.softness;
    sp.keep = 1.0f / sp.soft1;
    sp.allow_sleep = d.allow_sleep;
    sp.use_filter = w->use_filter ? 1u : 0u;
    sp.sleep_lin2 = d.sleep_linear * d.sleep_linear;
    sp.sleep_ang2 = d.sleep_angular * d.sleep_angular;
    sp.sleep_time = d.sleep_time;
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

axm_status clear_sticky(axm_world_s* w) {
    AXM_HIP(hipMemsetAsync(&w->cnt->sticky, 0, sizeof(uint32_t), w->stream));
    return AXM_SUCCESS;
}

// The bundle kernel's profiling events of slot pair p (created on first use).
axm_status bundle_events(axm_world_s* w, int p) {
    for (int k = 2 * p; k < 2 * p + 2; ++k)
        if (!w->bprof_ev[k]) AXM_HIP(hipEventCreate(&w->bprof_ev[k]));
    return AXM_SUCCESS;
}

// The next profiling event (created on first use), or null when not profiling.
axm_status prof_event(axm_world_s* w, size_t& nev, bool capturing, hipEvent_t& e) {
    e = nullptr;
    if (!w->prof_on || capturing) return AXM_SUCCESS;
    while (w->prof_ev.size() < nev + 1) {
        hipEvent_t x;
        AXM_HIP(hipEventCreate(&x));
        w->prof_ev.push_back(x);
    }
    e = w->prof_ev[nev++];
    return AXM_SUCCESS;
}
axm_status prof_mark(axm_world_s* w, size_t& nev, bool capturing) {
    hipEvent_t e;
    AXM_TRY(prof_event(w, nev, capturing, e));
    if (e) AXM_HIP(hipEventRecord(e, w->stream));
    return AXM_SUCCESS;
}

// A step is enqueued in three phases so that a profiled graph replay can launch
// the sweep kernel eagerly between two captured graphs, bracketed by events.
enum : uint32_t { PH_PRE = 1u, PH_SWEEP = 2u, PH_POST = 4u, PH_ALL = 7u };
constexpr uint32_t SWEEP_SPIN_LIMIT_US = 2000000u;  // bounded dependency wait of the sweep kernel

struct StepOpts {
    uint32_t iters = 10;
    uint32_t flags = 0;
    const float4* warm_in = nullptr;
    bool colour_launches = false;  // one sweep launch per colour (the host reads the colour ranges)
    bool capture = false;          // being captured into a graph: no host interaction
    hipEvent_t* ev = nullptr;      // optional stage events: start, broad, narrow, solve, end
    uint32_t phases = PH_ALL;      // which part of the step to enqueue
};

// ---- broadphase: solver velocities, AABBs, keys, sort, cell table, pairs
void enqueue_broad(axm_world_s* w, const StepParams& sp) {
    const uint32_t n = w->n;
    hipStream_t s = w->stream;
    // with the Morton bitmap (one world) the keys are the 30-bit Morton codes, sorted as
    // 32-bit keys in the same buffers (a third less data per pass than 64-bit ones)
    const bool k32 = w->ct.mbits != nullptr;
    uint32_t* key32 = reinterpret_cast<uint32_t*>(w->key);
    uint32_t* key32_alt = reinterpret_cast<uint32_t*>(w->key_alt);
    launch_begin(w->B, sp, n, w->key, k32 ? key32 : nullptr, w->val, w->bmax0, w->bmax1, w->bmask, w->bdone, w->cnt,
                 w->pkinds, w->ct, s);
    SortedKeys skey;
    int alt;
    if (k32) {
        alt = radix_sort_u32(key32, w->val, key32_alt, w->val_alt, n, 30, w->sort_tmp, s, nullptr, &w->cnt->sticky);
        skey.k32 = alt ? key32_alt : key32;
    } else {
        const int end_bit = 30 + bits_for(w->n_worlds > 1 ? w->n_worlds - 1 : 0);
        alt = radix_sort_u64(w->key, w->val, w->key_alt, w->val_alt, n, end_bit, w->sort_tmp, s, nullptr,
                             &w->cnt->sticky);
        skey.k64 = alt ? w->key_alt : w->key;
    }
    const uint32_t* sval = alt ? w->val_alt : w->val;
    launch_cell_clear(w->ct, s);
    launch_cell_table(w->B, skey, sval, n, w->ct, w->cnt, w->scan_tmp, s);
    // pair offsets in two levels: per-workgroup totals (256 parts each), scanned, then the
    // emit pass scans within each workgroup
    const uint32_t pwg = (BP_PARTS * n + 255u) / 256u;
    uint32_t* bsum = w->pair_off;
    uint32_t* boff = w->pair_off + pwg;
    launch_pair_count(w->B, sp, skey, sval, w->ct, n, w->pair_cnt, bsum, w->pair_stash, w->cnt, s);
    scan_exclusive(bsum, boff, pwg, w->scan_tmp, s, &w->cnt->pairs, nullptr, false,
                   &w->cnt->sticky);
    launch_pair_emit(w->B, sp, skey, sval, w->ct, n, w->pair_cnt, boff, w->cap_pairs, w->pa, w->pb, w->cnt, w->pair_stash,
                     s);
}

// Whether the step solves islands of at most ISLAND_LOCAL_MAX manifolds on their own
// (not with a per-sweep residual history, which is one over all manifolds).
uint32_t island_split(const axm_world_s* w) {
    return w->desc.island_solve && !w->desc.record_residual && w->n < ISLAND_ROOT_LIMIT ? 1u : 0u;
}

// ---- narrowphase into per-pair staging, compaction into manifolds (pair order)
void enqueue_narrow(axm_world_s* w, const StepParams& sp) {
    hipStream_t s = w->stream;
    launch_narrow(w->B, sp, w->pa, w->pb, w->cap_pairs, w->p_np, w->p_pt, w->p_nf, w->epa_list, w->p_kind, w->pkinds,
                  w->p_order, 4u * (uint32_t)w->cu_count, w->cnt, s);
    // manifold offsets: scan of (points > 0) over the pairs found this step
    scan_exclusive(w->p_np, w->man_off, w->cap_pairs, w->pscan_tmp, s, &w->cnt->manifolds, &w->cnt->p_eff, true,
                   &w->cnt->sticky);
    launch_compact(w->pa, w->pb, w->cap_pairs, w->p_np, w->p_pt, w->p_nf, w->man_off, w->cap_man, w->m_ab, w->m_np,
                   w->c_pt, w->c_nf, w->cnt, s);
}

// Broadphase, narrowphase, colouring and row preparation of one step.
axm_status enqueue_pre(axm_world_s* w, const StepParams& sp, const StepOpts& o) {
    const uint32_t n = w->n;
    hipStream_t s = w->stream;
    const int c = w->cur, pv = 1 - c;
    const SolverRows& R = w->R[c];
    const uint32_t rc = w->rcap[c];
    if (o.ev) AXM_HIP(hipEventRecord(o.ev[0], s));
    enqueue_broad(w, sp);  // its first launch resets the step's counters
    if (o.ev) AXM_HIP(hipEventRecord(o.ev[1], s));
    if (o.flags & AXM_STEP_KEEP_STAGES) {
        AXM_HIP(hipMemcpyAsync(w->pre_v, w->B.sv, 16ull * n, hipMemcpyDeviceToDevice, s));
        AXM_HIP(hipMemcpyAsync(w->pre_w, w->B.sw, 16ull * n, hipMemcpyDeviceToDevice, s));
    }
    enqueue_narrow(w, sp);
    if (o.ev) AXM_HIP(hipEventRecord(o.ev[2], s));
    // previous step's (a, b) -> slot table: warm start and incremental colouring
    PrevCacheArgs pca{};
    pca.key = w->pc_key[pv];
    pca.slot = w->pc_slot[pv];
    pca.mask = w->pc_mask[pv];
    pca.lam = w->R[pv].lam;
    pca.fid = w->R[pv].fid;
    pca.np = w->np_slot[pv];
    pca.colour = w->R[pv].colour;
    pca.cap = w->rcap[pv];
    pca.valid = (w->prev_valid && pca.key) ? 1u : 0u;
    // ---- islands, colouring and the solve order (by colour; the small islands after the
    // large ones, island by island)
    launch_isl_union(w->B, n, w->m_ab, w->cap_man, w->cnt, s);
    // bdone is the per-body manifold count of the colouring (cleared by k_begin)
    const uint32_t split = island_split(w);
    launch_colour(w->B, w->m_ab, w->cap_man, w->col, w->prio, w->bmax0, w->bmax1, w->bmask, w->bdone, w->clist,
                  w->bar, w->cnt, w->colour_blocks, pca, split, s);
    launch_col_sort_input(w->B, w->m_ab, w->col, w->cap_man, w->ckey, w->cval, w->cnt, split, s);
    radix_sort_u32(w->ckey, w->cval, w->ckey_alt, w->cval_alt, w->cap_man, split ? 32 : 8, w->csort_tmp, s,
                   &w->cnt->m_eff, &w->cnt->sticky);
    // an odd number of 8-bit passes leaves the result in the alternate buffers
    const uint32_t* scol = split ? w->ckey : w->ckey_alt;
    const uint32_t* order = split ? w->cval : w->cval_alt;
    w->solve_order = order;
    w->solve_keys = scol;
    // ---- solver
    if (w->desc.warm_start == AXM_WARM_NONE) pca.valid = 0;
    // the same launch forms the colour offsets, builds the small islands' work list, clears
    // the residual words (the per-sweep residuals, then the check arrivals) and empties this
    // row set's (a, b) -> slot table, which the step's integration launch refills
    PrepExtra X;
    X.skey = scol;
    X.col_off = w->col_off;
    X.max_k = MAX_SORT_COLOURS;
    X.list = split ? w->isl_list : nullptr;
    X.zero = w->res_bits;
    X.zero_n = (uint32_t)res_words(w->res_cap);
    X.ckey = w->pc_key[c];
    X.cpos = w->pc_pos[c];
    X.cset = (uint32_t)c;
    launch_prep(w->B, sp, order, rc, w->m_ab, w->m_np, w->c_pt, w->c_nf, R, pca, o.warm_in, w->col, w->bmask,
                w->bdone, w->cnt, w->want_feat ? w->feat : nullptr, X, s);
    if (w->mlp && w->desc.warm_start == AXM_WARM_MLP && !o.warm_in && w->feat) {
        launch_mlp_forward(w->mlp->dev(), w->feat, 4u * w->cap_man, &w->cnt->m_eff, 4u, w->ymlp,
                           2u * (uint32_t)w->cu_count, s);
        launch_mlp_apply(R, w->cnt, rc, w->ymlp, sp, s);
    }
    return AXM_SUCCESS;
}

// The persistent sweep's residual checks (a tolerance stop or a per-sweep residual
// history), when the step has any; the per-colour launches check on the host.
bool sweep_checks(const axm_world_s* w, bool colour_launches, SweepStop& st) {
    if (colour_launches || !(w->desc.residual_tol > 0.0f || w->desc.record_residual)) return false;
    st.bits = w->res_bits;
    st.arrive = w->res_bits + w->res_cap;
    st.tol = w->desc.residual_tol > 0.0f ? w->desc.residual_tol : -1.0f;
    st.every = w->desc.record_residual ? 1u : std::max<uint32_t>(w->desc.residual_every, 1u);
    return true;
}

// Enqueues one step.  Synchronises only for per-colour sweep launches (with a
// residual tolerance they check it on the host); the persistent sweep checks the
// residual itself, so a step with a tolerance enqueues without synchronising.
axm_status enqueue_step(axm_world_s* w, float dt, const StepOpts& o) {
    // every kernel below sizes its grid and row strides by the current set's capacity
    if (w->rcap[w->cur] < w->cap_man || w->rcap[1 - w->cur] < w->cap_man) return AXM_GPU_INVALID_OPERATION;
    const uint32_t n = w->n;
    hipStream_t s = w->stream;
    const StepParams sp = make_params(w, dt);
    size_t nev = 0;
    const bool pre = (o.phases & PH_PRE) != 0, sweep = (o.phases & PH_SWEEP) != 0, post = (o.phases & PH_POST) != 0;
    const int c = w->cur, pv = 1 - c;
    const SolverRows& R = w->R[c];
    const uint32_t rc = w->rcap[c];
    const bool tol_mode = w->desc.residual_tol > 0.0f && !o.capture;
    const uint32_t max_it =
        w->desc.residual_tol > 0.0f ? std::max<uint32_t>(w->desc.max_iterations, 1u) : o.iters;
    // residual checks inside the sweep launch: with a tolerance, or every sweep's residual recorded
    SweepStop stop;
    const bool dev_checks = sweep_checks(w, o.colour_launches, stop);
    w->last_iters_dev = dev_checks;
    std::vector<uint32_t> off;
    if (pre) AXM_TRY(enqueue_pre(w, sp, o));
    // the islands solved on their own run first; their launch also sets up the large
    // islands' sweep launch (its inboxes and next visits)
    const bool flow_inited = sweep && island_split(w) && !o.colour_launches;
    if (sweep && island_split(w)) {
        hipEvent_t b0 = nullptr, b1 = nullptr;
        if (w->prof_on && !o.capture) {
            AXM_TRY(bundle_events(w, 0));
            b0 = w->bprof_ev[0];
            b1 = w->bprof_ev[1];
        }
        launch_pgs_local(w->B, sp, R, w->cnt, rc, w->solve_keys, w->isl_list, w->desc.residual_tol, max_it,
                         2u * (uint32_t)w->cu_count, s, b0, b1, flow_inited ? w->bmask : nullptr);
    }
    const bool want_res = (o.flags & AXM_STEP_RESIDUAL) || w->desc.record_residual || tol_mode;
    const uint32_t every = std::max<uint32_t>(w->desc.residual_every, 1u);
    uint32_t it = sweep ? 0u : max_it;
    if (sweep && o.colour_launches) {
        AXM_TRY(read_counters(w));
        if (w->h_cnt->sticky != 0u) {
            off.assign(1, 0u);  // the step already failed: its rows were not prepared, launch nothing
        } else {
            const uint32_t K = std::min(w->h_cnt->max_colour, MAX_SORT_COLOURS);
            off.resize(K + 1);
            AXM_HIP(hipMemcpyAsync(off.data(), w->col_off, 4ull * (K + 1), hipMemcpyDeviceToHost, s));
            AXM_HIP(hipStreamSynchronize(s));
        }
    }
    if (dev_checks && sweep) {
        hipEvent_t e0, e1;
        AXM_TRY(prof_event(w, nev, o.capture, e0));
        AXM_TRY(prof_event(w, nev, o.capture, e1));
        launch_pgs_flow(w->B, sp, R, rc, w->bmask, 0, max_it, w->cnt, SWEEP_SPIN_LIMIT_US, w->sweep_groups,
                        w->sweep_lds, w->sweep_tune, s, e0, e1, stop, flow_inited);
        it = max_it;  // at most; the count the launch ran is read back with the counters
    }
    bool all_local = false;
    if (sweep && o.colour_launches && !off.empty() && off.back() == 0u && w->h_cnt->m_eff > 0u) {
        // every island was solved on its own: no sweep of the large ones runs
        all_local = true;
        it = max_it;
        launch_fill_u32(&w->cnt->tol_hit, 1u, 1u, s);
    }
    while (it < max_it && !dev_checks) {
        uint32_t seg = max_it - it;
        if (w->desc.record_residual) seg = 1;
        else if (tol_mode) seg = std::min(seg, every - (it % every));
        if (o.colour_launches) {
            AXM_TRY(prof_mark(w, nev, o.capture));
            for (uint32_t k = 0; k < seg; ++k)
                for (size_t col = 0; col + 1 < off.size(); ++col)
                    launch_pgs(w->B, sp, R, off[col], off[col + 1], rc, it + k == 0 ? 1u : 0u, s);
            AXM_TRY(prof_mark(w, nev, o.capture));
        } else {
            // the events bracket the sweep kernel itself
            hipEvent_t e0, e1;
            AXM_TRY(prof_event(w, nev, o.capture, e0));
            AXM_TRY(prof_event(w, nev, o.capture, e1));
            launch_pgs_flow(w->B, sp, R, rc, w->bmask, it, seg, w->cnt, SWEEP_SPIN_LIMIT_US,
                            w->sweep_groups, w->sweep_lds, w->sweep_tune, s, e0, e1, SweepStop{},
                            flow_inited && it == 0u);
        }
        it += seg;
        const bool last = it == max_it;
        const bool check = tol_mode && (it % every == 0);
        if (w->desc.record_residual || check || (last && want_res)) {
            launch_residual(w->B, sp, R, w->cnt, rc, w->res_bits + it - 1, s);
            if (check) {
                AXM_HIP(hipMemcpyAsync(&w->h_cnt->residual_bits, w->res_bits + it - 1, 4, hipMemcpyDeviceToHost, s));
                AXM_HIP(hipStreamSynchronize(s));
                float res;
                std::memcpy(&res, &w->h_cnt->residual_bits, 4);
                if (res <= w->desc.residual_tol) {
                    launch_fill_u32(&w->cnt->tol_hit, 1u, 1u, s);
                    break;
                }
            }
        }
    }
    if (all_local) it = 0;
    w->last_iters = it;
    if (!post) return AXM_SUCCESS;
    w->last_dt = dt;
    if (o.ev) AXM_HIP(hipEventRecord(o.ev[3], s));
    // the integration launch also builds the step's (a, b) -> slot table (the step's row
    // preparation emptied it)
    const CacheTable ctab{w->pc_key[c], w->pc_slot[c], w->pc_mask[c], w->np_slot[c], w->pc_pos[c], (uint32_t)c};
    launch_integrate(w->B, sp, n, w->cnt, w->last_iters_dev ? SWEEPS_ON_DEVICE : it, s, &R, &ctab);
    launch_islands(w->B, sp, n, w->cnt, s);
    if (o.ev) AXM_HIP(hipEventRecord(o.ev[4], s));
    if (w->prof_on && !o.capture) {
        AXM_HIP(hipStreamSynchronize(s));
        double ms = 0.0;
        for (size_t k = 0; k + 1 < nev; k += 2) {
            float e = 0.0f;
            AXM_HIP(hipEventElapsedTime(&e, w->prof_ev[k], w->prof_ev[k + 1]));
            ms += e;
        }
        AXM_TRY(read_counters(w));
        const uint32_t ran = w->last_iters_dev ? w->h_cnt->sweeps_run : it;
        w->prof.steps += 1;
        w->prof.pgs_launches += nev / 2;
        w->prof.pgs_ms += ms;
        w->prof.pgs_bytes += (double)spread_total(*w->h_cnt, &SpreadRow::pgs_bytes) * (double)ran;
        w->prof.sweeps += ran;
        if (sweep && island_split(w)) {
            float e = 0.0f;
            AXM_HIP(hipEventElapsedTime(&e, w->bprof_ev[0], w->bprof_ev[1]));
            w->prof.bundle_launches += 1;
            w->prof.bundle_ms += e;
            w->prof.bundle_point_sweeps += w->h_cnt->bundle_point_sweeps;
        }
    }
    w->cur = pv;
    w->have_step = true;
    w->prev_valid = true;
    w->have_pre = (o.flags & AXM_STEP_KEEP_STAGES) != 0;
    return AXM_SUCCESS;
}

void fill_stats(axm_world_s* w, axm_step_stats& S) {
    const Counters& h = *w->h_cnt;
    S.pairs = h.pairs;
    S.manifolds = h.manifolds;
    S.contacts = spread_total(h, &SpreadRow::contacts);
    S.colours = h.max_colour;
    S.colour_rounds = h.last_round;
    S.iterations = std::max(w->last_iters, h.loc_sweeps);  // the most sweeps any island ran
    S.gjk_fail = h.gjk_fail;
    S.epa_fail = h.epa_fail;
    S.warm_hits = spread_total(h, &SpreadRow::warm_hits);
    S.overflow = (h.sticky & STICKY_PAIRS ? 1u : 0u) | (h.sticky & STICKY_MANIFOLDS ? 2u : 0u);
    S.islands = spread_total(h, &SpreadRow::islands);
    S.sleeping = spread_total(h, &SpreadRow::sleeping);
}

// Handles a sticky failure of the step just run: grows capacities or switches
// to per-colour sweep launches.  Returns SUCCESS when the step should be run
// again, an error otherwise.  The body state and warm-start buffers are
// unchanged (every persistent write of the failed step was skipped).
axm_status recover(axm_world_s* w, StepOpts& o) {
    const uint32_t st = w->h_cnt->sticky;
    w->cur = 1 - w->cur;  // the failed step's buffer flip is undone
    if (st & STICKY_TIMEOUT) return AXM_GPU_TIMEOUT;
    if (st & STICKY_INTERNAL) return AXM_GPU_OPERATION_FAILED;
    if (st & STICKY_COLOURS) return AXM_CONSTRAINT_SOLVER_FAILURE;
    if (st & (STICKY_PAIRS | STICKY_MANIFOLDS)) {
        if (!w->desc.auto_grow) return AXM_OUT_OF_RANGE;
        if (st & STICKY_PAIRS) {
            const uint64_t P = w->h_cnt->pairs;
            AXM_TRY(alloc_pairs(w, (uint32_t)std::min<uint64_t>(0xFFFFFFF0ull, P + P / 2 + 1024)));
        }
        if (st & STICKY_MANIFOLDS) {
            const uint64_t M = w->h_cnt->manifolds;
            AXM_TRY(alloc_manifolds(w, (uint32_t)std::min<uint64_t>(0xFFFFFFF0ull, M + M / 2 + 1024)));
        }
        w->drop_graphs();
    }
    if (st & STICKY_FLOW_COLOURS) o.colour_launches = true;
    AXM_TRY(clear_sticky(w));
    return AXM_SUCCESS;
}

axm_status ensure_res(axm_world_s* w, uint32_t iters) {
    const uint32_t need = std::max<uint32_t>(iters, w->desc.residual_tol > 0.0f ? w->desc.max_iterations : 0u) + 1;
    if (w->res_cap >= need) return AXM_SUCCESS;
    w->mem.release(w->res_bits);
    AXM_TRY(w->mem.get(w->res_bits, res_words(need)));
    w->res_cap = need;
    w->drop_graphs();
    return AXM_SUCCESS;
}

// One step with recovery; synchronises once.
// Whether the last step met the residual tolerance (large islands and every island solved on
// its own); true without a tolerance.
bool step_reached(const axm_world_s* w) {
    if (!(w->desc.residual_tol > 0.0f)) return true;
    return w->h_cnt->tol_hit != 0u && w->h_cnt->loc_miss == 0u;
}
// With desc.strict, the status of steps that completed with GJK or EPA failures or with steps
// that stopped at max_iterations above the tolerance (their results are kept).
axm_status strict_status(const axm_world_s* w, uint32_t gjk_fail, uint32_t epa_fail, uint32_t missed) {
    if (!w->desc.strict) return AXM_SUCCESS;
    if (gjk_fail) return AXM_GJK_FAILED_TO_CONVERGE;
    if (epa_fail) return AXM_EPA_FAILED_TO_CONVERGE;
    if (missed) return AXM_CONSTRAINT_SOLVER_FAILURE;
    return AXM_SUCCESS;
}

axm_status step_sync(axm_world_s* w, float dt, StepOpts o, axm_step_stats* st) {
    if (!(dt > 0.0f) || !std::isfinite(dt)) return AXM_INVALID_PARAMETER;
    if (w->desc.solver_mode == 1) o.colour_launches = true;
    AXM_TRY(ensure_res(w, o.iters));
    axm_step_stats S{};
    if (w->n == 0) {
        w->have_step = true;
        if (st) *st = S;
        return AXM_SUCCESS;
    }
    const bool timing = (o.flags & AXM_STEP_TIMINGS) != 0;
    hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    if (timing) {
        for (auto& e : ev) AXM_HIP(hipEventCreate(&e));
        o.ev = ev;
    }
    auto free_ev = [&]() {
        for (auto e : ev)
            if (e) (void)hipEventDestroy(e);
    };
    for (int attempt = 0;; ++attempt) {
        const axm_status es = enqueue_step(w, dt, o);
        if (es != AXM_SUCCESS) {
            free_ev();
            return es;
        }
        AXM_TRY(read_counters(w));
        if (w->last_iters_dev) w->last_iters = w->h_cnt->sweeps_run;
        if (w->h_cnt->sticky == 0u) break;
        fill_stats(w, S);
        const axm_status r = recover(w, o);
        if (r != AXM_SUCCESS || attempt >= 3) {
            if (st) *st = S;
            free_ev();
            return r != AXM_SUCCESS ? r : AXM_GPU_OPERATION_FAILED;
        }
    }
    fill_stats(w, S);
    const bool want_res = (o.flags & AXM_STEP_RESIDUAL) || w->desc.record_residual || w->desc.residual_tol > 0.0f;
    w->h_res.clear();
    // the islands solved on their own are finished before the large ones' sweeps: every
    // entry of the history reports the larger of the two residuals
    float loc_res;
    std::memcpy(&loc_res, &w->h_cnt->loc_res_bits, 4);
    if (want_res && w->last_iters) {
        std::vector<uint32_t> bits(w->last_iters);
        AXM_HIP(hipMemcpy(bits.data(), w->res_bits, 4ull * w->last_iters, hipMemcpyDeviceToHost));
        for (uint32_t k = 0; k < w->last_iters; ++k) {
            float f;
            std::memcpy(&f, &bits[k], 4);
            if (bits[k] != 0u || w->desc.record_residual || k + 1 == w->last_iters)
                w->h_res.push_back(std::max(f, loc_res));
        }
    } else if (want_res && w->h_cnt->loc_sweeps) {
        w->h_res.push_back(loc_res);  // only islands solved on their own
    }
    if (!w->h_res.empty()) {
        S.residual = w->h_res.back();
        S.residual_first = w->h_res.front();
    }
    if (timing) {
        float t[4] = {0, 0, 0, 0};
        for (int k = 0; k < 4; ++k) (void)hipEventElapsedTime(&t[k], ev[k], ev[k + 1]);
        S.broadphase_ms = t[0];
        S.narrowphase_ms = t[1];
        S.solver_ms = t[2];
        S.integration_ms = t[3];
        (void)hipEventElapsedTime(&S.total_ms, ev[0], ev[4]);
        free_ev();
    }
    if (st) *st = S;
    return strict_status(w, S.gjk_fail, S.epa_fail, step_reached(w) ? 0u : 1u);
}

// Captures one step with the given warm-start buffer parity into a graph.
axm_status capture(axm_world_s* w, float dt, uint32_t iters, int parity, uint32_t phases, hipGraphExec_t* out) {
    const int saved_cur = w->cur;
    const bool saved_prev = w->prev_valid, saved_step = w->have_step, saved_pre = w->have_pre;
    const uint32_t saved_iters = w->last_iters;
    const bool saved_dev = w->last_iters_dev;
    w->cur = parity;
    w->prev_valid = true;
    StepOpts o;
    o.iters = iters;
    o.capture = true;
    o.phases = phases;
    AXM_HIP(hipStreamBeginCapture(w->stream, hipStreamCaptureModeThreadLocal));
    const axm_status st = enqueue_step(w, dt, o);
    hipGraph_t g = nullptr;
    const hipError_t ee = hipStreamEndCapture(w->stream, &g);
    w->cur = saved_cur;
    w->prev_valid = saved_prev;
    w->have_step = saved_step;
    w->have_pre = saved_pre;
    w->last_iters = saved_iters;
    w->last_iters_dev = saved_dev;
    if (st != AXM_SUCCESS || ee != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return st != AXM_SUCCESS ? st : AXM_GPU_OPERATION_FAILED;
    }
    hipGraphExec_t ex = nullptr;
    const hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ie != hipSuccess) return AXM_GPU_OPERATION_FAILED;
    *out = ex;
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
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, desc->device) == hipSuccess && prop.multiProcessorCount > 0)
        w->cu_count = prop.multiProcessorCount;
    // colouring kernel: one workgroup per CU, all resident (its grid barrier needs that);
    // the occupancy query confirms at least one 256-thread workgroup fits per CU
    w->colour_blocks = (uint32_t)w->cu_count * (uint32_t)std::min(colour_blocks_per_cu(), 1);
    // sweep kernel: the largest LDS a workgroup may have (at most 160 KiB), one workgroup per CU
    w->max_lds = (uint32_t)std::min<size_t>(prop.sharedMemPerBlock, 160u * 1024u);
    w->sweep_lds = w->max_lds & ~16383u;
    const int own_per_cu = pgs_own_setup(w->max_lds, w->sweep_lds);
    if (own_per_cu < 1 || !pgs_bundle_setup()) {
        delete w;
        return AXM_GPU_OPERATION_FAILED;
    }
    w->sweep_groups = (uint32_t)w->cu_count;
    if (hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking) != hipSuccess) {
        delete w;
        return AXM_GPU_OPERATION_FAILED;
    }
    if (hipHostMalloc((void**)&w->h_cnt, sizeof(Counters), hipHostMallocDefault) != hipSuccess) {
        delete w;
        return AXM_BUFFER_ALLOCATION_FAILED;
    }
    axm_status st = w->mem.get(w->cnt, 1);
    if (st == AXM_SUCCESS) st = w->mem.get(w->bar, 1);
    if (st == AXM_SUCCESS) st = w->mem.get(w->pkinds, 32);
    if (st == AXM_SUCCESS) st = w->mem.get(w->col_off, MAX_SORT_COLOURS + 2);
    if (st == AXM_SUCCESS) st = alloc_bodies(w, std::max<uint32_t>(desc->max_bodies, 16u));
    // explicit capacities are honoured as given; 0 means "sized from the body count"
    if (st == AXM_SUCCESS) st = alloc_pairs(w, desc->max_pairs ? desc->max_pairs : 1024u);
    if (st == AXM_SUCCESS) st = alloc_manifolds(w, desc->max_manifolds ? desc->max_manifolds : 1024u);
    if (st == AXM_SUCCESS && w->rcap[1 - w->cur] < w->cap_man) st = alloc_rows(w, 1 - w->cur, w->cap_man);
    if (st == AXM_SUCCESS && (hipMemset(w->cnt, 0, sizeof(Counters)) != hipSuccess ||
                              hipMemset(w->bar, 0, sizeof(GridBar)) != hipSuccess))
        st = AXM_GPU_OPERATION_FAILED;
    if (st != AXM_SUCCESS) {
        delete w;
        return st;
    }
    *out = w;
    return AXM_SUCCESS;
}

axm_status axm_world_destroy(axm_world w) {
    if (!w) return AXM_NULL_POINTER;
    (void)hipSetDevice(w->desc.device);
    delete w;
    return AXM_SUCCESS;
}

axm_status axm_world_body_count(axm_world w, uint32_t* n) {
    if (!w || !n) return AXM_NULL_POINTER;
    *n = w->n;
    return AXM_SUCCESS;
}

// Per-body arrays in device layout (x, q, vel, angv, ii, shape, mat float4; bfl uint4;
// world u32; world_first per world + 1; sleep words and wake stamps, empty for a fresh set).
struct HostBodies {
    std::vector<float4> x, q, v, av, ii, sh, mat;
    std::vector<uint4> bfl;
    std::vector<uint32_t> wd, first;
    std::vector<uint2> slp;
    std::vector<uint32_t> wake;
};
axm_status install_bodies(axm_world_s* w, uint32_t n, uint32_t nw, float cell, bool use_filter,
                          const HostBodies& hb, bool keep_graphs = false);

axm_status axm_world_set_bodies(axm_world w, const axm_body_soa* b, uint32_t n) {
    if (!w || !b) return AXM_NULL_POINTER;
    if (n && (!b->pos || !b->quat || !b->lin_vel || !b->ang_vel || !b->inv_mass || !b->inv_inertia || !b->shape_type ||
              !b->shape_params))
        return AXM_NULL_POINTER;
    if (n >= (1u << 29)) return AXM_OUT_OF_RANGE;
    (void)hipSetDevice(w->desc.device);
    std::vector<float4> x(n), q(n), v(n), av(n), ii(n), sh(n), mat(n);
    std::vector<uint4> bfl(n);
    std::vector<uint32_t> wd(n);
    const axm_world_desc& D = w->desc;
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
        float mu = D.friction, e = D.restitution, ld = D.lin_damp, ad = D.ang_damp;
        if (b->material) {
            const float* mt = b->material + 4 * i;
            if (!std::isfinite(mt[0]) || mt[0] < 0.0f || !std::isfinite(mt[1]) || mt[1] < 0.0f ||
                !std::isfinite(mt[2]) || !std::isfinite(mt[3]))
                return AXM_INVALID_RIGID_BODY;
            mu = mt[0];
            e = mt[1];
            ld = mt[2];
            ad = mt[3];
        }
        mat[i] = make_float4(mu, e, 1.0f - ld, 1.0f - ad);
        const uint32_t fl = b->flags ? b->flags[i] : 0u;
        if ((fl & BODY_KINEMATIC) && im != 0.0f) return AXM_INVALID_RIGID_BODY;  // kinematic bodies have no mass
        bfl[i] = b->filter ? make_uint4(fl, b->filter[3 * i], b->filter[3 * i + 1], b->filter[3 * i + 2])
                           : make_uint4(fl, 1u, 0xFFFFFFFFu, 0u);
    }
    if (nw > (1u << 22)) return AXM_OUT_OF_RANGE;
    std::vector<uint32_t> first(nw + 1, 0);
    for (uint32_t i = n; i-- > 0;) first[wd[i]] = i;
    const float need = 2.0f * (rmax + w->desc.margin);
    float cell = w->desc.cell_size;
    if (cell == 0.0f) cell = need > 0.0f ? need : 1.0f;
    if (cell < need) return AXM_INVALID_PARAMETER;
    HostBodies hb{std::move(x), std::move(q), std::move(v), std::move(av), std::move(ii), std::move(sh),
                  std::move(mat), std::move(bfl), std::move(wd), std::move(first), {}, {}};
    return install_bodies(w, n, std::max(nw, 1u), cell, b->filter != nullptr, hb);
}


// Uploads a body set and resets the per-world step state (counters, warm-start cache,
// occupancy bitmap).  With keep_graphs the captured steps stay (the caller has checked that
// nothing they hold changes: description, body count, worlds, cell, filters) unless a buffer
// has to grow.
axm_status install_bodies(axm_world_s* w, uint32_t n, uint32_t nw, float cell, bool use_filter,
                          const HostBodies& hb, bool keep_graphs) {
    const std::vector<float4>&x = hb.x, &q = hb.q, &v = hb.v, &av = hb.av, &ii = hb.ii, &sh = hb.sh, &mat = hb.mat;
    const std::vector<uint4>& bfl = hb.bfl;
    const std::vector<uint32_t>&wd = hb.wd, &first = hb.first;
    AXM_HIP(hipStreamSynchronize(w->stream));
    const bool grow = n > w->cap_n || (w->desc.max_pairs == 0 && w->cap_pairs < 8ull * n) ||
                      (w->desc.max_manifolds == 0 && w->cap_man < 4ull * n) || w->rcap[1 - w->cur] < w->cap_man;
    if (!keep_graphs || grow) w->drop_graphs();
    if (n > w->cap_n) AXM_TRY(alloc_bodies(w, n));
    if (w->desc.max_pairs == 0 && w->cap_pairs < 8ull * n) AXM_TRY(alloc_pairs(w, 8 * n));
    if (w->desc.max_manifolds == 0 && w->cap_man < 4ull * n) AXM_TRY(alloc_manifolds(w, 4 * n));
    if (w->rcap[1 - w->cur] < w->cap_man) AXM_TRY(alloc_rows(w, 1 - w->cur, w->cap_man));
    hipStream_t s = w->stream;
    if (n) {
        AXM_HIP(hipMemcpyAsync(w->B.x, x.data(), 16ull * n, hipMemcpyHostToDevice, s));
        AXM_HIP(hipMemcpyAsync(w->B.q, q.data(), 16ull * n, hipMemcpyHostToDevice, s));
        AXM_HIP(hipMemcpyAsync(w->B.vel, v.data(), 16ull * n, hipMemcpyHostToDevice, s));
        AXM_HIP(hipMemcpyAsync(w->B.angv, av.data(), 16ull * n, hipMemcpyHostToDevice, s));
        AXM_HIP(hipMemcpyAsync(w->B.ii, ii.data(), 16ull * n, hipMemcpyHostToDevice, s));
        AXM_HIP(hipMemcpyAsync(w->B.shape, sh.data(), 16ull * n, hipMemcpyHostToDevice, s));
        AXM_HIP(hipMemcpyAsync(w->B.world, wd.data(), 4ull * n, hipMemcpyHostToDevice, s));
        AXM_HIP(hipMemcpyAsync(w->B.mat, mat.data(), 16ull * n, hipMemcpyHostToDevice, s));
        AXM_HIP(hipMemcpyAsync(w->B.bfl, bfl.data(), 16ull * n, hipMemcpyHostToDevice, s));
        AXM_HIP(hipMemcpyAsync(w->B.world_first, first.data(), 4ull * first.size(), hipMemcpyHostToDevice, s));
        if (hb.slp.size() == n && hb.wake.size() == n) {
            AXM_HIP(hipMemcpyAsync(w->B.slp, hb.slp.data(), 8ull * n, hipMemcpyHostToDevice, s));
            AXM_HIP(hipMemcpyAsync(w->B.wake, hb.wake.data(), 4ull * n, hipMemcpyHostToDevice, s));
        } else {
            // every body starts awake with no rest time and no island; no island has a wake request
            const std::vector<uint2> slp(n, make_uint2(0u, LABEL_NONE));
            AXM_HIP(hipMemcpyAsync(w->B.slp, slp.data(), 8ull * n, hipMemcpyHostToDevice, s));
            AXM_HIP(hipMemsetAsync(w->B.wake, 0xFF, 4ull * n, s));
        }
    }
    AXM_HIP(hipMemsetAsync(w->cnt, 0, sizeof(Counters), s));
    // the tables' entry counts are zero now: empty the tables in full
    for (int set = 0; set < 2; ++set)
        if (w->pc_key[set]) AXM_HIP(hipMemsetAsync(w->pc_key[set], 0xFF, 8ull * (w->pc_mask[set] + 1), s));
    // one large world: occupancy bits at the cells' Morton codes (cleared here in full,
    // then per step only the words the previous step set)
    const bool morton = nw <= 1u && n >= MORTON_MIN_BODIES;
    if (morton && !w->ct.mbits) AXM_TRY(w->mem.get(w->ct.mbits, MORTON_WORDS));
    if (morton && !w->ct.wrank) AXM_TRY(w->mem.get(w->ct.wrank, MORTON_WORDS));
    if (!morton && w->ct.mbits) w->mem.release(w->ct.mbits);
    if (!morton && w->ct.wrank) w->mem.release(w->ct.wrank);
    w->ct.mlist_n = morton ? n : 0u;
    if (morton) {
        AXM_HIP(hipMemsetAsync(w->ct.mbits, 0, 4ull * MORTON_WORDS, s));
        AXM_HIP(hipMemsetAsync(w->ct.mlist, 0xFF, 4ull * n, s));
    }
    AXM_HIP(hipStreamSynchronize(s));
    w->n = n;
    w->use_filter = use_filter;
    w->n_worlds = std::max(nw, 1u);
    w->cell = cell;
    w->prev_valid = false;
    w->have_step = false;
    w->have_pre = false;
    return AXM_SUCCESS;
}

axm_status axm_world_get_bodies(axm_world w, axm_body_soa* out) {
    if (!w || !out) return AXM_NULL_POINTER;
    (void)hipSetDevice(w->desc.device);
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
    (void)hipSetDevice(w->desc.device);
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

axm_status axm_world_step(axm_world w, float dt, uint32_t iters, uint32_t flags, axm_step_stats* stats) {
    if (!w) return AXM_NULL_POINTER;
    (void)hipSetDevice(w->desc.device);
    StepOpts o;
    o.iters = iters;
    o.flags = flags;
    return step_sync(w, dt, o, stats);
}

axm_status axm_world_run(axm_world w, float dt, uint32_t iters, uint32_t steps, int graph, axm_step_stats* last) {
    if (!w) return AXM_NULL_POINTER;
    if (!(dt > 0.0f) || !std::isfinite(dt)) return AXM_INVALID_PARAMETER;
    (void)hipSetDevice(w->desc.device);
    if (steps == 0) return AXM_SUCCESS;
    const bool use_graph = graph && !w->desc.record_residual && w->desc.solver_mode == 0 && w->n > 0;
    if (!use_graph) {
        StepOpts o;
        o.iters = iters;
        for (uint32_t k = 0; k + 1 < steps; ++k) AXM_TRY(step_sync(w, dt, o, nullptr));
        return step_sync(w, dt, o, last);
    }
    AXM_TRY(ensure_res(w, iters));
    // run totals at the start, for the strict status of the whole run
    AXM_TRY(read_counters(w));
    const uint32_t acc0_steps = w->h_cnt->acc_steps, acc0_reached = w->h_cnt->acc_reached;
    const uint32_t acc0_gjk = w->h_cnt->acc_gjk_fail, acc0_epa = w->h_cnt->acc_epa_fail;
    uint32_t done = 0;
    if (!w->prev_valid) {
        // the first step after new bodies has no warm-start cache: run it eagerly
        StepOpts o;
        o.iters = iters;
        AXM_TRY(step_sync(w, dt, o, steps == 1 ? last : nullptr));
        done = 1;
    }
    while (done < steps) {
        if (w->graph_dt != dt || w->graph_iters != iters) {
            w->drop_graphs();
            w->graph_dt = dt;
            w->graph_iters = iters;
        }
        for (int p = 0; p < 2; ++p) {
            if (w->prof_on) {
                if (!w->gpre[p]) AXM_TRY(capture(w, dt, iters, p, PH_PRE, &w->gpre[p]));
                if (!w->gpost[p]) AXM_TRY(capture(w, dt, iters, p, PH_POST, &w->gpost[p]));
            } else if (!w->gexec[p]) {
                AXM_TRY(capture(w, dt, iters, p, PH_ALL, &w->gexec[p]));
            }
        }
        const bool split = island_split(w) != 0;
        if (w->prof_on) {
            for (auto& e : w->gprof_ev)
                if (!e) AXM_HIP(hipEventCreate(&e));
            AXM_TRY(bundle_events(w, 0));
            AXM_TRY(bundle_events(w, 1));
        }
        const StepParams sp = make_params(w, dt);
        AXM_TRY(read_counters(w));
        const uint32_t base = w->h_cnt->steps_done;
        const unsigned long long base_bytes = w->h_cnt->pgs_bytes_total;
        const unsigned long long base_sweeps = w->h_cnt->sweeps_total;
        const unsigned long long base_bundle = w->h_cnt->bundle_point_sweeps_total;
        const int cur0 = w->cur;
        const uint32_t todo = steps - done;
        // With profiling, the sweep kernel of step k-1 is timed while step k
        // already runs: the host waits on the previous step only, so the GPU
        // queue never drains.
        double prof_ms = 0.0, bundle_ms = 0.0;
        int prev = -1;
        for (uint32_t k = 0; k < todo; ++k) {
            const int c = w->cur;
            if (w->prof_on) {
                AXM_HIP(hipGraphLaunch(w->gpre[c], w->stream));
                SweepStop stop;
                const bool checks = sweep_checks(w, false, stop);
                const uint32_t nsw = w->desc.residual_tol > 0.0f ? std::max<uint32_t>(w->desc.max_iterations, 1u)
                                                                 : iters;
                if (split)
                    launch_pgs_local(w->B, sp, w->R[c], w->cnt, w->rcap[c], w->solve_keys, w->isl_list,
                                     w->desc.residual_tol, nsw, 2u * (uint32_t)w->cu_count, w->stream,
                                     w->bprof_ev[2 * c], w->bprof_ev[2 * c + 1], w->bmask);
                launch_pgs_flow(w->B, sp, w->R[c], w->rcap[c], w->bmask, 0, nsw, w->cnt, SWEEP_SPIN_LIMIT_US,
                                w->sweep_groups, w->sweep_lds, w->sweep_tune, w->stream, w->gprof_ev[2 * c],
                                w->gprof_ev[2 * c + 1], checks ? stop : SweepStop{}, split);
                AXM_HIP(hipGraphLaunch(w->gpost[c], w->stream));
            } else {
                AXM_HIP(hipGraphLaunch(w->gexec[c], w->stream));
            }
            if (w->prof_on && prev >= 0) {
                float e = 0.0f;
                AXM_HIP(hipEventSynchronize(w->gprof_ev[2 * prev + 1]));
                AXM_HIP(hipEventElapsedTime(&e, w->gprof_ev[2 * prev], w->gprof_ev[2 * prev + 1]));
                prof_ms += e;
                if (split) {
                    AXM_HIP(hipEventElapsedTime(&e, w->bprof_ev[2 * prev], w->bprof_ev[2 * prev + 1]));
                    bundle_ms += e;
                }
            }
            prev = w->cur;
            w->cur = 1 - w->cur;
        }
        AXM_TRY(read_counters(w));
        if (w->prof_on && prev >= 0) {
            float e = 0.0f;
            AXM_HIP(hipEventElapsedTime(&e, w->gprof_ev[2 * prev], w->gprof_ev[2 * prev + 1]));
            prof_ms += e;
            if (split) {
                AXM_HIP(hipEventElapsedTime(&e, w->bprof_ev[2 * prev], w->bprof_ev[2 * prev + 1]));
                bundle_ms += e;
            }
        }
        const uint32_t ok = w->h_cnt->steps_done - base;
        if (w->prof_on) {
            w->prof.steps += todo;
            w->prof.pgs_launches += todo;
            w->prof.pgs_ms += prof_ms;
            w->prof.pgs_bytes += (double)(w->h_cnt->pgs_bytes_total - base_bytes);
            w->prof.sweeps += w->h_cnt->sweeps_total - base_sweeps;
            if (split) {
                w->prof.bundle_launches += todo;
                w->prof.bundle_ms += bundle_ms;
                w->prof.bundle_point_sweeps += w->h_cnt->bundle_point_sweeps_total - base_bundle;
            }
        }
        done += ok;
        w->have_step = true;
        // with a residual tolerance each step ran the sweeps its own checks chose
        w->last_iters = w->desc.residual_tol > 0.0f ? w->h_cnt->sweeps_run : iters;
        if (ok >= todo) break;
        // step `ok` of this batch failed and every later one skipped: repair and replay it eagerly
        w->cur = (ok & 1u) ? 1 - cur0 : cur0;
        w->cur = 1 - w->cur;  // recover() undoes one flip
        StepOpts o;
        o.iters = iters;
        AXM_TRY(recover(w, o));
        AXM_TRY(step_sync(w, dt, o, nullptr));
        done += 1;
    }
    if (last) {
        axm_step_stats S{};
        fill_stats(w, S);
        *last = S;
    }
    AXM_TRY(read_counters(w));
    const Counters& h = *w->h_cnt;
    const uint32_t ran = h.acc_steps - acc0_steps, reached = h.acc_reached - acc0_reached;
    return strict_status(w, h.acc_gjk_fail - acc0_gjk, h.acc_epa_fail - acc0_epa,
                         w->desc.residual_tol > 0.0f ? ran - reached : 0u);
}

axm_status axm_world_get_desc(axm_world w, axm_world_desc* out) {
    if (!w || !out) return AXM_NULL_POINTER;
    *out = w->desc;
    return AXM_SUCCESS;
}

axm_status axm_world_configure(axm_world w, const axm_world_desc* desc) {
    if (!w || !desc) return AXM_NULL_POINTER;
    if (desc->num_planes > AXM_MAX_PLANES || !(desc->softness >= 0.0f) || !(desc->friction >= 0.0f) ||
        !std::isfinite(desc->gravity[0]) || !std::isfinite(desc->gravity[1]) || !std::isfinite(desc->gravity[2]))
        return AXM_INVALID_PARAMETER;
    axm_world_desc d = *desc;
    // what sizes or places the buffers stays
    d.cell_size = w->desc.cell_size;
    for (int k = 0; k < 3; ++k) d.origin[k] = w->desc.origin[k];
    d.max_bodies = w->desc.max_bodies;
    d.max_pairs = w->desc.max_pairs;
    d.max_manifolds = w->desc.max_manifolds;
    d.auto_grow = w->desc.auto_grow;
    d.device = w->desc.device;
    w->desc = d;
    w->drop_graphs();
    return AXM_SUCCESS;
}

axm_status axm_world_get_islands(axm_world w, uint32_t* label, uint32_t* asleep, float* rest_time) {
    if (!w) return AXM_NULL_POINTER;
    (void)hipSetDevice(w->desc.device);
    const uint32_t n = w->n;
    if (!n) return AXM_SUCCESS;
    std::vector<uint2> h(n);
    AXM_HIP(hipMemcpyAsync(h.data(), w->B.slp, 8ull * n, hipMemcpyDeviceToHost, w->stream));
    AXM_HIP(hipStreamSynchronize(w->stream));
    for (uint32_t i = 0; i < n; ++i) {
        if (label) label[i] = h[i].y & LABEL_MASK;
        if (asleep) asleep[i] = (h[i].y & ASLEEP_BIT) ? 1u : 0u;
        if (rest_time) std::memcpy(rest_time + i, &h[i].x, 4);
    }
    return AXM_SUCCESS;
}

axm_status axm_world_get_totals(axm_world w, axm_world_totals* out, int reset) {
    if (!w || !out) return AXM_NULL_POINTER;
    (void)hipSetDevice(w->desc.device);
    AXM_TRY(read_counters(w));
    const Counters& h = *w->h_cnt;
    out->steps = h.acc_steps;
    out->steps_at_tol = h.acc_reached;
    out->sweeps = h.acc_sweeps;
    out->contact_sweeps = h.acc_contact_sweeps;
    out->max_sweeps = h.acc_max_sweeps;
    out->pad = 0;
    out->local_contact_sweeps = h.acc_local_contact_sweeps;
    if (reset) {
        // the accumulators sit between steps_done's block and the spread rows
        const size_t off = offsetof(Counters, acc_steps);
        const size_t len = offsetof(Counters, acc_epa_fail) + sizeof(uint32_t) - off;
        launch_fill_u32(reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(w->cnt) + off), 0u,
                        (uint32_t)(len / 4), w->stream);
        AXM_HIP(hipStreamSynchronize(w->stream));
    }
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
    (void)hipSetDevice(w->desc.device);
    AXM_HIP(hipStreamSynchronize(w->stream));
    return AXM_SUCCESS;
}

axm_status axm_broadphase(axm_world w, uint32_t* count) {
    if (!w) return AXM_NULL_POINTER;
    (void)hipSetDevice(w->desc.device);
    const StepParams sp = make_params(w, 1.0f / 120.0f);
    for (int attempt = 0; attempt < 3; ++attempt) {
        launch_counters_begin(w->cnt, w->pkinds, w->stream);
        if (w->n) enqueue_broad(w, sp);
        AXM_TRY(read_counters(w));
        if (!(w->h_cnt->sticky & STICKY_PAIRS)) break;
        if (!w->desc.auto_grow) return AXM_OUT_OF_RANGE;
        const uint64_t P = w->h_cnt->pairs;
        AXM_TRY(alloc_pairs(w, (uint32_t)std::min<uint64_t>(0xFFFFFFF0ull, P + P / 2 + 1024)));
        w->drop_graphs();
        AXM_TRY(clear_sticky(w));
    }
    AXM_TRY(clear_sticky(w));
    w->have_step = true;
    if (count) *count = w->h_cnt->pairs;
    return AXM_SUCCESS;
}

axm_status axm_collide(axm_world w, uint32_t* manifold_count) {
    if (!w) return AXM_NULL_POINTER;
    (void)hipSetDevice(w->desc.device);
    const StepParams sp = make_params(w, 1.0f / 120.0f);
    for (int attempt = 0; attempt < 4; ++attempt) {
        launch_counters_begin(w->cnt, w->pkinds, w->stream);
        if (w->n) {
            enqueue_broad(w, sp);
            enqueue_narrow(w, sp);
        }
        AXM_TRY(read_counters(w));
        const uint32_t st = w->h_cnt->sticky;
        if (!(st & (STICKY_PAIRS | STICKY_MANIFOLDS))) break;
        if (!w->desc.auto_grow) return AXM_OUT_OF_RANGE;
        if (st & STICKY_PAIRS) {
            const uint64_t P = w->h_cnt->pairs;
            AXM_TRY(alloc_pairs(w, (uint32_t)std::min<uint64_t>(0xFFFFFFF0ull, P + P / 2 + 1024)));
        }
        if (st & STICKY_MANIFOLDS) {
            const uint64_t M = w->h_cnt->manifolds;
            AXM_TRY(alloc_manifolds(w, (uint32_t)std::min<uint64_t>(0xFFFFFFF0ull, M + M / 2 + 1024)));
        }
        w->drop_graphs();
        AXM_TRY(clear_sticky(w));
    }
    AXM_TRY(clear_sticky(w));
    w->have_step = true;
    if (manifold_count) *manifold_count = w->h_cnt->manifolds;
    return AXM_SUCCESS;
}

axm_status axm_world_get_pairs(axm_world w, uint32_t* a, uint32_t* b, uint32_t cap, uint32_t* count) {
    if (!w || !count) return AXM_NULL_POINTER;
    if (!w->have_step) return AXM_GPU_INVALID_OPERATION;
    (void)hipSetDevice(w->desc.device);
    AXM_TRY(read_counters(w));
    const uint32_t P = w->h_cnt->p_eff;
    *count = P;
    if (cap < P) return AXM_OUT_OF_RANGE;
    if (P) {
        if (a) AXM_HIP(hipMemcpy(a, w->pa, 4ull * P, hipMemcpyDeviceToHost));
        if (b) AXM_HIP(hipMemcpy(b, w->pb, 4ull * P, hipMemcpyDeviceToHost));
    }
    return AXM_SUCCESS;
}

axm_status axm_world_get_manifolds(axm_world w, uint32_t* a, uint32_t* b, uint32_t* npts, float* pos, float* normal,
                                   float* depth, uint32_t* feature, uint32_t cap, uint32_t* count) {
    if (!w || !count) return AXM_NULL_POINTER;
    if (!w->have_step) return AXM_GPU_INVALID_OPERATION;
    (void)hipSetDevice(w->desc.device);
    AXM_TRY(read_counters(w));
    const uint32_t M = w->h_cnt->m_eff;
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
    (void)hipSetDevice(w->desc.device);
    AXM_TRY(read_counters(w));
    const uint32_t M = w->h_cnt->m_eff;
    *count = M;
    if (num_colours) *num_colours = w->h_cnt->max_colour;
    if (cap < M) return AXM_OUT_OF_RANGE;
    if (M) {
        if (colour) AXM_HIP(hipMemcpy(colour, w->col, 4ull * M, hipMemcpyDeviceToHost));
        if (solve_order && w->solve_order)
            AXM_HIP(hipMemcpy(solve_order, w->solve_order, 4ull * M, hipMemcpyDeviceToHost));
    }
    return AXM_SUCCESS;
}

axm_status axm_world_get_impulses(axm_world w, float* initial, float* final_, uint32_t cap, uint32_t* count) {
    if (!w || !count) return AXM_NULL_POINTER;
    if (!w->have_step) return AXM_GPU_INVALID_OPERATION;
    (void)hipSetDevice(w->desc.device);
    AXM_TRY(read_counters(w));
    const uint32_t M = w->h_cnt->m_eff;
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

axm_status axm_world_get_friction_bounds(axm_world w, float* bound, uint32_t cap, uint32_t* count) {
    if (!w || !count) return AXM_NULL_POINTER;
    if (!w->have_step) return AXM_GPU_INVALID_OPERATION;
    (void)hipSetDevice(w->desc.device);
    AXM_TRY(read_counters(w));
    const uint32_t M = w->h_cnt->m_eff;
    *count = M;
    if (cap < M) return AXM_OUT_OF_RANGE;
    if (!M || !bound) return AXM_SUCCESS;
    const int set = 1 - w->cur;  // rows of the last step
    const SolverRows& R = w->R[set];
    const uint32_t rc = w->rcap[set];
    std::vector<uint4> hdr(M);
    AXM_HIP(hipMemcpy(hdr.data(), R.hdr, 16ull * M, hipMemcpyDeviceToHost));
    std::vector<float4> row(M);
    for (uint32_t k = 0; k < 4; ++k) {
        // row 8 of point k (t1's inverse-inertia row) carries the bound in .w
        AXM_HIP(hipMemcpy(row.data(), R.J + ((size_t)k * JROWS + 8) * rc, 16ull * M, hipMemcpyDeviceToHost));
        for (uint32_t s = 0; s < M; ++s) bound[(size_t)hdr[s].w * 4 + k] = k < hdr[s].z ? row[s].w : 0.0f;
    }
    return AXM_SUCCESS;
}

axm_status axm_world_set_sweep_config(axm_world w, uint32_t groups, uint32_t lds_kib, uint32_t policy) {
    if (!w) return AXM_NULL_POINTER;
    if ((policy & ~0x76u) != 0u) return AXM_INVALID_PARAMETER;
    const uint32_t lds = lds_kib ? lds_kib * 1024u : (w->max_lds & ~16383u);
    if (lds > w->max_lds) return AXM_INVALID_PARAMETER;
    (void)hipSetDevice(w->desc.device);
    // every workgroup of the sweep must be resident at once (its items wait on each other)
    const int per_cu = pgs_own_setup(w->max_lds, lds);
    if (per_cu < 1) return AXM_INVALID_PARAMETER;
    const uint32_t resident = (uint32_t)per_cu * (uint32_t)w->cu_count;
    if (groups == 0) groups = std::min<uint32_t>((uint32_t)w->cu_count, resident);
    if (groups > resident) return AXM_INVALID_PARAMETER;
    w->sweep_groups = groups;
    w->sweep_lds = lds;
    w->sweep_tune = policy;
    w->drop_graphs();
    return AXM_SUCCESS;
}

axm_status axm_world_autotune(axm_world w, float dt, uint32_t iters, uint32_t steps, axm_sweep_config* out) {
    if (!w) return AXM_NULL_POINTER;
    if (steps == 0 || iters == 0) return AXM_INVALID_PARAMETER;
    (void)hipSetDevice(w->desc.device);
    // candidates: 1, 2 and 4 workgroups per CU with the LDS split between them (more
    // lanes, fewer slots each: better when a step has many manifolds per lane)
    struct Cand {
        uint32_t groups, lds_kib, policy;
    };
    std::vector<Cand> cands;
    const uint32_t max_kib = (w->max_lds & ~16383u) / 1024u;
    for (uint32_t per_cu : {1u, 2u, 4u}) {
        const uint32_t kib = (max_kib / per_cu) & ~15u;
        if (kib < 16u) continue;
        const int fit = pgs_own_setup(w->max_lds, kib * 1024u);
        if (fit < (int)per_cu) continue;
        cands.push_back({per_cu * (uint32_t)w->cu_count, kib, 0x12u});
    }
    if (cands.empty()) return AXM_GPU_OPERATION_FAILED;
    const bool was_on = w->prof_on;
    double best = 1e300;
    Cand pick = cands[0];
    for (const Cand& c : cands) {
        AXM_TRY(axm_world_set_sweep_config(w, c.groups, c.lds_kib, c.policy));
        AXM_TRY(axm_world_run(w, dt, iters, 1, 1, nullptr));  // capture outside the timing
        AXM_TRY(axm_world_profile(w, 1, nullptr));
        AXM_TRY(axm_world_run(w, dt, iters, steps, 1, nullptr));
        AXM_HIP(hipStreamSynchronize(w->stream));
        axm_kernel_profile p{};
        AXM_TRY(axm_world_profile(w, 0, &p));
        const double ms = p.pgs_launches ? p.pgs_ms / (double)p.pgs_launches : 1e300;
        if (ms < best) {
            best = ms;
            pick = c;
        }
    }
    AXM_TRY(axm_world_set_sweep_config(w, pick.groups, pick.lds_kib, pick.policy));
    w->prof_on = was_on;
    if (out) {
        out->groups = pick.groups;
        out->lds_kib = pick.lds_kib;
        out->policy = pick.policy;
        out->sweep_ms = best;
    }
    return AXM_SUCCESS;
}

axm_status axm_world_set_mlp(axm_world w, axm_mlp m) {
    if (!w) return AXM_NULL_POINTER;
    if (m && m->device != w->desc.device) return AXM_INVALID_PARAMETER;
    (void)hipSetDevice(w->desc.device);
    AXM_HIP(hipStreamSynchronize(w->stream));
    w->mlp = m;
    const bool want = w->want_feat || m != nullptr;
    if (want != w->want_feat || (want && !w->feat)) {
        w->want_feat = want;
        AXM_TRY(alloc_feat(w));
    }
    w->drop_graphs();
    return AXM_SUCCESS;
}

axm_status axm_world_set_feature_logging(axm_world w, int enable) {
    if (!w) return AXM_NULL_POINTER;
    (void)hipSetDevice(w->desc.device);
    AXM_HIP(hipStreamSynchronize(w->stream));
    const bool want = enable != 0 || w->mlp != nullptr;
    if (want != w->want_feat || (want && !w->feat)) {
        w->want_feat = want;
        AXM_TRY(alloc_feat(w));
        w->drop_graphs();
    }
    return AXM_SUCCESS;
}

axm_status axm_world_get_features(axm_world w, float* X, float* Y, uint32_t cap, uint32_t* count) {
    if (!w || !count) return AXM_NULL_POINTER;
    if (!w->have_step) return AXM_GPU_INVALID_OPERATION;
    if (!w->want_feat || !w->feat) return AXM_GPU_INVALID_OPERATION;
    (void)hipSetDevice(w->desc.device);
    AXM_TRY(read_counters(w));
    const uint32_t M = w->h_cnt->m_eff;
    *count = M;
    if (cap < M) return AXM_OUT_OF_RANGE;
    if (!M) return AXM_SUCCESS;
    const int set = 1 - w->cur;  // rows of the last step
    const SolverRows& R = w->R[set];
    const uint32_t rc = w->rcap[set];
    std::vector<uint4> hdr(M);
    AXM_HIP(hipMemcpy(hdr.data(), R.hdr, 16ull * M, hipMemcpyDeviceToHost));
    std::vector<float> f((size_t)M * 4 * MLP_IN);
    AXM_HIP(hipMemcpy(f.data(), w->feat, sizeof(float) * f.size(), hipMemcpyDeviceToHost));
    std::vector<float4> l(M), r0(M);
    const float gdt = MLP_G_REF * w->last_dt;
    for (uint32_t k = 0; k < 4; ++k) {
        AXM_HIP(hipMemcpy(l.data(), R.lam + (size_t)k * rc, 16ull * M, hipMemcpyDeviceToHost));
        AXM_HIP(hipMemcpy(r0.data(), R.J + (size_t)k * JROWS * rc, 16ull * M, hipMemcpyDeviceToHost));  // (n, mN)
        for (uint32_t s = 0; s < M; ++s) {
            const uint32_t m = hdr[s].w;
            const bool on = k < hdr[s].z;
            const size_t row = (size_t)m * 4 + k;
            if (X)
                for (int c = 0; c < MLP_IN; ++c) X[row * MLP_IN + c] = on ? f[((size_t)4 * s + k) * MLP_IN + c] : 0.0f;
            if (Y) {
                // target: (final - previous impulse) / u; the previous impulse / u is input 18..20
                const float u = r0[s].w * gdt;
                const float iu = (on && u > 0.0f) ? 1.0f / u : 0.0f;
                const float* fp = &f[((size_t)4 * s + k) * MLP_IN + 18];
                Y[row * 3] = on ? l[s].x * iu - fp[0] : 0.0f;
                Y[row * 3 + 1] = on ? l[s].y * iu - fp[1] : 0.0f;
                Y[row * 3 + 2] = on ? l[s].z * iu - fp[2] : 0.0f;
            }
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


// ---------------------------------------------------------------- world snapshots
// File layout (version 1): magic "AXMWORLD", SnapHeader, the world description, the
// per-body arrays in device layout (x, q, vel, angv, ii, shape, world, mat, bfl, sleep
// words, wake stamps, world_first), then, after a step, the warm-start cache of that
// step (impulses and feature ids of point k = 0..3 over its m_prev slots, slot headers,
// colours) and the FNV-1a hash.  Loading it and stepping on gives bit-identical results
// to stepping on from where it was saved: every input of the next step is restored (the
// per-step buffers are rebuilt by the step itself).
}  // extern "C"
namespace {
constexpr char SNAP_MAGIC[8] = {'A', 'X', 'M', 'W', 'O', 'R', 'L', 'D'};
constexpr uint32_t SNAP_VERSION = 1;
struct SnapHeader {
    uint32_t version, desc_bytes;
    uint32_t n, n_worlds, use_filter, prev_valid;
    float cell;
    uint32_t steps_done, m_prev, reserved;
};
template <class T> axm_status pull_vec(std::vector<T>& v, const T* src, size_t n, hipStream_t s) {
    v.resize(n);
    if (n) AXM_HIP(hipMemcpyAsync(v.data(), src, sizeof(T) * n, hipMemcpyDeviceToHost, s));
    return AXM_SUCCESS;
}
}  // namespace
extern "C" {

axm_status axm_world_save(axm_world w, const char* path) {
    if (!w || !path) return AXM_NULL_POINTER;
    (void)hipSetDevice(w->desc.device);
    AXM_HIP(hipStreamSynchronize(w->stream));
    AXM_TRY(read_counters(w));
    const uint32_t n = w->n, nw = w->n_worlds;
    const int pv = 1 - w->cur;
    const uint32_t m_prev = w->prev_valid ? std::min(w->h_cnt->m_eff, w->rcap[pv]) : 0u;
    hipStream_t s = w->stream;
    HostBodies hb;
    AXM_TRY(pull_vec(hb.x, w->B.x, n, s));
    AXM_TRY(pull_vec(hb.q, w->B.q, n, s));
    AXM_TRY(pull_vec(hb.v, w->B.vel, n, s));
    AXM_TRY(pull_vec(hb.av, w->B.angv, n, s));
    AXM_TRY(pull_vec(hb.ii, w->B.ii, n, s));
    AXM_TRY(pull_vec(hb.sh, w->B.shape, n, s));
    AXM_TRY(pull_vec(hb.wd, w->B.world, n, s));
    AXM_TRY(pull_vec(hb.mat, w->B.mat, n, s));
    AXM_TRY(pull_vec(hb.bfl, w->B.bfl, n, s));
    AXM_TRY(pull_vec(hb.slp, w->B.slp, n, s));
    AXM_TRY(pull_vec(hb.wake, w->B.wake, n, s));
    AXM_TRY(pull_vec(hb.first, w->B.world_first, n ? (size_t)nw + 1 : 0, s));
    const SolverRows& R = w->R[pv];
    const size_t rc = w->rcap[pv];
    std::vector<float4> lam[4];
    std::vector<uint32_t> fid[4];
    std::vector<uint4> hdr;
    std::vector<uint32_t> col;
    if (m_prev) {
        for (int k = 0; k < 4; ++k) {
            AXM_TRY(pull_vec(lam[k], R.lam + (size_t)k * rc, m_prev, s));
            AXM_TRY(pull_vec(fid[k], R.fid + (size_t)k * rc, m_prev, s));
        }
        AXM_TRY(pull_vec(hdr, R.hdr, m_prev, s));
        AXM_TRY(pull_vec(col, R.colour, m_prev, s));
    }
    AXM_HIP(hipStreamSynchronize(s));
    BinWriter f(path);
    if (!f.ok()) return AXM_INVALID_PARAMETER;
    SnapHeader h{SNAP_VERSION, (uint32_t)sizeof(axm_world_desc), n, nw, w->use_filter ? 1u : 0u,
                 m_prev ? 1u : 0u, w->cell, w->h_cnt->steps_done, m_prev, 0u};
    f.put(SNAP_MAGIC, 8);
    f.pod(h);
    f.pod(w->desc);
    for (const auto* v : {&hb.x, &hb.q, &hb.v, &hb.av, &hb.ii, &hb.sh}) f.vec(*v);
    f.vec(hb.wd);
    f.vec(hb.mat);
    f.vec(hb.bfl);
    f.vec(hb.slp);
    f.vec(hb.wake);
    f.vec(hb.first);
    for (int k = 0; k < 4 && m_prev; ++k) {
        f.vec(lam[k]);
        f.vec(fid[k]);
    }
    f.vec(hdr);
    f.vec(col);
    return f.finish() ? AXM_SUCCESS : AXM_INVALID_PARAMETER;
}

axm_status axm_world_load(axm_world w, const char* path) {
    if (!w || !path) return AXM_NULL_POINTER;
    (void)hipSetDevice(w->desc.device);
    BinReader f(path);
    if (!f.ok()) return AXM_INVALID_PARAMETER;
    char magic[8];
    f.get(magic, 8);
    SnapHeader h{};
    f.pod(h);
    if (!f.ok() || std::memcmp(magic, SNAP_MAGIC, 8) != 0 || h.version != SNAP_VERSION ||
        h.desc_bytes != sizeof(axm_world_desc) || h.n >= (1u << 29) || h.n_worlds > (1u << 22) ||
        (h.n && h.n_worlds == 0) || h.m_prev > 4u * (h.n + 16u))
        return AXM_INVALID_PARAMETER;
    axm_world_desc d{};
    f.pod(d);
    const uint32_t n = h.n;
    HostBodies hb;
    for (auto* v : {&hb.x, &hb.q, &hb.v, &hb.av, &hb.ii, &hb.sh}) f.vec(*v, n);
    f.vec(hb.wd, n);
    f.vec(hb.mat, n);
    f.vec(hb.bfl, n);
    f.vec(hb.slp, n);
    f.vec(hb.wake, n);
    f.vec(hb.first, n ? (size_t)h.n_worlds + 1 : 0);
    const uint32_t mp = h.m_prev;
    std::vector<float4> lam[4];
    std::vector<uint32_t> fid[4];
    std::vector<uint4> hdr;
    std::vector<uint32_t> col;
    for (int k = 0; k < 4 && mp; ++k) {
        f.vec(lam[k], mp);
        f.vec(fid[k], mp);
    }
    f.vec(hdr, mp);
    f.vec(col, mp);
    if (!f.finish()) return AXM_INVALID_PARAMETER;
    // the world keeps its device; everything else comes from the file
    d.device = w->desc.device;
    AXM_HIP(hipStreamSynchronize(w->stream));
    // the captured steps stay valid when the file holds what they were captured with
    const bool same = std::memcmp(&d, &w->desc, sizeof(d)) == 0 && n == w->n &&
                      std::max(h.n_worlds, 1u) == w->n_worlds && h.cell == w->cell &&
                      (h.use_filter != 0) == w->use_filter;
    w->desc = d;
    AXM_TRY(install_bodies(w, n, std::max(h.n_worlds, 1u), h.cell, h.use_filter != 0, hb, same));
    hipStream_t s = w->stream;
    AXM_HIP(hipMemcpy(&w->cnt->steps_done, &h.steps_done, 4, hipMemcpyHostToDevice));
    if (mp) {
        // the previous step's cache in the set the next step reads, and its lookup table
        const int pv = 1 - w->cur;
        if (w->cap_man < mp || w->rcap[pv] < std::max(w->cap_man, mp)) w->drop_graphs();
        if (w->cap_man < mp) AXM_TRY(alloc_manifolds(w, mp));
        if (w->rcap[pv] < w->cap_man) AXM_TRY(alloc_rows(w, pv, w->cap_man));
        const SolverRows& R = w->R[pv];
        const size_t rc = w->rcap[pv];
        for (int k = 0; k < 4; ++k) {
            AXM_HIP(hipMemcpyAsync(R.lam + (size_t)k * rc, lam[k].data(), 16ull * mp, hipMemcpyHostToDevice, s));
            AXM_HIP(hipMemcpyAsync(R.fid + (size_t)k * rc, fid[k].data(), 4ull * mp, hipMemcpyHostToDevice, s));
        }
        AXM_HIP(hipMemcpyAsync(R.hdr, hdr.data(), 16ull * mp, hipMemcpyHostToDevice, s));
        AXM_HIP(hipMemcpyAsync(R.colour, col.data(), 4ull * mp, hipMemcpyHostToDevice, s));
        AXM_HIP(hipMemcpyAsync(&w->cnt->m_eff, &mp, 4, hipMemcpyHostToDevice, s));
        launch_cache_build(R, w->cnt, w->rcap[pv], w->pc_key[pv], w->pc_slot[pv], w->pc_mask[pv], w->np_slot[pv],
                           w->pc_pos[pv], (uint32_t)pv, s);
        AXM_HIP(hipStreamSynchronize(s));
        w->prev_valid = true;
    }
    return AXM_SUCCESS;
}

}  // extern "C"
