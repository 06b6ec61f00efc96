// Batched pairwise geometry entry points: narrowphase on arbitrary shape pairs
// and GJK / EPA on shape cores, one thread per pair.  These expose the
// collision routines the world step uses (dev_geom.hpp) on caller-chosen
// configurations.
#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "axiom/axm.h"
#include "dev_geom.hpp"

using namespace axm;

// ---------------------------------------------------------------- allocation-free staging
// Staging of axm_collide_pairs_noalloc: the inputs, outputs and per-pair status of at
// most NOALLOC_MAX_PAIRS pairs, one region of 32-bit words per field (each region holds
// that field for every pair slot: NOALLOC_MAX_PAIRS x width words).  One copy per
// device, held until the process ends; every call on a device shares it.
constexpr uint32_t NOALLOC_MAX_PAIRS = 16384;
// region offsets, in words per pair slot: ta 1, pa 3, xa 3, qa 4, tb 1, pb 3, xb 3, qb 4 (inputs),
// npts 1, pos 12, normal 12, depth 4, feature 4, status 1 (outputs)
constexpr uint32_t NA_TA = 0, NA_PA = 1, NA_XA = 4, NA_QA = 7, NA_TB = 11, NA_PB = 12, NA_XB = 15, NA_QB = 18,
                   NA_NP = 22, NA_POS = 23, NA_NRM = 35, NA_DEP = 47, NA_FID = 51, NA_ST = 55, NA_WORDS = 56;
static __device__ __align__(256) uint32_t scratch[(size_t)NOALLOC_MAX_PAIRS * NA_WORDS];

namespace {

struct PairIn {
    const uint32_t* ta;
    const float* pa;
    const float* xa;
    const float* qa;
    const uint32_t* tb;
    const float* pb;
    const float* xb;
    const float* qb;
};

__device__ Shp shape_at(const uint32_t* t, const float* p, const float* x, const float* q, uint32_t i) {
    Shp s;
    s.type = t[i];
    s.prm = f3{p[3 * i], p[3 * i + 1], p[3 * i + 2]};
    s.c = f3{x[3 * i], x[3 * i + 1], x[3 * i + 2]};
    s.R = quat_to_mat(q4{q[4 * i], q[4 * i + 1], q[4 * i + 2], q[4 * i + 3]});
    return s;
}

__global__ void __launch_bounds__(64) k_collide_pairs(PairIn in, uint32_t n, float margin, uint32_t* npts,
                                                      float* pos, float* nrm, float* depth, uint32_t* fid,
                                                      int32_t* status) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Shp A = shape_at(in.ta, in.pa, in.xa, in.qa, i);
    Man m;
    m.np = 0;
    int st = GEOM_OK;
    if (in.tb[i] == SHAPE_PLANE) {
        // plane given as params = normal, pos.x = offset d
        collide_plane(m, A, f3{in.pb[3 * i], in.pb[3 * i + 1], in.pb[3 * i + 2]}, in.xb[3 * i], margin);
    } else {
        const Shp B = shape_at(in.tb, in.pb, in.xb, in.qb, i);
        const bool swap = A.type > B.type;
        bool need_epa = false;
        st = swap ? collide_canon(m, B, A, margin, need_epa) : collide_canon(m, A, B, margin, need_epa);
        if (need_epa) {
            EpaWork W;
            m.np = 0;
            st = swap ? box_capsule_epa(m, B, A, W) : box_capsule_epa(m, A, B, W);
        }
        if (swap) flip_manifold(m);
    }
    npts[i] = m.np;
    status[i] = st;
    for (uint32_t k = 0; k < 4; ++k) {
        const bool on = k < m.np;
        const size_t s = (size_t)i * 4 + k;
        pos[3 * s] = on ? m.pt[k].p.x : 0.f;
        pos[3 * s + 1] = on ? m.pt[k].p.y : 0.f;
        pos[3 * s + 2] = on ? m.pt[k].p.z : 0.f;
        nrm[3 * s] = on ? m.pt[k].n.x : 0.f;
        nrm[3 * s + 1] = on ? m.pt[k].n.y : 0.f;
        nrm[3 * s + 2] = on ? m.pt[k].n.z : 0.f;
        depth[s] = on ? m.pt[k].depth : 0.f;
        fid[s] = on ? m.pt[k].fid : 0u;
    }
}

__global__ void __launch_bounds__(64) k_gjk_epa_pairs(PairIn in, uint32_t n, float* out, int32_t* status,
                                                       int gjk_iters, int epa_iters) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Shp A = shape_at(in.ta, in.pa, in.xa, in.qa, i);
    const Shp B = shape_at(in.tb, in.pb, in.xb, in.qb, i);
    const GjkOut g = gjk(A, B, gjk_iters);
    float* o = out + (size_t)i * 10;
    if (g.status == 1) {
        EpaWork W;
        const EpaOut e = epa(A, B, g, W, epa_iters);
        o[0] = e.depth;
        o[1] = e.pA.x; o[2] = e.pA.y; o[3] = e.pA.z;
        o[4] = e.pB.x; o[5] = e.pB.y; o[6] = e.pB.z;
        o[7] = e.n.x; o[8] = e.n.y; o[9] = e.n.z;
        status[i] = e.status == 0 ? 1 : GEOM_EPA_FAIL;
        return;
    }
    o[0] = g.dist;
    o[1] = g.pA.x; o[2] = g.pA.y; o[3] = g.pA.z;
    o[4] = g.pB.x; o[5] = g.pB.y; o[6] = g.pB.z;
    o[7] = o[8] = o[9] = 0.f;
    status[i] = g.status == 0 ? 0 : GEOM_GJK_FAIL;
}

// Small device arena for one batched call.
struct Arena {
    std::vector<void*> p;
    ~Arena() {
        for (void* q : p) (void)hipFree(q);
    }
    template <class T> T* up(const T* h, size_t n) {
        void* d = nullptr;
        if (hipMalloc(&d, n * sizeof(T) + 16) != hipSuccess) return nullptr;
        p.push_back(d);
        if (h && hipMemcpy(d, h, n * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) return nullptr;
        return (T*)d;
    }
};

bool upload_pairs(Arena& ar, uint32_t n, const uint32_t* ta, const float* pa, const float* xa, const float* qa,
                  const uint32_t* tb, const float* pb, const float* xb, const float* qb, PairIn& in) {
    in.ta = ar.up(ta, n);
    in.pa = ar.up(pa, 3ull * n);
    in.xa = ar.up(xa, 3ull * n);
    in.qa = ar.up(qa, 4ull * n);
    in.tb = ar.up(tb, n);
    in.pb = ar.up(pb, 3ull * n);
    in.xb = ar.up(xb, 3ull * n);
    in.qb = ar.up(qb, 4ull * n);
    return in.ta && in.pa && in.xa && in.qa && in.tb && in.pb && in.xb && in.qb;
}

// per device: the address of its copy of `scratch`, set on first use
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
uint32_t* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// The device address of `scratch` on `device`, looked up once per device with that device
// current (each device holds its own copy at its own address; the host-side name of the
// array is not a device address and is never passed to the device).
uint32_t* scratch_on(int32_t device) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return nullptr;
        g_scratch_ptr[device] = static_cast<uint32_t*>(p);
    }
    return g_scratch_ptr[device];
}

bool valid_types(uint32_t n, const uint32_t* ta, const uint32_t* tb, bool allow_plane_b) {
    for (uint32_t i = 0; i < n; ++i) {
        if (ta[i] > AXM_SHAPE_CAPSULE) return false;
        if (tb[i] > (allow_plane_b ? (uint32_t)AXM_SHAPE_PLANE : (uint32_t)AXM_SHAPE_CAPSULE)) return false;
    }
    return true;
}

}  // namespace

extern "C" {

axm_status axm_collide_pairs(int32_t device, uint32_t n, const uint32_t* type_a, const float* prm_a,
                             const float* pos_a, const float* quat_a, const uint32_t* type_b, const float* prm_b,
                             const float* pos_b, const float* quat_b, float margin, uint32_t* npts, float* pos,
                             float* normal, float* depth, uint32_t* feature) {
    if (!n) return AXM_SUCCESS;
    if (!type_a || !prm_a || !pos_a || !quat_a || !type_b || !prm_b || !pos_b || !quat_b || !npts || !pos || !normal ||
        !depth || !feature)
        return AXM_NULL_POINTER;
    if (!valid_types(n, type_a, type_b, true)) return AXM_INVALID_SHAPE;
    if (hipSetDevice(device) != hipSuccess) return AXM_GPU_OPERATION_FAILED;
    Arena ar;
    PairIn in;
    if (!upload_pairs(ar, n, type_a, prm_a, pos_a, quat_a, type_b, prm_b, pos_b, quat_b, in))
        return AXM_BUFFER_ALLOCATION_FAILED;
    uint32_t* d_np = ar.up<uint32_t>(nullptr, n);
    float* d_pos = ar.up<float>(nullptr, 12ull * n);
    float* d_n = ar.up<float>(nullptr, 12ull * n);
    float* d_d = ar.up<float>(nullptr, 4ull * n);
    uint32_t* d_f = ar.up<uint32_t>(nullptr, 4ull * n);
    int32_t* d_s = ar.up<int32_t>(nullptr, n);
    if (!d_np || !d_pos || !d_n || !d_d || !d_f || !d_s) return AXM_BUFFER_ALLOCATION_FAILED;
    hipLaunchKernelGGL(k_collide_pairs, dim3((n + 63) / 64), dim3(64), 0, 0, in, n, margin, d_np, d_pos, d_n, d_d,
                       d_f, d_s);
    if (hipDeviceSynchronize() != hipSuccess) return AXM_GPU_OPERATION_FAILED;
    std::vector<int32_t> st(n);
    if (hipMemcpy(npts, d_np, 4ull * n, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(pos, d_pos, 48ull * n, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(normal, d_n, 48ull * n, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(depth, d_d, 16ull * n, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(feature, d_f, 16ull * n, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(st.data(), d_s, 4ull * n, hipMemcpyDeviceToHost) != hipSuccess)
        return AXM_GPU_OPERATION_FAILED;
    for (uint32_t i = 0; i < n; ++i)
        if (st[i] != GEOM_OK) return st[i];
    return AXM_SUCCESS;
}

axm_status axm_collide_pairs_noalloc(int32_t device, uint32_t n, const uint32_t* type_a, const float* prm_a,
                                     const float* pos_a, const float* quat_a, const uint32_t* type_b,
                                     const float* prm_b, const float* pos_b, const float* quat_b, float margin,
                                     uint32_t* npts, float* pos, float* normal, float* depth, uint32_t* feature) {
    if (!n) return AXM_SUCCESS;
    if (!type_a || !prm_a || !pos_a || !quat_a || !type_b || !prm_b || !pos_b || !quat_b || !npts || !pos || !normal ||
        !depth || !feature)
        return AXM_NULL_POINTER;
    if (n > NOALLOC_MAX_PAIRS) return AXM_OUT_OF_RANGE;
    if (!valid_types(n, type_a, type_b, true)) return AXM_INVALID_SHAPE;
    if (hipSetDevice(device) != hipSuccess) return AXM_GPU_OPERATION_FAILED;
    uint32_t* base = scratch_on(device);
    if (!base) return AXM_GPU_OPERATION_FAILED;
    // region r starts at base + r * NOALLOC_MAX_PAIRS
    auto reg = [&](uint32_t r) { return base + (size_t)r * NOALLOC_MAX_PAIRS; };
    struct Up {
        uint32_t r;
        const void* h;
        size_t bytes;
    };
    const Up ups[] = {{NA_TA, type_a, 4ull * n}, {NA_PA, prm_a, 12ull * n}, {NA_XA, pos_a, 12ull * n},
                      {NA_QA, quat_a, 16ull * n}, {NA_TB, type_b, 4ull * n}, {NA_PB, prm_b, 12ull * n},
                      {NA_XB, pos_b, 12ull * n}, {NA_QB, quat_b, 16ull * n}};
    for (const Up& u : ups)
        if (hipMemcpy(reg(u.r), u.h, u.bytes, hipMemcpyHostToDevice) != hipSuccess) return AXM_GPU_OPERATION_FAILED;
    PairIn in;
    in.ta = reg(NA_TA);
    in.pa = reinterpret_cast<const float*>(reg(NA_PA));
    in.xa = reinterpret_cast<const float*>(reg(NA_XA));
    in.qa = reinterpret_cast<const float*>(reg(NA_QA));
    in.tb = reg(NA_TB);
    in.pb = reinterpret_cast<const float*>(reg(NA_PB));
    in.xb = reinterpret_cast<const float*>(reg(NA_XB));
    in.qb = reinterpret_cast<const float*>(reg(NA_QB));
    // the kernel writes every output word of the n pairs (zeros past a pair's point count),
    // so nothing left by an earlier call is read back
    hipLaunchKernelGGL(k_collide_pairs, dim3((n + 63) / 64), dim3(64), 0, 0, in, n, margin, reg(NA_NP),
                       reinterpret_cast<float*>(reg(NA_POS)), reinterpret_cast<float*>(reg(NA_NRM)),
                       reinterpret_cast<float*>(reg(NA_DEP)), reg(NA_FID), reinterpret_cast<int32_t*>(reg(NA_ST)));
    if (hipDeviceSynchronize() != hipSuccess) return AXM_GPU_OPERATION_FAILED;
    if (hipMemcpy(npts, reg(NA_NP), 4ull * n, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(pos, reg(NA_POS), 48ull * n, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(normal, reg(NA_NRM), 48ull * n, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(depth, reg(NA_DEP), 16ull * n, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(feature, reg(NA_FID), 16ull * n, hipMemcpyDeviceToHost) != hipSuccess)
        return AXM_GPU_OPERATION_FAILED;
    // the first failing pair's status, as axm_collide_pairs reports it; read 4096 at a time
    // into a fixed host buffer
    int32_t st[4096];
    for (uint32_t i0 = 0; i0 < n; i0 += 4096) {
        const uint32_t c = std::min<uint32_t>(4096u, n - i0);
        if (hipMemcpy(st, reg(NA_ST) + i0, 4ull * c, hipMemcpyDeviceToHost) != hipSuccess)
            return AXM_GPU_OPERATION_FAILED;
        for (uint32_t i = 0; i < c; ++i)
            if (st[i] != GEOM_OK) return st[i];
    }
    return AXM_SUCCESS;
}

axm_status axm_gjk_epa_pairs(int32_t device, uint32_t n, const uint32_t* type_a, const float* prm_a,
                             const float* pos_a, const float* quat_a, const uint32_t* type_b, const float* prm_b,
                             const float* pos_b, const float* quat_b, float* out, int32_t* status) {
    return axm_gjk_epa_pairs_capped(device, n, type_a, prm_a, pos_a, quat_a, type_b, prm_b, pos_b, quat_b, 0u, 0u, out,
                                    status);
}

axm_status axm_gjk_epa_pairs_capped(int32_t device, uint32_t n, const uint32_t* type_a, const float* prm_a,
                                    const float* pos_a, const float* quat_a, const uint32_t* type_b,
                                    const float* prm_b, const float* pos_b, const float* quat_b,
                                    uint32_t gjk_max_iters, uint32_t epa_max_iters, float* out, int32_t* status) {
    if (gjk_max_iters > 4096u || epa_max_iters > 4096u) return AXM_INVALID_PARAMETER;
    const int gi = gjk_max_iters ? (int)gjk_max_iters : 32, ei = epa_max_iters ? (int)epa_max_iters : 48;
    if (!n) return AXM_SUCCESS;
    if (!type_a || !prm_a || !pos_a || !quat_a || !type_b || !prm_b || !pos_b || !quat_b || !out || !status)
        return AXM_NULL_POINTER;
    if (!valid_types(n, type_a, type_b, false)) return AXM_INVALID_SHAPE;
    if (hipSetDevice(device) != hipSuccess) return AXM_GPU_OPERATION_FAILED;
    Arena ar;
    PairIn in;
    if (!upload_pairs(ar, n, type_a, prm_a, pos_a, quat_a, type_b, prm_b, pos_b, quat_b, in))
        return AXM_BUFFER_ALLOCATION_FAILED;
    float* d_o = ar.up<float>(nullptr, 10ull * n);
    int32_t* d_s = ar.up<int32_t>(nullptr, n);
    if (!d_o || !d_s) return AXM_BUFFER_ALLOCATION_FAILED;
    hipLaunchKernelGGL(k_gjk_epa_pairs, dim3((n + 63) / 64), dim3(64), 0, 0, in, n, d_o, d_s, gi, ei);
    if (hipDeviceSynchronize() != hipSuccess) return AXM_GPU_OPERATION_FAILED;
    if (hipMemcpy(out, d_o, 40ull * n, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(status, d_s, 4ull * n, hipMemcpyDeviceToHost) != hipSuccess)
        return AXM_GPU_OPERATION_FAILED;
    return AXM_SUCCESS;
}

}  // extern "C"
