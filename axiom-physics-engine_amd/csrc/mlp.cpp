// C ABI of the learned warm-start network (include/axiom/axm.h, axm_mlp_*):
// parameters, forward pass, loss gradient, Adam and a minibatch training loop,
// all on the GPU through the kernels of k_mlp.hip.  The reference has no
// learned component; the network and its training are this project's.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "file_io.hpp"
#include "mlp_impl.hpp"

using namespace axm;

namespace {

#define MLP_HIP(expr)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess) {                                                            \
            std::fprintf(stderr, "axm: %s failed: %s\n", #expr, hipGetErrorString(e_));  \
            return AXM_GPU_OPERATION_FAILED;                                               \
        }                                                                                  \
    } while (0)

constexpr uint32_t GRAD_BLOCKS = 256;  // 64 rows each per pass (grid-stride beyond)

template <class T> axm_status dalloc(T*& p, size_t count) {
    if (hipMalloc((void**)&p, count * sizeof(T)) != hipSuccess) {
        p = nullptr;
        return AXM_BUFFER_ALLOCATION_FAILED;
    }
    return AXM_SUCCESS;
}
template <class T> void dfree(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

axm_status ensure_rows(axm_mlp_s* m, uint32_t n) {
    if (m->cap_rows >= n) return AXM_SUCCESS;
    MLP_HIP(hipStreamSynchronize(m->stream));
    dfree(m->X);
    dfree(m->Y);
    uint32_t cap = m->cap_rows ? m->cap_rows : 4096u;
    while (cap < n) cap = cap > 0x7FFFFFFFu / 2 ? n : cap * 2;
    if (dalloc(m->X, (size_t)cap * MLP_IN) != AXM_SUCCESS || dalloc(m->Y, (size_t)cap * MLP_OUT) != AXM_SUCCESS) {
        m->cap_rows = 0;
        return AXM_BUFFER_ALLOCATION_FAILED;
    }
    m->cap_rows = cap;
    return AXM_SUCCESS;
}

// splitmix64: initialisation stream of the network weights
uint64_t splitmix(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

axm_status upload_params(axm_mlp_s* m, const float* p) {
    MLP_HIP(hipMemcpyAsync(m->prm, p, sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice, m->stream));
    launch_mlp_cast(m->prm, m->w1, m->w2, m->w3, m->stream);
    MLP_HIP(hipStreamSynchronize(m->stream));
    return AXM_SUCCESS;
}

}  // namespace

extern "C" {

axm_status axm_mlp_create(int32_t device, axm_mlp* out) {
    if (!out) return AXM_NULL_POINTER;
    *out = nullptr;
    int nd = 0;
    if (hipGetDeviceCount(&nd) != hipSuccess || nd <= 0) return AXM_GPU_OPERATION_FAILED;
    if (device < 0 || device >= nd) return AXM_INVALID_PARAMETER;
    if (hipSetDevice(device) != hipSuccess) return AXM_GPU_OPERATION_FAILED;
    if (mlp_prepare_device() != hipSuccess) return AXM_GPU_OPERATION_FAILED;
    axm_mlp_s* m = new (std::nothrow) axm_mlp_s;
    if (!m) return AXM_OUT_OF_MEMORY;
    m->device = device;
    axm_status st = AXM_SUCCESS;
    if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess) st = AXM_GPU_OPERATION_FAILED;
    if (st == AXM_SUCCESS) st = dalloc(m->prm, MLP_NPARAM);
    if (st == AXM_SUCCESS) st = dalloc(m->m, MLP_NPARAM);
    if (st == AXM_SUCCESS) st = dalloc(m->v, MLP_NPARAM);
    if (st == AXM_SUCCESS) st = dalloc(m->grad, MLP_NPARAM);
    if (st == AXM_SUCCESS) st = dalloc(m->w1, (size_t)MLP_H * MLP_INP);
    if (st == AXM_SUCCESS) st = dalloc(m->w2, (size_t)MLP_H * MLP_H);
    if (st == AXM_SUCCESS) st = dalloc(m->w3, (size_t)MLP_OUTP * MLP_H);
    if (st == AXM_SUCCESS) st = dalloc(m->partial, mlp_grad_partial_floats(GRAD_BLOCKS));
    if (st == AXM_SUCCESS) st = dalloc(m->loss_partial, GRAD_BLOCKS);
    if (st == AXM_SUCCESS) st = dalloc(m->loss, 1);
    if (st == AXM_SUCCESS) {
        m->blocks = GRAD_BLOCKS;
        std::vector<float> zero(MLP_NPARAM, 0.0f);
        if (hipMemcpy(m->m, zero.data(), sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(m->v, zero.data(), sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice) != hipSuccess)
            st = AXM_GPU_OPERATION_FAILED;
        else
            st = upload_params(m, zero.data());
    }
    if (st != AXM_SUCCESS) {
        axm_mlp_destroy(m);
        return st;
    }
    *out = m;
    return AXM_SUCCESS;
}

axm_status axm_mlp_destroy(axm_mlp m) {
    if (!m) return AXM_NULL_POINTER;
    (void)hipSetDevice(m->device);
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    dfree(m->prm); dfree(m->m); dfree(m->v); dfree(m->grad);
    dfree(m->w1); dfree(m->w2); dfree(m->w3);
    dfree(m->partial); dfree(m->loss_partial); dfree(m->loss);
    dfree(m->X); dfree(m->Y);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
    return AXM_SUCCESS;
}

axm_status axm_mlp_init(axm_mlp m, uint64_t seed) {
    if (!m) return AXM_NULL_POINTER;
    (void)hipSetDevice(m->device);
    // He-uniform weights (limit sqrt(6 / fan_in)), zero biases
    std::vector<float> p(MLP_NPARAM, 0.0f);
    uint64_t s = seed;
    auto fill = [&](int off, int count, int fan_in) {
        const double lim = std::sqrt(6.0 / fan_in);
        for (int i = 0; i < count; ++i) {
            const double u = (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0);
            p[off + i] = (float)((2.0 * u - 1.0) * lim);
        }
    };
    fill(MLP_OFF_W1, MLP_H * MLP_IN, MLP_IN);
    fill(MLP_OFF_W2, MLP_H * MLP_H, MLP_H);
    fill(MLP_OFF_W3, MLP_OUT * MLP_H, MLP_H);
    m->adam_t = 0;
    std::vector<float> zero(MLP_NPARAM, 0.0f);
    MLP_HIP(hipMemcpy(m->m, zero.data(), sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice));
    MLP_HIP(hipMemcpy(m->v, zero.data(), sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice));
    return upload_params(m, p.data());
}

axm_status axm_mlp_set_params(axm_mlp m, const float* params) {
    if (!m || !params) return AXM_NULL_POINTER;
    for (int i = 0; i < MLP_NPARAM; ++i)
        if (!std::isfinite(params[i])) return AXM_INVALID_PARAMETER;
    (void)hipSetDevice(m->device);
    return upload_params(m, params);
}

axm_status axm_mlp_get_params(axm_mlp m, float* params) {
    if (!m || !params) return AXM_NULL_POINTER;
    (void)hipSetDevice(m->device);
    MLP_HIP(hipMemcpyAsync(params, m->prm, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost, m->stream));
    MLP_HIP(hipStreamSynchronize(m->stream));
    return AXM_SUCCESS;
}

axm_status axm_mlp_forward(axm_mlp m, const float* X, uint32_t n, float* Y) {
    if (!m || (n && (!X || !Y))) return AXM_NULL_POINTER;
    if (!n) return AXM_SUCCESS;
    (void)hipSetDevice(m->device);
    axm_status st = ensure_rows(m, n);
    if (st != AXM_SUCCESS) return st;
    MLP_HIP(hipMemcpyAsync(m->X, X, sizeof(float) * MLP_IN * (size_t)n, hipMemcpyHostToDevice, m->stream));
    launch_mlp_forward(m->dev(), m->X, n, nullptr, 1, m->Y, 1024, m->stream);
    MLP_HIP(hipGetLastError());
    MLP_HIP(hipMemcpyAsync(Y, m->Y, sizeof(float) * MLP_OUT * (size_t)n, hipMemcpyDeviceToHost, m->stream));
    MLP_HIP(hipStreamSynchronize(m->stream));
    return AXM_SUCCESS;
}

axm_status axm_mlp_loss_grad(axm_mlp m, const float* X, const float* Y, uint32_t n, float* grad, double* loss) {
    if (!m || !X || !Y) return AXM_NULL_POINTER;
    if (!n) return AXM_INVALID_PARAMETER;
    (void)hipSetDevice(m->device);
    axm_status st = ensure_rows(m, n);
    if (st != AXM_SUCCESS) return st;
    MLP_HIP(hipMemcpyAsync(m->X, X, sizeof(float) * MLP_IN * (size_t)n, hipMemcpyHostToDevice, m->stream));
    MLP_HIP(hipMemcpyAsync(m->Y, Y, sizeof(float) * MLP_OUT * (size_t)n, hipMemcpyHostToDevice, m->stream));
    launch_mlp_grad(m->dev(), m->X, m->Y, n, m->partial, m->loss_partial, m->grad, m->loss, m->blocks, m->stream);
    MLP_HIP(hipGetLastError());
    if (grad) MLP_HIP(hipMemcpyAsync(grad, m->grad, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost, m->stream));
    if (loss) MLP_HIP(hipMemcpyAsync(loss, m->loss, sizeof(double), hipMemcpyDeviceToHost, m->stream));
    MLP_HIP(hipStreamSynchronize(m->stream));
    return AXM_SUCCESS;
}

axm_status axm_mlp_adam_step(axm_mlp m, const float* grad, float lr, float beta1, float beta2, float eps) {
    if (!m) return AXM_NULL_POINTER;
    if (!(lr > 0.0f) || !(beta1 >= 0.0f && beta1 < 1.0f) || !(beta2 >= 0.0f && beta2 < 1.0f) || !(eps > 0.0f))
        return AXM_INVALID_PARAMETER;
    (void)hipSetDevice(m->device);
    if (grad) MLP_HIP(hipMemcpyAsync(m->grad, grad, sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice, m->stream));
    m->adam_t += 1;
    const float c1 = 1.0f - std::pow(beta1, (float)m->adam_t);
    const float c2 = 1.0f - std::pow(beta2, (float)m->adam_t);
    launch_mlp_adam(m->prm, m->m, m->v, m->grad, lr, beta1, beta2, eps, c1, c2, m->stream);
    launch_mlp_cast(m->prm, m->w1, m->w2, m->w3, m->stream);
    MLP_HIP(hipGetLastError());
    MLP_HIP(hipStreamSynchronize(m->stream));
    return AXM_SUCCESS;
}

axm_status axm_mlp_train(axm_mlp m, const float* X, const float* Y, uint32_t n, uint32_t epochs, uint32_t batch,
                         float lr, float* epoch_loss) {
    if (!m || !X || !Y) return AXM_NULL_POINTER;
    if (!n || !batch || !(lr > 0.0f)) return AXM_INVALID_PARAMETER;
    (void)hipSetDevice(m->device);
    axm_status st = ensure_rows(m, n);
    if (st != AXM_SUCCESS) return st;
    MLP_HIP(hipMemcpyAsync(m->X, X, sizeof(float) * MLP_IN * (size_t)n, hipMemcpyHostToDevice, m->stream));
    MLP_HIP(hipMemcpyAsync(m->Y, Y, sizeof(float) * MLP_OUT * (size_t)n, hipMemcpyHostToDevice, m->stream));
    const uint32_t nb = (n + batch - 1) / batch;
    double* losses = nullptr;
    if (dalloc(losses, nb) != AXM_SUCCESS) return AXM_BUFFER_ALLOCATION_FAILED;
    std::vector<double> h(nb);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (uint32_t e = 0; e < epochs && st == AXM_SUCCESS; ++e) {
        for (uint32_t b = 0; b < nb; ++b) {
            const uint32_t r0 = b * batch;
            const uint32_t rows = (n - r0) < batch ? (n - r0) : batch;
            launch_mlp_grad(m->dev(), m->X + (size_t)r0 * MLP_IN, m->Y + (size_t)r0 * MLP_OUT, rows, m->partial,
                            m->loss_partial, m->grad, losses + b, m->blocks, m->stream);
            m->adam_t += 1;
            const float c1 = 1.0f - std::pow(b1, (float)m->adam_t);
            const float c2 = 1.0f - std::pow(b2, (float)m->adam_t);
            launch_mlp_adam(m->prm, m->m, m->v, m->grad, lr, b1, b2, eps, c1, c2, m->stream);
            launch_mlp_cast(m->prm, m->w1, m->w2, m->w3, m->stream);
        }
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(h.data(), losses, sizeof(double) * nb, hipMemcpyDeviceToHost, m->stream) != hipSuccess ||
            hipStreamSynchronize(m->stream) != hipSuccess) {
            st = AXM_GPU_OPERATION_FAILED;
            break;
        }
        if (epoch_loss) {
            double s = 0.0;
            for (uint32_t b = 0; b < nb; ++b) {
                const uint32_t rows = (n - b * batch) < batch ? (n - b * batch) : batch;
                s += h[b] * rows;
            }
            epoch_loss[e] = (float)(s / n);
        }
    }
    dfree(losses);
    return st;
}


// Network file (version 1): magic "AXMMLP01", version, parameter count, Adam step count,
// then the fp32 parameters, Adam's first and second moments (packing of axm_mlp_set_params)
// and the FNV-1a hash.  Loading restores the network and its optimiser state exactly, so
// training resumed from a file continues bit for bit.
axm_status axm_mlp_save(axm_mlp m, const char* path) {
    if (!m || !path) return AXM_NULL_POINTER;
    (void)hipSetDevice(m->device);
    std::vector<float> p(MLP_NPARAM), mo(MLP_NPARAM), vo(MLP_NPARAM);
    MLP_HIP(hipMemcpyAsync(p.data(), m->prm, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost, m->stream));
    MLP_HIP(hipMemcpyAsync(mo.data(), m->m, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost, m->stream));
    MLP_HIP(hipMemcpyAsync(vo.data(), m->v, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost, m->stream));
    MLP_HIP(hipStreamSynchronize(m->stream));
    axm::BinWriter f(path);
    if (!f.ok()) return AXM_INVALID_PARAMETER;
    const uint32_t head[3] = {1u, (uint32_t)MLP_NPARAM, m->adam_t};
    f.put("AXMMLP01", 8);
    f.put(head, sizeof(head));
    f.vec(p);
    f.vec(mo);
    f.vec(vo);
    return f.finish() ? AXM_SUCCESS : AXM_INVALID_PARAMETER;
}

axm_status axm_mlp_load(axm_mlp m, const char* path) {
    if (!m || !path) return AXM_NULL_POINTER;
    axm::BinReader f(path);
    if (!f.ok()) return AXM_INVALID_PARAMETER;
    char magic[8];
    uint32_t head[3];
    f.get(magic, 8);
    f.get(head, sizeof(head));
    if (!f.ok() || std::memcmp(magic, "AXMMLP01", 8) != 0 || head[0] != 1u || head[1] != (uint32_t)MLP_NPARAM)
        return AXM_INVALID_PARAMETER;
    std::vector<float> p, mo, vo;
    f.vec(p, MLP_NPARAM);
    f.vec(mo, MLP_NPARAM);
    f.vec(vo, MLP_NPARAM);
    if (!f.finish()) return AXM_INVALID_PARAMETER;
    for (int i = 0; i < MLP_NPARAM; ++i)
        if (!std::isfinite(p[i]) || !std::isfinite(mo[i]) || !std::isfinite(vo[i])) return AXM_INVALID_PARAMETER;
    (void)hipSetDevice(m->device);
    MLP_HIP(hipMemcpyAsync(m->m, mo.data(), sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice, m->stream));
    MLP_HIP(hipMemcpyAsync(m->v, vo.data(), sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice, m->stream));
    const axm_status st = upload_params(m, p.data());
    if (st == AXM_SUCCESS) m->adam_t = head[2];
    return st;
}

}  // extern "C"
