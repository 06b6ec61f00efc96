// Warm-start MLP kernels for gfx950: BF16 MFMA forward and backward with fp32
// accumulation, a fixed-order reduction of per-block gradients, and Adam.
//
// MFMA shape: gfx950's v_mfma_f32_16x16x32_bf16 (one wave computes a 16x16 fp32 tile
// from a 16x32 bf16 A and a 32x16 bf16 B tile).  Lane l (r = l % 16, q = l / 16) holds
//   A[r][8q + i], B[8q + i][r] (i = 0..7) and D[4q + i][r] (i = 0..3).
// A 64-row block is one workgroup of 4 waves, 16 rows per wave; hidden
// activations are staged through LDS between layers (row-major, bf16).
// The same rounding points as the CPU oracle's bf16 emulation: inputs and
// weights rounded to bf16 (nearest even), hidden activations rounded after the
// ReLU, propagated hidden gradients rounded before they feed an MFMA.
#include "k_mlp.hpp"

namespace axm {

typedef short bf16x8 __attribute__((ext_vector_type(8)));   // bf16 bit patterns
typedef __bf16 bf16x8v __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float x) {
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7F800000u) == 0x7F800000u) return (uint16_t)((u >> 16) | ((u & 0xFFFFu) ? 0x40u : 0u));
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ uint16_t relu_bf(float v) { return f2bf(v > 0.0f ? v : 0.0f); }

__device__ __forceinline__ f32x4 mfma32(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8v, a), __builtin_bit_cast(bf16x8v, b), c, 0,
                                                   0, 0);
}
// eight consecutive bf16 (16-byte aligned)
__device__ __forceinline__ bf16x8 ld8(const uint16_t* p) { return *(const bf16x8*)p; }
// eight bf16 `stride` elements apart
__device__ __forceinline__ bf16x8 ld8s(const uint16_t* p, int stride) {
    bf16x8 r;
#pragma unroll
    for (int i = 0; i < 8; ++i) r[i] = (short)p[i * stride];
    return r;
}
__device__ __forceinline__ f32x4 splat(float v) {
    f32x4 r;
    r.x = r.y = r.z = r.w = v;
    return r;
}

// ---------------------------------------------------------------- casts
__global__ void k_mlp_cast(const float* __restrict__ prm, uint16_t* __restrict__ w1, uint16_t* __restrict__ w2,
                           uint16_t* __restrict__ w3) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < (uint32_t)(MLP_H * MLP_INP)) {
        const uint32_t o = i / MLP_INP, k = i % MLP_INP;
        w1[i] = k < (uint32_t)MLP_IN ? f2bf(prm[MLP_OFF_W1 + o * MLP_IN + k]) : (uint16_t)0;
    }
    if (i < (uint32_t)(MLP_H * MLP_H)) w2[i] = f2bf(prm[MLP_OFF_W2 + i]);
    if (i < (uint32_t)(MLP_OUTP * MLP_H)) {
        const uint32_t o = i / MLP_H, k = i % MLP_H;
        w3[i] = o < (uint32_t)MLP_OUT ? f2bf(prm[MLP_OFF_W3 + o * MLP_H + k]) : (uint16_t)0;
    }
}

// Cooperative copy of the bf16 weights and fp32 biases into LDS.
__device__ __forceinline__ void load_weights(const MlpDev& P, uint16_t* sW1, uint16_t* sW2, uint16_t* sW3, float* sB) {
    const uint4* g1 = (const uint4*)P.w1;
    const uint4* g2 = (const uint4*)P.w2;
    const uint4* g3 = (const uint4*)P.w3;
    uint4* l1 = (uint4*)sW1;
    uint4* l2 = (uint4*)sW2;
    uint4* l3 = (uint4*)sW3;
    for (uint32_t i = threadIdx.x; i < (uint32_t)(MLP_H * MLP_INP / 8); i += blockDim.x) l1[i] = g1[i];
    for (uint32_t i = threadIdx.x; i < (uint32_t)(MLP_H * MLP_H / 8); i += blockDim.x) l2[i] = g2[i];
    for (uint32_t i = threadIdx.x; i < (uint32_t)(MLP_OUTP * MLP_H / 8); i += blockDim.x) l3[i] = g3[i];
    for (uint32_t i = threadIdx.x; i < (uint32_t)MLP_H; i += blockDim.x) {
        sB[i] = P.prm[MLP_OFF_B1 + i];
        sB[MLP_H + i] = P.prm[MLP_OFF_B2 + i];
    }
    if (threadIdx.x < 4u) sB[2 * MLP_H + threadIdx.x] = threadIdx.x < (uint32_t)MLP_OUT ? P.prm[MLP_OFF_B3 + threadIdx.x] : 0.0f;
}

// Input tile of one wave: the A operand of the one 32-wide k step (k >= 24 padded with 0).
__device__ __forceinline__ bf16x8 load_x(const float* __restrict__ X, uint32_t n, uint32_t row, int q) {
    const bool ok = row < n;
    const float* xr = X + (size_t)row * MLP_IN;
    bf16x8 a;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int k = 8 * q + i;
        a[i] = (short)((ok && k < MLP_IN) ? f2bf(xr[k]) : (uint16_t)0);
    }
    return a;
}

// Layer 1 of a 16-row tile: acc[j] = b1 + X W1^T for output tile j.
__device__ __forceinline__ void layer1(const uint16_t* sW1, const float* sB, bf16x8 a, int r, int q, f32x4 acc[8]) {
#pragma unroll
    for (int j = 0; j < 8; ++j) acc[j] = mfma32(a, ld8(&sW1[(16 * j + r) * MLP_INP + 8 * q]), splat(sB[16 * j + r]));
}
// Layer 2 of a 16-row tile from the wave's activations act[16][128].
__device__ __forceinline__ void layer2(const uint16_t* sW2, const float* sB, const uint16_t* act, int r, int q,
                                       f32x4 acc[8]) {
#pragma unroll
    for (int j = 0; j < 8; ++j) acc[j] = splat(sB[MLP_H + 16 * j + r]);
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
        const bf16x8 a = ld8(&act[r * MLP_H + 32 * kk + 8 * q]);
#pragma unroll
        for (int j = 0; j < 8; ++j) acc[j] = mfma32(a, ld8(&sW2[(16 * j + r) * MLP_H + 32 * kk + 8 * q]), acc[j]);
    }
}
__device__ __forceinline__ f32x4 layer3(const uint16_t* sW3, const float* sB, const uint16_t* act, int r, int q) {
    f32x4 y = splat(r < MLP_OUT ? sB[2 * MLP_H + r] : 0.0f);
#pragma unroll
    for (int kk = 0; kk < 4; ++kk)
        y = mfma32(ld8(&act[r * MLP_H + 32 * kk + 8 * q]), ld8(&sW3[r * MLP_H + 32 * kk + 8 * q]), y);
    return y;
}
// ReLU + bf16 of a 16x128 accumulator set into row-major act[16][128].
__device__ __forceinline__ void store_act(uint16_t* act, const f32x4 acc[8], int r, int q) {
#pragma unroll
    for (int j = 0; j < 8; ++j)
#pragma unroll
        for (int i = 0; i < 4; ++i) act[(4 * q + i) * MLP_H + 16 * j + r] = relu_bf(acc[j][i]);
}

// ---------------------------------------------------------------- forward
__global__ void __launch_bounds__(256) k_mlp_forward(MlpDev P, const float* __restrict__ X, uint32_t n_cap,
                                                     const uint32_t* __restrict__ n_dev, uint32_t n_mul,
                                                     float* __restrict__ Y) {
    __shared__ __align__(16) uint16_t sW1[MLP_H * MLP_INP];
    __shared__ __align__(16) uint16_t sW2[MLP_H * MLP_H];
    __shared__ __align__(16) uint16_t sW3[MLP_OUTP * MLP_H];
    __shared__ float sB[2 * MLP_H + 4];
    __shared__ __align__(16) uint16_t sAct[4][16 * MLP_H];
    uint32_t n = n_cap;
    if (n_dev) {
        const uint64_t v = (uint64_t)(*n_dev) * n_mul;
        n = v < n_cap ? (uint32_t)v : n_cap;
    }
    if (blockIdx.x * (uint32_t)MLP_ROWS_PER_BLOCK >= n) return;  // uniform per block
    load_weights(P, sW1, sW2, sW3, sB);
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 15, q = lane >> 4;
    uint16_t* act = sAct[wave];
    for (uint32_t base = blockIdx.x * MLP_ROWS_PER_BLOCK; base < n; base += gridDim.x * MLP_ROWS_PER_BLOCK) {
        const uint32_t r0 = base + 16u * wave;
        const bf16x8 a = load_x(X, n, r0 + r, q);
        f32x4 acc[8];
        layer1(sW1, sB, a, r, q, acc);
        store_act(act, acc, r, q);
        __syncthreads();
        layer2(sW2, sB, act, r, q, acc);
        __syncthreads();
        store_act(act, acc, r, q);
        __syncthreads();
        const f32x4 y = layer3(sW3, sB, act, r, q);
        if (r < MLP_OUT) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const uint32_t row = r0 + 4 * q + i;
                if (row < n) Y[(size_t)row * MLP_OUT + r] = y[i];
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- backward
// LDS layout of the gradient kernel (bytes): W1 8192, W2 32768, W3 4096,
// W3 as float 1536, biases 1056, X tile 4096, A1 16384 (h1, then d1),
// A2 16384 (h2, then d2), dy 1024.
constexpr size_t G_OFF_W2 = 8192, G_OFF_W3 = G_OFF_W2 + 32768, G_OFF_W3F = G_OFF_W3 + 4096,
                 G_OFF_B = G_OFF_W3F + 1536, G_OFF_X = G_OFF_B + 1056, G_OFF_A1 = G_OFF_X + 4096,
                 G_OFF_A2 = G_OFF_A1 + 16384, G_OFF_DY = G_OFF_A2 + 16384, G_LDS = G_OFF_DY + 1024;

__global__ void __launch_bounds__(256) k_mlp_grad(MlpDev P, const float* __restrict__ X, const float* __restrict__ Yt,
                                                  uint32_t n, float scale, float* __restrict__ partial,
                                                  double* __restrict__ loss_partial) {
    extern __shared__ __align__(16) unsigned char smem[];
    uint16_t* sW1 = (uint16_t*)smem;
    uint16_t* sW2 = (uint16_t*)(smem + G_OFF_W2);
    uint16_t* sW3 = (uint16_t*)(smem + G_OFF_W3);
    float* sW3f = (float*)(smem + G_OFF_W3F);
    float* sB = (float*)(smem + G_OFF_B);
    uint16_t* sX = (uint16_t*)(smem + G_OFF_X);
    uint16_t* sA1 = (uint16_t*)(smem + G_OFF_A1);
    uint16_t* sA2 = (uint16_t*)(smem + G_OFF_A2);
    float* sdy = (float*)(smem + G_OFF_DY);
    load_weights(P, sW1, sW2, sW3, sB);
    for (uint32_t i = threadIdx.x; i < (uint32_t)(MLP_OUT * MLP_H); i += blockDim.x) sW3f[i] = bf2f(P.w3[i]);
    __syncthreads();
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, r = lane & 15, q = lane >> 4;
    f32x4 gW2[2][8], gW1[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
#pragma unroll
        for (int b = 0; b < 8; ++b) gW2[a][b] = splat(0.0f);
        gW1[a][0] = gW1[a][1] = splat(0.0f);
    }
    float gW3a = 0.0f, gW3b = 0.0f, gb = 0.0f, gb3 = 0.0f;
    double lossacc = 0.0;
    for (uint32_t base = blockIdx.x * MLP_ROWS_PER_BLOCK; base < n; base += gridDim.x * MLP_ROWS_PER_BLOCK) {
        const uint32_t r0 = base + 16u * wave;
        // ---- forward (keep x, h1, h2 in LDS)
        const bf16x8 a = load_x(X, n, r0 + r, q);
#pragma unroll
        for (int i = 0; i < 8; ++i) sX[(16 * wave + r) * MLP_INP + 8 * q + i] = (uint16_t)a[i];
        f32x4 acc[8];
        uint16_t* h1w = sA1 + 16 * wave * MLP_H;
        uint16_t* h2w = sA2 + 16 * wave * MLP_H;
        layer1(sW1, sB, a, r, q, acc);
        store_act(h1w, acc, r, q);
        __syncthreads();
        layer2(sW2, sB, h1w, r, q, acc);
        store_act(h2w, acc, r, q);
        __syncthreads();
        const f32x4 y = layer3(sW3, sB, h2w, r, q);
        if (r < 4) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const uint32_t row = r0 + 4 * q + i;
                float d = 0.0f;
                if (r < MLP_OUT && row < n) {
                    const float e = y[i] - Yt[(size_t)row * MLP_OUT + r];
                    lossacc += (double)e * (double)e;
                    d = e * scale;
                }
                sdy[(16 * wave + 4 * q + i) * 4 + r] = d;
            }
        }
        __syncthreads();
        // ---- layer 3: dW3 += dy^T h2 (rows in order), db3 += dy
        for (int e = tid, slot = 0; e < MLP_OUT * MLP_H; e += 256, ++slot) {
            const int o = e / MLP_H, k = e % MLP_H;
            float s = slot ? gW3b : gW3a;
            for (int rr = 0; rr < MLP_ROWS_PER_BLOCK; ++rr) s += sdy[rr * 4 + o] * bf2f(sA2[rr * MLP_H + k]);
            if (slot) gW3b = s; else gW3a = s;
        }
        if (tid < MLP_OUT)
            for (int rr = 0; rr < MLP_ROWS_PER_BLOCK; ++rr) gb3 += sdy[rr * 4 + tid];
        __syncthreads();
        // d2 = relu'(h2) * bf16(sum_o W3[o][k] dy[o]), in place of h2
        for (int e = tid; e < MLP_ROWS_PER_BLOCK * MLP_H; e += 256) {
            const int rr = e / MLP_H, k = e % MLP_H;
            float s = 0.0f;
            s += sW3f[k] * sdy[rr * 4 + 0];
            s += sW3f[MLP_H + k] * sdy[rr * 4 + 1];
            s += sW3f[2 * MLP_H + k] * sdy[rr * 4 + 2];
            sA2[e] = bf2f(sA2[e]) > 0.0f ? f2bf(s) : (uint16_t)0;
        }
        __syncthreads();
        if (tid < MLP_H)
            for (int rr = 0; rr < MLP_ROWS_PER_BLOCK; ++rr) gb += bf2f(sA2[rr * MLP_H + tid]);
        // ---- layer 2: dW2[o][k] += sum_r d2[r][o] h1[r][k]; this wave owns o tiles 2w, 2w+1
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const bf16x8 A0 = ld8s(&sA2[(32 * c + 8 * q) * MLP_H + 16 * (2 * wave) + r], MLP_H);
            const bf16x8 A1 = ld8s(&sA2[(32 * c + 8 * q) * MLP_H + 16 * (2 * wave + 1) + r], MLP_H);
#pragma unroll
            for (int b = 0; b < 8; ++b) {
                const bf16x8 Bm = ld8s(&sA1[(32 * c + 8 * q) * MLP_H + 16 * b + r], MLP_H);
                gW2[0][b] = mfma32(A0, Bm, gW2[0][b]);
                gW2[1][b] = mfma32(A1, Bm, gW2[1][b]);
            }
        }
        // d1 (this wave's rows) = bf16(d2 W2), masked by relu'(h1) after every wave has read h1
        f32x4 d1[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) d1[j] = splat(0.0f);
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            const bf16x8 A = ld8(&sA2[(16 * wave + r) * MLP_H + 32 * kk + 8 * q]);
#pragma unroll
            for (int j = 0; j < 8; ++j) d1[j] = mfma32(A, ld8s(&sW2[(32 * kk + 8 * q) * MLP_H + 16 * j + r], MLP_H), d1[j]);
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < 8; ++j)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                uint16_t& h = h1w[(4 * q + i) * MLP_H + 16 * j + r];
                h = bf2f(h) > 0.0f ? f2bf(d1[j][i]) : (uint16_t)0;
            }
        __syncthreads();
        // ---- layer 1: dW1[o][k] += sum_r d1[r][o] x[r][k]
#pragma unroll
        for (int c = 0; c < 2; ++c) {
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                const bf16x8 A = ld8s(&sA1[(32 * c + 8 * q) * MLP_H + 16 * (2 * wave + a) + r], MLP_H);
#pragma unroll
                for (int b = 0; b < 2; ++b)
                    gW1[a][b] = mfma32(A, ld8s(&sX[(32 * c + 8 * q) * MLP_INP + 16 * b + r], MLP_INP), gW1[a][b]);
            }
        }
        if (tid >= MLP_H) {
            float s = 0.0f;
            for (int rr = 0; rr < MLP_ROWS_PER_BLOCK; ++rr) s += bf2f(sA1[rr * MLP_H + (tid - MLP_H)]);
            gb += s;
        }
        __syncthreads();
    }
    // ---- per-block partials
    float* out = partial + (size_t)blockIdx.x * MLP_NPARAM;
#pragma unroll
    for (int a = 0; a < 2; ++a) {
#pragma unroll
        for (int b = 0; b < 8; ++b)
#pragma unroll
            for (int i = 0; i < 4; ++i)
                out[MLP_OFF_W2 + (16 * (2 * wave + a) + 4 * q + i) * MLP_H + 16 * b + r] = gW2[a][b][i];
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int k = 16 * b + r;
                if (k < MLP_IN) out[MLP_OFF_W1 + (16 * (2 * wave + a) + 4 * q + i) * MLP_IN + k] = gW1[a][b][i];
            }
    }
    if (tid < MLP_H) out[MLP_OFF_B2 + tid] = gb;
    else out[MLP_OFF_B1 + tid - MLP_H] = gb;
    out[MLP_OFF_W3 + tid] = gW3a;
    if (tid + 256 < MLP_OUT * MLP_H) out[MLP_OFF_W3 + tid + 256] = gW3b;
    if (tid < MLP_OUT) out[MLP_OFF_B3 + tid] = gb3;
    // block sum of the loss (fixed order)
    __syncthreads();
    double* red = (double*)(smem + G_OFF_A1);
    red[tid] = lossacc;
    __syncthreads();
    if (tid == 0) {
        double s = 0.0;
        for (int t = 0; t < 256; ++t) s += red[t];
        loss_partial[blockIdx.x] = s;
    }
}

__global__ void k_mlp_reduce(const float* __restrict__ partial, uint32_t blocks, const double* __restrict__ loss_partial,
                             double inv_count, float* __restrict__ grad, double* __restrict__ loss) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < (uint32_t)MLP_NPARAM) {
        float s = 0.0f;
        for (uint32_t b = 0; b < blocks; ++b) s += partial[(size_t)b * MLP_NPARAM + p];
        grad[p] = s;
    }
    if (p == 0) {
        double s = 0.0;
        for (uint32_t b = 0; b < blocks; ++b) s += loss_partial[b];
        *loss = s * inv_count;
    }
}

__global__ void k_mlp_adam(float* __restrict__ prm, float* __restrict__ m, float* __restrict__ v,
                           const float* __restrict__ g, float lr, float b1, float b2, float eps, float c1, float c2) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint32_t)MLP_NPARAM) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / c1;
    const float vh = vi / c2;
    prm[i] -= lr * mh / (sqrtf(vh) + eps);
}

// ---------------------------------------------------------------- launchers
void launch_mlp_cast(const float* prm, uint16_t* w1, uint16_t* w2, uint16_t* w3, hipStream_t s) {
    hipLaunchKernelGGL(k_mlp_cast, dim3((MLP_H * MLP_H + 255) / 256), dim3(256), 0, s, prm, w1, w2, w3);
}

void launch_mlp_forward(const MlpDev& P, const float* X, uint32_t n_cap, const uint32_t* n_dev, uint32_t n_mul,
                        float* Y, uint32_t blocks, hipStream_t s) {
    if (n_cap == 0) return;
    const uint32_t need = (n_cap + MLP_ROWS_PER_BLOCK - 1) / MLP_ROWS_PER_BLOCK;
    const uint32_t g = blocks < need ? blocks : need;
    hipLaunchKernelGGL(k_mlp_forward, dim3(g ? g : 1), dim3(256), 0, s, P, X, n_cap, n_dev, n_mul, Y);
}

size_t mlp_grad_partial_floats(uint32_t blocks) { return (size_t)blocks * MLP_NPARAM; }
size_t mlp_grad_lds_bytes() { return G_LDS; }

hipError_t mlp_prepare_device() {
    return hipFuncSetAttribute((const void*)k_mlp_grad, hipFuncAttributeMaxDynamicSharedMemorySize, (int)G_LDS);
}

void launch_mlp_grad(const MlpDev& P, const float* X, const float* Y, uint32_t n, float* partial, double* loss_partial,
                     float* grad, double* loss, uint32_t blocks, hipStream_t s) {
    const uint32_t need = (n + MLP_ROWS_PER_BLOCK - 1) / MLP_ROWS_PER_BLOCK;
    const uint32_t g = blocks < need ? blocks : need;
    const float scale = 2.0f / (float)((double)n * MLP_OUT);
    hipLaunchKernelGGL(k_mlp_grad, dim3(g ? g : 1), dim3(256), G_LDS, s, P, X, Y, n, scale, partial, loss_partial);
    hipLaunchKernelGGL(k_mlp_reduce, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, s, partial, g ? g : 1u,
                       loss_partial, 1.0 / ((double)n * MLP_OUT), grad, loss);
}

void launch_mlp_adam(float* prm, float* m, float* v, const float* g, float lr, float b1, float b2, float eps, float c1,
                     float c2, hipStream_t s) {
    hipLaunchKernelGGL(k_mlp_adam, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, s, prm, m, v, g, lr, b1, b2, eps,
                       c1, c2);
}

}  // namespace axm
