// CPU oracle — TEST INFRASTRUCTURE ONLY.
//
// Reference forward / backward / Adam for the warm-start MLP
// (24 -> 128 -> 128 -> 3, ReLU).  The reference has no learned component; the
// network is specified by this project.  Two modes:
//   * bf16 emulation (float accumulation, operands rounded to bfloat16 with
//     round-to-nearest-even exactly where the GPU kernels round them), used as
//     the parity target of the MFMA kernels;
//   * plain double, used for the finite-difference gradient check.
// Adam follows Kingma & Ba (2015), bias-corrected.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace orc {

constexpr int MLP_IN = 24, MLP_H = 128, MLP_OUT = 3;
constexpr int MLP_OFF_W1 = 0;
constexpr int MLP_OFF_B1 = MLP_OFF_W1 + MLP_H * MLP_IN;
constexpr int MLP_OFF_W2 = MLP_OFF_B1 + MLP_H;
constexpr int MLP_OFF_B2 = MLP_OFF_W2 + MLP_H * MLP_H;
constexpr int MLP_OFF_W3 = MLP_OFF_B2 + MLP_H;
constexpr int MLP_OFF_B3 = MLP_OFF_W3 + MLP_OUT * MLP_H;
constexpr int MLP_NPARAM = MLP_OFF_B3 + MLP_OUT;  // 20099

inline float bf16r(float x) {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    if ((u & 0x7F800000u) == 0x7F800000u) return x;  // inf / nan unchanged
    u += 0x7FFFu + ((u >> 16) & 1u);
    u &= 0xFFFF0000u;
    float r;
    std::memcpy(&r, &u, 4);
    return r;
}

template <class T> struct MlpCache {
    std::vector<T> x, h1, h2, y;  // x rounded input, post-activation hidden, output
};

// y = W3 relu(W2 relu(W1 x + b1) + b2) + b3 for B rows.  `emu` rounds weights,
// inputs and hidden activations to bf16 (float path only).
template <class T> void mlp_forward(const float* prm, const float* X, int B, bool emu, MlpCache<T>& c) {
    auto rw = [&](float v) -> T { return emu ? T(bf16r(v)) : T(v); };
    c.x.assign((size_t)B * MLP_IN, T(0));
    c.h1.assign((size_t)B * MLP_H, T(0));
    c.h2.assign((size_t)B * MLP_H, T(0));
    c.y.assign((size_t)B * MLP_OUT, T(0));
    std::vector<T> W1(MLP_H * MLP_IN), W2(MLP_H * MLP_H), W3(MLP_OUT * MLP_H);
    for (int i = 0; i < MLP_H * MLP_IN; ++i) W1[i] = rw(prm[MLP_OFF_W1 + i]);
    for (int i = 0; i < MLP_H * MLP_H; ++i) W2[i] = rw(prm[MLP_OFF_W2 + i]);
    for (int i = 0; i < MLP_OUT * MLP_H; ++i) W3[i] = rw(prm[MLP_OFF_W3 + i]);
    for (int r = 0; r < B; ++r) {
        T* x = &c.x[(size_t)r * MLP_IN];
        for (int k = 0; k < MLP_IN; ++k) x[k] = rw(X[(size_t)r * MLP_IN + k]);
        T* h1 = &c.h1[(size_t)r * MLP_H];
        for (int o = 0; o < MLP_H; ++o) {
            T s = T(prm[MLP_OFF_B1 + o]);
            for (int k = 0; k < MLP_IN; ++k) s += W1[o * MLP_IN + k] * x[k];
            s = s > T(0) ? s : T(0);
            h1[o] = emu ? T(bf16r((float)s)) : s;
        }
        T* h2 = &c.h2[(size_t)r * MLP_H];
        for (int o = 0; o < MLP_H; ++o) {
            T s = T(prm[MLP_OFF_B2 + o]);
            for (int k = 0; k < MLP_H; ++k) s += W2[o * MLP_H + k] * h1[k];
            s = s > T(0) ? s : T(0);
            h2[o] = emu ? T(bf16r((float)s)) : s;
        }
        T* y = &c.y[(size_t)r * MLP_OUT];
        for (int o = 0; o < MLP_OUT; ++o) {
            T s = T(prm[MLP_OFF_B3 + o]);
            for (int k = 0; k < MLP_H; ++k) s += W3[o * MLP_H + k] * h2[k];
            y[o] = s;
        }
    }
}

// Mean-squared-error loss over B x 3 outputs and its gradient w.r.t. every
// parameter (same packing as prm).  With `emu`, the backward pass rounds the
// propagated hidden gradients to bf16 where the GPU kernels feed them to MFMA.
template <class T> double mlp_loss_grad(const float* prm, const float* X, const float* Y, int B, bool emu, T* grad) {
    MlpCache<T> c;
    mlp_forward<T>(prm, X, B, emu, c);
    auto rw = [&](float v) -> T { return emu ? T(bf16r(v)) : T(v); };
    auto rg = [&](T v) -> T { return emu ? T(bf16r((float)v)) : v; };
    for (int i = 0; i < MLP_NPARAM; ++i) grad[i] = T(0);
    double loss = 0.0;
    const T scale = T(2) / T(B * MLP_OUT);
    std::vector<T> dy((size_t)B * MLP_OUT), d2((size_t)B * MLP_H), d1((size_t)B * MLP_H);
    for (int r = 0; r < B; ++r)
        for (int o = 0; o < MLP_OUT; ++o) {
            const T e = c.y[(size_t)r * MLP_OUT + o] - T(Y[(size_t)r * MLP_OUT + o]);
            loss += (double)e * (double)e;
            dy[(size_t)r * MLP_OUT + o] = e * scale;
        }
    loss /= (double)(B * MLP_OUT);
    // layer 3
    for (int r = 0; r < B; ++r) {
        for (int o = 0; o < MLP_OUT; ++o) {
            const T g = dy[(size_t)r * MLP_OUT + o];
            grad[MLP_OFF_B3 + o] += g;
            for (int k = 0; k < MLP_H; ++k) grad[MLP_OFF_W3 + o * MLP_H + k] += g * c.h2[(size_t)r * MLP_H + k];
        }
        for (int k = 0; k < MLP_H; ++k) {
            T s = T(0);
            for (int o = 0; o < MLP_OUT; ++o) s += rw(prm[MLP_OFF_W3 + o * MLP_H + k]) * dy[(size_t)r * MLP_OUT + o];
            d2[(size_t)r * MLP_H + k] = c.h2[(size_t)r * MLP_H + k] > T(0) ? rg(s) : T(0);
        }
    }
    // layer 2
    for (int r = 0; r < B; ++r) {
        for (int o = 0; o < MLP_H; ++o) {
            const T g = d2[(size_t)r * MLP_H + o];
            grad[MLP_OFF_B2 + o] += g;
            for (int k = 0; k < MLP_H; ++k) grad[MLP_OFF_W2 + o * MLP_H + k] += g * c.h1[(size_t)r * MLP_H + k];
        }
        for (int k = 0; k < MLP_H; ++k) {
            T s = T(0);
            for (int o = 0; o < MLP_H; ++o) s += rw(prm[MLP_OFF_W2 + o * MLP_H + k]) * d2[(size_t)r * MLP_H + o];
            d1[(size_t)r * MLP_H + k] = c.h1[(size_t)r * MLP_H + k] > T(0) ? rg(s) : T(0);
        }
    }
    // layer 1
    for (int r = 0; r < B; ++r)
        for (int o = 0; o < MLP_H; ++o) {
            const T g = d1[(size_t)r * MLP_H + o];
            grad[MLP_OFF_B1 + o] += g;
            for (int k = 0; k < MLP_IN; ++k) grad[MLP_OFF_W1 + o * MLP_IN + k] += g * c.x[(size_t)r * MLP_IN + k];
        }
    return loss;
}

// One bias-corrected Adam update (step t >= 1).
inline void adam_step(float* prm, float* m, float* v, const float* g, int n, int t, float lr, float b1, float b2,
                      float eps) {
    const float c1 = 1.0f - std::pow(b1, (float)t);
    const float c2 = 1.0f - std::pow(b2, (float)t);
    for (int i = 0; i < n; ++i) {
        m[i] = b1 * m[i] + (1.0f - b1) * g[i];
        v[i] = b2 * v[i] + (1.0f - b2) * g[i] * g[i];
        const float mh = m[i] / c1;
        const float vh = v[i] / c2;
        prm[i] -= lr * mh / (std::sqrt(vh) + eps);
    }
}

}  // namespace orc
