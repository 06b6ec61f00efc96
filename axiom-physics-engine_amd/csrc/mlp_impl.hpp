// Device state of one warm-start network (shared by the network ABI in mlp.cpp
// and by worlds that run it in their step).
#pragma once
#include "axiom/axm.h"
#include "k_mlp.hpp"

struct axm_mlp_s {
    int device = 0;
    hipStream_t stream = nullptr;
    float* prm = nullptr;     // fp32 master parameters
    float* m = nullptr;       // Adam first moment
    float* v = nullptr;       // Adam second moment
    float* grad = nullptr;
    uint16_t* w1 = nullptr;   // bf16 weights used by the kernels
    uint16_t* w2 = nullptr;
    uint16_t* w3 = nullptr;
    float* partial = nullptr; // per-block gradient partials
    double* loss_partial = nullptr;
    double* loss = nullptr;   // device scalar
    uint32_t blocks = 0;      // gradient blocks (partials capacity)
    uint32_t adam_t = 0;      // Adam step count
    // staging for host-array calls and training data
    float* X = nullptr;
    float* Y = nullptr;
    uint32_t cap_rows = 0;

    axm::MlpDev dev() const { return axm::MlpDev{prm, w1, w2, w3}; }
};
