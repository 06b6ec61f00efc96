// Learned warm-start network (24 -> 128 -> 128 -> 3, ReLU): device buffers and
// launchers of the BF16 MFMA forward / backward kernels and of Adam.
//
// Parameter packing (fp32 master copy, 20,099 floats): W1 [128][24], b1 [128],
// W2 [128][128], b2 [128], W3 [3][128], b3 [3]; W is [out][in].  The kernels
// read bf16 copies of the weights (round to nearest even), rebuilt after every
// update: W1 padded to [128][32], W2 [128][128], W3 padded to [16][128].
// Inputs are rounded to bf16 where they enter an MFMA, hidden activations after
// the ReLU; accumulation is fp32 starting from the bias.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace axm {

constexpr int MLP_IN = 24, MLP_INP = 32, MLP_H = 128, MLP_OUT = 3, MLP_OUTP = 16;
constexpr int MLP_OFF_W1 = 0;
constexpr int MLP_OFF_B1 = MLP_OFF_W1 + MLP_H * MLP_IN;
constexpr int MLP_OFF_W2 = MLP_OFF_B1 + MLP_H;
constexpr int MLP_OFF_B2 = MLP_OFF_W2 + MLP_H * MLP_H;
constexpr int MLP_OFF_W3 = MLP_OFF_B2 + MLP_H;
constexpr int MLP_OFF_B3 = MLP_OFF_W3 + MLP_OUT * MLP_H;
constexpr int MLP_NPARAM = MLP_OFF_B3 + MLP_OUT;  // 20099
constexpr int MLP_ROWS_PER_BLOCK = 64;             // 4 waves x 16 rows
constexpr float MLP_G_REF = 9.81f;                 // impulse unit: effective mass * 9.81 m/s^2 * dt
constexpr float MLP_Y_MAX = 1024.0f;               // largest prediction used (impulse units)

struct MlpDev {
    const float* prm;      // fp32 master parameters (biases are read from here)
    const uint16_t* w1;    // bf16 [128][32]
    const uint16_t* w2;    // bf16 [128][128]
    const uint16_t* w3;    // bf16 [16][128]
};

// bf16 copies of the weights from the fp32 parameters.
void launch_mlp_cast(const float* prm, uint16_t* w1, uint16_t* w2, uint16_t* w3, hipStream_t s);
// Y[r][0..2] for rows r < n, X [n][24] fp32.  With n_dev, n = min(*n_dev * n_mul, n_cap).
void launch_mlp_forward(const MlpDev& P, const float* X, uint32_t n_cap, const uint32_t* n_dev, uint32_t n_mul,
                        float* Y, uint32_t blocks, hipStream_t s);
// Mean-squared-error loss gradient over n rows: per-block partials [blocks][MLP_NPARAM] and
// per-block loss sums (double), then a fixed-order reduction into grad and *loss.
size_t mlp_grad_partial_floats(uint32_t blocks);
void launch_mlp_grad(const MlpDev& P, const float* X, const float* Y, uint32_t n, float* partial, double* loss_partial,
                     float* grad, double* loss, uint32_t blocks, hipStream_t s);
// One bias-corrected Adam step (c1 = 1 - b1^t, c2 = 1 - b2^t computed by the caller in float).
void launch_mlp_adam(float* prm, float* m, float* v, const float* g, float lr, float b1, float b2, float eps, float c1,
                     float c2, hipStream_t s);
// Dynamic LDS bytes of the gradient kernel (set once per device).
size_t mlp_grad_lds_bytes();
hipError_t mlp_prepare_device();

}  // namespace axm
