/*
 * mpm_mlp.h — the learned return-map surrogate: a 13 -> 64 -> 64 -> 64 -> 7
 * MLP with SiLU, BF16 weights and activations, FP32 accumulation on MFMA.
 *
 * Parameter block (FP32 master copy, PARAM_COUNT floats), layer l maps
 * IN_l -> OUT_l with W_l stored row-major [IN_l][OUT_l]:
 *   W1 [16][64] (inputs 13..15 are zero padding)  b1 [64]
 *   W2 [64][64]                                    b2 [64]
 *   W3 [64][64]                                    b3 [64]
 *   W4 [64][16] (outputs 7..15 are zero padding)   b4 [16]
 */
#ifndef MPM_AMD_MLP_H
#define MPM_AMD_MLP_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#define MLP_IN 13
#define MLP_OUT 7
#define MLP_IN_PAD 16
#define MLP_HID 64
#define MLP_OUT_PAD 16

#define OFF_W1 0
#define OFF_B1 (OFF_W1 + MLP_IN_PAD * MLP_HID)
#define OFF_W2 (OFF_B1 + MLP_HID)
#define OFF_B2 (OFF_W2 + MLP_HID * MLP_HID)
#define OFF_W3 (OFF_B2 + MLP_HID)
#define OFF_B3 (OFF_W3 + MLP_HID * MLP_HID)
#define OFF_W4 (OFF_B3 + MLP_HID)
#define OFF_B4 (OFF_W4 + MLP_HID * MLP_OUT_PAD)
#define PARAM_COUNT (OFF_B4 + MLP_OUT_PAD)

/* number of BF16 weight entries (weights only, biases stay FP32) */
#define WEIGHT_BF16_COUNT (MLP_IN_PAD * MLP_HID + 2 * MLP_HID * MLP_HID + MLP_HID * MLP_OUT_PAD)

struct MLPNorm {
    float in_mean[MLP_IN], in_std[MLP_IN];
    float in_lo[MLP_IN], in_hi[MLP_IN];   /* envelope of the normalised inputs */
    float out_mean[MLP_OUT], out_std[MLP_OUT];
};

/* Inference on n normalised FP32 samples: feature k of sample i is
 * x[i * xs + k * xk] (13 features) and output k goes to y[i * ys + k * yk]
 * (7 outputs).  dcount (device, optional) caps n at launch time. */
hipError_t mlp_launch_forward(const float *params, const uint16_t *wbf16, const float *x, int64_t xs, int64_t xk,
                              float *y, int64_t ys, int64_t yk, int64_t n, hipStream_t s,
                              const uint32_t *dcount = nullptr);
hipError_t mlp_launch_pack(const float *params, uint16_t *wbf16, hipStream_t s);
/* Gradient of mean((y - t)^2 over the 7 outputs and n rows) w.r.t. params;
 * partial: scratch of ceil(n / 64) * PARAM_COUNT floats; loss: 1 float. */
hipError_t mlp_launch_gradient(const float *params, const uint16_t *wbf16, const float *x, const float *t,
                               int64_t n, float *partial, float *grad, float *loss, hipStream_t s);
hipError_t mlp_launch_adam(float *params, const float *grad, float *m, float *v, float lr, float beta1, float beta2,
                           float eps, int step, hipStream_t s);

#endif /* MPM_AMD_MLP_H */
