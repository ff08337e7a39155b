/*
 * mpm_surrogate.h — the surrogate object behind the opaque mpm_surrogate.
 */
#ifndef MPM_AMD_SURROGATE_H
#define MPM_AMD_SURROGATE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mpm_mlp.h"

struct mpm_surrogate {
    int device = 0;
    hipStream_t stream = nullptr;
    float *params = nullptr;      /* FP32 master parameters */
    uint16_t *wbf16 = nullptr;    /* BF16 weights used by the MFMA kernels */
    float *grad = nullptr, *adam_m = nullptr, *adam_v = nullptr, *loss = nullptr;
    float *partial = nullptr;
    int64_t partial_chunks = 0;
    float *tx = nullptr, *tt = nullptr;  /* training set */
    int64_t train_rows = 0;
    float *io_x = nullptr, *io_y = nullptr;
    int64_t io_rows = 0;
    int adam_step = 0;
    double stress_scale = 1.0, strain_scale = 1.0;
    MLPNorm norm{};
    float *scratch = nullptr;     /* this device's address of the library's fixed scratch array */

    ~mpm_surrogate();
    int ensure_partial(int64_t rows);
    int ensure_io(int64_t rows);
};

#endif /* MPM_AMD_SURROGATE_H */
