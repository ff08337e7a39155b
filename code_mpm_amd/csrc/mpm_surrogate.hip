/*
 * mpm_surrogate.hip — host side of the learned return-map surrogate (C ABI
 * in include/mpm.h): parameters, normalisation, inference, gradients and
 * Adam training on a device-resident training set.
 */
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <random>
#include <string>
#include <vector>

#include "../../include/mpm.h"
#include "mpm_surrogate.h"

/* Fixed gradient scratch of mpm_surrogate_gradient_noalloc: per 64-row chunk
 * PARAM_COUNT gradient partials, then one squared-error sum per chunk, for
 * up to MPM_SURROGATE_NOALLOC_MAX_ROWS rows (128 chunks, 5.35 MB).  Each GPU
 * the library runs on holds its own copy until the process ends.  In host
 * code the name is not a device address: kernels get the address that
 * scratch_address() looked up for the device instead. */
static constexpr int64_t NOALLOC_CHUNKS = (MPM_SURROGATE_NOALLOC_MAX_ROWS + 63) / 64;
static __device__ __align__(256) float scratch[NOALLOC_CHUNKS * (PARAM_COUNT + 1)];

namespace {
template <typename T>
void dfree(T *&p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

constexpr int MAX_DEVICES = 64;
std::mutex scratch_mu;
float *scratch_ptr[MAX_DEVICES] = {};

/* The device address of this device's scratch, looked up once per device
 * with that device current (called from mpm_surrogate_create). */
int scratch_address(int device, float **out) {
    if (device < 0 || device >= MAX_DEVICES) return MPM_ERR_RANGE;
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_ptr[device]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return MPM_ERR_HIP;
        scratch_ptr[device] = static_cast<float *>(p);
    }
    *out = scratch_ptr[device];
    return MPM_OK;
}
}  // namespace

mpm_surrogate::~mpm_surrogate() {
    (void)hipSetDevice(device);
    if (stream) (void)hipStreamSynchronize(stream);
    dfree(params); dfree(wbf16); dfree(grad); dfree(adam_m); dfree(adam_v); dfree(loss);
    dfree(partial); dfree(tx); dfree(tt); dfree(io_x); dfree(io_y);
    if (stream) (void)hipStreamDestroy(stream);
}

int mpm_surrogate::ensure_partial(int64_t rows) {
    int64_t chunks = (rows + 63) / 64;
    if (chunks <= partial_chunks) return MPM_OK;
    dfree(partial);
    if (hipMalloc(&partial, sizeof(float) * (size_t)chunks * (PARAM_COUNT + 1)) != hipSuccess) return MPM_ERR_NOMEM;
    partial_chunks = chunks;
    return MPM_OK;
}

int mpm_surrogate::ensure_io(int64_t rows) {
    if (rows <= io_rows) return MPM_OK;
    dfree(io_x);
    dfree(io_y);
    if (hipMalloc(&io_x, sizeof(float) * (size_t)rows * MLP_IN) != hipSuccess ||
        hipMalloc(&io_y, sizeof(float) * (size_t)rows * MLP_OUT) != hipSuccess)
        return MPM_ERR_NOMEM;
    io_rows = rows;
    return MPM_OK;
}

extern "C" {

int64_t mpm_surrogate_param_count(void) { return PARAM_COUNT; }

int mpm_surrogate_create(mpm_surrogate **out, int32_t device, uint64_t seed) {
    if (!out) return MPM_ERR_ARG;
    *out = nullptr;
    if (hipSetDevice(device) != hipSuccess) return MPM_ERR_HIP;
    mpm_surrogate *s = new mpm_surrogate();
    s->device = device;
    int src = scratch_address(device, &s->scratch);
    if (src) {
        delete s;
        return src;
    }
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc(&s->params, sizeof(float) * PARAM_COUNT) != hipSuccess ||
        hipMalloc(&s->wbf16, sizeof(uint16_t) * WEIGHT_BF16_COUNT) != hipSuccess ||
        hipMalloc(&s->grad, sizeof(float) * PARAM_COUNT) != hipSuccess ||
        hipMalloc(&s->adam_m, sizeof(float) * PARAM_COUNT) != hipSuccess ||
        hipMalloc(&s->adam_v, sizeof(float) * PARAM_COUNT) != hipSuccess ||
        hipMalloc(&s->loss, sizeof(float)) != hipSuccess) {
        delete s;
        return MPM_ERR_NOMEM;
    }
    /* Glorot-uniform weights from the seed, zero biases and padding */
    std::vector<float> p(PARAM_COUNT, 0.0f);
    std::mt19937_64 rng(seed);
    auto fill = [&](int off, int in, int out, int in_used, int out_used) {
        float a = std::sqrt(6.0f / (float)(in_used + out_used));
        std::uniform_real_distribution<float> u(-a, a);
        for (int i = 0; i < in; i++)
            for (int o = 0; o < out; o++) p[off + i * out + o] = (i < in_used && o < out_used) ? u(rng) : 0.0f;
    };
    fill(OFF_W1, MLP_IN_PAD, MLP_HID, MLP_IN, MLP_HID);
    fill(OFF_W2, MLP_HID, MLP_HID, MLP_HID, MLP_HID);
    fill(OFF_W3, MLP_HID, MLP_HID, MLP_HID, MLP_HID);
    fill(OFF_W4, MLP_HID, MLP_OUT_PAD, MLP_HID, MLP_OUT);
    for (int k = 0; k < MLP_IN; k++) {
        s->norm.in_mean[k] = 0.f;
        s->norm.in_std[k] = 1.f;
        s->norm.in_lo[k] = -1e30f;
        s->norm.in_hi[k] = 1e30f;
    }
    for (int k = 0; k < MLP_OUT; k++) {
        s->norm.out_mean[k] = 0.f;
        s->norm.out_std[k] = 1.f;
    }
    int rc = mpm_surrogate_set_params(s, p.data(), PARAM_COUNT);
    if (rc) {
        delete s;
        return rc;
    }
    *out = s;
    return MPM_OK;
}

int mpm_surrogate_destroy(mpm_surrogate *s) {
    delete s;
    return MPM_OK;
}

int mpm_surrogate_set_params(mpm_surrogate *s, const float *params, int64_t count) {
    if (!s || !params || count != PARAM_COUNT) return MPM_ERR_ARG;
    if (hipSetDevice(s->device) != hipSuccess) return MPM_ERR_HIP;
    if (hipMemcpy(s->params, params, sizeof(float) * PARAM_COUNT, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(s->adam_m, 0, sizeof(float) * PARAM_COUNT) != hipSuccess ||
        hipMemset(s->adam_v, 0, sizeof(float) * PARAM_COUNT) != hipSuccess ||
        mlp_launch_pack(s->params, s->wbf16, s->stream) != hipSuccess || hipStreamSynchronize(s->stream) != hipSuccess)
        return MPM_ERR_HIP;
    s->adam_step = 0;
    return MPM_OK;
}

int mpm_surrogate_get_params(mpm_surrogate *s, float *params, int64_t count) {
    if (!s || !params || count != PARAM_COUNT) return MPM_ERR_ARG;
    if (hipSetDevice(s->device) != hipSuccess || hipStreamSynchronize(s->stream) != hipSuccess ||
        hipMemcpy(params, s->params, sizeof(float) * PARAM_COUNT, hipMemcpyDeviceToHost) != hipSuccess)
        return MPM_ERR_HIP;
    return MPM_OK;
}

int mpm_surrogate_set_normalization(mpm_surrogate *s, double stress_scale, double strain_scale,
                                    const float in_mean[13], const float in_std[13], const float in_lo[13],
                                    const float in_hi[13], const float out_mean[7], const float out_std[7]) {
    if (!s || !in_mean || !in_std || !in_lo || !in_hi || !out_mean || !out_std) return MPM_ERR_ARG;
    if (!(stress_scale > 0.0) || !(strain_scale > 0.0)) return MPM_ERR_ARG;
    for (int k = 0; k < MLP_IN; k++)
        if (!(in_std[k] > 0.0f)) return MPM_ERR_ARG;
    s->stress_scale = stress_scale;
    s->strain_scale = strain_scale;
    std::memcpy(s->norm.in_mean, in_mean, sizeof(float) * MLP_IN);
    std::memcpy(s->norm.in_std, in_std, sizeof(float) * MLP_IN);
    std::memcpy(s->norm.in_lo, in_lo, sizeof(float) * MLP_IN);
    std::memcpy(s->norm.in_hi, in_hi, sizeof(float) * MLP_IN);
    std::memcpy(s->norm.out_mean, out_mean, sizeof(float) * MLP_OUT);
    std::memcpy(s->norm.out_std, out_std, sizeof(float) * MLP_OUT);
    return MPM_OK;
}

int mpm_surrogate_forward(mpm_surrogate *s, const float *x, float *y, int64_t n) {
    if (!s || n < 0 || (n > 0 && (!x || !y))) return MPM_ERR_ARG;
    if (n == 0) return MPM_OK;
    if (hipSetDevice(s->device) != hipSuccess) return MPM_ERR_HIP;
    int rc = s->ensure_io(n);
    if (rc) return rc;
    if (hipMemcpy(s->io_x, x, sizeof(float) * (size_t)n * MLP_IN, hipMemcpyHostToDevice) != hipSuccess ||
        mlp_launch_forward(s->params, s->wbf16, s->io_x, MLP_IN, 1, s->io_y, MLP_OUT, 1, n, s->stream) != hipSuccess ||
        hipStreamSynchronize(s->stream) != hipSuccess ||
        hipMemcpy(y, s->io_y, sizeof(float) * (size_t)n * MLP_OUT, hipMemcpyDeviceToHost) != hipSuccess)
        return MPM_ERR_HIP;
    return MPM_OK;
}

int mpm_surrogate_gradient(mpm_surrogate *s, const float *x, const float *t, int64_t n, float *grad, float *loss) {
    if (!s || n <= 0 || !x || !t || !grad) return MPM_ERR_ARG;
    if (hipSetDevice(s->device) != hipSuccess) return MPM_ERR_HIP;
    int rc = s->ensure_io(n);
    if (!rc) rc = s->ensure_partial(n);
    if (rc) return rc;
    if (hipMemcpy(s->io_x, x, sizeof(float) * (size_t)n * MLP_IN, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(s->io_y, t, sizeof(float) * (size_t)n * MLP_OUT, hipMemcpyHostToDevice) != hipSuccess ||
        mlp_launch_gradient(s->params, s->wbf16, s->io_x, s->io_y, n, s->partial, s->grad, s->loss, s->stream) !=
            hipSuccess ||
        hipStreamSynchronize(s->stream) != hipSuccess ||
        hipMemcpy(grad, s->grad, sizeof(float) * PARAM_COUNT, hipMemcpyDeviceToHost) != hipSuccess)
        return MPM_ERR_HIP;
    if (loss && hipMemcpy(loss, s->loss, sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) return MPM_ERR_HIP;
    return MPM_OK;
}

int mpm_surrogate_gradient_noalloc(mpm_surrogate *s, const float *x, const float *t, int64_t n, float *grad,
                                    float *loss) {
    if (!s || n <= 0 || !x || !t || !grad) return MPM_ERR_ARG;
    if (n > MPM_SURROGATE_NOALLOC_MAX_ROWS) return MPM_ERR_RANGE;
    if (!s->scratch) return MPM_ERR_STATE;
    if (hipSetDevice(s->device) != hipSuccess) return MPM_ERR_HIP;
    /* every call writes all partials of its chunks before the reduction
     * reads them, so nothing needs clearing between calls */
    if (mlp_launch_gradient(s->params, s->wbf16, x, t, n, s->scratch, grad, loss, s->stream) != hipSuccess ||
        hipStreamSynchronize(s->stream) != hipSuccess)
        return MPM_ERR_HIP;
    return MPM_OK;
}

int mpm_surrogate_load_training_set(mpm_surrogate *s, const float *x, const float *t, int64_t n) {
    if (!s || n <= 0 || !x || !t) return MPM_ERR_ARG;
    if (hipSetDevice(s->device) != hipSuccess) return MPM_ERR_HIP;
    dfree(s->tx);
    dfree(s->tt);
    s->train_rows = 0;
    if (hipMalloc(&s->tx, sizeof(float) * (size_t)n * MLP_IN) != hipSuccess ||
        hipMalloc(&s->tt, sizeof(float) * (size_t)n * MLP_OUT) != hipSuccess)
        return MPM_ERR_NOMEM;
    if (hipMemcpy(s->tx, x, sizeof(float) * (size_t)n * MLP_IN, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(s->tt, t, sizeof(float) * (size_t)n * MLP_OUT, hipMemcpyHostToDevice) != hipSuccess)
        return MPM_ERR_HIP;
    s->train_rows = n;
    return MPM_OK;
}

int mpm_surrogate_train(mpm_surrogate *s, int32_t steps, int32_t batch, float lr, uint64_t seed, float *loss_history) {
    if (!s || steps < 0 || batch < 64 || s->train_rows < batch) return MPM_ERR_ARG;
    if (hipSetDevice(s->device) != hipSuccess) return MPM_ERR_HIP;
    int rc = s->ensure_partial(batch);
    if (rc) return rc;
    std::mt19937_64 rng(seed);
    std::uniform_int_distribution<int64_t> pick(0, (s->train_rows - batch) / 64);
    for (int st = 0; st < steps; st++) {
        /* a 64-row-aligned window of the (pre-shuffled) training set */
        int64_t off = pick(rng) * 64;
        s->adam_step++;
        if (mlp_launch_gradient(s->params, s->wbf16, s->tx + off * MLP_IN, s->tt + off * MLP_OUT, batch, s->partial,
                                s->grad, s->loss, s->stream) != hipSuccess ||
            mlp_launch_adam(s->params, s->grad, s->adam_m, s->adam_v, lr, 0.9f, 0.999f, 1e-8f, s->adam_step,
                            s->stream) != hipSuccess ||
            mlp_launch_pack(s->params, s->wbf16, s->stream) != hipSuccess)
            return MPM_ERR_HIP;
        if (loss_history && hipMemcpyAsync(loss_history + st, s->loss, sizeof(float), hipMemcpyDeviceToHost,
                                           s->stream) != hipSuccess)
            return MPM_ERR_HIP;
    }
    if (hipStreamSynchronize(s->stream) != hipSuccess) return MPM_ERR_HIP;
    return MPM_OK;
}

}  // extern "C"
