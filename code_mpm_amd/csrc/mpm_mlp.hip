/*
 * mpm_mlp.hip — BF16 MFMA kernels of the learned return-map surrogate:
 * inference, forward+backward (mean squared error) and Adam.
 *
 * Matrix products use gfx950's v_mfma_f32_16x16x32_bf16 (wave64) wherever
 * the reduction length is a multiple of 32, and v_mfma_f32_16x16x16_bf16
 * where it is 16 (the 16 padded inputs, the 16 padded outputs).  Operand
 * layout, lane l = 0..63, g = l / 16, r = l % 16:
 *   x32: A (16 x 32, M x K): lane holds A[r][8g + j], j = 0..7
 *        B (32 x 16, K x N): lane holds B[8g + j][r]
 *   x16: A (16 x 16):        lane holds A[r][4g + j], j = 0..3
 *        B (16 x 16):        lane holds B[4g + j][r]
 *   D (16 x 16 FP32), both:  lane holds D[4g + j][r]
 * Training: weights live in LDS as row-major BF16 with padded row strides;
 * each wave owns 64 sample rows and keeps its activations in LDS.
 * Inference: see k_mlp_infer.
 */
#include "mpm_mlp.h"

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef short s16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

/* D += A (16 x 32) B (32 x 16), operands as BF16 bit patterns */
__device__ __forceinline__ f32x4 mfma32(s16x8 a, s16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0,
                                                   0, 0);
}

__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u); /* round to nearest even */
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ __forceinline__ float silu_grad(float z) {
    float s = 1.0f / (1.0f + __expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

/* LDS strides (BF16 elements) chosen so the 16 rows a fragment touches fall
 * in different banks. */
#define WS_H 72   /* rows of 64 outputs */
#define WS_O 24   /* rows of 16 outputs */
#define AS 72     /* activation rows (64 features) */
#define XS 24     /* input rows (16 features) */

#define LDS_W1 0
#define LDS_W2 (LDS_W1 + MLP_IN_PAD * WS_H)
#define LDS_W3 (LDS_W2 + MLP_HID * WS_H)
#define LDS_W4 (LDS_W3 + MLP_HID * WS_H)
#define LDS_W_TOTAL (LDS_W4 + MLP_HID * WS_O)

/* global packed BF16 weights: row-major, unpadded, W1 W2 W3 W4 */
#define G_W1 0
#define G_W2 (G_W1 + MLP_IN_PAD * MLP_HID)
#define G_W3 (G_W2 + MLP_HID * MLP_HID)
#define G_W4 (G_W3 + MLP_HID * MLP_HID)

__device__ __forceinline__ void load_weights_lds(const uint16_t *wbf16, const float *params, uint16_t *W, float *bias) {
    for (int i = threadIdx.x; i < MLP_IN_PAD * MLP_HID; i += blockDim.x)
        W[LDS_W1 + (i / MLP_HID) * WS_H + i % MLP_HID] = wbf16[G_W1 + i];
    for (int i = threadIdx.x; i < MLP_HID * MLP_HID; i += blockDim.x) {
        W[LDS_W2 + (i / MLP_HID) * WS_H + i % MLP_HID] = wbf16[G_W2 + i];
        W[LDS_W3 + (i / MLP_HID) * WS_H + i % MLP_HID] = wbf16[G_W3 + i];
    }
    for (int i = threadIdx.x; i < MLP_HID * MLP_OUT_PAD; i += blockDim.x)
        W[LDS_W4 + (i / MLP_OUT_PAD) * WS_O + i % MLP_OUT_PAD] = wbf16[G_W4 + i];
    for (int i = threadIdx.x; i < MLP_HID; i += blockDim.x) {
        bias[i] = params[OFF_B1 + i];
        bias[64 + i] = params[OFF_B2 + i];
        bias[128 + i] = params[OFF_B3 + i];
    }
    for (int i = threadIdx.x; i < MLP_OUT_PAD; i += blockDim.x) bias[192 + i] = params[OFF_B4 + i];
}

/* acc[MT][NT] = A(16MT x 16KT) * B(16KT x 16NT) over one wave; fa(row, k)
 * and fb(k, col) return BF16 bits.  KT even: 16x16x32 steps; KT == 1: one
 * 16x16x16 step. */
template <int MT, int NT, int KT, class FA, class FB>
__device__ __forceinline__ void wave_gemm(FA fa, FB fb, f32x4 (&acc)[MT][NT]) {
    static_assert(KT == 1 || KT % 2 == 0, "reduction of 16 or a multiple of 32");
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
#pragma unroll
    for (int m = 0; m < MT; m++)
#pragma unroll
        for (int n = 0; n < NT; n++) acc[m][n] = f32x4{0.f, 0.f, 0.f, 0.f};
    if (KT == 1) {
        bf16x4 a[MT], b[NT];
#pragma unroll
        for (int m = 0; m < MT; m++)
#pragma unroll
            for (int j = 0; j < 4; j++) a[m][j] = (short)fa(m * 16 + r, 4 * g + j);
#pragma unroll
        for (int n = 0; n < NT; n++)
#pragma unroll
            for (int j = 0; j < 4; j++) b[n][j] = (short)fb(4 * g + j, n * 16 + r);
#pragma unroll
        for (int m = 0; m < MT; m++)
#pragma unroll
            for (int n = 0; n < NT; n++)
                acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[m], b[n], acc[m][n], 0, 0, 0);
        return;
    }
#pragma unroll
    for (int k = 0; k < KT / 2; k++) {
        s16x8 a[MT], b[NT];
#pragma unroll
        for (int m = 0; m < MT; m++)
#pragma unroll
            for (int j = 0; j < 8; j++) a[m][j] = (short)fa(m * 16 + r, k * 32 + 8 * g + j);
#pragma unroll
        for (int n = 0; n < NT; n++)
#pragma unroll
            for (int j = 0; j < 8; j++) b[n][j] = (short)fb(k * 32 + 8 * g + j, n * 16 + r);
#pragma unroll
        for (int m = 0; m < MT; m++)
#pragma unroll
            for (int n = 0; n < NT; n++) acc[m][n] = mfma32(a[m], b[n], acc[m][n]);
    }
}

/* One wave: rows of X (64 x 16 BF16 in lx) through the 4 layers.  Hidden
 * pre-activations go to z1..z3 and activations to h1..h3 when those are
 * non-null (training); the output (64 x 16 FP32, bias added) goes to out. */
__device__ __forceinline__ void wave_forward(const uint16_t *W, const float *bias, const uint16_t *lx, uint16_t *z1, uint16_t *h1,
                             uint16_t *z2, uint16_t *h2, uint16_t *z3, uint16_t *h3, float *out) {
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    f32x4 acc[4][4];
    /* layer 1: X (16) -> 64 */
    wave_gemm<4, 4, 1>([=](int row, int k) { return lx[row * XS + k]; },
                       [=](int k, int col) { return W[LDS_W1 + k * WS_H + col]; }, acc);
#pragma unroll
    for (int m = 0; m < 4; m++)
#pragma unroll
        for (int n = 0; n < 4; n++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                int row = m * 16 + 4 * g + j, col = n * 16 + r;
                float z = acc[m][n][j] + bias[col];
                if (z1) z1[row * AS + col] = f2bf(z);
                h1[row * AS + col] = f2bf(silu(z));
            }
    __syncthreads();
    /* layer 2 */
    wave_gemm<4, 4, 4>([=](int row, int k) { return h1[row * AS + k]; },
                       [=](int k, int col) { return W[LDS_W2 + k * WS_H + col]; }, acc);
    __syncthreads();
#pragma unroll
    for (int m = 0; m < 4; m++)
#pragma unroll
        for (int n = 0; n < 4; n++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                int row = m * 16 + 4 * g + j, col = n * 16 + r;
                float z = acc[m][n][j] + bias[64 + col];
                if (z2) z2[row * AS + col] = f2bf(z);
                h2[row * AS + col] = f2bf(silu(z));
            }
    __syncthreads();
    /* layer 3 */
    wave_gemm<4, 4, 4>([=](int row, int k) { return h2[row * AS + k]; },
                       [=](int k, int col) { return W[LDS_W3 + k * WS_H + col]; }, acc);
    __syncthreads();
#pragma unroll
    for (int m = 0; m < 4; m++)
#pragma unroll
        for (int n = 0; n < 4; n++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                int row = m * 16 + 4 * g + j, col = n * 16 + r;
                float z = acc[m][n][j] + bias[128 + col];
                if (z3) z3[row * AS + col] = f2bf(z);
                h3[row * AS + col] = f2bf(silu(z));
            }
    __syncthreads();
    /* layer 4: 64 -> 16 (7 used), linear */
    f32x4 o[4][1];
    wave_gemm<4, 1, 4>([=](int row, int k) { return h3[row * AS + k]; },
                       [=](int k, int col) { return W[LDS_W4 + k * WS_O + col]; }, o);
#pragma unroll
    for (int m = 0; m < 4; m++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            int row = m * 16 + 4 * g + j;
            out[row * MLP_OUT_PAD + r] = o[m][0][j] + bias[192 + r];
        }
    __syncthreads();
}

/* Inference in the transposed form H^T = W^T X^T.  The weights are the
 * MFMA A operand and stay in registers for the whole kernel; 16 samples form
 * the N dimension.  A layer's FP32 output tiles (D layout: in tile mt, lane
 * holds feature 16 mt + 4g + j of sample r) feed the next layer's x32 steps
 * straight from registers: step t takes tiles 2t and 2t + 1, so lane (g, r)
 * holds, at reduction position 8g + j, feature 32t + 4g + j (j < 4) or
 * 32t + 16 + 4g + (j - 4) (j >= 4); the weight fragments are loaded in the
 * same feature order.  Sample s, feature k of x is x[s * xs + k * xk]; of y,
 * y[s * ys + k * yk]. */
#define INFER_WAVES 4
__device__ __forceinline__ int x32_feature(int t, int g, int j) {
    return 32 * t + (j < 4 ? 4 * g + j : 16 + 4 * g + (j - 4));
}
__global__ __launch_bounds__(64 * INFER_WAVES) void k_mlp_infer(const float *params, const uint16_t *wbf16,
                                                                const float *x, int64_t xs, int64_t xk, float *y,
                                                                int64_t ys, int64_t yk, int64_t n,
                                                                const uint32_t *dcount) {
    if (dcount) n = min(n, (int64_t)*dcount);
    __shared__ float bias[208];
    for (int i = threadIdx.x; i < MLP_HID; i += blockDim.x) {
        bias[i] = params[OFF_B1 + i];
        bias[64 + i] = params[OFF_B2 + i];
        bias[128 + i] = params[OFF_B3 + i];
    }
    for (int i = threadIdx.x; i < MLP_OUT_PAD; i += blockDim.x) bias[192 + i] = params[OFF_B4 + i];
    __syncthreads();
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    const int wave = threadIdx.x >> 6;
    /* A fragments of W^T: layer 1 (x16) lane holds W1[4g + j][16 mt + r];
     * the x32 layers lane holds W[x32_feature(t, g, j)][16 mt + r] */
    bf16x4 w1[4];
    s16x8 w2[4][2], w3[4][2], w4[2];
#pragma unroll
    for (int mt = 0; mt < 4; mt++)
#pragma unroll
        for (int j = 0; j < 4; j++) w1[mt][j] = (short)wbf16[G_W1 + (4 * g + j) * MLP_HID + mt * 16 + r];
#pragma unroll
    for (int t = 0; t < 2; t++)
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int f = x32_feature(t, g, j);
#pragma unroll
            for (int mt = 0; mt < 4; mt++) {
                w2[mt][t][j] = (short)wbf16[G_W2 + f * MLP_HID + mt * 16 + r];
                w3[mt][t][j] = (short)wbf16[G_W3 + f * MLP_HID + mt * 16 + r];
            }
            w4[t][j] = (short)wbf16[G_W4 + f * MLP_OUT_PAD + r];
        }
    const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
    /* a layer's SiLU outputs as the next layer's two x32 B operands */
    auto activate = [&](const f32x4 (&acc)[4], int boff, s16x8 (&bo)[2]) {
#pragma unroll
        for (int mt = 0; mt < 4; mt++)
#pragma unroll
            for (int j = 0; j < 4; j++)
                bo[mt >> 1][(mt & 1) * 4 + j] = (short)f2bf(silu(acc[mt][j] + bias[boff + mt * 16 + 4 * g + j]));
    };
    for (int64_t s0 = ((int64_t)blockIdx.x * INFER_WAVES + wave) * 16; s0 < n;
         s0 += (int64_t)gridDim.x * INFER_WAVES * 16) {
        const int64_t smp = s0 + r;
        const bool live = smp < n;
        /* layer 1 B operand: X^T[4g + j][r] = x(sample s0 + r, feature 4g + j) */
        bf16x4 xb;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int k = 4 * g + j;
            xb[j] = (short)f2bf((live && k < MLP_IN) ? x[smp * xs + k * xk] : 0.f);
        }
        f32x4 acc[4];
        s16x8 hb[2];
#pragma unroll
        for (int mt = 0; mt < 4; mt++) acc[mt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(w1[mt], xb, zero, 0, 0, 0);
        activate(acc, 0, hb);
        /* layer 2 */
#pragma unroll
        for (int mt = 0; mt < 4; mt++) acc[mt] = mfma32(w2[mt][1], hb[1], mfma32(w2[mt][0], hb[0], zero));
        activate(acc, 64, hb);
        /* layer 3 */
#pragma unroll
        for (int mt = 0; mt < 4; mt++) acc[mt] = mfma32(w3[mt][1], hb[1], mfma32(w3[mt][0], hb[0], zero));
        activate(acc, 128, hb);
        /* layer 4, linear: lane holds Y^T[4g + j][r] */
        const f32x4 o = mfma32(w4[1], hb[1], mfma32(w4[0], hb[0], zero));
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int f = 4 * g + j;
            if (live && f < MLP_OUT) y[smp * ys + f * yk] = o[j] + bias[192 + f];
        }
    }
}

/* FP32 params -> BF16 row-major weights. */
__global__ void k_mlp_pack(const float *params, uint16_t *wbf16) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < MLP_IN_PAD * MLP_HID) wbf16[G_W1 + i] = f2bf(params[OFF_W1 + i]);
    if (i < MLP_HID * MLP_HID) {
        wbf16[G_W2 + i] = f2bf(params[OFF_W2 + i]);
        wbf16[G_W3 + i] = f2bf(params[OFF_W3 + i]);
    }
    if (i < MLP_HID * MLP_OUT_PAD) wbf16[G_W4 + i] = f2bf(params[OFF_W4 + i]);
}

/* Forward + backward for one 64-row chunk per 64-thread block; writes the
 * chunk's gradient (PARAM_COUNT floats) and its summed squared error. */
__global__ __launch_bounds__(64) void k_mlp_grad_chunk(const float *params, const uint16_t *wbf16, const float *x,
                                                       const float *t, int64_t n, float inv_count, float *partial,
                                                       float *chunk_loss) {
    __shared__ uint16_t W[LDS_W_TOTAL];
    __shared__ float bias[208];
    __shared__ uint16_t lx[64 * XS];
    __shared__ uint16_t z1[64 * AS], h1[64 * AS], z2[64 * AS], h2[64 * AS], z3[64 * AS], h3[64 * AS];
    __shared__ uint16_t da[64 * AS];
    __shared__ float dz[64 * 64];
    __shared__ float out[64 * MLP_OUT_PAD];
    load_weights_lds(wbf16, params, W, bias);
    const int lane = threadIdx.x, r = lane & 15, g = lane >> 4;
    const int64_t chunk = blockIdx.x;
    const int64_t row = chunk * 64 + lane;
    const bool live = row < n;
#pragma unroll
    for (int k = 0; k < MLP_IN_PAD; k++) lx[lane * XS + k] = (live && k < MLP_IN) ? f2bf(x[row * MLP_IN + k]) : (uint16_t)0;
    __syncthreads();
    wave_forward(W, bias, lx, z1, h1, z2, h2, z3, h3, out);
    float *P = partial + chunk * (int64_t)PARAM_COUNT;
    /* dL/dy for L = inv_count * sum (y - t)^2, padded outputs and dead rows 0 */
    float se = 0.f;
    for (int k = 0; k < MLP_OUT_PAD; k++) {
        float d = 0.f;
        if (live && k < MLP_OUT) {
            float e = out[lane * MLP_OUT_PAD + k] - t[row * MLP_OUT + k];
            se += e * e;
            d = 2.f * e * inv_count;
        }
        dz[lane * 64 + k] = d;
        da[lane * AS + k] = f2bf(d);
    }
    /* chunk loss: wave sum */
    for (int o = 32; o > 0; o >>= 1) se += __shfl_xor(se, o, 64);
    if (lane == 0) chunk_loss[chunk] = se;
    __syncthreads();
    /* b4, W4 */
    if (lane < MLP_OUT_PAD) {
        float s = 0.f;
        for (int i = 0; i < 64; i++) s += dz[i * 64 + lane];
        P[OFF_B4 + lane] = s;
    }
    {
        f32x4 acc[4][1];
        wave_gemm<4, 1, 4>([=](int f, int k) { return h3[k * AS + f]; }, [=](int k, int o) { return da[k * AS + o]; },
                           acc);
#pragma unroll
        for (int m = 0; m < 4; m++)
#pragma unroll
            for (int j = 0; j < 4; j++) P[OFF_W4 + (m * 16 + 4 * g + j) * MLP_OUT_PAD + r] = acc[m][0][j];
    }
    /* dH3 = dY * W4^T, dZ3 = dH3 * silu'(z3) */
#pragma unroll
    for (int layer = 2; layer >= 0; layer--) {
        const uint16_t *zl = layer == 2 ? z3 : (layer == 1 ? z2 : z1);
        const uint16_t *in = layer == 2 ? h2 : (layer == 1 ? h1 : lx); /* input of this layer */
        const int wnext = layer == 1 ? LDS_W3 : LDS_W2;                   /* W of the layer above */
        const int boff = layer == 2 ? OFF_B3 : (layer == 1 ? OFF_B2 : OFF_B1);
        const int woff = layer == 2 ? OFF_W3 : (layer == 1 ? OFF_W2 : OFF_W1);
        f32x4 acc[4][4];
        if (layer == 2)
            wave_gemm<4, 4, 1>([=](int rr, int k) { return da[rr * AS + k]; },
                               [=](int k, int c) { return W[LDS_W4 + c * WS_O + k]; }, acc);
        else
            wave_gemm<4, 4, 4>([=](int rr, int k) { return da[rr * AS + k]; },
                               [=](int k, int c) { return W[wnext + c * WS_H + k]; }, acc);
        __syncthreads();
#pragma unroll
        for (int m = 0; m < 4; m++)
#pragma unroll
            for (int nn = 0; nn < 4; nn++)
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    int rr = m * 16 + 4 * g + j, c = nn * 16 + r;
                    float d = acc[m][nn][j] * silu_grad(bf2f(zl[rr * AS + c]));
                    if (chunk * 64 + rr >= n) d = 0.f;
                    dz[rr * 64 + c] = d;
                    da[rr * AS + c] = f2bf(d);
                }
        __syncthreads();
        /* bias gradient: column sums in row order */
        {
            float s = 0.f;
            for (int i = 0; i < 64; i++) s += dz[i * 64 + lane];
            P[boff + lane] = s;
        }
        /* weight gradient: input^T * dZ */
        if (layer == 0) {
            f32x4 gw[1][4];
            wave_gemm<1, 4, 4>([=](int f, int k) { return in[k * XS + f]; },
                               [=](int k, int c) { return da[k * AS + c]; }, gw);
#pragma unroll
            for (int nn = 0; nn < 4; nn++)
#pragma unroll
                for (int j = 0; j < 4; j++) P[woff + (4 * g + j) * MLP_HID + nn * 16 + r] = gw[0][nn][j];
        } else {
            f32x4 gw[4][4];
            wave_gemm<4, 4, 4>([=](int f, int k) { return in[k * AS + f]; },
                               [=](int k, int c) { return da[k * AS + c]; }, gw);
#pragma unroll
            for (int m = 0; m < 4; m++)
#pragma unroll
                for (int nn = 0; nn < 4; nn++)
#pragma unroll
                    for (int j = 0; j < 4; j++)
                        P[woff + (m * 16 + 4 * g + j) * MLP_HID + nn * 16 + r] = gw[m][nn][j];
        }
        __syncthreads();
    }
}

/* grad[p] = sum over chunks in chunk order (deterministic); loss likewise. */
__global__ void k_mlp_reduce(const float *partial, int64_t chunks, float *grad, const float *chunk_loss, float inv_count,
                             float *loss) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < PARAM_COUNT) {
        float s = 0.f;
        for (int64_t c = 0; c < chunks; c++) s += partial[c * PARAM_COUNT + p];
        grad[p] = s;
    }
    if (p == 0 && loss) {
        float s = 0.f;
        for (int64_t c = 0; c < chunks; c++) s += chunk_loss[c];
        *loss = s * inv_count;
    }
}

__global__ void k_mlp_adam(float *params, const float *grad, float *m, float *v, float lr, float b1, float b2,
                           float eps, float c1, float c2) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= PARAM_COUNT) return;
    float gr = grad[p];
    float mm = b1 * m[p] + (1.f - b1) * gr;
    float vv = b2 * v[p] + (1.f - b2) * gr * gr;
    m[p] = mm;
    v[p] = vv;
    params[p] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

hipError_t mlp_launch_forward(const float *params, const uint16_t *wbf16, const float *x, int64_t xs, int64_t xk,
                              float *y, int64_t ys, int64_t yk, int64_t n, hipStream_t s, const uint32_t *dcount) {
    if (n <= 0) return hipSuccess;
    int64_t tiles = (n + 15) / 16;
    int64_t blocks = (tiles + INFER_WAVES - 1) / INFER_WAVES;
    unsigned grid = (unsigned)(blocks < 2048 ? blocks : 2048);
    hipLaunchKernelGGL(k_mlp_infer, dim3(grid), dim3(64 * INFER_WAVES), 0, s, params, wbf16, x, xs, xk, y, ys, yk, n,
                       dcount);
    return hipGetLastError();
}

hipError_t mlp_launch_pack(const float *params, uint16_t *wbf16, hipStream_t s) {
    hipLaunchKernelGGL(k_mlp_pack, dim3((MLP_HID * MLP_HID + 255) / 256), dim3(256), 0, s, params, wbf16);
    return hipGetLastError();
}

hipError_t mlp_launch_gradient(const float *params, const uint16_t *wbf16, const float *x, const float *t, int64_t n,
                               float *partial, float *grad, float *loss, hipStream_t s) {
    if (n <= 0) return hipErrorInvalidValue;
    int64_t chunks = (n + 63) / 64;
    float inv = 1.0f / (float)(n * MLP_OUT);
    float *chunk_loss = partial + chunks * (int64_t)PARAM_COUNT;
    hipLaunchKernelGGL(k_mlp_grad_chunk, dim3((unsigned)chunks), dim3(64), 0, s, params, wbf16, x, t, n, inv, partial,
                       chunk_loss);
    hipLaunchKernelGGL(k_mlp_reduce, dim3((PARAM_COUNT + 255) / 256), dim3(256), 0, s, partial, chunks, grad,
                       chunk_loss, 1.0f / (float)(n * MLP_OUT), loss);
    return hipGetLastError();
}

hipError_t mlp_launch_adam(float *params, const float *grad, float *m, float *v, float lr, float beta1, float beta2,
                           float eps, int step, hipStream_t s) {
    float c1 = 1.f - powf(beta1, (float)step), c2 = 1.f - powf(beta2, (float)step);
    hipLaunchKernelGGL(k_mlp_adam, dim3((PARAM_COUNT + 255) / 256), dim3(256), 0, s, params, grad, m, v, lr, beta1,
                       beta2, eps, c1, c2);
    return hipGetLastError();
}
