/*
 * mpm_eso.hip — the voxel-objective smoothing filter of the BESO driver,
 * Canvas2D_CC::filterObjective_Smooth (Canvas2D_CC.cpp:302-341), on the GPU.
 *
 * The reference compares every active voxel with every active voxel
 * (O(N^2) distance evaluations, ~3e9 for the 500 x 500 L-problem canvas).
 * Only voxels closer than the radius contribute, and they contribute in
 * ascending voxel index, so each thread here visits the window of rows and
 * columns that can lie within the radius, rows ascending and columns
 * ascending inside a row: the same terms in the same order, hence the same
 * doubles.  One thread per voxel; the canvas (a few MB) stays in L2.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mpm.h"

namespace {

constexpr int ESO_BLOCK = 256;

/* Voxel centre, Canvas2D_CC.cpp:22: dvec2(index) * offset + 0.5 * dvec2(offset). */
__device__ __forceinline__ double voxel_centre(int i, double off) { return (double)i * off + 0.5 * off; }

__global__ __launch_bounds__(ESO_BLOCK) void k_eso_filter(const uint8_t *active, const uint8_t *eso,
                                                          const double *obj, int nx, int ny, double off,
                                                          double radius, int w, double *out) {
    const int64_t v = (int64_t)blockIdx.x * ESO_BLOCK + threadIdx.x;
    if (v >= (int64_t)nx * ny) return;
    if (!active[v]) {
        out[v] = obj[v];
        return;
    }
    const int x = (int)(v % nx), y = (int)(v / nx);
    const double px = voxel_centre(x, off), py = voxel_centre(y, off);
    const int y0 = max(0, y - w), y1 = min(ny - 1, y + w);
    const int x0 = max(0, x - w), x1 = min(nx - 1, x + w);
    double sum = 0.0, wsum = 0.0;
    for (int yy = y0; yy <= y1; yy++) {
        const double dy = py - voxel_centre(yy, off);
        const int64_t row = (int64_t)yy * nx;
        for (int xx = x0; xx <= x1; xx++) {
            const int64_t u = row + xx;
            if (!active[u] || !eso[u]) continue;
            const double dx = px - voxel_centre(xx, off);
            const double d = sqrt(dx * dx + dy * dy); /* glm::length(a - b) */
            if (d < radius) {
                const double wt = (radius - d) / radius;
                wsum += wt;
                sum += wt * obj[u];
            }
        }
    }
    out[v] = sum / wsum;
}

}  // namespace

extern "C" {

int mpm_eso_filter(int32_t device, const uint8_t *active, const uint8_t *eso, const double *objective, int32_t nx,
                   int32_t ny, double offset, double radius, double *out) {
    if (nx < 0 || ny < 0) return MPM_ERR_ARG;
    const int64_t nv = (int64_t)nx * ny;
    if (nv == 0) return MPM_OK;
    if (!active || !eso || !objective || !out) return MPM_ERR_ARG;
    if (!(offset > 0.0) || !(radius >= 0.0)) return MPM_ERR_ARG;
    if (hipSetDevice(device) != hipSuccess) return MPM_ERR_HIP;
    /* window half-width: a voxel w or more columns (rows) away is at least
     * (R / off + 1) offsets away, beyond the radius whatever the rounding */
    const double wd = radius / offset + 2.0;
    const int w = wd > (double)(nx > ny ? nx : ny) ? (nx > ny ? nx : ny) : (int)wd;
    uint8_t *dflags = nullptr;
    double *dobj = nullptr;
    if (hipMalloc(&dflags, 2 * (size_t)nv) != hipSuccess) return MPM_ERR_NOMEM;
    if (hipMalloc(&dobj, 16 * (size_t)nv) != hipSuccess) {
        (void)hipFree(dflags);
        return MPM_ERR_NOMEM;
    }
    int rc = MPM_OK;
    if (hipMemcpy(dflags, active, (size_t)nv, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dflags + nv, eso, (size_t)nv, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dobj, objective, 8 * (size_t)nv, hipMemcpyHostToDevice) != hipSuccess)
        rc = MPM_ERR_HIP;
    if (rc == MPM_OK) {
        const unsigned blocks = (unsigned)((nv + ESO_BLOCK - 1) / ESO_BLOCK);
        hipLaunchKernelGGL(k_eso_filter, dim3(blocks), dim3(ESO_BLOCK), 0, 0, dflags, dflags + nv, dobj, nx, ny,
                           offset, radius, w, dobj + nv);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
            hipMemcpy(out, dobj + nv, 8 * (size_t)nv, hipMemcpyDeviceToHost) != hipSuccess)
            rc = MPM_ERR_HIP;
    }
    (void)hipFree(dflags);
    (void)hipFree(dobj);
    return rc;
}

}  // extern "C"
