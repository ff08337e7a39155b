/*
 * mpm_device.h — per-particle FP64 numerics of the timestep, device side.
 *
 * Every expression keeps the reference's operand order and grouping so that,
 * compiled with -ffp-contract=off (no FMA fusion; IEEE division and square
 * root), each result is the same double the reference computes.
 */
#ifndef MPM_AMD_DEVICE_H
#define MPM_AMD_DEVICE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#define MPM_MAX_MATERIALS 16
#define MPM_NONE 0xFFFFFFFFu

/* Grid geometry.  Nodes: n = cells + 1 per axis; binning cells cover
 * [-1, cells] per axis, i.e. cells + 2, so every particle with at least one
 * on-grid node has a bin. */
struct GridGeom {
    double L[3];
    double h[3];
    int cells[3];
    int nn[3];
    int ec[3];
    int64_t nnodes;
    int64_t necells;
};

/* Material_BC table (Material_BC.h:13-35). */
struct MatTable {
    int n;
    int type[MPM_MAX_MATERIALS];
    double E[MPM_MAX_MATERIALS];
    double nu[MPM_MAX_MATERIALS];
    double yield[MPM_MAX_MATERIALS];
    double c0[MPM_MAX_MATERIALS];
    double c1[MPM_MAX_MATERIALS];
};

/* Quotient a / b, correctly rounded, from y = RN(1 / b) (one IEEE division
 * per divisor instead of one per quotient).  q0 = RN(a y) is within 1.5 ulp
 * of a / b; the first residual correction q1 = RN(q0 + RN(a - b q0) y)
 * brings it within one ulp (the error of the correction is below 2^-50 ulp);
 * for a faithful q1 and y within half an ulp of 1 / b, r = a - b q1 is exact
 * in one FMA and RN(q1 + r y) is RN(a / b) (Markstein's theorem,
 * P. Markstein, IBM J. Res. Dev. 34(1), 1990; Muller et al., Handbook of
 * Floating-Point Arithmetic, division by FMA).  copysign restores the sign
 * of a zero quotient.  Exact when b > 0 with 2^-100 <= b <= 2^100, and
 * a = +-0 or 2^-860 <= |a| <= 2^860 (no quotient, residual or product
 * leaves the normal range); callers keep to that. */
__host__ __device__ __forceinline__ double div_rcp(double a, double b, double y) {
    double q = a * y;
    double r = fma(-q, b, a);
    q = fma(r, y, q);
    r = fma(-q, b, a);
    return copysign(fma(r, y, q), a);
}

/* Node position along one axis: (i * L) / cells, GridPoint_Factory.h:53-59. */
__device__ __forceinline__ double node_coord(int i, double L, int cells) {
    return (double)i * L / (double)cells;
}

__device__ __forceinline__ double glm_sign(double x) {
    return (double)(0.0 < x) - (double)(x < 0.0);
}

/* One axis of GridPoint_Mediator::calculateBases_Classic
 * (GridPoint_Mediator.cpp:95-144): s = clipped 1 - |d/h|, and whether the
 * gradient along this axis is non-zero (1e-12 < |d/h| < 1 - 1e-12). */
struct AxisShape {
    double s;      /* shape factor */
    double sgn;    /* sign(d) */
    bool grad;     /* gradient component taken */
};

__device__ __forceinline__ AxisShape axis_shape(double xp, double xn, double h) {
    AxisShape a;
    double d = xp - xn;
    double q = d / h;
    double s = 1.0 - fabs(q);
    if (s < 0.0 || s > 1.0) s = 0.0;
    a.s = s;
    double aq = fabs(q);
    a.grad = (aq > 1.0e-12) && (aq < 1.0 - 1.0e-12);
    a.sgn = glm_sign(d);
    return a;
}

/* N = sx*sy*sz and grad N_k = -s_j * s_l * sign(d_k) / h_k, GridPoint_Mediator.cpp:114,138-143. */
__device__ __forceinline__ void shape_full(const AxisShape &ax, const AxisShape &ay, const AxisShape &az,
                                           const double h[3], double &N, double &gx, double &gy, double &gz) {
    N = ax.s * ay.s * az.s;
    gx = ax.grad ? (-ay.s * az.s * ax.sgn / h[0]) : 0.0;
    gy = ay.grad ? (-ax.s * az.s * ay.sgn / h[1]) : 0.0;
    gz = az.grad ? (-ax.s * ay.s * az.sgn / h[2]) : 0.0;
}

/* Base cell of a position, GridPoint_Mediator.cpp:8 (floor(x / h)), and the
 * layer range of :14-26 (tolerance 0). Positions whose floor lies outside
 * [-4, n + 4] (including NaN) are clamped there: they touch no node, as in
 * the reference where every such node index is out of range. */
__device__ __forceinline__ void base_cell(const double x, const double h, int nn, int &b, int &lo, int &hi) {
    double q = x / h;
    double f = floor(q);
    if (!(f >= -4.0)) f = -4.0;
    if (f > (double)nn + 4.0) f = (double)nn + 4.0;
    b = (int)f;
    lo = 0;
    hi = 2;
    if (((double)b + 1.0 - x / h) < 0.0) lo = 1;
    if (((x / h) - (double)b) < 0.0) hi = 1;
}

/* Outputs of ConstitutiveRelation (ConstitutiveRelation.h:22-24). */
struct CRResult {
    double dsig[6];
    double depsp[6];
    double dR;
    bool yielded;
};

/* ConstitutiveRelation_calculateIncrement_Elastic.cpp:4-17. */
__device__ __forceinline__ void trial_increment(double E, double nu, const double de[6], double dtr[6]) {
    double C = E / (1.0 + nu) / (1.0 - 2.0 * nu);
    dtr[0] = C * ((1.0 - nu) * de[0] + nu * de[1] + nu * de[2]);
    dtr[1] = C * ((1.0 - nu) * de[1] + nu * de[0] + nu * de[2]);
    dtr[2] = C * ((1.0 - nu) * de[2] + nu * de[0] + nu * de[1]);
    dtr[3] = C * ((0.5 - nu) * de[3]);
    dtr[4] = C * ((0.5 - nu) * de[4]);
    dtr[5] = C * ((0.5 - nu) * de[5]);
}

/* Radial return of ConstitutiveRelation_calculateIncrement_PerfectlyPlastic.cpp:4-76
 * (K = sigma_y / sqrt(3)) and of ..._VonMisesHardening.cpp:4-92 (K = K0 + R / sqrt(2),
 * then the explicit saturating hardening increment of :81-90). */
__device__ __forceinline__ void radial_return(int type, double E, double nu, double yield, double c0, double c1,
                                              double R, const double sig[6], const double de[6], CRResult &o) {
    const bool hardening = (type == 4);
    double dtr[6], str[6], dev[6];
    trial_increment(E, nu, de, dtr);
    for (int k = 0; k < 6; k++) str[k] = sig[k] + dtr[k];
    double hyd = 1.0 / 3.0 * (str[0] + str[1] + str[2]);
    for (int k = 0; k < 6; k++) dev[k] = str[k];
    dev[0] -= hyd;
    dev[1] -= hyd;
    dev[2] -= hyd;
    double J2 = 0.0;
    J2 += 1.0 / 6.0 * (str[0] - str[1]) * (str[0] - str[1]);
    J2 += 1.0 / 6.0 * (str[1] - str[2]) * (str[1] - str[2]);
    J2 += 1.0 / 6.0 * (str[2] - str[0]) * (str[2] - str[0]);
    J2 += str[3] * str[3];
    J2 += str[4] * str[4];
    J2 += str[5] * str[5];
    double K0 = yield / sqrt(3.0);
    double K = K0;
    if (hardening) K = K0 + R / sqrt(2.0);
    o.dR = 0.0;
    if (J2 <= K * K) {
        for (int k = 0; k < 6; k++) o.dsig[k] = dtr[k];
        for (int k = 0; k < 6; k++) o.depsp[k] = 0.0;
        o.yielded = false;
        return;
    }
    o.yielded = true;
    double scale = K / sqrt(J2);
    for (int k = 0; k < 6; k++) dev[k] *= scale;
    for (int k = 0; k < 6; k++) o.dsig[k] = dev[k] - sig[k];
    o.dsig[0] += hyd;
    o.dsig[1] += hyd;
    o.dsig[2] += hyd;
    double G = 0.5 * E / (1.0 + nu);
    double dee[6];
    dee[0] = o.dsig[0] / E - o.dsig[1] * nu / E - o.dsig[2] * nu / E;
    dee[1] = o.dsig[1] / E - o.dsig[0] * nu / E - o.dsig[2] * nu / E;
    dee[2] = o.dsig[2] / E - o.dsig[0] * nu / E - o.dsig[1] * nu / E;
    dee[3] = o.dsig[3] / G;
    dee[4] = o.dsig[4] / G;
    dee[5] = o.dsig[5] / G;
    for (int k = 0; k < 6; k++) o.depsp[k] = de[k] - dee[k];
    if (hardening) {
        double ep = 0.0;
        for (int k = 0; k < 6; k++) ep += o.depsp[k] * o.depsp[k];
        ep *= 2.0 / 3.0;
        ep = sqrt(ep);
        if ((sqrt(2.0 / 3.0) * c1 - R) > 0.0)
            o.dR = c0 * (sqrt(2.0 / 3.0) * c1 - R) * ep;
        else
            o.dR = 0.0;
    }
}

/* Material dispatch, PhysicsEngine_runSimulation_Classic_DoublePass_MPLocks.cpp:232-239:
 * 0 elastic, 3 perfectly plastic, 4 hardening, anything else perfectly plastic. */
__device__ __forceinline__ void constitutive(const MatTable &mt, int m, double R, const double sig[6],
                                             const double de[6], CRResult &o) {
    int type = mt.type[m];
    if (type == 0) {
        trial_increment(mt.E[m], mt.nu[m], de, o.dsig);
        for (int k = 0; k < 6; k++) o.depsp[k] = 0.0;
        o.dR = 0.0;
        o.yielded = false;
        return;
    }
    if (type != 4) type = 3;
    radial_return(type, mt.E[m], mt.nu[m], mt.yield[m], mt.c0[m], mt.c1[m], R, sig, de, o);
}

#endif /* MPM_AMD_DEVICE_H */
