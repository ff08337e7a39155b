/*
 * mpm_oracle.cpp — CPU FP64 restatement of the reference's Classic MPM
 * DoublePass timestep.  TEST INFRASTRUCTURE ONLY: it is the checker for the
 * HIP path (tests, smoke) and the timed CPU baseline (bench.py).  Nothing in
 * the product links or calls it.
 *
 * It keeps the reference's structure on purpose, because it is also the
 * baseline that the GPU is measured against:
 *   - heap-allocated AoS particles and nodes behind pointer vectors
 *     (PhysicsEngine.h:89-91);
 *   - a per-step rebuild of each particle's adjacent-grid-point list
 *     (PhysicsEngine_runSimulation_Classic_DoublePass_MPLocks.cpp:23-47);
 *   - one OpenMP lock per node guarding the two particle->grid scatters
 *     (same file :66/86 and :556/562) when more than one thread runs;
 *   - the reference's phase order (same file :462-790).
 * Run with one thread it sums every node quantity in ascending particle
 * order, which is the order the reference produces with one thread.
 *
 * Arithmetic follows the reference expression by expression (operand order
 * and grouping as C++ evaluates them).  The reference does its vector algebra
 * with glm, which is not vendored; the glm operations used here are restated
 * from glm's documented semantics (column-major mat3 m[col][row],
 * outerProduct(c, r)[i][j] = c[j] * r[i], dot = x*x + y*y + z*z left to
 * right, normalize(v) = v * (1 / sqrt(dot(v, v))), sign(x) = (0 < x) - (x < 0),
 * mat3 product summed over k = 0,1,2 left to right, 3x3 determinant by
 * cofactors of the first column).  Those glm bit-level choices are "parity
 * unpinned" against real glm: no glm is available to check them.
 *
 * Build: oracle/Makefile (g++ -O2 -fopenmp -ffp-contract=off, no FMA).
 */
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <utility>
#include <vector>
#include <omp.h>

#include "../include/mpm.h"

namespace {

struct V3 {
    double x = 0.0, y = 0.0, z = 0.0;
};
static inline V3 v3(double x, double y, double z) { V3 r; r.x = x; r.y = y; r.z = z; return r; }

/* glm::dmat3, column-major: m[col][row]. */
struct M3 {
    double m[3][3];
};
static inline M3 m3_identity() {
    M3 r;
    for (int c = 0; c < 3; c++)
        for (int k = 0; k < 3; k++) r.m[c][k] = (c == k) ? 1.0 : 0.0;
    return r;
}
static inline M3 m3_zero() {
    M3 r;
    for (int c = 0; c < 3; c++)
        for (int k = 0; k < 3; k++) r.m[c][k] = 0.0;
    return r;
}
/* glm operator*(mat3, mat3): Result[c][r] = m1[0][r]*m2[c][0] + m1[1][r]*m2[c][1] + m1[2][r]*m2[c][2]. */
static inline M3 m3_mul(const M3 &a, const M3 &b) {
    M3 r;
    for (int c = 0; c < 3; c++)
        for (int k = 0; k < 3; k++)
            r.m[c][k] = a.m[0][k] * b.m[c][0] + a.m[1][k] * b.m[c][1] + a.m[2][k] * b.m[c][2];
    return r;
}
/* glm determinant(mat3). */
static inline double m3_det(const M3 &m) {
    return +m.m[0][0] * (m.m[1][1] * m.m[2][2] - m.m[2][1] * m.m[1][2])
           - m.m[1][0] * (m.m[0][1] * m.m[2][2] - m.m[2][1] * m.m[0][2])
           + m.m[2][0] * (m.m[0][1] * m.m[1][2] - m.m[1][1] * m.m[0][2]);
}
static inline M3 m3_transpose(const M3 &a) {
    M3 r;
    for (int c = 0; c < 3; c++)
        for (int k = 0; k < 3; k++) r.m[c][k] = a.m[k][c];
    return r;
}

/* getFactorial, ConstitutiveRelation_MatrixAlgebra.cpp:3-11. */
static double series_factorial(int i) {
    double r = 1.0;
    for (int k = 1; k <= i; k++) r *= k;
    return r;
}
/* getPower, ConstitutiveRelation_MatrixAlgebra.cpp:12-23: the identity
 * multiplied on the right by A, p times (I * A is formed, not taken as A). */
static M3 m3_power(const M3 &A, int p) {
    M3 r = m3_identity();
    for (int k = 1; k <= p; k++) r = m3_mul(r, A);
    return r;
}
/* getExp, ConstitutiveRelation_MatrixAlgebra.cpp:25-37: the sum over
 * t = 0..8 of A^t / t!, added term by term, element by element. */
static M3 m3_exp_series(const M3 &A) {
    M3 r = m3_zero();
    for (int t = 0; t <= 8; t++) {
        M3 P = m3_power(A, t);
        double f = series_factorial(t);
        for (int c = 0; c < 3; c++)
            for (int k = 0; k < 3; k++) r.m[c][k] += P.m[c][k] / f;
    }
    return r;
}
/* getLog, ConstitutiveRelation_MatrixAlgebra.cpp:39-53: the sum over
 * t = 1..8 of ((A - I)^t * (1/t)) * (-1)^(t+1). */
static M3 m3_log_series(const M3 &A) {
    M3 I = m3_identity(), D;
    for (int c = 0; c < 3; c++)
        for (int k = 0; k < 3; k++) D.m[c][k] = A.m[c][k] - I.m[c][k];
    M3 r = m3_zero();
    for (int t = 1; t <= 8; t++) {
        M3 P = m3_power(D, t);
        double s = 1.0 / t;
        double sg = std::pow(-1.0, t + 1);
        for (int c = 0; c < 3; c++)
            for (int k = 0; k < 3; k++) r.m[c][k] += P.m[c][k] * s * sg;
    }
    return r;
}
/* 0.5 * getLog(A * transpose(A)), the strain measure of DP:323 and DP:342-343. */
static M3 m3_half_log_aat(const M3 &A) {
    M3 r = m3_log_series(m3_mul(A, m3_transpose(A)));
    for (int c = 0; c < 3; c++)
        for (int k = 0; k < 3; k++) r.m[c][k] = 0.5 * r.m[c][k];
    return r;
}
/* Voigt vector with engineering shears, DP:325-330 / DP:348-361. */
static void m3_voigt(const M3 &m, double v[6]) {
    v[0] = m.m[0][0];
    v[1] = m.m[1][1];
    v[2] = m.m[2][2];
    v[3] = m.m[0][1] + m.m[1][0];
    v[4] = m.m[1][2] + m.m[2][1];
    v[5] = m.m[2][0] + m.m[0][2];
}

static inline double glm_sign(double x) { return (double)(0.0 < x) - (double)(x < 0.0); }
static inline double glm_dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

struct Material {
    int type = MPM_MAT_ELASTIC;
    double density = 0, E = 0, nu = 0, yield = 0, viscosity = 0, c0 = 0, c1 = 0;
};

/* AGPstruct, MaterialPoint_BC.h:18-23. */
struct AGP {
    unsigned int index = 0;
    double N = 0.0;
    V3 g;
};

/* MaterialPoint_BC, MaterialPoint_BC.h:25-85 (fields on the timestep only). */
struct Particle {
    int material = -1;
    bool dc = false; /* on the displacement-control list */
    uint8_t dc_axes = 0;
    float dc_mult = 1.0f;
    double m = 0.0, V = 0.0, V0 = 0.0;
    V3 x, v, fext;
    M3 F = m3_identity();
    M3 Fe = m3_identity(); /* d33_DeformationGradient_Elastic, large-strain pass only */
    double eps[6] = {0, 0, 0, 0, 0, 0};
    double sig[6] = {0, 0, 0, 0, 0, 0};
    double epsp[6] = {0, 0, 0, 0, 0, 0};
    double Ws = 0.0, Wmax = 0.0, Wti = 0.0, Wp = 0.0, R = 0.0;
    std::vector<AGP> agp;
};

/* GridPoint, GridPoint.h:6-45. */
struct Node {
    bool active = false;
    uint8_t fixed = 0;
    V3 pos;
    double friction = 0.0, m = 0.0;
    V3 p, v, f, ft;
};

/* ConstitutiveRelation outputs (ConstitutiveRelation.h:22-24); a fresh
 * object per particle starts them at zero (DP:204). */
struct CR {
    double dsig[6] = {0, 0, 0, 0, 0, 0};
    double depsp[6] = {0, 0, 0, 0, 0, 0};
    double dR = 0.0;
    bool yielded = false;
};

/* ConstitutiveRelation_calculateIncrement_Elastic.cpp:4-17. */
static void cr_elastic(CR &o, const Material &mt, const double de[6]) {
    double E = mt.E, nu = mt.nu;
    double C = E / (1.0 + nu) / (1.0 - 2.0 * nu);
    o.dsig[0] = C * ((1.0 - nu) * de[0] + nu * de[1] + nu * de[2]);
    o.dsig[1] = C * ((1.0 - nu) * de[1] + nu * de[0] + nu * de[2]);
    o.dsig[2] = C * ((1.0 - nu) * de[2] + nu * de[0] + nu * de[1]);
    o.dsig[3] = C * ((0.5 - nu) * de[3]);
    o.dsig[4] = C * ((0.5 - nu) * de[4]);
    o.dsig[5] = C * ((0.5 - nu) * de[5]);
}

/* Shared body of the radial return: trial stress, J2, and the scaled
 * deviator, ConstitutiveRelation_calculateIncrement_PerfectlyPlastic.cpp:4-76
 * and ConstitutiveRelation_calculateIncrement_VonMisesHardening.cpp:4-92.
 * radius = K0 (perfectly plastic) or K0 + R/sqrt(2) (hardening).  Returns
 * true when the trial state is outside the yield surface. */
static bool cr_radial(CR &o, const Material &mt, const double sig[6], const double de[6], double radius_k0,
                      double radius_r, bool hardening) {
    double E = mt.E, nu = mt.nu;
    double dtr[6], str[6];
    double C = E / (1.0 + nu) / (1.0 - 2.0 * nu);
    dtr[0] = C * ((1.0 - nu) * de[0] + nu * de[1] + nu * de[2]);
    dtr[1] = C * ((1.0 - nu) * de[1] + nu * de[0] + nu * de[2]);
    dtr[2] = C * ((1.0 - nu) * de[2] + nu * de[0] + nu * de[1]);
    dtr[3] = C * ((0.5 - nu) * de[3]);
    dtr[4] = C * ((0.5 - nu) * de[4]);
    dtr[5] = C * ((0.5 - nu) * de[5]);
    for (int k = 0; k < 6; k++) str[k] = sig[k] + dtr[k];
    double hyd = 1.0 / 3.0 * (str[0] + str[1] + str[2]);
    double dev[6];
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
    double K = hardening ? (radius_k0 + radius_r) : radius_k0;
    if (J2 <= K * K) {
        for (int k = 0; k < 6; k++) o.dsig[k] = dtr[k];
        for (int k = 0; k < 6; k++) o.depsp[k] = 0.0;
        if (hardening) o.dR = 0.0;
        return false;
    }
    for (int k = 0; k < 6; k++) dev[k] *= K / std::sqrt(J2);
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
    return true;
}

/* ConstitutiveRelation_calculateIncrement_PerfectlyPlastic.cpp:4-76. */
static void cr_plastic(CR &o, const Material &mt, const double sig[6], const double de[6]) {
    double K0 = mt.yield / std::sqrt(3.0);
    o.yielded = cr_radial(o, mt, sig, de, K0, 0.0, false);
}

/* ConstitutiveRelation_calculateIncrement_VonMisesHardening.cpp:4-92. */
static void cr_vmh(CR &o, const Material &mt, double R, const double sig[6], const double de[6]) {
    double K0 = mt.yield / std::sqrt(3.0);
    double Rr = R / std::sqrt(2.0);
    o.yielded = cr_radial(o, mt, sig, de, K0, Rr, true);
    if (!o.yielded) return;
    double ep = 0.0;
    for (int k = 0; k < 6; k++) ep += o.depsp[k] * o.depsp[k];
    ep *= 2.0 / 3.0;
    ep = std::sqrt(ep);
    if ((std::sqrt(2.0 / 3.0) * mt.c1 - R) > 0.0)
        o.dR = mt.c0 * (std::sqrt(2.0 / 3.0) * mt.c1 - R) * ep;
    else
        o.dR = 0.0;
}

/* Dispatch, PhysicsEngine_runSimulation_Classic_DoublePass_MPLocks.cpp:232-239. */
static void cr_dispatch(CR &o, const Material &mt, double R, const double sig[6], const double de[6]) {
    if (mt.type == MPM_MAT_ELASTIC)
        cr_elastic(o, mt, de);
    else if (mt.type == MPM_MAT_PLASTIC)
        cr_plastic(o, mt, sig, de);
    else if (mt.type == MPM_MAT_VONMISESHARDENING)
        cr_vmh(o, mt, R, sig, de);
    else
        cr_plastic(o, mt, sig, de);
}

/* runSimulation_G2P_P2_LargeStrain after the node gather, DP:303-435, for
 * one particle with velocity gradient L; mt is NULL for a particle with no
 * material (DP:373-377: its stress was already reset, nothing else
 * changes).  Returns false for that case. */
static bool large_strain_point(Particle *p, const Material *mt, const M3 &L, double dt, CR &cr) {
    for (int k = 0; k < 6; k++) p->sig[k] = 0.0; /* DP:305-306 */
    M3 I = m3_identity(), dF;
    for (int c = 0; c < 3; c++)
        for (int r = 0; r < 3; r++) dF.m[c][r] = I.m[c][r] + L.m[c][r] * dt; /* DP:308 */
    double etot[6], etr[6], einc[6];
    m3_voigt(m3_half_log_aat(p->F), etot);          /* DP:320-331, F before the step */
    M3 FeT = m3_mul(dF, p->Fe);                     /* DP:337-339 */
    m3_voigt(m3_half_log_aat(FeT), etr);            /* DP:342, DP:348-353 */
    m3_voigt(m3_half_log_aat(dF), einc);            /* DP:343, DP:355-360 */
    if (!mt) return false;
    cr_dispatch(cr, *mt, p->R, p->sig, etr);        /* DP:379-386, from zero stress */
    double eel[6];
    for (int k = 0; k < 6; k++) eel[k] = etr[k] - cr.depsp[k]; /* DP:390-392 */
    M3 Sel = m3_zero();
    Sel.m[0][0] = eel[0];
    Sel.m[1][1] = eel[1];
    Sel.m[2][2] = eel[2];
    Sel.m[0][1] = 0.5 * eel[3];
    Sel.m[1][0] = 0.5 * eel[3];
    Sel.m[1][2] = 0.5 * eel[4];
    Sel.m[2][1] = 0.5 * eel[4];
    Sel.m[2][0] = 0.5 * eel[5];
    Sel.m[0][2] = 0.5 * eel[5];
    p->Fe = m3_exp_series(Sel);                     /* DP:410 */
    p->F = m3_mul(p->F, dF);                        /* DP:412 */
    double det = m3_det(p->F);                      /* DP:414-415 */
    p->V = det * p->V0;
    for (int k = 0; k < 6; k++) p->eps[k] = etot[k];                     /* DP:417-418 */
    for (int k = 0; k < 6; k++) p->sig[k] = (1.0 / det) * cr.dsig[k];   /* DP:423-424 */
    for (int k = 0; k < 6; k++) p->epsp[k] += cr.depsp[k];              /* DP:426-427 */
    for (int k = 0; k < 6; k++) p->Ws += p->sig[k] * einc[k] * p->V;    /* DP:429-430 */
    for (int k = 0; k < 6; k++) p->Wp += p->sig[k] * cr.depsp[k] * p->V; /* DP:432-433 */
    p->R += cr.dR;                                  /* DP:435 */
    return true;
}

struct Timeline {
    std::vector<double> t;
    std::vector<V3> v;
    /* TimeLine::getVelocity, TimeLine.h:33-46: the last matching interval wins. */
    V3 velocity(double time) const {
        V3 r = v3(0.0, 0.0, 0.0);
        if (t.empty()) return r;
        for (size_t i = 0; i + 1 < t.size(); i++) {
            if (t[i] <= time && time < t[i + 1]) {
                double s = (time - t[i]) / (t[i + 1] - t[i]);
                r.x = (1.0 - s) * v[i].x + (s)*v[i + 1].x;
                r.y = (1.0 - s) * v[i].y + (s)*v[i + 1].y;
                r.z = (1.0 - s) * v[i].z + (s)*v[i + 1].z;
            }
        }
        return r;
    }
};

}  // namespace

struct oracle_engine {
    V3 length, h;
    int cells[3] = {0, 0, 0};
    int nn[3] = {0, 0, 0};
    std::vector<Node *> nodes;
    std::vector<omp_lock_t *> locks;
    std::vector<Material> materials;
    std::vector<Particle *> particles;
    std::vector<int64_t> dc_list;
    std::vector<int64_t> monitor[MPM_MON_COUNT];
    Timeline timeline;
    double time = 0.0, dt_max = 1.0e-3, time_end = 10.0;
    double damping = 0.0, mass_min = 0.0;
    int64_t cycle = 0;
    mpm_health health{};
    std::string err;

    ~oracle_engine() { clear_nodes(); clear_particles(); }
    void clear_nodes() {
        for (Node *n : nodes) delete n;
        nodes.clear();
        for (omp_lock_t *l : locks) {
            omp_destroy_lock(l);
            delete l;
        }
        locks.clear();
    }
    void clear_particles() {
        for (Particle *p : particles) delete p;
        particles.clear();
    }

    /* GridPoint_Mediator::findAdjacentGridPoints, GridPoint_Mediator.cpp:4-54. */
    void find_adjacent(V3 x, std::vector<unsigned int> &out) const {
        out.clear();
        double q[3] = {x.x / h.x, x.y / h.y, x.z / h.z};
        double hh[3] = {h.x, h.y, h.z};
        double xx[3] = {x.x, x.y, x.z};
        int b[3];
        for (int a = 0; a < 3; a++) {
            double f = std::floor(q[a]);
            /* out-of-int-range and NaN positions touch no node (the int
               conversion would be undefined); clamping keeps them off-grid */
            if (!(f >= -4.0)) f = -4.0;
            if (f > (double)nn[a] + 4.0) f = (double)nn[a] + 4.0;
            b[a] = (int)f;
        }
        int start[3] = {0, 0, 0}, end[3] = {2, 2, 2};
        double tol = 0.0;
        for (int a = 0; a < 3; a++)
            if (((double)b[a] + 1.0 - xx[a] / hh[a]) < tol) start[a] = 1;
        for (int a = 0; a < 3; a++)
            if (((xx[a] / hh[a]) - (double)b[a]) < tol) end[a] = 1;
        for (int ix = start[0]; ix < end[0]; ix++)
            for (int iy = start[1]; iy < end[1]; iy++)
                for (int iz = start[2]; iz < end[2]; iz++) {
                    int n[3] = {b[0] + ix, b[1] + iy, b[2] + iz};
                    if (n[0] < 0 || n[1] < 0 || n[2] < 0) continue;
                    if (n[0] > nn[0] - 1 || n[1] > nn[1] - 1 || n[2] > nn[2] - 1) continue;
                    out.push_back((unsigned int)(n[0] + n[1] * nn[0] + n[2] * (nn[0] * nn[1])));
                }
    }

    /* GridPoint_Mediator::calculateBases_Classic, GridPoint_Mediator.cpp:95-144. */
    void bases(V3 xp, V3 xn, double &N, V3 &g) const {
        V3 d = v3(xp.x - xn.x, xp.y - xn.y, xp.z - xn.z);
        double sx = 1.0 - std::fabs(d.x / h.x);
        double sy = 1.0 - std::fabs(d.y / h.y);
        double sz = 1.0 - std::fabs(d.z / h.z);
        if (sx < 0.0 || sx > 1.0) sx = 0.0;
        if (sy < 0.0 || sy > 1.0) sy = 0.0;
        if (sz < 0.0 || sz > 1.0) sz = 0.0;
        N = sx * sy * sz;
        g = v3(0.0, 0.0, 0.0);
        double tol = 1.0e-12;
        if (std::fabs(d.x / h.x) > tol && std::fabs(d.x / h.x) < 1.0 - tol) g.x = -sy * sz * glm_sign(d.x) / h.x;
        if (std::fabs(d.y / h.y) > tol && std::fabs(d.y / h.y) < 1.0 - tol) g.y = -sx * sz * glm_sign(d.y) / h.y;
        if (std::fabs(d.z / h.z) > tol && std::fabs(d.z / h.z) < 1.0 - tol) g.z = -sx * sy * glm_sign(d.z) / h.z;
    }

    /* runSimulation_ResetGrid, DP:4-22. */
    void reset_grid() {
#pragma omp for
        for (long i = 0; i < (long)nodes.size(); i++) {
            Node *n = nodes[i];
            if (!n->active) continue;
            n->active = false;
            n->m = 0.0;
            n->v = v3(0, 0, 0);
            n->p = v3(0, 0, 0);
            n->f = v3(0, 0, 0);
            n->ft = v3(0, 0, 0);
        }
    }

    /* runSimulation_FindAGPs, DP:23-47. */
    void find_agps() {
        std::vector<unsigned int> adj;
#pragma omp for
        for (long i = 0; i < (long)particles.size(); i++) {
            Particle *p = particles[i];
            p->agp.clear();
            find_adjacent(p->x, adj);
            for (size_t k = 0; k < adj.size(); k++) {
                Node *n = nodes[adj[k]];
                n->active = true;
                AGP a;
                a.index = adj[k];
                bases(p->x, n->pos, a.N, a.g);
                p->agp.push_back(a);
            }
        }
    }

    /* runSimulation_M2G, DP:48-89 (DPC:110-150).  skip_dc leaves out the
     * displacement-controlled particles (SinglePass). */
    void m2g(int nthreads, bool skip_dc = false) {
#pragma omp for
        for (long i = 0; i < (long)particles.size(); i++) {
            Particle *p = particles[i];
            if (skip_dc && p->dc) continue;
            for (size_t k = 0; k < p->agp.size(); k++) {
                const AGP &a = p->agp[k];
                Node *n = nodes[a.index];
                double N = a.N;
                V3 g = a.g;
                if (nthreads > 1) omp_set_lock(locks[a.index]);
                n->m += N * p->m;
                n->p.x += N * (p->m * p->v.x);
                n->p.y += N * (p->m * p->v.y);
                n->p.z += N * (p->m * p->v.z);
                double V = p->V;
                n->f.x += -V * (g.x * p->sig[0] + g.y * p->sig[3] + g.z * p->sig[5]);
                n->f.y += -V * (g.y * p->sig[1] + g.x * p->sig[3] + g.z * p->sig[4]);
                n->f.z += -V * (g.z * p->sig[2] + g.x * p->sig[5] + g.y * p->sig[4]);
                n->f.x += N * p->fext.x;
                n->f.y += N * p->fext.y;
                n->f.z += N * p->fext.z;
                if (nthreads > 1) omp_unset_lock(locks[a.index]);
            }
        }
    }

    /* ---- cell summation order (the engine's MPM_ORDER_CELL) ----
     * Node sums are taken as sum over the node's 8 neighbour cells k = 0..7
     * (x outer, z inner; cell = node - 1 + k per axis) of per-cell partial
     * sums over the cell's particles in ascending index.  Same terms as the
     * reference's scatter, another association. */
    int order = MPM_ORDER_PARTICLE;
    std::vector<uint32_t> bin_start, bin_ids;
    int ec[3] = {0, 0, 0};

    void build_bins() {
        for (int a = 0; a < 3; a++) ec[a] = cells[a] + 2;
        size_t necells = (size_t)ec[0] * ec[1] * ec[2];
        std::vector<long> key(particles.size(), -1);
        bin_start.assign(necells + 1, 0);
        for (size_t i = 0; i < particles.size(); i++) {
            const V3 &x = particles[i]->x;
            double q[3] = {x.x / h.x, x.y / h.y, x.z / h.z};
            long e[3];
            bool ok = true;
            for (int a = 0; a < 3; a++) {
                double f = std::floor(q[a]);
                if (!(f >= -1.0 && f <= (double)cells[a])) ok = false;
                else e[a] = (long)f + 1;
            }
            if (!ok) continue;
            key[i] = e[0] + ec[0] * (e[1] + (long)ec[1] * e[2]);
            bin_start[key[i] + 1]++;
        }
        for (size_t c = 0; c < necells; c++) bin_start[c + 1] += bin_start[c];
        bin_ids.assign(bin_start[necells], 0);
        std::vector<uint32_t> fill(bin_start.begin(), bin_start.end() - 1);
        for (size_t i = 0; i < particles.size(); i++)
            if (key[i] >= 0) bin_ids[fill[key[i]]++] = (uint32_t)i; /* ascending index within a cell */
    }

    const AGP *agp_for(const Particle *p, unsigned node) const {
        for (const AGP &a : p->agp)
            if (a.index == node) return &a;
        return nullptr;
    }

    /* The node's neighbour cells in the order the node adds their partial
     * sums, and where that order closes a group (the node adds each group's
     * sum, itself summed from 0 in order).  Cell order: k = 0..7 (x outer,
     * z inner; binning cell = node + k per axis), each cell its own group.
     * Tile order: the cells grouped by summation tile (MPM_SUM_TILE_X x Y x Z
     * binning cells) in ascending tile index, and within a tile in parity
     * class order, class = 4 (e_x & 1) + 2 (e_y & 1) + (e_z & 1).  Pair
     * order: cells k and k + 4 (the lower and upper x cell of the same y, z
     * offsets) form a group, lower first, and the groups follow k = 0..3. */
    void neighbour_sequence(long ni, long cell[8], bool last_of_group[8]) const {
        int ix = (int)(ni % nn[0]), iy = (int)((ni / nn[0]) % nn[1]), iz = (int)(ni / ((long)nn[0] * nn[1]));
        long key[8];
        for (int k = 0; k < 8; k++) {
            int e[3] = {ix + (k >> 2), iy + ((k >> 1) & 1), iz + (k & 1)};
            cell[k] = e[0] + ec[0] * (e[1] + (long)ec[1] * e[2]);
            if (order == MPM_ORDER_TILE) {
                int ntx = (ec[0] + MPM_SUM_TILE_X - 1) / MPM_SUM_TILE_X;
                int nty = (ec[1] + MPM_SUM_TILE_Y - 1) / MPM_SUM_TILE_Y;
                long tile = e[0] / MPM_SUM_TILE_X +
                            ntx * ((long)(e[1] / MPM_SUM_TILE_Y) + (long)nty * (e[2] / MPM_SUM_TILE_Z));
                int cls = 4 * (e[0] & 1) + 2 * (e[1] & 1) + (e[2] & 1);
                key[k] = tile * 8 + cls;
            } else if (order == MPM_ORDER_PAIR) {
                key[k] = 2 * (k & 3) + (k >> 2);
            } else {
                key[k] = k;
            }
        }
        /* the 8 classes are distinct, so the keys are too */
        for (int a = 1; a < 8; a++)
            for (int b = a; b > 0 && key[b - 1] > key[b]; b--) {
                std::swap(key[b - 1], key[b]);
                std::swap(cell[b - 1], cell[b]);
            }
        for (int k = 0; k < 8; k++)
            last_of_group[k] = order == MPM_ORDER_PAIR ? (k & 1) != 0
                                                       : (order != MPM_ORDER_TILE || k == 7 || key[k] / 8 != key[k + 1] / 8);
    }

    void m2g_cells(bool skip_dc = false) {
#pragma omp for
        for (long ni = 0; ni < (long)nodes.size(); ni++) {
            Node *n = nodes[ni];
            if (!n->active) continue;
            long seq[8];
            bool close[8];
            neighbour_sequence(ni, seq, close);
            double gm = 0.0;
            V3 gp, gf;
            for (int k = 0; k < 8; k++) {
                long c = seq[k];
                if (bin_start[c] != bin_start[c + 1]) {
                    double sm = 0.0;
                    V3 sp, sf;
                    for (uint32_t q = bin_start[c]; q < bin_start[c + 1]; q++) {
                        const Particle *p = particles[bin_ids[q]];
                        if (skip_dc && p->dc) continue;
                        const AGP *a = agp_for(p, (unsigned)ni);
                        double N = a->N;
                        V3 g = a->g;
                        sm += N * p->m;
                        sp.x += N * (p->m * p->v.x);
                        sp.y += N * (p->m * p->v.y);
                        sp.z += N * (p->m * p->v.z);
                        double V = p->V;
                        sf.x += -V * (g.x * p->sig[0] + g.y * p->sig[3] + g.z * p->sig[5]);
                        sf.y += -V * (g.y * p->sig[1] + g.x * p->sig[3] + g.z * p->sig[4]);
                        sf.z += -V * (g.z * p->sig[2] + g.x * p->sig[5] + g.y * p->sig[4]);
                        sf.x += N * p->fext.x;
                        sf.y += N * p->fext.y;
                        sf.z += N * p->fext.z;
                    }
                    gm += sm;
                    gp.x += sp.x; gp.y += sp.y; gp.z += sp.z;
                    gf.x += sf.x; gf.y += sf.y; gf.z += sf.z;
                }
                if (close[k]) {
                    n->m += gm;
                    n->p.x += gp.x; n->p.y += gp.y; n->p.z += gp.z;
                    n->f.x += gf.x; n->f.y += gf.y; n->f.z += gf.z;
                    gm = 0.0;
                    gp = V3();
                    gf = V3();
                }
            }
        }
    }

    void remap_cells(bool skip_dc = false) {
#pragma omp for
        for (long ni = 0; ni < (long)nodes.size(); ni++) {
            Node *n = nodes[ni];
            if (!n->active || !(n->m > mass_min)) continue;
            long seq[8];
            bool close[8];
            neighbour_sequence(ni, seq, close);
            V3 gv;
            for (int k = 0; k < 8; k++) {
                long c = seq[k];
                if (bin_start[c] != bin_start[c + 1]) {
                    V3 sv;
                    for (uint32_t q = bin_start[c]; q < bin_start[c + 1]; q++) {
                        const Particle *p = particles[bin_ids[q]];
                        if (skip_dc && p->dc) continue;
                        double N = agp_for(p, (unsigned)ni)->N;
                        sv.x += N * (p->m * p->v.x) / n->m;
                        sv.y += N * (p->m * p->v.y) / n->m;
                        sv.z += N * (p->m * p->v.z) / n->m;
                    }
                    gv.x += sv.x; gv.y += sv.y; gv.z += sv.z;
                }
                if (close[k]) {
                    n->v.x += gv.x; n->v.y += gv.y; n->v.z += gv.z;
                    gv = V3();
                }
            }
        }
    }

    /* runSimulation_DisplacementControl, DP:143-174. */
    void displacement_control() {
#pragma omp for
        for (long i = 0; i < (long)dc_list.size(); i++) {
            Particle *p = particles[dc_list[i]];
            for (size_t k = 0; k < p->agp.size(); k++) {
                Node *n = nodes[p->agp[k].index];
                if (p->dc_axes & 1) {
                    n->v.x = p->v.x;
                    n->p.x = n->m * p->v.x;
                    n->f.x = 0.0;
                }
                if (p->dc_axes & 2) {
                    n->v.y = p->v.y;
                    n->p.y = n->m * p->v.y;
                    n->f.y = 0.0;
                }
                if (p->dc_axes & 4) {
                    n->v.z = p->v.z;
                    n->p.z = n->m * p->v.z;
                    n->f.z = 0.0;
                }
            }
        }
    }

    /* runSimulation_IntegrateGrid, DP:90-142. */
    void integrate_grid(double dt) {
#pragma omp for
        for (long i = 0; i < (long)nodes.size(); i++) {
            Node *n = nodes[i];
            if (!n->active) continue;
            if (std::sqrt(glm_dot(n->p, n->p)) > 1.0e-24) {
                double inv = 1.0 / std::sqrt(glm_dot(n->p, n->p));
                V3 nrm = v3(n->p.x * inv, n->p.y * inv, n->p.z * inv);
                double s = damping * std::fabs(glm_dot(n->f, nrm));
                n->f.x -= s * nrm.x;
                n->f.y -= s * nrm.y;
                n->f.z -= s * nrm.z;
            }
            if (n->fixed & 1) {
                V3 pt = v3(n->p.x + n->f.x * dt, n->p.y + n->f.y * dt, n->p.z + n->f.z * dt);
                if (std::fabs(pt.y * dt) > 1.0e-24 && n->f.x > 0.0) {
                    double ffy = n->friction * glm_sign(-pt.y) * std::fabs(n->f.x);
                    if (std::fabs(ffy * dt) > std::fabs(pt.y)) ffy *= std::fabs(pt.y / dt) / std::fabs(ffy);
                    n->f.y += ffy;
                }
                n->v.x = 0.0;
                n->p.x = 0.0;
                n->f.x = 0.0;
            }
            if (n->fixed & 2) {
                n->v.y = 0.0;
                n->p.y = 0.0;
                n->ft.y += n->f.y;
                n->f.y = 0.0;
            }
            if (n->fixed & 4) {
                n->v.z = 0.0;
                n->p.z = 0.0;
                n->f.z = 0.0;
            }
            n->p.x += n->f.x * dt;
            n->p.y += n->f.y * dt;
            n->p.z += n->f.z * dt;
        }
    }

    /* G2P pass 1, DP:490-510 (DPC:304-324 and the velocity line of
     * SP:293-294 skip the displacement-controlled particles). */
    void g2p_velocity(double dt, bool skip_dc = false) {
#pragma omp for
        for (long i = 0; i < (long)particles.size(); i++) {
            Particle *p = particles[i];
            if (skip_dc && p->dc) continue;
            for (size_t k = 0; k < p->agp.size(); k++) {
                const AGP &a = p->agp[k];
                Node *n = nodes[a.index];
                if (n->m > mass_min) {
                    p->v.x += a.N * (n->f.x / n->m) * dt;
                    p->v.y += a.N * (n->f.y / n->m) * dt;
                    p->v.z += a.N * (n->f.z / n->m) * dt;
                }
            }
        }
    }

    /* Velocity remap P2G, DP:539-564 (DPC:353-378 skips the
     * displacement-controlled particles). */
    void remap(int nthreads, bool skip_dc = false) {
#pragma omp for
        for (long i = 0; i < (long)particles.size(); i++) {
            Particle *p = particles[i];
            if (skip_dc && p->dc) continue;
            for (size_t k = 0; k < p->agp.size(); k++) {
                const AGP &a = p->agp[k];
                Node *n = nodes[a.index];
                if (nthreads > 1) omp_set_lock(locks[a.index]);
                if (n->m > mass_min) {
                    n->v.x += a.N * (p->m * p->v.x) / n->m;
                    n->v.y += a.N * (p->m * p->v.y) / n->m;
                    n->v.z += a.N * (p->m * p->v.z) / n->m;
                }
                if (nthreads > 1) omp_unset_lock(locks[a.index]);
            }
        }
    }

    /* Boundary conditions re-applied, DP:573-599. */
    void reapply_bc() {
#pragma omp for
        for (long i = 0; i < (long)nodes.size(); i++) {
            Node *n = nodes[i];
            if (!n->active) continue;
            if (n->fixed & 1) { n->v.x = 0.0; n->p.x = 0.0; n->f.x = 0.0; }
            if (n->fixed & 2) { n->v.y = 0.0; n->p.y = 0.0; n->f.y = 0.0; }
            if (n->fixed & 4) { n->v.z = 0.0; n->p.z = 0.0; n->f.z = 0.0; }
        }
    }

    /* runSimulation_G2P_P2_SmallStrain, DP:175-280. */
    void g2p_strain(double dt, int64_t *null_count, int64_t *plastic_count, int64_t *outside_count,
                    int64_t *nonfinite_count) {
#pragma omp for
        for (long i = 0; i < (long)particles.size(); i++) {
            Particle *p = particles[i];
            if (p->agp.empty()) {
#pragma omp atomic
                (*outside_count)++;
            }
            M3 L = m3_zero();
            for (size_t k = 0; k < p->agp.size(); k++) {
                const AGP &a = p->agp[k];
                Node *n = nodes[a.index];
                if (n->m > mass_min) {
                    p->x.x += a.N * (n->p.x / n->m) * dt;
                    p->x.y += a.N * (n->p.y / n->m) * dt;
                    p->x.z += a.N * (n->p.z / n->m) * dt;
                }
                /* outerProduct(v, g): result[c][r] = v[r] * g[c] */
                double vv[3] = {n->v.x, n->v.y, n->v.z};
                double gg[3] = {a.g.x, a.g.y, a.g.z};
                for (int c = 0; c < 3; c++)
                    for (int r = 0; r < 3; r++) L.m[c][r] += vv[r] * gg[c];
            }
            CR cr;
            M3 dF;
            M3 I = m3_identity();
            for (int c = 0; c < 3; c++)
                for (int r = 0; r < 3; r++) dF.m[c][r] = I.m[c][r] + L.m[c][r] * dt;
            M3 E;
            for (int c = 0; c < 3; c++)
                for (int r = 0; r < 3; r++) E.m[c][r] = 0.5 * (dF.m[r][c] + dF.m[c][r]) - I.m[c][r];
            double de[6];
            de[0] = E.m[0][0];
            de[1] = E.m[1][1];
            de[2] = E.m[2][2];
            de[3] = E.m[0][1] + E.m[1][0];
            de[4] = E.m[1][2] + E.m[2][1];
            de[5] = E.m[2][0] + E.m[0][2];
            if (p->material < 0 || p->material >= (int)materials.size()) {
#pragma omp atomic
                (*null_count)++;
                if (!finite_state(p)) {
#pragma omp atomic
                    (*nonfinite_count)++;
                }
                continue;
            }
            cr_dispatch(cr, materials[p->material], p->R, p->sig, de);
            if (cr.yielded) {
#pragma omp atomic
                (*plastic_count)++;
            }
            p->F = m3_mul(p->F, dF);
            double det = m3_det(p->F);
            p->V = det * p->V0;
            for (int k = 0; k < 6; k++) p->eps[k] += de[k];
            for (int k = 0; k < 6; k++) p->sig[k] += cr.dsig[k];
            for (int k = 0; k < 6; k++) p->epsp[k] += cr.depsp[k];
            double wi = 0.0;
            for (int k = 0; k < 6; k++) wi += (p->sig[k] - 0.5 * cr.dsig[k]) * de[k] * p->V;
            p->Ws += wi;
            p->Wti += p->Ws * dt;
            if (p->Ws > p->Wmax) p->Wmax = p->Ws;
            double wp = 0.0;
            for (int k = 0; k < 6; k++) wp += (p->sig[k] - 0.5 * cr.dsig[k]) * cr.depsp[k] * p->V;
            p->Wp += wp;
            p->R += cr.dR;
            if (!finite_state(p)) {
#pragma omp atomic
                (*nonfinite_count)++;
            }
        }
    }

    static bool finite_state(const Particle *p) {
        bool ok = std::isfinite(p->x.x) && std::isfinite(p->x.y) && std::isfinite(p->x.z);
        for (int k = 0; k < 6; k++) ok = ok && std::isfinite(p->sig[k]);
        return ok;
    }

    /* Displacement-control velocity schedule, DP:766-789. */
    void dc_schedule() {
#pragma omp for
        for (long i = 0; i < (long)dc_list.size(); i++) {
            Particle *p = particles[dc_list[i]];
            if (timeline.t.size() != 0) {
                if (p->dc_axes & 1) p->v.x = p->dc_mult * timeline.velocity(time).x;
                if (p->dc_axes & 2) p->v.y = p->dc_mult * timeline.velocity(time).y;
                if (p->dc_axes & 4) p->v.z = p->dc_mult * timeline.velocity(time).z;
            }
        }
    }


    /* ---- the two other step loops ------------------------------------------ */
    int variant = MPM_VARIANT_DOUBLEPASS;
    int strain = MPM_STRAIN_SMALL;

    /* runSimulation_G2P_P2_LargeStrain, DP:281-437: the gather of x and L
     * as in the small-strain pass (DP:287-303), then large_strain_point. */
    void g2p_strain_large(double dt, int64_t *null_count, int64_t *plastic_count, int64_t *outside_count,
                          int64_t *nonfinite_count) {
#pragma omp for
        for (long i = 0; i < (long)particles.size(); i++) {
            Particle *p = particles[i];
            if (p->agp.empty()) {
#pragma omp atomic
                (*outside_count)++;
            }
            M3 L = m3_zero();
            for (size_t k = 0; k < p->agp.size(); k++) {
                const AGP &a = p->agp[k];
                Node *n = nodes[a.index];
                if (n->m > mass_min) {
                    p->x.x += a.N * (n->p.x / n->m) * dt;
                    p->x.y += a.N * (n->p.y / n->m) * dt;
                    p->x.z += a.N * (n->p.z / n->m) * dt;
                }
                double vv[3] = {n->v.x, n->v.y, n->v.z};
                double gg[3] = {a.g.x, a.g.y, a.g.z};
                for (int c = 0; c < 3; c++)
                    for (int r = 0; r < 3; r++) L.m[c][r] += vv[r] * gg[c];
            }
            const bool valid = p->material >= 0 && p->material < (int)materials.size();
            CR cr;
            large_strain_point(p, valid ? &materials[p->material] : nullptr, L, dt, cr);
            if (!valid) {
#pragma omp atomic
                (*null_count)++;
            } else if (cr.yielded) {
#pragma omp atomic
                (*plastic_count)++;
            }
            if (!finite_state(p)) {
#pragma omp atomic
                (*nonfinite_count)++;
            }
        }
    }

    /* Grid update of the contact variant, DPC:242-277: damping along p (as
     * DP:100-101), momentum integrated first, then the fixed components
     * zeroed with the y reaction taken before (no friction). */
    void integrate_grid_dpc(double dt) {
#pragma omp for
        for (long i = 0; i < (long)nodes.size(); i++) {
            Node *n = nodes[i];
            if (!n->active) continue;
            if (std::sqrt(glm_dot(n->p, n->p)) > 1.0e-24) {
                double inv = 1.0 / std::sqrt(glm_dot(n->p, n->p));
                V3 nrm = v3(n->p.x * inv, n->p.y * inv, n->p.z * inv);
                double s = damping * std::fabs(glm_dot(n->f, nrm));
                n->f.x -= s * nrm.x;
                n->f.y -= s * nrm.y;
                n->f.z -= s * nrm.z;
            }
            n->p.x += n->f.x * dt;
            n->p.y += n->f.y * dt;
            n->p.z += n->f.z * dt;
            if (n->fixed & 1) { n->v.x = 0.0; n->p.x = 0.0; n->f.x = 0.0; }
            if (n->fixed & 2) { n->v.y = 0.0; n->p.y = 0.0; n->ft.y += n->f.y; n->f.y = 0.0; }
            if (n->fixed & 4) { n->v.z = 0.0; n->p.z = 0.0; n->f.z = 0.0; }
        }
    }

    /* SinglePass particle-to-grid, SP:108-204, as one thread runs it: mass
     * first (SP:111-130), then velocity (mass-weighted, only on nodes whose
     * mass exceeds the minimum), momentum (likewise) and force (SP:152-185).
     * Displacement-controlled particles are left out. */
    void p2g_singlepass(int nthreads) {
#pragma omp for
        for (long i = 0; i < (long)particles.size(); i++) {
            Particle *p = particles[i];
            if (p->dc) continue;
            for (size_t k = 0; k < p->agp.size(); k++) {
                const AGP &a = p->agp[k];
                Node *n = nodes[a.index];
                if (nthreads > 1) omp_set_lock(locks[a.index]);
                n->m += a.N * p->m;
                if (nthreads > 1) omp_unset_lock(locks[a.index]);
            }
        }
#pragma omp barrier
#pragma omp for
        for (long i = 0; i < (long)particles.size(); i++) {
            Particle *p = particles[i];
            if (p->dc) continue;
            for (size_t k = 0; k < p->agp.size(); k++) {
                const AGP &a = p->agp[k];
                Node *n = nodes[a.index];
                double N = a.N;
                V3 g = a.g;
                if (nthreads > 1) omp_set_lock(locks[a.index]);
                if (n->m > mass_min) {
                    n->v.x += N * (p->m * p->v.x) / n->m;
                    n->v.y += N * (p->m * p->v.y) / n->m;
                    n->v.z += N * (p->m * p->v.z) / n->m;
                    n->p.x += N * (p->m * p->v.x);
                    n->p.y += N * (p->m * p->v.y);
                    n->p.z += N * (p->m * p->v.z);
                }
                double V = p->V;
                n->f.x += -V * (g.x * p->sig[0] + g.y * p->sig[3] + g.z * p->sig[5]);
                n->f.y += -V * (g.y * p->sig[1] + g.x * p->sig[3] + g.z * p->sig[4]);
                n->f.z += -V * (g.z * p->sig[2] + g.x * p->sig[5] + g.y * p->sig[4]);
                n->f.x += N * p->fext.x;
                n->f.y += N * p->fext.y;
                n->f.z += N * p->fext.z;
                if (nthreads > 1) omp_unset_lock(locks[a.index]);
            }
        }
    }

    /* Cell order of the same sums: mass, momentum and force per-cell
     * partials in one pass, momentum dropped on nodes at or below the mass
     * minimum, then the mass-weighted velocity as in remap_cells. */
    void p2g_singlepass_cells() {
        m2g_cells(true);
#pragma omp barrier
#pragma omp for
        for (long ni = 0; ni < (long)nodes.size(); ni++) {
            Node *n = nodes[ni];
            if (n->active && !(n->m > mass_min)) n->p = v3(0.0, 0.0, 0.0);
        }
#pragma omp barrier
        remap_cells(true);
    }

    /* Grid update of the single pass, SP:210-245: damping along v with the
     * force magnitude, v += f / m dt on nodes above the mass minimum,
     * p += f dt, fixed components zeroed with the y reaction taken before. */
    void integrate_grid_singlepass(double dt) {
#pragma omp for
        for (long i = 0; i < (long)nodes.size(); i++) {
            Node *n = nodes[i];
            if (!n->active) continue;
            if (std::sqrt(glm_dot(n->v, n->v)) > 1.0e-9) {
                double lf = std::sqrt(glm_dot(n->f, n->f));
                double inv = 1.0 / std::sqrt(glm_dot(n->v, n->v));
                V3 nrm = v3(n->v.x * inv, n->v.y * inv, n->v.z * inv);
                double s = damping * lf;
                n->f.x -= s * nrm.x;
                n->f.y -= s * nrm.y;
                n->f.z -= s * nrm.z;
            }
            if (n->m > mass_min) {
                n->v.x += n->f.x / n->m * dt;
                n->v.y += n->f.y / n->m * dt;
                n->v.z += n->f.z / n->m * dt;
            }
            n->p.x += n->f.x * dt;
            n->p.y += n->f.y * dt;
            n->p.z += n->f.z * dt;
            if (n->fixed & 1) { n->v.x = 0.0; n->p.x = 0.0; n->f.x = 0.0; }
            if (n->fixed & 2) { n->v.y = 0.0; n->p.y = 0.0; n->ft.y += n->f.y; n->f.y = 0.0; }
            if (n->fixed & 4) { n->v.z = 0.0; n->p.z = 0.0; n->f.z = 0.0; }
        }
    }

    /* Displacement control on all three axes, DPC:283-297 and 416-432,
     * SP:251-268: each node of a listed particle takes its velocity, the
     * matching momentum and zero force; later list entries win. */
    void displacement_control_all() {
#pragma omp single
        for (size_t i = 0; i < dc_list.size(); i++) {
            Particle *p = particles[dc_list[i]];
            for (size_t k = 0; k < p->agp.size(); k++) {
                Node *n = nodes[p->agp[k].index];
                n->v = p->v;
                n->p = v3(n->m * p->v.x, n->m * p->v.y, n->m * p->v.z);
                n->f = v3(0.0, 0.0, 0.0);
            }
        }
    }

    /* Grid-to-particle strain pass of the contact variant (DPC:438-525) and
     * of the single pass (SP:274-363), displacement-controlled particles
     * skipped: x from p/m, L from the grid velocity, F += (L F) dt,
     * dstrain = dF - I in Voigt form straight from dF, strain updated before
     * the material check (DPC) or the per-particle material defaults used
     * (SP), energies from the new stress. */
    void g2p_strain_variant(double dt, int64_t *null_count, int64_t *plastic_count, int64_t *outside_count,
                            int64_t *nonfinite_count) {
#pragma omp for
        for (long i = 0; i < (long)particles.size(); i++) {
            Particle *p = particles[i];
            if (p->agp.empty()) {
#pragma omp atomic
                (*outside_count)++;
            }
            if (p->dc) continue;
            M3 L = m3_zero();
            for (size_t k = 0; k < p->agp.size(); k++) {
                const AGP &a = p->agp[k];
                Node *n = nodes[a.index];
                if (n->m > mass_min) {
                    p->x.x += a.N * (n->p.x / n->m) * dt;
                    p->x.y += a.N * (n->p.y / n->m) * dt;
                    p->x.z += a.N * (n->p.z / n->m) * dt;
                }
                double vv[3] = {n->v.x, n->v.y, n->v.z};
                double gg[3] = {a.g.x, a.g.y, a.g.z};
                for (int c = 0; c < 3; c++)
                    for (int r = 0; r < 3; r++) L.m[c][r] += vv[r] * gg[c];
            }
            /* F += (L * F) * dt: glm mat3 product, then scaled, then added */
            M3 LF = m3_mul(L, p->F);
            for (int c = 0; c < 3; c++)
                for (int r = 0; r < 3; r++) p->F.m[c][r] = p->F.m[c][r] + LF.m[c][r] * dt;
            double det = m3_det(p->F);
            p->V = det * p->V0;
            M3 dF;
            M3 I = m3_identity();
            for (int c = 0; c < 3; c++)
                for (int r = 0; r < 3; r++) dF.m[c][r] = I.m[c][r] + L.m[c][r] * dt;
            double de[6];
            de[0] = dF.m[0][0] - 1.0;
            de[1] = dF.m[1][1] - 1.0;
            de[2] = dF.m[2][2] - 1.0;
            de[3] = dF.m[0][1] + dF.m[1][0];
            de[4] = dF.m[1][2] + dF.m[2][1];
            de[5] = dF.m[2][0] + dF.m[0][2];
            for (int k = 0; k < 6; k++) p->eps[k] += de[k];
            const bool valid = p->material >= 0 && p->material < (int)materials.size();
            if (!valid) {
#pragma omp atomic
                (*null_count)++;
                if (variant == MPM_VARIANT_DOUBLEPASS_CONTACT) { /* DPC:491-495: message and skip */
                    if (!finite_state(p)) {
#pragma omp atomic
                        (*nonfinite_count)++;
                    }
                    continue;
                }
            }
            Material defaults; /* MaterialPoint_BC.h:48-57: elastic, E = nu = 0 */
            defaults.type = MPM_MAT_ELASTIC;
            const Material &mt = valid ? materials[p->material] : defaults;
            CR cr;
            cr_dispatch(cr, mt, p->R, p->sig, de);
            if (cr.yielded) {
#pragma omp atomic
                (*plastic_count)++;
            }
            for (int k = 0; k < 6; k++) p->sig[k] += cr.dsig[k];
            for (int k = 0; k < 6; k++) p->epsp[k] += cr.depsp[k];
            for (int k = 0; k < 6; k++) p->Ws += p->sig[k] * de[k] * p->V;
            for (int k = 0; k < 6; k++) p->Wp += p->sig[k] * cr.depsp[k] * p->V;
            p->R += cr.dR;
            if (!finite_state(p)) {
#pragma omp atomic
                (*nonfinite_count)++;
            }
        }
    }

    /* Displacement-control schedule of DPC:531-540 and SP:369-378: all three
     * axes from the timeline (when it has points), then x += v dt. */
    void dc_schedule_move(double dt) {
#pragma omp for
        for (long i = 0; i < (long)dc_list.size(); i++) {
            Particle *p = particles[dc_list[i]];
            if (timeline.t.size() != 0) {
                V3 vt = timeline.velocity(time);
                double mult = (double)p->dc_mult;
                p->v = v3(mult * vt.x, mult * vt.y, mult * vt.z);
            }
            p->x.x += p->v.x * dt;
            p->x.y += p->v.y * dt;
            p->x.z += p->v.z * dt;
        }
    }

    /* One pass of the while-loop body, DP:462-803. */
    void step(double dt, int nthreads) {
        if (variant == MPM_VARIANT_DOUBLEPASS_CONTACT) return step_contact(dt, nthreads);
        if (variant == MPM_VARIANT_SINGLEPASS) return step_singlepass(dt, nthreads);
        int64_t nulls = 0, plastic = 0, outside = 0, nonfinite = 0;
#pragma omp parallel num_threads(nthreads)
        {
            int nt = omp_get_num_threads();
#pragma omp barrier
            reset_grid();
#pragma omp barrier
            find_agps();
#pragma omp barrier
            if (order != MPM_ORDER_PARTICLE) {
#pragma omp single
                build_bins();
                m2g_cells();
            } else {
                m2g(nt);
            }
#pragma omp barrier
            displacement_control();
#pragma omp barrier
            integrate_grid(dt);
#pragma omp barrier
            g2p_velocity(dt);
#pragma omp barrier
            if (order != MPM_ORDER_PARTICLE)
                remap_cells();
            else
                remap(nt);
#pragma omp barrier
            displacement_control();
#pragma omp barrier
            reapply_bc();
#pragma omp barrier
            if (strain == MPM_STRAIN_LARGE)
                g2p_strain_large(dt, &nulls, &plastic, &outside, &nonfinite);
            else
                g2p_strain(dt, &nulls, &plastic, &outside, &nonfinite);
#pragma omp barrier
            dc_schedule();
        }
        finish_step(dt, nulls, plastic, outside, nonfinite);
    }


    /* One pass of the contact variant's loop body, DPC:27-543. */
    void step_contact(double dt, int nthreads) {
        int64_t nulls = 0, plastic = 0, outside = 0, nonfinite = 0;
#pragma omp parallel num_threads(nthreads)
        {
            int nt = omp_get_num_threads();
            reset_grid();
#pragma omp barrier
            find_agps();
#pragma omp barrier
            if (order != MPM_ORDER_PARTICLE) {
#pragma omp single
                build_bins();
                m2g_cells();
            } else {
                m2g(nt);
            }
#pragma omp barrier
            integrate_grid_dpc(dt);
#pragma omp barrier
            displacement_control_all();
#pragma omp barrier
            g2p_velocity(dt, true);
#pragma omp barrier
            if (order != MPM_ORDER_PARTICLE)
                remap_cells(true);
            else
                remap(nt, true);
#pragma omp barrier
            reapply_bc();
#pragma omp barrier
            displacement_control_all();
#pragma omp barrier
            g2p_strain_variant(dt, &nulls, &plastic, &outside, &nonfinite);
#pragma omp barrier
            dc_schedule_move(dt);
        }
        finish_step(dt, nulls, plastic, outside, nonfinite);
    }

    /* One pass of the single-pass loop body, SP:27-381. */
    void step_singlepass(double dt, int nthreads) {
        int64_t nulls = 0, plastic = 0, outside = 0, nonfinite = 0;
#pragma omp parallel num_threads(nthreads)
        {
            int nt = omp_get_num_threads();
            reset_grid();
#pragma omp barrier
            find_agps();
#pragma omp barrier
            if (order != MPM_ORDER_PARTICLE) {
#pragma omp single
                build_bins();
                p2g_singlepass_cells();
            } else {
                p2g_singlepass(nt);
            }
#pragma omp barrier
            integrate_grid_singlepass(dt);
#pragma omp barrier
            displacement_control_all();
#pragma omp barrier
            g2p_velocity(dt, true);
#pragma omp barrier
            g2p_strain_variant(dt, &nulls, &plastic, &outside, &nonfinite);
#pragma omp barrier
            dc_schedule_move(dt);
        }
        finish_step(dt, nulls, plastic, outside, nonfinite);
    }

    void finish_step(double dt, int64_t nulls, int64_t plastic, int64_t outside, int64_t nonfinite) {
        health.steps++;
        health.null_material += nulls;
        health.plastic_points += plastic;
        health.outside_grid += outside;
        health.nonfinite_particles += nonfinite;
        time += dt;
        cycle++;
    }
};

extern "C" {

int oracle_create(oracle_engine **out) {
    if (!out) return MPM_ERR_ARG;
    *out = new oracle_engine();
    return MPM_OK;
}

int oracle_destroy(oracle_engine *e) {
    delete e;
    return MPM_OK;
}

/* GridPoint_Factory::createGrid, GridPoint_Factory.h:33-68. */
int oracle_set_grid(oracle_engine *e, const double length[3], const int32_t cells[3]) {
    if (!e || !length || !cells) return MPM_ERR_ARG;
    for (int a = 0; a < 3; a++)
        if (cells[a] < 1) return MPM_ERR_ARG;
    e->clear_nodes();
    e->length = v3(length[0], length[1], length[2]);
    e->h = v3(length[0] / (double)cells[0], length[1] / (double)cells[1], length[2] / (double)cells[2]);
    for (int a = 0; a < 3; a++) {
        e->cells[a] = cells[a];
        e->nn[a] = cells[a] + 1;
    }
    size_t count = (size_t)e->nn[0] * e->nn[1] * e->nn[2];
    e->nodes.resize(count);
    for (int ix = 0; ix < e->nn[0]; ix++)
        for (int iy = 0; iy < e->nn[1]; iy++)
            for (int iz = 0; iz < e->nn[2]; iz++) {
                Node *n = new Node();
                n->pos.x = ix * length[0] / cells[0];
                n->pos.y = iy * length[1] / cells[1];
                n->pos.z = iz * length[2] / cells[2];
                e->nodes[ix + iy * e->nn[0] + iz * (e->nn[0] * e->nn[1])] = n;
            }
    e->locks.resize(count);
    for (size_t i = 0; i < count; i++) {
        e->locks[i] = new omp_lock_t;
        omp_init_lock(e->locks[i]);
    }
    return MPM_OK;
}

int oracle_set_node_boundary(oracle_engine *e, const uint8_t *fixed, const double *friction, int64_t n) {
    if (!e || n != (int64_t)e->nodes.size()) return MPM_ERR_ARG;
    for (int64_t i = 0; i < n; i++) {
        e->nodes[i]->fixed = fixed ? fixed[i] : 0;
        e->nodes[i]->friction = friction ? friction[i] : 0.0;
    }
    return MPM_OK;
}

int oracle_set_materials(oracle_engine *e, const mpm_material *m, int32_t n) {
    if (!e || (n > 0 && !m) || n < 0) return MPM_ERR_ARG;
    e->materials.clear();
    for (int i = 0; i < n; i++) {
        Material mt;
        mt.type = m[i].type;
        mt.density = m[i].density;
        mt.E = m[i].elastic_modulus;
        mt.nu = m[i].poisson_ratio;
        mt.yield = m[i].yield_stress;
        mt.viscosity = m[i].viscosity;
        mt.c0 = m[i].hardening_c0;
        mt.c1 = m[i].hardening_c1;
        e->materials.push_back(mt);
    }
    return MPM_OK;
}

int oracle_set_particle_count(oracle_engine *e, int64_t n) {
    if (!e || n < 0) return MPM_ERR_ARG;
    e->clear_particles();
    e->particles.resize(n);
    for (int64_t i = 0; i < n; i++) e->particles[i] = new Particle();
    e->dc_list.clear();
    for (int k = 0; k < MPM_MON_COUNT; k++) e->monitor[k].clear();
    return MPM_OK;
}

static int field_components(int field) {
    static const int nc[MPM_PF_COUNT] = {3, 3, 3, 1, 1, 1, 9, 6, 6, 6, 1, 1, 1, 1, 1, 9};
    if (field < 0 || field >= MPM_PF_COUNT) return -1;
    return nc[field];
}

static double *particle_slot(Particle *p, int field, int c) {
    switch (field) {
    case MPM_PF_POSITION: return c == 0 ? &p->x.x : c == 1 ? &p->x.y : &p->x.z;
    case MPM_PF_VELOCITY: return c == 0 ? &p->v.x : c == 1 ? &p->v.y : &p->v.z;
    case MPM_PF_FORCE_EXTERNAL: return c == 0 ? &p->fext.x : c == 1 ? &p->fext.y : &p->fext.z;
    case MPM_PF_MASS: return &p->m;
    case MPM_PF_VOLUME: return &p->V;
    case MPM_PF_VOLUME_INITIAL: return &p->V0;
    case MPM_PF_DEFORMATION_GRADIENT: return &p->F.m[c / 3][c % 3];
    case MPM_PF_STRAIN: return &p->eps[c];
    case MPM_PF_STRESS: return &p->sig[c];
    case MPM_PF_STRAIN_PLASTIC: return &p->epsp[c];
    case MPM_PF_BACKSTRESS: return &p->R;
    case MPM_PF_ENERGY_STRAIN: return &p->Ws;
    case MPM_PF_ENERGY_STRAIN_MAX: return &p->Wmax;
    case MPM_PF_ENERGY_STRAIN_TI: return &p->Wti;
    case MPM_PF_ENERGY_PLASTIC: return &p->Wp;
    case MPM_PF_DEFORMATION_GRADIENT_ELASTIC: return &p->Fe.m[c / 3][c % 3];
    }
    return nullptr;
}

int oracle_set_particle_field(oracle_engine *e, int32_t field, const double *src, int64_t n) {
    int nc = field_components(field);
    if (!e || !src || nc < 0 || n != (int64_t)e->particles.size()) return MPM_ERR_ARG;
    for (int c = 0; c < nc; c++)
        for (int64_t i = 0; i < n; i++) *particle_slot(e->particles[i], field, c) = src[c * n + i];
    return MPM_OK;
}

int oracle_get_particle_field(oracle_engine *e, int32_t field, double *dst, int64_t n) {
    int nc = field_components(field);
    if (!e || !dst || nc < 0 || n != (int64_t)e->particles.size()) return MPM_ERR_ARG;
    for (int c = 0; c < nc; c++)
        for (int64_t i = 0; i < n; i++) dst[c * n + i] = *particle_slot(e->particles[i], field, c);
    return MPM_OK;
}

int oracle_set_particle_material(oracle_engine *e, const int32_t *mat, int64_t n) {
    if (!e || !mat || n != (int64_t)e->particles.size()) return MPM_ERR_ARG;
    for (int64_t i = 0; i < n; i++) e->particles[i]->material = mat[i];
    return MPM_OK;
}

int oracle_set_displacement_control(oracle_engine *e, const int64_t *idx, const uint8_t *axes, const float *mult,
                                    int64_t count) {
    if (!e || count < 0 || (count > 0 && (!idx || !axes || !mult))) return MPM_ERR_ARG;
    for (int64_t i = 0; i < count; i++)
        if (idx[i] < 0 || idx[i] >= (int64_t)e->particles.size()) return MPM_ERR_ARG;
    e->dc_list.assign(idx, idx + count);
    for (Particle *p : e->particles) p->dc = false;
    for (int64_t i = 0; i < count; i++) {
        e->particles[idx[i]]->dc = true;
        e->particles[idx[i]]->dc_axes = axes[i];
        e->particles[idx[i]]->dc_mult = mult[i];
    }
    return MPM_OK;
}

int oracle_set_monitor_list(oracle_engine *e, int32_t kind, const int64_t *idx, int64_t count) {
    if (!e || kind < 0 || kind >= MPM_MON_COUNT || count < 0 || (count > 0 && !idx)) return MPM_ERR_ARG;
    for (int64_t i = 0; i < count; i++)
        if (idx[i] < 0 || idx[i] >= (int64_t)e->particles.size()) return MPM_ERR_ARG;
    e->monitor[kind].assign(idx, idx + count);
    return MPM_OK;
}

int oracle_set_timeline(oracle_engine *e, const double *t, const double *v3p, int32_t n) {
    if (!e || n < 0 || (n > 0 && (!t || !v3p))) return MPM_ERR_ARG;
    e->timeline.t.assign(t, t + n);
    e->timeline.v.resize(n);
    for (int i = 0; i < n; i++) e->timeline.v[i] = v3(v3p[3 * i], v3p[3 * i + 1], v3p[3 * i + 2]);
    return MPM_OK;
}

int oracle_set_time(oracle_engine *e, double time, double dt_max, double time_end) {
    if (!e) return MPM_ERR_ARG;
    e->time = time;
    e->dt_max = dt_max;
    e->time_end = time_end;
    return MPM_OK;
}

int oracle_set_damping(oracle_engine *e, double damping, double mass_min) {
    if (!e) return MPM_ERR_ARG;
    e->damping = damping;
    e->mass_min = mass_min;
    return MPM_OK;
}

int oracle_set_order(oracle_engine *e, int32_t order) {
    if (!e || (order != MPM_ORDER_PARTICLE && order != MPM_ORDER_CELL && order != MPM_ORDER_TILE &&
               order != MPM_ORDER_PAIR))
        return MPM_ERR_ARG;
    e->order = order;
    return MPM_OK;
}

int oracle_set_variant(oracle_engine *e, int32_t variant) {
    if (!e || variant < MPM_VARIANT_DOUBLEPASS || variant > MPM_VARIANT_SINGLEPASS) return MPM_ERR_ARG;
    e->variant = variant;
    return MPM_OK;
}

int oracle_set_strain_formulation(oracle_engine *e, int32_t f) {
    if (!e || f < MPM_STRAIN_SMALL || f > MPM_STRAIN_LARGE) return MPM_ERR_ARG;
    e->strain = f;
    return MPM_OK;
}

int oracle_step(oracle_engine *e, double dt, int32_t nsteps, int32_t nthreads) {
    if (!e || nsteps < 0 || nthreads < 1 || e->nodes.empty()) return MPM_ERR_ARG;
    if (e->strain == MPM_STRAIN_LARGE && e->variant != MPM_VARIANT_DOUBLEPASS) return MPM_ERR_STATE;
    for (int s = 0; s < nsteps; s++) e->step(dt, nthreads);
    return MPM_OK;
}

/* runSimulation_Classic_DoublePass_MPLocks loop control, DP:439-458, 802-814
 * (console report left to the caller). */
int oracle_run(oracle_engine *e, double dt_total, int32_t nthreads, int32_t *finished) {
    if (!e || nthreads < 1 || e->nodes.empty()) return MPM_ERR_ARG;
    if (e->strain == MPM_STRAIN_LARGE && e->variant != MPM_VARIANT_DOUBLEPASS) return MPM_ERR_STATE;
    double acc = 0.0;
    while (acc < dt_total) {
        double dt = e->dt_max;
        if (e->dt_max > (dt_total - acc)) dt = dt_total - acc;
        acc += dt;
        e->step(dt, nthreads);
    }
    if (finished) *finished = (e->time < e->time_end) ? 0 : 1;
    return MPM_OK;
}

int oracle_get_time(oracle_engine *e, double *time, int64_t *cycle) {
    if (!e) return MPM_ERR_ARG;
    if (time) *time = e->time;
    if (cycle) *cycle = e->cycle;
    return MPM_OK;
}

int oracle_get_node_count(oracle_engine *e, int64_t *n) {
    if (!e || !n) return MPM_ERR_ARG;
    *n = (int64_t)e->nodes.size();
    return MPM_OK;
}

int oracle_get_node_field(oracle_engine *e, int32_t field, double *dst, int64_t n) {
    if (!e || !dst || n != (int64_t)e->nodes.size()) return MPM_ERR_ARG;
    for (int64_t i = 0; i < n; i++) {
        Node *nd = e->nodes[i];
        switch (field) {
        case MPM_NF_MASS: dst[i] = nd->m; break;
        case MPM_NF_MOMENTUM: dst[i] = nd->p.x; dst[n + i] = nd->p.y; dst[2 * n + i] = nd->p.z; break;
        case MPM_NF_VELOCITY: dst[i] = nd->v.x; dst[n + i] = nd->v.y; dst[2 * n + i] = nd->v.z; break;
        case MPM_NF_FORCE: dst[i] = nd->f.x; dst[n + i] = nd->f.y; dst[2 * n + i] = nd->f.z; break;
        case MPM_NF_FORCE_TEMP: dst[i] = nd->ft.x; dst[n + i] = nd->ft.y; dst[2 * n + i] = nd->ft.z; break;
        case MPM_NF_ACTIVE: dst[i] = nd->active ? 1.0 : 0.0; break;
        default: return MPM_ERR_ARG;
        }
    }
    return MPM_OK;
}

/* PhysicsEngine::reportConsole, PhysicsEngine.cpp:40-183 (principal monitor omitted). */
int oracle_get_monitors(oracle_engine *e, mpm_monitors *out) {
    if (!e || !out) return MPM_ERR_ARG;
    std::memset(out, 0, sizeof(*out));
    out->time = e->time;
    const std::vector<int64_t> &dm = e->monitor[MPM_MON_DISPLACEMENT];
    if (dm.size() > 0) {
        V3 s;
        for (size_t i = 0; i < dm.size(); i++) {
            s.x += e->particles[dm[i]]->x.x;
            s.y += e->particles[dm[i]]->x.y;
            s.z += e->particles[dm[i]]->x.z;
        }
        out->position_y = s.y / dm.size();
        out->has_position = 1;
    }
    {
        V3 f;
        for (size_t i = 0; i < e->nodes.size(); i++) {
            f.x += e->nodes[i]->ft.x;
            f.y += e->nodes[i]->ft.y;
            f.z += e->nodes[i]->ft.z;
        }
        out->force_y = f.y;
    }
    const std::vector<int64_t> &mm = e->monitor[MPM_MON_MOMENTUM];
    if (mm.size() > 0) {
        V3 s;
        for (size_t i = 0; i < mm.size(); i++) {
            Particle *p = e->particles[mm[i]];
            s.x += p->v.x * p->m;
            s.y += p->v.y * p->m;
            s.z += p->v.z * p->m;
        }
        out->momentum_y = s.y;
        out->has_momentum = 1;
    }
    const std::vector<int64_t> &sm = e->monitor[MPM_MON_STRESS];
    if (sm.size() > 0) {
        V3 st, sn;
        for (size_t i = 0; i < sm.size(); i++) {
            Particle *p = e->particles[sm[i]];
            st.x += p->sig[0]; st.y += p->sig[1]; st.z += p->sig[2];
            sn.x += p->eps[0]; sn.y += p->eps[1]; sn.z += p->eps[2];
        }
        double c = (double)sm.size();
        out->stress_y = st.y / c;
        out->strain_y = sn.y / c;
        out->has_stress = 1;
    }
    const std::vector<int64_t> &em = e->monitor[MPM_MON_ENERGY];
    if (em.size() > 0) {
        double ke = 0.0, ws = 0.0, wp = 0.0;
        for (size_t i = 0; i < em.size(); i++) {
            Particle *p = e->particles[em[i]];
            ke += 0.5 * p->m * std::pow(std::sqrt(glm_dot(p->v, p->v)), 2.0);
            ws += p->Ws;
            wp += p->Wp;
        }
        out->kinetic_energy = ke;
        out->strain_energy = ws;
        out->plastic_energy = wp;
        out->has_energy = 1;
    }
    return MPM_OK;
}

int oracle_get_health(oracle_engine *e, mpm_health *out) {
    if (!e || !out) return MPM_ERR_ARG;
    *out = e->health;
    return MPM_OK;
}

/* Known-answer entry points: the mediator and the return maps one call at a time. */
int oracle_find_adjacent(const double length[3], const int32_t cells[3], const double x[3], int32_t *out_index,
                         int32_t *out_count) {
    oracle_engine e;
    e.length = v3(length[0], length[1], length[2]);
    e.h = v3(length[0] / (double)cells[0], length[1] / (double)cells[1], length[2] / (double)cells[2]);
    for (int a = 0; a < 3; a++) {
        e.cells[a] = cells[a];
        e.nn[a] = cells[a] + 1;
    }
    std::vector<unsigned int> adj;
    e.find_adjacent(v3(x[0], x[1], x[2]), adj);
    for (size_t i = 0; i < adj.size(); i++) out_index[i] = (int32_t)adj[i];
    *out_count = (int32_t)adj.size();
    return MPM_OK;
}

int oracle_bases(const double h[3], const double xp[3], const double xn[3], double *N, double g[3]) {
    oracle_engine e;
    e.h = v3(h[0], h[1], h[2]);
    V3 gg;
    e.bases(v3(xp[0], xp[1], xp[2]), v3(xn[0], xn[1], xn[2]), *N, gg);
    g[0] = gg.x;
    g[1] = gg.y;
    g[2] = gg.z;
    return MPM_OK;
}

int oracle_constitutive_batch(const mpm_material *m, const double *R, const double *sig, const double *de,
                              double *dsig, double *depsp, double *dR, int64_t n) {
    if (!m) return MPM_ERR_ARG;
    Material mt;
    mt.type = m->type;
    mt.density = m->density;
    mt.E = m->elastic_modulus;
    mt.nu = m->poisson_ratio;
    mt.yield = m->yield_stress;
    mt.c0 = m->hardening_c0;
    mt.c1 = m->hardening_c1;
    for (int64_t i = 0; i < n; i++) {
        CR o;
        cr_dispatch(o, mt, R[i], sig + 6 * i, de + 6 * i);
        for (int k = 0; k < 6; k++) {
            dsig[6 * i + k] = o.dsig[k];
            depsp[6 * i + k] = o.depsp[k];
        }
        dR[i] = o.dR;
    }
    return MPM_OK;
}

/* getLog / getExp of one glm column-major 3x3 (component 3 * col + row). */
static M3 m3_from(const double *a) {
    M3 m;
    for (int c = 0; c < 3; c++)
        for (int r = 0; r < 3; r++) m.m[c][r] = a[3 * c + r];
    return m;
}
static void m3_to(const M3 &m, double *a) {
    for (int c = 0; c < 3; c++)
        for (int r = 0; r < 3; r++) a[3 * c + r] = m.m[c][r];
}
int oracle_matrix_log(const double *a, double *out) {
    if (!a || !out) return MPM_ERR_ARG;
    m3_to(m3_log_series(m3_from(a)), out);
    return MPM_OK;
}
int oracle_matrix_exp(const double *a, double *out) {
    if (!a || !out) return MPM_ERR_ARG;
    m3_to(m3_exp_series(m3_from(a)), out);
    return MPM_OK;
}

/* The large-strain update of one particle with velocity gradient L (glm
 * column-major) and step dt.  state (in/out, 41 doubles): F 9, Fe 9,
 * sigma 6, strain 6, plastic strain 6, R, V, V0, strain energy, plastic
 * energy.  m = NULL: a particle with no material.  *yielded: the return
 * map yielded. */
int oracle_large_strain_point(const mpm_material *m, const double *L, double dt, double *state, int32_t *yielded) {
    if (!L || !state) return MPM_ERR_ARG;
    Material mt;
    if (m) {
        mt.type = m->type;
        mt.density = m->density;
        mt.E = m->elastic_modulus;
        mt.nu = m->poisson_ratio;
        mt.yield = m->yield_stress;
        mt.c0 = m->hardening_c0;
        mt.c1 = m->hardening_c1;
    }
    Particle p;
    p.F = m3_from(state);
    p.Fe = m3_from(state + 9);
    for (int k = 0; k < 6; k++) {
        p.sig[k] = state[18 + k];
        p.eps[k] = state[24 + k];
        p.epsp[k] = state[30 + k];
    }
    p.R = state[36];
    p.V = state[37];
    p.V0 = state[38];
    p.Ws = state[39];
    p.Wp = state[40];
    CR cr;
    large_strain_point(&p, m ? &mt : nullptr, m3_from(L), dt, cr);
    m3_to(p.F, state);
    m3_to(p.Fe, state + 9);
    for (int k = 0; k < 6; k++) {
        state[18 + k] = p.sig[k];
        state[24 + k] = p.eps[k];
        state[30 + k] = p.epsp[k];
    }
    state[36] = p.R;
    state[37] = p.V;
    state[38] = p.V0;
    state[39] = p.Ws;
    state[40] = p.Wp;
    if (yielded) *yielded = cr.yielded ? 1 : 0;
    return MPM_OK;
}

/* Canvas2D_CC::filterObjective_Smooth, Canvas2D_CC.cpp:302-341, as the
 * reference runs it: the active voxels in ascending index (getVoxels_Active,
 * :262-276), and for each one every active voxel in that order, the ESO ones
 * closer than the radius weighted by (R - d) / R; centres per
 * Canvas2D_CC.cpp:22, distance glm::length(a - b). */
int oracle_eso_filter(const uint8_t *active, const uint8_t *eso, const double *obj, int32_t nx, int32_t ny,
                      double off, double radius, double *out) {
    if (nx < 0 || ny < 0) return MPM_ERR_ARG;
    const int64_t nv = (int64_t)nx * ny;
    if (nv && (!active || !eso || !obj || !out)) return MPM_ERR_ARG;
    std::vector<int64_t> act;
    for (int64_t v = 0; v < nv; v++) {
        out[v] = obj[v];
        if (active[v]) act.push_back(v);
    }
    std::vector<double> px(act.size()), py(act.size()), res(act.size());
    for (size_t k = 0; k < act.size(); k++) {
        px[k] = (double)(act[k] % nx) * off + 0.5 * off;
        py[k] = (double)(act[k] / nx) * off + 0.5 * off;
    }
    for (size_t k = 0; k < act.size(); k++) {
        double sum = 0.0, wsum = 0.0;
        for (size_t j = 0; j < act.size(); j++) {
            double dx = px[k] - px[j], dy = py[k] - py[j];
            double d = std::sqrt(dx * dx + dy * dy);
            if (!eso[act[j]]) continue;
            if (d < radius) {
                double w = (radius - d) / radius;
                wsum += w;
                sum += w * obj[act[j]];
            }
        }
        res[k] = sum / wsum;
    }
    for (size_t k = 0; k < act.size(); k++) out[act[k]] = res[k];
    return MPM_OK;
}

int oracle_timeline_velocity(const double *t, const double *v3p, int32_t n, double time, double out[3]) {
    Timeline tl;
    tl.t.assign(t, t + n);
    tl.v.resize(n);
    for (int i = 0; i < n; i++) tl.v[i] = v3(v3p[3 * i], v3p[3 * i + 1], v3p[3 * i + 2]);
    V3 r = tl.velocity(time);
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
    return MPM_OK;
}

}  // extern "C"
