/*
 * mpm_kernels.h — device views and kernel launchers shared by the engine.
 */
#ifndef MPM_AMD_KERNELS_H
#define MPM_AMD_KERNELS_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mpm_device.h"

/* Node tile of one 256-thread block: 8 x 8 x 4 nodes, x fastest. */
#define TILE_X 8
#define TILE_Y 8
#define TILE_Z 4

enum { HEALTH_OUTSIDE = 0, HEALTH_NULL = 1, HEALTH_PLASTIC = 2, HEALTH_NONFINITE = 3, HEALTH_SLOTS = 4 };

/* Particle state, structure of arrays; vector/tensor fields are
 * component-major (component c of particle i at [c * n + i]). */
struct PView {
    double *x, *v, *fext;          /* 3n each */
    double *m, *V, *V0;            /* n */
    double *F;                     /* 9n, glm column-major components */
    double *eps, *sig, *epsp;      /* 6n */
    double *R, *Ws, *Wmax, *Wti, *Wp;
    int32_t *mat;                  /* material index, -1 = none */
    int32_t *dc_rank;              /* latest position in the control list, -1 = not controlled */
    uint8_t *dc_axes;              /* bit0 x, bit1 y, bit2 z */
    float *dc_mult;                /* f_DisplacementControl_Multiplier */
    int64_t n;
};

/* Node state; 3-vectors component-major over nnodes. */
struct NView {
    double *m, *p, *v, *f;
    double *a;     /* f / m after the grid update (G2P pass 1) */
    double *u;     /* p / m after the remap (G2P pass 2)       */
    double *fty;   /* support reaction, Force_Temp.y           */
    uint8_t *active, *fixed;
    double *friction;
};

/* Cell binning: key per particle, count / start per ext cell, ids sorted by cell. */
struct BView {
    uint32_t *key;
    uint32_t *count;
    uint32_t *start;   /* necells + 1 */
    uint32_t *sorted;
    int64_t necells;
};

struct StepParams {
    double dt;
    double mass_min;
    double damping;
    double vt[3];            /* TimeLine::getVelocity(time at step start) */
    int has_fext;
    int timeline_nonempty;
};

struct MonLists {
    const int64_t *list[4];   /* displacement, momentum, energy, stress */
    int64_t count[4];
    const int64_t *fixed_y;   /* fixed-y nodes, ascending */
    int64_t nfy;
};

hipError_t launch_keys(const GridGeom &g, const PView &P, const BView &B, hipStream_t s);
int64_t scan_block_count(int64_t necells);
hipError_t launch_bin(const GridGeom &g, const PView &P, const BView &B, uint32_t *block_sum, hipStream_t s);
hipError_t launch_step_kernels(const GridGeom &g, const PView &P, const NView &N, const BView &B,
                               const StepParams &sp, const MatTable &mt, unsigned long long *health,
                               hipStream_t s);
hipError_t launch_monitors(const GridGeom &g, const PView &P, const NView &N, const MonLists &ml, double *out,
                           hipStream_t s);
hipError_t launch_node_export(const NView &N, int field, int64_t nn, double *dst, hipStream_t s);
hipError_t launch_constitutive(const MatTable &mt, const double *R, const double *sig, const double *de,
                               double *dsig, double *depsp, double *dR, int64_t n, hipStream_t s);

#endif /* MPM_AMD_KERNELS_H */
