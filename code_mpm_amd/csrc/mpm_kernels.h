/*
 * mpm_kernels.h — device views and kernel launchers shared by the engine.
 */
#ifndef MPM_AMD_KERNELS_H
#define MPM_AMD_KERNELS_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mpm.h"
#include "mpm_device.h"
#include "mpm_mlp.h"

/* Node tiles: one 64-thread block (one wave) covers tx x ty x tz nodes (x
 * fastest); the shape is chosen at run time (default 32 x 2 x 1, or by the
 * autotuner).  A tile that holds no corner of an occupied cell costs its
 * wave one byte read, so small tiles skip more of a sparse body's empty
 * space: on the 8.0 M-particle honeycomb, 64-node tiles leave 2.1 to 2.5 M
 * node threads in live tiles against 3.6 M with 256-node tiles, for 1.37 M
 * active nodes (node passes 0.215 / 0.190 -> 0.198 / 0.180 ms). */
#define NODE_BLOCK 64

enum { HEALTH_OUTSIDE = 0, HEALTH_NULL = 1, HEALTH_PLASTIC = 2, HEALTH_NONFINITE = 3, HEALTH_SURROGATE_CALLS = 4,
       HEALTH_SURROGATE_FALLBACKS = 5, HEALTH_SLOTS = 6 };
/* The health counters are kept in HEALTH_STRIPES copies, each on its own
 * 128-byte line (HEALTH_STRIDE counters apart); block b adds to copy
 * b % HEALTH_STRIPES and a read sums the copies.  With one copy, every wave
 * of a step where most points yield adds to the same address, and those
 * atomics serialise (2.3 ms of a 3.0 ms pass on 8 M yielding points). */
#define HEALTH_STRIPES 64
#define HEALTH_STRIDE 16
#define HEALTH_WORDS (HEALTH_STRIPES * HEALTH_STRIDE)

enum { MODE_EXACT = 0, MODE_CAPTURE = 1, MODE_SURROGATE = 2 };

/* Return-map mode of the G2P pass 2 kernel, with the surrogate's queue and
 * the sample-capture buffer. */
struct SurrogateArgs {
    int mode;
    int material;               /* material index the network serves */
    double stress_scale;        /* sigma_y of that material */
    double strain_scale;        /* sigma_y / E */
    double tol;                 /* accepted yield residual, in units of sigma_y */
    MLPNorm norm;
    uint32_t *qcount;           /* queue length (device) */
    uint32_t *qid;              /* particle of each queue slot */
    float *qx;                  /* 13 x capacity normalised inputs, feature-major */
    float *qy;                  /* 7 x capacity normalised outputs, feature-major */
    uint8_t *qenv;              /* inputs inside the training envelope */
    double *cap;                /* capture rows: sigma6 de6 R material */
    unsigned long long *capcount;
    int64_t capcap;
};

/* Particle state, structure of arrays; vector/tensor fields are
 * component-major (component c of particle i at [c * n + i]). */
struct PView {
    double *x, *v, *fext;          /* 3n each */
    double *m, *V, *V0;            /* n */
    double *F;                     /* 9n, glm column-major components */
    double *Fe;                    /* 9n, elastic deformation gradient (large-strain pass) */
    double *eps, *sig, *epsp;      /* 6n */
    double *R, *Ws, *Wmax, *Wti, *Wp;
    int32_t *mat;                  /* material index, -1 = none */
    int32_t *dc_rank;              /* latest position in the control list, -1 = not controlled */
    uint8_t *dc_axes;              /* bit0 x, bit1 y, bit2 z */
    float *dc_mult;                /* f_DisplacementControl_Multiplier */
    double *de;                    /* 6n: this step's strain increment (kinematics -> return map) */
    uint8_t *flag;                 /* bit0: non-finite position after the kinematics */
    int64_t n;
};

/* Node state; 3-vectors component-major over nnodes. */
struct NView {
    double *m, *p, *v, *f;
    double *g1;    /* per node {m, f/m}: what G2P pass 1 reads, 32 B  */
    double *g2;    /* per node {m, p/m, v, 0}: G2P pass 2, 64 B       */
    double *fty;   /* support reaction, Force_Temp.y           */
    uint8_t *active, *fixed;
    double *friction;
};

/* Cell binning: key per particle, count / start per ext cell, ids sorted by
 * cell; slot = rank of an occupied cell among occupied cells, occ = the
 * occupied cells in ascending order, nocc = their number (device). */
struct BView {
    uint32_t *key;
    uint32_t *count;
    uint32_t *start;   /* necells + 1 */
    uint32_t *sorted;
    uint32_t *slot;    /* necells, valid for occupied cells */
    uint32_t *occ;     /* up to min(necells, particles) */
    uint32_t *ostart;  /* start of each occupied cell (by rank), then the total */
    uint32_t *nocc;
    int64_t necells;
};

/* Node summation order. */
enum { ORDER_PARTICLE = 0, ORDER_CELL = 1, ORDER_TILE = 2, ORDER_PAIR = 3 };

/* Tile order: binning cells (e = cell + 1 per axis) grouped in summation
 * tiles of ST_X x ST_Y x ST_Z; a tile's node box is the (ST_X + 1) x
 * (ST_Y + 1) x (ST_Z + 1) nodes its cells touch, of which the ST_BND on the
 * box faces also take sums from neighbouring tiles (one record per tile and
 * box-face node, ST_REC doubles). */
#define ST_X MPM_SUM_TILE_X
#define ST_Y MPM_SUM_TILE_Y
#define ST_Z MPM_SUM_TILE_Z
#define ST_CELLS (ST_X * ST_Y * ST_Z)
#define ST_BX (ST_X + 1)
#define ST_BY (ST_Y + 1)
#define ST_BZ (ST_Z + 1)
#define ST_BOX (ST_BX * ST_BY * ST_BZ)
#define ST_BND (ST_BOX - (ST_X - 1) * (ST_Y - 1) * (ST_Z - 1))
#define ST_REC 8
/* step loops: DoublePass, DoublePass_Contact, SinglePass (same codes as MPM_VARIANT_*) */
enum { VAR_DP = 0, VAR_DPC = 1, VAR_SP = 2 };
/* G2P pass 2 of the DoublePass loop (same codes as MPM_STRAIN_*) */
enum { STRAIN_SMALL = 0, STRAIN_LARGE = 1 };

/* Cell-order partial sums: one record per (occupied cell, corner), stored
 * corner-major, item t = corner * cap + occupied rank, so that neighbouring
 * nodes, which read the same neighbour index of neighbouring cells, read
 * neighbouring records; 8 doubles (m, p, f, pad) in the P2G pass, 4 (v, pad)
 * in the remap pass. */
struct CellParts {
    double *w;
    int64_t cap; /* occupied-cell capacity */
    /* particle order: each binned particle's P2G terms at its corner nodes,
     * cell-major (list_term), then one row per particle (P2G: m, m v,
     * f_ext; the remap overwrites it with the updated m v): at most 40
     * doubles per particle */
    double *list;
    int64_t lcap; /* particle capacity */
};

/* slot[] of a cell: its rank among occupied cells, MPM_NONE when empty,
 * with SLOT_DC set when it holds a displacement-controlled particle. */
#define SLOT_DC 0x80000000u

/* One instance of a batch (a single scene is a batch of one): geometry,
 * where its nodes / cells / tiles / particles start in the shared arrays,
 * and its step parameters.  time and vt live on the device so that steps
 * can be replayed from a graph without host writes: k_prepare sets vt from
 * the instance's timeline at step start, k_advance adds dt at step end. */
struct Inst {
    GridGeom g;
    int64_t node_off, cell_off, part_off, npart;
    uint32_t tile_off, tiles;
    uint32_t stile_off, stiles; /* summation tiles (tile order) */
    double dt, damping, mass_min, time;
    double vt[3];            /* TimeLine::getVelocity(time at step start) */
    int32_t tl_off, tl_count;
    int32_t has_fext, pad;
};

struct Batch {
    Inst *inst;               /* device table, ninst entries */
    const uint32_t *tile_off; /* ninst + 1 prefix of node tiles */
    const uint32_t *pinst;    /* instance of each particle; null when ninst == 1 */
    const uint32_t *wave;     /* batch only: per wave of 64 consecutive particles of one
                                 instance, (instance, first particle); null when ninst == 1 */
    uint32_t nwaves;
    const double *tl;         /* timeline points (t, vx, vy, vz) of all instances */
    int ninst;
    uint32_t ntiles;
    int64_t nnodes;           /* all instances: component stride of node arrays */
    int tx, ty, tz;           /* node tile shape, tx * ty * tz == NODE_BLOCK */
    int variant;              /* VAR_* */
    int strain;               /* STRAIN_SMALL / STRAIN_LARGE: G2P pass 2 of the DoublePass loop */
    int fext;                 /* any particle carries an external force */
    int nmat;                 /* materials in the table */
    const uint32_t *stile_off; /* ninst + 1 prefix of summation tiles (tile order) */
    uint32_t nstiles;
    uint8_t *sflag;           /* per summation tile: 1 when it holds an occupied cell (tile order) */
    double *srec;             /* per summation tile, ST_BND face-node records of ST_REC doubles */
    uint8_t *tflag;           /* per node tile: 1 when a corner of an occupied cell lies in it
                                 (set by the P2G partials, cleared by the remap node pass);
                                 null in the particle order */
};

struct MonLists {
    const int64_t *list[4];   /* displacement, momentum, energy, stress */
    int64_t count[4];
    const int64_t *fixed_y;   /* fixed-y nodes, ascending */
    int64_t nfy;
};

hipError_t launch_keys(const Batch &bt, const PView &P, const BView &B, hipStream_t s);
int64_t scan_block_count(int64_t necells);
/* Cells the count, slot (and, plus one, start) arrays hold: whole scan tiles. */
int64_t scan_padded_cells(int64_t necells);
hipError_t launch_bin(const PView &P, const BView &B, uint64_t *block_sum, hipStream_t s);
/* Step phases after binning, in launch order (PHASE_BIN is launch_bin). */
enum { PHASE_BIN = 0, PHASE_P2G_GRID = 1, PHASE_G2P_VELOCITY = 2, PHASE_REMAP_GRID = 3, PHASE_G2P_STRAIN = 4,
       PHASE_COUNT = 5 };
hipError_t launch_phase(int phase, const Batch &bt, const PView &P, const NView &N, const BView &B,
                        const CellParts &C, int order, const MatTable &mt, unsigned long long *health,
                        const SurrogateArgs &sg, hipStream_t s,
                        hipEvent_t mid = nullptr);
hipError_t launch_zero_u32(uint32_t *p, int n, hipStream_t s);
hipError_t launch_surrogate_post(const Batch &bt, const PView &P, const MatTable &mt, unsigned long long *health,
                                 const SurrogateArgs &sg, hipStream_t s);
/* step bracket: vt from each instance's timeline; time += dt; dt of one instance */
hipError_t launch_prepare(const Batch &bt, hipStream_t s);
hipError_t launch_advance(const Batch &bt, hipStream_t s);
hipError_t launch_set_dt(const Batch &bt, int inst, double dt, hipStream_t s);
hipError_t launch_monitors(const PView &P, const NView &N, int64_t nnodes, const MonLists &ml, double *out,
                           hipStream_t s);
hipError_t launch_node_export(const NView &N, int field, int64_t nn, double *dst, hipStream_t s);
hipError_t launch_constitutive(const MatTable &mt, const double *R, const double *sig, const double *de,
                               double *dsig, double *depsp, double *dR, int64_t n, hipStream_t s);

#endif /* MPM_AMD_KERNELS_H */
