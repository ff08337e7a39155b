/*
 * mpm_kernels.hip — gfx950 kernels of the Classic MPM DoublePass timestep.
 *
 * One step (PhysicsEngine_runSimulation_Classic_DoublePass_MPLocks.cpp:462-790)
 * becomes eight launches on one stream:
 *
 *   scan_reduce / scan_blocks / scan_apply   exclusive scan of the per-cell
 *                                            particle counts -> cell_start
 *   scatter        particle ids into their cell's slots (count -> 0)
 *   cell_sort      each cell's ids ascending
 *   p2g_grid       per node: gather of every touching particle in ascending
 *                  particle order (8-way merge of the 8 adjacent cell lists),
 *                  then the fused displacement-control override, damping,
 *                  friction, fixed BCs, support reaction and momentum update
 *                  (DP:48-142, 143-174)
 *   g2p_velocity   per particle: v += N (f/m) dt (DP:490-510)
 *   remap_grid     per node: velocity remap gather in ascending particle order,
 *                  then the second displacement-control override and the
 *                  BC re-application; hoists p/m (DP:539-599)
 *   g2p_strain     per particle: position, velocity gradient, strain, return
 *                  map, F, V, energies, hardening, the displacement-control
 *                  schedule, and the cell key + count for the next step
 *                  (DP:175-280, 766-789)
 *
 * Each node sum is accumulated by one thread in the order the reference's
 * single-threaded scatter produces (ascending particle index), so results do
 * not depend on scheduling and need no floating-point atomics.
 */
#include "mpm_kernels.h"

#define BLOCK 256

/* Blocks are dealt round-robin to the 8 XCDs; give each XCD one contiguous
 * range of tiles so neighbouring tiles share that XCD's L2. */
__device__ __forceinline__ uint32_t xcd_remap(uint32_t bid, uint32_t nblk) {
    uint32_t q = nblk / 8u, r = nblk % 8u;
    uint32_t x = bid % 8u, k = bid / 8u;
    return (x < r) ? x * (q + 1u) + k : r * (q + 1u) + (x - r) * q + k;
}

/* Ext-cell key of a position: cell floor(x/h) + 1 per axis when inside
 * [-1, cells], else MPM_NONE. */
__device__ __forceinline__ uint32_t cell_key(const GridGeom &g, double x, double y, double z) {
    double p[3] = {x, y, z};
    int c[3];
    for (int a = 0; a < 3; a++) {
        double f = floor(p[a] / g.h[a]);
        if (!(f >= -1.0 && f <= (double)g.cells[a])) return MPM_NONE;
        c[a] = (int)f + 1;
    }
    return (uint32_t)c[0] + (uint32_t)g.ec[0] * ((uint32_t)c[1] + (uint32_t)g.ec[1] * (uint32_t)c[2]);
}

__global__ __launch_bounds__(BLOCK) void k_keys(GridGeom g, PView P, BView B) {
    uint32_t i = xcd_remap(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
    if (i >= P.n) return;
    uint32_t k = cell_key(g, P.x[i], P.x[P.n + i], P.x[2 * P.n + i]);
    B.key[i] = k;
    if (k != MPM_NONE) atomicAdd(&B.count[k], 1u);
}

/* ---- exclusive scan of cell counts (reduce, scan of block sums, apply) ---- */

#define SCAN_ITEMS 16
#define SCAN_TILE (BLOCK * SCAN_ITEMS)

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t *lds, uint32_t &total) {
    /* wave-level inclusive scan, then across the 4 waves */
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) lds[wave] = x;
    __syncthreads();
    uint32_t off = 0;
    for (int w = 0; w < wave; w++) off += lds[w];
    total = lds[0] + lds[1] + lds[2] + lds[3];
    __syncthreads();
    return off + x - v;
}

__global__ __launch_bounds__(BLOCK) void k_scan_reduce(const uint32_t *count, int64_t n, uint32_t *block_sum) {
    __shared__ uint32_t lds[4];
    int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; j++)
        if (base + j < n) s += count[base + j];
    uint32_t total;
    block_exclusive_scan(s, lds, total);
    if (threadIdx.x == 0) block_sum[blockIdx.x] = total;
}

/* One block scans all block sums in place (exclusive) in chunks of BLOCK. */
__global__ __launch_bounds__(BLOCK) void k_scan_blocks(uint32_t *block_sum, int64_t nblocks) {
    __shared__ uint32_t lds[4];
    uint32_t carry = 0;
    for (int64_t base = 0; base < nblocks; base += BLOCK) {
        int64_t i = base + threadIdx.x;
        uint32_t v = (i < nblocks) ? block_sum[i] : 0u;
        uint32_t total;
        uint32_t ex = block_exclusive_scan(v, lds, total);
        if (i < nblocks) block_sum[i] = carry + ex;
        carry += total;
    }
}

__global__ __launch_bounds__(BLOCK) void k_scan_apply(const uint32_t *count, int64_t n, const uint32_t *block_sum,
                                                     uint32_t *start) {
    __shared__ uint32_t lds[4];
    int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS];
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; j++) {
        v[j] = (base + j < n) ? count[base + j] : 0u;
        s += v[j];
    }
    uint32_t total;
    uint32_t ex = block_exclusive_scan(s, lds, total) + block_sum[blockIdx.x];
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; j++) {
        if (base + j < n) start[base + j] = ex;
        ex += v[j];
    }
    if (base + SCAN_ITEMS >= n && base < n) start[n] = ex; /* the thread holding the last cell */
}

__global__ __launch_bounds__(BLOCK) void k_scatter(PView P, BView B) {
    uint32_t i = xcd_remap(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
    if (i >= P.n) return;
    uint32_t k = B.key[i];
    if (k == MPM_NONE) return;
    uint32_t slot = B.start[k] + atomicSub(&B.count[k], 1u) - 1u;
    B.sorted[slot] = i;
}

/* Ascending ids inside each cell: the order the node gathers merge in. */
__global__ __launch_bounds__(BLOCK) void k_cell_sort(BView B) {
    int64_t c = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (c >= B.necells) return;
    uint32_t s = B.start[c], e = B.start[c + 1];
    if (e - s < 2) return;
    for (uint32_t i = s + 1; i < e; i++) {
        uint32_t v = B.sorted[i];
        uint32_t j = i;
        while (j > s && B.sorted[j - 1] > v) {
            B.sorted[j] = B.sorted[j - 1];
            j--;
        }
        B.sorted[j] = v;
    }
}

/* ---- 8-way merge of a node's adjacent cell lists in ascending id order ---- */

struct Merge8 {
    uint32_t p[8], e[8], h[8];
};

__device__ __forceinline__ void merge_init(Merge8 &M, const BView &B, const GridGeom &g, int ix, int iy, int iz) {
    /* node i touches cells i-1 and i per axis = ext cells i and i+1 */
#pragma unroll
    for (int k = 0; k < 8; k++) {
        int cx = ix + ((k >> 2) & 1), cy = iy + ((k >> 1) & 1), cz = iz + (k & 1);
        uint32_t c = (uint32_t)cx + (uint32_t)g.ec[0] * ((uint32_t)cy + (uint32_t)g.ec[1] * (uint32_t)cz);
        M.p[k] = B.start[c];
        M.e[k] = B.start[c + 1];
        M.h[k] = (M.p[k] < M.e[k]) ? B.sorted[M.p[k]] : MPM_NONE;
    }
}

__device__ __forceinline__ uint32_t merge_min(const Merge8 &M) {
    uint32_t a = min(min(M.h[0], M.h[1]), min(M.h[2], M.h[3]));
    uint32_t b = min(min(M.h[4], M.h[5]), min(M.h[6], M.h[7]));
    return min(a, b);
}

__device__ __forceinline__ void merge_advance(Merge8 &M, const BView &B, uint32_t id) {
#pragma unroll
    for (int k = 0; k < 8; k++) {
        if (M.h[k] == id) {
            M.p[k]++;
            M.h[k] = (M.p[k] < M.e[k]) ? B.sorted[M.p[k]] : MPM_NONE;
        }
    }
}

__device__ __forceinline__ void node_of_tile(const GridGeom &g, uint32_t tile, int &ix, int &iy, int &iz) {
    int tnx = (g.nn[0] + TILE_X - 1) / TILE_X, tny = (g.nn[1] + TILE_Y - 1) / TILE_Y;
    int tx = (int)(tile % (uint32_t)tnx);
    int ty = (int)((tile / (uint32_t)tnx) % (uint32_t)tny);
    int tz = (int)(tile / ((uint32_t)tnx * (uint32_t)tny));
    ix = tx * TILE_X + (int)(threadIdx.x % TILE_X);
    iy = ty * TILE_Y + (int)((threadIdx.x / TILE_X) % TILE_Y);
    iz = tz * TILE_Z + (int)(threadIdx.x / (TILE_X * TILE_Y));
}

/* Track, per axis, the displacement-controlled particle latest in the
 * control list among those touching the node (the last writer of
 * runSimulation_DisplacementControl, DP:143-174, when run in list order). */
struct DCWin {
    int rank[3];
    double v[3];
};

__device__ __forceinline__ void dc_consider(DCWin &w, const PView &P, uint32_t id) {
    int r = P.dc_rank[id];
    if (r < 0) return;
    uint8_t ax = P.dc_axes[id];
#pragma unroll
    for (int a = 0; a < 3; a++)
        if ((ax >> a) & 1)
            if (r > w.rank[a]) {
                w.rank[a] = r;
                w.v[a] = P.v[(size_t)a * P.n + id];
            }
}

__global__ __launch_bounds__(BLOCK) void k_p2g_grid(GridGeom g, PView P, NView Nd, BView B, StepParams sp) {
    int ix, iy, iz;
    node_of_tile(g, xcd_remap(blockIdx.x, gridDim.x), ix, iy, iz);
    if (ix >= g.nn[0] || iy >= g.nn[1] || iz >= g.nn[2]) return;
    const int64_t node = (int64_t)ix + (int64_t)g.nn[0] * ((int64_t)iy + (int64_t)g.nn[1] * iz);
    const size_t n = P.n;
    const double xn[3] = {node_coord(ix, g.L[0], g.cells[0]), node_coord(iy, g.L[1], g.cells[1]),
                          node_coord(iz, g.L[2], g.cells[2])};
    Merge8 M;
    merge_init(M, B, g, ix, iy, iz);
    uint32_t id = merge_min(M);
    if (id == MPM_NONE) {
        Nd.active[node] = 0;
        return;
    }
    double m = 0.0, p[3] = {0.0, 0.0, 0.0}, f[3] = {0.0, 0.0, 0.0};
    DCWin w;
    w.rank[0] = w.rank[1] = w.rank[2] = -1;
    w.v[0] = w.v[1] = w.v[2] = 0.0;
    while (id != MPM_NONE) {
        /* runSimulation_M2G body, DP:69-84 */
        AxisShape ax = axis_shape(P.x[id], xn[0], g.h[0]);
        AxisShape ay = axis_shape(P.x[n + id], xn[1], g.h[1]);
        AxisShape az = axis_shape(P.x[2 * n + id], xn[2], g.h[2]);
        double N, gx, gy, gz;
        shape_full(ax, ay, az, g.h, N, gx, gy, gz);
        double mp = P.m[id];
        m += N * mp;
        p[0] += N * (mp * P.v[id]);
        p[1] += N * (mp * P.v[n + id]);
        p[2] += N * (mp * P.v[2 * n + id]);
        double V = P.V[id];
        const double *s = P.sig;
        double s0 = s[id], s1 = s[n + id], s2 = s[2 * n + id], s3 = s[3 * n + id], s4 = s[4 * n + id],
               s5 = s[5 * n + id];
        f[0] += -V * (gx * s0 + gy * s3 + gz * s5);
        f[1] += -V * (gy * s1 + gx * s3 + gz * s4);
        f[2] += -V * (gz * s2 + gx * s5 + gy * s4);
        if (sp.has_fext) {
            f[0] += N * P.fext[id];
            f[1] += N * P.fext[n + id];
            f[2] += N * P.fext[2 * n + id];
        }
        dc_consider(w, P, id);
        merge_advance(M, B, id);
        id = merge_min(M);
    }
    double v[3] = {0.0, 0.0, 0.0};
    /* runSimulation_DisplacementControl, DP:143-174 */
#pragma unroll
    for (int a = 0; a < 3; a++)
        if (w.rank[a] >= 0) {
            v[a] = w.v[a];
            p[a] = m * w.v[a];
            f[a] = 0.0;
        }
    /* runSimulation_IntegrateGrid, DP:90-142 */
    const double dt = sp.dt;
    {
        double pp = p[0] * p[0] + p[1] * p[1] + p[2] * p[2];
        if (sqrt(pp) > 1.0e-24) {
            double inv = 1.0 / sqrt(pp);
            double nx = p[0] * inv, ny = p[1] * inv, nz = p[2] * inv;
            double sc = sp.damping * fabs(f[0] * nx + f[1] * ny + f[2] * nz);
            f[0] -= sc * nx;
            f[1] -= sc * ny;
            f[2] -= sc * nz;
        }
    }
    const uint8_t fixed = Nd.fixed[node];
    double fty = 0.0;
    if (fixed & 1) {
        double pty = p[1] + f[1] * dt;
        if (fabs(pty * dt) > 1.0e-24 && f[0] > 0.0) {
            double ffy = Nd.friction[node] * glm_sign(-pty) * fabs(f[0]);
            if (fabs(ffy * dt) > fabs(pty)) ffy *= fabs(pty / dt) / fabs(ffy);
            f[1] += ffy;
        }
        v[0] = 0.0;
        p[0] = 0.0;
        f[0] = 0.0;
    }
    if (fixed & 2) {
        v[1] = 0.0;
        p[1] = 0.0;
        fty = 0.0 + f[1];
        f[1] = 0.0;
    }
    if (fixed & 4) {
        v[2] = 0.0;
        p[2] = 0.0;
        f[2] = 0.0;
    }
    p[0] += f[0] * dt;
    p[1] += f[1] * dt;
    p[2] += f[2] * dt;
    const size_t nn = (size_t)g.nnodes;
    Nd.active[node] = 1;
    Nd.m[node] = m;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        Nd.p[a * nn + node] = p[a];
        Nd.f[a * nn + node] = f[a];
        Nd.a[a * nn + node] = f[a] / m; /* hoisted f/m of DP:508 */
    }
    Nd.fty[node] = fty;
}

/* G2P pass 1: v_p += N * (f_i / m_i) * dt over the adjacent nodes in
 * GridPoint_Mediator order (x, then y, then z innermost), DP:490-510. */
__global__ __launch_bounds__(BLOCK) void k_g2p_velocity(GridGeom g, PView P, NView Nd, StepParams sp) {
    uint32_t i = xcd_remap(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
    if (i >= P.n) return;
    const size_t n = P.n, nn = (size_t)g.nnodes;
    double x[3] = {P.x[i], P.x[n + i], P.x[2 * n + i]};
    int b[3], lo[3], hi[3];
    double s[3][2];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        base_cell(x[a], g.h[a], g.nn[a], b[a], lo[a], hi[a]);
#pragma unroll
        for (int o = 0; o < 2; o++) s[a][o] = axis_shape(x[a], node_coord(b[a] + o, g.L[a], g.cells[a]), g.h[a]).s;
    }
    double v[3] = {P.v[i], P.v[n + i], P.v[2 * n + i]};
    const double dt = sp.dt;
    for (int ox = lo[0]; ox < hi[0]; ox++) {
        int jx = b[0] + ox;
        if (jx < 0 || jx > g.nn[0] - 1) continue;
        for (int oy = lo[1]; oy < hi[1]; oy++) {
            int jy = b[1] + oy;
            if (jy < 0 || jy > g.nn[1] - 1) continue;
            for (int oz = lo[2]; oz < hi[2]; oz++) {
                int jz = b[2] + oz;
                if (jz < 0 || jz > g.nn[2] - 1) continue;
                size_t node = (size_t)jx + (size_t)g.nn[0] * ((size_t)jy + (size_t)g.nn[1] * jz);
                if (Nd.m[node] > sp.mass_min) {
                    double N = s[0][ox] * s[1][oy] * s[2][oz];
                    v[0] += N * Nd.a[node] * dt;
                    v[1] += N * Nd.a[nn + node] * dt;
                    v[2] += N * Nd.a[2 * nn + node] * dt;
                }
            }
        }
    }
    P.v[i] = v[0];
    P.v[n + i] = v[1];
    P.v[2 * n + i] = v[2];
}

/* Velocity remap (DP:539-564) gathered per node in ascending particle
 * order, then the second displacement-control override (DP:568) and the BC
 * re-application (DP:573-599). Stores the final v, p, the zeroed force
 * components and the hoisted p/m used by G2P pass 2 (DP:198). */
__global__ __launch_bounds__(BLOCK) void k_remap_grid(GridGeom g, PView P, NView Nd, BView B, StepParams sp) {
    int ix, iy, iz;
    node_of_tile(g, xcd_remap(blockIdx.x, gridDim.x), ix, iy, iz);
    if (ix >= g.nn[0] || iy >= g.nn[1] || iz >= g.nn[2]) return;
    const int64_t node = (int64_t)ix + (int64_t)g.nn[0] * ((int64_t)iy + (int64_t)g.nn[1] * iz);
    if (!Nd.active[node]) return;
    const size_t n = P.n, nn = (size_t)g.nnodes;
    const double xn[3] = {node_coord(ix, g.L[0], g.cells[0]), node_coord(iy, g.L[1], g.cells[1]),
                          node_coord(iz, g.L[2], g.cells[2])};
    const double m = Nd.m[node];
    const bool massive = m > sp.mass_min;
    Merge8 M;
    merge_init(M, B, g, ix, iy, iz);
    double v[3] = {0.0, 0.0, 0.0};
    DCWin w;
    w.rank[0] = w.rank[1] = w.rank[2] = -1;
    w.v[0] = w.v[1] = w.v[2] = 0.0;
    for (uint32_t id = merge_min(M); id != MPM_NONE; id = merge_min(M)) {
        if (massive) {
            double N = axis_shape(P.x[id], xn[0], g.h[0]).s * axis_shape(P.x[n + id], xn[1], g.h[1]).s *
                       axis_shape(P.x[2 * n + id], xn[2], g.h[2]).s;
            double mp = P.m[id];
            v[0] += N * (mp * P.v[id]) / m;
            v[1] += N * (mp * P.v[n + id]) / m;
            v[2] += N * (mp * P.v[2 * n + id]) / m;
        }
        dc_consider(w, P, id);
        merge_advance(M, B, id);
    }
    double p[3] = {Nd.p[node], Nd.p[nn + node], Nd.p[2 * nn + node]};
    const uint8_t fixed = Nd.fixed[node];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        if (w.rank[a] >= 0) {
            v[a] = w.v[a];
            p[a] = m * w.v[a];
            Nd.f[a * nn + node] = 0.0;
        }
        if ((fixed >> a) & 1) {
            v[a] = 0.0;
            p[a] = 0.0;
        }
        Nd.v[a * nn + node] = v[a];
        Nd.p[a * nn + node] = p[a];
        Nd.u[a * nn + node] = p[a] / m;
    }
}

__device__ __forceinline__ void block_count_add(unsigned long long *dst, uint32_t flag) {
    /* wave ballot, then one atomic per wave */
    uint64_t b = __ballot(flag != 0);
    if ((threadIdx.x & 63) == 0 && b) atomicAdd(dst, (unsigned long long)__popcll(b));
}

/* runSimulation_G2P_P2_SmallStrain (DP:175-280) + the displacement-control
 * schedule (DP:766-789) + next step's binning key. */
__global__ __launch_bounds__(BLOCK) void k_g2p_strain(GridGeom g, PView P, NView Nd, BView B, StepParams sp,
                                                     MatTable mt, unsigned long long *health) {
    uint32_t i = xcd_remap(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
    const bool live = i < P.n;
    uint32_t f_outside = 0, f_null = 0, f_plastic = 0, f_nonfinite = 0;
    if (live) {
        const size_t n = P.n, nn = (size_t)g.nnodes;
        const double x0[3] = {P.x[i], P.x[n + i], P.x[2 * n + i]};
        double x[3] = {x0[0], x0[1], x0[2]};
        int b[3], lo[3], hi[3];
        AxisShape as[3][2];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            base_cell(x0[a], g.h[a], g.nn[a], b[a], lo[a], hi[a]);
#pragma unroll
            for (int o = 0; o < 2; o++) as[a][o] = axis_shape(x0[a], node_coord(b[a] + o, g.L[a], g.cells[a]), g.h[a]);
        }
        double L[3][3]; /* glm column-major L[col][row] */
#pragma unroll
        for (int c = 0; c < 3; c++)
#pragma unroll
            for (int r = 0; r < 3; r++) L[c][r] = 0.0;
        const double dt = sp.dt;
        int touched = 0;
        for (int ox = lo[0]; ox < hi[0]; ox++) {
            int jx = b[0] + ox;
            if (jx < 0 || jx > g.nn[0] - 1) continue;
            for (int oy = lo[1]; oy < hi[1]; oy++) {
                int jy = b[1] + oy;
                if (jy < 0 || jy > g.nn[1] - 1) continue;
                for (int oz = lo[2]; oz < hi[2]; oz++) {
                    int jz = b[2] + oz;
                    if (jz < 0 || jz > g.nn[2] - 1) continue;
                    touched++;
                    size_t node = (size_t)jx + (size_t)g.nn[0] * ((size_t)jy + (size_t)g.nn[1] * jz);
                    double N, gr[3];
                    shape_full(as[0][ox], as[1][oy], as[2][oz], g.h, N, gr[0], gr[1], gr[2]);
                    if (Nd.m[node] > sp.mass_min) {
                        x[0] += N * Nd.u[node] * dt;
                        x[1] += N * Nd.u[nn + node] * dt;
                        x[2] += N * Nd.u[2 * nn + node] * dt;
                    }
                    double vv[3] = {Nd.v[node], Nd.v[nn + node], Nd.v[2 * nn + node]};
#pragma unroll
                    for (int c = 0; c < 3; c++)
#pragma unroll
                        for (int r = 0; r < 3; r++) L[c][r] += vv[r] * gr[c];
                }
            }
        }
        f_outside = (touched == 0);
        P.x[i] = x[0];
        P.x[n + i] = x[1];
        P.x[2 * n + i] = x[2];
        const int mat = P.mat[i];
        if (mat < 0 || mat >= mt.n) {
            f_null = 1;
        } else {
            double dF[3][3];
#pragma unroll
            for (int c = 0; c < 3; c++)
#pragma unroll
                for (int r = 0; r < 3; r++) dF[c][r] = (c == r ? 1.0 : 0.0) + L[c][r] * dt;
            double E[3][3];
#pragma unroll
            for (int c = 0; c < 3; c++)
#pragma unroll
                for (int r = 0; r < 3; r++) E[c][r] = 0.5 * (dF[r][c] + dF[c][r]) - (c == r ? 1.0 : 0.0);
            double de[6];
            de[0] = E[0][0];
            de[1] = E[1][1];
            de[2] = E[2][2];
            de[3] = E[0][1] + E[1][0];
            de[4] = E[1][2] + E[2][1];
            de[5] = E[2][0] + E[0][2];
            double sig[6];
#pragma unroll
            for (int k = 0; k < 6; k++) sig[k] = P.sig[k * n + i];
            const double R = P.R[i];
            CRResult cr;
            constitutive(mt, mat, R, sig, de, cr);
            f_plastic = cr.yielded;
            /* F = F * dF (DP:242) */
            double F[3][3], Fn[3][3];
#pragma unroll
            for (int c = 0; c < 3; c++)
#pragma unroll
                for (int r = 0; r < 3; r++) F[c][r] = P.F[(3 * c + r) * n + i];
#pragma unroll
            for (int c = 0; c < 3; c++)
#pragma unroll
                for (int r = 0; r < 3; r++) Fn[c][r] = F[0][r] * dF[c][0] + F[1][r] * dF[c][1] + F[2][r] * dF[c][2];
#pragma unroll
            for (int c = 0; c < 3; c++)
#pragma unroll
                for (int r = 0; r < 3; r++) P.F[(3 * c + r) * n + i] = Fn[c][r];
            double det = +Fn[0][0] * (Fn[1][1] * Fn[2][2] - Fn[2][1] * Fn[1][2]) -
                         Fn[1][0] * (Fn[0][1] * Fn[2][2] - Fn[2][1] * Fn[0][2]) +
                         Fn[2][0] * (Fn[0][1] * Fn[1][2] - Fn[1][1] * Fn[0][2]);
            const double V = det * P.V0[i];
            P.V[i] = V;
#pragma unroll
            for (int k = 0; k < 6; k++) P.eps[k * n + i] += de[k];
#pragma unroll
            for (int k = 0; k < 6; k++) sig[k] += cr.dsig[k];
#pragma unroll
            for (int k = 0; k < 6; k++) P.sig[k * n + i] = sig[k];
#pragma unroll
            for (int k = 0; k < 6; k++) P.epsp[k * n + i] += cr.depsp[k];
            double wi = 0.0;
#pragma unroll
            for (int k = 0; k < 6; k++) wi += (sig[k] - 0.5 * cr.dsig[k]) * de[k] * V;
            double Ws = P.Ws[i] + wi;
            P.Ws[i] = Ws;
            P.Wti[i] += Ws * dt;
            if (Ws > P.Wmax[i]) P.Wmax[i] = Ws;
            double wp = 0.0;
#pragma unroll
            for (int k = 0; k < 6; k++) wp += (sig[k] - 0.5 * cr.dsig[k]) * cr.depsp[k] * V;
            P.Wp[i] += wp;
            P.R[i] = R + cr.dR;
            for (int k = 0; k < 6; k++) f_nonfinite |= !isfinite(sig[k]);
        }
        f_nonfinite |= !(isfinite(x[0]) && isfinite(x[1]) && isfinite(x[2]));
        if (f_null) {
            for (int k = 0; k < 6; k++) f_nonfinite |= !isfinite(P.sig[k * n + i]);
        }
        /* displacement-control schedule, DP:766-789 */
        if (sp.timeline_nonempty && P.dc_rank[i] >= 0) {
            uint8_t axm = P.dc_axes[i];
            double mult = (double)P.dc_mult[i];
#pragma unroll
            for (int a = 0; a < 3; a++)
                if ((axm >> a) & 1) P.v[a * n + i] = mult * sp.vt[a];
        }
        /* next step's binning */
        uint32_t k = cell_key(g, x[0], x[1], x[2]);
        B.key[i] = k;
        if (k != MPM_NONE) atomicAdd(&B.count[k], 1u);
    }
    block_count_add(&health[HEALTH_OUTSIDE], f_outside);
    block_count_add(&health[HEALTH_NULL], f_null);
    block_count_add(&health[HEALTH_PLASTIC], f_plastic);
    block_count_add(&health[HEALTH_NONFINITE], f_nonfinite);
}

/* ---- reportConsole sums (PhysicsEngine.cpp:40-183) in list order ---- */

#define MON_CHUNK 1024

/* One block. Lanes stage a chunk of values in LDS; lane 0 adds them in list
 * order, so every sum is the reference's sequential sum. */
__global__ __launch_bounds__(BLOCK) void k_monitors(GridGeom g, PView P, NView Nd, MonLists ml, double *out) {
    __shared__ double a0[MON_CHUNK], a1[MON_CHUNK], a2[MON_CHUNK];
    const size_t n = P.n, nn = (size_t)g.nnodes;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    /* Position_y */
    for (int64_t base = 0; base < ml.count[0]; base += MON_CHUNK) {
        for (int64_t j = threadIdx.x; j < MON_CHUNK && base + j < ml.count[0]; j += BLOCK)
            a0[j] = P.x[n + ml.list[0][base + j]];
        __syncthreads();
        if (threadIdx.x == 0)
            for (int64_t j = 0; j < MON_CHUNK && base + j < ml.count[0]; j++) s0 += a0[j];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = s0 / (double)(uint64_t)ml.count[0];
    /* Force_y over the fixed-y nodes (every other node holds 0) */
    s0 = 0.0;
    for (int64_t base = 0; base < ml.nfy; base += MON_CHUNK) {
        for (int64_t j = threadIdx.x; j < MON_CHUNK && base + j < ml.nfy; j += BLOCK) {
            int64_t node = ml.fixed_y[base + j];
            a0[j] = Nd.active[node] ? Nd.fty[node] : 0.0;
        }
        __syncthreads();
        if (threadIdx.x == 0)
            for (int64_t j = 0; j < MON_CHUNK && base + j < ml.nfy; j++) s0 += a0[j];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[1] = s0;
    /* momentum_y */
    s0 = 0.0;
    for (int64_t base = 0; base < ml.count[1]; base += MON_CHUNK) {
        for (int64_t j = threadIdx.x; j < MON_CHUNK && base + j < ml.count[1]; j += BLOCK) {
            int64_t q = ml.list[1][base + j];
            a0[j] = P.v[n + q] * P.m[q];
        }
        __syncthreads();
        if (threadIdx.x == 0)
            for (int64_t j = 0; j < MON_CHUNK && base + j < ml.count[1]; j++) s0 += a0[j];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[2] = s0;
    /* energies */
    s0 = s1 = s2 = 0.0;
    for (int64_t base = 0; base < ml.count[2]; base += MON_CHUNK) {
        for (int64_t j = threadIdx.x; j < MON_CHUNK && base + j < ml.count[2]; j += BLOCK) {
            int64_t q = ml.list[2][base + j];
            double vx = P.v[q], vy = P.v[n + q], vz = P.v[2 * n + q];
            double len = sqrt(vx * vx + vy * vy + vz * vz);
            a0[j] = 0.5 * P.m[q] * (len * len);
            a1[j] = P.Ws[q];
            a2[j] = P.Wp[q];
        }
        __syncthreads();
        if (threadIdx.x == 0)
            for (int64_t j = 0; j < MON_CHUNK && base + j < ml.count[2]; j++) {
                s0 += a0[j];
                s1 += a1[j];
                s2 += a2[j];
            }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[3] = s0;
        out[4] = s1;
        out[5] = s2;
    }
    /* stress monitor means */
    s0 = s1 = 0.0;
    for (int64_t base = 0; base < ml.count[3]; base += MON_CHUNK) {
        for (int64_t j = threadIdx.x; j < MON_CHUNK && base + j < ml.count[3]; j += BLOCK) {
            int64_t q = ml.list[3][base + j];
            a0[j] = P.sig[n + q];
            a1[j] = P.eps[n + q];
        }
        __syncthreads();
        if (threadIdx.x == 0)
            for (int64_t j = 0; j < MON_CHUNK && base + j < ml.count[3]; j++) {
                s0 += a0[j];
                s1 += a1[j];
            }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double c = (double)(uint64_t)ml.count[3];
        out[6] = s0 / c;
        out[7] = s1 / c;
    }
}

/* Node download view: zero where the node was not touched in the last step. */
__global__ __launch_bounds__(BLOCK) void k_node_export(NView Nd, int field, int64_t nn, double *dst) {
    int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= nn) return;
    const bool act = Nd.active[i] != 0;
    switch (field) {
    case 0: dst[i] = act ? Nd.m[i] : 0.0; break;
    case 1:
        for (int a = 0; a < 3; a++) dst[a * nn + i] = act ? Nd.p[a * nn + i] : 0.0;
        break;
    case 2:
        for (int a = 0; a < 3; a++) dst[a * nn + i] = act ? Nd.v[a * nn + i] : 0.0;
        break;
    case 3:
        for (int a = 0; a < 3; a++) dst[a * nn + i] = act ? Nd.f[a * nn + i] : 0.0;
        break;
    case 4:
        dst[i] = 0.0;
        dst[nn + i] = act ? Nd.fty[i] : 0.0;
        dst[2 * nn + i] = 0.0;
        break;
    case 5: dst[i] = act ? 1.0 : 0.0; break;
    }
}

/* Stand-alone return maps over n points (row-major n x 6). */
__global__ __launch_bounds__(BLOCK) void k_constitutive(MatTable mt, const double *R, const double *sig,
                                                       const double *de, double *dsig, double *depsp, double *dR,
                                                       int64_t n) {
    int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    double s[6], d[6];
    for (int k = 0; k < 6; k++) {
        s[k] = sig[6 * i + k];
        d[k] = de[6 * i + k];
    }
    CRResult o;
    constitutive(mt, 0, R[i], s, d, o);
    for (int k = 0; k < 6; k++) {
        dsig[6 * i + k] = o.dsig[k];
        depsp[6 * i + k] = o.depsp[k];
    }
    dR[i] = o.dR;
}

/* ---- launchers ---- */

static inline unsigned blocks_for(int64_t n, int64_t per) { return (unsigned)((n + per - 1) / per); }

hipError_t launch_keys(const GridGeom &g, const PView &P, const BView &B, hipStream_t s) {
    if (P.n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_keys, dim3(blocks_for(P.n, BLOCK)), dim3(BLOCK), 0, s, g, P, B);
    return hipGetLastError();
}

int64_t scan_block_count(int64_t necells) { return (necells + SCAN_TILE - 1) / SCAN_TILE; }

hipError_t launch_bin(const GridGeom &g, const PView &P, const BView &B, uint32_t *block_sum, hipStream_t s) {
    int64_t nb = scan_block_count(B.necells);
    hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nb), dim3(BLOCK), 0, s, B.count, B.necells, block_sum);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(BLOCK), 0, s, block_sum, nb);
    hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nb), dim3(BLOCK), 0, s, B.count, B.necells, block_sum, B.start);
    if (P.n > 0) hipLaunchKernelGGL(k_scatter, dim3(blocks_for(P.n, BLOCK)), dim3(BLOCK), 0, s, P, B);
    hipLaunchKernelGGL(k_cell_sort, dim3(blocks_for(B.necells, BLOCK)), dim3(BLOCK), 0, s, B);
    return hipGetLastError();
}

static unsigned node_tiles(const GridGeom &g) {
    int64_t tnx = (g.nn[0] + TILE_X - 1) / TILE_X, tny = (g.nn[1] + TILE_Y - 1) / TILE_Y,
            tnz = (g.nn[2] + TILE_Z - 1) / TILE_Z;
    return (unsigned)(tnx * tny * tnz);
}

hipError_t launch_step_kernels(const GridGeom &g, const PView &P, const NView &N, const BView &B,
                               const StepParams &sp, const MatTable &mt, unsigned long long *health,
                               hipStream_t s) {
    unsigned tiles = node_tiles(g);
    hipLaunchKernelGGL(k_p2g_grid, dim3(tiles), dim3(BLOCK), 0, s, g, P, N, B, sp);
    if (P.n > 0)
        hipLaunchKernelGGL(k_g2p_velocity, dim3(blocks_for(P.n, BLOCK)), dim3(BLOCK), 0, s, g, P, N, sp);
    hipLaunchKernelGGL(k_remap_grid, dim3(tiles), dim3(BLOCK), 0, s, g, P, N, B, sp);
    if (P.n > 0)
        hipLaunchKernelGGL(k_g2p_strain, dim3(blocks_for(P.n, BLOCK)), dim3(BLOCK), 0, s, g, P, N, B, sp, mt,
                           health);
    return hipGetLastError();
}

hipError_t launch_monitors(const GridGeom &g, const PView &P, const NView &N, const MonLists &ml, double *out,
                           hipStream_t s) {
    hipLaunchKernelGGL(k_monitors, dim3(1), dim3(BLOCK), 0, s, g, P, N, ml, out);
    return hipGetLastError();
}

hipError_t launch_node_export(const NView &N, int field, int64_t nn, double *dst, hipStream_t s) {
    if (nn == 0) return hipSuccess;
    hipLaunchKernelGGL(k_node_export, dim3(blocks_for(nn, BLOCK)), dim3(BLOCK), 0, s, N, field, nn, dst);
    return hipGetLastError();
}

hipError_t launch_constitutive(const MatTable &mt, const double *R, const double *sig, const double *de,
                               double *dsig, double *depsp, double *dR, int64_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_constitutive, dim3(blocks_for(n, BLOCK)), dim3(BLOCK), 0, s, mt, R, sig, de, dsig, depsp,
                       dR, n);
    return hipGetLastError();
}
