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

#include <algorithm>

#define BLOCK 256

/* Blocks are dealt round-robin to the 8 XCDs; give each XCD one contiguous
 * range of tiles so neighbouring tiles share that XCD's L2. */
__device__ __forceinline__ uint32_t xcd_remap(uint32_t bid, uint32_t nblk) {
    uint32_t q = nblk / 8u, r = nblk % 8u;
    uint32_t x = bid % 8u, k = bid / 8u;
    return (x < r) ? x * (q + 1u) + k : r * (q + 1u) + (x - r) * q + k;
}

__device__ __forceinline__ uint32_t inst_of_particle(const Batch &bt, uint32_t i) {
    return bt.pinst ? bt.pinst[i] : 0u;
}

/* instance owning global node tile t: last k with tile_off[k] <= t */
__device__ __forceinline__ int inst_of_tile(const Batch &bt, uint32_t t) {
    int lo = 0, hi = bt.ninst - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (bt.tile_off[mid] <= t)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

/* Instance lookups for the heavy kernels, specialised at launch on whether
 * the batch has more than one instance.  With one, the instance is the
 * constant 0: its table row is read through scalar loads into SGPRs instead
 * of occupying vector registers of every lane. */
template <bool MULTI>
__device__ __forceinline__ const Inst &inst_p(const Batch &bt, uint32_t i) {
    return MULTI ? bt.inst[bt.pinst[i]] : bt.inst[0];
}
template <bool MULTI>
__device__ __forceinline__ const Inst &inst_t(const Batch &bt, uint32_t t) {
    return MULTI ? bt.inst[inst_of_tile(bt, t)] : bt.inst[0];
}

/* This thread's particle and instance in the per-particle kernels that
 * keep the instance row in scalar registers.  One instance: the particle
 * from the (XCD-remapped) block and thread, instance 0.  A batch: each wave
 * takes the 64 consecutive particles of one instance that its wave-table
 * entry names, so the instance index is wave-uniform and its row is read
 * through scalar loads, as for one instance.  Returns false for a wave
 * beyond the table (such a wave has no particle and may leave at once:
 * these kernels have no block-wide barrier). */
template <bool MULTI>
__device__ __forceinline__ bool particle_of_thread(const Batch &bt, uint32_t n, uint32_t &i, bool &live,
                                                   uint32_t &inst) {
    if (!MULTI) {
        i = xcd_remap(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
        live = i < n;
        inst = 0;
        return true;
    }
    const uint32_t w = (uint32_t)__builtin_amdgcn_readfirstlane(
        (int)(xcd_remap(blockIdx.x, gridDim.x) * (BLOCK / 64) + (threadIdx.x >> 6)));
    if (w >= bt.nwaves) return false;
    inst = bt.wave[2 * (size_t)w];
    const uint32_t first = bt.wave[2 * (size_t)w + 1];
    const Inst &I = bt.inst[inst];
    i = first + (threadIdx.x & 63u);
    live = (int64_t)i < I.part_off + I.npart;
    return true;
}

/* TimeLine::getVelocity (TimeLine.h:33-46): the last interval [t_i, t_i+1)
 * containing the time wins; zero outside every interval. */
__device__ __forceinline__ void timeline_velocity(const double *tl, int off, int count, double time, double out[3]) {
    out[0] = out[1] = out[2] = 0.0;
    for (int i = 0; i + 1 < count; i++) {
        const double *a = tl + 4 * (off + i), *b = a + 4;
        if (a[0] <= time && time < b[0]) {
            double s = (time - a[0]) / (b[0] - a[0]);
            for (int k = 0; k < 3; k++) out[k] = (1.0 - s) * a[1 + k] + (s)*b[1 + k];
        }
    }
}

__global__ void k_prepare(Batch bt) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= bt.ninst) return;
    Inst &I = bt.inst[k];
    double vt[3];
    timeline_velocity(bt.tl, I.tl_off, I.tl_count, I.time, vt);
    I.vt[0] = vt[0];
    I.vt[1] = vt[1];
    I.vt[2] = vt[2];
}

__global__ void k_advance(Batch bt) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= bt.ninst) return;
    bt.inst[k].time += bt.inst[k].dt;
}

__global__ void k_set_dt(Batch bt, int k, double dt) {
    if (threadIdx.x == 0 && blockIdx.x == 0) bt.inst[k].dt = dt;
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

/* A slot in an append-only queue, one atomic per block: lanes with want
 * set take consecutive slots in thread order after the block's base.  Every
 * thread of the block must call it (it synchronises the block). */
template <typename T>
__device__ __forceinline__ T block_append(T *counter, bool want) {
    __shared__ T s_off[BLOCK / 64], s_base;
    const uint64_t mask = __ballot(want);
    const int lane = (int)(threadIdx.x & 63), wave = (int)(threadIdx.x >> 6);
    if (lane == 0) s_off[wave] = (T)__popcll(mask);
    __syncthreads();
    if (threadIdx.x == 0) {
        T tot = 0;
        for (int w = 0; w < BLOCK / 64; w++) {
            const T c = s_off[w];
            s_off[w] = tot;
            tot += c;
        }
        s_base = tot ? atomicAdd(counter, tot) : (T)0;
    }
    __syncthreads();
    const T slot = s_base + s_off[wave] + (T)__popcll(mask & ((1ull << lane) - 1ull));
    __syncthreads();
    return slot;
}

/* Runs of equal keys over adjacent lanes of a wave: each lane's run head
 * (first lane) and run length, from one ballot.  Neighbouring particle ids
 * mostly share a cell, so a wave holds few runs and their heads can issue
 * one atomic each in a single wave-wide instruction.  A key that shows up in
 * two separate runs gets two atomics, which is still exact.  Every lane of
 * the wave must call it. */
struct KeyRun {
    int head;
    uint32_t len;
};

__device__ __forceinline__ KeyRun wave_key_run(uint32_t key) {
    const int lane = (int)(threadIdx.x & 63);
    const uint32_t prev = __shfl_up(key, 1, 64);
    const uint64_t starts = __ballot(lane == 0 || prev != key);
    const uint64_t upto = lane == 63 ? ~0ull : ((1ull << (lane + 1)) - 1ull);
    const uint64_t above = starts & ~upto;
    KeyRun r;
    r.head = 63 - __clzll((long long)(starts & upto)); /* bit 0 is always set */
    r.len = (uint32_t)((above ? __ffsll((unsigned long long)above) - 1 : 64) - r.head);
    return r;
}

/* count[key] += 1 for every lane with want set, one atomic per run of equal
 * keys.  Every lane of the wave must call it. */
__device__ __forceinline__ void wave_key_add(uint32_t *count, uint32_t key, bool want) {
    const uint32_t k = want ? key : MPM_NONE;
    const KeyRun r = wave_key_run(k);
    if (want && (int)(threadIdx.x & 63) == r.head) atomicAdd(&count[k], r.len);
}

__global__ __launch_bounds__(BLOCK) void k_keys(Batch bt, PView P, BView B) {
    uint32_t i = xcd_remap(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
    uint32_t k = MPM_NONE;
    if (i < P.n) {
        const Inst &I = bt.inst[inst_of_particle(bt, i)];
        k = cell_key(I.g, P.x[i], P.x[P.n + i], P.x[2 * P.n + i]);
        if (k != MPM_NONE) k += (uint32_t)I.cell_off;
        B.key[i] = k;
    }
    wave_key_add(B.count, k, k != MPM_NONE);
}

/* ---- exclusive scan of cell counts (reduce, scan of block sums, apply) ---- */

#define SCAN_ITEMS 16
#define SCAN_TILE (BLOCK * SCAN_ITEMS)

/* Cell counts are scanned as one 64-bit value per cell: particle count in
 * the low word, 1 if the cell is occupied in the high word.  The exclusive
 * prefix gives the cell's first slot in the sorted id list (low word) and
 * its rank among occupied cells (high word). */
__device__ __forceinline__ uint64_t pack_count(uint32_t c) { return (uint64_t)c | ((uint64_t)(c != 0u) << 32); }

__device__ __forceinline__ uint64_t block_exclusive_scan(uint64_t v, uint64_t *lds, uint64_t &total) {
    /* wave-level inclusive scan, then across the 4 waves */
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint64_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint64_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) lds[wave] = x;
    __syncthreads();
    uint64_t off = 0;
    for (int w = 0; w < wave; w++) off += lds[w];
    total = lds[0] + lds[1] + lds[2] + lds[3];
    __syncthreads();
    return off + x - v;
}

/* The count, start and slot arrays are padded to whole tiles (zero counts
 * past the last cell), so every load and store below is an unconditional
 * 16-byte access.  A tile is SCAN_SUB sub-tiles of BLOCK x 4 cells; thread t
 * holds cells 4t..4t+3 of each sub-tile, so each of the four loads of a wave
 * reads one contiguous KiB. */
#define SCAN_SUB (SCAN_ITEMS / 4)

__device__ __forceinline__ uint64_t pack_sum4(const uint4 &c) {
    return pack_count(c.x) + pack_count(c.y) + pack_count(c.z) + pack_count(c.w);
}

__global__ __launch_bounds__(BLOCK) void k_scan_reduce(const uint32_t *count, uint64_t *block_sum) {
    __shared__ uint64_t lds[4];
    const uint4 *c4 = reinterpret_cast<const uint4 *>(count + (int64_t)blockIdx.x * SCAN_TILE) + threadIdx.x;
    uint4 c[SCAN_SUB];
#pragma unroll
    for (int q = 0; q < SCAN_SUB; q++) c[q] = c4[q * BLOCK];
    uint64_t s = 0;
#pragma unroll
    for (int q = 0; q < SCAN_SUB; q++) s += pack_sum4(c[q]);
    uint64_t total;
    block_exclusive_scan(s, lds, total);
    if (threadIdx.x == 0) block_sum[blockIdx.x] = total;
}

/* One block scans all block sums in place (exclusive) in chunks of BLOCK. */
__global__ __launch_bounds__(BLOCK) void k_scan_blocks(uint64_t *block_sum, int64_t nblocks) {
    __shared__ uint64_t lds[4];
    uint64_t carry = 0;
    for (int64_t base = 0; base < nblocks; base += BLOCK) {
        int64_t i = base + threadIdx.x;
        uint64_t v = (i < nblocks) ? block_sum[i] : 0u;
        uint64_t total;
        uint64_t ex = block_exclusive_scan(v, lds, total);
        if (i < nblocks) block_sum[i] = carry + ex;
        carry += total;
    }
}

__device__ __forceinline__ void scan_emit(uint32_t c, int64_t cell, uint64_t &ex, uint32_t &st, uint32_t &sl,
                                          uint32_t *occ, uint32_t *ostart) {
    st = (uint32_t)ex;
    const uint32_t r = (uint32_t)(ex >> 32);
    sl = c ? r : MPM_NONE;
    if (c) {
        occ[r] = (uint32_t)cell;
        ostart[r] = st;
    }
    ex += pack_count(c);
}

/* Cell starts, slots, the occupied-cell list and the occupied cells' starts
 * by rank (so that a group of consecutive occupied cells finds its particle
 * ranges with one load each, not a cell lookup then a start lookup).  The
 * last thread of the last block also writes start[padded], ostart[nocc]
 * and the occupied count, which are the grand totals (start[necells]
 * equals them too, padding cells being empty). */
__global__ __launch_bounds__(BLOCK) void k_scan_apply(const uint32_t *count, const uint64_t *block_sum,
                                                     uint32_t *start, uint32_t *slot, uint32_t *occ, uint32_t *ostart,
                                                     uint32_t *nocc) {
    __shared__ uint64_t lds[SCAN_SUB][4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t tile = (int64_t)blockIdx.x * SCAN_TILE;
    const uint4 *c4 = reinterpret_cast<const uint4 *>(count + tile) + threadIdx.x;
    uint4 c[SCAN_SUB];
#pragma unroll
    for (int q = 0; q < SCAN_SUB; q++) c[q] = c4[q * BLOCK];
    uint64_t s[SCAN_SUB], x[SCAN_SUB];
#pragma unroll
    for (int q = 0; q < SCAN_SUB; q++) {
        s[q] = pack_sum4(c[q]);
        x[q] = s[q];
    }
    /* wave-level inclusive scans of the SCAN_SUB sub-tiles side by side */
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
#pragma unroll
        for (int q = 0; q < SCAN_SUB; q++) {
            const uint64_t y = __shfl_up(x[q], o, 64);
            if (lane >= o) x[q] += y;
        }
    }
    if (lane == 63) {
#pragma unroll
        for (int q = 0; q < SCAN_SUB; q++) lds[q][wave] = x[q];
    }
    __syncthreads();
    uint64_t off = block_sum[blockIdx.x];
    uint4 *st4 = reinterpret_cast<uint4 *>(start + tile) + threadIdx.x;
    uint4 *sl4 = reinterpret_cast<uint4 *>(slot + tile) + threadIdx.x;
#pragma unroll
    for (int q = 0; q < SCAN_SUB; q++) {
        uint64_t before = 0, sub = 0;
#pragma unroll
        for (int w = 0; w < 4; w++) {
            const uint64_t t = lds[q][w];
            before += w < wave ? t : 0u;
            sub += t;
        }
        uint64_t ex = off + before + x[q] - s[q];
        const int64_t cell = tile + (int64_t)q * BLOCK * 4 + (int64_t)threadIdx.x * 4;
        uint4 st, sl;
        scan_emit(c[q].x, cell + 0, ex, st.x, sl.x, occ, ostart);
        scan_emit(c[q].y, cell + 1, ex, st.y, sl.y, occ, ostart);
        scan_emit(c[q].z, cell + 2, ex, st.z, sl.z, occ, ostart);
        scan_emit(c[q].w, cell + 3, ex, st.w, sl.w, occ, ostart);
        st4[q * BLOCK] = st;
        sl4[q * BLOCK] = sl;
        off += sub;
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == BLOCK - 1) {
        start[tile + SCAN_TILE] = (uint32_t)off;
        ostart[off >> 32] = (uint32_t)off;
        *nocc = (uint32_t)(off >> 32);
    }
}

/* Ascending sort of N keys held in registers (bitonic network: every index
 * is a compile-time constant, so the keys never leave registers). */
template <int N>
__device__ __forceinline__ void bitonic_sort(uint32_t (&v)[N]) {
#pragma unroll
    for (int size = 2; size <= N; size <<= 1)
#pragma unroll
        for (int stride = size >> 1; stride > 0; stride >>= 1)
#pragma unroll
            for (int i = 0; i < N; i++) {
                const int j = i ^ stride;
                if (j > i) {
                    const bool up = (i & size) == 0;
                    const uint32_t lo = min(v[i], v[j]), hi = max(v[i], v[j]);
                    v[i] = up ? lo : hi;
                    v[j] = up ? hi : lo;
                }
            }
}

/* The ids of occupied cell t (rank) in ascending order.  Up to SORT_REG ids
 * are loaded together, sorted in registers and stored back; a larger cell
 * takes an insertion sort in place. */
#define SORT_REG 16
__device__ __forceinline__ void sort_cell(const BView &B, uint32_t t) {
    const uint32_t s = B.ostart[t], e = B.ostart[t + 1];
    if (e - s < 2) return;
    if (e - s <= SORT_REG) {
        uint32_t v[SORT_REG];
#pragma unroll
        for (int k = 0; k < SORT_REG; k++) v[k] = s + k < e ? B.sorted[s + k] : MPM_NONE;
        bitonic_sort(v); /* MPM_NONE (no particle id) sorts last */
#pragma unroll
        for (int k = 0; k < SORT_REG; k++)
            if (s + k < e) B.sorted[s + k] = v[k];
        return;
    }
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


/* Particle ids into their cell's range of the sorted list.  A wave takes
 * SCATTER_RUNS groups of 64 consecutive particles; in each, every run of
 * equal keys takes consecutive slots in id order with one atomic decrement
 * of the cell's count (which also leaves the counts at zero for the next
 * step).  The wave's key loads, then its atomics and start loads, are all
 * issued before any result is used, so their latencies overlap (one group
 * per wave waited for each in turn).  k_cell_sort orders what different
 * runs put in one cell. */
#define SCATTER_RUNS 4
__global__ __launch_bounds__(BLOCK) void k_scatter(PView P, BView B) {
    const int lane = (int)(threadIdx.x & 63);
    const uint32_t w = xcd_remap(blockIdx.x, gridDim.x) * (BLOCK / 64) + (threadIdx.x >> 6);
    const uint32_t base = w * 64u * SCATTER_RUNS;
    uint32_t id[SCATTER_RUNS], k[SCATTER_RUNS], top[SCATTER_RUNS], st[SCATTER_RUNS];
    int32_t dcr[SCATTER_RUNS];
#pragma unroll
    for (int u = 0; u < SCATTER_RUNS; u++) {
        id[u] = base + (uint32_t)(u * 64 + lane);
        const bool live = id[u] < P.n;
        k[u] = live ? B.key[id[u]] : MPM_NONE;
        dcr[u] = live ? P.dc_rank[id[u]] : -1;
    }
    KeyRun r[SCATTER_RUNS];
#pragma unroll
    for (int u = 0; u < SCATTER_RUNS; u++) {
        r[u] = wave_key_run(k[u]);
        top[u] = 0;
        if (k[u] != MPM_NONE && lane == r[u].head) top[u] = atomicSub(&B.count[k[u]], r[u].len);
        st[u] = k[u] != MPM_NONE ? B.start[k[u]] : 0u;
    }
#pragma unroll
    for (int u = 0; u < SCATTER_RUNS; u++) {
        top[u] = __shfl(top[u], r[u].head, 64);
        if (k[u] == MPM_NONE) continue;
        B.sorted[st[u] + top[u] - r[u].len + (uint32_t)(lane - r[u].head)] = id[u];
        if (dcr[u] >= 0) atomicOr(&B.slot[k[u]], SLOT_DC); /* the cell holds a controlled particle */
    }
}

/* Ascending ids inside each occupied cell: the order every node sum uses. */
__global__ __launch_bounds__(BLOCK) void k_cell_sort(BView B) {
    const uint32_t nocc = *B.nocc;
    for (uint32_t t = blockIdx.x * BLOCK + threadIdx.x; t < nocc; t += gridDim.x * BLOCK) sort_cell(B, t);
}
/* ---- 8-way merge of a node's adjacent cell lists in ascending id order ---- */

/* s / p / e: the start, next and end position of each cell's list in the
 * sorted ids, h: the id at p (MPM_NONE past the end); dc: bit k set when
 * cell k holds a displacement-controlled particle (its slot's SLOT_DC
 * flag). */
struct Merge8 {
    uint32_t s[8], p[8], e[8], h[8];
    uint32_t dc;
};

/* Every list's bounds and slot in one round of loads, then every head in a
 * second (a round per cell would chain 8 waits). */
__device__ __forceinline__ void merge_init(Merge8 &M, const BView &B, const GridGeom &g, int64_t cell_off, int ix,
                                           int iy, int iz) {
    /* node i touches cells i-1 and i per axis = ext cells i and i+1 */
    uint32_t sl[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
        int cx = ix + ((k >> 2) & 1), cy = iy + ((k >> 1) & 1), cz = iz + (k & 1);
        uint32_t c = (uint32_t)cell_off +
                     (uint32_t)cx + (uint32_t)g.ec[0] * ((uint32_t)cy + (uint32_t)g.ec[1] * (uint32_t)cz);
        M.s[k] = M.p[k] = B.start[c];
        M.e[k] = B.start[c + 1];
        sl[k] = B.slot[c];
    }
    M.dc = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        M.h[k] = (M.p[k] < M.e[k]) ? B.sorted[M.p[k]] : MPM_NONE;
        M.dc |= (sl[k] != MPM_NONE && (sl[k] & SLOT_DC)) ? 1u << k : 0u;
    }
}

__device__ __forceinline__ uint32_t merge_min(const Merge8 &M) {
    uint32_t a = min(min(M.h[0], M.h[1]), min(M.h[2], M.h[3]));
    uint32_t b = min(min(M.h[4], M.h[5]), min(M.h[6], M.h[7]));
    return min(a, b);
}

/* The list holding id (the minimum head): its index k, the id's position
 * and the list's start and length. */
__device__ __forceinline__ void merge_find(const Merge8 &M, uint32_t id, uint32_t &k, uint32_t &pos, uint32_t &s0,
                                           uint32_t &cnt) {
    k = 0;
    pos = s0 = cnt = 0;
#pragma unroll
    for (int j = 0; j < 8; j++)
        if (M.h[j] == id) {
            k = (uint32_t)j;
            pos = M.p[j];
            s0 = M.s[j];
            cnt = M.e[j] - M.s[j];
        }
}

/* Past the head of list k: one load (the list's next id, or again the
 * current one at the end of the list), not under a branch, so that it
 * issues with the caller's other loads of the step and they wait together. */
__device__ __forceinline__ void merge_step(Merge8 &M, const BView &B, uint32_t k, uint32_t pos) {
    uint32_t e = 0;
#pragma unroll
    for (int j = 0; j < 8; j++)
        if ((uint32_t)j == k) e = M.e[j];
    const uint32_t q = pos + 1u;
    const uint32_t nx = B.sorted[q < e ? q : pos];
    const uint32_t h = q < e ? nx : MPM_NONE;
#pragma unroll
    for (int j = 0; j < 8; j++)
        if ((uint32_t)j == k) {
            M.p[j] = q;
            M.h[j] = h;
        }
}

/* Inverse of node_of_tile: the instance-local tile holding node (ix, iy, iz). */
__device__ __forceinline__ uint32_t tile_of_node(const Batch &bt, const GridGeom &g, int ix, int iy, int iz) {
    const uint32_t tnx = (uint32_t)((g.nn[0] + bt.tx - 1) / bt.tx), tny = (uint32_t)((g.nn[1] + bt.ty - 1) / bt.ty);
    return (uint32_t)(ix / bt.tx) + tnx * ((uint32_t)(iy / bt.ty) + tny * (uint32_t)(iz / bt.tz));
}

__device__ __forceinline__ void node_of_tile(const Batch &bt, const GridGeom &g, uint32_t tile, int &ix, int &iy,
                                             int &iz) {
    const int TX = bt.tx, TY = bt.ty;
    int tnx = (g.nn[0] + TX - 1) / TX, tny = (g.nn[1] + TY - 1) / TY;
    int tx = (int)(tile % (uint32_t)tnx);
    int ty = (int)((tile / (uint32_t)tnx) % (uint32_t)tny);
    int tz = (int)(tile / ((uint32_t)tnx * (uint32_t)tny));
    ix = tx * TX + (int)(threadIdx.x % TX);
    iy = ty * TY + (int)((threadIdx.x / TX) % TY);
    iz = tz * bt.tz + (int)(threadIdx.x / (TX * TY));
}


/* Track, per axis, the displacement-controlled particle latest in the
 * control list among those touching the node (the last writer of
 * runSimulation_DisplacementControl, DP:143-174, when run in list order). */
struct DCWin {
    int rank[3];
    double v[3];
};

/* A controlled particle claims its flagged axes (DoublePass, DP:143-174) or
 * all three (DPC:283-297, SP:251-268); the highest list rank wins. */
__device__ __forceinline__ void dc_consider(DCWin &w, const PView &P, uint32_t id, bool all_axes) {
    int r = P.dc_rank[id];
    if (r < 0) return;
    uint8_t ax = all_axes ? (uint8_t)7 : P.dc_axes[id];
#pragma unroll
    for (int a = 0; a < 3; a++)
        if ((ax >> a) & 1)
            if (r > w.rank[a]) {
                w.rank[a] = r;
                w.v[a] = P.v[(size_t)a * P.n + id];
            }
}

/* Packed node records read by the grid-to-particle passes (one 32 / 64
 * byte access per adjacent node instead of seven scattered ones). */
__device__ __forceinline__ void store_g1(const NView &Nd, int64_t node, double m, const double f[3]) {
    double2 *o = reinterpret_cast<double2 *>(Nd.g1 + (size_t)node * 4u);
    o[0] = make_double2(m, f[0] / m); /* hoisted f/m of DP:508 */
    o[1] = make_double2(f[1] / m, f[2] / m);
}
__device__ __forceinline__ void store_g2(const NView &Nd, int64_t node, double m, const double p[3],
                                         const double v[3]) {
    double2 *o = reinterpret_cast<double2 *>(Nd.g2 + (size_t)node * 8u);
    o[0] = make_double2(m, p[0] / m); /* hoisted p/m of DP:198 */
    o[1] = make_double2(p[1] / m, p[2] / m);
    o[2] = make_double2(v[0], v[1]);
    o[3] = make_double2(v[2], 0.0);
}

/* What the node updates read besides the sums: instance scalars, the
 * node's fixed mask and friction (P2G), its mass and momentum (remap) and,
 * for the single pass, its force sums.  Callers read it before their first
 * store, with the node's cell slots, so that every load is in flight at
 * once (a load after a store cannot be proven unaliased and waits on its
 * own). */
struct NodeIn {
    double dt, damping, mass_min;
    double friction;
    double m, p[3], f[3];
    uint8_t fixed;
};

template <bool REMAP>
__device__ __forceinline__ NodeIn node_in(const Batch &bt, const Inst &I, const NView &Nd, int64_t node) {
    const size_t nn = (size_t)bt.nnodes;
    NodeIn r;
    r.dt = I.dt;
    r.damping = I.damping;
    r.mass_min = I.mass_min;
    r.fixed = Nd.fixed[node];
    r.friction = 0.0;
    r.m = 0.0;
#pragma unroll
    for (int a = 0; a < 3; a++) r.p[a] = r.f[a] = 0.0;
    if (!REMAP) {
        if (bt.variant == VAR_DP) r.friction = Nd.friction[node];
    } else {
        r.m = Nd.m[node];
#pragma unroll
        for (int a = 0; a < 3; a++) r.p[a] = Nd.p[a * nn + node];
        if (bt.variant == VAR_SP) {
#pragma unroll
            for (int a = 0; a < 3; a++) r.f[a] = Nd.f[a * nn + node];
        }
    }
    return r;
}

/* DC override then runSimulation_IntegrateGrid (DP:90-142) on one node's
 * sums, in registers; stores m, p, f, f/m and the reaction. */
__device__ __forceinline__ void node_integrate(const NodeIn &in, const NView &Nd, int64_t node, size_t nn, double m,
                                               double p[3], double f[3], const DCWin &w) {
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
    const double dt = in.dt;
    {
        double pp = p[0] * p[0] + p[1] * p[1] + p[2] * p[2];
        if (sqrt(pp) > 1.0e-24) {
            double inv = 1.0 / sqrt(pp);
            double nx = p[0] * inv, ny = p[1] * inv, nz = p[2] * inv;
            double sc = in.damping * fabs(f[0] * nx + f[1] * ny + f[2] * nz);
            f[0] -= sc * nx;
            f[1] -= sc * ny;
            f[2] -= sc * nz;
        }
    }
    const uint8_t fixed = in.fixed;
    double fty = 0.0;
    if (fixed & 1) {
        double pty = p[1] + f[1] * dt;
        if (fabs(pty * dt) > 1.0e-24 && f[0] > 0.0) {
            double ffy = in.friction * glm_sign(-pty) * fabs(f[0]);
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
    Nd.active[node] = 1;
    Nd.m[node] = m;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        Nd.p[a * nn + node] = p[a];
        Nd.f[a * nn + node] = f[a];
    }
    store_g1(Nd, node, m, f);
    Nd.fty[node] = fty;
}

/* Grid update of the contact loop on one node's sums, DPC:242-297:
 * damping along p (as DP:100-101), p += f dt, the fixed components zeroed
 * with the y reaction taken first (no friction), then displacement control
 * on all axes.  Stores m, p, f, f/m and the reaction. */
__device__ __forceinline__ void node_integrate_contact(const NodeIn &in, const NView &Nd, int64_t node, size_t nn,
                                                       double m, double p[3], double f[3], const DCWin &w) {
    const double dt = in.dt;
    {
        double pp = p[0] * p[0] + p[1] * p[1] + p[2] * p[2];
        if (sqrt(pp) > 1.0e-24) {
            double inv = 1.0 / sqrt(pp);
            double nx = p[0] * inv, ny = p[1] * inv, nz = p[2] * inv;
            double sc = in.damping * fabs(f[0] * nx + f[1] * ny + f[2] * nz);
            f[0] -= sc * nx;
            f[1] -= sc * ny;
            f[2] -= sc * nz;
        }
    }
#pragma unroll
    for (int a = 0; a < 3; a++) p[a] += f[a] * dt;
    const uint8_t fixed = in.fixed;
    double fty = 0.0;
    if (fixed & 1) {
        p[0] = 0.0;
        f[0] = 0.0;
    }
    if (fixed & 2) {
        p[1] = 0.0;
        fty = 0.0 + f[1];
        f[1] = 0.0;
    }
    if (fixed & 4) {
        p[2] = 0.0;
        f[2] = 0.0;
    }
#pragma unroll
    for (int a = 0; a < 3; a++)
        if (w.rank[a] >= 0) {
            p[a] = m * w.v[a];
            f[a] = 0.0;
        }
    Nd.active[node] = 1;
    Nd.m[node] = m;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        Nd.p[a * nn + node] = p[a];
        Nd.f[a * nn + node] = f[a];
    }
    store_g1(Nd, node, m, f);
    Nd.fty[node] = fty;
}

/* Single pass, first half of the grid work (SP:108-204): the node keeps
 * the sums; momentum only where the mass exceeds the minimum (SP:170-174). */
__device__ __forceinline__ void node_store_singlepass(const NodeIn &in, const NView &Nd, int64_t node, size_t nn,
                                                      double m, double p[3], double f[3]) {
    const bool massive = m > in.mass_min;
    Nd.active[node] = 1;
    Nd.m[node] = m;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        Nd.p[a * nn + node] = massive ? p[a] : 0.0;
        Nd.f[a * nn + node] = f[a];
    }
}

/* Single pass, grid update on the mass-weighted velocity v (SP:210-268):
 * damping along v with |f|, v += f / m dt above the mass minimum, p += f dt,
 * fixed components zeroed with the y reaction taken first, then
 * displacement control on all axes.  Reads the sums stored by
 * node_store_singlepass; stores v, p, f, f/m, p/m and the reaction. */
__device__ __forceinline__ void node_update_singlepass(const NodeIn &in, const NView &Nd, int64_t node, size_t nn,
                                                       double m, double v[3], const DCWin &w) {
    const double dt = in.dt;
    double p[3] = {in.p[0], in.p[1], in.p[2]};
    double f[3] = {in.f[0], in.f[1], in.f[2]};
    {
        double vv = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
        if (sqrt(vv) > 1.0e-9) {
            double lf = sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
            double inv = 1.0 / sqrt(vv);
            double nx = v[0] * inv, ny = v[1] * inv, nz = v[2] * inv;
            double sc = in.damping * lf;
            f[0] -= sc * nx;
            f[1] -= sc * ny;
            f[2] -= sc * nz;
        }
    }
    if (m > in.mass_min) {
#pragma unroll
        for (int a = 0; a < 3; a++) v[a] += f[a] / m * dt;
    }
#pragma unroll
    for (int a = 0; a < 3; a++) p[a] += f[a] * dt;
    const uint8_t fixed = in.fixed;
    double fty = 0.0;
    if (fixed & 1) {
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
#pragma unroll
    for (int a = 0; a < 3; a++)
        if (w.rank[a] >= 0) {
            v[a] = w.v[a];
            p[a] = m * w.v[a];
            f[a] = 0.0;
        }
#pragma unroll
    for (int a = 0; a < 3; a++) {
        Nd.v[a * nn + node] = v[a];
        Nd.p[a * nn + node] = p[a];
        Nd.f[a * nn + node] = f[a];
    }
    store_g1(Nd, node, m, f);
    store_g2(Nd, node, m, p, v);
    Nd.fty[node] = fty;
}

/* The node side of the particle-to-grid phase for each loop. */
__device__ __forceinline__ void node_p2g_finish(const Batch &bt, const NodeIn &in, const NView &Nd, int64_t node,
                                                double m, double p[3], double f[3], const DCWin &w) {
    const size_t nn = (size_t)bt.nnodes;
    if (bt.variant == VAR_DPC)
        node_integrate_contact(in, Nd, node, nn, m, p, f, w);
    else if (bt.variant == VAR_SP)
        node_store_singlepass(in, Nd, node, nn, m, p, f);
    else
        node_integrate(in, Nd, node, nn, m, p, f, w);
}

/* G2P pass 1 for one particle: v_p += N * (f_i / m_i) * dt over the
 * adjacent nodes in GridPoint_Mediator order (x, then y, then z innermost),
 * DP:490-510. */
template <class REC>
__device__ __forceinline__ void g2p_velocity_update_rec(const Inst &I, REC rec, const double x[3], double v[3]) {
    const GridGeom &g = I.g;
    /* grid fields read once, outside any branch */
    const double gL[3] = {g.L[0], g.L[1], g.L[2]}, gh[3] = {g.h[0], g.h[1], g.h[2]};
    const int gc[3] = {g.cells[0], g.cells[1], g.cells[2]}, gn[3] = {g.nn[0], g.nn[1], g.nn[2]};
    int b[3], lo[3], hi[3];
    double s[3][2];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        base_cell(x[a], gh[a], gn[a], b[a], lo[a], hi[a]);
#pragma unroll
        for (int o = 0; o < 2; o++) s[a][o] = axis_shape(x[a], node_coord(b[a] + o, gL[a], gc[a]), gh[a]).s;
    }
    bool ok[3][2];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int o = 0; o < 2; o++)
            ok[a][o] = (o >= lo[a]) & (o < hi[a]) & (b[a] + o >= 0) & (b[a] + o <= gn[a] - 1);
    /* instance fields read once: inside the per-corner branches they would
     * be reloaded (and waited for) corner by corner */
    const double dt = I.dt, mass_min = I.mass_min;
    const size_t node_off = (size_t)I.node_off, nx = (size_t)gn[0], ny = (size_t)gn[1];
    /* branch-free over the 8 corners (off-grid corners read the instance's
     * first node and add nothing), so that all gathers can be in flight */
    double2 ra[8], rb[8];
    bool in[8];
#pragma unroll
    for (int c = 0; c < 8; c++) {
        const int ox = c >> 2, oy = (c >> 1) & 1, oz = c & 1;
        in[c] = ok[0][ox] & ok[1][oy] & ok[2][oz];
        const size_t node =
            in[c] ? node_off + (size_t)(b[0] + ox) + nx * ((size_t)(b[1] + oy) + ny * (size_t)(b[2] + oz)) : node_off;
        rec(c, node, ra[c], rb[c]);
    }
#pragma unroll
    for (int c = 0; c < 8; c++) {
        const int ox = c >> 2, oy = (c >> 1) & 1, oz = c & 1;
        const bool take = in[c] & (ra[c].x > mass_min);
        const double N = s[0][ox] * s[1][oy] * s[2][oz];
        const double t0 = v[0] + N * ra[c].y * dt, t1 = v[1] + N * rb[c].x * dt, t2 = v[2] + N * rb[c].y * dt;
        v[0] = take ? t0 : v[0];
        v[1] = take ? t1 : v[1];
        v[2] = take ? t2 : v[2];
    }
}

__device__ __forceinline__ void g2p_velocity_update(const Inst &I, const NView &Nd, const double x[3], double v[3]) {
    const double *g1 = Nd.g1;
    g2p_velocity_update_rec(I,
                            [=](int, size_t node, double2 &a, double2 &b) {
                                const double2 *r = reinterpret_cast<const double2 *>(g1 + node * 4u);
                                a = r[0];
                                b = r[1];
                            },
                            x, v);
}

template <bool MULTI>
__global__ __launch_bounds__(BLOCK) void k_g2p_velocity(Batch bt, PView P, NView Nd) {
    uint32_t i, k;
    bool live;
    if (!particle_of_thread<MULTI>(bt, (uint32_t)P.n, i, live, k) || !live) return;
    if (bt.variant != VAR_DP && P.dc_rank[i] >= 0) return; /* DPC:309-310, SP:279-280 */
    const size_t n = P.n;
    const double x[3] = {P.x[i], P.x[n + i], P.x[2 * n + i]};
    double v[3] = {P.v[i], P.v[n + i], P.v[2 * n + i]};
    g2p_velocity_update(bt.inst[k], Nd, x, v);
    P.v[i] = v[0];
    P.v[n + i] = v[1];
    P.v[2 * n + i] = v[2];
}

/* Second DC override (DP:568) and BC re-application (DP:573-599) on the
 * remapped velocity; stores v, the final p, zeroed forces and p/m. */
__device__ __forceinline__ void node_remap_finish(const NodeIn &in, const NView &Nd, int64_t node, size_t nn,
                                                  double m, double v[3], const DCWin &w) {
    double p[3] = {in.p[0], in.p[1], in.p[2]};
    const uint8_t fixed = in.fixed;
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
    }
    store_g2(Nd, node, m, p, v);
}

/* Contact loop: BC re-applied (DPC:380-411), then displacement control on
 * all axes (DPC:413-433), so control wins over fixed components. */
__device__ __forceinline__ void node_remap_finish_contact(const NodeIn &in, const NView &Nd, int64_t node, size_t nn,
                                                          double m, double v[3], const DCWin &w) {
    double p[3] = {in.p[0], in.p[1], in.p[2]};
    const uint8_t fixed = in.fixed;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        if ((fixed >> a) & 1) {
            v[a] = 0.0;
            p[a] = 0.0;
        }
        if (w.rank[a] >= 0) {
            v[a] = w.v[a];
            p[a] = m * w.v[a];
            Nd.f[a * nn + node] = 0.0;
        }
        Nd.v[a * nn + node] = v[a];
        Nd.p[a * nn + node] = p[a];
    }
    store_g2(Nd, node, m, p, v);
}

/* The node side of the remap phase for each loop (the single pass runs it
 * before its only grid-to-particle pass, as its velocity sum). */
__device__ __forceinline__ void node_remap_dispatch(const Batch &bt, const NodeIn &in, const NView &Nd, int64_t node,
                                                    double v[3], const DCWin &w);

/* Velocity remap (DP:539-564) gathered per node in ascending particle
 * order, then the second displacement-control override (DP:568) and the BC
 * re-application (DP:573-599). Stores the final v, p, the zeroed force
 * components and the hoisted p/m used by G2P pass 2 (DP:198). */
__device__ __forceinline__ void node_remap_dispatch(const Batch &bt, const NodeIn &in, const NView &Nd, int64_t node,
                                                    double v[3], const DCWin &w) {
    const size_t nn = (size_t)bt.nnodes;
    if (bt.variant == VAR_DPC)
        node_remap_finish_contact(in, Nd, node, nn, in.m, v, w);
    else if (bt.variant == VAR_SP)
        node_update_singlepass(in, Nd, node, nn, in.m, v, w);
    else
        node_remap_finish(in, Nd, node, nn, in.m, v, w);
}

/* ---- cell order: per (occupied cell, corner) partial sums, then nodes ----
 *
 * A node's sum is the sum, in a fixed order of its 8 adjacent cells, of
 * per-cell partial sums, each accumulated over the cell's particles in
 * ascending id.  Thread t handles corner t % 8 of occupied cell t / 8: the 8
 * lanes of a cell read the same particles, so loads stay coalesced.  The
 * cell -> corner mapping: local cell c = ext - 1, corner bits (x, y, z) =
 * (k >> 2, k >> 1, k) & 1, node = c + corner; the node sees that cell as its
 * (7 - k)-th neighbour. */

/* Each block takes groups of PART_CELLS consecutive occupied cells (8
 * lanes per cell, one per corner).  The group's particles are contiguous in
 * the sorted list; they are staged through LDS in chunks with one coalesced
 * pass, and every lane then sums its cell's particles for its corner from
 * LDS in ascending id, the same order and arithmetic as reading them from
 * memory.
 *
 * In the remap pass, what is common to a particle's 8 corners is computed
 * once, by the thread that stages it: the six axis shape factors s[a][o] of
 * its cell's corner nodes (axis a, offset o); corner (ox, oy, oz) takes
 * N = s_x[ox] s_y[oy] s_z[oz], the product the lane would form.  The staged
 * particle's base cell is its bin (floor(x / h), as the binning computed it
 * from the same x), so its corners are exactly the lanes' nodes.  In the
 * P2G pass, which also needs the twelve gradient quotients (s_j s_l) / h_k
 * (too much LDS per staged particle), the 8 lanes of a cell share that work
 * for the particle at hand instead (see the P2G loop). */
#define PART_CELLS (BLOCK / 8)
#define PART_STAGE 256
/* P2G staging chunk: small enough that a block's LDS (staged fields plus
 * the lanes' exchange area) stays under 32 KB, so five blocks share a CU */
#define PART_STAGE_P2G 240

/* Shape factors s[a][o] of a particle's base-cell corner nodes (axis a,
 * offset o), the values axis_shape gives each corner lane. */
__device__ __forceinline__ void particle_factors(const GridGeom &g, const double x[3], double s[3][2]) {
#pragma unroll
    for (int a = 0; a < 3; a++) {
        int b, lo, hi;
        base_cell(x[a], g.h[a], g.nn[a], b, lo, hi);
#pragma unroll
        for (int o = 0; o < 2; o++)
            s[a][o] = axis_shape(x[a], node_coord(b + o, g.L[a], g.cells[a]), g.h[a]).s;
    }
}

/* PAIR (the pair order): a block's groups take 31 occupied cells and the
 * next occupied cell as a halo, so that the +x neighbour of every cell a
 * group owns is in the group when it is occupied.  After the partial sums,
 * the lane of a cell's +x corner adds the partial of its +x neighbour's
 * matching -x corner (the same node) to its own, lower cell first, as the
 * pair order adds them: from the lane 4 above it in the same wave, or, for
 * the last cell of a wave, through LDS from the first cell of the next wave.
 * Its record (corner ox = 1) then holds the pair sum; a -x corner writes
 * its own record only when its -x neighbour cell is empty (else that
 * cell's record holds the pair).  The halo cell writes nothing (the next
 * group owns it). */
template <bool REMAP, bool MULTI, bool FEXT = false, bool PAIR = false>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(REMAP ? 6 : 1))) void k_cell_partials(Batch bt, PView P, NView Nd, BView B, CellParts C) {
    /* staged per particle: m v (3) and either the shape factors s (6, remap)
     * or x (3, P2G), then for P2G m, V, sigma (6) and, with FEXT, f_ext (3);
     * remap rows padded by one double so that the corner-dependent rows of
     * one particle fall in different banks */
    constexpr int PS = REMAP ? PART_STAGE : PART_STAGE_P2G, PR = REMAP ? PS + 1 : PS;
    constexpr bool SF = !REMAP && FEXT;
    __shared__ double sS[REMAP ? 6 : 3][PR], spv[3][PR];
    __shared__ double sm[REMAP ? 1 : PR], sV[REMAP ? 1 : PR], ss[REMAP ? 1 : 6][REMAP ? 1 : PR];
    __shared__ double sf[SF ? 3 : 1][SF ? PR : 1];
    __shared__ int8_t sdc[PS];
    /* P2G: what the 8 lanes of a cell share for the particle at hand: the
     * six axis shapes (factor, gradient-taken and sign bits) and the twelve
     * gradient quotients */
    __shared__ double xs_s[REMAP ? 1 : PART_CELLS][6], xs_q[REMAP ? 1 : PART_CELLS][12];
    __shared__ uint8_t xs_b[REMAP ? 1 : PART_CELLS][6];
    /* PAIR remap: the -x corner partials a wave's first cell passes to the
     * wave below (P2G passes them in its own wave's rows of xs_q) */
    __shared__ double xw[REMAP && PAIR ? BLOCK / 64 : 1][4][3];
    const uint32_t nocc = *B.nocc;
    const size_t n = P.n;
    const int lc = threadIdx.x >> 3, corner = threadIdx.x & 7;
    const int ox = corner >> 2, oy = (corner >> 1) & 1, oz = corner & 1;
    /* remap, one instance: its row read once, in uniform control flow and
     * ahead of every store (scalar registers); read per group and lane it is
     * a chain of vector loads that each wait on their own (0.295 -> 0.247 ms
     * on the 8.0 M-particle crush).  The P2G pass keeps the per-group reads:
     * with the row held, it needs more vector registers than five waves per
     * SIMD allow (0.440 -> 0.464 ms). */
    Inst I0;
    if (!MULTI && REMAP) I0 = bt.inst[0];
    auto inst_of = [&](uint32_t id) -> const Inst & { return (MULTI || !REMAP) ? inst_p<MULTI>(bt, id) : I0; };
    /* P2G sharing: the axis-shape pair of lanes 0..5, and the (type, index)
     * of the gradient quotients each lane computes (type 0 x, 1 y, 2 z; -1
     * none); every quotient is computed by a lane that holds its factors */
    const int pa = corner < 6 ? corner >> 1 : 0, po = corner & 1;
    /* two bits per corner, 3 = none: first {x z z y y y z -}, second
     * {y x x x z - - -} */
    const int qt1 = (int)((0xE568u >> (2 * corner)) & 3u), qt2 = (int)((0xFE01u >> (2 * corner)) & 3u);
    auto qindex = [&](int t) { return t == 0 ? 2 * oy + oz : (t == 1 ? 4 + 2 * ox + oz : 8 + 2 * ox + oy); };
    const int qk1 = qindex(qt1), qk2 = qindex(qt2);
    const bool skip_dc = REMAP ? bt.variant != VAR_DP : bt.variant == VAR_SP;
    constexpr bool fext = SF;
    /* cells a group owns (PAIR: one fewer, the last lane group is a halo) */
    constexpr uint32_t GS = PAIR ? PART_CELLS - 1 : PART_CELLS;
    for (uint32_t g0 = blockIdx.x * GS; g0 < nocc; g0 += gridDim.x * GS) {
        const uint32_t g1 = min(g0 + (uint32_t)PART_CELLS, nocc);
        const uint32_t r0 = B.ostart[g0], r1 = B.ostart[g1];
        /* this lane's (cell, corner) */
        const uint32_t slot = g0 + (uint32_t)lc;
        const bool cell = slot < g1;
        bool have = cell;
        /* PAIR: the cell's -x / +x neighbour cells are occupied (the +x one
         * then sits in the next lane group) */
        bool xprev = false, xnext = false;
        uint32_t s0 = 0, s1 = 0;
        double mn = 0.0, xpair = 0.0, hpair = 1.0, hx_ = 1.0, hy_ = 1.0, hz_ = 1.0, rmn = 0.0;
        bool massive = false;
        /* remap: every lane reads a valid rank, so the loads below are not
         * under a branch and issue together (a lane past the group's end
         * only has have = false; 0.267 -> 0.254 ms).  P2G keeps the branch:
         * the loads it adds take its registers past five waves per SIMD
         * (0.441 -> 0.464 ms). */
        const uint32_t sr = REMAP ? (cell ? slot : g0) : slot;
        if (REMAP || cell) {
            const uint32_t c = B.occ[sr];
            s0 = B.ostart[sr];
            s1 = B.ostart[sr + 1];
            const Inst &I = inst_of(MULTI ? B.sorted[s0] : 0u);
            const GridGeom &g = I.g;
            const uint32_t e = c - (uint32_t)I.cell_off;
            const int ex = (int)(e % (uint32_t)g.ec[0]);
            const int ey = (int)((e / (uint32_t)g.ec[0]) % (uint32_t)g.ec[1]);
            const int ez = (int)(e / ((uint32_t)g.ec[0] * (uint32_t)g.ec[1]));
            const int ix = ex - 1 + ox, iy = ey - 1 + oy, iz = ez - 1 + oz;
            have = cell & !((ix < 0) | (iy < 0) | (iz < 0) | (ix >= g.nn[0]) | (iy >= g.nn[1]) | (iz >= g.nn[2]));
            if (PAIR && REMAP) {
                const uint32_t cprev = B.occ[sr > 0 ? sr - 1 : sr], cnext = B.occ[sr + 1 < g1 ? sr + 1 : sr];
                xprev = cell & (ex > 0) & (sr > 0) & (cprev == c - 1u);
                xnext = cell & (ex < g.ec[0] - 1) & (sr + 1 < g1) & (cnext == c + 1u);
            } else if (PAIR) {
                xprev = ex > 0 && sr > 0 && B.occ[sr - 1] == c - 1u;
                xnext = ex < g.ec[0] - 1 && sr + 1 < g1 && B.occ[sr + 1] == c + 1u;
            }
            if (!REMAP) {
                /* the node coordinate of this lane's shape pair: axis pa,
                 * the cell's node pa-index e - 1 + po */
                const int ep = pa == 0 ? ex : (pa == 1 ? ey : ez);
                const double Lp = pa == 0 ? g.L[0] : (pa == 1 ? g.L[1] : g.L[2]);
                const int cp = pa == 0 ? g.cells[0] : (pa == 1 ? g.cells[1] : g.cells[2]);
                xpair = node_coord(ep - 1 + po, Lp, cp);
                hx_ = g.h[0];
                hy_ = g.h[1];
                hz_ = g.h[2];
                hpair = pa == 0 ? hx_ : (pa == 1 ? hy_ : hz_);
            }
            if (REMAP && have) {
                const int64_t node = I.node_off + (int64_t)ix + (int64_t)g.nn[0] * ((int64_t)iy + (int64_t)g.nn[1] * iz);
                mn = Nd.m[node];
                massive = mn > I.mass_min;
                /* the remap's per-pair division by m_i through div_rcp (its
                 * range: rmn = 0 marks a mass outside it) */
                rmn = (massive && mn >= 0x1p-100 && mn <= 0x1p100) ? 1.0 / mn : 0.0;
            }
            if (!REMAP && have && bt.tflag) bt.tflag[I.tile_off + tile_of_node(bt, g, ix, iy, iz)] = 1;
        }
        double m = 0.0, p[3] = {0.0, 0.0, 0.0}, f[3] = {0.0, 0.0, 0.0};
        for (uint32_t base = r0; base < r1; base += PS) {
            const uint32_t top = min(base + (uint32_t)PS, r1);
            __syncthreads();
            for (uint32_t q = base + threadIdx.x; q < top; q += BLOCK) {
                const uint32_t id = B.sorted[q], j = q - base;
                const double x[3] = {P.x[id], P.x[n + id], P.x[2 * n + id]};
                const double v[3] = {P.v[id], P.v[n + id], P.v[2 * n + id]};
                const double mp = P.m[id];
                const int dcr = skip_dc ? P.dc_rank[id] : -1;
                if (!REMAP) {
#pragma unroll
                    for (int a = 0; a < 3; a++) sS[a][j] = x[a];
                    sm[j] = mp;
                    sV[j] = P.V[id];
#pragma unroll
                    for (int k = 0; k < 6; k++) ss[k][j] = P.sig[k * n + id];
                    if (fext) {
                        sf[0][j] = P.fext[id];
                        sf[1][j] = P.fext[n + id];
                        sf[2][j] = P.fext[2 * n + id];
                    }
                } else {
                    double sh[3][2];
                    particle_factors(inst_of(id).g, x, sh);
#pragma unroll
                    for (int a = 0; a < 3; a++)
#pragma unroll
                        for (int o = 0; o < 2; o++) sS[2 * a + o][j] = sh[a][o];
                }
                /* flags: bit 0 left out of the scatter, bit 1 every component
                 * of m v within div_rcp's range for the remap (0 or within
                 * [2^-600, 2^600]: with N 0 or within [2^-159, 1], each
                 * numerator N m v is 0 or within [2^-759, 2^600]) */
                bool pv_ok = true;
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    spv[k][j] = mp * v[k];
                    const double am = fabs(spv[k][j]);
                    pv_ok &= (am >= 0x1p-600 && am <= 0x1p600) || am == 0.0;
                }
                sdc[j] = (int8_t)(((skip_dc && dcr >= 0) ? 1 : 0) | (pv_ok ? 2 : 0));
            }
            __syncthreads();
            if (REMAP ? !have : !cell) continue;
            const uint32_t a = max(s0, base), b = min(s1, top);
            if (!REMAP) {
                for (uint32_t q = a; q < b; q++) {
                    const uint32_t j = q - base;
                    /* only the single-pass loop leaves particles out here (the
                     * flag's read and wait skipped in the others) */
                    if (skip_dc && (sdc[j] & 1)) continue; /* SP:116-117, 157-158 */
                    /* lanes 0..5 of the cell: the axis shape of pair (axis
                     * corner >> 1, offset corner & 1), as axis_shape */
                    if (corner < 6) {
                        const AxisShape A = axis_shape(sS[pa][j], xpair, hpair);
                        xs_s[lc][corner] = A.s;
                        xs_b[lc][corner] = (uint8_t)((A.grad ? 1u : 0u) | (A.sgn > 0.0 ? 2u : 0u));
                    }
                    const double sx = xs_s[lc][0 + ox], sy = xs_s[lc][2 + oy], sz = xs_s[lc][4 + oz];
                    const uint32_t bx = xs_b[lc][0 + ox], by = xs_b[lc][2 + oy], bz = xs_b[lc][4 + oz];
                    /* the twelve quotients, at most two per lane, each from
                     * the lane's own factors: x (s_y s_z) / h_x, y (s_x s_z)
                     * / h_y, z (s_x s_y) / h_z */
                    {
                        const double pyz = sy * sz, pxz = sx * sz, pxy = sx * sy;
                        const double n1 = qt1 == 0 ? pyz : (qt1 == 1 ? pxz : pxy), d1 = qt1 == 0 ? hx_ : (qt1 == 1 ? hy_ : hz_);
                        const double n2 = qt2 == 0 ? pyz : (qt2 == 1 ? pxz : pxy), d2 = qt2 == 0 ? hx_ : (qt2 == 1 ? hy_ : hz_);
                        if (qt1 != 3) xs_q[lc][qk1] = n1 / d1;
                        if (qt2 != 3) xs_q[lc][qk2] = n2 / d2;
                    }
                    /* shape_full's gradient (-s_j s_l sgn) / h is exactly
                     * -sgn ((s_j s_l) / h): sgn = +-1 where it is taken */
                    const double qx = xs_q[lc][0 + 2 * oy + oz], qy = xs_q[lc][4 + 2 * ox + oz],
                                 qz = xs_q[lc][8 + 2 * ox + oy];
                    const double gx = (bx & 1u) ? ((bx & 2u) ? -qx : qx) : 0.0;
                    const double gy = (by & 1u) ? ((by & 2u) ? -qy : qy) : 0.0;
                    const double gz = (bz & 1u) ? ((bz & 2u) ? -qz : qz) : 0.0;
                    const double N = sx * sy * sz;
                    m += N * sm[j];
                    p[0] += N * spv[0][j];
                    p[1] += N * spv[1][j];
                    p[2] += N * spv[2][j];
                    const double V = sV[j];
                    f[0] += -V * (gx * ss[0][j] + gy * ss[3][j] + gz * ss[5][j]);
                    f[1] += -V * (gy * ss[1][j] + gx * ss[3][j] + gz * ss[4][j]);
                    f[2] += -V * (gz * ss[2][j] + gx * ss[5][j] + gy * ss[4][j]);
                    if (fext) {
                        f[0] += N * sf[0][j];
                        f[1] += N * sf[1][j];
                        f[2] += N * sf[2][j];
                    }
                }
            } else if (massive) {
                for (uint32_t q = a; q < b; q++) {
                    const uint32_t j = q - base;
                    const int8_t fl = sdc[j];
                    if (fl & 1) continue; /* DPC:358-359, SP:157-158 */
                    const double N = sS[0 + ox][j] * sS[2 + oy][j] * sS[4 + oz][j];
                    const double a0 = N * spv[0][j], a1 = N * spv[1][j], a2 = N * spv[2][j];
                    double t0, t1, t2;
                    /* N (m_p v_p) / m_i (DP:560): the same doubles as the IEEE
                     * division, which a wave takes when any of its pairs is
                     * outside div_rcp's range (0.252 -> 0.230 ms on the 8.0 M-
                     * particle crush; the divisions were 80 % of this loop's
                     * FP64 instructions) */
                    if (__any(!((fl & 2) && rmn != 0.0))) {
                        t0 = a0 / mn;
                        t1 = a1 / mn;
                        t2 = a2 / mn;
                    } else {
                        t0 = div_rcp(a0, mn, rmn);
                        t1 = div_rcp(a1, mn, rmn);
                        t2 = div_rcp(a2, mn, rmn);
                    }
                    f[0] += t0;
                    f[1] += t1;
                    f[2] += t2;
                }
            }
        }
        if (PAIR) {
            /* the partial of the +x neighbour's matching -x corner: lane + 4
             * (cell lc + 1, corner - 4) in the same wave; for a wave's last
             * cell, the next wave's first cell hands it over through LDS
             * (P2G: that wave's own rows of xs_q, which no other wave reads
             * in the particle loop; the wave has left its loop) */
            constexpr int NV = REMAP ? 3 : 7;
            double mine[NV], nb[NV];
            if (!REMAP) {
                mine[0] = m;
                mine[1] = p[0];
                mine[2] = p[1];
                mine[3] = p[2];
            }
            mine[NV - 3] = f[0];
            mine[NV - 2] = f[1];
            mine[NV - 1] = f[2];
#pragma unroll
            for (int v = 0; v < NV; v++) nb[v] = __shfl_down(mine[v], 4);
            double *xwv = REMAP ? &xw[0][0][0] : &xs_q[0][0];
            /* wave w's slot: P2G rows 8w .. 8w + 2 of xs_q (36 doubles); remap xw[w] */
            auto xslot = [&](int w, int c4) { return xwv + (REMAP ? (w * 4 + c4) * 3 : (w * 8 * 12 + c4 * 7)); };
            if ((lc & 7) == 0 && lc > 0 && ox == 0 && cell) {
                double *d = xslot(lc >> 3, corner & 3);
#pragma unroll
                for (int v = 0; v < NV; v++) d[v] = mine[v];
            }
            __syncthreads();
            if ((lc & 7) == 7 && lc < PART_CELLS - 1 && ox == 1 && xnext) {
                const double *d = xslot((lc >> 3) + 1, corner & 3);
#pragma unroll
                for (int v = 0; v < NV; v++) nb[v] = d[v];
            }
            if (lc == PART_CELLS - 1 || (ox == 0 && xprev)) continue;
            if (ox == 1 && xnext && (REMAP ? massive : true)) {
                if (!REMAP) {
                    m = m + nb[0];
                    p[0] = p[0] + nb[1];
                    p[1] = p[1] + nb[2];
                    p[2] = p[2] + nb[3];
                }
                f[0] = f[0] + nb[NV - 3];
                f[1] = f[1] + nb[NV - 2];
                f[2] = f[2] + nb[NV - 1];
            }
        }
        if (!have) continue;
        const size_t t = (size_t)corner * (size_t)C.cap + (size_t)slot;
        if (!REMAP) {
            /* one 64-byte record per (cell, corner): m, p, f */
            double2 *o = reinterpret_cast<double2 *>(C.w + t * 8u);
            o[0] = make_double2(m, p[0]);
            o[1] = make_double2(p[1], p[2]);
            o[2] = make_double2(f[0], f[1]);
            o[3] = make_double2(f[2], 0.0);
        } else {
            /* one 32-byte record per (cell, corner): v */
            double2 *o = reinterpret_cast<double2 *>(C.w + t * 4u);
            o[0] = make_double2(f[0], f[1]);
            o[1] = make_double2(f[2], 0.0);
        }
    }
}

/* Displacement-control winner among the controlled particles of one cell
 * (cells holding one are flagged in their slot by the binning). */
__device__ __forceinline__ void dc_scan_cell(DCWin &w, const PView &P, const BView &B, uint32_t c, bool all_axes) {
    for (uint32_t q = B.start[c]; q < B.start[c + 1]; q++) dc_consider(w, P, B.sorted[q], all_axes);
}

/* ---- particle order: per-particle terms, then ascending-id node sums ----
 *
 * Index of the stored term of the particle at sorted position q of a cell
 * whose ids occupy sorted positions s0 .. s0 + cnt - 1, at the cell's
 * corner node `corner` (k_list_terms writes, k_p2g_list / k_remap_list
 * read).  Cell-major: a cell's 8 cnt terms together, corner by corner, its
 * particles in ascending id within a corner.  A node then reads the
 * particles of one cell at one corner from consecutive terms, and the
 * threads of one cell's particles store one corner's terms consecutively.
 *
 * P2G term (4 doubles, 32 bytes): N and the internal force -V (sigma .
 * grad N), whose expression needs the particle's stress; N m and N m v
 * (and N f_ext) are formed by the node from N and the particle's row (4
 * doubles m, m v; 8 with f_ext: 3 more and a pad), one per sorted
 * position after the 8 lcap terms, so that a particle's mass and momentum
 * are stored once and not per corner.  The remap reuses the P2G terms' N
 * (the same shape values: the particles have not moved, and the binning
 * cell is the base cell of both) with a row of the particle's updated m v
 * in place of the P2G row; the node forms N (m v) / m_i itself. */
__device__ __forceinline__ size_t list_term(uint32_t s0, uint32_t cnt, uint32_t corner, uint32_t q) {
    return (size_t)s0 * 8u + (size_t)corner * (size_t)cnt + (size_t)(q - s0);
}
template <bool FEXT>
__device__ __forceinline__ double *list_row(const CellParts &C, uint32_t q) {
    return C.list + (size_t)C.lcap * 32u + (size_t)q * (FEXT ? 8u : 4u);
}

/* Particle side: one thread per binned particle (by sorted position) forms
 * its terms at its 8 corner nodes with the expressions of the grouped
 * orders' partial sums: P2G N and the internal force -V (sigma . grad N)
 * per corner and the particle's row (m, m v, f_ext), from which the node
 * forms m N, N m v and N f_ext, DP:48-89; the remap only the row of m v,
 * from which the node forms N (m v) / m_i with the P2G N, DP:539-564.
 * Terms go only to corner nodes inside the grid.  A particle the pass
 * leaves out of the scatter (SP:116-117, 157-158; DPC:358-359) stores
 * none: the node side skips it too (a particle the remap includes the
 * P2G pass includes as well). */
template <bool REMAP, bool MULTI, bool FEXT>
__global__ __launch_bounds__(BLOCK) void k_list_terms(Batch bt, PView P, BView B, CellParts C) {
    const size_t n = P.n;
    const uint32_t q = xcd_remap(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
    if (q >= B.ostart[*B.nocc]) return; /* past the binned particles */
    const uint32_t id = B.sorted[q], c = B.key[id];
    const uint32_t s0 = B.start[c], cnt = B.start[c + 1] - s0;
    const bool skip_dc = REMAP ? bt.variant != VAR_DP : bt.variant == VAR_SP;
    if (skip_dc && P.dc_rank[id] >= 0) return;
    const double mp = P.m[id];
    double pv[3];
#pragma unroll
    for (int a = 0; a < 3; a++) pv[a] = mp * P.v[a * n + id];
    if (REMAP) {
        double2 *t = reinterpret_cast<double2 *>(list_row<false>(C, q));
        t[0] = make_double2(pv[0], pv[1]);
        t[1] = make_double2(pv[2], 0.0);
        return;
    }
    const Inst &I = inst_p<MULTI>(bt, id);
    const GridGeom &g = I.g;
    const uint32_t e = c - (uint32_t)I.cell_off;
    const int ec[3] = {(int)(e % (uint32_t)g.ec[0]), (int)((e / (uint32_t)g.ec[0]) % (uint32_t)g.ec[1]),
                       (int)(e / ((uint32_t)g.ec[0] * (uint32_t)g.ec[1]))};
    const double x[3] = {P.x[id], P.x[n + id], P.x[2 * n + id]};
    /* the six axis shapes (GridPoint_Mediator.cpp:95-144) and the twelve
     * gradient quotients (s_j s_l) / h_k; shape_full's gradient
     * (-s_j s_l sgn) / h is exactly -sgn ((s_j s_l) / h) */
    double s[3][2];
    bool gr[3][2], neg[3][2];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int o = 0; o < 2; o++) {
            const AxisShape A = axis_shape(x[a], node_coord(ec[a] - 1 + o, g.L[a], g.cells[a]), g.h[a]);
            s[a][o] = A.s;
            gr[a][o] = A.grad;
            neg[a][o] = A.sgn > 0.0;
        }
    const double V = P.V[id];
    double sig[6], fx[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < 6; k++) sig[k] = P.sig[k * n + id];
    if (FEXT) {
#pragma unroll
        for (int a = 0; a < 3; a++) fx[a] = P.fext[a * n + id];
    }
#pragma unroll
    for (int corner = 0; corner < 8; corner++) {
        const int ox = corner >> 2, oy = (corner >> 1) & 1, oz = corner & 1;
        const int ix = ec[0] - 1 + ox, iy = ec[1] - 1 + oy, iz = ec[2] - 1 + oz;
        if ((ix < 0) | (iy < 0) | (iz < 0) | (ix >= g.nn[0]) | (iy >= g.nn[1]) | (iz >= g.nn[2])) continue;
        const double sx = s[0][ox], sy = s[1][oy], sz = s[2][oz];
        const double qx = (sy * sz) / g.h[0], qy = (sx * sz) / g.h[1], qz = (sx * sy) / g.h[2];
        const double gx = gr[0][ox] ? (neg[0][ox] ? -qx : qx) : 0.0;
        const double gy = gr[1][oy] ? (neg[1][oy] ? -qy : qy) : 0.0;
        const double gz = gr[2][oz] ? (neg[2][oz] ? -qz : qz) : 0.0;
        const double N = sx * sy * sz;
        double2 *r = reinterpret_cast<double2 *>(C.list + list_term(s0, cnt, (uint32_t)corner, q) * 4u);
        r[0] = make_double2(N, -V * (gx * sig[0] + gy * sig[3] + gz * sig[5]));
        r[1] = make_double2(-V * (gy * sig[1] + gx * sig[3] + gz * sig[4]), -V * (gz * sig[2] + gx * sig[5] + gy * sig[4]));
    }
    /* the particle's row: m, m v (and f_ext) */
    double2 *t = reinterpret_cast<double2 *>(list_row<FEXT>(C, q));
    t[0] = make_double2(mp, pv[0]);
    t[1] = make_double2(pv[1], pv[2]);
    if (FEXT) {
        t[FEXT ? 2 : 0] = make_double2(fx[0], fx[1]);
        t[FEXT ? 3 : 0] = make_double2(fx[2], 0.0);
    }
}

/* Node side: the node merges its 8 cells' sorted id lists and adds each
 * particle's stored terms at this node (its corner 7 - k in cell k) in
 * ascending id, the order of the reference's one-thread M2G loop
 * (PhysicsEngine_runSimulation_Classic_DoublePass_MPLocks.cpp:48-89). */
template <bool FEXT>
__global__ __launch_bounds__(NODE_BLOCK) void k_p2g_list(Batch bt, PView P, NView Nd, BView B, CellParts C) {
    const uint32_t tile = xcd_remap(blockIdx.x, gridDim.x);
    const Inst &I = bt.inst[inst_of_tile(bt, tile)];
    const GridGeom &g = I.g;
    int ix, iy, iz;
    node_of_tile(bt, g, tile - I.tile_off, ix, iy, iz);
    if (ix >= g.nn[0] || iy >= g.nn[1] || iz >= g.nn[2]) return;
    const int64_t node = I.node_off + (int64_t)ix + (int64_t)g.nn[0] * ((int64_t)iy + (int64_t)g.nn[1] * iz);
    const NodeIn in = node_in<false>(bt, I, Nd, node);
    Merge8 M;
    merge_init(M, B, g, I.cell_off, ix, iy, iz);
    uint32_t id = merge_min(M);
    if (id == MPM_NONE) {
        Nd.active[node] = 0;
        return;
    }
    double m = 0.0, p[3] = {0.0, 0.0, 0.0}, f[3] = {0.0, 0.0, 0.0};
    DCWin w;
    w.rank[0] = w.rank[1] = w.rank[2] = -1;
    w.v[0] = w.v[1] = w.v[2] = 0.0;
    const bool all_axes = bt.variant != VAR_DP, skip_dc = bt.variant == VAR_SP;
    const bool fext = FEXT && I.has_fext;
    /* one round of loads per particle: the list's next id and this
     * particle's terms issue together; only a cell holding a controlled
     * particle reads per-particle control data */
    /* without f_ext: the next particle of the same cell's term and row are
     * loaded one step ahead (the merge usually stays in a cell for its run
     * of ids), so a step waits on its id load only; ck / cpos name what is
     * held (1.72 -> 1.59 ms on the 8.0 M-particle crush, at four waves per
     * SIMD instead of five) */
    uint32_t ck = 8u, cpos = 0u;
    double2 ca0 = make_double2(0.0, 0.0), ca1 = ca0, cb0 = ca0, cb1 = ca0;
    while (id != MPM_NONE) {
        uint32_t k, pos, s0, cnt;
        merge_find(M, id, k, pos, s0, cnt);
        merge_step(M, B, k, pos);
        const double2 *r = reinterpret_cast<const double2 *>(C.list + list_term(s0, cnt, 7u - k, pos) * 4u);
        const double2 *pr = reinterpret_cast<const double2 *>(list_row<FEXT>(C, pos));
        double2 a0, a1, b0, b1;
        if (!FEXT && k == ck && pos == cpos) {
            a0 = ca0; a1 = ca1; b0 = cb0; b1 = cb1;
        } else {
            a0 = r[0]; a1 = r[1]; b0 = pr[0]; b1 = pr[1];
        }
        if (!FEXT) {
            if (pos + 1u < s0 + cnt) {
                ca0 = r[2]; ca1 = r[3]; cb0 = pr[2]; cb1 = pr[3];
                ck = k;
                cpos = pos + 1u;
            } else {
                ck = 8u;
            }
        }
        const double2 b2 = fext ? pr[FEXT ? 2 : 0] : make_double2(0.0, 0.0);
        const double2 b3 = fext ? pr[FEXT ? 3 : 0] : make_double2(0.0, 0.0);
        bool add = true;
        if ((M.dc >> k) & 1u) {
            add = !(skip_dc && P.dc_rank[id] >= 0); /* SP:116-117, 157-158 */
            dc_consider(w, P, id, all_axes);
        }
        if (add) {
            /* the terms k_cell_partials sums: m N and N m v from the
             * corner's N and the particle's row, the internal force as
             * stored, N f_ext */
            const double N = a0.x;
            m += N * b0.x;
            p[0] += N * b0.y;
            p[1] += N * b1.x;
            p[2] += N * b1.y;
            f[0] += a0.y;
            f[1] += a1.x;
            f[2] += a1.y;
            if (fext) {
                f[0] += N * b2.x;
                f[1] += N * b2.y;
                f[2] += N * b3.x;
            }
        }
        id = merge_min(M);
    }
    node_p2g_finish(bt, in, Nd, node, m, p, f, w);
}

/* The remap's node side in the particle order from the stored terms (the
 * ascending-id sums of the reference's one-thread loop). */
__global__ __launch_bounds__(NODE_BLOCK) void k_remap_list(Batch bt, PView P, NView Nd, BView B, CellParts C) {
    const uint32_t tile = xcd_remap(blockIdx.x, gridDim.x);
    const Inst &I = bt.inst[inst_of_tile(bt, tile)];
    const GridGeom &g = I.g;
    int ix, iy, iz;
    node_of_tile(bt, g, tile - I.tile_off, ix, iy, iz);
    if (ix >= g.nn[0] || iy >= g.nn[1] || iz >= g.nn[2]) return;
    const int64_t node = I.node_off + (int64_t)ix + (int64_t)g.nn[0] * ((int64_t)iy + (int64_t)g.nn[1] * iz);
    if (!Nd.active[node]) return;
    const NodeIn in = node_in<true>(bt, I, Nd, node);
    const bool massive = in.m > in.mass_min;
    Merge8 M;
    merge_init(M, B, g, I.cell_off, ix, iy, iz);
    double v[3] = {0.0, 0.0, 0.0};
    DCWin w;
    w.rank[0] = w.rank[1] = w.rank[2] = -1;
    w.v[0] = w.v[1] = w.v[2] = 0.0;
    const bool all_axes = bt.variant != VAR_DP, skip_dc = bt.variant != VAR_DP;
    for (uint32_t id = merge_min(M); id != MPM_NONE; id = merge_min(M)) {
        uint32_t k, pos, s0, cnt;
        merge_find(M, id, k, pos, s0, cnt);
        merge_step(M, B, k, pos);
        double t[3] = {0.0, 0.0, 0.0};
        if (massive) {
            const double N = C.list[list_term(s0, cnt, 7u - k, pos) * 4u];
            const double2 *pr = reinterpret_cast<const double2 *>(list_row<false>(C, pos));
            const double2 b0 = pr[0], b1 = pr[1];
            t[0] = N * b0.x / in.m;
            t[1] = N * b0.y / in.m;
            t[2] = N * b1.x / in.m;
        }
        bool add = massive;
        if ((M.dc >> k) & 1u) {
            add = add && !(skip_dc && P.dc_rank[id] >= 0); /* DPC:358-359, SP:157-158 */
            dc_consider(w, P, id, all_axes);
        }
        if (add) {
            v[0] += t[0];
            v[1] += t[1];
            v[2] += t[2];
        }
    }
    node_remap_dispatch(bt, in, Nd, node, v, w);
}

/* Node side of the cell order: adds the 8 partials in neighbour order
 * k = 0..7 (x outer, z inner), then the same grid update as the particle
 * order (P2G) or the same remap finish (REMAP).  PAIR (the pair order):
 * adds 4 records, j = 0..3: the pair sum the lower x cell k = j holds at
 * its +x corner, or, when that cell is empty, the upper cell k = j + 4's
 * own partial at its -x corner. */
template <bool REMAP, bool MULTI, bool PAIR = false>
__global__ __launch_bounds__(NODE_BLOCK) __attribute__((amdgpu_waves_per_eu(PAIR ? (REMAP ? 6 : 5) : 1))) void k_node_cells(Batch bt, PView P, NView Nd, BView B, CellParts C) {
    const uint32_t tile = xcd_remap(blockIdx.x, gridDim.x);
    /* tiles with no corner of an occupied cell: all their nodes are inactive
     * (the remap pass also re-arms the flag for the next step, after every
     * read of this kernel is issued: a store ahead of them would turn them
     * into vector loads that each wait on their own) */
    __shared__ uint8_t s_live;
    if (threadIdx.x == 0) s_live = bt.tflag[tile];
    __syncthreads();
    const bool live_tile = s_live != 0;
    const Inst &I = inst_t<MULTI>(bt, tile);
    const GridGeom &g = I.g;
    const uint32_t cell_off = (uint32_t)I.cell_off, ecx = (uint32_t)g.ec[0], ecy = (uint32_t)g.ec[1];
    int ix, iy, iz;
    node_of_tile(bt, g, tile - I.tile_off, ix, iy, iz);
    const bool in_grid = ix < g.nn[0] && iy < g.nn[1] && iz < g.nn[2];
    const int64_t node = I.node_off + (int64_t)ix + (int64_t)g.nn[0] * ((int64_t)iy + (int64_t)g.nn[1] * iz);
    const bool work = in_grid && live_tile;
    bool act = true;
    uint32_t sl[8];
    NodeIn in;
    if (work) {
        if (REMAP) act = Nd.active[node] != 0;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const uint32_t c = cell_off + (uint32_t)(ix + (k >> 2)) +
                               ecx * ((uint32_t)(iy + ((k >> 1) & 1)) + ecy * (uint32_t)(iz + (k & 1)));
            sl[k] = B.slot[c];
        }
        in = node_in<REMAP>(bt, I, Nd, node);
    }
    if (REMAP && threadIdx.x == 0) bt.tflag[tile] = 0;
    if (!in_grid) return;
    if (!live_tile) {
        if (!REMAP) Nd.active[node] = 0;
        return;
    }
    if (REMAP && !act) return;
    double m = 0.0, p[3] = {0.0, 0.0, 0.0}, f[3] = {0.0, 0.0, 0.0};
    DCWin w;
    w.rank[0] = w.rank[1] = w.rank[2] = -1;
    w.v[0] = w.v[1] = w.v[2] = 0.0;
    const bool all_axes = bt.variant != VAR_DP;
    bool any = false;
#pragma unroll
    for (int k = 0; k < 8; k++) any |= sl[k] != MPM_NONE;
    if (!REMAP && !any) {
        Nd.active[node] = 0;
        return;
    }
    /* all partial records first (an empty cell reads record 0 and adds
     * nothing), then the sums in neighbour order */
    if (PAIR) {
        constexpr int RW = REMAP ? 4 : 8; /* doubles per record */
        double2 ra[4], rb[4], rc[4], rd[4];
        bool o[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint32_t lo = sl[j], hi = sl[j + 4];
            o[j] = lo != MPM_NONE || hi != MPM_NONE;
            const size_t t = lo != MPM_NONE ? (size_t)(7 - j) * (size_t)C.cap + (size_t)(lo & ~SLOT_DC)
                                            : (hi != MPM_NONE ? (size_t)(3 - j) * (size_t)C.cap + (size_t)(hi & ~SLOT_DC)
                                                              : 0);
            const double2 *q = reinterpret_cast<const double2 *>(C.w + t * RW);
            ra[j] = q[0];
            rb[j] = q[1];
            if (!REMAP) {
                rc[j] = q[2];
                rd[j] = q[3];
            }
        }
#pragma unroll
        for (int j = 0; j < 4; j++) {
            if (!REMAP) {
                const double t0 = m + ra[j].x, t1 = p[0] + ra[j].y, t2 = p[1] + rb[j].x, t3 = p[2] + rb[j].y;
                const double t4 = f[0] + rc[j].x, t5 = f[1] + rc[j].y, t6 = f[2] + rd[j].x;
                m = o[j] ? t0 : m;
                p[0] = o[j] ? t1 : p[0];
                p[1] = o[j] ? t2 : p[1];
                p[2] = o[j] ? t3 : p[2];
                f[0] = o[j] ? t4 : f[0];
                f[1] = o[j] ? t5 : f[1];
                f[2] = o[j] ? t6 : f[2];
            } else {
                const double t0 = f[0] + ra[j].x, t1 = f[1] + ra[j].y, t2 = f[2] + rb[j].x;
                f[0] = o[j] ? t0 : f[0];
                f[1] = o[j] ? t1 : f[1];
                f[2] = o[j] ? t2 : f[2];
            }
        }
    } else if (!REMAP) {
        double2 ra[8], rb[8], rc[8], rd[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const size_t j = sl[k] == MPM_NONE ? 0 : (size_t)(7 - k) * (size_t)C.cap + (size_t)(sl[k] & ~SLOT_DC);
            const double2 *q = reinterpret_cast<const double2 *>(C.w + j * 8u);
            ra[k] = q[0];
            rb[k] = q[1];
            rc[k] = q[2];
            rd[k] = q[3];
        }
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const bool o = sl[k] != MPM_NONE;
            const double t0 = m + ra[k].x, t1 = p[0] + ra[k].y, t2 = p[1] + rb[k].x, t3 = p[2] + rb[k].y;
            const double t4 = f[0] + rc[k].x, t5 = f[1] + rc[k].y, t6 = f[2] + rd[k].x;
            m = o ? t0 : m;
            p[0] = o ? t1 : p[0];
            p[1] = o ? t2 : p[1];
            p[2] = o ? t3 : p[2];
            f[0] = o ? t4 : f[0];
            f[1] = o ? t5 : f[1];
            f[2] = o ? t6 : f[2];
        }
    } else {
        double2 ra[8], rb[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const size_t j = sl[k] == MPM_NONE ? 0 : (size_t)(7 - k) * (size_t)C.cap + (size_t)(sl[k] & ~SLOT_DC);
            const double2 *q = reinterpret_cast<const double2 *>(C.w + j * 4u);
            ra[k] = q[0];
            rb[k] = q[1];
        }
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const bool o = sl[k] != MPM_NONE;
            const double t0 = f[0] + ra[k].x, t1 = f[1] + ra[k].y, t2 = f[2] + rb[k].x;
            f[0] = o ? t0 : f[0];
            f[1] = o ? t1 : f[1];
            f[2] = o ? t2 : f[2];
        }
    }
#pragma unroll
    for (int k = 0; k < 8; k++)
        if (sl[k] != MPM_NONE && (sl[k] & SLOT_DC)) {
            const uint32_t c = cell_off + (uint32_t)(ix + (k >> 2)) +
                               ecx * ((uint32_t)(iy + ((k >> 1) & 1)) + ecy * (uint32_t)(iz + (k & 1)));
            dc_scan_cell(w, P, B, c, all_axes);
        }
    if (!REMAP) {
        node_p2g_finish(bt, in, Nd, node, m, p, f, w);
    } else {
        node_remap_dispatch(bt, in, Nd, node, f, w);
    }
}

/* ---- tile order: per summation tile, the node sums accumulated in LDS ----
 *
 * The binning cells are grouped in fixed summation tiles of ST_X x ST_Y x
 * ST_Z cells.  A node adds, over the tiles that hold its 8 adjacent cells in
 * ascending tile index, that tile's sum of those cells' partial sums (each
 * over the cell's particles in ascending id), the cells of one tile taken in
 * parity-class order, class = 4 (e_x & 1) + 2 (e_y & 1) + (e_z & 1).
 *
 * One block per tile.  Cells of one class never share a node (two cells
 * sharing a node differ by at most one along every axis, so some axis
 * parity differs), so the block runs the classes one after the other, each
 * with 8 lanes per cell (one per corner) as in the cell order, and every
 * lane adds its corner partial straight into its node's accumulator in LDS:
 * no conflicts, and each node receives its classes in class order.  A node
 * whose 8 cells all lie in the tile is then complete and is finished in the
 * block (displacement control, grid update, stores); a node on the face of
 * the tile's node box gets one record of the tile's sum, and the boundary
 * pass (k_tile_bnd) adds the records of its 2, 4 or 8 tiles in ascending tile
 * index and finishes it.  No partial sums leave the block except those of
 * face nodes, and no floating-point atomics are used. */

/* instance owning global summation tile t: last k with stile_off[k] <= t */
__device__ __forceinline__ int inst_of_stile(const Batch &bt, uint32_t t) {
    int lo = 0, hi = bt.ninst - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (bt.stile_off[mid] <= t)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ void stile_grid(const GridGeom &g, int &ntx, int &nty) {
    ntx = (g.ec[0] + ST_X - 1) / ST_X;
    nty = (g.ec[1] + ST_Y - 1) / ST_Y;
}

/* Index of a face node of a tile's node box (box coordinates lx < ST_BX,
 * ly < ST_BY, lz < ST_BZ) among the ST_BND face nodes: the two z faces,
 * then the two y faces without them, then the two x faces without both. */
__device__ __forceinline__ int box_face_index(int lx, int ly, int lz) {
    constexpr int ZF = ST_BX * ST_BY, YF = ST_BX * (ST_BZ - 2), XF = (ST_BY - 2) * (ST_BZ - 2);
    if (lz == 0) return lx + ST_BX * ly;
    if (lz == ST_BZ - 1) return ZF + lx + ST_BX * ly;
    if (ly == 0) return 2 * ZF + lx + ST_BX * (lz - 1);
    if (ly == ST_BY - 1) return 2 * ZF + YF + lx + ST_BX * (lz - 1);
    if (lx == 0) return 2 * ZF + 2 * YF + (ly - 1) + (ST_BY - 2) * (lz - 1);
    return 2 * ZF + 2 * YF + XF + (ly - 1) + (ST_BY - 2) * (lz - 1);
}

/* Particles staged per chunk, and the tile's particle ids kept on chip
 * (tiles holding more take their ids from memory class by class). */
#define ST_STAGE 256
#define ST_IDS 2048

/* What one staging thread carries from memory to LDS for one particle. */
template <bool REMAP, bool FEXT>
struct TileStage {
    double x[3], v[3], m, V, sig[REMAP ? 1 : 6], fx[FEXT ? 3 : 1];
    int dcr;
};

template <bool REMAP, bool MULTI, bool FEXT>
__global__ __launch_bounds__(BLOCK) void k_tile_p2g(Batch bt, PView P, NView Nd, BView B) {
    static_assert(ST_CELLS == BLOCK, "one thread per tile cell");
    constexpr int NC = REMAP ? 3 : 7;               /* accumulated components */
    constexpr int PS = ST_STAGE, PR = REMAP ? PS + 1 : PS;
    constexpr bool SF = !REMAP && FEXT;
    constexpr int CPC = ST_CELLS / 8;                /* cells per class */
    static_assert(CPC * 8 == BLOCK, "8 lanes for every cell of a class");
    static_assert(PS == BLOCK, "one staged particle per thread and chunk");
    __shared__ double acc[NC][ST_BOX];
    __shared__ uint32_t s_slot[ST_CELLS], s_first[ST_CELLS], s_cnt[ST_CELLS];
    __shared__ uint8_t s_cls[8][CPC], s_rank[ST_CELLS];
    __shared__ uint32_t s_ncls[8], s_pre[8][CPC + 1], s_base[9], s_cbase[9];
    __shared__ uint8_t s_gcell[ST_CELLS], s_gcls[ST_CELLS];
    __shared__ uint32_t s_gpre[ST_CELLS + 1];
    __shared__ uint32_t s_ids[ST_IDS];
    __shared__ double sS[REMAP ? 6 : 3][PR], spv[3][PR];
    __shared__ double sm[REMAP ? 1 : PR], sV[REMAP ? 1 : PR], ss[REMAP ? 1 : 6][REMAP ? 1 : PR];
    __shared__ double sf[SF ? 3 : 1][SF ? PR : 1];
    __shared__ int8_t sdc[PS];
    __shared__ double xs_s[REMAP ? 1 : CPC][6], xs_q[REMAP ? 1 : CPC][12];
    __shared__ uint8_t xs_b[REMAP ? 1 : CPC][6];

    const uint32_t tile = xcd_remap(blockIdx.x, gridDim.x);
    const Inst &I = MULTI ? bt.inst[inst_of_stile(bt, tile)] : bt.inst[0];
    const GridGeom &g = I.g;
    int ntx, nty;
    stile_grid(g, ntx, nty);
    const uint32_t lt = tile - I.stile_off;
    const int ta = (int)(lt % (uint32_t)ntx), tb = (int)((lt / (uint32_t)ntx) % (uint32_t)nty),
              tc = (int)(lt / ((uint32_t)ntx * (uint32_t)nty));
    const int ex0 = ST_X * ta, ey0 = ST_Y * tb, ez0 = ST_Z * tc;
    const int t = (int)threadIdx.x;
    const size_t n = P.n;
    auto cell_index = [&](int lcell) {
        const int ex = ex0 + lcell % ST_X, ey = ey0 + (lcell / ST_X) % ST_Y, ez = ez0 + lcell / (ST_X * ST_Y);
        return (size_t)I.cell_off + (size_t)ex + (size_t)g.ec[0] * ((size_t)ey + (size_t)g.ec[1] * ez);
    };

    /* the tile's cells, one per thread: slot, first sorted position, count */
    {
        const int cx = t % ST_X, cy = (t / ST_X) % ST_Y, cz = t / (ST_X * ST_Y);
        uint32_t sl = MPM_NONE, first = 0, cnt = 0;
        if (ex0 + cx < g.ec[0] && ey0 + cy < g.ec[1] && ez0 + cz < g.ec[2]) {
            const size_t c = cell_index(t);
            sl = B.slot[c];
            first = B.start[c];
            cnt = B.start[c + 1] - first;
        }
        s_slot[t] = sl;
        s_first[t] = first;
        s_cnt[t] = sl != MPM_NONE ? cnt : 0u;
        for (int j = t; j < NC * ST_BOX; j += BLOCK) (&acc[0][0])[j] = 0.0;
    }
    __syncthreads();
    /* class lists: wave w builds classes 2w and 2w + 1, one half-wave each;
     * cell q of a class is (2 i + px, 2 j + py, 2 k + pz), q = i + 4 j + 16 k,
     * kept in ascending q (any order would do: a class's cells share no node) */
    {
        const int lane = t & 63, half = lane >> 5, q = lane & 31;
        const int cls = 2 * (t >> 6) + half;
        const int px = cls >> 2, py = (cls >> 1) & 1, pz = cls & 1;
        const int cx = 2 * (q & 3) + px, cy = 2 * ((q >> 2) & 3) + py, cz = 2 * (q >> 4) + pz;
        const int lcell = cx + ST_X * (cy + ST_Y * cz);
        const bool occ = s_slot[lcell] != MPM_NONE;
        const uint32_t cnt = s_cnt[lcell];
        const uint64_t mask = __ballot(occ);
        const uint32_t hm = (uint32_t)(half ? (mask >> 32) : (mask & 0xFFFFFFFFull));
        const uint32_t below = hm & ((1u << q) - 1u);
        const int rank = __popc(below);
        /* inclusive scan of the counts over the half-wave, in q order */
        uint32_t x = cnt;
#pragma unroll
        for (int o = 1; o < 32; o <<= 1) {
            const uint32_t y = __shfl_up(x, o, 32);
            if (q >= o) x += y;
        }
        if (occ) {
            s_cls[cls][rank] = (uint8_t)lcell;
            s_rank[lcell] = (uint8_t)rank;
            s_pre[cls][rank + 1] = x; /* unoccupied cells add 0, so x is the prefix up to this cell */
        }
        if (q == 0) {
            s_pre[cls][0] = 0;
            s_ncls[cls] = (uint32_t)__popc(hm);
        }
    }
    __syncthreads();
    if (t == 0) {
        uint32_t b = 0;
        for (int c = 0; c < 8; c++) {
            s_base[c] = b;
            b += s_pre[c][s_ncls[c]];
        }
        s_base[8] = b;
    }
    __syncthreads();
    const uint32_t total_all = s_base[8];
    if (!REMAP && t == 0) bt.sflag[tile] = total_all ? 1 : 0;
    if (total_all == 0) {
        /* uniform; the nodes whose cells all lie in this tile were touched
         * by no particle (the face nodes are left to k_tile_bnd) */
        if (!REMAP)
            for (int j = t; j < ST_BOX; j += BLOCK) {
                const int lx = j % ST_BX, ly = (j / ST_BX) % ST_BY, lz = j / (ST_BX * ST_BY);
                const int ix = ex0 - 1 + lx, iy = ey0 - 1 + ly, iz = ez0 - 1 + lz;
                const bool inner = lx >= 1 && lx <= ST_X - 1 && ly >= 1 && ly <= ST_Y - 1 && lz >= 1 && lz <= ST_Z - 1;
                if (inner && ix < g.nn[0] && iy < g.nn[1] && iz < g.nn[2])
                    Nd.active[I.node_off + (int64_t)ix + (int64_t)g.nn[0] * ((int64_t)iy + (int64_t)g.nn[1] * iz)] = 0;
            }
        return;
    }
    /* the tile's occupied cells as one list, class by class (cells of a
     * class in list order): cell g, its first particle in the list, its
     * class; and the tile's particle ids in the same order */
    if (t == 0) {
        uint32_t b = 0;
        for (int c = 0; c < 8; c++) {
            s_cbase[c] = b;
            b += s_ncls[c];
        }
        s_cbase[8] = b;
        s_gpre[b] = total_all;
    }
    __syncthreads();
    const int ncell = (int)s_cbase[8];
    const bool ids_on_chip = total_all <= ST_IDS;
    if (s_slot[t] != MPM_NONE) {
        const int cx = t % ST_X, cy = (t / ST_X) % ST_Y, cz = t / (ST_X * ST_Y);
        const int cls = 4 * (cx & 1) + 2 * (cy & 1) + (cz & 1);
        const int rank = s_rank[t];
        const int gi = (int)s_cbase[cls] + rank;
        const uint32_t dst = s_base[cls] + s_pre[cls][rank];
        s_gcell[gi] = (uint8_t)t;
        s_gcls[gi] = (uint8_t)cls;
        s_gpre[gi] = dst;
        if (ids_on_chip) {
            const uint32_t first = s_first[t], cnt = s_cnt[t];
            for (uint32_t i0 = 0; i0 < cnt; i0 += 8) {
                uint32_t r[8];
#pragma unroll
                for (int i = 0; i < 8; i++) r[i] = i0 + i < cnt ? B.sorted[first + i0 + i] : 0u;
#pragma unroll
                for (int i = 0; i < 8; i++)
                    if (i0 + i < cnt) s_ids[dst + i0 + i] = r[i];
            }
        }
    }
    __syncthreads();

    const int lc = t >> 3, corner = t & 7;
    const int ox = corner >> 2, oy = (corner >> 1) & 1, oz = corner & 1;
    const int pa = corner < 6 ? corner >> 1 : 0, po = corner & 1;
    const int qt1 = (int)((0xE568u >> (2 * corner)) & 3u), qt2 = (int)((0xFE01u >> (2 * corner)) & 3u);
    auto qindex = [&](int qq) { return qq == 0 ? 2 * oy + oz : (qq == 1 ? 4 + 2 * ox + oz : 8 + 2 * ox + oy); };
    const int qk1 = qindex(qt1), qk2 = qindex(qt2);
    const bool skip_dc = REMAP ? bt.variant != VAR_DP : bt.variant == VAR_SP;
    const double hx_ = g.h[0], hy_ = g.h[1], hz_ = g.h[2];
    const double hpair = pa == 0 ? hx_ : (pa == 1 ? hy_ : hz_);

    /* the id at list position q */
    auto id_of = [&](uint32_t q) -> uint32_t {
        if (ids_on_chip) return s_ids[q];
        int lo = 0, hi = ncell - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (s_gpre[mid] <= q)
                lo = mid;
            else
                hi = mid - 1;
        }
        return B.sorted[s_first[s_gcell[lo]] + (q - s_gpre[lo])];
    };
    /* one particle's fields, loaded into registers ahead of their chunk */
    TileStage<REMAP, SF> pf;
    auto load = [&](uint32_t q) {
        const uint32_t id = id_of(q);
        pf.x[0] = P.x[id];
        pf.x[1] = P.x[n + id];
        pf.x[2] = P.x[2 * n + id];
        pf.v[0] = P.v[id];
        pf.v[1] = P.v[n + id];
        pf.v[2] = P.v[2 * n + id];
        pf.m = P.m[id];
        if (!REMAP) {
            pf.V = P.V[id];
#pragma unroll
            for (int k = 0; k < (REMAP ? 1 : 6); k++) pf.sig[k] = P.sig[k * n + id];
        }
        if (SF) {
            pf.fx[0] = P.fext[id];
            pf.fx[1] = P.fext[n + id];
            pf.fx[2] = P.fext[2 * n + id];
        }
        pf.dcr = skip_dc ? P.dc_rank[id] : -1;
    };
    /* Rounds of CPC consecutive cells of the list (8 lanes per cell, as in
     * the cell order), their particles staged in chunks of PS; the chunk
     * after (round r0, first particle base) */
    auto round_end = [&](int r0) { return s_gpre[min(r0 + CPC, ncell)]; };
    auto next = [&](int &r0, uint32_t &base) {
        base += PS;
        if (base >= round_end(r0)) {
            r0 += CPC;
            base = r0 < ncell ? s_gpre[r0] : 0u;
        }
    };
    int cr0 = 0;
    uint32_t cbase = 0;
    if ((uint32_t)t < round_end(0)) load((uint32_t)t);
    double m = 0.0, p[3] = {0.0, 0.0, 0.0}, f[3] = {0.0, 0.0, 0.0};
    while (cr0 < ncell) {
        const int r0 = cr0;
        const uint32_t base = cbase;
        const int g1 = min(r0 + CPC, ncell);
        const uint32_t p0 = s_gpre[r0], top_round = s_gpre[g1];
        const uint32_t top = min(base + (uint32_t)PS, top_round);
        __syncthreads(); /* staging free */
        if (base + (uint32_t)t < top) {
            const uint32_t j = (uint32_t)t;
            if (!REMAP) {
#pragma unroll
                for (int a = 0; a < 3; a++) sS[a][j] = pf.x[a];
                sm[j] = pf.m;
                sV[j] = pf.V;
#pragma unroll
                for (int k = 0; k < (REMAP ? 1 : 6); k++) ss[k][j] = pf.sig[k];
                if (SF) {
                    sf[0][j] = pf.fx[0];
                    sf[1][j] = pf.fx[1];
                    sf[2][j] = pf.fx[2];
                }
            } else {
                double sh[3][2];
                particle_factors(g, pf.x, sh);
#pragma unroll
                for (int a = 0; a < 3; a++)
#pragma unroll
                    for (int o = 0; o < 2; o++) sS[2 * a + o][j] = sh[a][o];
            }
#pragma unroll
            for (int k = 0; k < 3; k++) spv[k][j] = pf.m * pf.v[k];
            sdc[j] = (skip_dc && pf.dcr >= 0) ? 1 : 0;
        }
        __syncthreads();
        /* the next chunk's loads, in flight while this one is summed */
        next(cr0, cbase);
        if (cr0 < ncell && cbase + (uint32_t)t < round_end(cr0)) load(cbase + (uint32_t)t);
        /* this lane's (cell, corner) in this round */
        if (base == p0) {
            m = 0.0;
            p[0] = p[1] = p[2] = f[0] = f[1] = f[2] = 0.0;
        }
        const int gi = r0 + lc;
        const bool cell = gi < g1;
        bool have = cell, massive = false;
        int box = 0, mycls = -1;
        double mn = 0.0, xpair = 0.0;
        uint32_t v0 = 0, v1 = 0;
        if (cell) {
            const int lcell = s_gcell[gi];
            mycls = s_gcls[gi];
            const int cx = lcell % ST_X, cy = (lcell / ST_X) % ST_Y, cz = lcell / (ST_X * ST_Y);
            const int ix = ex0 + cx - 1 + ox, iy = ey0 + cy - 1 + oy, iz = ez0 + cz - 1 + oz;
            box = (cx + ox) + ST_BX * ((cy + oy) + ST_BY * (cz + oz));
            v0 = s_gpre[gi];
            v1 = s_gpre[gi + 1];
            if (!REMAP) {
                const int ep = pa == 0 ? ex0 + cx : (pa == 1 ? ey0 + cy : ez0 + cz);
                const double Lp = pa == 0 ? g.L[0] : (pa == 1 ? g.L[1] : g.L[2]);
                const int cp = pa == 0 ? g.cells[0] : (pa == 1 ? g.cells[1] : g.cells[2]);
                xpair = node_coord(ep - 1 + po, Lp, cp);
            }
            if (ix < 0 || iy < 0 || iz < 0 || ix >= g.nn[0] || iy >= g.nn[1] || iz >= g.nn[2]) have = false;
            if (REMAP && have) {
                mn = Nd.m[I.node_off + (int64_t)ix + (int64_t)g.nn[0] * ((int64_t)iy + (int64_t)g.nn[1] * iz)];
                massive = mn > I.mass_min;
            }
        }
        if (REMAP ? (have && massive) : cell) {
            const uint32_t a = max(v0, base), b = min(v1, top);
            if (!REMAP) {
                for (uint32_t q = a; q < b; q++) {
                    const uint32_t j = q - base;
                    if (sdc[j]) continue; /* SP:116-117, 157-158 */
                    if (corner < 6) {
                        const AxisShape A = axis_shape(sS[pa][j], xpair, hpair);
                        xs_s[lc][corner] = A.s;
                        xs_b[lc][corner] = (uint8_t)((A.grad ? 1u : 0u) | (A.sgn > 0.0 ? 2u : 0u));
                    }
                    const double sx = xs_s[lc][0 + ox], sy = xs_s[lc][2 + oy], sz = xs_s[lc][4 + oz];
                    const uint32_t bx = xs_b[lc][0 + ox], by = xs_b[lc][2 + oy], bz = xs_b[lc][4 + oz];
                    {
                        const double pyz = sy * sz, pxz = sx * sz, pxy = sx * sy;
                        const double n1 = qt1 == 0 ? pyz : (qt1 == 1 ? pxz : pxy), d1 = qt1 == 0 ? hx_ : (qt1 == 1 ? hy_ : hz_);
                        const double n2 = qt2 == 0 ? pyz : (qt2 == 1 ? pxz : pxy), d2 = qt2 == 0 ? hx_ : (qt2 == 1 ? hy_ : hz_);
                        if (qt1 != 3) xs_q[lc][qk1] = n1 / d1;
                        if (qt2 != 3) xs_q[lc][qk2] = n2 / d2;
                    }
                    const double qx = xs_q[lc][0 + 2 * oy + oz], qy = xs_q[lc][4 + 2 * ox + oz],
                                 qz = xs_q[lc][8 + 2 * ox + oy];
                    const double gx = (bx & 1u) ? ((bx & 2u) ? -qx : qx) : 0.0;
                    const double gy = (by & 1u) ? ((by & 2u) ? -qy : qy) : 0.0;
                    const double gz = (bz & 1u) ? ((bz & 2u) ? -qz : qz) : 0.0;
                    const double N = sx * sy * sz;
                    m += N * sm[j];
                    p[0] += N * spv[0][j];
                    p[1] += N * spv[1][j];
                    p[2] += N * spv[2][j];
                    const double V = sV[j];
                    f[0] += -V * (gx * ss[0][j] + gy * ss[3][j] + gz * ss[5][j]);
                    f[1] += -V * (gy * ss[1][j] + gx * ss[3][j] + gz * ss[4][j]);
                    f[2] += -V * (gz * ss[2][j] + gx * ss[5][j] + gy * ss[4][j]);
                    if (SF) {
                        f[0] += N * sf[0][j];
                        f[1] += N * sf[1][j];
                        f[2] += N * sf[2][j];
                    }
                }
            } else {
                for (uint32_t q = a; q < b; q++) {
                    const uint32_t j = q - base;
                    if (sdc[j]) continue; /* DPC:358-359, SP:157-158 */
                    const double N = sS[0 + ox][j] * sS[2 + oy][j] * sS[4 + oz][j];
                    f[0] += N * spv[0][j] / mn;
                    f[1] += N * spv[1][j] / mn;
                    f[2] += N * spv[2][j] / mn;
                }
            }
        }
        /* the round's last chunk: the corner partials into their nodes'
         * accumulators, class by class (cells of one class share no node;
         * a class a round holds only in part adds the rest in the next
         * round, before any later class) */
        if (top == top_round) {
            const int c0 = s_gcls[r0], c1 = s_gcls[g1 - 1];
            const bool add = REMAP ? (have && massive) : have;
            for (int c = c0; c <= c1; c++) {
                if (c > c0) __syncthreads();
                if (add && mycls == c) {
                    if (!REMAP) {
                        acc[0][box] += m;
                        acc[1][box] += p[0];
                        acc[2][box] += p[1];
                        acc[3][box] += p[2];
                        acc[4][box] += f[0];
                        acc[5][box] += f[1];
                        acc[6][box] += f[2];
                    } else {
                        acc[0][box] += f[0];
                        acc[1][box] += f[1];
                        acc[2][box] += f[2];
                    }
                }
            }
        }
    }
    __syncthreads();

    /* finish the nodes whose 8 cells lie in the tile; record the face nodes */
    const bool all_axes = bt.variant != VAR_DP;
    for (int j = t; j < ST_BOX; j += BLOCK) {
        const int lx = j % ST_BX, ly = (j / ST_BX) % ST_BY, lz = j / (ST_BX * ST_BY);
        const int ix = ex0 - 1 + lx, iy = ey0 - 1 + ly, iz = ez0 - 1 + lz;
        if (ix < 0 || iy < 0 || iz < 0 || ix >= g.nn[0] || iy >= g.nn[1] || iz >= g.nn[2]) continue;
        const int64_t nd = I.node_off + (int64_t)ix + (int64_t)g.nn[0] * ((int64_t)iy + (int64_t)g.nn[1] * iz);
        const bool inner = lx >= 1 && lx <= ST_X - 1 && ly >= 1 && ly <= ST_Y - 1 && lz >= 1 && lz <= ST_Z - 1;
        /* the node's cells that lie in this tile: local (lx - 1 + dx, ...) */
        bool any = false, dc = false;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const int cx = lx - 1 + (k >> 2), cy = ly - 1 + ((k >> 1) & 1), cz = lz - 1 + (k & 1);
            if (cx < 0 || cy < 0 || cz < 0 || cx >= ST_X || cy >= ST_Y || cz >= ST_Z) continue;
            const uint32_t sl = s_slot[cx + ST_X * (cy + ST_Y * cz)];
            if (sl != MPM_NONE) {
                any = true;
                dc |= (sl & SLOT_DC) != 0;
            }
        }
        if (!inner) {
            double *r = bt.srec + ((size_t)tile * ST_BND + (size_t)box_face_index(lx, ly, lz)) * ST_REC;
            double2 *o = reinterpret_cast<double2 *>(r);
            if (!REMAP) {
                o[0] = make_double2(acc[0][j], acc[1][j]);
                o[1] = make_double2(acc[2][j], acc[3][j]);
                o[2] = make_double2(acc[4][j], acc[5][j]);
                o[3] = make_double2(acc[6][j], any ? 1.0 : 0.0);
            } else {
                o[0] = make_double2(acc[0][j], acc[1][j]);
                o[1] = make_double2(acc[2][j], 0.0);
            }
            continue;
        }
        if (!REMAP) {
            if (!any) {
                Nd.active[nd] = 0;
                continue;
            }
        } else if (!Nd.active[nd]) {
            continue;
        }
        const NodeIn in = node_in<REMAP>(bt, I, Nd, nd);
        DCWin w;
        w.rank[0] = w.rank[1] = w.rank[2] = -1;
        w.v[0] = w.v[1] = w.v[2] = 0.0;
        if (dc) {
#pragma unroll 1
            for (int k = 0; k < 8; k++) {
                const int cx = lx - 1 + (k >> 2), cy = ly - 1 + ((k >> 1) & 1), cz = lz - 1 + (k & 1);
                const uint32_t sl = s_slot[cx + ST_X * (cy + ST_Y * cz)];
                if (sl != MPM_NONE && (sl & SLOT_DC)) {
                    const size_t c = (size_t)I.cell_off + (size_t)(ex0 + cx) +
                                     (size_t)g.ec[0] * ((size_t)(ey0 + cy) + (size_t)g.ec[1] * (ez0 + cz));
                    dc_scan_cell(w, P, B, (uint32_t)c, all_axes);
                }
            }
        }
        if (!REMAP) {
            double mm = acc[0][j], pp[3] = {acc[1][j], acc[2][j], acc[3][j]}, ff[3] = {acc[4][j], acc[5][j], acc[6][j]};
            node_p2g_finish(bt, in, Nd, nd, mm, pp, ff, w);
        } else {
            double vv[3] = {acc[0][j], acc[1][j], acc[2][j]};
            node_remap_dispatch(bt, in, Nd, nd, vv, w);
        }
    }
}

/* Face nodes of the tile order: each tile finishes the face nodes it owns
 * (the nodes ix in [ST_X a, ST_X a + ST_X) etc. on its upper faces, the
 * other nodes of that range being complete in k_tile_p2g): the records of
 * the tiles holding the node's cells, in ascending tile index, added from
 * 0, then the same node finish as the cell order. */
template <bool REMAP, bool MULTI>
__global__ __launch_bounds__(BLOCK) void k_tile_bnd(Batch bt, PView P, NView Nd, BView B) {
    static_assert(ST_CELLS == BLOCK, "one thread per owned node");
    const uint32_t tile = xcd_remap(blockIdx.x, gridDim.x);
    const Inst &I = MULTI ? bt.inst[inst_of_stile(bt, tile)] : bt.inst[0];
    const GridGeom &g = I.g;
    int ntx, nty;
    stile_grid(g, ntx, nty);
    const uint32_t lt = tile - I.stile_off;
    const int ta = (int)(lt % (uint32_t)ntx), tb = (int)((lt / (uint32_t)ntx) % (uint32_t)nty),
              tc = (int)(lt / ((uint32_t)ntx * (uint32_t)nty));
    const int t = (int)threadIdx.x;
    const int ox = t % ST_X, oy = (t / ST_X) % ST_Y, oz = t / (ST_X * ST_Y);
    const bool fx = ox == ST_X - 1, fy = oy == ST_Y - 1, fz = oz == ST_Z - 1;
    if (!(fx || fy || fz)) return;
    const int ix = ST_X * ta + ox, iy = ST_Y * tb + oy, iz = ST_Z * tc + oz;
    if (ix >= g.nn[0] || iy >= g.nn[1] || iz >= g.nn[2]) return;
    const int64_t nd = I.node_off + (int64_t)ix + (int64_t)g.nn[0] * ((int64_t)iy + (int64_t)g.nn[1] * iz);
    /* source tiles in ascending index: d = dx + 2 dy + 4 dz */
    double s[REMAP ? 3 : 7];
#pragma unroll
    for (int k = 0; k < (REMAP ? 3 : 7); k++) s[k] = 0.0;
    bool any = false;
#pragma unroll
    for (int d = 0; d < 8; d++) {
        const int dx = d & 1, dy = (d >> 1) & 1, dz = d >> 2;
        if ((dx && !fx) || (dy && !fy) || (dz && !fz)) continue;
        const uint32_t src = tile + (uint32_t)dx + (uint32_t)ntx * ((uint32_t)dy + (uint32_t)nty * (uint32_t)dz);
        if (!bt.sflag[src]) continue;
        const int lx = ix - (ST_X * (ta + dx) - 1), ly = iy - (ST_Y * (tb + dy) - 1), lz = iz - (ST_Z * (tc + dz) - 1);
        const double2 *r = reinterpret_cast<const double2 *>(
            bt.srec + ((size_t)src * ST_BND + (size_t)box_face_index(lx, ly, lz)) * ST_REC);
        if (!REMAP) {
            const double2 r0 = r[0], r1 = r[1], r2 = r[2], r3 = r[3];
            s[0] += r0.x;
            s[1] += r0.y;
            s[2] += r1.x;
            s[3] += r1.y;
            s[4] += r2.x;
            s[5] += r2.y;
            s[6] += r3.x;
            any |= r3.y != 0.0;
        } else {
            const double2 r0 = r[0], r1 = r[1];
            s[0] += r0.x;
            s[1] += r0.y;
            s[2] += r1.x;
        }
    }
    if (!REMAP) {
        if (!any) {
            Nd.active[nd] = 0;
            return;
        }
    } else if (!Nd.active[nd]) {
        return;
    }
    const NodeIn in = node_in<REMAP>(bt, I, Nd, nd);
    DCWin w;
    w.rank[0] = w.rank[1] = w.rank[2] = -1;
    w.v[0] = w.v[1] = w.v[2] = 0.0;
    const bool all_axes = bt.variant != VAR_DP;
    const uint32_t cell_off = (uint32_t)I.cell_off, ecx = (uint32_t)g.ec[0], ecy = (uint32_t)g.ec[1];
#pragma unroll 1
    for (int k = 0; k < 8; k++) {
        const uint32_t c = cell_off + (uint32_t)(ix + (k >> 2)) +
                           ecx * ((uint32_t)(iy + ((k >> 1) & 1)) + ecy * (uint32_t)(iz + (k & 1)));
        const uint32_t sl = B.slot[c];
        if (sl != MPM_NONE && (sl & SLOT_DC)) dc_scan_cell(w, P, B, c, all_axes);
    }
    if (!REMAP) {
        double pp[3] = {s[1], s[2], s[3]}, ff[3] = {s[4], s[5], s[6]};
        node_p2g_finish(bt, in, Nd, nd, s[0], pp, ff, w);
    } else {
        double vv[3] = {s[0], s[1], s[2]};
        node_remap_dispatch(bt, in, Nd, nd, vv, w);
    }
}

__device__ __forceinline__ void block_count_add(unsigned long long *health, int slot, uint32_t flag) {
    /* wave ballot, then one atomic per wave, into this block's copy */
    uint64_t b = __ballot(flag != 0);
    if ((threadIdx.x & 63) == 0 && b)
        atomicAdd(&health[(blockIdx.x % HEALTH_STRIPES) * HEALTH_STRIDE + slot], (unsigned long long)__popcll(b));
}

/* The accumulated fields the state update reads and writes back.  They
 * are read with the rest of the particle's state, before the kernel's
 * first store: a read-modify-write in place would make each load wait on
 * the store before it (the compiler cannot prove the arrays apart).
 *
 * The plastic strain, the plastic work and the hardening variable change
 * only when the point yields: otherwise dεp = +0, dR = +0 and the plastic
 * work increment is ±0, and adding them gives back the stored doubles (a
 * stored -0.0 would become +0.0 in the reference; these fields start at
 * +0.0 and a sum is -0.0 only when both terms are, so that needs a caller
 * to store -0.0 there).  Those three are read and written only for points
 * that yield or whose plastic work increment is not zero (NaN included),
 * which is most of the state update's traffic on an elastic step. */
struct PAccum {
    double eps[6], Ws, Wti, Wmax;
};

/* DP reads all of them; the contact and single-pass updates (new stress
 * energies, strain updated before the material check) need no history
 * maximum or time integral. */
__device__ __forceinline__ void load_accum(const PView &P, size_t i, bool dp, PAccum &a) {
    const size_t n = P.n;
#pragma unroll
    for (int k = 0; k < 6; k++) a.eps[k] = __builtin_nontemporal_load(&P.eps[k * n + i]);
    a.Ws = __builtin_nontemporal_load(&P.Ws[i]);
    a.Wti = dp ? __builtin_nontemporal_load(&P.Wti[i]) : 0.0;
    a.Wmax = dp ? __builtin_nontemporal_load(&P.Wmax[i]) : 0.0;
}

/* The plastic fields, when they change: reads εp and Wp, writes εp + dεp,
 * Wp + wp and R + dR. */
__device__ __forceinline__ void update_plastic(const PView &P, size_t i, const CRResult &cr, double R, double wp) {
    const size_t n = P.n;
    const bool change = cr.yielded || !(wp == 0.0) || !(cr.dR == 0.0);
    if (!change) return;
    double epsp[6];
#pragma unroll
    for (int k = 0; k < 6; k++) epsp[k] = __builtin_nontemporal_load(&P.epsp[k * n + i]);
    const double Wp = __builtin_nontemporal_load(&P.Wp[i]);
#pragma unroll
    for (int k = 0; k < 6; k++) __builtin_nontemporal_store(epsp[k] + cr.depsp[k], &P.epsp[k * n + i]);
    __builtin_nontemporal_store(Wp + wp, &P.Wp[i]);
    __builtin_nontemporal_store(R + cr.dR, &P.R[i]);
}

/* State update after the return map (DP:242-278 minus F and V, which the
 * caller has done): eps, sigma, plastic strain, energies, hardening. */
__device__ __forceinline__ uint32_t finish_update(const PView &P, size_t i, const double de[6], double sig[6],
                                                  const CRResult &cr, double R, double V, double dt,
                                                  const PAccum &a) {
    const size_t n = P.n;
#pragma unroll
    for (int k = 0; k < 6; k++) sig[k] += cr.dsig[k];
    double wi = 0.0;
#pragma unroll
    for (int k = 0; k < 6; k++) wi += (sig[k] - 0.5 * cr.dsig[k]) * de[k] * V;
    const double Ws = a.Ws + wi;
    double wp = 0.0;
#pragma unroll
    for (int k = 0; k < 6; k++) wp += (sig[k] - 0.5 * cr.dsig[k]) * cr.depsp[k] * V;
    update_plastic(P, i, cr, R, wp);
#pragma unroll
    for (int k = 0; k < 6; k++) __builtin_nontemporal_store(a.eps[k] + de[k], &P.eps[k * n + i]);
#pragma unroll
    for (int k = 0; k < 6; k++) __builtin_nontemporal_store(sig[k], &P.sig[k * n + i]);
    __builtin_nontemporal_store(Ws, &P.Ws[i]);
    __builtin_nontemporal_store(a.Wti + Ws * dt, &P.Wti[i]);
    if (Ws > a.Wmax) __builtin_nontemporal_store(Ws, &P.Wmax[i]);
    uint32_t bad = 0;
    for (int k = 0; k < 6; k++) bad |= !isfinite(sig[k]);
    return bad;
}


/* Stress, plastic strain, energies and backstress of the contact and
 * single-pass loops (DPC:514-524, SP:351-362): energies accumulate
 * sigma_new[k] * de[k] * V one component at a time; no history maximum or
 * time integral.  The strain was already updated. */
__device__ __forceinline__ uint32_t finish_update_new_stress(const PView &P, size_t i, double sig[6],
                                                             const CRResult &cr, double R, double V,
                                                             const double de[6], const PAccum &a) {
    const size_t n = P.n;
#pragma unroll
    for (int k = 0; k < 6; k++) sig[k] += cr.dsig[k];
    double Ws = a.Ws;
#pragma unroll
    for (int k = 0; k < 6; k++) Ws += sig[k] * de[k] * V;
    /* the plastic work increment, added to Wp one component at a time as
     * the reference does (Wp + t0 + t1 ...), is formed here as a sum of the
     * terms only to decide whether any is non-zero; update_plastic is given
     * the exact sequential result below */
    double t[6];
    bool any = false;
#pragma unroll
    for (int k = 0; k < 6; k++) {
        t[k] = sig[k] * cr.depsp[k] * V;
        any |= !(t[k] == 0.0);
    }
    if (cr.yielded || any || !(cr.dR == 0.0)) {
        double epsp[6];
#pragma unroll
        for (int k = 0; k < 6; k++) epsp[k] = P.epsp[k * n + i];
        double Wp = P.Wp[i];
#pragma unroll
        for (int k = 0; k < 6; k++) Wp += t[k];
#pragma unroll
        for (int k = 0; k < 6; k++) P.epsp[k * n + i] = epsp[k] + cr.depsp[k];
        P.Wp[i] = Wp;
        P.R[i] = R + cr.dR;
    }
#pragma unroll
    for (int k = 0; k < 6; k++) P.sig[k * n + i] = sig[k];
    P.Ws[i] = Ws;
    uint32_t bad = 0;
    for (int k = 0; k < 6; k++) bad |= !isfinite(sig[k]);
    return bad;
}

/* Trial state of the radial return: J2 of sigma + C:de and the yield radius
 * (the first half of ConstitutiveRelation_calculateIncrement_{PerfectlyPlastic,
 * VonMisesHardening}.cpp, identical expressions). */
__device__ __forceinline__ void trial_state(const MatTable &mt, int m, double R, const double sig[6],
                                            const double de[6], double dtr[6], double &J2, double &K) {
    double str[6];
    trial_increment(mt.E[m], mt.nu[m], de, dtr);
    for (int k = 0; k < 6; k++) str[k] = sig[k] + dtr[k];
    J2 = 0.0;
    J2 += 1.0 / 6.0 * (str[0] - str[1]) * (str[0] - str[1]);
    J2 += 1.0 / 6.0 * (str[1] - str[2]) * (str[1] - str[2]);
    J2 += 1.0 / 6.0 * (str[2] - str[0]) * (str[2] - str[0]);
    J2 += str[3] * str[3];
    J2 += str[4] * str[4];
    J2 += str[5] * str[5];
    K = mt.yield[m] / sqrt(3.0);
    if (mt.type[m] == 4) K = K + R / sqrt(2.0);
}

/* The trial deviator s_tr = dev(sigma + C : de) and the network's inputs:
 * the direction s_tr / sqrt(J2) (6), sqrt(J2) / K, the overstress
 * (sqrt(J2) - K) / sigma_y, K / sigma_y, the trial pressure / sigma_y,
 * R / sigma_y, |de| / (sigma_y / E) and sqrt(J2) / sigma_y.  All are formed
 * in FP64 from the trial state, so the quantities the radial return depends
 * on reach the network at FP32 resolution before standardisation, instead of
 * as small differences of raw stresses. */
__device__ __forceinline__ void surrogate_features(const MatTable &mt, int m, double R, const double sig[6],
                                                   const double de[6], double stress_scale, double strain_scale,
                                                   double dtr[6], double dev[6], double &K, double feat[MLP_IN]) {
    double J2;
    trial_state(mt, m, R, sig, de, dtr, J2, K);
    double str[6];
    for (int k = 0; k < 6; k++) str[k] = sig[k] + dtr[k];
    const double p = (str[0] + str[1] + str[2]) / 3.0;
    for (int k = 0; k < 6; k++) dev[k] = str[k] - (k < 3 ? p : 0.0);
    const double q = sqrt(J2);
    for (int k = 0; k < 6; k++) feat[k] = dev[k] / q;
    feat[6] = q / K;
    feat[7] = (q - K) / stress_scale;
    feat[8] = K / stress_scale;
    feat[9] = p / stress_scale;
    feat[10] = R / stress_scale;
    double e2 = 0.0;
    for (int k = 0; k < 6; k++) e2 += de[k] * de[k];
    feat[11] = sqrt(e2) / strain_scale;
    feat[12] = q / stress_scale;
}

/* runSimulation_G2P_P2_SmallStrain (DP:175-280), in two kernels so that
 * each keeps enough loads in flight: the kinematics (node gather, x, L, the
 * strain increment, F and V, the displacement-control schedule of
 * DP:766-789 and the next step's binning key) and then the return map and
 * state update, a pure per-particle stream.  The strain increment passes
 * between them in P.de. */
/* The kinematics of one particle: x, L, the strain increment, F, V, the
 * control schedule and the particle's next cell key.  rec(corner, node,
 * r0..r3) fetches the packed g2 record of one adjacent node.  Every
 * instance field and every particle field the pass needs is read before
 * its first store: reads after a store cannot be proven unaliased, become
 * vector loads and each waits on its own. */
/* What the kinematics hands to the state update. */
struct KinOut {
    double de[6]; /* strain increment */
    double V;     /* volume after the pass */
    double L[3][3], F[3][3], V0; /* large-strain pass: velocity gradient, F before the step, V0 */
    double dt;
    int mat;
    uint32_t flag; /* position not finite */
    bool ctl;      /* left out of the pass (controlled, contact and single-pass loops) */
};

/* HANDOFF: store the strain increment and the flag for k_g2p_constitutive
 * (the fused pass keeps them in registers).  LARGE: the large-strain pass,
 * which takes L, F and V0 in ko and does the F and V update itself. */
template <int VAR, bool HANDOFF, bool LARGE = false, class REC>
__device__ __forceinline__ void g2p_kinematics_particle(const Batch &bt, const PView &P, const Inst &I, uint32_t i,
                                                        REC rec, uint32_t &next_key, uint32_t &f_outside,
                                                        KinOut &ko) {
    const size_t n = P.n;
    const GridGeom g = I.g;
    const double dt = I.dt, mass_min = I.mass_min;
    const size_t node_off = (size_t)I.node_off;
    const uint32_t cell_off = (uint32_t)I.cell_off;
    const int32_t tl_count = I.tl_count;
    const double vt[3] = {I.vt[0], I.vt[1], I.vt[2]};
    const double x0[3] = {__builtin_nontemporal_load(&P.x[i]), __builtin_nontemporal_load(&P.x[n + i]), __builtin_nontemporal_load(&P.x[2 * n + i])};
    const int32_t dcr = P.dc_rank[i];
    const int mat = P.mat[i];
    double F[3][3];
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
        for (int r = 0; r < 3; r++) F[c][r] = __builtin_nontemporal_load(&P.F[(3 * c + r) * n + i]);
    const double V0 = __builtin_nontemporal_load(&P.V0[i]);
    const double hx[3] = {g.h[0], g.h[1], g.h[2]};
    const int gn[3] = {g.nn[0], g.nn[1], g.nn[2]};
    double x[3] = {x0[0], x0[1], x0[2]};
    int b[3], lo[3], hi[3];
    AxisShape as[3][2];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        base_cell(x0[a], hx[a], gn[a], b[a], lo[a], hi[a]);
#pragma unroll
        for (int o = 0; o < 2; o++)
            as[a][o] = axis_shape(x0[a], node_coord(b[a] + o, g.L[a], g.cells[a]), hx[a]);
    }
    double L[3][3]; /* glm column-major L[col][row] */
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
        for (int r = 0; r < 3; r++) L[c][r] = 0.0;
    bool ok[3][2];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int o = 0; o < 2; o++)
            ok[a][o] = (o >= lo[a]) & (o < hi[a]) & (b[a] + o >= 0) & (b[a] + o <= gn[a] - 1);
    f_outside = !((ok[0][0] || ok[0][1]) && (ok[1][0] || ok[1][1]) && (ok[2][0] || ok[2][1]));
    /* the contact and single-pass loops leave controlled particles out
     * of the grid-to-particle pass (DPC:443-444, SP:279-280) */
    const bool ctl = VAR != VAR_DP && dcr >= 0;
    uint8_t flag = 0;
    ko.ctl = ctl;
    ko.mat = mat;
    ko.dt = dt;
    ko.V = 0.0;
    if (!ctl) {
        const size_t nx = (size_t)gn[0], ny = (size_t)gn[1];
        /* corners in GridPoint_Mediator order: x outer, z inner;
         * branch-free over the 8 corners (off-grid corners read the
         * instance's first node and add nothing), so that all gathers can
         * be in flight */
#pragma unroll
        for (int cc = 0; cc < 8; cc++) {
            const int ox = cc >> 2, oy = (cc >> 1) & 1, oz = cc & 1;
            const bool in = ok[0][ox] && ok[1][oy] && ok[2][oz];
            const size_t node =
                in ? node_off + (size_t)(b[0] + ox) + nx * ((size_t)(b[1] + oy) + ny * (size_t)(b[2] + oz))
                   : node_off;
            double N, gr[3];
            shape_full(as[0][ox], as[1][oy], as[2][oz], hx, N, gr[0], gr[1], gr[2]);
            double2 r0, r1, r2, r3;
            rec(cc, node, r0, r1, r2, r3);
            const bool take = in && r0.x > mass_min;
            const double t0 = x[0] + N * r0.y * dt, t1 = x[1] + N * r1.x * dt, t2 = x[2] + N * r1.y * dt;
            x[0] = take ? t0 : x[0];
            x[1] = take ? t1 : x[1];
            x[2] = take ? t2 : x[2];
            const double vv[3] = {r2.x, r2.y, r3.x};
#pragma unroll
            for (int c = 0; c < 3; c++)
#pragma unroll
                for (int r = 0; r < 3; r++) {
                    const double t = L[c][r] + vv[r] * gr[c];
                    L[c][r] = in ? t : L[c][r];
                }
        }
        __builtin_nontemporal_store(x[0], &P.x[i]);
        __builtin_nontemporal_store(x[1], &P.x[n + i]);
        __builtin_nontemporal_store(x[2], &P.x[2 * n + i]);
        if (LARGE) {
#pragma unroll
            for (int c = 0; c < 3; c++)
#pragma unroll
                for (int r = 0; r < 3; r++) {
                    ko.L[c][r] = L[c][r];
                    ko.F[c][r] = F[c][r];
                }
            ko.V0 = V0;
        } else {
        const bool valid = mat >= 0 && mat < bt.nmat;
        double dF[3][3];
#pragma unroll
        for (int c = 0; c < 3; c++)
#pragma unroll
            for (int r = 0; r < 3; r++) dF[c][r] = (c == r ? 1.0 : 0.0) + L[c][r] * dt;
        double de[6];
        if (VAR == VAR_DP) {
            double E[3][3];
#pragma unroll
            for (int c = 0; c < 3; c++)
#pragma unroll
                for (int r = 0; r < 3; r++) E[c][r] = 0.5 * (dF[r][c] + dF[c][r]) - (c == r ? 1.0 : 0.0);
            de[0] = E[0][0];
            de[1] = E[1][1];
            de[2] = E[2][2];
            de[3] = E[0][1] + E[1][0];
            de[4] = E[1][2] + E[2][1];
            de[5] = E[2][0] + E[0][2];
        } else {
            /* DPC:469-477, SP:310-318: straight from dF */
            de[0] = dF[0][0] - 1.0;
            de[1] = dF[1][1] - 1.0;
            de[2] = dF[2][2] - 1.0;
            de[3] = dF[0][1] + dF[1][0];
            de[4] = dF[1][2] + dF[2][1];
            de[5] = dF[2][0] + dF[0][2];
        }
#pragma unroll
        for (int k = 0; k < 6; k++) {
            ko.de[k] = de[k];
            if (HANDOFF) P.de[k * n + i] = de[k];
        }
        /* DP: F = F dF after the material check (DP:242-245); DPC / SP:
         * F += (L F) dt before it (DPC:464-467, SP:305-308) */
        if (VAR != VAR_DP || valid) {
            double Fn[3][3];
            if (VAR == VAR_DP) {
#pragma unroll
                for (int c = 0; c < 3; c++)
#pragma unroll
                    for (int r = 0; r < 3; r++)
                        Fn[c][r] = F[0][r] * dF[c][0] + F[1][r] * dF[c][1] + F[2][r] * dF[c][2];
            } else {
#pragma unroll
                for (int c = 0; c < 3; c++)
#pragma unroll
                    for (int r = 0; r < 3; r++)
                        Fn[c][r] = F[c][r] + (L[0][r] * F[c][0] + L[1][r] * F[c][1] + L[2][r] * F[c][2]) * dt;
            }
#pragma unroll
            for (int c = 0; c < 3; c++)
#pragma unroll
                for (int r = 0; r < 3; r++) __builtin_nontemporal_store(Fn[c][r], &P.F[(3 * c + r) * n + i]);
            double det = +Fn[0][0] * (Fn[1][1] * Fn[2][2] - Fn[2][1] * Fn[1][2]) -
                         Fn[1][0] * (Fn[0][1] * Fn[2][2] - Fn[2][1] * Fn[0][2]) +
                         Fn[2][0] * (Fn[0][1] * Fn[1][2] - Fn[1][1] * Fn[0][2]);
            ko.V = det * V0;
            __builtin_nontemporal_store(ko.V, &P.V[i]);
        }
        }
        flag = !(isfinite(x[0]) && isfinite(x[1]) && isfinite(x[2])) ? 1 : 0;
    }
    ko.flag = flag;
    if (HANDOFF) P.flag[i] = flag;
    /* displacement-control schedule: DP:766-789 (flagged axes), or
     * DPC:531-540 / SP:369-378 (all axes, then x += v dt) */
    if (VAR == VAR_DP) {
        if (tl_count != 0 && dcr >= 0) {
            uint8_t axm = P.dc_axes[i];
            double mult = (double)P.dc_mult[i];
#pragma unroll
            for (int a = 0; a < 3; a++)
                if ((axm >> a) & 1) P.v[a * n + i] = mult * vt[a];
        }
    } else if (ctl) {
        double v[3] = {P.v[i], P.v[n + i], P.v[2 * n + i]};
        if (tl_count != 0) {
            double mult = (double)P.dc_mult[i];
#pragma unroll
            for (int a = 0; a < 3; a++) {
                v[a] = mult * vt[a];
                P.v[a * n + i] = v[a];
            }
        }
#pragma unroll
        for (int a = 0; a < 3; a++) {
            x[a] = x0[a] + v[a] * dt;
            P.x[a * n + i] = x[a];
        }
    }
    next_key = cell_key(g, x[0], x[1], x[2]);
    if (next_key != MPM_NONE) next_key += cell_off;
}

template <int VAR, bool MULTI>
__global__ __launch_bounds__(BLOCK) void k_g2p_kinematics(Batch bt, PView P, NView Nd, BView B,
                                                         unsigned long long *health) {
    uint32_t i = xcd_remap(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
    const bool live = i < P.n;
    uint32_t f_outside = 0;
    uint32_t next_key = MPM_NONE;
    if (live) {
        const double *g2 = Nd.g2;
        KinOut ko;
        g2p_kinematics_particle<VAR, true>(bt, P, inst_p<MULTI>(bt, i), i,
                                     [=](int, size_t node, double2 &r0, double2 &r1, double2 &r2, double2 &r3) {
                                         const double2 *q = reinterpret_cast<const double2 *>(g2 + node * 8u);
                                         r0 = q[0];
                                         r1 = q[1];
                                         r2 = q[2];
                                         r3 = q[3];
                                     },
                                     next_key, f_outside, ko);
        B.key[i] = next_key;
    }
    wave_key_add(B.count, next_key, next_key != MPM_NONE);
    block_count_add(health, HEALTH_OUTSIDE, f_outside);
}

/* The return map and state update of the same pass (DP:204-278, DPC:479-524,
 * SP:320-362) on the strain increment left by k_g2p_kinematics.
 *   MODE_EXACT      the exact FP64 return map for every particle;
 *   MODE_CAPTURE    as exact, and every yielding point's (sigma, de, R,
 *                   material) is appended to the sample buffer;
 *   MODE_SURROGATE  points of the surrogate's material whose trial state
 *                   yields are queued for the MLP; everything else is
 *                   finished here. */
template <int MODE, int VAR>
__global__ __launch_bounds__(BLOCK) void k_g2p_constitutive(Batch bt, PView P, MatTable mt,
                                                           unsigned long long *health, SurrogateArgs sg) {
    uint32_t i = xcd_remap(blockIdx.x, gridDim.x) * BLOCK + threadIdx.x;
    uint32_t f_null = 0, f_plastic = 0, f_nonfinite = 0, f_queued = 0;
    const size_t n = P.n;
    const bool live = i < P.n && !(VAR != VAR_DP && P.dc_rank[i] >= 0);
    double de[6], sig[6], R = 0.0, V = 0.0, dt = 0.0;
    PAccum acc;
    int mat = -1;
    bool valid = false;
    if (live) {
        dt = bt.inst[inst_of_particle(bt, i)].dt;
        mat = P.mat[i];
        valid = mat >= 0 && mat < mt.n;
#pragma unroll
        for (int k = 0; k < 6; k++) de[k] = P.de[k * n + i];
        V = P.V[i];
#pragma unroll
        for (int k = 0; k < 6; k++) sig[k] = P.sig[k * n + i];
        R = P.R[i];
        f_nonfinite = P.flag[i] & 1u;
        load_accum(P, i, VAR == VAR_DP, acc);
        if (!valid) f_null = 1;
    }
    /* the learned map takes the points of its material whose trial state yields */
    bool queue = false;
    if (MODE == MODE_SURROGATE && live && valid && mat == sg.material) {
        double dtr[6], J2, K;
        trial_state(mt, mat, R, sig, de, dtr, J2, K);
        queue = !(J2 <= K * K);
    }
    if (MODE == MODE_SURROGATE) {
        const uint32_t slot = block_append<uint32_t>(sg.qcount, queue);
        if (queue) {
            sg.qid[slot] = i;
            const size_t cap = (size_t)P.n;
            double dtr[6], dev[6], K, feat[MLP_IN];
            surrogate_features(mt, mat, R, sig, de, sg.stress_scale, sg.strain_scale, dtr, dev, K, feat);
            uint8_t inside = 1;
            for (int k = 0; k < MLP_IN; k++) {
                float z = ((float)feat[k] - sg.norm.in_mean[k]) / sg.norm.in_std[k];
                inside &= (z >= sg.norm.in_lo[k] && z <= sg.norm.in_hi[k]) ? 1 : 0;
                sg.qx[(size_t)k * cap + slot] = z; /* feature-major: coalesced */
            }
            sg.qenv[slot] = inside;
            f_queued = 1;
            f_plastic = 1;
            /* stress and the rest are finished by the post kernel */
        }
    }
    bool capture = false;
    CRResult cr;
    if (live && !queue) {
        if (!valid && VAR != VAR_SP) {
            /* DP:226-230, DPC:491-495: message and skip */
            for (int k = 0; k < 6; k++) f_nonfinite |= !isfinite(sig[k]);
        } else {
            if (!valid) {
                /* SP with no material: the per-particle defaults (elastic,
                 * E = nu = 0, MaterialPoint_BC.h:48-57) */
                trial_increment(0.0, 0.0, de, cr.dsig);
                for (int k = 0; k < 6; k++) cr.depsp[k] = 0.0;
                cr.dR = 0.0;
                cr.yielded = false;
            } else {
                constitutive(mt, mat, R, sig, de, cr);
                f_plastic = cr.yielded;
                capture = MODE == MODE_CAPTURE && cr.yielded;
            }
        }
    }
    if (MODE == MODE_CAPTURE) {
        const unsigned long long slot = block_append<unsigned long long>(sg.capcount, capture);
        if (capture && slot < (unsigned long long)sg.capcap) {
            double *row = sg.cap + slot * 14;
            for (int k = 0; k < 6; k++) row[k] = sig[k];
            for (int k = 0; k < 6; k++) row[6 + k] = de[k];
            row[12] = R;
            row[13] = (double)mat;
        }
    }
    if (live && !queue && (valid || VAR == VAR_SP)) {
        if (VAR == VAR_DP)
            f_nonfinite |= finish_update(P, i, de, sig, cr, R, V, dt, acc);
        else
            f_nonfinite |= finish_update_new_stress(P, i, sig, cr, R, V, de, acc);
    }
    if (live && VAR != VAR_DP) {
        /* the strain is updated before the material check, for every point
         * the pass visits (stored last: no read follows it) */
#pragma unroll
        for (int k = 0; k < 6; k++) P.eps[k * n + i] = acc.eps[k] + de[k];
    }
    block_count_add(health, HEALTH_NULL, f_null);
    block_count_add(health, HEALTH_PLASTIC, f_plastic);
    block_count_add(health, HEALTH_NONFINITE, f_nonfinite);
    if (MODE == MODE_SURROGATE) block_count_add(health, HEALTH_SURROGATE_CALLS, f_queued);
}

/* The exact state update of one particle that the pass visits, from the
 * kinematics' outputs in registers: the same steps as k_g2p_constitutive's
 * exact mode (material check, return map, finish, strain). */
template <int VAR>
__device__ __forceinline__ void state_update_particle(const PView &P, const MatTable &mt, uint32_t i, const KinOut &ko,
                                                      uint32_t &f_null, uint32_t &f_plastic, uint32_t &f_nonfinite) {
    const size_t n = P.n;
    const int mat = ko.mat;
    const bool valid = mat >= 0 && mat < mt.n;
    double sig[6];
#pragma unroll
    for (int k = 0; k < 6; k++) sig[k] = __builtin_nontemporal_load(&P.sig[k * n + i]);
    const double R = __builtin_nontemporal_load(&P.R[i]);
    PAccum acc;
    load_accum(P, i, VAR == VAR_DP, acc);
    f_nonfinite = ko.flag;
    if (!valid) f_null = 1;
    if (!valid && VAR != VAR_SP) {
        /* DP:226-230, DPC:491-495: message and skip */
        for (int k = 0; k < 6; k++) f_nonfinite |= !isfinite(sig[k]);
    } else {
        CRResult cr;
        if (!valid) {
            /* SP with no material: the per-particle defaults (elastic,
             * E = nu = 0, MaterialPoint_BC.h:48-57) */
            trial_increment(0.0, 0.0, ko.de, cr.dsig);
            for (int k = 0; k < 6; k++) cr.depsp[k] = 0.0;
            cr.dR = 0.0;
            cr.yielded = false;
        } else {
            constitutive(mt, mat, R, sig, ko.de, cr);
            f_plastic = cr.yielded;
        }
        if (VAR == VAR_DP)
            f_nonfinite |= finish_update(P, i, ko.de, sig, cr, R, ko.V, ko.dt, acc);
        else
            f_nonfinite |= finish_update_new_stress(P, i, sig, cr, R, ko.V, ko.de, acc);
    }
    if (VAR != VAR_DP) {
#pragma unroll
        for (int k = 0; k < 6; k++) P.eps[k * n + i] = acc.eps[k] + ko.de[k];
    }
}

/* G2P pass 2 in one kernel for the exact return map: kinematics, then the
 * state update of the same particle, with the strain increment, the new
 * volume and the flags kept in registers (the two-kernel form hands them
 * over through memory; the surrogate and capture modes use it for their
 * queues). */
template <int VAR, bool MULTI>
__global__ __launch_bounds__(BLOCK) void k_g2p_fused(Batch bt, PView P, NView Nd, BView B, MatTable mt,
                                                    unsigned long long *health) {
    uint32_t i, k;
    bool live;
    if (!particle_of_thread<MULTI>(bt, (uint32_t)P.n, i, live, k)) return;
    uint32_t f_outside = 0, f_null = 0, f_plastic = 0, f_nonfinite = 0;
    uint32_t next_key = MPM_NONE;
    if (live) {
        const double *g2 = Nd.g2;
        KinOut ko;
        g2p_kinematics_particle<VAR, false>(bt, P, bt.inst[k], i,
                                            [=](int, size_t node, double2 &r0, double2 &r1, double2 &r2, double2 &r3) {
                                                const double2 *q = reinterpret_cast<const double2 *>(g2 + node * 8u);
                                                r0 = q[0];
                                                r1 = q[1];
                                                r2 = q[2];
                                                r3 = q[3];
                                            },
                                            next_key, f_outside, ko);
        B.key[i] = next_key;
        if (!ko.ctl) state_update_particle<VAR>(P, mt, i, ko, f_null, f_plastic, f_nonfinite);
    }
    wave_key_add(B.count, next_key, next_key != MPM_NONE);
    block_count_add(health, HEALTH_OUTSIDE, f_outside);
    block_count_add(health, HEALTH_NULL, f_null);
    block_count_add(health, HEALTH_PLASTIC, f_plastic);
    block_count_add(health, HEALTH_NONFINITE, f_nonfinite);
}

/* G2P pass 2 of the DoublePass loop with the learned map, in one kernel for
 * a single instance: the kinematics as in k_g2p_fused, then, for a point of
 * the surrogate's material whose trial state yields, its inputs go to the
 * network's queue (and its strain increment to P.de, which the post kernel
 * reads); every other point takes the exact state update in registers.
 * Only the queued points' increments cross memory, so with few yielding
 * points the pass moves what the exact fused pass moves. */
__global__ __launch_bounds__(BLOCK) void k_g2p_fused_queue(Batch bt, PView P, NView Nd, BView B, MatTable mt,
                                                          unsigned long long *health, SurrogateArgs sg) {
    uint32_t i, k;
    bool live;
    particle_of_thread<false>(bt, (uint32_t)P.n, i, live, k);
    uint32_t f_outside = 0, f_null = 0, f_plastic = 0, f_nonfinite = 0, f_queued = 0;
    uint32_t next_key = MPM_NONE;
    bool queue = false;
    KinOut ko;
    double sig[6], R = 0.0;
    if (live) {
        const double *g2 = Nd.g2;
        g2p_kinematics_particle<VAR_DP, false>(bt, P, bt.inst[0], i,
                                               [=](int, size_t node, double2 &r0, double2 &r1, double2 &r2,
                                                   double2 &r3) {
                                                   const double2 *q = reinterpret_cast<const double2 *>(g2 + node * 8u);
                                                   r0 = q[0];
                                                   r1 = q[1];
                                                   r2 = q[2];
                                                   r3 = q[3];
                                               },
                                               next_key, f_outside, ko);
        B.key[i] = next_key;
        if (!ko.ctl && ko.mat == sg.material && ko.mat >= 0 && ko.mat < mt.n) {
            const size_t n = P.n;
            for (int c = 0; c < 6; c++) sig[c] = P.sig[c * n + i];
            R = P.R[i];
            double dtr[6], J2, K;
            trial_state(mt, ko.mat, R, sig, ko.de, dtr, J2, K);
            queue = !(J2 <= K * K);
        }
    }
    const uint32_t slot = block_append<uint32_t>(sg.qcount, queue);
    if (queue) {
        const size_t n = P.n, cap = (size_t)P.n;
        sg.qid[slot] = i;
        for (int c = 0; c < 6; c++) P.de[c * n + i] = ko.de[c];
        double dtr[6], dev[6], K, feat[MLP_IN];
        surrogate_features(mt, ko.mat, R, sig, ko.de, sg.stress_scale, sg.strain_scale, dtr, dev, K, feat);
        uint8_t inside = 1;
        for (int c = 0; c < MLP_IN; c++) {
            float z = ((float)feat[c] - sg.norm.in_mean[c]) / sg.norm.in_std[c];
            inside &= (z >= sg.norm.in_lo[c] && z <= sg.norm.in_hi[c]) ? 1 : 0;
            sg.qx[(size_t)c * cap + slot] = z;
        }
        sg.qenv[slot] = inside;
        f_queued = 1;
        f_plastic = 1;
        f_nonfinite = ko.flag;
    } else if (live && !ko.ctl) {
        state_update_particle<VAR_DP>(P, mt, i, ko, f_null, f_plastic, f_nonfinite);
    }
    wave_key_add(B.count, next_key, next_key != MPM_NONE);
    block_count_add(health, HEALTH_OUTSIDE, f_outside);
    block_count_add(health, HEALTH_NULL, f_null);
    block_count_add(health, HEALTH_PLASTIC, f_plastic);
    block_count_add(health, HEALTH_NONFINITE, f_nonfinite);
    block_count_add(health, HEALTH_SURROGATE_CALLS, f_queued);
}

/* ---- large-strain pass ---------------------------------------------------- */

/* glm mat3 product a * b: Result[c][r] = a[0][r]*b[c][0] + a[1][r]*b[c][1] + a[2][r]*b[c][2]. */
__device__ __forceinline__ void m3_mul(const double a[3][3], const double b[3][3], double out[3][3]) {
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
        for (int r = 0; r < 3; r++) out[c][r] = a[0][r] * b[c][0] + a[1][r] * b[c][1] + a[2][r] * b[c][2];
}

/* 0.5 * getLog(A * transpose(A)) (DP:323, DP:342-343): getLog is the series
 * sum over t = 1..8 of ((B - I)^t * (1/t)) * (-1)^(t+1), the powers formed as
 * I * D * D ... (getPower), added term by term
 * (ConstitutiveRelation_MatrixAlgebra.cpp:12-23, 39-53).  P^t is P^(t-1) * D,
 * the same products getPower repeats. */
__device__ __forceinline__ void half_log_aat(const double A[3][3], double out[3][3]) {
    double D[3][3], P[3][3], res[3][3];
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
        for (int r = 0; r < 3; r++) {
            D[c][r] = (A[0][r] * A[0][c] + A[1][r] * A[1][c] + A[2][r] * A[2][c]) - (c == r ? 1.0 : 0.0);
            P[c][r] = c == r ? 1.0 : 0.0;
            res[c][r] = 0.0;
        }
#pragma unroll
    for (int t = 1; t <= 8; t++) {
        double Q[3][3];
        m3_mul(P, D, Q);
        const double s = 1.0 / (double)t;
        const double sg = (t & 1) ? 1.0 : -1.0;
#pragma unroll
        for (int c = 0; c < 3; c++)
#pragma unroll
            for (int r = 0; r < 3; r++) {
                P[c][r] = Q[c][r];
                res[c][r] += Q[c][r] * s * sg;
            }
    }
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
        for (int r = 0; r < 3; r++) out[c][r] = 0.5 * res[c][r];
}

/* getExp (ConstitutiveRelation_MatrixAlgebra.cpp:25-37): the sum over
 * t = 0..8 of A^t / t!, added term by term. */
__device__ __forceinline__ void exp_series(const double A[3][3], double out[3][3]) {
    double P[3][3];
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
        for (int r = 0; r < 3; r++) {
            P[c][r] = c == r ? 1.0 : 0.0;
            out[c][r] = 0.0 + P[c][r] / 1.0;
        }
    double f = 1.0;
#pragma unroll
    for (int t = 1; t <= 8; t++) {
        double Q[3][3];
        m3_mul(P, A, Q);
        f *= (double)t;
#pragma unroll
        for (int c = 0; c < 3; c++)
#pragma unroll
            for (int r = 0; r < 3; r++) {
                P[c][r] = Q[c][r];
                out[c][r] += Q[c][r] / f;
            }
    }
}

/* Voigt vector with engineering shears (DP:325-330, 348-361). */
__device__ __forceinline__ void m3_voigt(const double m[3][3], double v[6]) {
    v[0] = m[0][0];
    v[1] = m[1][1];
    v[2] = m[2][2];
    v[3] = m[0][1] + m[1][0];
    v[4] = m[1][2] + m[2][1];
    v[5] = m[2][0] + m[0][2];
}

/* runSimulation_G2P_P2_LargeStrain after the node gather (DP:303-435) for
 * one particle: stress reset, dF = I + L dt, the log strains of F (before
 * the step), of dF Fe and of dF; then, for a particle with a material, the
 * return map on the elastic trial strain from zero stress, Fe = exp(trial
 * minus plastic strain, tensor shears), F = F dF, V = det F V0, strain =
 * log strain of the old F, stress = returned stress / det F, plastic strain,
 * energies (sigma[k] * increment[k] * V added one component at a time) and
 * hardening.  A particle with no material keeps everything but the zeroed
 * stress (DP:373-377). */
__device__ __forceinline__ void large_strain_update(const PView &P, const MatTable &mt, uint32_t i, const KinOut &ko,
                                                    uint32_t &f_null, uint32_t &f_plastic, uint32_t &f_nonfinite) {
    const size_t n = P.n;
    const int mat = ko.mat;
    const bool valid = mat >= 0 && mat < mt.n;
    f_nonfinite = ko.flag;
    if (!valid) {
        f_null = 1;
#pragma unroll
        for (int k = 0; k < 6; k++) P.sig[k * n + i] = 0.0;
        return;
    }
    double Fe[3][3];
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
        for (int r = 0; r < 3; r++) Fe[c][r] = P.Fe[(3 * c + r) * n + i];
    const double R = P.R[i];
    double epsp[6];
#pragma unroll
    for (int k = 0; k < 6; k++) epsp[k] = P.epsp[k * n + i];
    const double Ws0 = P.Ws[i], Wp0 = P.Wp[i];
    const double dt = ko.dt;
    double dF[3][3];
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
        for (int r = 0; r < 3; r++) dF[c][r] = (c == r ? 1.0 : 0.0) + ko.L[c][r] * dt;
    double S[3][3], etot[6], etr[6], einc[6];
    half_log_aat(ko.F, S);
    m3_voigt(S, etot);
    {
        double FeT[3][3];
        m3_mul(dF, Fe, FeT);
        half_log_aat(FeT, S);
        m3_voigt(S, etr);
    }
    half_log_aat(dF, S);
    m3_voigt(S, einc);
    const double zero[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    CRResult cr;
    constitutive(mt, mat, R, zero, etr, cr);
    f_plastic = cr.yielded;
    double eel[6];
#pragma unroll
    for (int k = 0; k < 6; k++) eel[k] = etr[k] - cr.depsp[k];
    {
        double Se[3][3];
        Se[0][0] = eel[0];
        Se[1][1] = eel[1];
        Se[2][2] = eel[2];
        Se[0][1] = 0.5 * eel[3];
        Se[1][0] = 0.5 * eel[3];
        Se[1][2] = 0.5 * eel[4];
        Se[2][1] = 0.5 * eel[4];
        Se[2][0] = 0.5 * eel[5];
        Se[0][2] = 0.5 * eel[5];
        exp_series(Se, Fe);
    }
    double Fn[3][3];
    m3_mul(ko.F, dF, Fn);
    const double det = +Fn[0][0] * (Fn[1][1] * Fn[2][2] - Fn[2][1] * Fn[1][2]) -
                       Fn[1][0] * (Fn[0][1] * Fn[2][2] - Fn[2][1] * Fn[0][2]) +
                       Fn[2][0] * (Fn[0][1] * Fn[1][2] - Fn[1][1] * Fn[0][2]);
    const double V = det * ko.V0;
    double sig[6];
#pragma unroll
    for (int k = 0; k < 6; k++) sig[k] = (1.0 / det) * cr.dsig[k];
    double Ws = Ws0, Wp = Wp0;
#pragma unroll
    for (int k = 0; k < 6; k++) Ws += sig[k] * einc[k] * V;
#pragma unroll
    for (int k = 0; k < 6; k++) Wp += sig[k] * cr.depsp[k] * V;
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
        for (int r = 0; r < 3; r++) {
            P.Fe[(3 * c + r) * n + i] = Fe[c][r];
            P.F[(3 * c + r) * n + i] = Fn[c][r];
        }
    P.V[i] = V;
#pragma unroll
    for (int k = 0; k < 6; k++) {
        P.eps[k * n + i] = etot[k];
        P.sig[k * n + i] = sig[k];
        P.epsp[k * n + i] = epsp[k] + cr.depsp[k];
    }
    P.Ws[i] = Ws;
    P.Wp[i] = Wp;
    P.R[i] = R + cr.dR;
    uint32_t bad = 0;
    for (int k = 0; k < 6; k++) bad |= !isfinite(sig[k]);
    f_nonfinite |= bad;
}

/* G2P pass 2 of the DoublePass loop with the large-strain formulation: the
 * kinematics of k_g2p_fused (x, L, control schedule, next cell) and then
 * large_strain_update in the same thread. */
template <bool MULTI>
__global__ __launch_bounds__(BLOCK) void k_g2p_large(Batch bt, PView P, NView Nd, BView B, MatTable mt,
                                                    unsigned long long *health) {
    uint32_t i, k;
    bool live;
    if (!particle_of_thread<MULTI>(bt, (uint32_t)P.n, i, live, k)) return;
    uint32_t f_outside = 0, f_null = 0, f_plastic = 0, f_nonfinite = 0;
    uint32_t next_key = MPM_NONE;
    if (live) {
        const double *g2 = Nd.g2;
        KinOut ko;
        g2p_kinematics_particle<VAR_DP, false, true>(bt, P, bt.inst[k], i,
                                                     [=](int, size_t node, double2 &r0, double2 &r1, double2 &r2,
                                                         double2 &r3) {
                                                         const double2 *q =
                                                             reinterpret_cast<const double2 *>(g2 + node * 8u);
                                                         r0 = q[0];
                                                         r1 = q[1];
                                                         r2 = q[2];
                                                         r3 = q[3];
                                                     },
                                                     next_key, f_outside, ko);
        B.key[i] = next_key;
        large_strain_update(P, mt, i, ko, f_null, f_plastic, f_nonfinite);
    }
    wave_key_add(B.count, next_key, next_key != MPM_NONE);
    block_count_add(health, HEALTH_OUTSIDE, f_outside);
    block_count_add(health, HEALTH_NULL, f_null);
    block_count_add(health, HEALTH_PLASTIC, f_plastic);
    block_count_add(health, HEALTH_NONFINITE, f_nonfinite);
}

/* Applies the network's output to the queued points.  The von Mises return
 * scales the trial deviator
 * (ConstitutiveRelation_calculateIncrement_VonMisesHardening.cpp:59-67):
 * the stress increment is C : de - alpha s_tr with the trial pressure kept,
 * alpha = 1 - K / sqrt(J2).  The network predicts the scalar alpha (its
 * output 0); s_tr is formed here in FP64, so a direction error is not
 * possible and the yield-surface check below bounds the error of alpha,
 * and with it the stress error, by the tolerance.  The answer is accepted
 * when the inputs were inside the training envelope and the new stress lies
 * on the yield surface within tol * sigma_y; otherwise the exact radial
 * return runs.  The plastic strain and hardening increments follow from the
 * accepted stress increment as in
 * ConstitutiveRelation_calculateIncrement_VonMisesHardening.cpp:69-90. */
__global__ __launch_bounds__(BLOCK) void k_surrogate_post(Batch bt, PView P, MatTable mt, unsigned long long *health,
                                                         SurrogateArgs sg) {
    const uint32_t count = *sg.qcount;
    uint32_t f_fallback = 0, f_nonfinite = 0;
    uint32_t slot = blockIdx.x * BLOCK + threadIdx.x;
    for (; slot < ((count + BLOCK - 1) / BLOCK) * BLOCK; slot += gridDim.x * BLOCK) {
        f_fallback = 0;
        f_nonfinite = 0;
        if (slot < count) {
            const size_t n = P.n;
            const uint32_t i = sg.qid[slot];
            const int m = P.mat[i];
            double de[6], sig[6];
            for (int k = 0; k < 6; k++) de[k] = P.de[k * n + i];
            for (int k = 0; k < 6; k++) sig[k] = P.sig[k * n + i];
            const double R = P.R[i];
            double dtr[6], dev[6], K, feat[MLP_IN];
            surrogate_features(mt, m, R, sig, de, sg.stress_scale, sg.strain_scale, dtr, dev, K, feat);
            CRResult cr;
            bool accept = sg.qenv[slot] != 0;
            const double alpha = (double)sg.qy[slot] * (double)sg.norm.out_std[0] + (double)sg.norm.out_mean[0];
            for (int k = 0; k < 6; k++) cr.dsig[k] = dtr[k] - alpha * dev[k];
            if (accept) {
                double s2[6];
                for (int k = 0; k < 6; k++) s2[k] = sig[k] + cr.dsig[k];
                double J = 0.0;
                J += 1.0 / 6.0 * (s2[0] - s2[1]) * (s2[0] - s2[1]);
                J += 1.0 / 6.0 * (s2[1] - s2[2]) * (s2[1] - s2[2]);
                J += 1.0 / 6.0 * (s2[2] - s2[0]) * (s2[2] - s2[0]);
                J += s2[3] * s2[3] + s2[4] * s2[4] + s2[5] * s2[5];
                double resid = fabs(sqrt(J) - K);
                accept = isfinite(resid) && resid <= sg.tol * mt.yield[m];
            }
            if (accept) {
                const double E = mt.E[m], nu = mt.nu[m];
                double G = 0.5 * E / (1.0 + nu);
                double dee[6];
                dee[0] = cr.dsig[0] / E - cr.dsig[1] * nu / E - cr.dsig[2] * nu / E;
                dee[1] = cr.dsig[1] / E - cr.dsig[0] * nu / E - cr.dsig[2] * nu / E;
                dee[2] = cr.dsig[2] / E - cr.dsig[0] * nu / E - cr.dsig[1] * nu / E;
                dee[3] = cr.dsig[3] / G;
                dee[4] = cr.dsig[4] / G;
                dee[5] = cr.dsig[5] / G;
                for (int k = 0; k < 6; k++) cr.depsp[k] = de[k] - dee[k];
                cr.dR = 0.0;
                if (mt.type[m] == 4) {
                    double ep = 0.0;
                    for (int k = 0; k < 6; k++) ep += cr.depsp[k] * cr.depsp[k];
                    ep *= 2.0 / 3.0;
                    ep = sqrt(ep);
                    if ((sqrt(2.0 / 3.0) * mt.c1[m] - R) > 0.0) cr.dR = mt.c0[m] * (sqrt(2.0 / 3.0) * mt.c1[m] - R) * ep;
                }
                cr.yielded = true;
            } else {
                constitutive(mt, m, R, sig, de, cr);
                f_fallback = 1;
            }
            PAccum acc;
            load_accum(P, i, bt.variant == VAR_DP, acc);
            if (bt.variant == VAR_DP)
                f_nonfinite = finish_update(P, i, de, sig, cr, R, P.V[i], bt.inst[inst_of_particle(bt, i)].dt, acc);
            else
                f_nonfinite = finish_update_new_stress(P, i, sig, cr, R, P.V[i], de, acc);
        }
        block_count_add(health, HEALTH_SURROGATE_FALLBACKS, f_fallback);
        block_count_add(health, HEALTH_NONFINITE, f_nonfinite);
    }
}

/* ---- reportConsole sums (PhysicsEngine.cpp:40-183) in list order ---- */

#define MON_CHUNK 1024

/* One block. Lanes stage a chunk of values in LDS; lane 0 adds them in list
 * order, so every sum is the reference's sequential sum. */
__global__ __launch_bounds__(BLOCK) void k_monitors(PView P, NView Nd, MonLists ml, double *out) {
    __shared__ double a0[MON_CHUNK], a1[MON_CHUNK], a2[MON_CHUNK];
    const size_t n = P.n;
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

/* Graph-safe clear of a small counter array (no memset node). */
__global__ void k_zero_u32(uint32_t *p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

/* ---- launchers ---- */

static inline unsigned blocks_for(int64_t n, int64_t per) { return (unsigned)((n + per - 1) / per); }

hipError_t launch_keys(const Batch &bt, const PView &P, const BView &B, hipStream_t s) {
    if (P.n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_keys, dim3(blocks_for(P.n, BLOCK)), dim3(BLOCK), 0, s, bt, P, B);
    return hipGetLastError();
}

int64_t scan_block_count(int64_t necells) { return (necells + SCAN_TILE - 1) / SCAN_TILE; }
int64_t scan_padded_cells(int64_t necells) { return std::max<int64_t>(scan_block_count(necells), 1) * SCAN_TILE; }

hipError_t launch_bin(const PView &P, const BView &B, uint64_t *block_sum, hipStream_t s) {
    const int64_t nb = scan_padded_cells(B.necells) / SCAN_TILE;
    hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nb), dim3(BLOCK), 0, s, B.count, block_sum);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(BLOCK), 0, s, block_sum, nb);
    hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nb), dim3(BLOCK), 0, s, B.count, block_sum, B.start, B.slot,
                       B.occ, B.ostart, B.nocc);
    if (P.n > 0)
        hipLaunchKernelGGL(k_scatter, dim3(blocks_for(P.n, (int64_t)BLOCK * SCATTER_RUNS)), dim3(BLOCK), 0, s, P, B);
    hipLaunchKernelGGL(k_cell_sort, dim3((unsigned)std::min<int64_t>(blocks_for(std::max<int64_t>(P.n, 1), BLOCK), 4096)),
                       dim3(BLOCK), 0, s, B);
    return hipGetLastError();
}

/* Blocks for the cell partials: one group of PART_CELLS occupied cells per
 * block and pass, at most as many groups as particles allow. */
static unsigned partial_grid(int64_t n) {
    int64_t b = (std::max<int64_t>(n, 1) + PART_CELLS - 1) / PART_CELLS;
    return (unsigned)std::min<int64_t>(b, 8192);
}

/* Blocks of the per-particle kernels that take a batch wave by wave: one
 * wave per wave-table entry. */
static unsigned wave_grid(const Batch &bt) {
    return (unsigned)std::max<int64_t>(((int64_t)bt.nwaves + BLOCK / 64 - 1) / (BLOCK / 64), 1);
}

/* Launches KERNEL with M = whether the batch has more than one instance. */
#define LAUNCH_MULTI(bt, KERNEL, ...)                                   \
    do {                                                                 \
        if ((bt).ninst > 1) {                                            \
            constexpr bool M = true;                                     \
            hipLaunchKernelGGL(KERNEL, __VA_ARGS__);                     \
        } else {                                                         \
            constexpr bool M = false;                                    \
            hipLaunchKernelGGL(KERNEL, __VA_ARGS__);                     \
        }                                                                \
    } while (0)

template <int VAR>
static void launch_g2p_strain(int mode, dim3 grid, hipStream_t s, const Batch &bt, const PView &P, const NView &N,
                              const BView &B, const MatTable &mt, unsigned long long *health, const SurrogateArgs &sg,
                              hipEvent_t mid) {
    if (VAR == VAR_DP && bt.strain == STRAIN_LARGE) {
        /* the engine allows the large-strain pass with the exact map only */
        LAUNCH_MULTI(bt, (k_g2p_large<M>), dim3(M ? wave_grid(bt) : grid.x), dim3(BLOCK), 0, s, bt, P, N, B, mt,
                     health);
        if (mid) (void)hipEventRecord(mid, s);
        return;
    }
    /* exact map: the fused pass (a batch's waves each take one instance,
     * whose row then sits in scalar registers as for a single instance) */
    if (mode == MODE_EXACT) {
        if (bt.ninst > 1)
            hipLaunchKernelGGL((k_g2p_fused<VAR, true>), dim3(wave_grid(bt)), dim3(BLOCK), 0, s, bt, P, N, B, mt,
                               health);
        else
            hipLaunchKernelGGL((k_g2p_fused<VAR, false>), grid, dim3(BLOCK), 0, s, bt, P, N, B, mt, health);
        if (mid) (void)hipEventRecord(mid, s);
        return;
    }
    if (mode == MODE_SURROGATE && VAR == VAR_DP && bt.ninst == 1) {
        hipLaunchKernelGGL(k_g2p_fused_queue, grid, dim3(BLOCK), 0, s, bt, P, N, B, mt, health, sg);
        if (mid) (void)hipEventRecord(mid, s);
        return;
    }
    if (bt.ninst > 1)
        hipLaunchKernelGGL((k_g2p_kinematics<VAR, true>), grid, dim3(BLOCK), 0, s, bt, P, N, B, health);
    else
        hipLaunchKernelGGL((k_g2p_kinematics<VAR, false>), grid, dim3(BLOCK), 0, s, bt, P, N, B, health);
    if (mid) (void)hipEventRecord(mid, s);
    if (mode == MODE_SURROGATE)
        hipLaunchKernelGGL((k_g2p_constitutive<MODE_SURROGATE, VAR>), grid, dim3(BLOCK), 0, s, bt, P, mt, health, sg);
    else if (mode == MODE_CAPTURE)
        hipLaunchKernelGGL((k_g2p_constitutive<MODE_CAPTURE, VAR>), grid, dim3(BLOCK), 0, s, bt, P, mt, health, sg);
    else
        hipLaunchKernelGGL((k_g2p_constitutive<MODE_EXACT, VAR>), grid, dim3(BLOCK), 0, s, bt, P, mt, health, sg);
}

hipError_t launch_phase(int phase, const Batch &bt, const PView &P, const NView &N, const BView &B,
                        const CellParts &C, int order, const MatTable &mt, unsigned long long *health,
                        const SurrogateArgs &sg, hipStream_t s, hipEvent_t mid) {
    switch (phase) {
    case PHASE_P2G_GRID:
        if (!bt.ntiles) {
            if (mid) (void)hipEventRecord(mid, s);
            break;
        }
        if (order == ORDER_TILE) {
            if (!bt.nstiles || !bt.srec || !bt.sflag) return hipErrorInvalidValue;
            if (bt.fext)
                LAUNCH_MULTI(bt, (k_tile_p2g<false, M, true>), dim3(bt.nstiles), dim3(BLOCK), 0, s, bt, P, N, B);
            else
                LAUNCH_MULTI(bt, (k_tile_p2g<false, M, false>), dim3(bt.nstiles), dim3(BLOCK), 0, s, bt, P, N, B);
            if (mid) (void)hipEventRecord(mid, s);
            LAUNCH_MULTI(bt, (k_tile_bnd<false, M>), dim3(bt.nstiles), dim3(BLOCK), 0, s, bt, P, N, B);
        } else if (order == ORDER_CELL) {
            if (P.n > 0) {
                if (bt.fext)
                    LAUNCH_MULTI(bt, (k_cell_partials<false, M, true>), dim3(partial_grid(P.n)), dim3(BLOCK), 0, s,
                                 bt, P, N, B, C);
                else
                    LAUNCH_MULTI(bt, (k_cell_partials<false, M, false>), dim3(partial_grid(P.n)), dim3(BLOCK), 0, s,
                                 bt, P, N, B, C);
            }
            if (mid) (void)hipEventRecord(mid, s);
            LAUNCH_MULTI(bt, (k_node_cells<false, M>), dim3(bt.ntiles), dim3(NODE_BLOCK), 0, s, bt, P, N, B, C);
        } else if (order == ORDER_PAIR) {
            if (P.n > 0) {
                if (bt.fext)
                    LAUNCH_MULTI(bt, (k_cell_partials<false, M, true, true>), dim3(partial_grid(P.n)),
                                 dim3(BLOCK), 0, s, bt, P, N, B, C);
                else
                    LAUNCH_MULTI(bt, (k_cell_partials<false, M, false, true>), dim3(partial_grid(P.n)),
                                 dim3(BLOCK), 0, s, bt, P, N, B, C);
            }
            if (mid) (void)hipEventRecord(mid, s);
            LAUNCH_MULTI(bt, (k_node_cells<false, M, true>), dim3(bt.ntiles), dim3(NODE_BLOCK), 0, s, bt, P, N, B, C);
        } else {
            if (!C.list) return hipErrorInvalidValue;
            if (P.n > 0) {
                if (bt.fext)
                    LAUNCH_MULTI(bt, (k_list_terms<false, M, true>), dim3(blocks_for(P.n, BLOCK)), dim3(BLOCK), 0, s,
                                 bt, P, B, C);
                else
                    LAUNCH_MULTI(bt, (k_list_terms<false, M, false>), dim3(blocks_for(P.n, BLOCK)), dim3(BLOCK), 0, s,
                                 bt, P, B, C);
            }
            if (mid) (void)hipEventRecord(mid, s);
            if (bt.fext)
                hipLaunchKernelGGL(k_p2g_list<true>, dim3(bt.ntiles), dim3(NODE_BLOCK), 0, s, bt, P, N, B, C);
            else
                hipLaunchKernelGGL(k_p2g_list<false>, dim3(bt.ntiles), dim3(NODE_BLOCK), 0, s, bt, P, N, B, C);
        }
        break;
    case PHASE_G2P_VELOCITY:
        if (P.n > 0)
            LAUNCH_MULTI(bt, (k_g2p_velocity<M>), dim3(M ? wave_grid(bt) : blocks_for(P.n, BLOCK)), dim3(BLOCK), 0,
                         s, bt, P, N);
        if (mid) (void)hipEventRecord(mid, s);
        break;
    case PHASE_REMAP_GRID:
        if (!bt.ntiles) {
            if (mid) (void)hipEventRecord(mid, s);
            break;
        }
        if (order == ORDER_TILE) {
            if (!bt.nstiles || !bt.srec || !bt.sflag) return hipErrorInvalidValue;
            LAUNCH_MULTI(bt, (k_tile_p2g<true, M, false>), dim3(bt.nstiles), dim3(BLOCK), 0, s, bt, P, N, B);
            if (mid) (void)hipEventRecord(mid, s);
            LAUNCH_MULTI(bt, (k_tile_bnd<true, M>), dim3(bt.nstiles), dim3(BLOCK), 0, s, bt, P, N, B);
        } else if (order == ORDER_CELL) {
            if (P.n > 0) {
                LAUNCH_MULTI(bt, (k_cell_partials<true, M>), dim3(partial_grid(P.n)), dim3(BLOCK), 0, s, bt, P, N,
                             B, C);
            }
            if (mid) (void)hipEventRecord(mid, s);
            LAUNCH_MULTI(bt, (k_node_cells<true, M>), dim3(bt.ntiles), dim3(NODE_BLOCK), 0, s, bt, P, N, B, C);
        } else if (order == ORDER_PAIR) {
            if (P.n > 0) {
                LAUNCH_MULTI(bt, (k_cell_partials<true, M, false, true>), dim3(partial_grid(P.n)),
                             dim3(BLOCK), 0, s, bt, P, N, B, C);
            }
            if (mid) (void)hipEventRecord(mid, s);
            LAUNCH_MULTI(bt, (k_node_cells<true, M, true>), dim3(bt.ntiles), dim3(NODE_BLOCK), 0, s, bt, P, N, B, C);
        } else {
            if (!C.list) return hipErrorInvalidValue;
            if (P.n > 0)
                LAUNCH_MULTI(bt, (k_list_terms<true, M, false>), dim3(blocks_for(P.n, BLOCK)), dim3(BLOCK), 0, s,
                             bt, P, B, C);
            if (mid) (void)hipEventRecord(mid, s);
            hipLaunchKernelGGL(k_remap_list, dim3(bt.ntiles), dim3(NODE_BLOCK), 0, s, bt, P, N, B, C);
        }
        break;
    case PHASE_G2P_STRAIN:
        if (P.n == 0 && mid) (void)hipEventRecord(mid, s);
        if (P.n > 0) {
            dim3 grid(blocks_for(P.n, BLOCK));
            if (bt.variant == VAR_DPC)
                launch_g2p_strain<VAR_DPC>(sg.mode, grid, s, bt, P, N, B, mt, health, sg, mid);
            else if (bt.variant == VAR_SP)
                launch_g2p_strain<VAR_SP>(sg.mode, grid, s, bt, P, N, B, mt, health, sg, mid);
            else
                launch_g2p_strain<VAR_DP>(sg.mode, grid, s, bt, P, N, B, mt, health, sg, mid);
        }
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_zero_u32(uint32_t *p, int n, hipStream_t s) {
    hipLaunchKernelGGL(k_zero_u32, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

hipError_t launch_surrogate_post(const Batch &bt, const PView &P, const MatTable &mt, unsigned long long *health,
                                 const SurrogateArgs &sg, hipStream_t s) {
    if (P.n == 0) return hipSuccess;
    /* the queue length is only known on the device: a grid that fills the
     * chip strides over a full queue; with a short queue the surplus exits
     * at once */
    unsigned grid = (unsigned)std::min<int64_t>(blocks_for(P.n, BLOCK), 2048);
    hipLaunchKernelGGL(k_surrogate_post, dim3(grid), dim3(BLOCK), 0, s, bt, P, mt, health, sg);
    return hipGetLastError();
}

hipError_t launch_prepare(const Batch &bt, hipStream_t s) {
    hipLaunchKernelGGL(k_prepare, dim3((bt.ninst + 63) / 64), dim3(64), 0, s, bt);
    return hipGetLastError();
}

hipError_t launch_advance(const Batch &bt, hipStream_t s) {
    hipLaunchKernelGGL(k_advance, dim3((bt.ninst + 63) / 64), dim3(64), 0, s, bt);
    return hipGetLastError();
}

hipError_t launch_set_dt(const Batch &bt, int inst, double dt, hipStream_t s) {
    hipLaunchKernelGGL(k_set_dt, dim3(1), dim3(1), 0, s, bt, inst, dt);
    return hipGetLastError();
}

hipError_t launch_monitors(const PView &P, const NView &N, int64_t nnodes, const MonLists &ml, double *out,
                           hipStream_t s) {
    (void)nnodes;
    hipLaunchKernelGGL(k_monitors, dim3(1), dim3(BLOCK), 0, s, P, N, ml, out);
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
