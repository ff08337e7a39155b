/*
 * mpm_engine.hip — host side of the C ABI declared in include/mpm.h.
 *
 * The engine owns every device buffer (particles, nodes, cell bins) and one
 * HIP stream; the reference's PhysicsEngine owns the same objects as heap
 * AoS (PhysicsEngine.h:82-117).  An engine holds one or more independent
 * instances (scenes) laid out back to back in the same arrays; each step is
 * one set of launches over all of them.  Per-instance geometry, step size,
 * damping, time and timeline sit in a device table (Inst); time advances on
 * the device, and the host keeps a mirror that adds the same dt in the same
 * order.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/mpm.h"
#include "mpm_kernels.h"
#include "mpm_surrogate.h"

namespace {

const int kFieldComponents[MPM_PF_COUNT] = {3, 3, 3, 1, 1, 1, 9, 6, 6, 6, 1, 1, 1, 1, 1, 9};

std::string g_create_error = "no error";

template <typename T>
void dfree(T *&p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

/* Host-side state of one instance. */
struct HostInst {
    GridGeom g{};
    int64_t node_off = 0, cell_off = 0, part_off = 0, npart = 0;
    uint32_t tile_off = 0, tiles = 0;
    uint32_t stile_off = 0, stiles = 0; /* summation tiles of the tile order */
    double time = 0.0, dt_max = 1.0e-3, time_end = 10.0;
    double damping = 0.0, mass_min = 0.0;
    std::vector<double> tl; /* t, vx, vy, vz per point */
};

bool make_geom(const double L[3], const int32_t cells[3], GridGeom &g) {
    for (int a = 0; a < 3; a++) {
        if (cells[a] < 1 || !(L[a] > 0.0) || !std::isfinite(L[a])) return false;
        g.L[a] = L[a];
        g.h[a] = L[a] / (double)cells[a];
        g.cells[a] = cells[a];
        g.nn[a] = cells[a] + 1;
        g.ec[a] = cells[a] + 2;
    }
    g.nnodes = (int64_t)g.nn[0] * g.nn[1] * g.nn[2];
    g.necells = (int64_t)g.ec[0] * g.ec[1] * g.ec[2];
    return true;
}

/* Summation tiles of the tile order over the instance's binning cells. */
uint32_t stiles_of(const GridGeom &g) {
    int64_t t = (int64_t)((g.ec[0] + ST_X - 1) / ST_X) * ((g.ec[1] + ST_Y - 1) / ST_Y) * ((g.ec[2] + ST_Z - 1) / ST_Z);
    return (uint32_t)t;
}

uint32_t tiles_of(const GridGeom &g, const int shape[3]) {
    int64_t t = (int64_t)((g.nn[0] + shape[0] - 1) / shape[0]) * ((g.nn[1] + shape[1] - 1) / shape[1]) *
                ((g.nn[2] + shape[2] - 1) / shape[2]);
    return (uint32_t)t;
}

/* Node tile shapes the autotuner tries (all NODE_BLOCK nodes).  64 x 1 x 1
 * is left out: the repeated particle-to-grid launches the autotuner times
 * can rank it first on the 8.0 M-particle crush, but in whole steps both
 * node passes run slower with it than with 32 x 2 x 1 (0.210 / 0.191 against
 * 0.198 / 0.180 ms). */
const int kTileShapes[][3] = {{4, 4, 4}, {8, 4, 2}, {4, 8, 2}, {8, 8, 1}, {16, 4, 1},
                              {4, 16, 1}, {2, 8, 4}, {16, 2, 2}, {32, 2, 1}};
const int kNumTileShapes = sizeof(kTileShapes) / sizeof(kTileShapes[0]);

}  // namespace

struct mpm_engine {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err = "no error";

    /* instances */
    std::vector<HostInst> inst;
    int64_t nnodes = 0, ncells = 0;
    uint32_t ntiles = 0;
    Inst *dinst = nullptr;
    uint32_t *dtile_off = nullptr, *dpinst = nullptr;
    uint32_t *dwave = nullptr; /* batch wave table, see Batch::wave */
    uint32_t nwaves = 0;
    double *dtl = nullptr;
    bool table_dirty = true;
    bool layout_set = false;

    /* nodes */
    double *nm = nullptr, *np = nullptr, *nv = nullptr, *nf = nullptr, *na = nullptr, *nu = nullptr,
           *nfty = nullptr, *nfric = nullptr;
    uint8_t *nact = nullptr, *nfix = nullptr;
    std::vector<int64_t> fixed_y; /* host copy, ascending */

    /* particles */
    int64_t n = 0;
    double *pf[MPM_PF_COUNT] = {};
    int32_t *pmat = nullptr, *pdc_rank = nullptr;
    uint8_t *pdc_axes = nullptr;
    float *pdc_mult = nullptr;
    double *pde = nullptr;
    uint8_t *pflag = nullptr;
    bool has_fext = false;

    /* bins */
    uint32_t *bkey = nullptr, *bcount = nullptr, *bstart = nullptr, *bsorted = nullptr;
    uint32_t *bslot = nullptr, *bocc = nullptr, *bnocc = nullptr;
    uint64_t *bblock = nullptr;
    bool keys_valid = false;

    /* node tile shape */
    int tile[3] = {32, 2, 1};

    /* node summation order and the cell-order partial sums */
    int order = ORDER_PAIR;
    int variant = VAR_DP;
    int strain = STRAIN_SMALL;
    uint8_t *dtflag = nullptr; /* node tile flags, see Batch::tflag */
    int64_t tflag_cap = 0;
    /* tile order: summation tile offsets, occupancy flags and face records */
    uint32_t nstiles = 0;
    uint32_t *dstile_off = nullptr;
    uint8_t *dsflag = nullptr;
    double *dsrec = nullptr;
    CellParts parts{};
    int64_t parts_cap = 0;

    /* materials, lists */
    MatTable mt{};
    std::vector<int64_t> mon[MPM_MON_COUNT];
    int64_t cycle = 0;

    unsigned long long *d_health = nullptr;
    double *d_mon_out = nullptr;
    int64_t *d_monbuf = nullptr;
    int64_t monbuf_len = 0;
    int64_t steps = 0;

    /* surrogate and sample capture */
    mpm_surrogate *sur = nullptr;
    int sur_material = -1;
    double sur_tol = 1.0e-4;
    uint32_t *qcount = nullptr, *qid = nullptr;
    float *qx = nullptr, *qy = nullptr;
    uint8_t *qenv = nullptr;
    int64_t qcap = 0;
    double *cap = nullptr;
    unsigned long long *capcount = nullptr;
    int64_t capcap = 0;

    /* captured step graphs, keyed by step count */
    std::map<int, hipGraphExec_t> graphs;

    int fail(int code, const char *fmt, ...) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        err = buf;
        return code;
    }
    int hip(hipError_t st, const char *what) {
        if (st == hipSuccess) return MPM_OK;
        return fail(st == hipErrorOutOfMemory ? MPM_ERR_NOMEM : MPM_ERR_HIP, "%s: %s", what, hipGetErrorString(st));
    }

    void drop_graphs() {
        for (auto &kv : graphs) (void)hipGraphExecDestroy(kv.second);
        graphs.clear();
    }
    void free_nodes() {
        dfree(dtflag);
        tflag_cap = 0;
        dfree(nm); dfree(np); dfree(nv); dfree(nf); dfree(na); dfree(nu); dfree(nfty); dfree(nfric);
        dfree(nact); dfree(nfix);
        fixed_y.clear();
        dfree(bcount); dfree(bstart); dfree(bblock); dfree(bslot); dfree(bnocc);
        dfree(dinst); dfree(dtile_off); dfree(dtl);
        free_stiles();
        free_parts();
    }
    void free_stiles() {
        dfree(dstile_off); dfree(dsflag); dfree(dsrec);
    }
    /* Tile order: the face records (ST_BND per summation tile, 64 bytes
     * each) and occupancy flags, allocated on first use. */
    int ensure_stiles() {
        if (dsrec) return MPM_OK;
        int rc;
        const size_t nt = std::max<size_t>(nstiles, 1);
        if ((rc = hip(hipMalloc(&dsflag, nt), "alloc summation tile flags")) ||
            (rc = hip(hipMemset(dsflag, 0, nt), "clear summation tile flags")) ||
            (rc = hip(hipMalloc(&dsrec, nt * (size_t)ST_BND * ST_REC * sizeof(double)), "alloc tile face records"))) {
            dfree(dsflag);
            dfree(dsrec);
            return rc;
        }
        drop_graphs();
        return MPM_OK;
    }
    void free_parts() {
        dfree(parts.w); dfree(bocc); dfree(parts.list);
        parts = CellParts{};
        parts_cap = 0;
        occ_cap = 0;
    }
    /* The occupied-cell list written by every scan, and the particle start
     * of each occupied cell in rank order (one more entry: the total):
     * occupied cells never exceed min(cells, particles). */
    int64_t occ_cap = 0;
    int ensure_occ() {
        int64_t cap = std::max<int64_t>(1, std::min<int64_t>(ncells, n));
        if (cap == occ_cap && bocc) return MPM_OK;
        dfree(bocc);
        occ_cap = 0;
        int rc = hip(hipMalloc(&bocc, 4 * (2 * (size_t)cap + 1)), "alloc occupied cells");
        if (rc) return rc;
        occ_cap = cap;
        drop_graphs();
        return MPM_OK;
    }
    int ensure_parts() {
        int64_t cap = occ_cap;
        if (cap == parts_cap && parts.w) return MPM_OK;
        dfree(parts.w);
        parts.w = nullptr;
        parts.cap = 0;
        parts_cap = 0;
        const size_t c8 = 8 * (size_t)cap;
        int rc;
        if ((rc = hip(hipMalloc(&parts.w, 64 * c8), "alloc partials"))) {
            free_parts();
            return rc;
        }
        parts.cap = cap;
        parts_cap = cap;
        drop_graphs();
        return MPM_OK;
    }
    /* Particle order: the per-(particle, corner) terms, 4 doubles each in
     * P2G (3 in the remap), and a P2G row of 8 doubles (4 without f_ext)
     * per particle: 40 doubles per particle at most. */
    int ensure_list() {
        const int64_t cap = std::max<int64_t>(n, 1);
        if (parts.list && parts.lcap == cap) return MPM_OK;
        dfree(parts.list);
        parts.list = nullptr;
        parts.lcap = 0;
        int rc;
        if ((rc = hip(hipMalloc(&parts.list, (size_t)cap * 40u * sizeof(double)), "alloc particle-order terms")))
            return rc;
        parts.lcap = cap;
        drop_graphs();
        return MPM_OK;
    }
    void free_particles() {
        for (int f = 0; f < MPM_PF_COUNT; f++) dfree(pf[f]);
        dfree(pmat); dfree(pdc_rank); dfree(pdc_axes); dfree(pdc_mult); dfree(pde); dfree(pflag);
        dfree(bkey); dfree(bsorted); dfree(dpinst); dfree(dwave);
        nwaves = 0;
        for (int k = 0; k < MPM_MON_COUNT; k++) mon[k].clear();
    }
    void free_queue() {
        dfree(qid); dfree(qx); dfree(qy); dfree(qenv);
        qcap = 0;
    }
    int ensure_queue() {
        if (qcap == n && qid) return MPM_OK;
        free_queue();
        if (n == 0) return MPM_OK;
        const size_t sn = (size_t)n;
        int rc;
        if ((rc = hip(hipMalloc(&qid, 4 * sn), "alloc queue")) ||
            (rc = hip(hipMalloc(&qx, 4 * MLP_IN * sn), "alloc queue")) ||
            (rc = hip(hipMalloc(&qy, 4 * MLP_OUT * sn), "alloc queue")) ||
            (rc = hip(hipMalloc(&qenv, sn), "alloc queue"))) {
            free_queue();
            return rc;
        }
        qcap = n;
        return MPM_OK;
    }
    SurrogateArgs sargs() const {
        SurrogateArgs a{};
        a.mode = sur ? MODE_SURROGATE : (cap ? MODE_CAPTURE : MODE_EXACT);
        a.material = sur_material;
        a.tol = sur_tol;
        if (sur) {
            a.stress_scale = sur->stress_scale;
            a.strain_scale = sur->strain_scale;
            a.norm = sur->norm;
        }
        a.qcount = qcount; a.qid = qid; a.qx = qx; a.qy = qy; a.qenv = qenv;
        a.cap = cap; a.capcount = capcount; a.capcap = capcap;
        return a;
    }

    PView pview() const {
        PView P;
        P.x = pf[MPM_PF_POSITION];
        P.v = pf[MPM_PF_VELOCITY];
        P.fext = pf[MPM_PF_FORCE_EXTERNAL];
        P.m = pf[MPM_PF_MASS];
        P.V = pf[MPM_PF_VOLUME];
        P.V0 = pf[MPM_PF_VOLUME_INITIAL];
        P.F = pf[MPM_PF_DEFORMATION_GRADIENT];
        P.Fe = pf[MPM_PF_DEFORMATION_GRADIENT_ELASTIC];
        P.eps = pf[MPM_PF_STRAIN];
        P.sig = pf[MPM_PF_STRESS];
        P.epsp = pf[MPM_PF_STRAIN_PLASTIC];
        P.R = pf[MPM_PF_BACKSTRESS];
        P.Ws = pf[MPM_PF_ENERGY_STRAIN];
        P.Wmax = pf[MPM_PF_ENERGY_STRAIN_MAX];
        P.Wti = pf[MPM_PF_ENERGY_STRAIN_TI];
        P.Wp = pf[MPM_PF_ENERGY_PLASTIC];
        P.mat = pmat;
        P.dc_rank = pdc_rank;
        P.dc_axes = pdc_axes;
        P.dc_mult = pdc_mult;
        P.de = pde;
        P.flag = pflag;
        P.n = n;
        return P;
    }
    NView nview() const {
        NView N;
        N.m = nm; N.p = np; N.v = nv; N.f = nf; N.g1 = na; N.g2 = nu; N.fty = nfty;
        N.active = nact; N.fixed = nfix; N.friction = nfric;
        return N;
    }
    BView bview() const {
        BView B;
        B.key = bkey; B.count = bcount; B.start = bstart; B.sorted = bsorted;
        B.slot = bslot; B.occ = bocc; B.nocc = bnocc;
        B.ostart = bocc ? bocc + occ_cap : nullptr;
        B.necells = ncells;
        return B;
    }
    Batch batch() const {
        Batch b;
        b.inst = dinst;
        b.tile_off = dtile_off;
        b.pinst = inst.size() > 1 ? dpinst : nullptr;
        b.wave = inst.size() > 1 ? dwave : nullptr;
        b.nwaves = inst.size() > 1 ? nwaves : 0;
        b.tl = dtl;
        b.ninst = (int)inst.size();
        b.ntiles = ntiles;
        b.nnodes = nnodes;
        b.tx = tile[0];
        b.ty = tile[1];
        b.tz = tile[2];
        b.variant = variant;
        b.strain = strain;
        b.fext = has_fext ? 1 : 0;
        b.nmat = mt.n;
        b.tflag = order == ORDER_CELL || order == ORDER_PAIR ? dtflag : nullptr;
        b.stile_off = dstile_off;
        b.nstiles = nstiles;
        b.sflag = dsflag;
        b.srec = dsrec;
        return b;
    }

    /* Phase order of the step loop: DoublePass and its contact variant run
     * P2G + grid, G2P velocity, remap + BC, G2P strain; the single pass
     * runs its velocity sum + grid update (the remap phase) before its one
     * grid-to-particle pass (velocity, then strain). */
    void phase_order(int ph[4]) const {
        ph[0] = PHASE_P2G_GRID;
        ph[1] = variant == VAR_SP ? PHASE_REMAP_GRID : PHASE_G2P_VELOCITY;
        ph[2] = variant == VAR_SP ? PHASE_G2P_VELOCITY : PHASE_REMAP_GRID;
        ph[3] = PHASE_G2P_STRAIN;
    }

    /* Tile offsets of every instance for the current tile shape. */
    int layout_tiles() {
        ntiles = 0;
        for (auto &I : inst) {
            I.tile_off = ntiles;
            I.tiles = tiles_of(I.g, tile);
            if ((int64_t)ntiles + I.tiles >= (int64_t)0xFFFFFFFFll) return fail(MPM_ERR_RANGE, "too many node tiles");
            ntiles += I.tiles;
        }
        if ((int64_t)ntiles > tflag_cap) {
            dfree(dtflag);
            tflag_cap = 0;
            int rc = hip(hipMalloc(&dtflag, std::max<size_t>(ntiles, 1)), "alloc tile flags");
            if (!rc) rc = hip(hipMemset(dtflag, 0, std::max<size_t>(ntiles, 1)), "clear tile flags");
            if (rc) return rc;
            tflag_cap = ntiles;
        }
        std::vector<uint32_t> to(inst.size() + 1);
        for (size_t k = 0; k < inst.size(); k++) to[k] = inst[k].tile_off;
        to[inst.size()] = ntiles;
        table_dirty = true;
        return hip(hipMemcpy(dtile_off, to.data(), 4 * to.size(), hipMemcpyHostToDevice), "upload tile offsets");
    }

    /* Allocates nodes, bins and the instance table for the current `inst`. */
    int build_layout() {
        (void)hipStreamSynchronize(stream);
        drop_graphs();
        free_nodes();
        layout_set = false;
        nnodes = ncells = 0;
        int64_t st = 0;
        for (auto &I : inst) {
            I.node_off = nnodes;
            I.cell_off = ncells;
            I.stile_off = (uint32_t)st;
            I.stiles = stiles_of(I.g);
            nnodes += I.g.nnodes;
            ncells += I.g.necells;
            st += I.stiles;
        }
        if (st >= (int64_t)0xFFFFFFFFll) return fail(MPM_ERR_RANGE, "too many summation tiles");
        nstiles = (uint32_t)st;
        if (scan_padded_cells(ncells) >= (int64_t)0xFFFFFFFFll) return fail(MPM_ERR_RANGE, "grid too large: %lld cells", (long long)ncells);
        const size_t nn = (size_t)nnodes;
        int rc = 0;
        if ((rc = hip(hipMalloc(&nm, 8 * nn), "alloc node mass")) ||
            (rc = hip(hipMalloc(&np, 24 * nn), "alloc node momentum")) ||
            (rc = hip(hipMalloc(&nv, 24 * nn), "alloc node velocity")) ||
            (rc = hip(hipMalloc(&nf, 24 * nn), "alloc node force")) ||
            (rc = hip(hipMalloc(&na, 32 * nn), "alloc node G2P records")) ||
            (rc = hip(hipMalloc(&nu, 64 * nn), "alloc node G2P records")) ||
            (rc = hip(hipMalloc(&nfty, 8 * nn), "alloc node reaction")) ||
            (rc = hip(hipMalloc(&nfric, 8 * nn), "alloc node friction")) ||
            (rc = hip(hipMalloc(&nact, nn), "alloc node active")) ||
            (rc = hip(hipMalloc(&nfix, nn), "alloc node fixed")) ||
            (rc = hip(hipMalloc(&bcount, 4 * (size_t)scan_padded_cells(ncells)), "alloc cell counts")) ||
            (rc = hip(hipMalloc(&bstart, 4 * ((size_t)scan_padded_cells(ncells) + 1)), "alloc cell starts")) ||
            (rc = hip(hipMalloc(&bblock, 8 * (size_t)scan_block_count(scan_padded_cells(ncells))), "alloc scan blocks")) ||
            (rc = hip(hipMalloc(&bslot, 4 * (size_t)scan_padded_cells(ncells)), "alloc cell slots")) ||
            (rc = hip(hipMalloc(&bnocc, 4), "alloc occupied count")) ||
            (rc = hip(hipMalloc(&dinst, sizeof(Inst) * inst.size()), "alloc instance table")) ||
            (rc = hip(hipMalloc(&dtile_off, 4 * (inst.size() + 1)), "alloc tile offsets")) ||
            (rc = hip(hipMalloc(&dstile_off, 4 * (inst.size() + 1)), "alloc summation tile offsets"))) {
            free_nodes();
            return rc;
        }
        if ((rc = hip(hipMemset(nact, 0, nn), "clear active")) || (rc = hip(hipMemset(nfix, 0, nn), "clear fixed")) ||
            (rc = hip(hipMemset(nfric, 0, 8 * nn), "clear friction")) || (rc = hip(hipMemset(nm, 0, 8 * nn), "clear mass")))
            return rc;
        {
            std::vector<uint32_t> so(inst.size() + 1);
            for (size_t k = 0; k < inst.size(); k++) so[k] = inst[k].stile_off;
            so[inst.size()] = nstiles;
            if ((rc = hip(hipMemcpy(dstile_off, so.data(), 4 * so.size(), hipMemcpyHostToDevice),
                          "upload summation tile offsets")))
                return rc;
        }
        if ((rc = layout_tiles())) return rc;
        layout_set = true;
        keys_valid = false;
        table_dirty = true;
        return MPM_OK;
    }

    /* Uploads timelines and the instance table (the host mirror is authoritative). */
    int upload_table() {
        if (!table_dirty) return MPM_OK;
        int rc = hip(hipStreamSynchronize(stream), "table sync");
        if (rc) return rc;
        std::vector<double> tl;
        std::vector<Inst> t(inst.size());
        for (size_t k = 0; k < inst.size(); k++) {
            const HostInst &H = inst[k];
            Inst &I = t[k];
            std::memset(&I, 0, sizeof(I));
            I.g = H.g;
            I.node_off = H.node_off;
            I.cell_off = H.cell_off;
            I.part_off = H.part_off;
            I.npart = H.npart;
            I.tile_off = H.tile_off;
            I.tiles = H.tiles;
            I.stile_off = H.stile_off;
            I.stiles = H.stiles;
            I.dt = H.dt_max;
            I.damping = H.damping;
            I.mass_min = H.mass_min;
            I.time = H.time;
            I.tl_off = (int32_t)(tl.size() / 4);
            I.tl_count = (int32_t)(H.tl.size() / 4);
            I.has_fext = has_fext ? 1 : 0;
            tl.insert(tl.end(), H.tl.begin(), H.tl.end());
        }
        dfree(dtl);
        if (!tl.empty()) {
            if ((rc = hip(hipMalloc(&dtl, 8 * tl.size()), "alloc timelines"))) return rc;
            if ((rc = hip(hipMemcpy(dtl, tl.data(), 8 * tl.size(), hipMemcpyHostToDevice), "upload timelines"))) return rc;
        }
        if ((rc = hip(hipMemcpy(dinst, t.data(), sizeof(Inst) * t.size(), hipMemcpyHostToDevice), "upload table")))
            return rc;
        table_dirty = false;
        drop_graphs(); /* graphs hold the old timeline pointer */
        return MPM_OK;
    }

    /* Cell keys and counts for the current positions (after an upload). */
    int ensure_keys() {
        if (keys_valid) return MPM_OK;
        int rc = hip(hipMemsetAsync(bcount, 0, sizeof(uint32_t) * scan_padded_cells(ncells), stream), "clear cell counts");
        if (rc) return rc;
        rc = hip(launch_keys(batch(), pview(), bview(), stream), "cell keys");
        if (rc) return rc;
        keys_valid = true;
        return MPM_OK;
    }

    /* The kernels of one step (no host work).  With ev != nullptr an event is
     * recorded before the binning (ev[0]), before each later phase in launch
     * order (ev[1..4], see phase_order) and after the last (ev[PHASE_COUNT]). */
    int enqueue_step(hipEvent_t *ev = nullptr, hipEvent_t *mid = nullptr) {
        const Batch b = batch();
        int rc;
        if ((rc = hip(launch_prepare(b, stream), "step prepare"))) return rc;
        if (ev && (rc = hip(hipEventRecord(ev[0], stream), "event"))) return rc;
        if ((rc = hip(launch_bin(pview(), bview(), bblock, stream), "cell binning"))) return rc;
        const SurrogateArgs sa = sargs();
        int order_of[4];
        phase_order(order_of);
        for (int q = 0; q < 4; q++) {
            const int ph = order_of[q];
            if (ev && (rc = hip(hipEventRecord(ev[1 + q], stream), "event"))) return rc;
            if (ph == PHASE_G2P_STRAIN && sur && (rc = hip(launch_zero_u32(qcount, 1, stream), "queue reset")))
                return rc;
            if ((rc = hip(launch_phase(ph, b, pview(), nview(), bview(), parts, order, mt, d_health, sa, stream,
                                       mid ? mid[1 + q] : nullptr),
                          "step kernels")))
                return rc;
        }
        if (sur) {
            /* the network on the queued points, then acceptance / fallback */
            rc = hip(mlp_launch_forward(sur->params, sur->wbf16, qx, 1, n, qy, 1, n, n, stream, qcount), "surrogate forward");
            if (!rc) rc = hip(launch_surrogate_post(b, pview(), mt, d_health, sa, stream), "surrogate post");
            if (rc) return rc;
        }
        if (ev && (rc = hip(hipEventRecord(ev[PHASE_COUNT], stream), "event"))) return rc;
        return hip(launch_advance(b, stream), "step advance");
    }

    int begin_steps() {
        if (!layout_set) return fail(MPM_ERR_STATE, "step before set_grid / set_instances");
        if (strain == STRAIN_LARGE && (variant != VAR_DP || sur || cap))
            return fail(MPM_ERR_STATE, "the large-strain pass runs with the DoublePass loop and the exact return map only");
        int rc;
        if ((rc = upload_table())) return rc;
        if (sur && (rc = ensure_queue())) return rc;
        if ((rc = ensure_occ())) return rc;
        if ((order == ORDER_CELL || order == ORDER_PAIR) && (rc = ensure_parts())) return rc;
        if (order == ORDER_PARTICLE && (rc = ensure_list())) return rc;
        if (order == ORDER_TILE && (rc = ensure_stiles())) return rc;
        return ensure_keys();
    }

    /* Host mirror of the device clock after one step. */
    void advance_mirror(const double *dt0) {
        for (size_t k = 0; k < inst.size(); k++) inst[k].time += (k == 0 && dt0) ? *dt0 : inst[k].dt_max;
        cycle++;
        steps++;
    }

    /* One eager step; dt0 (when given) replaces instance 0's dt_max. */
    int step_once(const double *dt0, hipEvent_t *ev = nullptr, hipEvent_t *mid = nullptr) {
        int rc = begin_steps();
        if (rc) return rc;
        if ((rc = hip(launch_set_dt(batch(), 0, dt0 ? *dt0 : inst[0].dt_max, stream), "set dt"))) return rc;
        if ((rc = enqueue_step(ev, mid))) return rc;
        keys_valid = true; /* g2p_strain wrote the next keys and counts */
        advance_mirror(dt0);
        return MPM_OK;
    }

    /* A step taken phase by phase (mpm_step_phase): phase_step is the index,
     * in phase_order, of the next phase, or -1 when no such step is open. */
    int phase_step = -1;
    bool phase_dt_given = false;
    double phase_dt = 0.0;

    int step_phase(int phase, double dt) {
        int rc;
        if (phase_step < 0) {
            if (phase != PHASE_BIN) return fail(MPM_ERR_STATE, "a step taken phase by phase opens with the binning phase");
            if ((rc = begin_steps())) return rc;
            phase_dt_given = dt > 0.0;
            phase_dt = dt;
            if ((rc = hip(launch_set_dt(batch(), 0, phase_dt_given ? dt : inst[0].dt_max, stream), "set dt")) ||
                (rc = hip(launch_prepare(batch(), stream), "step prepare")) ||
                (rc = hip(launch_bin(pview(), bview(), bblock, stream), "cell binning")))
                return rc;
            phase_step = 0;
            return MPM_OK;
        }
        int order_of[4];
        phase_order(order_of);
        if (phase != order_of[phase_step])
            return fail(MPM_ERR_STATE, "phase %d out of order (the open step's next phase is %d)", phase,
                        order_of[phase_step]);
        /* a failed launch abandons the step */
        const int at = phase_step;
        phase_step = -1;
        const Batch b = batch();
        const SurrogateArgs sa = sargs();
        if (phase == PHASE_G2P_STRAIN && sur && (rc = hip(launch_zero_u32(qcount, 1, stream), "queue reset")))
            return rc;
        if ((rc = hip(launch_phase(phase, b, pview(), nview(), bview(), parts, order, mt, d_health, sa, stream),
                      "step kernels")))
            return rc;
        if (at + 1 < 4) {
            phase_step = at + 1;
            return MPM_OK;
        }
        if (sur) {
            rc = hip(mlp_launch_forward(sur->params, sur->wbf16, qx, 1, n, qy, 1, n, n, stream, qcount), "surrogate forward");
            if (!rc) rc = hip(launch_surrogate_post(b, pview(), mt, d_health, sa, stream), "surrogate post");
            if (rc) return rc;
        }
        if ((rc = hip(launch_advance(b, stream), "step advance"))) return rc;
        keys_valid = true;
        advance_mirror(phase_dt_given ? &phase_dt : nullptr);
        return MPM_OK;
    }

    /* Mean milliseconds of `reps` launches of one phase (or of the phase
     * pair phase, phase2 when phase2 >= 0). */
    int time_phase(int phase, int reps, const SurrogateArgs &sa, hipEvent_t e0, hipEvent_t e1, double &ms,
                   int phase2 = -1) {
        int rc;
        if ((rc = hip(hipEventRecord(e0, stream), "event"))) return rc;
        for (int r = 0; r < reps; r++) {
            if ((rc = hip(launch_phase(phase, batch(), pview(), nview(), bview(), parts, order, mt, d_health, sa, stream),
                          "autotune phase")))
                return rc;
            if (phase2 >= 0 &&
                (rc = hip(launch_phase(phase2, batch(), pview(), nview(), bview(), parts, order, mt, d_health, sa,
                                       stream), "autotune phase")))
                return rc;
        }
        if ((rc = hip(hipEventRecord(e1, stream), "event")) || (rc = hip(hipEventSynchronize(e1), "event sync")))
            return rc;
        float f = 0.f;
        if ((rc = hip(hipEventElapsedTime(&f, e0, e1), "event time"))) return rc;
        ms = f / reps;
        return MPM_OK;
    }
    int set_tile(const int shape[3]) {
        int rc = hip(hipStreamSynchronize(stream), "tile sync");
        if (rc) return rc;
        tile[0] = shape[0];
        tile[1] = shape[1];
        tile[2] = shape[2];
        drop_graphs();
        if ((rc = layout_tiles())) return rc;
        if ((rc = upload_table())) return rc;
        /* the table upload cleared vt: recompute it from the device clock */
        return hip(launch_prepare(batch(), stream), "step prepare");
    }
    /* One step in which the particle-to-grid phase, which only reads
     * particle and bin state and overwrites node state (its result does not
     * depend on the tile shape), is timed under every candidate node tile
     * shape; the fastest shape is kept and the step then completes exactly
     * as an ordinary one. */
    int autotune_step(int reps, double *best_ms) {
        int rc = begin_steps();
        if (rc) return rc;
        if ((rc = hip(launch_set_dt(batch(), 0, inst[0].dt_max, stream), "set dt"))) return rc;
        hipEvent_t e0, e1;
        if ((rc = hip(hipEventCreate(&e0), "event")) || (rc = hip(hipEventCreate(&e1), "event"))) return rc;
        const SurrogateArgs sa = sargs();
        std::vector<double> t_p2g(kNumTileShapes, 0.0);
        rc = hip(launch_prepare(batch(), stream), "step prepare");
        if (!rc) rc = hip(launch_bin(pview(), bview(), bblock, stream), "cell binning");
        for (int k = 0; k < kNumTileShapes && !rc; k++) {
            rc = set_tile(kTileShapes[k]);
            if (!rc) rc = time_phase(PHASE_P2G_GRID, reps, sa, e0, e1, t_p2g[k]);
        }
        int best = 0;
        for (int k = 1; k < kNumTileShapes; k++)
            if (t_p2g[k] < t_p2g[best]) best = k;
        if (!rc) rc = set_tile(kTileShapes[best]);
        /* once more with the chosen shape: the tile flags follow its numbering */
        double once = 0.0;
        if (!rc) rc = time_phase(PHASE_P2G_GRID, 1, sa, e0, e1, once);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (rc) return rc;
        int order_of[4];
        phase_order(order_of);
        for (int q = 1; q < 4; q++) {
            if (order_of[q] == PHASE_G2P_STRAIN && sur && (rc = hip(launch_zero_u32(qcount, 1, stream), "queue reset")))
                return rc;
            if ((rc = hip(launch_phase(order_of[q], batch(), pview(), nview(), bview(), parts, order, mt, d_health, sa,
                                       stream), "step kernels")))
                return rc;
        }
        if (sur) {
            rc = hip(mlp_launch_forward(sur->params, sur->wbf16, qx, 1, n, qy, 1, n, n, stream, qcount), "surrogate forward");
            if (!rc) rc = hip(launch_surrogate_post(batch(), pview(), mt, d_health, sa, stream), "surrogate post");
            if (rc) return rc;
        }
        if ((rc = hip(launch_advance(batch(), stream), "step advance"))) return rc;
        keys_valid = true;
        advance_mirror(nullptr);
        if (best_ms) *best_ms = t_p2g[best];
        return MPM_OK;
    }

    int step_graph(int nsteps) {
        int rc = begin_steps();
        if (rc) return rc;
        /* instance 0 back to its dt_max (outside any graph) */
        if ((rc = hip(launch_set_dt(batch(), 0, inst[0].dt_max, stream), "set dt"))) return rc;
        auto it = graphs.find(nsteps);
        if (it == graphs.end()) {
            if ((rc = hip(hipStreamSynchronize(stream), "pre-capture sync"))) return rc;
            hipGraph_t graph;
            if ((rc = hip(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "begin capture"))) return rc;
            for (int s = 0; s < nsteps && !rc; s++) rc = enqueue_step();
            hipError_t ce = hipStreamEndCapture(stream, &graph);
            if (rc) return rc;
            if ((rc = hip(ce, "end capture"))) return rc;
            hipGraphExec_t exec;
            rc = hip(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "graph instantiate");
            (void)hipGraphDestroy(graph);
            if (rc) return rc;
            it = graphs.emplace(nsteps, exec).first;
        }
        if ((rc = hip(hipGraphLaunch(it->second, stream), "graph launch"))) return rc;
        keys_valid = true;
        for (int s = 0; s < nsteps; s++) advance_mirror(nullptr);
        return MPM_OK;
    }

    /* reportConsole sums over monitor entries in [p0, p1) and fixed-y nodes in [n0, n1). */
    int monitors(int64_t p0, int64_t p1, int64_t n0, int64_t n1, double time, mpm_monitors *out) {
        std::memset(out, 0, sizeof(*out));
        out->time = time;
        std::vector<int64_t> lists[MPM_MON_COUNT + 1];
        for (int k = 0; k < MPM_MON_COUNT; k++)
            for (int64_t q : mon[k])
                if (q >= p0 && q < p1) lists[k].push_back(q);
        for (int64_t q : fixed_y)
            if (q >= n0 && q < n1) lists[MPM_MON_COUNT].push_back(q);
        int64_t total = 0;
        for (auto &l : lists) total += (int64_t)l.size();
        int rc;
        if ((rc = hip(hipStreamSynchronize(stream), "monitor sync"))) return rc;
        if (total > monbuf_len) {
            dfree(d_monbuf);
            if ((rc = hip(hipMalloc(&d_monbuf, 8 * (size_t)total), "alloc monitor lists"))) return rc;
            monbuf_len = total;
        }
        MonLists ml;
        int64_t off = 0;
        for (int k = 0; k <= MPM_MON_COUNT; k++) {
            if (!lists[k].empty() &&
                (rc = hip(hipMemcpy(d_monbuf + off, lists[k].data(), 8 * lists[k].size(), hipMemcpyHostToDevice),
                          "upload monitor list")))
                return rc;
            if (k < MPM_MON_COUNT) {
                ml.list[k] = d_monbuf + off;
                ml.count[k] = (int64_t)lists[k].size();
            } else {
                ml.fixed_y = d_monbuf + off;
                ml.nfy = (int64_t)lists[k].size();
            }
            off += (int64_t)lists[k].size();
        }
        if ((rc = hip(launch_monitors(pview(), nview(), nnodes, ml, d_mon_out, stream), "monitors"))) return rc;
        double v[8];
        if ((rc = hip(hipMemcpyAsync(v, d_mon_out, sizeof(v), hipMemcpyDeviceToHost, stream), "monitor copy")))
            return rc;
        if ((rc = hip(hipStreamSynchronize(stream), "monitor sync"))) return rc;
        out->has_position = !lists[MPM_MON_DISPLACEMENT].empty();
        out->has_momentum = !lists[MPM_MON_MOMENTUM].empty();
        out->has_energy = !lists[MPM_MON_ENERGY].empty();
        out->has_stress = !lists[MPM_MON_STRESS].empty();
        out->position_y = out->has_position ? v[0] : 0.0;
        out->force_y = v[1];
        out->momentum_y = out->has_momentum ? v[2] : 0.0;
        if (out->has_energy) {
            out->kinetic_energy = v[3];
            out->strain_energy = v[4];
            out->plastic_energy = v[5];
        }
        if (out->has_stress) {
            out->stress_y = v[6];
            out->strain_y = v[7];
        }
        return MPM_OK;
    }

    int alloc_particles(int64_t count) {
        (void)hipStreamSynchronize(stream);
        drop_graphs();
        free_particles();
        free_queue();
        free_parts();
        n = count;
        has_fext = false;
        keys_valid = false;
        table_dirty = true;
        if (n == 0) return MPM_OK;
        const size_t sn = (size_t)n;
        int rc;
        for (int f = 0; f < MPM_PF_COUNT; f++) {
            if ((rc = hip(hipMalloc(&pf[f], 8 * sn * kFieldComponents[f]), "alloc particle field"))) {
                free_particles();
                n = 0;
                return rc;
            }
            if ((rc = hip(hipMemset(pf[f], 0, 8 * sn * kFieldComponents[f]), "clear particle field"))) return rc;
        }
        if ((rc = hip(hipMalloc(&pmat, 4 * sn), "alloc material ids")) ||
            (rc = hip(hipMalloc(&pdc_rank, 4 * sn), "alloc control ranks")) ||
            (rc = hip(hipMalloc(&pdc_axes, sn), "alloc control axes")) ||
            (rc = hip(hipMalloc(&pdc_mult, 4 * sn), "alloc control multipliers")) ||
            (rc = hip(hipMalloc(&pde, 48 * sn), "alloc strain increments")) ||
            (rc = hip(hipMalloc(&pflag, sn), "alloc particle flags")) ||
            (rc = hip(hipMalloc(&bkey, 4 * sn), "alloc keys")) ||
            (rc = hip(hipMalloc(&bsorted, 4 * sn), "alloc sorted ids"))) {
            free_particles();
            n = 0;
            return rc;
        }
        /* defaults of MaterialPoint_BC (MaterialPoint_BC.h:31-82) */
        std::vector<double> F(9 * sn, 0.0);
        for (int d = 0; d < 3; d++) std::fill(F.begin() + (4 * d) * sn, F.begin() + (4 * d + 1) * sn, 1.0);
        std::vector<int32_t> minus1(sn, -1);
        std::vector<float> one(sn, 1.0f);
        if ((rc = hip(hipMemcpy(pf[MPM_PF_DEFORMATION_GRADIENT], F.data(), 8 * 9 * sn, hipMemcpyHostToDevice), "init F")) ||
            (rc = hip(hipMemcpy(pf[MPM_PF_DEFORMATION_GRADIENT_ELASTIC], F.data(), 8 * 9 * sn, hipMemcpyHostToDevice),
                      "init Fe")) ||
            (rc = hip(hipMemcpy(pmat, minus1.data(), 4 * sn, hipMemcpyHostToDevice), "init materials")) ||
            (rc = hip(hipMemcpy(pdc_rank, minus1.data(), 4 * sn, hipMemcpyHostToDevice), "init control ranks")) ||
            (rc = hip(hipMemset(pdc_axes, 0, sn), "init control axes")) ||
            (rc = hip(hipMemcpy(pdc_mult, one.data(), 4 * sn, hipMemcpyHostToDevice), "init multipliers")))
            return rc;
        return MPM_OK;
    }
};

#define CHECK_E(e)                                                                 \
    do {                                                                           \
        if (!(e)) return MPM_ERR_ARG;                                              \
        if (hipSetDevice((e)->device) != hipSuccess)                               \
            return (e)->fail(MPM_ERR_HIP, "hipSetDevice(%d) failed", (e)->device); \
    } while (0)

extern "C" {

int mpm_create(mpm_engine **out, int32_t device) {
    if (!out) return MPM_ERR_ARG;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        g_create_error = "no HIP device available";
        return MPM_ERR_HIP;
    }
    if (device < 0 || device >= count) {
        g_create_error = "device ordinal out of range";
        return MPM_ERR_ARG;
    }
    if (hipSetDevice(device) != hipSuccess) {
        g_create_error = "hipSetDevice failed";
        return MPM_ERR_HIP;
    }
    mpm_engine *e = new mpm_engine();
    e->device = device;
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc(&e->d_health, sizeof(unsigned long long) * HEALTH_WORDS) != hipSuccess ||
        hipMalloc(&e->d_mon_out, sizeof(double) * 8) != hipSuccess ||
        hipMalloc(&e->qcount, sizeof(uint32_t)) != hipSuccess ||
        hipMalloc(&e->capcount, sizeof(unsigned long long)) != hipSuccess ||
        hipMemset(e->capcount, 0, sizeof(unsigned long long)) != hipSuccess ||
        hipMemset(e->d_health, 0, sizeof(unsigned long long) * HEALTH_WORDS) != hipSuccess) {
        g_create_error = "stream or buffer creation failed";
        mpm_destroy(e);
        return MPM_ERR_HIP;
    }
    *out = e;
    return MPM_OK;
}

int mpm_destroy(mpm_engine *e) {
    if (!e) return MPM_OK;
    (void)hipSetDevice(e->device);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    e->drop_graphs();
    e->free_nodes();
    e->free_particles();
    e->free_queue();
    dfree(e->d_health);
    dfree(e->d_mon_out);
    dfree(e->d_monbuf);
    dfree(e->qcount);
    dfree(e->cap);
    dfree(e->capcount);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
    return MPM_OK;
}

const char *mpm_last_error(const mpm_engine *e) { return e ? e->err.c_str() : g_create_error.c_str(); }

int mpm_set_grid(mpm_engine *e, const double length[3], const int32_t cells[3]) {
    if (e && e->phase_step >= 0) return e->fail(MPM_ERR_STATE, "a step taken phase by phase is open");
    CHECK_E(e);
    if (!length || !cells) return e->fail(MPM_ERR_ARG, "null grid argument");
    HostInst I;
    if (!make_geom(length, cells, I.g)) return e->fail(MPM_ERR_ARG, "bad grid (cells >= 1, lengths > 0)");
    if (e->inst.size() == 1) {
        /* keep the scalar settings of a single-instance engine */
        I.time = e->inst[0].time;
        I.dt_max = e->inst[0].dt_max;
        I.time_end = e->inst[0].time_end;
        I.damping = e->inst[0].damping;
        I.mass_min = e->inst[0].mass_min;
        I.tl = e->inst[0].tl;
    } else if (e->inst.size() > 1) {
        /* leaving multi-instance mode: the particle arrays belonged to the batch */
        e->alloc_particles(0);
    }
    I.npart = e->n;
    e->inst.assign(1, I);
    dfree(e->dpinst);
    dfree(e->dwave);
    e->nwaves = 0;
    return e->build_layout();
}

int mpm_set_instances(mpm_engine *e, int32_t count, const double *lengths, const int32_t *cells,
                      const int64_t *particle_counts) {
    if (e && e->phase_step >= 0) return e->fail(MPM_ERR_STATE, "a step taken phase by phase is open");
    CHECK_E(e);
    if (count < 1 || !lengths || !cells || !particle_counts) return e->fail(MPM_ERR_ARG, "bad instance list");
    std::vector<HostInst> v(count);
    int64_t total = 0;
    for (int k = 0; k < count; k++) {
        if (!make_geom(lengths + 3 * k, cells + 3 * k, v[k].g)) return e->fail(MPM_ERR_ARG, "bad grid of instance %d", k);
        if (particle_counts[k] < 0) return e->fail(MPM_ERR_ARG, "negative particle count");
        v[k].part_off = total;
        v[k].npart = particle_counts[k];
        total += particle_counts[k];
    }
    if (total >= (int64_t)0xFFFFFFFFll) return e->fail(MPM_ERR_RANGE, "too many particles");
    e->inst = v;
    int rc = e->build_layout();
    if (rc) return rc;
    if ((rc = e->alloc_particles(total))) return rc;
    std::vector<uint32_t> pi((size_t)total);
    for (int k = 0; k < count; k++)
        std::fill(pi.begin() + v[k].part_off, pi.begin() + v[k].part_off + v[k].npart, (uint32_t)k);
    if (total > 0) {
        if ((rc = e->hip(hipMalloc(&e->dpinst, 4 * (size_t)total), "alloc particle instances"))) return rc;
        if ((rc = e->hip(hipMemcpy(e->dpinst, pi.data(), 4 * (size_t)total, hipMemcpyHostToDevice), "upload instances")))
            return rc;
    }
    /* the wave table: one (instance, first particle) entry per 64 consecutive
     * particles of an instance */
    std::vector<uint32_t> wt;
    for (int k = 0; k < count; k++)
        for (int64_t q = 0; q < v[k].npart; q += 64) {
            wt.push_back((uint32_t)k);
            wt.push_back((uint32_t)(v[k].part_off + q));
        }
    e->nwaves = (uint32_t)(wt.size() / 2);
    if (!wt.empty()) {
        if ((rc = e->hip(hipMalloc(&e->dwave, 4 * wt.size()), "alloc wave table"))) return rc;
        if ((rc = e->hip(hipMemcpy(e->dwave, wt.data(), 4 * wt.size(), hipMemcpyHostToDevice), "upload wave table")))
            return rc;
    }
    return MPM_OK;
}

int mpm_get_instance_layout(const mpm_engine *e, int32_t instance, int64_t *particle_offset, int64_t *node_offset,
                            int64_t *node_count) {
    if (!e || instance < 0 || instance >= (int32_t)e->inst.size()) return MPM_ERR_ARG;
    const HostInst &I = e->inst[instance];
    if (particle_offset) *particle_offset = I.part_off;
    if (node_offset) *node_offset = I.node_off;
    if (node_count) *node_count = I.g.nnodes;
    return MPM_OK;
}

int mpm_set_node_boundary(mpm_engine *e, const uint8_t *fixed_mask, const double *friction, int64_t n_nodes) {
    if (e && e->phase_step >= 0) return e->fail(MPM_ERR_STATE, "a step taken phase by phase is open");
    CHECK_E(e);
    if (!e->layout_set) return e->fail(MPM_ERR_STATE, "set_node_boundary before set_grid");
    if (n_nodes != e->nnodes) return e->fail(MPM_ERR_ARG, "node count %lld != %lld", (long long)n_nodes, (long long)e->nnodes);
    const size_t nn = (size_t)n_nodes;
    (void)hipStreamSynchronize(e->stream);
    int rc;
    e->fixed_y.clear();
    if (fixed_mask) {
        if ((rc = e->hip(hipMemcpy(e->nfix, fixed_mask, nn, hipMemcpyHostToDevice), "upload fixed"))) return rc;
        for (size_t i = 0; i < nn; i++)
            if (fixed_mask[i] & 2) e->fixed_y.push_back((int64_t)i);
    } else if ((rc = e->hip(hipMemset(e->nfix, 0, nn), "clear fixed"))) {
        return rc;
    }
    if (friction) {
        if ((rc = e->hip(hipMemcpy(e->nfric, friction, 8 * nn, hipMemcpyHostToDevice), "upload friction"))) return rc;
    } else if ((rc = e->hip(hipMemset(e->nfric, 0, 8 * nn), "clear friction"))) {
        return rc;
    }
    return MPM_OK;
}

int mpm_set_materials(mpm_engine *e, const mpm_material *materials, int32_t n) {
    if (e && e->phase_step >= 0) return e->fail(MPM_ERR_STATE, "a step taken phase by phase is open");
    CHECK_E(e);
    if (n < 0 || n > MPM_MAX_MATERIALS) return e->fail(MPM_ERR_RANGE, "material count %d outside [0, %d]", n, MPM_MAX_MATERIALS);
    if (n > 0 && !materials) return e->fail(MPM_ERR_ARG, "null material table");
    MatTable mt{};
    mt.n = n;
    for (int i = 0; i < n; i++) {
        mt.type[i] = materials[i].type;
        mt.E[i] = materials[i].elastic_modulus;
        mt.nu[i] = materials[i].poisson_ratio;
        mt.yield[i] = materials[i].yield_stress;
        mt.c0[i] = materials[i].hardening_c0;
        mt.c1[i] = materials[i].hardening_c1;
    }
    (void)hipStreamSynchronize(e->stream);
    e->mt = mt;
    e->drop_graphs(); /* graphs hold the material table by value */
    return MPM_OK;
}

int mpm_set_particle_count(mpm_engine *e, int64_t n) {
    if (e && e->phase_step >= 0) return e->fail(MPM_ERR_STATE, "a step taken phase by phase is open");
    CHECK_E(e);
    if (n < 0 || n >= (int64_t)0xFFFFFFFFll) return e->fail(MPM_ERR_RANGE, "particle count %lld unsupported", (long long)n);
    if (e->inst.size() > 1) {
        if (n != e->n) return e->fail(MPM_ERR_ARG, "a multi-instance engine's particle count is set by set_instances");
        return MPM_OK;
    }
    int rc = e->alloc_particles(n);
    if (!rc && !e->inst.empty()) {
        e->inst[0].part_off = 0;
        e->inst[0].npart = n;
    }
    return rc;
}

int mpm_get_particle_count(const mpm_engine *e, int64_t *n) {
    if (!e || !n) return MPM_ERR_ARG;
    *n = e->n;
    return MPM_OK;
}

int mpm_set_particle_field(mpm_engine *e, int32_t field, const double *src, int64_t n) {
    if (e && e->phase_step >= 0) return e->fail(MPM_ERR_STATE, "a step taken phase by phase is open");
    CHECK_E(e);
    if (field < 0 || field >= MPM_PF_COUNT) return e->fail(MPM_ERR_ARG, "unknown particle field %d", field);
    if (n != e->n) return e->fail(MPM_ERR_ARG, "field length %lld != particle count %lld", (long long)n, (long long)e->n);
    if (n == 0) return MPM_OK;
    if (!src) return e->fail(MPM_ERR_ARG, "null source");
    (void)hipStreamSynchronize(e->stream);
    size_t bytes = 8 * (size_t)n * kFieldComponents[field];
    int rc = e->hip(hipMemcpy(e->pf[field], src, bytes, hipMemcpyHostToDevice), "upload particle field");
    if (rc) return rc;
    if (field == MPM_PF_POSITION) e->keys_valid = false;
    if (field == MPM_PF_FORCE_EXTERNAL) {
        /* N * 0 added to a sum never changes it, so an all-zero field is skipped */
        bool any = false;
        for (size_t i = 0; i < 3 * (size_t)n && !any; i++) any = (src[i] != 0.0);
        if (any != e->has_fext) e->table_dirty = true;
        e->has_fext = any;
    }
    return MPM_OK;
}

int mpm_get_particle_field(mpm_engine *e, int32_t field, double *dst, int64_t n) {
    CHECK_E(e);
    if (field < 0 || field >= MPM_PF_COUNT) return e->fail(MPM_ERR_ARG, "unknown particle field %d", field);
    if (n != e->n) return e->fail(MPM_ERR_ARG, "field length %lld != particle count %lld", (long long)n, (long long)e->n);
    if (n == 0) return MPM_OK;
    if (!dst) return e->fail(MPM_ERR_ARG, "null destination");
    int rc = e->hip(hipStreamSynchronize(e->stream), "sync before download");
    if (rc) return rc;
    return e->hip(hipMemcpy(dst, e->pf[field], 8 * (size_t)n * kFieldComponents[field], hipMemcpyDeviceToHost),
                  "download particle field");
}

int mpm_set_particle_material(mpm_engine *e, const int32_t *material, int64_t n) {
    if (e && e->phase_step >= 0) return e->fail(MPM_ERR_STATE, "a step taken phase by phase is open");
    CHECK_E(e);
    if (n != e->n) return e->fail(MPM_ERR_ARG, "material length mismatch");
    if (n == 0) return MPM_OK;
    if (!material) return e->fail(MPM_ERR_ARG, "null material ids");
    (void)hipStreamSynchronize(e->stream);
    return e->hip(hipMemcpy(e->pmat, material, 4 * (size_t)n, hipMemcpyHostToDevice), "upload material ids");
}

int mpm_set_displacement_control(mpm_engine *e, const int64_t *index, const uint8_t *axes, const float *multiplier,
                                 int64_t count) {
    if (e && e->phase_step >= 0) return e->fail(MPM_ERR_STATE, "a step taken phase by phase is open");
    CHECK_E(e);
    if (count < 0 || (count > 0 && (!index || !axes || !multiplier))) return e->fail(MPM_ERR_ARG, "bad control list");
    if (count >= (int64_t)0x7FFFFFFF) return e->fail(MPM_ERR_RANGE, "control list too long");
    const size_t sn = (size_t)e->n;
    std::vector<int32_t> rank(sn, -1);
    std::vector<uint8_t> ax(sn, 0);
    std::vector<float> mult(sn, 1.0f);
    for (int64_t k = 0; k < count; k++) {
        if (index[k] < 0 || index[k] >= e->n) return e->fail(MPM_ERR_ARG, "control index %lld out of range", (long long)index[k]);
        rank[index[k]] = (int32_t)k; /* a later list position wins */
        ax[index[k]] = axes[k] & 7;
        mult[index[k]] = multiplier[k];
    }
    if (sn == 0) return MPM_OK;
    (void)hipStreamSynchronize(e->stream);
    int rc;
    if ((rc = e->hip(hipMemcpy(e->pdc_rank, rank.data(), 4 * sn, hipMemcpyHostToDevice), "upload control ranks")) ||
        (rc = e->hip(hipMemcpy(e->pdc_axes, ax.data(), sn, hipMemcpyHostToDevice), "upload control axes")) ||
        (rc = e->hip(hipMemcpy(e->pdc_mult, mult.data(), 4 * sn, hipMemcpyHostToDevice), "upload multipliers")))
        return rc;
    return MPM_OK;
}

int mpm_set_monitor_list(mpm_engine *e, int32_t kind, const int64_t *index, int64_t count) {
    if (e && e->phase_step >= 0) return e->fail(MPM_ERR_STATE, "a step taken phase by phase is open");
    CHECK_E(e);
    if (kind < 0 || kind >= MPM_MON_COUNT) return e->fail(MPM_ERR_ARG, "unknown monitor kind %d", kind);
    if (count < 0 || (count > 0 && !index)) return e->fail(MPM_ERR_ARG, "bad monitor list");
    for (int64_t k = 0; k < count; k++)
        if (index[k] < 0 || index[k] >= e->n) return e->fail(MPM_ERR_ARG, "monitor index out of range");
    e->mon[kind].assign(index, index + count);
    return MPM_OK;
}

int mpm_set_instance_timeline(mpm_engine *e, int32_t instance, const double *time, const double *v3, int32_t n) {
    if (e && e->phase_step >= 0) return e->fail(MPM_ERR_STATE, "a step taken phase by phase is open");
    if (!e) return MPM_ERR_ARG;
    if (instance < 0 || instance >= (int32_t)e->inst.size()) return e->fail(MPM_ERR_ARG, "bad instance");
    if (n < 0 || (n > 0 && (!time || !v3))) return e->fail(MPM_ERR_ARG, "bad timeline");
    std::vector<double> tl(4 * (size_t)n);
    for (int i = 0; i < n; i++) {
        tl[4 * i] = time[i];
        tl[4 * i + 1] = v3[3 * i];
        tl[4 * i + 2] = v3[3 * i + 1];
        tl[4 * i + 3] = v3[3 * i + 2];
    }
    e->inst[instance].tl = tl;
    e->table_dirty = true;
    return MPM_OK;
}

int mpm_set_timeline(mpm_engine *e, const double *time, const double *v3, int32_t n) {
    if (e && e->phase_step >= 0) return e->fail(MPM_ERR_STATE, "a step taken phase by phase is open");
    if (!e) return MPM_ERR_ARG;
    if (e->inst.empty()) return e->fail(MPM_ERR_STATE, "set_timeline before set_grid");
    for (int32_t k = 0; k < (int32_t)e->inst.size(); k++) {
        int rc = mpm_set_instance_timeline(e, k, time, v3, n);
        if (rc) return rc;
    }
    return MPM_OK;
}

int mpm_set_instance_time(mpm_engine *e, int32_t instance, double time, double dt_max, double time_end) {
    if (e && e->phase_step >= 0) return e->fail(MPM_ERR_STATE, "a step taken phase by phase is open");
    if (!e) return MPM_ERR_ARG;
    if (instance < 0 || instance >= (int32_t)e->inst.size()) return e->fail(MPM_ERR_ARG, "bad instance");
    e->inst[instance].time = time;
    e->inst[instance].dt_max = dt_max;
    e->inst[instance].time_end = time_end;
    e->table_dirty = true;
    return MPM_OK;
}

int mpm_set_time(mpm_engine *e, double time, double dt_max, double time_end) {
    if (e && e->phase_step >= 0) return e->fail(MPM_ERR_STATE, "a step taken phase by phase is open");
    if (!e) return MPM_ERR_ARG;
    if (e->inst.empty()) return e->fail(MPM_ERR_STATE, "set_time before set_grid");
    for (auto &I : e->inst) {
        I.time = time;
        I.dt_max = dt_max;
        I.time_end = time_end;
    }
    e->table_dirty = true;
    return MPM_OK;
}

int mpm_set_damping(mpm_engine *e, double damping, double mass_minimum) {
    if (e && e->phase_step >= 0) return e->fail(MPM_ERR_STATE, "a step taken phase by phase is open");
    if (!e) return MPM_ERR_ARG;
    if (e->inst.empty()) return e->fail(MPM_ERR_STATE, "set_damping before set_grid");
    for (auto &I : e->inst) {
        I.damping = damping;
        I.mass_min = mass_minimum;
    }
    e->table_dirty = true;
    return MPM_OK;
}

int mpm_step(mpm_engine *e, double dt, int32_t nsteps) {
    if (e && e->phase_step >= 0) return e->fail(MPM_ERR_STATE, "a step taken phase by phase is open");
    CHECK_E(e);
    if (nsteps < 0) return e->fail(MPM_ERR_ARG, "negative step count");
    if (e->inst.size() > 1 && dt > 0.0)
        return e->fail(MPM_ERR_ARG, "a multi-instance engine steps each instance by its own dt_max (pass dt <= 0)");
    for (int s = 0; s < nsteps; s++) {
        int rc = e->step_once(dt > 0.0 ? &dt : nullptr);
        if (rc) return rc;
    }
    return MPM_OK;
}

int mpm_set_variant(mpm_engine *e, int32_t variant) {
    if (e && e->phase_step >= 0) return e->fail(MPM_ERR_STATE, "a step taken phase by phase is open");
    CHECK_E(e);
    if (variant < MPM_VARIANT_DOUBLEPASS || variant > MPM_VARIANT_SINGLEPASS)
        return e->fail(MPM_ERR_ARG, "unknown variant %d", variant);
    static_assert(MPM_VARIANT_DOUBLEPASS == VAR_DP && MPM_VARIANT_DOUBLEPASS_CONTACT == VAR_DPC &&
                      MPM_VARIANT_SINGLEPASS == VAR_SP, "variant codes disagree");
    (void)hipStreamSynchronize(e->stream);
    e->drop_graphs();
    e->variant = variant;
    return MPM_OK;
}

int mpm_set_strain_formulation(mpm_engine *e, int32_t formulation) {
    if (e && e->phase_step >= 0) return e->fail(MPM_ERR_STATE, "a step taken phase by phase is open");
    CHECK_E(e);
    if (formulation != MPM_STRAIN_SMALL && formulation != MPM_STRAIN_LARGE)
        return e->fail(MPM_ERR_ARG, "unknown strain formulation %d", formulation);
    static_assert(MPM_STRAIN_SMALL == STRAIN_SMALL && MPM_STRAIN_LARGE == STRAIN_LARGE, "strain codes disagree");
    if (formulation != e->strain) {
        (void)hipStreamSynchronize(e->stream);
        e->drop_graphs();
    }
    e->strain = formulation;
    return MPM_OK;
}

int mpm_get_strain_formulation(const mpm_engine *e, int32_t *formulation) {
    if (!e || !formulation) return MPM_ERR_ARG;
    *formulation = e->strain;
    return MPM_OK;
}

int mpm_get_variant(const mpm_engine *e, int32_t *variant) {
    if (!e || !variant) return MPM_ERR_ARG;
    *variant = e->variant;
    return MPM_OK;
}

int mpm_set_summation_order(mpm_engine *e, int32_t order) {
    if (e && e->phase_step >= 0) return e->fail(MPM_ERR_STATE, "a step taken phase by phase is open");
    CHECK_E(e);
    if (order != MPM_ORDER_PARTICLE && order != MPM_ORDER_CELL && order != MPM_ORDER_TILE && order != MPM_ORDER_PAIR)
        return e->fail(MPM_ERR_ARG, "unknown order %d", order);
    (void)hipStreamSynchronize(e->stream);
    e->drop_graphs();
    e->order = order == MPM_ORDER_CELL   ? ORDER_CELL
               : order == MPM_ORDER_TILE ? ORDER_TILE
               : order == MPM_ORDER_PAIR ? ORDER_PAIR
                                         : ORDER_PARTICLE;
    return MPM_OK;
}

int mpm_set_tile_shape(mpm_engine *e, const int32_t shape[3]) {
    if (e && e->phase_step >= 0) return e->fail(MPM_ERR_STATE, "a step taken phase by phase is open");
    CHECK_E(e);
    if (!shape || shape[0] < 1 || shape[1] < 1 || shape[2] < 1 || shape[0] * shape[1] * shape[2] != NODE_BLOCK)
        return e->fail(MPM_ERR_ARG, "tile shape must multiply to %d", NODE_BLOCK);
    if (!e->layout_set) return e->fail(MPM_ERR_STATE, "no grid");
    int s[3] = {shape[0], shape[1], shape[2]};
    return e->set_tile(s);
}

int mpm_get_tile_shape(const mpm_engine *e, int32_t shape[3]) {
    if (!e || !shape) return MPM_ERR_ARG;
    for (int a = 0; a < 3; a++) shape[a] = e->tile[a];
    return MPM_OK;
}

int mpm_autotune(mpm_engine *e, int32_t reps, int32_t shape[3], double *node_ms) {
    if (e && e->phase_step >= 0) return e->fail(MPM_ERR_STATE, "a step taken phase by phase is open");
    CHECK_E(e);
    if (reps < 1 || reps > 100) return e->fail(MPM_ERR_ARG, "reps outside [1, 100]");
    int rc = e->autotune_step(reps, node_ms);
    if (!rc && shape)
        for (int a = 0; a < 3; a++) shape[a] = e->tile[a];
    return rc;
}

int mpm_step_graph(mpm_engine *e, int32_t nsteps) {
    if (e && e->phase_step >= 0) return e->fail(MPM_ERR_STATE, "a step taken phase by phase is open");
    CHECK_E(e);
    if (nsteps < 1 || nsteps > 100000) return e->fail(MPM_ERR_ARG, "graph step count out of range");
    return e->step_graph(nsteps);
}

int mpm_step_timed(mpm_engine *e, double dt, int32_t nsteps, double phase_ms[MPM_PHASE_COUNT]) {
    if (e && e->phase_step >= 0) return e->fail(MPM_ERR_STATE, "a step taken phase by phase is open");
    CHECK_E(e);
    if (nsteps < 0 || !phase_ms) return e->fail(MPM_ERR_ARG, "bad timed-step arguments");
    if (e->inst.size() > 1 && dt > 0.0) return e->fail(MPM_ERR_ARG, "multi-instance engines take dt <= 0");
    static_assert(MPM_PHASE_COUNT == PHASE_COUNT, "phase enums disagree");
    for (int k = 0; k < MPM_PHASE_COUNT; k++) phase_ms[k] = 0.0;
    hipEvent_t ev[PHASE_COUNT + 1] = {};
    int rc = MPM_OK;
    for (int k = 0; k <= PHASE_COUNT && !rc; k++) rc = e->hip(hipEventCreate(&ev[k]), "event create");
    for (int s = 0; s < nsteps && !rc; s++) {
        rc = e->step_once(dt > 0.0 ? &dt : nullptr, ev);
        if (!rc) rc = e->hip(hipEventSynchronize(ev[PHASE_COUNT]), "event sync");
        int order_of[4];
        e->phase_order(order_of);
        for (int k = 0; k < PHASE_COUNT && !rc; k++) {
            float ms = 0.f;
            rc = e->hip(hipEventElapsedTime(&ms, ev[k], ev[k + 1]), "event time");
            phase_ms[k == 0 ? PHASE_BIN : order_of[k - 1]] += ms;
        }
    }
    for (int k = 0; k <= PHASE_COUNT; k++)
        if (ev[k]) (void)hipEventDestroy(ev[k]);
    return rc;
}

int mpm_step_timed_kernels(mpm_engine *e, double dt, int32_t nsteps, double kernel_ms[MPM_KSLOT_COUNT]) {
    if (e && e->phase_step >= 0) return e->fail(MPM_ERR_STATE, "a step taken phase by phase is open");
    CHECK_E(e);
    if (nsteps < 0 || !kernel_ms) return e->fail(MPM_ERR_ARG, "bad timed-step arguments");
    if (e->inst.size() > 1 && dt > 0.0) return e->fail(MPM_ERR_ARG, "multi-instance engines take dt <= 0");
    for (int k = 0; k < MPM_KSLOT_COUNT; k++) kernel_ms[k] = 0.0;
    hipEvent_t ev[PHASE_COUNT + 1] = {}, mid[PHASE_COUNT + 1] = {};
    int rc = MPM_OK;
    for (int k = 0; k <= PHASE_COUNT && !rc; k++) rc = e->hip(hipEventCreate(&ev[k]), "event create");
    for (int k = 1; k < PHASE_COUNT && !rc; k++) rc = e->hip(hipEventCreate(&mid[k]), "event create");
    /* first kernel slot of each phase; the phase's second kernel takes the next slot */
    const int first[PHASE_COUNT] = {MPM_KSLOT_BIN, MPM_KSLOT_P2G_PARTIALS, MPM_KSLOT_G2P_VELOCITY,
                                    MPM_KSLOT_REMAP_PARTIALS, MPM_KSLOT_G2P_KINEMATICS};
    for (int s = 0; s < nsteps && !rc; s++) {
        rc = e->step_once(dt > 0.0 ? &dt : nullptr, ev, mid);
        if (!rc) rc = e->hip(hipEventSynchronize(ev[PHASE_COUNT]), "event sync");
        int order_of[4];
        e->phase_order(order_of);
        float ms = 0.f;
        if (!rc) rc = e->hip(hipEventElapsedTime(&ms, ev[0], ev[1]), "event time");
        kernel_ms[MPM_KSLOT_BIN] += ms;
        for (int q = 0; q < 4 && !rc; q++) {
            const int ph = order_of[q];
            float a = 0.f, b = 0.f;
            rc = e->hip(hipEventElapsedTime(&a, ev[1 + q], mid[1 + q]), "event time");
            if (!rc) rc = e->hip(hipEventElapsedTime(&b, mid[1 + q], ev[2 + q]), "event time");
            if (ph == PHASE_G2P_VELOCITY) {
                kernel_ms[first[ph]] += a + b;
            } else {
                kernel_ms[first[ph]] += a;
                kernel_ms[first[ph] + 1] += b;
            }
        }
    }
    for (int k = 0; k <= PHASE_COUNT; k++) {
        if (ev[k]) (void)hipEventDestroy(ev[k]);
        if (mid[k]) (void)hipEventDestroy(mid[k]);
    }
    return rc;
}

int mpm_step_phase(mpm_engine *e, int32_t phase, double dt) {
    CHECK_E(e);
    if (phase < 0 || phase >= MPM_PHASE_COUNT) return e->fail(MPM_ERR_ARG, "unknown phase %d", phase);
    if (e->phase_step < 0 && e->inst.size() > 1 && dt > 0.0)
        return e->fail(MPM_ERR_ARG, "a multi-instance engine steps each instance by its own dt_max (pass dt <= 0)");
    return e->step_phase(phase, dt);
}

int mpm_run(mpm_engine *e, double dt_total, int32_t *finished) {
    if (e && e->phase_step >= 0) return e->fail(MPM_ERR_STATE, "a step taken phase by phase is open");
    CHECK_E(e);
    if (e->inst.size() != 1) return e->fail(MPM_ERR_STATE, "mpm_run drives a single-instance engine");
    double acc = 0.0;
    while (acc < dt_total) {
        double dt = e->inst[0].dt_max;
        if (e->inst[0].dt_max > (dt_total - acc)) dt = dt_total - acc;
        acc += dt;
        int rc = e->step_once(&dt);
        if (rc) return rc;
    }
    if (finished) *finished = (e->inst[0].time < e->inst[0].time_end) ? 0 : 1;
    return MPM_OK;
}

int mpm_synchronize(mpm_engine *e) {
    CHECK_E(e);
    return e->hip(hipStreamSynchronize(e->stream), "stream sync");
}

int mpm_get_time(const mpm_engine *e, double *time, int64_t *cycle) {
    if (!e) return MPM_ERR_ARG;
    if (time) *time = e->inst.empty() ? 0.0 : e->inst[0].time;
    if (cycle) *cycle = e->cycle;
    return MPM_OK;
}

int mpm_get_instance_time(const mpm_engine *e, int32_t instance, double *time) {
    if (!e || !time || instance < 0 || instance >= (int32_t)e->inst.size()) return MPM_ERR_ARG;
    *time = e->inst[instance].time;
    return MPM_OK;
}

int mpm_get_node_count(const mpm_engine *e, int64_t *n) {
    if (!e || !n) return MPM_ERR_ARG;
    *n = e->layout_set ? e->nnodes : 0;
    return MPM_OK;
}

int mpm_get_node_field(mpm_engine *e, int32_t field, double *dst, int64_t n) {
    CHECK_E(e);
    if (!e->layout_set) return e->fail(MPM_ERR_STATE, "no grid");
    if (field < 0 || field >= MPM_NF_COUNT) return e->fail(MPM_ERR_ARG, "unknown node field %d", field);
    if (n != e->nnodes || !dst) return e->fail(MPM_ERR_ARG, "node field length mismatch");
    const int comps = (field == MPM_NF_MASS || field == MPM_NF_ACTIVE) ? 1 : 3;
    double *tmp = nullptr;
    size_t bytes = 8 * (size_t)n * comps;
    int rc = e->hip(hipMalloc(&tmp, bytes), "alloc node export");
    if (rc) return rc;
    rc = e->hip(launch_node_export(e->nview(), field, n, tmp, e->stream), "node export");
    if (!rc) rc = e->hip(hipStreamSynchronize(e->stream), "node export sync");
    if (!rc) rc = e->hip(hipMemcpy(dst, tmp, bytes, hipMemcpyDeviceToHost), "download node field");
    (void)hipFree(tmp);
    return rc;
}

int mpm_get_monitors(mpm_engine *e, mpm_monitors *out) {
    CHECK_E(e);
    if (!out) return e->fail(MPM_ERR_ARG, "null monitor output");
    if (!e->layout_set) return e->fail(MPM_ERR_STATE, "no grid");
    return e->monitors(0, e->n, 0, e->nnodes, e->inst[0].time, out);
}

int mpm_get_instance_monitors(mpm_engine *e, int32_t instance, mpm_monitors *out) {
    CHECK_E(e);
    if (!out || instance < 0 || instance >= (int32_t)e->inst.size()) return e->fail(MPM_ERR_ARG, "bad instance");
    const HostInst &I = e->inst[instance];
    return e->monitors(I.part_off, I.part_off + I.npart, I.node_off, I.node_off + I.g.nnodes, I.time, out);
}

int mpm_get_health(mpm_engine *e, mpm_health *out) {
    CHECK_E(e);
    if (!out) return e->fail(MPM_ERR_ARG, "null health output");
    std::vector<unsigned long long> w(HEALTH_WORDS);
    int rc = e->hip(hipStreamSynchronize(e->stream), "health sync");
    if (!rc) rc = e->hip(hipMemcpy(w.data(), e->d_health, sizeof(unsigned long long) * HEALTH_WORDS,
                                   hipMemcpyDeviceToHost), "health copy");
    if (rc) return rc;
    unsigned long long h[HEALTH_SLOTS] = {};
    for (int s = 0; s < HEALTH_STRIPES; s++)
        for (int k = 0; k < HEALTH_SLOTS; k++) h[k] += w[(size_t)s * HEALTH_STRIDE + k];
    out->steps = e->steps;
    out->outside_grid = (int64_t)h[HEALTH_OUTSIDE];
    out->null_material = (int64_t)h[HEALTH_NULL];
    out->plastic_points = (int64_t)h[HEALTH_PLASTIC];
    out->nonfinite_particles = (int64_t)h[HEALTH_NONFINITE];
    out->surrogate_calls = (int64_t)h[HEALTH_SURROGATE_CALLS];
    out->surrogate_fallbacks = (int64_t)h[HEALTH_SURROGATE_FALLBACKS];
    return MPM_OK;
}

int mpm_reset_health(mpm_engine *e) {
    CHECK_E(e);
    e->steps = 0;
    int rc = e->hip(hipStreamSynchronize(e->stream), "health sync");
    if (rc) return rc;
    return e->hip(hipMemset(e->d_health, 0, sizeof(unsigned long long) * HEALTH_WORDS), "health reset");
}

int mpm_set_surrogate(mpm_engine *e, mpm_surrogate *s, int32_t material, double residual_tolerance) {
    if (e && e->phase_step >= 0) return e->fail(MPM_ERR_STATE, "a step taken phase by phase is open");
    CHECK_E(e);
    if (s && s->device != e->device) return e->fail(MPM_ERR_ARG, "surrogate lives on another device");
    if (s && (material < 0 || material >= e->mt.n)) return e->fail(MPM_ERR_ARG, "surrogate material out of range");
    if (s && !(residual_tolerance >= 0.0)) return e->fail(MPM_ERR_ARG, "negative residual tolerance");
    (void)hipStreamSynchronize(e->stream);
    if (s) (void)hipStreamSynchronize(s->stream);
    e->drop_graphs();
    e->sur = s;
    e->sur_material = s ? material : -1;
    e->sur_tol = residual_tolerance;
    if (!s) e->free_queue();
    return MPM_OK;
}

int mpm_set_sample_capture(mpm_engine *e, int64_t capacity) {
    if (e && e->phase_step >= 0) return e->fail(MPM_ERR_STATE, "a step taken phase by phase is open");
    CHECK_E(e);
    if (capacity < 0) return e->fail(MPM_ERR_ARG, "negative capture capacity");
    (void)hipStreamSynchronize(e->stream);
    e->drop_graphs();
    dfree(e->cap);
    e->capcap = 0;
    int rc = e->hip(hipMemset(e->capcount, 0, sizeof(unsigned long long)), "capture reset");
    if (rc || capacity == 0) return rc;
    if ((rc = e->hip(hipMalloc(&e->cap, sizeof(double) * 14 * (size_t)capacity), "alloc capture"))) return rc;
    e->capcap = capacity;
    return MPM_OK;
}

int mpm_get_captured_samples(mpm_engine *e, double *rows, int64_t capacity, int64_t *count) {
    CHECK_E(e);
    if (!count) return e->fail(MPM_ERR_ARG, "null count");
    unsigned long long c = 0;
    int rc = e->hip(hipStreamSynchronize(e->stream), "capture sync");
    if (!rc) rc = e->hip(hipMemcpy(&c, e->capcount, sizeof(c), hipMemcpyDeviceToHost), "capture count");
    if (rc) return rc;
    int64_t have = std::min<int64_t>((int64_t)c, e->capcap);
    *count = (int64_t)c;
    if (rows && capacity > 0 && have > 0)
        rc = e->hip(hipMemcpy(rows, e->cap, sizeof(double) * 14 * (size_t)std::min(have, capacity), hipMemcpyDeviceToHost),
                    "capture copy");
    return rc;
}

int mpm_constitutive_batch(int32_t device, const mpm_material *material, const double *backstress, const double *stress,
                           const double *dstrain, double *dstress, double *dstrain_plastic, double *dbackstress,
                           int64_t n) {
    if (!material || n < 0) return MPM_ERR_ARG;
    if (n == 0) return MPM_OK;
    if (!backstress || !stress || !dstrain || !dstress || !dstrain_plastic || !dbackstress) return MPM_ERR_ARG;
    if (hipSetDevice(device) != hipSuccess) return MPM_ERR_HIP;
    MatTable mt{};
    mt.n = 1;
    mt.type[0] = material->type;
    mt.E[0] = material->elastic_modulus;
    mt.nu[0] = material->poisson_ratio;
    mt.yield[0] = material->yield_stress;
    mt.c0[0] = material->hardening_c0;
    mt.c1[0] = material->hardening_c1;
    const size_t sn = (size_t)n;
    double *buf = nullptr;
    if (hipMalloc(&buf, 8 * sn * 26) != hipSuccess) return MPM_ERR_NOMEM;
    double *R = buf, *s = buf + sn, *d = s + 6 * sn, *ds = d + 6 * sn, *dp = ds + 6 * sn, *dR = dp + 6 * sn;
    int rc = MPM_OK;
    if (hipMemcpy(R, backstress, 8 * sn, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(s, stress, 48 * sn, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d, dstrain, 48 * sn, hipMemcpyHostToDevice) != hipSuccess ||
        launch_constitutive(mt, R, s, d, ds, dp, dR, n, 0) != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(dstress, ds, 48 * sn, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(dstrain_plastic, dp, 48 * sn, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(dbackstress, dR, 8 * sn, hipMemcpyDeviceToHost) != hipSuccess)
        rc = MPM_ERR_HIP;
    (void)hipFree(buf);
    return rc;
}

}  // extern "C"
