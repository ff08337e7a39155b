/*
 * mpm_engine.hip — host side of the C ABI declared in include/mpm.h.
 *
 * The engine owns every device buffer (particles, nodes, cell bins) and one
 * HIP stream; the reference's PhysicsEngine owns the same objects as heap
 * AoS (PhysicsEngine.h:82-117).  Scalars that steer a step (time, dt, the
 * timeline velocity) stay on the host and reach the kernels as arguments.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mpm.h"
#include "mpm_kernels.h"
#include "mpm_surrogate.h"

namespace {

const int kFieldComponents[MPM_PF_COUNT] = {3, 3, 3, 1, 1, 1, 9, 6, 6, 6, 1, 1, 1, 1, 1};

std::string g_create_error = "no error";

/* TimeLine (TimeLine.h:27-46): piecewise-linear velocity; the last interval
 * [t_i, t_{i+1}) containing t wins; zero outside every interval. */
struct HostTimeline {
    std::vector<double> t;
    std::vector<double> v;  /* 3 per point */
    void velocity(double time, double out[3]) const {
        out[0] = out[1] = out[2] = 0.0;
        for (size_t i = 0; i + 1 < t.size(); i++) {
            if (t[i] <= time && time < t[i + 1]) {
                double s = (time - t[i]) / (t[i + 1] - t[i]);
                for (int a = 0; a < 3; a++) out[a] = (1.0 - s) * v[3 * i + a] + (s)*v[3 * (i + 1) + a];
            }
        }
    }
};

template <typename T>
void dfree(T *&p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

}  // namespace

struct mpm_engine {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err = "no error";

    bool grid_set = false;
    GridGeom g{};

    /* nodes */
    double *nm = nullptr, *np = nullptr, *nv = nullptr, *nf = nullptr, *na = nullptr, *nu = nullptr,
           *nfty = nullptr, *nfric = nullptr;
    uint8_t *nact = nullptr, *nfix = nullptr;
    int64_t *d_fixed_y = nullptr;
    int64_t nfy = 0;

    /* particles */
    int64_t n = 0;
    double *pf[MPM_PF_COUNT] = {};
    int32_t *pmat = nullptr, *pdc_rank = nullptr;
    uint8_t *pdc_axes = nullptr;
    float *pdc_mult = nullptr;
    bool has_fext = false;

    /* bins */
    uint32_t *bkey = nullptr, *bcount = nullptr, *bstart = nullptr, *bsorted = nullptr, *bblock = nullptr;
    bool keys_valid = false;

    /* materials, lists, control */
    MatTable mt{};
    int64_t *d_mon[MPM_MON_COUNT] = {};
    int64_t mon_count[MPM_MON_COUNT] = {};
    HostTimeline timeline;
    double time = 0.0, dt_max = 1.0e-3, time_end = 10.0;
    double damping = 0.0, mass_min = 0.0;
    int64_t cycle = 0;

    unsigned long long *d_health = nullptr;
    double *d_mon_out = nullptr;
    int64_t steps = 0;

    /* surrogate and sample capture */
    mpm_surrogate *sur = nullptr;
    int sur_material = -1;
    double sur_tol = 1.0e-4;
    uint32_t *qcount = nullptr, *qid = nullptr;
    double *qde = nullptr;
    float *qx = nullptr, *qy = nullptr;
    uint8_t *qenv = nullptr;
    int64_t qcap = 0;
    double *cap = nullptr;
    unsigned long long *capcount = nullptr;
    int64_t capcap = 0;

    void free_queue() {
        dfree(qid); dfree(qde); dfree(qx); dfree(qy); dfree(qenv);
        qcap = 0;
    }
    int ensure_queue() {
        if (qcap == n && qid) return MPM_OK;
        free_queue();
        if (n == 0) return MPM_OK;
        const size_t sn = (size_t)n;
        int rc;
        if ((rc = hip(hipMalloc(&qid, 4 * sn), "alloc queue")) || (rc = hip(hipMalloc(&qde, 48 * sn), "alloc queue")) ||
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
        a.qcount = qcount; a.qid = qid; a.qde = qde; a.qx = qx; a.qy = qy; a.qenv = qenv;
        a.cap = cap; a.capcount = capcount; a.capcap = capcap;
        return a;
    }

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

    void free_nodes() {
        dfree(nm); dfree(np); dfree(nv); dfree(nf); dfree(na); dfree(nu); dfree(nfty); dfree(nfric);
        dfree(nact); dfree(nfix); dfree(d_fixed_y);
        nfy = 0;
        dfree(bcount); dfree(bstart); dfree(bblock);
    }
    void free_particles() {
        for (int f = 0; f < MPM_PF_COUNT; f++) dfree(pf[f]);
        dfree(pmat); dfree(pdc_rank); dfree(pdc_axes); dfree(pdc_mult);
        dfree(bkey); dfree(bsorted);
        for (int k = 0; k < MPM_MON_COUNT; k++) {
            dfree(d_mon[k]);
            mon_count[k] = 0;
        }
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
        P.n = n;
        return P;
    }
    NView nview() const {
        NView N;
        N.m = nm; N.p = np; N.v = nv; N.f = nf; N.a = na; N.u = nu; N.fty = nfty;
        N.active = nact; N.fixed = nfix; N.friction = nfric;
        return N;
    }
    BView bview() const {
        BView B;
        B.key = bkey; B.count = bcount; B.start = bstart; B.sorted = bsorted;
        B.necells = g.necells;
        return B;
    }

    /* Cell keys and counts for the current positions (after an upload). */
    int ensure_keys() {
        if (keys_valid) return MPM_OK;
        int rc = hip(hipMemsetAsync(bcount, 0, sizeof(uint32_t) * g.necells, stream), "clear cell counts");
        if (rc) return rc;
        rc = hip(launch_keys(g, pview(), bview(), stream), "cell keys");
        if (rc) return rc;
        keys_valid = true;
        return MPM_OK;
    }

    /* One step; with ev != nullptr an event is recorded before every phase
     * and after the last (ev[0..PHASE_COUNT]). */
    int step_once(double dt, hipEvent_t *ev = nullptr) {
        StepParams sp;
        sp.dt = dt;
        sp.mass_min = mass_min;
        sp.damping = damping;
        timeline.velocity(time, sp.vt);
        sp.has_fext = has_fext ? 1 : 0;
        sp.timeline_nonempty = timeline.t.empty() ? 0 : 1;
        int rc = ensure_keys();
        if (rc) return rc;
        if (ev && (rc = hip(hipEventRecord(ev[0], stream), "event"))) return rc;
        rc = hip(launch_bin(g, pview(), bview(), bblock, stream), "cell binning");
        if (rc) return rc;
        const SurrogateArgs sa = sargs();
        if (sur && (rc = ensure_queue())) return rc;
        for (int ph = PHASE_P2G_GRID; ph < PHASE_COUNT; ph++) {
            if (ev && (rc = hip(hipEventRecord(ev[ph], stream), "event"))) return rc;
            if (ph == PHASE_G2P_STRAIN && sur && (rc = hip(launch_zero_u32(qcount, 1, stream), "queue reset")))
                return rc;
            rc = hip(launch_phase(ph, g, pview(), nview(), bview(), sp, mt, d_health, sa, stream), "step kernels");
            if (rc) return rc;
        }
        if (sur) {
            /* the network on the queued points, then acceptance / fallback */
            rc = hip(mlp_launch_forward(sur->params, sur->wbf16, qx, qy, n, stream, qcount), "surrogate forward");
            if (!rc) rc = hip(launch_surrogate_post(pview(), sp, mt, d_health, sa, stream), "surrogate post");
            if (rc) return rc;
        }
        if (ev && (rc = hip(hipEventRecord(ev[PHASE_COUNT], stream), "event"))) return rc;
        keys_valid = true; /* g2p_strain wrote the next keys and counts */
        time += dt;
        cycle++;
        steps++;
        return MPM_OK;
    }
};

#define CHECK_E(e)                                   \
    do {                                             \
        if (!(e)) return MPM_ERR_ARG;                \
        if (hipSetDevice((e)->device) != hipSuccess) \
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
        hipMalloc(&e->d_health, sizeof(unsigned long long) * HEALTH_SLOTS) != hipSuccess ||
        hipMalloc(&e->d_mon_out, sizeof(double) * 8) != hipSuccess ||
        hipMalloc(&e->qcount, sizeof(uint32_t)) != hipSuccess ||
        hipMalloc(&e->capcount, sizeof(unsigned long long)) != hipSuccess ||
        hipMemset(e->d_health, 0, sizeof(unsigned long long) * HEALTH_SLOTS) != hipSuccess) {
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
    e->free_nodes();
    e->free_particles();
    dfree(e->d_health);
    dfree(e->d_mon_out);
    e->free_queue();
    dfree(e->qcount);
    dfree(e->cap);
    dfree(e->capcount);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
    return MPM_OK;
}

const char *mpm_last_error(const mpm_engine *e) { return e ? e->err.c_str() : g_create_error.c_str(); }

int mpm_set_grid(mpm_engine *e, const double length[3], const int32_t cells[3]) {
    CHECK_E(e);
    if (!length || !cells) return e->fail(MPM_ERR_ARG, "null grid argument");
    for (int a = 0; a < 3; a++) {
        if (cells[a] < 1) return e->fail(MPM_ERR_ARG, "cells[%d] = %d < 1", a, cells[a]);
        if (!(length[a] > 0.0) || !std::isfinite(length[a])) return e->fail(MPM_ERR_ARG, "length[%d] not positive", a);
    }
    GridGeom g{};
    for (int a = 0; a < 3; a++) {
        g.L[a] = length[a];
        g.h[a] = length[a] / (double)cells[a];
        g.cells[a] = cells[a];
        g.nn[a] = cells[a] + 1;
        g.ec[a] = cells[a] + 2;
    }
    g.nnodes = (int64_t)g.nn[0] * g.nn[1] * g.nn[2];
    g.necells = (int64_t)g.ec[0] * g.ec[1] * g.ec[2];
    if (g.necells >= (int64_t)0xFFFFFFFFll) return e->fail(MPM_ERR_RANGE, "grid too large: %lld cells", (long long)g.necells);
    (void)hipStreamSynchronize(e->stream);
    e->free_nodes();
    e->g = g;
    const size_t nn = (size_t)g.nnodes;
    int rc = 0;
    if ((rc = e->hip(hipMalloc(&e->nm, 8 * nn), "alloc node mass")) ||
        (rc = e->hip(hipMalloc(&e->np, 24 * nn), "alloc node momentum")) ||
        (rc = e->hip(hipMalloc(&e->nv, 24 * nn), "alloc node velocity")) ||
        (rc = e->hip(hipMalloc(&e->nf, 24 * nn), "alloc node force")) ||
        (rc = e->hip(hipMalloc(&e->na, 24 * nn), "alloc node f/m")) ||
        (rc = e->hip(hipMalloc(&e->nu, 24 * nn), "alloc node p/m")) ||
        (rc = e->hip(hipMalloc(&e->nfty, 8 * nn), "alloc node reaction")) ||
        (rc = e->hip(hipMalloc(&e->nfric, 8 * nn), "alloc node friction")) ||
        (rc = e->hip(hipMalloc(&e->nact, nn), "alloc node active")) ||
        (rc = e->hip(hipMalloc(&e->nfix, nn), "alloc node fixed")) ||
        (rc = e->hip(hipMalloc(&e->bcount, 4 * (size_t)g.necells), "alloc cell counts")) ||
        (rc = e->hip(hipMalloc(&e->bstart, 4 * ((size_t)g.necells + 1)), "alloc cell starts")) ||
        (rc = e->hip(hipMalloc(&e->bblock, 4 * (size_t)scan_block_count(g.necells)), "alloc scan blocks"))) {
        e->free_nodes();
        return rc;
    }
    if ((rc = e->hip(hipMemsetAsync(e->nact, 0, nn, e->stream), "clear active")) ||
        (rc = e->hip(hipMemsetAsync(e->nfix, 0, nn, e->stream), "clear fixed")) ||
        (rc = e->hip(hipMemsetAsync(e->nfric, 0, 8 * nn, e->stream), "clear friction")) ||
        (rc = e->hip(hipMemsetAsync(e->nm, 0, 8 * nn, e->stream), "clear mass")))
        return rc;
    e->grid_set = true;
    e->keys_valid = false;
    return e->hip(hipStreamSynchronize(e->stream), "set_grid sync");
}

int mpm_set_node_boundary(mpm_engine *e, const uint8_t *fixed_mask, const double *friction, int64_t n_nodes) {
    CHECK_E(e);
    if (!e->grid_set) return e->fail(MPM_ERR_STATE, "set_node_boundary before set_grid");
    if (n_nodes != e->g.nnodes) return e->fail(MPM_ERR_ARG, "node count %lld != %lld", (long long)n_nodes, (long long)e->g.nnodes);
    const size_t nn = (size_t)n_nodes;
    int rc;
    std::vector<int64_t> fy;
    if (fixed_mask) {
        if ((rc = e->hip(hipMemcpy(e->nfix, fixed_mask, nn, hipMemcpyHostToDevice), "upload fixed"))) return rc;
        for (size_t i = 0; i < nn; i++)
            if (fixed_mask[i] & 2) fy.push_back((int64_t)i);
    } else if ((rc = e->hip(hipMemset(e->nfix, 0, nn), "clear fixed"))) {
        return rc;
    }
    if (friction) {
        if ((rc = e->hip(hipMemcpy(e->nfric, friction, 8 * nn, hipMemcpyHostToDevice), "upload friction"))) return rc;
    } else if ((rc = e->hip(hipMemset(e->nfric, 0, 8 * nn), "clear friction"))) {
        return rc;
    }
    dfree(e->d_fixed_y);
    e->nfy = (int64_t)fy.size();
    if (!fy.empty()) {
        if ((rc = e->hip(hipMalloc(&e->d_fixed_y, 8 * fy.size()), "alloc fixed-y list"))) return rc;
        if ((rc = e->hip(hipMemcpy(e->d_fixed_y, fy.data(), 8 * fy.size(), hipMemcpyHostToDevice), "upload fixed-y list")))
            return rc;
    }
    return MPM_OK;
}

int mpm_set_materials(mpm_engine *e, const mpm_material *materials, int32_t n) {
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
    e->mt = mt;
    return MPM_OK;
}

int mpm_set_particle_count(mpm_engine *e, int64_t n) {
    CHECK_E(e);
    if (n < 0 || n >= (int64_t)0xFFFFFFFFll) return e->fail(MPM_ERR_RANGE, "particle count %lld unsupported", (long long)n);
    (void)hipStreamSynchronize(e->stream);
    e->free_particles();
    e->n = n;
    e->has_fext = false;
    e->keys_valid = false;
    if (n == 0) return MPM_OK;
    const size_t sn = (size_t)n;
    int rc;
    for (int f = 0; f < MPM_PF_COUNT; f++) {
        if ((rc = e->hip(hipMalloc(&e->pf[f], 8 * sn * kFieldComponents[f]), "alloc particle field"))) {
            e->free_particles();
            e->n = 0;
            return rc;
        }
        if ((rc = e->hip(hipMemset(e->pf[f], 0, 8 * sn * kFieldComponents[f]), "clear particle field")))
            return rc;
    }
    if ((rc = e->hip(hipMalloc(&e->pmat, 4 * sn), "alloc material ids")) ||
        (rc = e->hip(hipMalloc(&e->pdc_rank, 4 * sn), "alloc control ranks")) ||
        (rc = e->hip(hipMalloc(&e->pdc_axes, sn), "alloc control axes")) ||
        (rc = e->hip(hipMalloc(&e->pdc_mult, 4 * sn), "alloc control multipliers")) ||
        (rc = e->hip(hipMalloc(&e->bkey, 4 * sn), "alloc keys")) ||
        (rc = e->hip(hipMalloc(&e->bsorted, 4 * sn), "alloc sorted ids"))) {
        e->free_particles();
        e->n = 0;
        return rc;
    }
    /* defaults of MaterialPoint_BC (MaterialPoint_BC.h:31-82) */
    std::vector<double> F(9 * sn, 0.0);
    for (int d = 0; d < 3; d++) std::fill(F.begin() + (4 * d) * sn, F.begin() + (4 * d + 1) * sn, 1.0);
    std::vector<int32_t> minus1(sn, -1);
    std::vector<float> one(sn, 1.0f);
    if ((rc = e->hip(hipMemcpy(e->pf[MPM_PF_DEFORMATION_GRADIENT], F.data(), 8 * 9 * sn, hipMemcpyHostToDevice), "init F")) ||
        (rc = e->hip(hipMemcpy(e->pmat, minus1.data(), 4 * sn, hipMemcpyHostToDevice), "init materials")) ||
        (rc = e->hip(hipMemcpy(e->pdc_rank, minus1.data(), 4 * sn, hipMemcpyHostToDevice), "init control ranks")) ||
        (rc = e->hip(hipMemset(e->pdc_axes, 0, sn), "init control axes")) ||
        (rc = e->hip(hipMemcpy(e->pdc_mult, one.data(), 4 * sn, hipMemcpyHostToDevice), "init multipliers")))
        return rc;
    return e->hip(hipStreamSynchronize(e->stream), "particle init sync");
}

int mpm_get_particle_count(const mpm_engine *e, int64_t *n) {
    if (!e || !n) return MPM_ERR_ARG;
    *n = e->n;
    return MPM_OK;
}

int mpm_set_particle_field(mpm_engine *e, int32_t field, const double *src, int64_t n) {
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
    CHECK_E(e);
    if (n != e->n) return e->fail(MPM_ERR_ARG, "material length mismatch");
    if (n == 0) return MPM_OK;
    if (!material) return e->fail(MPM_ERR_ARG, "null material ids");
    (void)hipStreamSynchronize(e->stream);
    return e->hip(hipMemcpy(e->pmat, material, 4 * (size_t)n, hipMemcpyHostToDevice), "upload material ids");
}

int mpm_set_displacement_control(mpm_engine *e, const int64_t *index, const uint8_t *axes, const float *multiplier,
                                 int64_t count) {
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
    CHECK_E(e);
    if (kind < 0 || kind >= MPM_MON_COUNT) return e->fail(MPM_ERR_ARG, "unknown monitor kind %d", kind);
    if (count < 0 || (count > 0 && !index)) return e->fail(MPM_ERR_ARG, "bad monitor list");
    for (int64_t k = 0; k < count; k++)
        if (index[k] < 0 || index[k] >= e->n) return e->fail(MPM_ERR_ARG, "monitor index out of range");
    (void)hipStreamSynchronize(e->stream);
    dfree(e->d_mon[kind]);
    e->mon_count[kind] = count;
    if (count == 0) return MPM_OK;
    int rc = e->hip(hipMalloc(&e->d_mon[kind], 8 * (size_t)count), "alloc monitor list");
    if (rc) return rc;
    return e->hip(hipMemcpy(e->d_mon[kind], index, 8 * (size_t)count, hipMemcpyHostToDevice), "upload monitor list");
}

int mpm_set_timeline(mpm_engine *e, const double *time, const double *v3, int32_t n) {
    if (!e) return MPM_ERR_ARG;
    if (n < 0 || (n > 0 && (!time || !v3))) return e->fail(MPM_ERR_ARG, "bad timeline");
    e->timeline.t.assign(time, time + n);
    e->timeline.v.assign(v3, v3 + 3 * (size_t)n);
    return MPM_OK;
}

int mpm_set_time(mpm_engine *e, double time, double dt_max, double time_end) {
    if (!e) return MPM_ERR_ARG;
    e->time = time;
    e->dt_max = dt_max;
    e->time_end = time_end;
    return MPM_OK;
}

int mpm_set_damping(mpm_engine *e, double damping, double mass_minimum) {
    if (!e) return MPM_ERR_ARG;
    e->damping = damping;
    e->mass_min = mass_minimum;
    return MPM_OK;
}

int mpm_step(mpm_engine *e, double dt, int32_t nsteps) {
    CHECK_E(e);
    if (!e->grid_set) return e->fail(MPM_ERR_STATE, "step before set_grid");
    if (nsteps < 0) return e->fail(MPM_ERR_ARG, "negative step count");
    for (int s = 0; s < nsteps; s++) {
        int rc = e->step_once(dt);
        if (rc) return rc;
    }
    return MPM_OK;
}

int mpm_step_timed(mpm_engine *e, double dt, int32_t nsteps, double phase_ms[MPM_PHASE_COUNT]) {
    CHECK_E(e);
    if (!e->grid_set) return e->fail(MPM_ERR_STATE, "step before set_grid");
    if (nsteps < 0 || !phase_ms) return e->fail(MPM_ERR_ARG, "bad timed-step arguments");
    static_assert(MPM_PHASE_COUNT == PHASE_COUNT, "phase enums disagree");
    for (int k = 0; k < MPM_PHASE_COUNT; k++) phase_ms[k] = 0.0;
    hipEvent_t ev[PHASE_COUNT + 1] = {};
    int rc = MPM_OK;
    for (int k = 0; k <= PHASE_COUNT && !rc; k++) rc = e->hip(hipEventCreate(&ev[k]), "event create");
    for (int s = 0; s < nsteps && !rc; s++) {
        rc = e->step_once(dt, ev);
        if (!rc) rc = e->hip(hipEventSynchronize(ev[PHASE_COUNT]), "event sync");
        for (int k = 0; k < PHASE_COUNT && !rc; k++) {
            float ms = 0.f;
            rc = e->hip(hipEventElapsedTime(&ms, ev[k], ev[k + 1]), "event time");
            phase_ms[k] += ms;
        }
    }
    for (int k = 0; k <= PHASE_COUNT; k++)
        if (ev[k]) (void)hipEventDestroy(ev[k]);
    return rc;
}

int mpm_run(mpm_engine *e, double dt_total, int32_t *finished) {
    CHECK_E(e);
    if (!e->grid_set) return e->fail(MPM_ERR_STATE, "run before set_grid");
    double acc = 0.0;
    while (acc < dt_total) {
        double dt = e->dt_max;
        if (e->dt_max > (dt_total - acc)) dt = dt_total - acc;
        acc += dt;
        int rc = e->step_once(dt);
        if (rc) return rc;
    }
    if (finished) *finished = (e->time < e->time_end) ? 0 : 1;
    return MPM_OK;
}

int mpm_synchronize(mpm_engine *e) {
    CHECK_E(e);
    return e->hip(hipStreamSynchronize(e->stream), "stream sync");
}

int mpm_get_time(const mpm_engine *e, double *time, int64_t *cycle) {
    if (!e) return MPM_ERR_ARG;
    if (time) *time = e->time;
    if (cycle) *cycle = e->cycle;
    return MPM_OK;
}

int mpm_get_node_count(const mpm_engine *e, int64_t *n) {
    if (!e || !n) return MPM_ERR_ARG;
    *n = e->grid_set ? e->g.nnodes : 0;
    return MPM_OK;
}

int mpm_get_node_field(mpm_engine *e, int32_t field, double *dst, int64_t n) {
    CHECK_E(e);
    if (!e->grid_set) return e->fail(MPM_ERR_STATE, "no grid");
    if (field < 0 || field >= MPM_NF_COUNT) return e->fail(MPM_ERR_ARG, "unknown node field %d", field);
    if (n != e->g.nnodes || !dst) return e->fail(MPM_ERR_ARG, "node field length mismatch");
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
    if (!e->grid_set) return e->fail(MPM_ERR_STATE, "no grid");
    std::memset(out, 0, sizeof(*out));
    out->time = e->time;
    MonLists ml;
    for (int k = 0; k < MPM_MON_COUNT; k++) {
        ml.list[k] = e->d_mon[k];
        ml.count[k] = e->mon_count[k];
    }
    ml.fixed_y = e->d_fixed_y;
    ml.nfy = e->nfy;
    int rc = e->hip(launch_monitors(e->g, e->pview(), e->nview(), ml, e->d_mon_out, e->stream), "monitors");
    if (rc) return rc;
    double v[8];
    if ((rc = e->hip(hipMemcpyAsync(v, e->d_mon_out, sizeof(v), hipMemcpyDeviceToHost, e->stream), "monitor copy")))
        return rc;
    if ((rc = e->hip(hipStreamSynchronize(e->stream), "monitor sync"))) return rc;
    out->has_position = e->mon_count[MPM_MON_DISPLACEMENT] > 0;
    out->has_momentum = e->mon_count[MPM_MON_MOMENTUM] > 0;
    out->has_energy = e->mon_count[MPM_MON_ENERGY] > 0;
    out->has_stress = e->mon_count[MPM_MON_STRESS] > 0;
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

int mpm_get_health(mpm_engine *e, mpm_health *out) {
    CHECK_E(e);
    if (!out) return e->fail(MPM_ERR_ARG, "null health output");
    unsigned long long h[HEALTH_SLOTS];
    int rc = e->hip(hipStreamSynchronize(e->stream), "health sync");
    if (!rc) rc = e->hip(hipMemcpy(h, e->d_health, sizeof(h), hipMemcpyDeviceToHost), "health copy");
    if (rc) return rc;
    out->steps = e->steps;
    out->outside_grid = (int64_t)h[HEALTH_OUTSIDE];
    out->null_material = (int64_t)h[HEALTH_NULL];
    out->plastic_points = (int64_t)h[HEALTH_PLASTIC];
    out->nonfinite_particles = (int64_t)h[HEALTH_NONFINITE];
    out->surrogate_calls = (int64_t)h[HEALTH_SURROGATE_CALLS];
    out->surrogate_fallbacks = (int64_t)h[HEALTH_SURROGATE_FALLBACKS];
    return MPM_OK;
}

int mpm_set_surrogate(mpm_engine *e, mpm_surrogate *s, int32_t material, double residual_tolerance) {
    CHECK_E(e);
    if (s && s->device != e->device) return e->fail(MPM_ERR_ARG, "surrogate lives on another device");
    if (s && (material < 0 || material >= e->mt.n)) return e->fail(MPM_ERR_ARG, "surrogate material out of range");
    if (s && !(residual_tolerance >= 0.0)) return e->fail(MPM_ERR_ARG, "negative residual tolerance");
    (void)hipStreamSynchronize(e->stream);
    if (s) (void)hipStreamSynchronize(s->stream);
    e->sur = s;
    e->sur_material = s ? material : -1;
    e->sur_tol = residual_tolerance;
    if (!s) e->free_queue();
    return MPM_OK;
}

int mpm_set_sample_capture(mpm_engine *e, int64_t capacity) {
    CHECK_E(e);
    if (capacity < 0) return e->fail(MPM_ERR_ARG, "negative capture capacity");
    (void)hipStreamSynchronize(e->stream);
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

int mpm_reset_health(mpm_engine *e) {
    CHECK_E(e);
    e->steps = 0;
    int rc = e->hip(hipStreamSynchronize(e->stream), "health sync");
    if (rc) return rc;
    return e->hip(hipMemset(e->d_health, 0, sizeof(unsigned long long) * HEALTH_SLOTS), "health reset");
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
