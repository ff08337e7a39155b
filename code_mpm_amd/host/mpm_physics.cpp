/*
 * mpm_physics.cpp — the reference's host classes over the C ABI of
 * include/mpm.h (see mpm_physics.h).
 */
#include "mpm_physics.h"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <ctime>
#include <fstream>
#include <iostream>
#include <unordered_map>

namespace mpmhost {

/* ---- factories, time line ----------------------------------------------- */

unsigned int GridPoint_Factory::getIndex(ivec3 i, ivec3 nodes) {
    return (unsigned int)((long)i.x + (long)i.y * nodes.x + (long)i.z * nodes.x * nodes.y);
}

std::vector<GridPoint *> GridPoint_Factory::createGrid(dvec3 length, ivec3 cells) {
    const ivec3 nn(cells.x + 1, cells.y + 1, cells.z + 1);
    std::vector<GridPoint *> out((size_t)nn.x * nn.y * nn.z);
    for (int k = 0; k < nn.z; k++)
        for (int j = 0; j < nn.y; j++)
            for (int i = 0; i < nn.x; i++) {
                GridPoint *g = new GridPoint();
                g->d3_Position = dvec3((double)i * length.x / (double)cells.x, (double)j * length.y / (double)cells.y,
                                       (double)k * length.z / (double)cells.z);
                out[getIndex(ivec3(i, j, k), nn)] = g;
            }
    return out;
}

MaterialPoint_BC *MaterialPoint_Factory_Classic_CC::createMaterialPoint(dvec3 position) {
    MaterialPoint_BC *p = new MaterialPoint_Classic_CC();
    p->d3_Position = position;
    return p;
}

std::vector<MaterialPoint_BC *> MaterialPoint_Factory_Classic_CC::createDomain_Cuboid(dvec3 c, dvec3 dim,
                                                                                       double off) {
    std::vector<MaterialPoint_BC *> out;
    const bool single = dim.z <= off;
    const double z0 = single ? 0.0 : 0.5 * off;
    /* the quarter x, y > 0 (z >= 0), each point mirrored into the other
     * quadrants in the order (+,+), (-,+), (-,-), (+,-), then below z = 0;
     * the loop variables accumulate in floating point */
    for (double dx = 0.5 * off; dx <= 0.5 * dim.x; dx += off)
        for (double dy = 0.5 * off; dy <= 0.5 * dim.y; dy += off)
            for (double dz = z0; dz <= 0.5 * dim.z; dz += off) {
                const double sx[4] = {+dx, -dx, -dx, +dx}, sy[4] = {+dy, +dy, -dy, -dy};
                for (int q = 0; q < 4; q++) out.push_back(createMaterialPoint(c + dvec3(sx[q], sy[q], +dz)));
                if (!single)
                    for (int q = 0; q < 4; q++) out.push_back(createMaterialPoint(c + dvec3(sx[q], sy[q], -dz)));
            }
    return out;
}

void TimeLine::addTimePoint(double time, dvec3 velocity) {
    v_Time.push_back(time);
    v_Velocity.push_back(velocity);
}

dvec3 TimeLine::getVelocity(double time) const {
    dvec3 r;
    for (size_t k = 0; k + 1 < v_Time.size(); k++)
        if (v_Time[k] <= time && time < v_Time[k + 1]) {
            const double s = (time - v_Time[k]) / (v_Time[k + 1] - v_Time[k]);
            r = (1.0 - s) * v_Velocity[k] + s * v_Velocity[k + 1];
        }
    return r;
}

/* ---- shape functions ------------------------------------------------------ */

void GridPoint_Mediator::findAdjacentGridPoints(dvec3 x) {
    v_adjacentGridPoints.clear();
    const double xs[3] = {x.x, x.y, x.z}, hs[3] = {d3_Length_Cell.x, d3_Length_Cell.y, d3_Length_Cell.z};
    const int nn[3] = {i3_Node_Count.x, i3_Node_Count.y, i3_Node_Count.z};
    int b[3], lo[3], hi[3];
    for (int a = 0; a < 3; a++) {
        const double q = xs[a] / hs[a];
        b[a] = (int)std::floor(q);
        /* tolerance 0: a coordinate on a node plane keeps one layer */
        lo[a] = ((double)b[a] + 1.0 - xs[a] / hs[a]) < 0.0 ? 1 : 0;
        hi[a] = ((xs[a] / hs[a]) - (double)b[a]) < 0.0 ? 1 : 2;
    }
    for (int i = lo[0]; i < hi[0]; i++)
        for (int j = lo[1]; j < hi[1]; j++)
            for (int k = lo[2]; k < hi[2]; k++) {
                const ivec3 n(b[0] + i, b[1] + j, b[2] + k);
                if (n.x < 0 || n.y < 0 || n.z < 0 || n.x > nn[0] - 1 || n.y > nn[1] - 1 || n.z > nn[2] - 1) continue;
                v_adjacentGridPoints.push_back(GridPoint_Factory::getIndex(n, i3_Node_Count));
            }
}

void GridPoint_Mediator::calculateBases_Classic(dvec3 mp, dvec3 gp) {
    const double d[3] = {mp.x - gp.x, mp.y - gp.y, mp.z - gp.z};
    const double h[3] = {d3_Length_Cell.x, d3_Length_Cell.y, d3_Length_Cell.z};
    double s[3], g[3] = {0.0, 0.0, 0.0};
    for (int a = 0; a < 3; a++) {
        s[a] = 1.0 - std::fabs(d[a] / h[a]);
        if (s[a] < 0.0 || s[a] > 1.0) s[a] = 0.0;
    }
    d_ShapeValue = s[0] * s[1] * s[2];
    /* the gradient along an axis only strictly inside (1e-12, 1 - 1e-12) of
     * a cell length: -s_j * s_l * sign(d) / h, (j, l) the other two axes in
     * order */
    const double tol = 1.0e-12;
    const int other[3][2] = {{1, 2}, {0, 2}, {0, 1}};
    for (int a = 0; a < 3; a++) {
        const double r = std::fabs(d[a] / h[a]);
        if (r > tol && r < 1.0 - tol) {
            const double sg = (double)(0.0 < d[a]) - (double)(d[a] < 0.0);
            g[a] = -s[other[a][0]] * s[other[a][1]] * sg / h[a];
        }
    }
    d3_ShapeGradient = dvec3(g[0], g[1], g[2]);
}

/* ---- return maps ----------------------------------------------------------- */

static mpm_material to_abi(const Material_BC &m) {
    mpm_material r;
    r.type = m.i_MaterialType;
    r.id = m.i_ID;
    r.density = m.d_Density;
    r.elastic_modulus = m.d_ElasticModulus;
    r.poisson_ratio = m.d_PoissonRatio;
    r.yield_stress = m.d_YieldStress;
    r.viscosity = m.d_Viscosity;
    r.hardening_c0 = m.d_Hardening_Isotropic_C0;
    r.hardening_c1 = m.d_Hardening_Isotropic_C1;
    return r;
}

int ConstitutiveRelation::calculateIncrement_Batch(const Material_BC &material, int64_t n, const double *backstress,
                                                  const double *stress, const double *strain_increment,
                                                  double *stress_increment, double *plastic_strain_increment,
                                                  double *backstress_increment, int device) {
    const mpm_material m = to_abi(material);
    return mpm_constitutive_batch(device, &m, backstress, stress, strain_increment, stress_increment,
                                  plastic_strain_increment, backstress_increment, n);
}

void ConstitutiveRelation::run(Material_BC *material, double backstress, const double *stress,
                               const double *strain_increment) {
    if (!material) return;
    const double zero[6] = {0, 0, 0, 0, 0, 0};
    calculateIncrement_Batch(*material, 1, &backstress, stress ? stress : zero, strain_increment, d6StressIncrement,
                             d6PlasticStrainIncrement, &dBackstress_IsotropicIncrement);
}

void ConstitutiveRelation::calculateIncrement_Elastic(Material_BC *material, double strain_increment[6]) {
    if (!material) return;
    Material_BC m = *material;
    m.i_MaterialType = __ELASTIC;
    run(&m, 0.0, nullptr, strain_increment);
}

void ConstitutiveRelation::calculateIncrement_Plastic(Material_BC *material, double stress[6],
                                                      double strain_increment[6]) {
    if (!material) return;
    Material_BC m = *material;
    m.i_MaterialType = __PLASTIC;
    run(&m, 0.0, stress, strain_increment);
}

void ConstitutiveRelation::calculateIncrement_VonMisesHardening(Material_BC *material, double backstress,
                                                                double stress[6], double strain_increment[6]) {
    if (!material) return;
    Material_BC m = *material;
    m.i_MaterialType = __VONMISESHARDENING;
    run(&m, backstress, stress, strain_increment);
}

/* ---- the engine -------------------------------------------------------------- */

PhysicsEngine::PhysicsEngine(int device) {
    if (mpm_create(&gpu, device) != MPM_OK) {
        s_Error = mpm_last_error(nullptr);
        gpu = nullptr;
    }
}

PhysicsEngine::~PhysicsEngine() {
    if (gpu) mpm_destroy(gpu);
    for (MaterialPoint_BC *p : allMaterialPoint) delete p;
    for (GridPoint *g : allGridPoint) delete g;
    for (Material_BC *m : v_allMaterial) delete m;
}

const char *PhysicsEngine::last_error() const { return s_Error.c_str(); }

int PhysicsEngine::fail(int rc) {
    s_Error = gpu ? mpm_last_error(gpu) : "no engine";
    return rc ? rc : MPM_ERR_STATE;
}

void PhysicsEngine::initializeWorld_Classic_Bar() {
    const ivec3 cells(64, 8, 8);
    const double h0 = 1.5625e-3, E = 200.0e9, nu = 0.3, rho = 7850.0;
    d3_Length_World = dvec3(cells.x * h0, cells.y * h0, cells.z * h0);
    i3_Cells = cells;
    const double h = d3_Length_World.x / (double)cells.x;
    GridPoint_Factory gf;
    allGridPoint = gf.createGrid(d3_Length_World, cells);
    for (GridPoint *g : allGridPoint)
        if (g->d3_Position.x == 0.0) g->b3_Fixed = bvec3(true, true, true);
    Material_BC *steel = new Material_BC();
    steel->i_MaterialType = __ELASTIC;
    steel->d_Density = rho;
    steel->d_ElasticModulus = E;
    steel->d_PoissonRatio = nu;
    v_allMaterial.push_back(steel);
    const double off = h / 2.0;
    const double V0 = std::pow(off, 3.0);
    MaterialPoint_Factory_Classic_CC mf;
    /* cells with x fastest, then y, then z; in each cell the 2 x 2 x 2
     * points with x fastest */
    for (int cz = 0; cz < cells.z; cz++)
        for (int cy = 0; cy < cells.y; cy++)
            for (int cx = 0; cx < cells.x; cx++)
                for (int oz = 0; oz < 2; oz++)
                    for (int oy = 0; oy < 2; oy++)
                        for (int ox = 0; ox < 2; ox++) {
                            const dvec3 x(((double)cx + (ox ? 0.75 : 0.25)) * h, ((double)cy + (oy ? 0.75 : 0.25)) * h,
                                          ((double)cz + (oz ? 0.75 : 0.25)) * h);
                            MaterialPoint_BC *p = mf.createMaterialPoint(x);
                            p->i_ID = allMaterialPoint.size();
                            p->p_Material = steel;
                            p->d_Volume_Initial = V0;
                            p->d_Volume = V0;
                            p->d_Mass = rho * V0;
                            if (x.x > (double)(cells.x - 1) * h) {
                                p->d3_Velocity = dvec3(0.0, 0.0, -1.0);
                                v_MarkedMaterialPoints_Displacement_Monitor.push_back(p);
                            }
                            if (x.x < 2.0 * h) v_MarkedMaterialPoints_Stress_Monitor.push_back(p);
                            v_MarkedMaterialPoints_Momentum.push_back(p);
                            v_MarkedMaterialPoints_Monitor_Energy.push_back(p);
                            allMaterialPoint.push_back(p);
                        }
    const double c = std::sqrt(E / rho);
    d_TimeIncrement_Maximum = 0.1 * h / c;
    d_TimeEnd = 1000.0 * d_TimeIncrement_Maximum;
    d_TimeConsole_Interval = 100.0 * d_TimeIncrement_Maximum;
    d_Time = 0.0;
    d_Runtime_Total = 0.0;
    b_Uploaded = false;
    /* the description and the log file, as the reference's scene functions
     * write them (e.g. PhysicsEngine_initializeWorld_Classic_Bar.cpp:200-255) */
    double mass = 0.0;
    for (const MaterialPoint_BC *p : allMaterialPoint) mass += p->d_Mass;
    char stamp[80], started[80];
    const std::time_t now = std::time(nullptr);
    std::strftime(started, sizeof(started), "%d-%m-%Y %H:%M:%S", std::localtime(&now));
    std::strftime(stamp, sizeof(stamp), "%Y%m%d%H%M%S", std::localtime(&now));
    std::string d;
    d += "-------------------------------------------------------------\n";
    d += "Classic formulation, cantilever bar ---------------------------\n";
    d += "Process started on: " + std::string(started) + "\n";
    d += "-------------------------------------------------------------\n";
    d += "Device: MI355X, one GPU\n";
    d += "Time increment: " + Script(d_TimeIncrement_Maximum, 6) + "\n";
    d += "Material Point count: " + Script((long long)allMaterialPoint.size()) + "\n";
    d += "Mass: " + Script(mass, 6) + "\n";
    d += "-------------------------------------------------------------\n";
    d += "Grid Resolution: (" + Script((long long)cells.x) + "," + Script((long long)cells.y) + "," +
         Script((long long)cells.z) + ")(" + Script(h, 3) + ")\n";
    d += "dOffset: " + Script(off, 4) + "\n";
    d += "Modulus: " + Script(E, 3) + " N/m^2" + "\n";
    d += "Global Damping: " + Script(d_DampingCoefficient, 3) + "\n";
    d_TimeConsole_Last = 0.0;
    str_Log_FileName = std::string("./bmp/log_") + stamp + ".txt";
    reportConsole(d);
}

std::string Script(double r, int precision) {
    char buf[64];
    std::snprintf(buf, sizeof(buf), "%+.*e", precision, r);
    return buf;
}

std::string Script(long long i) { return std::to_string(i); }

std::string consoleLine(const ConsoleReport &r, const std::string &sDescription) {
    std::string s = sDescription;
    s += "\ttime: " + Script(r.time, 4);
    s += "\tRuntime: " + Script(r.runtime, 3);
    if (r.has_position) s += "\tPosition_y: " + Script(r.position_y, 14);
    s += "\tForce_y: " + Script(r.force_y, 4);
    if (r.has_momentum) s += "\t momentum_y: " + Script(r.momentum_y, 4);
    if (r.has_stress) {
        s += "\tStrain_y: " + Script(r.strain_y, 4);
        s += "\tStress_y: " + Script(r.stress_y, 4);
    }
    if (r.has_energy) {
        s += "\tKinetic Energy: " + Script(r.kinetic_energy, 4);
        s += "\tStrain Energy: " + Script(r.strain_energy, 4);
        s += "\tPlastic Energy: " + Script(r.plastic_energy, 4);
    }
    s += "\n";
    return s;
}

int PhysicsEngine::upload() {
    if (!gpu) return fail(MPM_ERR_STATE);
    const double L[3] = {d3_Length_World.x, d3_Length_World.y, d3_Length_World.z};
    const int32_t C[3] = {i3_Cells.x, i3_Cells.y, i3_Cells.z};
    if (mpm_set_grid(gpu, L, C) != MPM_OK) return fail(MPM_ERR_ARG);
    int64_t nn = 0;
    mpm_get_node_count(gpu, &nn);
    if ((int64_t)allGridPoint.size() != nn) {
        s_Error = "allGridPoint does not match the grid of d3_Length_World / i3_Cells";
        return MPM_ERR_ARG;
    }
    std::vector<uint8_t> fixed((size_t)nn);
    std::vector<double> mu((size_t)nn);
    for (size_t i = 0; i < (size_t)nn; i++) {
        const GridPoint *g = allGridPoint[i];
        fixed[i] = (uint8_t)((g->b3_Fixed.x ? 1 : 0) | (g->b3_Fixed.y ? 2 : 0) | (g->b3_Fixed.z ? 4 : 0));
        mu[i] = g->d_FrictionCoefficient;
    }
    if (mpm_set_node_boundary(gpu, fixed.data(), mu.data(), nn) != MPM_OK) return fail(MPM_ERR_ARG);
    std::vector<mpm_material> mats;
    std::unordered_map<const Material_BC *, int32_t> mat_index;
    for (const Material_BC *m : v_allMaterial) {
        mat_index[m] = (int32_t)mats.size();
        mats.push_back(to_abi(*m));
    }
    if (mpm_set_materials(gpu, mats.data(), (int32_t)mats.size()) != MPM_OK) return fail(MPM_ERR_ARG);
    const int64_t n = (int64_t)allMaterialPoint.size();
    if (mpm_set_particle_count(gpu, n) != MPM_OK) return fail(MPM_ERR_ARG);
    std::unordered_map<const MaterialPoint_BC *, int64_t> index;
    for (int64_t i = 0; i < n; i++) index[allMaterialPoint[i]] = i;
    if (n > 0) {
        std::vector<double> buf;
        auto put = [&](int32_t field, int nc, auto get) -> int {
            buf.assign((size_t)nc * n, 0.0);
            for (int64_t i = 0; i < n; i++)
                for (int c = 0; c < nc; c++) buf[(size_t)c * n + i] = get(allMaterialPoint[i], c);
            return mpm_set_particle_field(gpu, field, buf.data(), n);
        };
        auto v3 = [](const dvec3 &v, int c) { return c == 0 ? v.x : (c == 1 ? v.y : v.z); };
        int rc = MPM_OK;
        rc |= put(MPM_PF_POSITION, 3, [&](MaterialPoint_BC *p, int c) { return v3(p->d3_Position, c); });
        rc |= put(MPM_PF_VELOCITY, 3, [&](MaterialPoint_BC *p, int c) { return v3(p->d3_Velocity, c); });
        rc |= put(MPM_PF_FORCE_EXTERNAL, 3, [&](MaterialPoint_BC *p, int c) { return v3(p->d3_Force_External, c); });
        rc |= put(MPM_PF_MASS, 1, [](MaterialPoint_BC *p, int) { return p->d_Mass; });
        rc |= put(MPM_PF_VOLUME, 1, [](MaterialPoint_BC *p, int) { return p->d_Volume; });
        rc |= put(MPM_PF_VOLUME_INITIAL, 1, [](MaterialPoint_BC *p, int) { return p->d_Volume_Initial; });
        rc |= put(MPM_PF_DEFORMATION_GRADIENT, 9,
                  [](MaterialPoint_BC *p, int c) { return p->d33_DeformationGradient.m[c / 3][c % 3]; });
        rc |= put(MPM_PF_DEFORMATION_GRADIENT_ELASTIC, 9,
                  [](MaterialPoint_BC *p, int c) { return p->d33_DeformationGradient_Elastic.m[c / 3][c % 3]; });
        rc |= put(MPM_PF_STRAIN, 6, [](MaterialPoint_BC *p, int c) { return p->d6_Strain[c]; });
        rc |= put(MPM_PF_STRESS, 6, [](MaterialPoint_BC *p, int c) { return p->d6_Stress[c]; });
        rc |= put(MPM_PF_STRAIN_PLASTIC, 6, [](MaterialPoint_BC *p, int c) { return p->d6_Strain_Plastic[c]; });
        rc |= put(MPM_PF_BACKSTRESS, 1, [](MaterialPoint_BC *p, int) { return p->d_BackStress_Isotropic; });
        rc |= put(MPM_PF_ENERGY_STRAIN, 1, [](MaterialPoint_BC *p, int) { return p->d_Energy_Strain; });
        rc |= put(MPM_PF_ENERGY_STRAIN_MAX, 1, [](MaterialPoint_BC *p, int) { return p->d_Energy_Strain_MaxHistory; });
        rc |= put(MPM_PF_ENERGY_STRAIN_TI, 1,
                  [](MaterialPoint_BC *p, int) { return p->d_Energy_Strain_TimeIntegral; });
        rc |= put(MPM_PF_ENERGY_PLASTIC, 1, [](MaterialPoint_BC *p, int) { return p->d_Energy_Plastic; });
        if (rc != MPM_OK) return fail(MPM_ERR_ARG);
        std::vector<int32_t> mi((size_t)n);
        for (int64_t i = 0; i < n; i++) {
            auto it = mat_index.find(allMaterialPoint[i]->p_Material);
            mi[(size_t)i] = it == mat_index.end() ? -1 : it->second; /* NULL or foreign: no material */
        }
        if (mpm_set_particle_material(gpu, mi.data(), n) != MPM_OK) return fail(MPM_ERR_ARG);
    }
    auto indices = [&](const std::vector<MaterialPoint_BC *> &list, std::vector<int64_t> &out) -> bool {
        out.clear();
        for (const MaterialPoint_BC *p : list) {
            auto it = index.find(p);
            if (it == index.end()) return false;
            out.push_back(it->second);
        }
        return true;
    };
    std::vector<int64_t> idx;
    if (!indices(v_MarkedMaterialPoints_Displacement_Control, idx)) {
        s_Error = "a displacement-controlled point is not in allMaterialPoint";
        return MPM_ERR_ARG;
    }
    {
        std::vector<uint8_t> ax;
        std::vector<float> mult;
        for (const MaterialPoint_BC *p : v_MarkedMaterialPoints_Displacement_Control) {
            ax.push_back((uint8_t)((p->b3_DisplacementControl.x ? 1 : 0) | (p->b3_DisplacementControl.y ? 2 : 0) |
                                   (p->b3_DisplacementControl.z ? 4 : 0)));
            mult.push_back(p->f_DisplacementControl_Multiplier);
        }
        if (mpm_set_displacement_control(gpu, idx.data(), ax.data(), mult.data(), (int64_t)idx.size()) != MPM_OK)
            return fail(MPM_ERR_ARG);
    }
    const std::vector<MaterialPoint_BC *> *lists[MPM_MON_COUNT] = {
        &v_MarkedMaterialPoints_Displacement_Monitor, &v_MarkedMaterialPoints_Momentum,
        &v_MarkedMaterialPoints_Monitor_Energy, &v_MarkedMaterialPoints_Stress_Monitor};
    for (int k = 0; k < MPM_MON_COUNT; k++) {
        if (!indices(*lists[k], idx)) {
            s_Error = "a monitored point is not in allMaterialPoint";
            return MPM_ERR_ARG;
        }
        if (mpm_set_monitor_list(gpu, k, idx.data(), (int64_t)idx.size()) != MPM_OK) return fail(MPM_ERR_ARG);
    }
    std::vector<double> tv;
    for (const dvec3 &v : m_TimeLine.v_Velocity) {
        tv.push_back(v.x);
        tv.push_back(v.y);
        tv.push_back(v.z);
    }
    if (mpm_set_timeline(gpu, m_TimeLine.v_Time.data(), tv.data(), (int32_t)m_TimeLine.v_Time.size()) != MPM_OK ||
        mpm_set_time(gpu, d_Time, d_TimeIncrement_Maximum, d_TimeEnd) != MPM_OK ||
        mpm_set_damping(gpu, d_DampingCoefficient, d_Mass_Minimum) != MPM_OK ||
        mpm_set_summation_order(gpu, i_SummationOrder) != MPM_OK)
        return fail(MPM_ERR_ARG);
    b_Uploaded = true;
    b_ParticlesFresh = b_GridFresh = true;
    return MPM_OK;
}

int PhysicsEngine::run_loop(double total, int32_t variant) {
    if (!gpu) return -1;
    if (!b_Uploaded && upload() != MPM_OK) return -1;
    if (mpm_set_variant(gpu, variant) != MPM_OK || mpm_set_summation_order(gpu, i_SummationOrder) != MPM_OK)
        return fail(0), -1;
    const double dt_max = d_TimeIncrement_Maximum;
    /* d_Runtime_Total (PhysicsEngine_runSimulation_Classic_DoublePass_MPLocks.cpp:461, 794):
     * wall time of the steps, the console reports left out */
    using clock = std::chrono::steady_clock;
    clock::time_point mark = clock::now();
    auto settle = [&]() {
        mpm_synchronize(gpu);
        const clock::time_point t = clock::now();
        d_Runtime_Total += std::chrono::duration<double>(t - mark).count();
        mark = t;
    };
    int32_t pending = 0; /* full steps not yet sent to the device */
    auto flush = [&]() -> bool {
        if (!pending) return true;
        const int rc = mpm_step(gpu, dt_max, pending);
        pending = 0;
        return rc == MPM_OK;
    };
    double acc = 0.0;
    while (acc < total) {
        double dt = dt_max;
        if (dt_max > (total - acc)) dt = total - acc;
        acc += dt;
        if (dt == dt_max) {
            pending++;
        } else if (!flush() || mpm_step(gpu, dt, 1) != MPM_OK) {
            return fail(0), -1;
        }
        b_ParticlesFresh = b_GridFresh = false;
        /* the console check sees the time before this step's increment */
        const double t_before = d_Time;
        d_Time += dt;
        if (d_TimeConsole_Interval > 0.0 && t_before - d_TimeConsole_Last > d_TimeConsole_Interval) {
            if (!flush()) return fail(0), -1;
            settle();
            d_TimeConsole_Last = t_before;
            /* the reference reports before d_Time takes this step's increment */
            d_Time = t_before;
            reportConsole();
            d_Time = t_before + dt;
            mark = clock::now();
        }
    }
    if (!flush()) return fail(0), -1;
    settle();
    if (d_Time < d_TimeEnd) return 0;
    d_TimeConsole_Last = d_Time;
    reportConsole();
    return 1;
}

int PhysicsEngine::runSimulation_Phase(int32_t variant, int32_t phase, double dt) {
    if (!gpu) return -1;
    if (!b_Uploaded && upload() != MPM_OK) return -1;
    if (phase == MPM_PHASE_BIN) {
        if (mpm_set_variant(gpu, variant) != MPM_OK || mpm_set_summation_order(gpu, i_SummationOrder) != MPM_OK)
            return fail(0), -1;
        d_PhaseIncrement = dt;
    }
    if (mpm_step_phase(gpu, phase, dt) != MPM_OK) return fail(0), -1;
    b_ParticlesFresh = b_GridFresh = false;
    if (phase == MPM_PHASE_G2P_STRAIN) d_Time += d_PhaseIncrement;
    return 0;
}

int PhysicsEngine::runSimulation_Classic_DoublePass_MPLocks(double total) {
    return run_loop(total, MPM_VARIANT_DOUBLEPASS);
}
int PhysicsEngine::runSimulation_Classic_DoublePass_MPLocks_Contact(double total) {
    return run_loop(total, MPM_VARIANT_DOUBLEPASS_CONTACT);
}
int PhysicsEngine::runSimulation_Classic_SinglePass_MP(double total) {
    return run_loop(total, MPM_VARIANT_SINGLEPASS);
}

ConsoleReport PhysicsEngine::reportConsole(const std::string &sDescription) {
    ConsoleReport r{};
    r.time = d_Time;
    r.runtime = d_Runtime_Total;
    mpm_monitors m{};
    /* the monitors of the world as it stands (uploaded first if need be:
     * the scene functions report before the first step) */
    if (gpu && (b_Uploaded || upload() == MPM_OK) && mpm_get_monitors(gpu, &m) == MPM_OK) {
        r.position_y = m.position_y;
        r.force_y = m.force_y;
        r.momentum_y = m.momentum_y;
        r.kinetic_energy = m.kinetic_energy;
        r.strain_energy = m.strain_energy;
        r.plastic_energy = m.plastic_energy;
        r.strain_y = m.strain_y;
        r.stress_y = m.stress_y;
        r.has_position = m.has_position != 0;
        r.has_momentum = m.has_momentum != 0;
        r.has_stress = m.has_stress != 0;
        r.has_energy = m.has_energy != 0;
    }
    r.line = consoleLine(r, sDescription);
    if (b_Console_Print) std::cout << r.line << std::flush;
    if (!str_Log_FileName.empty()) {
        /* appended as the reference does; a log directory that does not
         * exist leaves the file unwritten, as there */
        std::ofstream log(str_Log_FileName.c_str(), std::ios_base::app);
        log << r.line;
    }
    v_ConsoleLog.push_back(r);
    return r;
}

int PhysicsEngine::refresh_particles() {
    if (b_ParticlesFresh || !gpu || !b_Uploaded) return MPM_OK;
    const int64_t n = (int64_t)allMaterialPoint.size();
    std::vector<double> buf;
    auto get = [&](int32_t field, int nc, auto set) -> int {
        buf.resize((size_t)nc * n);
        const int rc = mpm_get_particle_field(gpu, field, buf.data(), n);
        if (rc != MPM_OK) return rc;
        for (int64_t i = 0; i < n; i++)
            for (int c = 0; c < nc; c++) set(allMaterialPoint[i], c, buf[(size_t)c * n + i]);
        return MPM_OK;
    };
    auto v3 = [](dvec3 &v, int c, double x) { (c == 0 ? v.x : (c == 1 ? v.y : v.z)) = x; };
    int rc = MPM_OK;
    rc |= get(MPM_PF_POSITION, 3, [&](MaterialPoint_BC *p, int c, double x) { v3(p->d3_Position, c, x); });
    rc |= get(MPM_PF_VELOCITY, 3, [&](MaterialPoint_BC *p, int c, double x) { v3(p->d3_Velocity, c, x); });
    rc |= get(MPM_PF_VOLUME, 1, [](MaterialPoint_BC *p, int, double x) { p->d_Volume = x; });
    rc |= get(MPM_PF_DEFORMATION_GRADIENT, 9,
              [](MaterialPoint_BC *p, int c, double x) { p->d33_DeformationGradient.m[c / 3][c % 3] = x; });
    rc |= get(MPM_PF_DEFORMATION_GRADIENT_ELASTIC, 9,
              [](MaterialPoint_BC *p, int c, double x) { p->d33_DeformationGradient_Elastic.m[c / 3][c % 3] = x; });
    rc |= get(MPM_PF_STRAIN, 6, [](MaterialPoint_BC *p, int c, double x) { p->d6_Strain[c] = x; });
    rc |= get(MPM_PF_STRESS, 6, [](MaterialPoint_BC *p, int c, double x) { p->d6_Stress[c] = x; });
    rc |= get(MPM_PF_STRAIN_PLASTIC, 6, [](MaterialPoint_BC *p, int c, double x) { p->d6_Strain_Plastic[c] = x; });
    rc |= get(MPM_PF_BACKSTRESS, 1, [](MaterialPoint_BC *p, int, double x) { p->d_BackStress_Isotropic = x; });
    rc |= get(MPM_PF_ENERGY_STRAIN, 1, [](MaterialPoint_BC *p, int, double x) { p->d_Energy_Strain = x; });
    rc |= get(MPM_PF_ENERGY_STRAIN_MAX, 1,
              [](MaterialPoint_BC *p, int, double x) { p->d_Energy_Strain_MaxHistory = x; });
    rc |= get(MPM_PF_ENERGY_STRAIN_TI, 1,
              [](MaterialPoint_BC *p, int, double x) { p->d_Energy_Strain_TimeIntegral = x; });
    rc |= get(MPM_PF_ENERGY_PLASTIC, 1, [](MaterialPoint_BC *p, int, double x) { p->d_Energy_Plastic = x; });
    if (rc != MPM_OK) return fail(rc);
    b_ParticlesFresh = true;
    return MPM_OK;
}

int PhysicsEngine::refresh_grid() {
    if (b_GridFresh || !gpu || !b_Uploaded) return MPM_OK;
    const int64_t nn = (int64_t)allGridPoint.size();
    std::vector<double> buf;
    auto get = [&](int32_t field, int nc, auto set) -> int {
        buf.resize((size_t)nc * nn);
        const int rc = mpm_get_node_field(gpu, field, buf.data(), nn);
        if (rc != MPM_OK) return rc;
        for (int64_t i = 0; i < nn; i++)
            for (int c = 0; c < nc; c++) set(allGridPoint[i], c, buf[(size_t)c * nn + i]);
        return MPM_OK;
    };
    auto v3 = [](dvec3 &v, int c, double x) { (c == 0 ? v.x : (c == 1 ? v.y : v.z)) = x; };
    int rc = MPM_OK;
    rc |= get(MPM_NF_MASS, 1, [](GridPoint *g, int, double x) { g->d_Mass = x; });
    rc |= get(MPM_NF_MOMENTUM, 3, [&](GridPoint *g, int c, double x) { v3(g->d3_Momentum, c, x); });
    rc |= get(MPM_NF_VELOCITY, 3, [&](GridPoint *g, int c, double x) { v3(g->d3_Velocity, c, x); });
    rc |= get(MPM_NF_FORCE, 3, [&](GridPoint *g, int c, double x) { v3(g->d3_Force, c, x); });
    rc |= get(MPM_NF_FORCE_TEMP, 3, [&](GridPoint *g, int c, double x) { v3(g->d3_Force_Temp, c, x); });
    rc |= get(MPM_NF_ACTIVE, 1, [](GridPoint *g, int, double x) { g->b_Active = x != 0.0; });
    if (rc != MPM_OK) return fail(rc);
    b_GridFresh = true;
    return MPM_OK;
}

std::vector<MaterialPoint_BC *> PhysicsEngine::getMaterialPoints() {
    refresh_particles();
    return allMaterialPoint;
}

std::vector<GridPoint *> PhysicsEngine::getGridPoints() {
    refresh_grid();
    return allGridPoint;
}

}  // namespace mpmhost
