/*
 * mpm_physics.h — C++ host side of the MI355X MPM timestep.
 *
 * The classes the reference's scene and driver code talk to, with their
 * names, members and meaning (paths are in Sina-mon/code_MPM):
 *   Material_BC, MaterialPoint_BC, GridPoint       (host objects, AoS, as built by scene code)
 *   GridPoint_Factory, MaterialPoint_Factory_Classic_CC, TimeLine
 *   GridPoint_Mediator                             (shape functions on the host, for scene code)
 *   ConstitutiveRelation                           (return maps, evaluated on the GPU)
 *   PhysicsEngine                                  (the step loops, on the GPU)
 * over the C ABI of include/mpm.h.  Scene code fills allMaterialPoint,
 * allGridPoint, v_allMaterial, the marked lists and the time settings as it
 * does against the reference; the first runSimulation_* call uploads them;
 * the steps run on the device; getMaterialPoints() / getGridPoints() refresh
 * the host objects from the device when steps have run since the last
 * refresh.  After changing host objects between runs, call world_changed()
 * so that the next run uploads them again.
 *
 * The reference's vector types come from glm, which is not part of this
 * build; dvec3 / ivec3 / bvec3 / dmat3 below carry the same components
 * (dmat3 column-major, m[col][row], as glm's).
 */
#ifndef MPM_AMD_PHYSICS_H
#define MPM_AMD_PHYSICS_H

#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/mpm.h"

namespace mpmhost {

struct dvec3 {
    double x = 0.0, y = 0.0, z = 0.0;
    dvec3() = default;
    dvec3(double a, double b, double c) : x(a), y(b), z(c) {}
};
inline dvec3 operator+(const dvec3 &a, const dvec3 &b) { return dvec3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline dvec3 operator*(double s, const dvec3 &a) { return dvec3(s * a.x, s * a.y, s * a.z); }

struct ivec3 {
    int x = 0, y = 0, z = 0;
    ivec3() = default;
    ivec3(int a, int b, int c) : x(a), y(b), z(c) {}
};

struct bvec3 {
    bool x = false, y = false, z = false;
    bvec3() = default;
    bvec3(bool a, bool b, bool c) : x(a), y(b), z(c) {}
};

struct dmat3 {
    double m[3][3] = {{1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}}; /* identity, as glm::dmat3(1) */
};

/* Material type codes, Material_BC.h:6-11. */
enum { __ELASTIC = MPM_MAT_ELASTIC, __NEOHOOKEAN = MPM_MAT_NEOHOOKEAN, __VISCOELASTIC = MPM_MAT_VISCOELASTIC,
       __PLASTIC = MPM_MAT_PLASTIC, __VONMISESHARDENING = MPM_MAT_VONMISESHARDENING, __IDEALGASS = MPM_MAT_GASS };

/* Material_BC.h:13-35. */
struct Material_BC {
    int i_ID = 0;
    int i_MaterialType = __ELASTIC;
    double d_Density = 0.0;
    double d_ElasticModulus = 0.0;
    double d_PoissonRatio = 0.0;
    double d_YieldStress = 0.0;
    double d_Viscosity = 0.0;
    double d_Hardening_Isotropic_C0 = 0.0;
    double d_Hardening_Isotropic_C1 = 0.0;
};

/* MaterialPoint_BC.h:25-85: the fields the timestep reads and writes. */
struct MaterialPoint_BC {
    virtual ~MaterialPoint_BC() = default;
    unsigned long int i_ID = 0;
    Material_BC *p_Material = nullptr;
    dvec3 d3_Position, d3_Velocity, d3_Force_External;
    double d_Mass = 0.0, d_Volume = 0.0, d_Volume_Initial = 0.0;
    dmat3 d33_DeformationGradient;
    dmat3 d33_DeformationGradient_Elastic;
    double d6_Strain[6] = {0, 0, 0, 0, 0, 0};
    double d6_Stress[6] = {0, 0, 0, 0, 0, 0};
    double d6_Strain_Plastic[6] = {0, 0, 0, 0, 0, 0};
    double d_BackStress_Isotropic = 0.0;
    double d_Energy_Strain = 0.0;
    double d_Energy_Strain_MaxHistory = 0.0;
    double d_Energy_Strain_TimeIntegral = 0.0;
    double d_Energy_Plastic = 0.0;
    bvec3 b3_DisplacementControl;
    float f_DisplacementControl_Multiplier = 1.0f;
};
/* MaterialPoint_Classic_CC.h: the classic point adds nothing the step uses. */
struct MaterialPoint_Classic_CC : MaterialPoint_BC {};

/* GridPoint.h:6-45. */
struct GridPoint {
    dvec3 d3_Position;
    bool b_Active = false;
    bvec3 b3_Fixed;
    double d_FrictionCoefficient = 0.0;
    double d_Mass = 0.0;
    dvec3 d3_Momentum, d3_Velocity, d3_Force, d3_Force_Temp;
};

/* GridPoint_Factory.h:12-104. */
class GridPoint_Factory {
  public:
    /* ix + iy * nx + iz * nx * ny, GridPoint_Factory.h:18-32. */
    static unsigned int getIndex(ivec3 index, ivec3 nodes);
    /* Nodes at (i * L) / cells, GridPoint_Factory.h:33-68; the caller owns them. */
    std::vector<GridPoint *> createGrid(dvec3 length, ivec3 cells);
};

/* MaterialPoint_Factory_Classic_CC.h:6-25. */
class MaterialPoint_Factory_Classic_CC {
  public:
    /* createMaterialPoint.cpp:4-11: a default point at the position. */
    MaterialPoint_BC *createMaterialPoint(dvec3 position);
    /* createDomain_Cuboid.cpp:4-57: points offset apart filling a cuboid,
     * generated a quarter at a time and mirrored about the centre (one layer
     * at z = centre when the z dimension is at most one offset). */
    std::vector<MaterialPoint_BC *> createDomain_Cuboid(dvec3 center, dvec3 dimension, double offset);
};

/* TimeLine.h:12-52. */
class TimeLine {
  public:
    void addTimePoint(double time, dvec3 velocity);
    /* Linear interpolation within the last interval [t_i, t_i+1) holding
     * the time; zero outside every interval. */
    dvec3 getVelocity(double time) const;
    std::vector<dvec3> v_Velocity;
    std::vector<double> v_Time;
};

/* GridPoint_Mediator.h:14-41: the adjacent nodes and the linear hat
 * functions of a position, on the host (the device has its own copy of the
 * same arithmetic inside the step kernels). */
class GridPoint_Mediator {
  public:
    dvec3 d3_Length_Cell;
    ivec3 i3_Node_Count;
    std::vector<unsigned int> v_adjacentGridPoints;
    double d_ShapeValue = 0.0;
    dvec3 d3_ShapeGradient;
    /* GridPoint_Mediator.cpp:4-54: up to 8 node indices, x outer, z inner. */
    void findAdjacentGridPoints(dvec3 position);
    /* GridPoint_Mediator.cpp:95-144. */
    void calculateBases_Classic(dvec3 material_point, dvec3 grid_point);
};

/* ConstitutiveRelation.h:16-62.  The per-point calls evaluate on the GPU
 * (mpm_constitutive_batch with one point: correct, but a device round trip
 * per call); calculateIncrement_Batch takes many points at once. */
class ConstitutiveRelation {
  public:
    double d6StressIncrement[6] = {0, 0, 0, 0, 0, 0};
    double d6PlasticStrainIncrement[6] = {0, 0, 0, 0, 0, 0};
    double dBackstress_IsotropicIncrement = 0.0;
    void calculateIncrement_Elastic(Material_BC *material, double strain_increment[6]);
    void calculateIncrement_Plastic(Material_BC *material, double stress[6], double strain_increment[6]);
    void calculateIncrement_VonMisesHardening(Material_BC *material, double backstress, double stress[6],
                                              double strain_increment[6]);
    /* n points, row-major n x 6 arrays; the material's type picks the map. */
    static int calculateIncrement_Batch(const Material_BC &material, int64_t n, const double *backstress,
                                        const double *stress, const double *strain_increment, double *stress_increment,
                                        double *plastic_strain_increment, double *backstress_increment,
                                        int device = 0);

  private:
    void run(Material_BC *material, double backstress, const double *stress, const double *strain_increment);
};

/* One console report (PhysicsEngine::reportConsole, PhysicsEngine.cpp:40-183):
 * the values, which monitor lists were non-empty, and the line as written. */
struct ConsoleReport {
    double time, position_y, force_y, momentum_y, kinetic_energy, strain_energy, plastic_energy, strain_y, stress_y;
    double runtime; /* d_Runtime_Total: wall seconds of the steps so far */
    bool has_position, has_momentum, has_stress, has_energy;
    std::string line;
};

/* Script(double, precision) of Definitions.h:46-60: scientific notation with
 * that many digits after the point and an explicit sign. */
std::string Script(double r, int precision);
/* Script(int) of Definitions.h:25-41: decimal, no padding. */
std::string Script(long long i);
/* The line reportConsole writes (PhysicsEngine.cpp:42-151): the
 * description, then tab-separated "name: value" columns for time and
 * runtime, Position_y when the displacement monitor is set, Force_y,
 * momentum_y, Strain_y / Stress_y, the three energies, and a newline.  The
 * principal-value columns are not produced (their monitor list is not part
 * of this build). */
std::string consoleLine(const ConsoleReport &r, const std::string &sDescription);

/* PhysicsEngine.h:30-122, the timestep part. */
class PhysicsEngine {
  public:
    explicit PhysicsEngine(int device = 0);
    virtual ~PhysicsEngine(); /* deletes every point, node and material, PhysicsEngine.cpp:4-38 */
    PhysicsEngine(const PhysicsEngine &) = delete; /* owns its objects and its device engine */
    PhysicsEngine &operator=(const PhysicsEngine &) = delete;

    /* The cantilever bar of this build's C0 workload: 64 x 8 x 8 cells of
     * h = 1.5625e-3 m, 2 x 2 x 2 points per cell at (i + 0.25) h and
     * (i + 0.75) h, steel (E 200 GPa, nu 0.3, rho 7850, elastic), the x = 0
     * node plane fixed, the last cell column moving at (0, 0, -1) m/s,
     * dt = 0.1 h / sqrt(E / rho).  Ends as the reference's scene functions
     * do: str_Log_FileName = ./bmp/log_<YYYYmmddHHMMSS>.txt and a first
     * console report headed by the scene's description. */
    void initializeWorld_Classic_Bar();

    /* The three step loops (PhysicsEngine_runSimulation_Classic_*.cpp): steps
     * of d_TimeIncrement_Maximum with the last one clamped to the request, a
     * console report when the interval has passed; 0 while d_Time <
     * d_TimeEnd, else 1.  -1 if the device reports an error (see
     * last_error()); the reference has no error path here. */
    int runSimulation_Classic_DoublePass_MPLocks(double dTimeIncrement_Total);
    int runSimulation_Classic_DoublePass_MPLocks_Contact(double dTimeIncrement_Total);
    int runSimulation_Classic_SinglePass_MP(double dTimeIncrement_Total);
    /* The loops' components (PhysicsEngine.h:57-63, with the inline pass 1
     * and remap) as the device step's phases, for scene code that drives a
     * step itself: MPM_PHASE_BIN (reset + adjacent-node search) opens a step
     * of dTimeIncrement in the loop `variant` (MPM_VARIANT_*), the others
     * follow in that loop's order (mpm_step_phase) and MPM_PHASE_G2P_STRAIN
     * closes it and advances d_Time.  0, or -1 with last_error() set. */
    int runSimulation_Phase(int32_t variant, int32_t phase, double dTimeIncrement);

    double getTime_Current() const { return d_Time; }
    double getTime_End() const { return d_TimeEnd; }
    double getTime_Increment() const { return d_TimeIncrement_Maximum; }
    double getTime_ConsoleInterval() const { return d_TimeConsole_Interval; }
    unsigned int getCount_MaterialPoint() const { return (unsigned int)allMaterialPoint.size(); }
    unsigned int getCount_GridPoint() const { return (unsigned int)allGridPoint.size(); }
    /* Copies of the pointer vectors (PhysicsEngine.h:66-78), the objects
     * refreshed from the device first. */
    std::vector<MaterialPoint_BC *> getMaterialPoints();
    std::vector<GridPoint *> getGridPoints();
    /* PhysicsEngine::reportConsole (PhysicsEngine.cpp:40-183): the
     * monitors now, as one line (consoleLine) printed to stdout when
     * b_Console_Print is set (the reference always prints) and appended to
     * str_Log_FileName when that is set; the report is also kept in
     * v_ConsoleLog. */
    ConsoleReport reportConsole(const std::string &sDescription = "");

    void world_changed() { b_Uploaded = false; }
    const char *last_error() const;
    mpm_engine *engine() { return gpu; }

    /* the world, as scene code fills it (PhysicsEngine.h:80-122) */
    dvec3 d3_Length_World;
    ivec3 i3_Cells;
    std::vector<MaterialPoint_BC *> allMaterialPoint;
    std::vector<GridPoint *> allGridPoint;
    std::vector<Material_BC *> v_allMaterial;
    std::vector<MaterialPoint_BC *> v_MarkedMaterialPoints_Displacement_Control;
    std::vector<MaterialPoint_BC *> v_MarkedMaterialPoints_Displacement_Monitor;
    std::vector<MaterialPoint_BC *> v_MarkedMaterialPoints_Momentum;
    std::vector<MaterialPoint_BC *> v_MarkedMaterialPoints_Monitor_Energy;
    std::vector<MaterialPoint_BC *> v_MarkedMaterialPoints_Stress_Monitor;
    TimeLine m_TimeLine;
    double d_Time = 0.0;
    double d_TimeIncrement_Maximum = 1.0e-3;
    double d_TimeEnd = 10.0;
    double d_TimeConsole_Interval = 0.0;
    double d_TimeConsole_Last = 0.0;
    double d_DampingCoefficient = 0.0;
    double d_Mass_Minimum = 0.0;
    int32_t i_SummationOrder = MPM_ORDER_PAIR;
    double d_Runtime_Total = 0.0;     /* PhysicsEngine.h:116, wall seconds of the steps */
    std::string str_Log_FileName;     /* PhysicsEngine.h:82; empty: no log file */
    bool b_Console_Print = true;
    std::vector<ConsoleReport> v_ConsoleLog;

  private:
    int run_loop(double total, int32_t variant);
    int upload();
    int refresh_particles();
    int refresh_grid();
    int fail(int rc);

    mpm_engine *gpu = nullptr;
    bool b_Uploaded = false;
    double d_PhaseIncrement = 0.0; /* the open phase-by-phase step's dt */
    bool b_ParticlesFresh = true, b_GridFresh = true;
    std::string s_Error;
};

}  // namespace mpmhost

#endif /* MPM_AMD_PHYSICS_H */
