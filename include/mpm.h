/*
 * mpm.h — C ABI of the MI355X-native Classic MPM timestep (libmpm_amd.so).
 *
 * One opaque engine per simulation.  All state lives in device memory as
 * structure-of-arrays FP64; the host passes plain pointers and sizes.  Every
 * entry point returns MPM_OK (0) or a negative MPM_ERR_* code and never calls
 * exit(); mpm_last_error() gives the message of the last failure.
 *
 * Each declaration names the reference interface it replaces (paths are in
 * Sina-mon/code_MPM).  Layouts:
 *   - particle field arrays are component-major: component c of particle i is
 *     src[c * n + i];
 *   - the deformation gradient has 9 components in glm column-major order,
 *     component (3 * col + row) = F[col][row];
 *   - Voigt order is [xx, yy, zz, xy, yz, zx]; strains carry engineering
 *     shears (gamma = 2 eps), stresses tensor shears;
 *   - nodes are indexed ix + iy * nx + iz * nx * ny with n = cells + 1 per
 *     axis (GridPoint_Factory.h:18-22), node position (i * L) / cells.
 *
 * Threading: an engine is not re-entrant (neither is the reference's
 * PhysicsEngine, PhysicsEngine.h:53); different engines may be driven from
 * different host threads.  Each engine owns one HIP stream.
 */
#ifndef MPM_AMD_MPM_H
#define MPM_AMD_MPM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mpm_engine mpm_engine;

enum {
    MPM_OK = 0,
    MPM_ERR_ARG = -1,     /* bad argument (null pointer, size mismatch, ...) */
    MPM_ERR_HIP = -2,     /* a HIP runtime call failed                       */
    MPM_ERR_STATE = -3,   /* call out of order (e.g. step before set_grid)   */
    MPM_ERR_NOMEM = -4,   /* device allocation failed                        */
    MPM_ERR_RANGE = -5    /* size beyond what the engine supports            */
};

/* Material type codes, Material_BC.h:6-11. Unknown codes use the
 * perfectly-plastic map, as PhysicsEngine_runSimulation_Classic_DoublePass_MPLocks.cpp:238-239. */
enum {
    MPM_MAT_ELASTIC = 0,
    MPM_MAT_NEOHOOKEAN = 1,
    MPM_MAT_VISCOELASTIC = 2,
    MPM_MAT_PLASTIC = 3,
    MPM_MAT_VONMISESHARDENING = 4,
    MPM_MAT_GASS = 5
};

/* Material_BC (Material_BC.h:13-35). */
typedef struct mpm_material {
    int32_t type;
    int32_t id;
    double density;
    double elastic_modulus;
    double poisson_ratio;
    double yield_stress;
    double viscosity;
    double hardening_c0;
    double hardening_c1;
} mpm_material;

/* Particle fields (MaterialPoint_BC, MaterialPoint_BC.h:25-85), FP64. */
enum {
    MPM_PF_POSITION = 0,            /* 3 */
    MPM_PF_VELOCITY = 1,            /* 3 */
    MPM_PF_FORCE_EXTERNAL = 2,      /* 3 */
    MPM_PF_MASS = 3,                /* 1 */
    MPM_PF_VOLUME = 4,              /* 1 */
    MPM_PF_VOLUME_INITIAL = 5,      /* 1 */
    MPM_PF_DEFORMATION_GRADIENT = 6,/* 9 */
    MPM_PF_STRAIN = 7,              /* 6 */
    MPM_PF_STRESS = 8,              /* 6 */
    MPM_PF_STRAIN_PLASTIC = 9,      /* 6 */
    MPM_PF_BACKSTRESS = 10,         /* 1  d_BackStress_Isotropic */
    MPM_PF_ENERGY_STRAIN = 11,      /* 1 */
    MPM_PF_ENERGY_STRAIN_MAX = 12,  /* 1  d_Energy_Strain_MaxHistory */
    MPM_PF_ENERGY_STRAIN_TI = 13,   /* 1  d_Energy_Strain_TimeIntegral */
    MPM_PF_ENERGY_PLASTIC = 14,     /* 1 */
    MPM_PF_DEFORMATION_GRADIENT_ELASTIC = 15, /* 9  d33_DeformationGradient_Elastic (MaterialPoint_BC.h:65),
                                                 used by the large-strain pass only */
    MPM_PF_COUNT = 16
};

/* Node fields (GridPoint, GridPoint.h:6-45), FP64, state after the last step.
 * Nodes no particle touched in the last step read as zero, as after
 * runSimulation_ResetGrid (PhysicsEngine_runSimulation_Classic_DoublePass_MPLocks.cpp:4-22). */
enum {
    MPM_NF_MASS = 0,        /* 1 */
    MPM_NF_MOMENTUM = 1,    /* 3 */
    MPM_NF_VELOCITY = 2,    /* 3 */
    MPM_NF_FORCE = 3,       /* 3 */
    MPM_NF_FORCE_TEMP = 4,  /* 3  support reaction; only y is ever non-zero */
    MPM_NF_ACTIVE = 5,      /* 1  1.0 if touched in the last step */
    MPM_NF_COUNT = 6
};

/* Monitor lists of PhysicsEngine (PhysicsEngine.h:92-98). */
enum {
    MPM_MON_DISPLACEMENT = 0, /* v_MarkedMaterialPoints_Displacement_Monitor */
    MPM_MON_MOMENTUM = 1,     /* v_MarkedMaterialPoints_Momentum */
    MPM_MON_ENERGY = 2,       /* v_MarkedMaterialPoints_Monitor_Energy */
    MPM_MON_STRESS = 3,       /* v_MarkedMaterialPoints_Stress_Monitor */
    MPM_MON_COUNT = 4
};

/* reportConsole columns (PhysicsEngine.cpp:40-183) without the principal
 * monitor.  A column whose list is empty reports has_* = 0. */
typedef struct mpm_monitors {
    double time;
    double position_y;      /* mean platen y, PhysicsEngine.cpp:48-59 */
    double force_y;         /* sum of Force_Temp.y, PhysicsEngine.cpp:61-70 */
    double momentum_y;      /* PhysicsEngine.cpp:72-80 */
    double kinetic_energy;  /* PhysicsEngine.cpp:140-156 */
    double strain_energy;
    double plastic_energy;
    double strain_y;        /* mean over the stress monitor, PhysicsEngine.cpp:122-138 */
    double stress_y;
    int32_t has_position;
    int32_t has_momentum;
    int32_t has_energy;
    int32_t has_stress;
} mpm_monitors;

/* Per-engine health counters, cumulative since the last reset.  The
 * reference checks none of these; a null material prints and skips the
 * particle (PhysicsEngine_runSimulation_Classic_DoublePass_MPLocks.cpp:226-230). */
typedef struct mpm_health {
    int64_t steps;
    int64_t nonfinite_particles;  /* particle-steps ending with a non-finite position or stress */
    int64_t outside_grid;         /* particle-steps whose adjacent nodes all lie off the grid */
    int64_t null_material;        /* particle-steps skipped for a missing material */
    int64_t plastic_points;       /* particle-steps whose return map yielded */
    int64_t surrogate_calls;      /* particle-steps sent to the learned return map */
    int64_t surrogate_fallbacks;  /* of those, answered by the exact map instead */
} mpm_health;

/* --- lifetime ------------------------------------------------------------ */

/* PhysicsEngine::PhysicsEngine() (PhysicsEngine.h:33). device = HIP ordinal. */
int mpm_create(mpm_engine **out, int32_t device);
/* PhysicsEngine::~PhysicsEngine() (PhysicsEngine.cpp:4-38). */
int mpm_destroy(mpm_engine *e);
/* Message of the last failing call on e (or of the last failing mpm_create
 * when e is NULL).  Never NULL. */
const char *mpm_last_error(const mpm_engine *e);

/* --- world set-up (replaces the scene code's direct member writes) -------- */

/* GridPoint_Factory::createGrid (GridPoint_Factory.h:33-68) plus the
 * mediator's d3_Length_Cell / i3_Node_Count (GridPoint_Mediator.h:25-28).
 * Resets all node state and boundary conditions. */
int mpm_set_grid(mpm_engine *e, const double length[3], const int32_t cells[3]);
/* GridPoint::b3_Fixed (bit0 x, bit1 y, bit2 z) and d_FrictionCoefficient
 * (GridPoint.h:20-23) for every node; either pointer may be NULL (all zero). */
int mpm_set_node_boundary(mpm_engine *e, const uint8_t *fixed_mask,
                          const double *friction, int64_t n_nodes);
/* PhysicsEngine::v_allMaterial (PhysicsEngine.h:90). */
int mpm_set_materials(mpm_engine *e, const mpm_material *materials, int32_t n);
/* Resizes the particle set; every particle gets the MaterialPoint_BC
 * defaults (F = I, everything else 0, multiplier 1, no material). */
int mpm_set_particle_count(mpm_engine *e, int64_t n);
int mpm_get_particle_count(const mpm_engine *e, int64_t *n);
/* Host -> device copy of one particle field (n must equal the particle count). */
int mpm_set_particle_field(mpm_engine *e, int32_t field, const double *src, int64_t n);
/* Device -> host copy; the lazy mirror behind PhysicsEngine::getMaterialPoints (PhysicsEngine.h:76). */
int mpm_get_particle_field(mpm_engine *e, int32_t field, double *dst, int64_t n);
/* MaterialPoint_BC::p_Material as an index into the material table; -1 = NULL. */
int mpm_set_particle_material(mpm_engine *e, const int32_t *material, int64_t n);
/* v_MarkedMaterialPoints_Displacement_Control (PhysicsEngine.h:93) in list
 * order, with b3_DisplacementControl (bit0 x, bit1 y, bit2 z) and
 * f_DisplacementControl_Multiplier of each listed particle.  When two
 * listed particles write the same node the later one in the list wins, as in
 * runSimulation_DisplacementControl run by one thread. */
int mpm_set_displacement_control(mpm_engine *e, const int64_t *index, const uint8_t *axes,
                                 const float *multiplier, int64_t count);
/* The monitor lists of PhysicsEngine.h:92-98 (kind = MPM_MON_*), in list order. */
int mpm_set_monitor_list(mpm_engine *e, int32_t kind, const int64_t *index, int64_t count);
/* TimeLine::addTimePoint sequence (TimeLine.h:27-31); v3 = n x (vx, vy, vz). */
int mpm_set_timeline(mpm_engine *e, const double *time, const double *v3, int32_t n);
/* d_Time, d_TimeIncrement_Maximum, d_TimeEnd (PhysicsEngine.h:110-113). */
int mpm_set_time(mpm_engine *e, double time, double dt_max, double time_end);
/* d_DampingCoefficient and d_Mass_Minimum (PhysicsEngine.h:83-84). */
int mpm_set_damping(mpm_engine *e, double damping, double mass_minimum);

/* --- batches of independent instances (no counterpart in the reference) ---- */

/* Replaces the grid and particle set by `count` independent instances, each
 * a separate world as set up by one initializeWorld_* call (the BESO driver
 * builds one engine per iteration, Main.cpp:227-232).  Instance k has grid
 * lengths[3k..3k+2] / cells[3k..3k+2] and particle_counts[k] particles; all
 * instances share the particle, node and material arrays back to back
 * (particles of instance k start at the sum of the earlier counts, nodes
 * likewise, see mpm_get_instance_layout) and step together, one launch per
 * phase for all of them.  Particle, node-boundary and control uploads then
 * take the concatenated arrays; control and monitor lists use global
 * particle indices. */
int mpm_set_instances(mpm_engine *e, int32_t count, const double *lengths, const int32_t *cells,
                      const int64_t *particle_counts);
int mpm_get_instance_layout(const mpm_engine *e, int32_t instance, int64_t *particle_offset, int64_t *node_offset,
                            int64_t *node_count);
/* Per-instance TimeLine, d_Time / d_TimeIncrement_Maximum / d_TimeEnd. */
int mpm_set_instance_timeline(mpm_engine *e, int32_t instance, const double *time, const double *v3, int32_t n);
int mpm_set_instance_time(mpm_engine *e, int32_t instance, double time, double dt_max, double time_end);
int mpm_get_instance_time(const mpm_engine *e, int32_t instance, double *time);
/* reportConsole values of one instance. */
int mpm_get_instance_monitors(mpm_engine *e, int32_t instance, mpm_monitors *out);

/* --- node summation order -------------------------------------------------- */

/* All four orders are deterministic and use no floating-point atomics.
 * MPM_ORDER_PARTICLE: each node adds its particles' contributions in
 *   ascending particle index, the order of the reference's scatters run by
 *   one thread (PhysicsEngine_runSimulation_Classic_DoublePass_MPLocks.cpp:48-89,
 *   539-564); results equal that run bit for bit.  Steps in this order
 *   hold a device buffer of 320 bytes per particle (each particle's terms
 *   at its 8 corner nodes).
 * MPM_ORDER_CELL: each node adds 8 per-cell partial sums in a
 *   fixed neighbour order, each partial taken over the cell's particles in
 *   ascending index; the same terms in another association, so results
 *   differ from the particle order only by rounding, and much faster on the
 *   GPU (neighbouring lanes read the same particles).
 * MPM_ORDER_TILE: the binning cells (cell index + 1 per axis, cells -1 .. n)
 *   are grouped in fixed summation tiles of MPM_SUM_TILE_X x _Y x _Z; each
 *   node adds, over the tiles that hold its 8 neighbour cells in ascending
 *   tile index (x fastest), the tile's sum of those cells' per-cell partial
 *   sums taken in parity-class order, class 4 (e_x & 1) + 2 (e_y & 1) +
 *   (e_z & 1), each partial over the cell's particles in ascending index.
 *   The same terms in another association again; on the GPU a tile's node
 *   sums are accumulated in on-chip memory and only the sums of nodes on a
 *   tile face pass through device memory.  Steps in this order hold a
 *   device buffer of 16.1 KB per summation tile.
 * MPM_ORDER_PAIR (default): as the cell order, but each node first adds the partial
 *   sums of its two neighbour cells that differ only along x (lower x cell
 *   first), then adds the four pair sums in the cell order's neighbour
 *   order of their lower cells.  On the GPU the cells of a block are
 *   consecutive along x, so a pair's two partials meet in on-chip memory
 *   and a run of occupied cells along x passes one record per node and
 *   (y, z) corner through device memory instead of two. */
enum { MPM_ORDER_PARTICLE = 0, MPM_ORDER_CELL = 1, MPM_ORDER_TILE = 2, MPM_ORDER_PAIR = 3 };
#define MPM_SUM_TILE_X 8
#define MPM_SUM_TILE_Y 8
#define MPM_SUM_TILE_Z 4
int mpm_set_summation_order(mpm_engine *e, int32_t order);

/* --- timestep variant ------------------------------------------------------- */

/* Which of the reference's three step loops mpm_step / mpm_run execute:
 * MPM_VARIANT_DOUBLEPASS (default)
 *   runSimulation_Classic_DoublePass_MPLocks
 *   (PhysicsEngine_runSimulation_Classic_DoublePass_MPLocks.cpp:439-815);
 * MPM_VARIANT_DOUBLEPASS_CONTACT
 *   runSimulation_Classic_DoublePass_MPLocks_Contact
 *   (PhysicsEngine_runSimulation_Classic_DoublePass_MPLocks_Contact.cpp:4-566):
 *   momentum integrated before the boundary conditions and no friction;
 *   displacement control on all three axes after them; control particles
 *   skip both grid-to-particle passes and the remap and move by x += v dt;
 *   F += L F dt; energies from the new stress.  Its mass-gradient contact
 *   flags (:152-237) feed nothing back into the state and are not computed;
 * MPM_VARIANT_SINGLEPASS
 *   runSimulation_Classic_SinglePass_MP
 *   (PhysicsEngine_runSimulation_Classic_SinglePass_MP.cpp:4-404): USL;
 *   control particles are left out of the particle-to-grid sums; grid
 *   velocity is the mass-weighted particle velocity plus f/m dt, damping
 *   along it; one grid-to-particle pass; F += L F dt.  Per-particle
 *   material parameters (:328-345) are the particle's material's; a particle
 *   with no valid material gets the per-particle defaults (elastic, E = 0).
 * In every variant, "displacement-controlled particle" means a particle on
 * the mpm_set_displacement_control list. */
enum { MPM_VARIANT_DOUBLEPASS = 0, MPM_VARIANT_DOUBLEPASS_CONTACT = 1, MPM_VARIANT_SINGLEPASS = 2 };
int mpm_set_variant(mpm_engine *e, int32_t variant);
int mpm_get_variant(const mpm_engine *e, int32_t *variant);

/* Which grid-to-particle pass 2 the DoublePass loop runs:
 * MPM_STRAIN_SMALL (default)
 *   runSimulation_G2P_P2_SmallStrain
 *   (PhysicsEngine_runSimulation_Classic_DoublePass_MPLocks.cpp:175-280);
 * MPM_STRAIN_LARGE
 *   runSimulation_G2P_P2_LargeStrain (same file :281-437; the loop's call
 *   to it, :605, is commented out in the reference): logarithmic strains
 *   from 8-term series (ConstitutiveRelation_MatrixAlgebra.cpp:25-53), the
 *   return map on the elastic trial strain from zero stress, Fe = exp(the
 *   elastic strain), stress = returned stress / det F, strain = the log
 *   strain of F before the step.  A particle with no material gets zero
 *   stress and is otherwise left as the reference leaves it.
 * The large-strain pass exists for the DoublePass loop and the exact return
 * map only: a step with another variant, a surrogate or sample capture set
 * fails with MPM_ERR_STATE. */
enum { MPM_STRAIN_SMALL = 0, MPM_STRAIN_LARGE = 1 };
int mpm_set_strain_formulation(mpm_engine *e, int32_t formulation);
int mpm_get_strain_formulation(const mpm_engine *e, int32_t *formulation);

/* --- node tiles and the autotuner (no counterpart in the reference) -------- */

/* Nodes are processed in tiles of shape[0] x shape[1] x shape[2] = 64 nodes
 * per workgroup (default 32 x 2 x 1). */
int mpm_set_tile_shape(mpm_engine *e, const int32_t shape[3]);
int mpm_get_tile_shape(const mpm_engine *e, int32_t shape[3]);
/* Performs one step during which the particle-to-grid phase is launched
 * `reps` times under each candidate tile shape (it reads particle and bin
 * state and overwrites node state, so repeating it changes nothing) and
 * keeps the fastest shape; the step's results equal an ordinary mpm_step of
 * dt_max.  Returns the chosen shape and its particle-to-grid milliseconds. */
int mpm_autotune(mpm_engine *e, int32_t reps, int32_t shape[3], double *node_ms);

/* --- the hot path ---------------------------------------------------------- */

/* nsteps steps; each instance advances by its own dt_max.  The first call
 * for a given nsteps records the steps' kernels in a HIP graph (time and
 * the timeline velocity advance on the device, so no host work is needed
 * between steps); later calls replay it.  Set-up calls invalidate it. */
int mpm_step_graph(mpm_engine *e, int32_t nsteps);
/* nsteps steps of size dt, each the body of the while loop of
 * runSimulation_Classic_DoublePass_MPLocks
 * (PhysicsEngine_runSimulation_Classic_DoublePass_MPLocks.cpp:462-803):
 * time += dt and cycle += 1 after each. Asynchronous on the engine's stream.
 * A multi-instance engine takes dt <= 0 and steps each instance by its own
 * dt_max. */
int mpm_step(mpm_engine *e, double dt, int32_t nsteps);
/* runSimulation_Classic_DoublePass_MPLocks(dTimeIncrement_Total)
 * (PhysicsEngine_runSimulation_Classic_DoublePass_MPLocks.cpp:439-815)
 * without the console report: steps of d_TimeIncrement_Maximum, the last one
 * clamped to the remainder; *finished = 0 while time < time_end, else 1. */
int mpm_run(mpm_engine *e, double dt_total, int32_t *finished);
/* Step phases timed by mpm_step_timed, in launch order. */
enum {
    MPM_PHASE_BIN = 0,          /* cell binning (scan, scatter, in-cell sort) */
    MPM_PHASE_P2G_GRID = 1,     /* P2G + displacement control + grid integration */
    MPM_PHASE_G2P_VELOCITY = 2, /* G2P pass 1 */
    MPM_PHASE_REMAP_GRID = 3,   /* velocity remap + control + BC re-application */
    MPM_PHASE_G2P_STRAIN = 4,   /* G2P pass 2 + return map + schedule */
    MPM_PHASE_COUNT = 5
};
/* mpm_step with each phase bracketed by HIP events on the engine's stream;
 * phase_ms receives the summed milliseconds per phase.  Replaces the
 * reference's per-block wall-clock timers a_Runtime (PhysicsEngine.h:87,
 * PhysicsEngine_runSimulation_Classic_DoublePass_MPLocks.cpp:487-790).
 * Synchronises after every step. */
int mpm_step_timed(mpm_engine *e, double dt, int32_t nsteps, double phase_ms[MPM_PHASE_COUNT]);
/* One step taken phase by phase, for callers that run the step loop's
 * components themselves (PhysicsEngine.h:57-63: runSimulation_ResetGrid,
 * _FindAGPs, _M2G, _DisplacementControl, _IntegrateGrid and
 * _G2P_P2_SmallStrain / _LargeStrain, with the inline pass 1 and remap of
 * PhysicsEngine_runSimulation_Classic_DoublePass_MPLocks.cpp:462-803).  On
 * the device they come as these phases: MPM_PHASE_BIN (grid reset and the
 * adjacent-node search: cell binning) opens the step and takes its dt (as
 * mpm_step's; <= 0 for each instance's dt_max); then, in the loop's order,
 * MPM_PHASE_P2G_GRID (M2G, displacement control and grid integration),
 * MPM_PHASE_G2P_VELOCITY and MPM_PHASE_REMAP_GRID (DoublePass and contact
 * loops: pass 1, then the remap; single pass: the velocity sum, then its G2P
 * velocity), and MPM_PHASE_G2P_STRAIN, which closes the step (time += dt).
 * dt is ignored after the first phase.  The phases of one step equal one
 * mpm_step.  A phase out of order returns MPM_ERR_STATE; while a step is
 * open, set-up and other stepping calls return MPM_ERR_STATE (getters
 * work).  A failed launch abandons the step. */
int mpm_step_phase(mpm_engine *e, int32_t phase, double dt);
/* The same with one time per kernel (HIP events on the engine's stream),
 * summed over nsteps, in these slots.  A phase that runs one kernel in the
 * particle order (P2G, remap) books it in its first slot; with a surrogate
 * set, its network and post kernels count in the constitutive slot.  With
 * the exact return map (one instance or a batch), G2P pass 2 is one fused kernel
 * (kinematics and state update of each particle in one thread), booked in
 * the kinematics slot; the constitutive slot then times nothing. */
enum {
    MPM_KSLOT_BIN = 0,                /* scan, scatter, per-cell sort */
    MPM_KSLOT_P2G_PARTIALS = 1,       /* cell partials (cell order) or the node gather (particle order) */
    MPM_KSLOT_P2G_NODES = 2,          /* node sums + grid update (cell order) */
    MPM_KSLOT_G2P_VELOCITY = 3,
    MPM_KSLOT_REMAP_PARTIALS = 4,
    MPM_KSLOT_REMAP_NODES = 5,
    MPM_KSLOT_G2P_KINEMATICS = 6,     /* x, L, strain increment, F, V, control schedule, next cell */
    MPM_KSLOT_G2P_CONSTITUTIVE = 7,   /* return map, stress, energies */
    MPM_KSLOT_COUNT = 8
};
int mpm_step_timed_kernels(mpm_engine *e, double dt, int32_t nsteps, double kernel_ms[MPM_KSLOT_COUNT]);
/* Waits for the engine's stream. */
int mpm_synchronize(mpm_engine *e);

/* --- observation ------------------------------------------------------------ */

/* getTime_Current and the cycle counter i_TimeCycle (PhysicsEngine.h:67,113). */
int mpm_get_time(const mpm_engine *e, double *time, int64_t *cycle);
int mpm_get_node_count(const mpm_engine *e, int64_t *n);
int mpm_get_node_field(mpm_engine *e, int32_t field, double *dst, int64_t n);
/* reportConsole values (PhysicsEngine.cpp:40-183), reduced on the device in
 * the reference's summation order. */
int mpm_get_monitors(mpm_engine *e, mpm_monitors *out);
int mpm_get_health(mpm_engine *e, mpm_health *out);
int mpm_reset_health(mpm_engine *e);

/* --- stand-alone constitutive update ----------------------------------------- */

/* ConstitutiveRelation::calculateIncrement_{Elastic,Plastic,VonMisesHardening}
 * (ConstitutiveRelation.h:31-34; ConstitutiveRelation_calculateIncrement_*.cpp)
 * for n independent points on device `device`, host arrays in and out:
 * stress / dstrain / dstress / dstrain_plastic are n x 6 row-major,
 * backstress / dbackstress are n.  The material type selects the map as the
 * step does. */
int mpm_constitutive_batch(int32_t device, const mpm_material *material, const double *backstress,
                           const double *stress, const double *dstrain, double *dstress,
                           double *dstrain_plastic, double *dbackstress, int64_t n);

/* --- BESO objective filter ---------------------------------------------------- */

/* Canvas2D_CC::filterObjective_Smooth (Canvas2D_CC.cpp:302-341) on device
 * `device`, host arrays in and out, for an nx x ny voxel canvas (voxel
 * index x + y * nx, centre (i + 0.5) * offset per axis, Canvas2D_CC.cpp:13-28):
 * each active voxel's objective becomes the weighted mean, weights
 * (radius - d) / radius, of the objectives of the active ESO voxels closer
 * than radius, summed in ascending voxel index (the reference's order), so
 * the result equals the reference's double for double.  Inactive voxels
 * keep their objective.  A voxel with no such neighbour gets 0/0 (NaN), as
 * in the reference. */
int mpm_eso_filter(int32_t device, const uint8_t *active, const uint8_t *eso, const double *objective, int32_t nx,
                   int32_t ny, double offset, double radius, double *out);

/* --- learned return-map surrogate (no counterpart in the reference) ---------- */

/* A 13 -> 64 -> 64 -> 64 -> 7 SiLU MLP with BF16 weights and FP32
 * accumulation on MFMA, standing in for
 * ConstitutiveRelation::calculateIncrement_VonMisesHardening /
 * calculateIncrement_Plastic (ConstitutiveRelation_calculateIncrement_*.cpp)
 * at trial states outside the yield surface.  Inputs (13, formed in FP64
 * from the trial state sigma + C : de, then standardised in FP32): the trial
 * deviator's direction s_tr / sqrt(J2) (6), sqrt(J2) / K, the overstress
 * (sqrt(J2) - K) / s_sigma, K / s_sigma, the trial pressure / s_sigma,
 * R / s_sigma, |de| / s_eps and sqrt(J2) / s_sigma, with K = K0 + R/sqrt(2),
 * s_sigma the stress scale and s_eps the strain scale.  Outputs (7,
 * standardised): the radial-scale correction alpha, the fraction of the
 * trial deviator the return removes (the returned increment is
 * C : de - alpha s_tr), the equivalent plastic strain increment / s_eps, and
 * five unused outputs trained to zero.  Parameters:
 * mpm_surrogate_param_count() FP32 values,
 * W1[16][64] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[64][16] b4[16]. */
typedef struct mpm_surrogate mpm_surrogate;
int64_t mpm_surrogate_param_count(void);
/* Glorot-uniform initial weights drawn from seed. */
int mpm_surrogate_create(mpm_surrogate **out, int32_t device, uint64_t seed);
int mpm_surrogate_destroy(mpm_surrogate *s);
int mpm_surrogate_set_params(mpm_surrogate *s, const float *params, int64_t count);
int mpm_surrogate_get_params(mpm_surrogate *s, float *params, int64_t count);
/* Scales and standardisation; in_lo / in_hi bound the standardised inputs
 * (the training envelope): a point outside uses the exact map. */
int mpm_surrogate_set_normalization(mpm_surrogate *s, double stress_scale, double strain_scale,
                                    const float in_mean[13], const float in_std[13], const float in_lo[13],
                                    const float in_hi[13], const float out_mean[7], const float out_std[7]);
/* Standardised rows, host arrays: x n x 13 -> y n x 7. */
int mpm_surrogate_forward(mpm_surrogate *s, const float *x, float *y, int64_t n);
/* Gradient of mean((y - t)^2) over n rows and the 7 outputs. */
int mpm_surrogate_gradient(mpm_surrogate *s, const float *x, const float *t, int64_t n, float *grad, float *loss);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight training or serving loops,
 * harnesses that forbid hipMalloc): mpm_surrogate_gradient on rows already
 * in device memory, allocating nothing.  x (n x 13), t (n x 7), grad
 * (mpm_surrogate_param_count() floats) and loss (1 float, may be NULL) are
 * device pointers on the surrogate's device.  The per-chunk gradient
 * partials live in a fixed library array (5.35 MB per GPU, held until the
 * process ends), so n is limited to MPM_SURROGATE_NOALLOC_MAX_ROWS; a larger
 * n returns MPM_ERR_RANGE.  All calls on one GPU share that array: calls
 * that overlap on one GPU (several host threads or surrogates) race.
 * Returns after the result is written (waits for the surrogate's stream).
 * Equal, float for float, to mpm_surrogate_gradient on the same rows. */
#define MPM_SURROGATE_NOALLOC_MAX_ROWS 8192
int mpm_surrogate_gradient_noalloc(mpm_surrogate *s, const float *x, const float *t, int64_t n, float *grad,
                                   float *loss);
/* Copies a standardised training set to the device. */
int mpm_surrogate_load_training_set(mpm_surrogate *s, const float *x, const float *t, int64_t n);
/* Adam on random 64-row-aligned windows of `batch` rows; loss_history
 * (optional) receives each step's batch loss. */
int mpm_surrogate_train(mpm_surrogate *s, int32_t steps, int32_t batch, float lr, uint64_t seed, float *loss_history);

/* Routes the return map of material `material` through the surrogate (NULL
 * restores the exact map).  A prediction is accepted when its inputs lie in
 * the training envelope and |sqrt(J2(sigma_new)) - (K0 + R/sqrt(2))| <=
 * residual_tolerance * sigma_y; otherwise the exact map runs. */
int mpm_set_surrogate(mpm_engine *e, mpm_surrogate *s, int32_t material, double residual_tolerance);
/* Appends every yielding point of the exact map (sigma 6, dstrain 6, R,
 * material index: 14 doubles) to a device buffer of `capacity` rows; 0 stops. */
int mpm_set_sample_capture(mpm_engine *e, int64_t capacity);
/* Copies up to `capacity` captured rows; *count = rows produced (may exceed the buffer). */
int mpm_get_captured_samples(mpm_engine *e, double *rows, int64_t capacity, int64_t *count);

#ifdef __cplusplus
}
#endif

#endif /* MPM_AMD_MPM_H */
