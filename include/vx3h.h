/*
 * vx3h.h -- C ABI of the MI355X Voxelyze3 physics engine.
 *
 * This is the seam between the C++ host layer (VXA/VXD loader, simulation
 * manager, worker CLI, serving loop) and the HIP device engine.  Everything
 * crossing it is plain C: opaque handles, POD structs, pointers and sizes.
 * All functions return a vx3h_status; none aborts.  The last error message of
 * the calling thread is available from vx3h_last_error().
 *
 * Reference interfaces each entry point replaces (paths relative to the
 * gpuVoxels tree):
 *   vx3h_vxa_parse        VX3_SimulationManager::readVXD
 *                         (Voxelyze3/src/VX3/VX3_SimulationManager.cu:277-380)
 *                         + CVX_Sim::ReadVXA/Import (Voxelyze3/src/VXA/VX_Sim.cpp:59-183)
 *   vx3h_batch_create     VX3_SimulationManager ctor + device allocation
 *                         (VX3_SimulationManager.cu:123-130,283)
 *   vx3h_sim_admit        VX3_VoxelyzeKernel(CVX_Sim*) host ctor + cudaMemcpy
 *                         (Voxelyze3/src/VX3/VX3_VoxelyzeKernel.cu:27-105,
 *                          VX3_SimulationManager.cu:377)
 *   vx3h_batch_step       CUDA_Simulation loop body calling
 *                         VX3_VoxelyzeKernel::doTimeStep
 *                         (VX3_SimulationManager.cu:62-69, VX3_VoxelyzeKernel.cu:237-359)
 *   vx3h_sim_step         VX3_VoxelyzeKernel::doTimeStep(float dt) of one simulation
 *                         (VX3_VoxelyzeKernel.cuh:30, VX3_VoxelyzeKernel.cu:237-359)
 *   vx3h_batch_poll       VX3_SimulationManager::collectResults
 *                         (VX3_SimulationManager.cu:416-458)
 *   vx3h_sim_read_state   direct reads of the device voxel/link arrays done by
 *                         collectResults (VX3_SimulationManager.cu:435-452)
 *   vx3h_sim_retire       per-simulation cleanup (VX3_VoxelyzeKernel.cu:107-119)
 *   vx3h_batch_destroy    VX3_SimulationManager dtor (VX3_SimulationManager.cu:131-136)
 *   vx3h_report_xml       report writer of vx3_node_worker
 *                         (Voxelyze3/src/Executables/vx3_node_worker.cu:98-141)
 */
#ifndef VX3H_H
#define VX3H_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
typedef enum vx3h_status {
    VX3H_OK = 0,
    VX3H_EINVAL = 1,      /* bad argument or malformed model */
    VX3H_ENOMEM = 2,      /* the model does not fit the batch's device memory budget */
    VX3H_EDEVICE = 3,     /* HIP runtime error */
    VX3H_EPARSE = 4,      /* VXA/VXD could not be parsed */
    VX3H_EUNSUPPORTED = 5,/* VXA enables a feature this engine does not implement */
    VX3H_ENOTFOUND = 6    /* slot or handle not found */
} vx3h_status;

const char* vx3h_last_error(void);
vx3h_status vx3h_device_count(int* count);

/* ---------------------------------------------------------------- model description
 * The raw model a simulation is built from: the same information CVX_Sim::Import
 * consumes (palette, lattice, structure, phase offsets, boundary conditions,
 * environment) plus the VX3-only simulator keys read in readVXD.  Materials hold
 * the VXA palette values before Voxelyze's setters clamp them; the engine applies
 * the setters itself (CVX_Sim::CopyMat, Voxelyze3/src/VXA/VX_Sim.cpp:368-420).
 */
#define VX3H_MAX_CURVE 16      /* stress/strain data points per material (MDL_DATA) */
#define VX3H_MAX_TOKENS 256    /* MathTree tokens per program */

enum { VX3H_MDL_LINEAR = 0, VX3H_MDL_LINEAR_FAIL = 1, VX3H_MDL_BILINEAR = 2, VX3H_MDL_DATA = 3 };

typedef struct vx3h_material_desc {
    int    mat_model;           /* VX3H_MDL_*  (VXA/VX_Object.h:346) */
    double elastic_mod;         /* Pa */
    double plastic_mod;         /* Pa, bilinear */
    double yield_stress;        /* Pa, bilinear */
    double fail_stress;         /* Pa, LINEAR_FAIL / BILINEAR; 0 = none given */
    double density;             /* kg/m^3 */
    double poissons_ratio;
    double cte;                 /* 1/degC */
    double u_static, u_dynamic; /* floor friction coefficients */
    double zeta_internal;       /* BondDampingZ */
    double zeta_global;         /* SlowDampingZ */
    double zeta_collision;      /* ColDampingZ */
    int    n_data;              /* MDL_DATA points (strain_data/stress_data) */
    double strain_data[VX3H_MAX_CURVE];
    double stress_data[VX3H_MAX_CURVE];
    int    is_measured;         /* counts toward center of mass and fitness */
    int    fixed;               /* VX3 material-fixed voxels do not move */
    double thermal_on_after;    /* TurnOnThermalExpansionAfterThisManySeconds */
    int    matid;               /* palette id reported in the report's <mats> */
    int    color[4];            /* Display Red/Green/Blue/Alpha as 0-255 (int)(x*255), clamped
                                   (VXA/VX_Object.h:419-425, old/VX_Material.cpp:245-283) */
    double cilia;               /* Cilia: scale of the voxel's cilia force (VX3_Voxel.cu:376) */
    double cilia_on_after;      /* TurnOnCiliaAfterThisManySeconds */
    double remove_after;        /* RemoveFromSimulationAfterThisManySeconds (secondary experiment) */
    /* signals (VX3_Voxel.cu:279-348; VXA defaults VX_Object.cpp:1420-1425) */
    int    is_pacemaker;        /* isPaceMaker [0]: fires a signal of 100 every pacemaker_period */
    double pacemaker_period;    /* PaceMakerPeriod [0] */
    double signal_value_decay;  /* signalValueDecay [0.9]: value passed on = value received * decay */
    double signal_time_delay;   /* signalTimeDelay [0.03] s before a neighbour receives it */
    double inactive_period;     /* inactivePeriod [0.03] s deaf after receiving */
    int    is_target;           /* isTarget [0]: contacts with other materials count as hits
                                   (VX3_VoxelyzeKernel.cu:711-722) */
    int    sticky;              /* Sticky [0]: surface voxels of this material attach to each other
                                   (VX3_VoxelyzeKernel.cu:739) */
} vx3h_material_desc;

/* degree-of-freedom bits (Voxelyze3/src/old/types.h:48-55) */
enum { VX3H_DOF_TX = 1, VX3H_DOF_TY = 2, VX3H_DOF_TZ = 4, VX3H_DOF_RX = 8, VX3H_DOF_RY = 16, VX3H_DOF_RZ = 32 };

typedef struct vx3h_external {
    int    voxel;        /* index in the voxel list */
    int    dof_fixed;    /* VX3H_DOF_* bits */
    double translate[3]; /* prescribed displacement of fixed translational dofs (m) */
    double rotate[3];    /* prescribed rotation vector of fixed rotational dofs (rad) */
    double force[3];     /* N, global frame */
    double moment[3];    /* N*m, global frame */
} vx3h_external;

typedef struct vx3h_token { int op; double value; } vx3h_token;   /* VX3_MathTreeToken */

typedef struct vx3h_sim_desc {
    double voxel_size;                      /* lattice dimension (m) */
    int    n_materials;
    const vx3h_material_desc* materials;
    int    n_voxels;
    const int*    ijk;       /* 3*n_voxels lattice indices, in insertion order */
    const int*    material;  /* n_voxels indices into materials */
    const double* phase;     /* n_voxels phase offsets in cycles, or NULL for all 0 */
    int    n_externals;
    const vx3h_external* externals;
    /* environment */
    float  gravity_mult;     /* gravity in g (CVoxelyze::grav); force = -m*9.80665*g */
    int    floor_enabled;
    int    temp_enabled;     /* Environment/Thermal/TempEnabled */
    int    vary_temp_enabled;
    double temp_base, temp_amplitude, temp_period;
    int    enable_expansion; /* VX3: allow positive temperature (expansion) */
    /* simulator */
    double dt_frac;
    double dt_fixed;         /* > 0: every step uses this dt (s) instead of dt_frac x recommended */
    int    volume_effects;   /* VolumeEffectsEnabled; 0 zeroes every Poisson's ratio */
    int    collisions_enabled;
    double bounding_radius;  /* voxel units (default 0.75) */
    double watch_distance;   /* voxel units (default 1.0) */
    int    n_stop_tokens;    const vx3h_token* stop_tokens;     /* StopConditionFormula */
    int    n_fitness_tokens; const vx3h_token* fitness_tokens;  /* FitnessFunction */
    int    max_steps;        /* step budget; 0 = unlimited (the reference caps at 1e6) */
    int    save_position_of_all_voxels;
    /* VoxCAD's stop-condition family and equilibrium mode (VoxCAD/src/VX1/VX_Sim.cpp:862-891,
     * 1015-1160; StopConditionType of VXA/VX_Sim.cpp:260-277).  The Voxelyze3 driver reads
     * these keys but ignores them (VX3_VoxelyzeKernel.cu:162-182), so the loader leaves
     * them 0 here and reports the VXA's values in vx3h_vxa_info; a caller that wants
     * VoxCAD's behaviour copies them in.  Checked after the MathTree stop condition;
     * none of them stops a simulation before 1000 steps.  Simulations using 4-6 or
     * equilibrium mode run on the stream path. */
    int    stop_condition_type;   /* VX3H_SC_* */
    double stop_condition_value;
    int    equilibrium_mode;      /* zero all motion whenever the kinetic energy falls */
    /* VX3 cilia (Simulator.EnableCilia; VX3_VoxelyzeKernel.cu:296-303,846-859): each
     * surface voxel whose material has Cilia != 0, from TurnOnCiliaAfterThisManySeconds
     * on, is pushed by orient.RotateVec3D(base + localSignal*shift) * Cilia, with
     * localSignal the voxel's signal (0 unless enable_signals).  The arrays hold 3 per voxel
     * (Structure.BaseCiliaForce / ShiftCiliaForce) or are NULL for zeros. */
    int    enable_cilia;
    const double* base_cilia_force;
    const double* shift_cilia_force;
    /* VX3 force field (Simulator.ForceField.{x,y,z}_forcefield; VX3_Voxel.cu:178-184):
     * a MathTree per axis of the voxel position, added to every moving voxel's force;
     * 0 tokens = no force on that axis. */
    int    n_force_field_tokens[3];
    const vx3h_token* force_field_tokens[3];
    /* VX3 detach (Simulator.AttachDetach.EnableDetach; gpu_update_detach,
     * VX3_VoxelyzeKernel.cu:946-967): a link whose maximum strain passes its material's
     * failure strain leaves both voxels after the link update of that step, before the
     * voxel update.  With collisions, the nearby lists (voxels within the link-hop radius,
     * never paired) follow the links present at the next step, as after every topology
     * change (attach, detach, voxel removal). */
    int    enable_detach;
    /* VX3 secondary experiment (Simulator.SecondaryExperiment; VX3_VoxelyzeKernel.cu:
     * 336-350, 365-397): at the end of each step, every material whose remove_after is
     * positive and earlier than the step's start time loses its voxels (frozen in
     * place, still counted in the center of mass) and their links; once the start time
     * passes reinit_after (ReinitializeInitialPositionAfterThisManySeconds), the
     * initial center of mass becomes the last sampled one and the initial positions
     * the current ones. */
    int    secondary_experiment;
    double reinit_after;
    /* VX3 signals (Simulator.EnableSignals; VX3_Voxel.cu:271-348): after its update each
     * moving voxel passes a pending signal to its linked neighbours, fires if it is a
     * pacemaker, and lets its local signal decay; the local signal scales the cilia
     * shift force.  The reference's voxels write their neighbours' state concurrently;
     * here every voxel first updates its own state and posts what it sends, then
     * every voxel takes what its neighbours posted, in link-direction order
     * (X+, X-, Y+, Y-, Z+, Z-). */
    int    enable_signals;
    /* With collisions, a surface voxel pair of one target and one other material whose
     * centres are within 0.625 x the sum of their sizes counts one hit per step
     * (collisionCount, the MathTree variable hit) and, with signals, sends 100 to the
     * non-target voxel (handle_collision_attachment, VX3_VoxelyzeKernel.cu:682-722).
     * Pairs come from this engine's contact lists (the watch radius, 2.5 voxels by
     * default, covers the 1.25-voxel envelope; voxels within the nearby link-hop
     * radius of each other are not paired). */
    /* VX3 attach (Simulator.AttachDetach.EnableAttach; handle_collision_attachment,
     * VX3_VoxelyzeKernel.cu:724-826): at each step's start, two surface voxels of one
     * sticky, non-fixed material whose centres are within 0.625 x the sum of their
     * sizes, which are not within five links of each other and whose attach conditions
     * both hold, get a new link along the dominant axis of their offset in the first
     * voxel's frame, if both faces are free.  The new link carries no force in the step
     * it is made and 1/100 of its force for the next safety_guard steps.  Pairs are
     * taken in ascending (higher, lower) voxel order, each seeing the links made before
     * it.  Attach conditions: up to five MathTrees of the voxel position, hit, t and
     * angle, evaluated after each voxel update; all must be > 0 (0 tokens = no
     * condition). */
    int    enable_attach;
    int    n_attach_tokens[5];
    const vx3h_token* attach_tokens[5];
    int    safety_guard;          /* SafetyGuard [500] */
    /* Initial linear velocity of every voxel (3 per voxel, m/s, global frame), or NULL
     * for bodies at rest: the voxel starts with linear momentum mass * velocity.  No VXA
     * key carries it (CVX_Voxel::reset zeroes linMom, old/VX_Voxel.cpp:48-57); it is
     * for synthetic workloads such as bodies thrown into an arena. */
    const double* init_velocity;
    /* MaxDistInVoxelLengthsToCountAsPair [0] (VX3_SimulationManager.cu:358): when > 0,
     * at every center-of-mass sample the voxels of isTarget materials are paired, and
     * numClosePairs counts the pairs closer than this many voxel sizes while
     * targetCloseness sums 1/distance over all pairs (computeTargetCloseness,
     * VX3_VoxelyzeKernel.cu:545-562); both feed the stop, fitness, force-field and
     * attach-condition programs and the report. */
    double max_dist_close_pairs;
} vx3h_sim_desc;

enum {                         /* StopCondition (Voxelyze3/src/old/VX_Enums.h:55-63) */
    VX3H_SC_NONE = 0,
    VX3H_SC_MAX_TIME_STEPS = 1,     /* steps > (int)(value + 0.5) */
    VX3H_SC_MAX_SIM_TIME = 2,       /* time > value */
    VX3H_SC_TEMP_CYCLES = 3,        /* time > temperature period * value */
    VX3H_SC_CONST_MAXENERGY = 4,    /* total energy per voxel within 0.001*value of its value 50..450 steps ago */
    VX3H_SC_MIN_KE = 5,             /* kinetic energy per voxel below 0.001*value at 10 samples 50 steps apart */
    VX3H_SC_MIN_MAXMOVE = 6         /* largest voxel move per step below 0.001*value at those samples */
};

/* ---------------------------------------------------------------- VXA / VXD loader */
typedef struct vx3h_vxa vx3h_vxa;   /* owns a parsed description */
/* vxd_text may be NULL.  Children of <VXD> carrying replace="VXA.a.b" replace that
 * path of the VXA tree before it is read (Voxelyze3/src/Utils/ctool.h:49-57). */
vx3h_status vx3h_vxa_parse(const char* vxa_text, const char* vxd_text, vx3h_vxa** out);
const vx3h_sim_desc* vx3h_vxa_desc(const vx3h_vxa* v);
void        vx3h_vxa_free(vx3h_vxa* v);
/* VX3 simulator keys the loader read that the physics description does not carry
 * (VX3_SimulationManager.cu:326-367): features this engine does not run, history
 * recording, and the RawPrint text the reference echoes.  raw_print may be NULL;
 * the string stays valid until vx3h_vxa_free. */
typedef struct vx3h_vxa_info {
    int    enable_attach, enable_detach;   /* AttachDetach.EnableAttach / EnableDetach */
    int    enable_cilia, enable_signals;   /* EnableCilia / EnableSignals */
    int    secondary_experiment;           /* SecondaryExperiment */
    int    record_step_size;               /* RecordHistory.RecordStepSize */
    double max_dist_close_pairs;           /* MaxDistInVoxelLengthsToCountAsPair */
    int    has_force_field;                /* a ForceField element is present */
    int    record_voxel, record_link;      /* RecordHistory.RecordVoxel [1] / RecordLink [0] */
    int    stop_condition_type;            /* StopCondition.StopConditionType [0] */
    double stop_condition_value;           /* StopCondition.StopConditionValue [0] */
    int    equilibrium_mode;               /* EquilibriumMode.EquilibriumModeEnabled [0] */
} vx3h_vxa_info;
vx3h_status vx3h_vxa_get_info(const vx3h_vxa* v, vx3h_vxa_info* info, const char** raw_print);

/* ---------------------------------------------------------------- batch engine */
enum {
    VX3H_PATH_STREAM = 0,    /* state in HBM; link pass + voxel pass per step over every resident sim */
    VX3H_PATH_RESIDENT = 1,  /* one workgroup per sim, state held in LDS/registers across many steps
                                (a sim without collisions too large for one workgroup runs as several
                                tiles, one workgroup each, with a barrier per step); admitting a sim
                                it cannot take fails with VX3H_EUNSUPPORTED */
    VX3H_PATH_AUTO = 2       /* resident where the sim fits, stream otherwise */
};
enum { VX3H_PREC_F32 = 0, VX3H_PREC_F64 = 1 };

typedef struct vx3h_batch_cfg {
    size_t max_device_bytes;  /* admission budget for simulation state; 0 = 90% of free HBM */
    int    precision;         /* VX3H_PREC_*: F64 state meets the reference's trajectories over the
                                 full horizon; F32 state is faster and meets them over short horizons */
    int    path;              /* VX3H_PATH_* */
    int    block_size;        /* stream kernels' workgroup size; 0 = tuned default */
    int    steps_per_launch;  /* resident kernel steps per launch; 0 = tuned default */
    int    collision_rebuild_interval; /* broadphase rebuild period in steps; 0 = on staleness */
    int    use_graph;         /* capture the stream step sequence in a hipGraph */
    int    collision_table_entries; /* initial collision partner table of each simulation, in entries;
                                 0 = 64 per surface voxel.  A rebuild that needs more stops that
                                 simulation in place (no state change) until the next
                                 vx3h_batch_poll enlarges its table */
} vx3h_batch_cfg;

typedef struct vx3h_batch vx3h_batch;

enum { VX3H_SIM_RUNNING = 0, VX3H_SIM_DONE = 1, VX3H_SIM_DIVERGED = 2, VX3H_SIM_EMPTY = 3,
       VX3H_SIM_FAILED = 4   /* a collision table overflowed its capacity, or a tiled simulation's
                                workgroups were not all running at once; the results are not valid */ };

typedef struct vx3h_sim_status {
    int    slot;
    int    state;                 /* VX3H_SIM_* */
    int    steps;                 /* CurStepCount */
    int    num_voxel;
    int    num_measured_voxel;
    double current_time;          /* s (double accumulator, VX3_VoxelyzeKernel.cuh:54) */
    double dt;                    /* step size used (DtFrac * recommended) */
    double initial_com[3];        /* f64 center of mass of measured voxels */
    double current_com[3];
    double fitness;               /* FitnessFunction(current_com - initial_com, ...) */
    double total_distance_of_all_voxels;
    double recent_angle;
    double recommended_dt;        /* optimal step before DtFrac (recommendedTimeStep, VX3_VoxelyzeKernel.cu:184-217);
                                     0 until the first step */
    int    collision_count;       /* target contacts so far (collisionCount, the MathTree variable hit) */
    int    n_links;               /* links so far: the admitted ones plus those made by attach */
    int    num_close_pairs;       /* numClosePairs at the last center-of-mass sample (report; MathTree) */
    double target_closeness;      /* targetCloseness at the last sample: sum of 1/distance over target pairs */
} vx3h_sim_status;

vx3h_status vx3h_batch_create(int device, const vx3h_batch_cfg* cfg, vx3h_batch** out);
vx3h_status vx3h_batch_destroy(vx3h_batch* b);
/* Builds the model (topology, materials, dt) on the host and copies it to the device. */
vx3h_status vx3h_sim_admit(vx3h_batch* b, const vx3h_sim_desc* desc, int* slot);
/* Device bytes the model would take once admitted. */
vx3h_status vx3h_sim_footprint(const vx3h_sim_desc* desc, int precision, size_t* bytes);
/* Advances every running sim by up to n_steps steps (each sim stops at its own
 * step budget, stop condition or divergence).  Asynchronous on the batch stream. */
vx3h_status vx3h_batch_step(vx3h_batch* b, int n_steps);
vx3h_status vx3h_batch_sync(vx3h_batch* b);
/* One time step of one simulation of the batch, synchronously:
 * VX3_VoxelyzeKernel::doTimeStep(float dt = -1) (VX3_VoxelyzeKernel.cuh:30, semantics at
 * VX3_VoxelyzeKernel.cu:237-359).  dt < 0 steps by DtFrac x the recommended step, cached
 * at the simulation's first step (or by its fixed dt); dt == 0 only counts a step
 * (CurStepCount++) and returns; dt > 0 steps by dt, with the center-of-mass sampling
 * period int(TempPeriod / dt) of that dt.  *ok receives 0 when a link diverged (strain
 * > 100): the reference's false return, after which the simulation is
 * VX3H_SIM_DIVERGED and further calls return 0 without stepping; 1 otherwise.  The
 * step is taken by the per-step kernels for this simulation alone, whichever path it
 * was admitted to, and regardless of its stop condition: as in the reference, the
 * caller's loop tests the stop condition (vx3h_batch_poll: state VX3H_SIM_DONE). */
vx3h_status vx3h_sim_step(vx3h_batch* b, int slot, float dt, int* ok);
/* No counterpart in the reference: vx3h_batch_step for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Every running simulation of the batch, resident ones included,
 * takes n_steps steps on the per-step kernels (the same results as vx3h_batch_step),
 * asynchronously on the batch's stream; no device memory, graph, event or stream is
 * created.  The kernels' one scratch buffer, the link end loads of a step (36 values per
 * voxel in the state precision), is a static device array of VX3H_NOALLOC_SCRATCH_BYTES
 * per GPU, held until the process ends: a batch whose voxels (up to the highest one
 * admitted) need more is refused with VX3H_EINVAL -- 466,033 voxels in FP64, 932,067 in
 * FP32.  Every call on one GPU uses that GPU's array and leaves its contents behind:
 * calls that overlap on one GPU (other batches, streams or threads) race. */
#define VX3H_NOALLOC_SCRATCH_BYTES (128u << 20)
vx3h_status vx3h_batch_step_noalloc(vx3h_batch* b, int n_steps);
/* Status of every admitted sim (synchronizes). */
vx3h_status vx3h_batch_poll(vx3h_batch* b, vx3h_sim_status* out, int cap, int* n);
vx3h_status vx3h_sim_retire(vx3h_batch* b, int slot);

/* Per-voxel / per-link state readback for parity checks (synchronizes).
 * Any pointer may be NULL.  Arrays are in voxel/link list order, f64. */
typedef struct vx3h_state_view {
    double* pos;       /* 3*V */
    double* orient;    /* 4*V (w,x,y,z) */
    double* lin_mom;   /* 3*V */
    double* ang_mom;   /* 3*V */
    int*    vox_flags; /* V: bit 1 = interior (all six links), bit 3 = floor static friction,
                          bit 4 = removed (secondary experiment) */
    double* temp;      /* V */
    double* link_strain;  /* L */
    double* link_max_strain; /* L */
    int*    link_small_angle; /* L */
    double* init_pos;  /* 3*V initial positions (the lattice positions, or those saved by a
                          secondary experiment's re-initialization) */
    int     link_capacity;  /* 0: the link arrays hold the admitted links; > 0: they hold this
                               many entries, and links made by attach follow the admitted ones
                               in the order they were made */
} vx3h_state_view;
vx3h_status vx3h_sim_read_state(vx3h_batch* b, int slot, vx3h_state_view* view);

typedef struct vx3h_topology_view {
    int n_voxels, n_links;
    int* link_vneg;   /* L */
    int* link_vpos;   /* L */
    int* link_axis;   /* L */
    int* voxel_links; /* 6*V link index per direction X_POS..Z_NEG, -1 if none */
} vx3h_topology_view;
/* Host-side topology build only (no device); n_* are filled when arrays are NULL. */
vx3h_status vx3h_build_topology(const vx3h_sim_desc* desc, vx3h_topology_view* view);

/* Measurement hook (no counterpart in the reference, whose only timing is the
 * commented-out clock() in VX3_VoxelyzeKernel.cu:238-239,353-357): runs n_steps
 * steps without graph capture, with HIP events recorded on the batch's stream
 * around every kernel pass, and returns each pass's mean duration in ms
 * (avg_ms[k], k = 0 Poisson (now part of the voxel pass: always 0), 1 links,
 * 2 contacts, 3 voxels, 4 control, 5 resident multi-step kernel; 0 when the pass
 * did not run). */
#define VX3H_N_PASSES 6
vx3h_status vx3h_batch_kernel_times(vx3h_batch* b, int n_steps, double* avg_ms);

/* Which kernels step a simulation (no counterpart in the reference, which has one
 * path): *workgroups = 0 for the per-step kernels, 1 for one workgroup of the resident
 * kernel, n > 1 for n tiles of the tiled resident kernel. */
vx3h_status vx3h_sim_workgroups(vx3h_batch* b, int slot, int* workgroups);

/* Launch timing of ordinary stepping (no counterpart in the reference): while
 * enabled, vx3h_batch_step records a HIP event pair on the batch's stream around
 * every resident multi-step launch.  vx3h_batch_timing synchronizes the stream and
 * returns, since the last call, the summed duration in ms, the number of launches
 * and the number of steps those launches were asked to run; it then resets them. */
vx3h_status vx3h_batch_set_timing(vx3h_batch* b, int enable);
vx3h_status vx3h_batch_timing(vx3h_batch* b, double* total_ms, int* launches, long long* steps);

/* Plain device-to-device copy kernel of `bytes` bytes, run reps times (no
 * counterpart in the reference): the known-traffic reference for HBM counters and
 * the bandwidth a simple streaming kernel reaches; returns the mean ms per copy. */
vx3h_status vx3h_calib_copy(int device, size_t bytes, int reps, double* ms_per_copy);

/* Collision pairs a simulation currently watches, as sim-local voxel index pairs
 * (i < j) in ascending order -- CVoxelyze's collision list after its last
 * regeneration (Voxelyze3/src/old/Voxelyze.cpp:726-751).  pairs holds 2*cap ints and
 * may be NULL to query *n. */
vx3h_status vx3h_sim_collision_pairs(vx3h_batch* b, int slot, int* pairs, int cap, int* n);

/* ---------------------------------------------------------------- autotuner
 * No counterpart in the reference, whose launch shapes are fixed in the source
 * (VX3_VoxelyzeKernel.cu:259-269,566-623).  Runs the given simulations from their
 * initial state once per candidate configuration, in a fresh batch on `device`:
 * warmup_steps untimed, then trial_steps timed; *best receives the fastest (in
 * voxel-steps/s) of base with block_size in {0, 128, 256, 512 (, 1024 for F32)},
 * steps_per_launch in {250, 500, 1000, 2000} and, only when tune_rebuild_interval is
 * set and some simulation has collisions, collision_rebuild_interval in
 * {0, 4, 8, 16}.  None of them changes a result: the first two only decide which
 * threads compute what and when a launch ends; extra collision rebuilds only add
 * watched pairs that cannot touch before the staleness rule would rebuild anyway
 * (they contribute zero force).  trials (may be NULL) receives up to cap measured
 * candidates, *n_trials their count. */
typedef struct vx3h_tune_trial {
    int    block_size, steps_per_launch, collision_rebuild_interval;
    double voxel_steps_per_s;
} vx3h_tune_trial;
vx3h_status vx3h_autotune(int device, const vx3h_sim_desc* const* descs, int n_descs, const vx3h_batch_cfg* base,
                          int warmup_steps, int trial_steps, int tune_rebuild_interval, vx3h_batch_cfg* best,
                          vx3h_tune_trial* trials, int cap, int* n_trials);

/* ---------------------------------------------------------------- serving
 * Continuous batching over one batch (replaces the file-at-a-time worker loop of
 * Voxelyze3/src/Executables/vx3_node_worker.cu:63-96 and
 * VX3_SimulationManager::start, VX3_SimulationManager.cu:138-156): requests are
 * queued as they arrive (any thread), admitted first-come-first-served whenever
 * their device footprint fits the budget, stepped together, and retired as they
 * finish, between rounds of `quantum_steps` steps.  vx3h_server_round runs one
 * scheduling round; results report per-request submit/admit/finish times. */
typedef struct vx3h_server vx3h_server;
typedef struct vx3h_server_cfg {
    vx3h_batch_cfg batch;     /* engine configuration (its max_device_bytes is the admission budget) */
    int    quantum_steps;     /* steps per round; 0 = 1000 */
    int    max_sims;          /* concurrent simulations; 0 = unbounded */
} vx3h_server_cfg;
typedef struct vx3h_server_result {
    long long ticket;
    vx3h_sim_status status;   /* final status (slot is the slot it ran in) */
    double t_submit, t_admit, t_done;   /* seconds on the server's monotonic clock */
    double voxel_steps;       /* voxels x steps simulated */
} vx3h_server_result;
vx3h_status vx3h_server_create(int device, const vx3h_server_cfg* cfg, vx3h_server** out);
vx3h_status vx3h_server_destroy(vx3h_server* sv);
/* Copies desc; max_steps > 0 overrides the description's step budget.  Thread-safe.  The
 * request's host model and tile plan are then built on the server's own threads (host
 * work only) while rounds run; round, results and destroy belong to one thread. */
vx3h_status vx3h_server_submit(vx3h_server* sv, const vx3h_sim_desc* desc, int max_steps, long long* ticket);
/* One round: admit what fits, step every running simulation quantum_steps, retire the
 * finished ones.  *active receives running + queued requests after the round. */
vx3h_status vx3h_server_round(vx3h_server* sv, int* active);
/* Finished requests since the last call (cap may be 0 to query *n). */
vx3h_status vx3h_server_results(vx3h_server* sv, vx3h_server_result* out, int cap, int* n);
/* Seconds since the server was created, on the clock of t_submit/t_admit/t_done. */
double vx3h_server_clock(const vx3h_server* sv);

/* ---------------------------------------------------------------- history
 * The VoxCAD playback stream the reference prints while a simulation with
 * RecordHistory.RecordStepSize > 0 runs (VX3_SimulationManager.cu:40-114).  Text is
 * written NUL terminated into buf; *len receives the size needed (buf may be NULL).
 * The header holds the {{{setting}}} lines (rescale, material colours of the
 * simulation's palette in order, voxel size).  A record is one line,
 * "<<<Step%d Time:%f>>>" + one entry per surface voxel (position and corner
 * offsets in mm, orientation angle and axis, material id, signal) + "<<<>>>", and,
 * when record_link is set, "|[[[%d]]]" + the two end positions of every link +
 * "[[[]]]"; step_label is the reference's loop index (steps done - 1).  Reads the
 * device state (synchronizes). */
vx3h_status vx3h_sim_history_header(vx3h_batch* b, int slot, char* buf, size_t cap, size_t* len);
vx3h_status vx3h_sim_history_record(vx3h_batch* b, int slot, int step_label, int record_voxel, int record_link,
                                    char* buf, size_t cap, size_t* len);

/* ---------------------------------------------------------------- report */
typedef struct vx3h_report_entry {
    const char* vxd_name;        /* file name of the sim's VXD */
    const vx3h_sim_status* status;
    double voxel_size;
    int    save_positions;       /* write init_pos/pos/mats */
    const double* init_pos;      /* 3*V when save_positions */
    const double* pos;           /* 3*V when save_positions */
    const int* mats;             /* V when save_positions */
} vx3h_report_entry;
/* Writes the report XML into buf (NUL terminated); *len receives the size needed. */
vx3h_status vx3h_report_xml(const char* input_dir_name, const vx3h_report_entry* entries, int n,
                            char* buf, size_t cap, size_t* len);

#ifdef __cplusplus
}
#endif
#endif /* VX3H_H */
