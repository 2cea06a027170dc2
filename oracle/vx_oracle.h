/*
 * vx_oracle.h -- C API of the CPU oracle (TEST INFRASTRUCTURE ONLY).
 *
 * The oracle is a CPU restatement of the reference Voxelyze CPU physics
 * (CVoxelyze, Voxelyze3/src/old/) plus the VX3 driver semantics
 * (VX3_VoxelyzeKernel, Voxelyze3/src/VX3/).  It exists to check the HIP engine;
 * nothing in the product links it.  Only tests/, the smoke check and the
 * benchmark's CPU-baseline leg use it.
 *
 * The same API is exported by oracle/_ref/libvx_ref.so, which wraps the
 * reference's own CPU sources compiled from /root/reference, so every test can
 * run against both and check that the restatement is bit-identical to the
 * reference.
 */
#ifndef VX_ORACLE_H
#define VX_ORACLE_H
#include "../include/vx3h.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vxo_sim vxo_sim;

/* Builds a simulation the way CVX_Sim::Import does (VXA/VX_Sim.cpp:59-153). */
vxo_sim* vxo_create(const vx3h_sim_desc* desc);
void     vxo_free(vxo_sim* s);
const char* vxo_last_error(void);

void vxo_counts(vxo_sim* s, int* n_voxels, int* n_links);
/* link endpoints/axis in link-list order; voxel_links: 6 per voxel, -1 if none */
void vxo_topology(vxo_sim* s, int* vneg, int* vpos, int* axis, int* voxel_links);

/* ---- CVoxelyze-level API (Voxelyze3/src/old/Voxelyze.cpp) ---- */
int   vxo_do_time_step(vxo_sim* s, float dt);   /* CVoxelyze::doTimeStep: 1 ok, 0 diverged */
float vxo_recommended_dt(vxo_sim* s);           /* CVoxelyze::recommendedTimeStep */
void  vxo_reset_time(vxo_sim* s);               /* CVoxelyze::resetTime */
void  vxo_set_gravity(vxo_sim* s, float g);
void  vxo_enable_floor(vxo_sim* s, int en);
void  vxo_enable_collisions(vxo_sim* s, int en);
void  vxo_set_ambient_temperature(vxo_sim* s, float t, int all_voxels);
void  vxo_set_voxel_temperature(vxo_sim* s, int v, float t);

/* externals (VX_External.h) */
void vxo_ext_set_force(vxo_sim* s, int v, float fx, float fy, float fz);
void vxo_ext_set_moment(vxo_sim* s, int v, float mx, float my, float mz);
void vxo_ext_set_fixed(vxo_sim* s, int v, int dof_mask);          /* setFixed(6 bools) */
void vxo_ext_set_fixed_all(vxo_sim* s, int v, int fixed);         /* setFixedAll */
void vxo_ext_set_displacement_all(vxo_sim* s, int v, double tx, double ty, double tz,
                                  double rx, double ry, double rz);
void vxo_external_force(vxo_sim* s, int v, float* out3);          /* CVX_Voxel::externalForce */

/* material edits (CVX_Material setters; which: 0 internal damping, 1 global damping,
 * 2 collision damping, 3 static friction, 4 kinetic friction, 5 cte, 6 poisson) */
void vxo_material_set(vxo_sim* s, int m, int which, float value);
/* CVoxelyze::setVoxel on an existing voxel (replaceVoxel) */
void vxo_replace_voxel_material(vxo_sim* s, int v, int m);

/* ---- VX3 driver (VX3_SimulationManager.cu:54-117, VX3_VoxelyzeKernel.cu:237-359) ---- */
int  vxo_vx3_step(vxo_sim* s, int n_steps);    /* returns steps done */
void vxo_vx3_status(vxo_sim* s, vx3h_sim_status* st);

/* ---- state readback ---- */
void vxo_voxel_state(vxo_sim* s, double* pos, double* orient, double* lin_mom, double* ang_mom,
                     int* flags, float* temp);
void vxo_link_state(vxo_sim* s, float* strain, float* max_strain, float* strain_offset,
                    int* small_angle, int* bool_states, double* rest_length, double* forces12);
void vxo_corner_offset(vxo_sim* s, int v, int corner, float* out3);
/* VoxCAD stop-condition histories, 500 entries each, newest first (-1 = not computed) */
void vxo_voxcad_history(vxo_sim* s, double* kinetic, double* total, double* max_move);
/* VX3 local signal of each voxel (0 where the engine has no signals) */
void vxo_local_signal(vxo_sim* s, double* out);
int  vxo_collision_pairs(vxo_sim* s, int* pairs, int cap);        /* returns count */
/* n > 0: the VX3 driver also regenerates the collision lists at every step whose
 * zero-based index is a multiple of n, besides CVoxelyze's staleness rule (a fixed
 * broadphase period; the engine's collision_rebuild_interval) */
void vxo_set_collision_interval(vxo_sim* s, int n);
/* derived material constants, for the host builder's parity checks.
 * voxel mats: 14 floats each; link mats (list order): 20 floats each */
int  vxo_voxel_material_consts(vxo_sim* s, float* out, int cap);
int  vxo_link_material_consts(vxo_sim* s, int* link_mat_index, float* out, int cap);

/* MathTree evaluator (Utils/VX3_MathTree.h:50-192) */
double vxo_mathtree_eval(const vx3h_token* toks, int n, double x, double y, double z, double hit,
                         double t, double angle, double closeness, int num_close_pairs, int num_voxel);

#ifdef __cplusplus
}
#endif
#endif
