// vx3_batch.h -- device pointers of one batch, passed by value to every kernel.
#ifndef VX3_BATCH_H
#define VX3_BATCH_H
#include "vx3_types.h"

namespace vx3 {

template <typename T> struct DevBatch {
    // voxel state [Vcap]
    T *px, *py, *pz;          // position
    T *mx, *my, *mz;          // linear momentum
    T *ow, *ox, *oy, *oz;     // orientation quaternion
    T *ax, *ay, *az;          // angular momentum
    int* vflags;              // VF_* incl. link-presence mask
    float* tfac;              // 1 + temperature*CTE for the step about to run
    double* phase;            // actuation phase offset (cycles)
    int* vmat;                // index into vmatTab
    int* vsim;                // owning sim slot, -1 if free
    int* vext;                // index into extTab or -1
    int* vijk;                // 3*Vcap lattice index (externals / reporting)
    float *psx, *psy, *psz;   // Poisson strain of the voxel (CVX_Voxel::poissonsStrain)
    T* F;                     // link end loads of the step: F[(d*6+c)*Fs + v], 36 per voxel
    size_t Fs;                // stride of F: Vcap for the batch's own buffer
    double* cil;              // 8*Vcap cilia: base force xyz, shift force xyz (voxel frame), Cilia scale,
                              // switch-on time; cil[k*Vcap + v]
    T* xf;                    // 6*Vcap this step's global cilia force xyz and force-field force xyz
    double* ip;               // 3*Vcap initial positions saved by a secondary experiment's reset
    double* rt;               // Vcap removal time of the voxel's material (secondary experiment), 0 = never
    double* sg;               // 8*Vcap signal state: local signal, its decay time, next pacemaker pulse,
                              // deaf until, pending value and time, posted value and time
    double* sp;               // 5*Vcap signal material: pacemaker, period, value decay, time delay, inactive period
    int* nbr;                 // 6*Vcap neighbour voxel per link direction (signal simulations), -1 = none
    float* es;                // 6*Vcap axial strain of the link end in slot d: es[d*Vcap + v]
    // link state [Lcap]
    int *lneg, *lpos;
    int* lmeta;               // axis | (link material << 2)
    int* lsim;                // owning sim, -1 if free
    int* lflags;              // LF_*
    T *q2x, *q2y, *q2z;       // previous local pos2
    T *a1x, *a1y, *a1z;       // previous angle1v
    T *a2x, *a2y, *a2z;       // previous angle2v
    float *strain, *maxStrain, *strainOffset;
    int* lnew;                // steps left in which a link made by attach carries 1/100 of its force
    // collisions: per voxel the watched partners in ascending index order (both
    // directions), the order CVX_Voxel::force walks its colWatch list
    float* lastWatch;         // 3*Vcap position at the last rebuild (CVX_Voxel::lastColWatchPosition)
    // contact snapshot: position, linear momentum and thermal factor of every voxel of a
    // collision sim as they are before the voxel pass, which reads its partners' values
    // from here while it updates the live arrays in place
    T *cpx, *cpy, *cpz, *cmx, *cmy, *cmz;
    float* ctf;
    int *nbStart, *nbCount;   // per voxel range in nbList
    int *nearStart, *nearCount;  // per voxel range in nearList: surface voxels within the link-hop radius
    int* nearList;            // sim-local indices, ascending
    int* nbList;              // global voxel indices
    int* hashCnt;             // per bucket: surface voxels in it
    int* hashStart;           // per bucket: offset of its first voxel within its block of VX3_COL_HBLOCK buckets
    int* hashSim;             // owning sim of each bucket, -1 if free
    int* hashBlk;             // per bucket block: voxels in the block
    int* hashBlkOff;          // per bucket block: offset of its first voxel in cellVox
    int* cellVox;             // surface voxels sorted by bucket, a range of nv entries per sim
    int *vbk, *vrank;         // per voxel: its bucket and its rank within the bucket (last rebuild)
    int* tgtList;             // voxels of isTarget materials per sim (close pairs), a range per sim
    int* attCand;             // attach scratch per sim: attachable voxels, then their pair offsets
    unsigned long long* attPair;   // attach scratch per sim: candidate pairs (higher << 32 | lower), ascending
    unsigned* nearMask;       // [VX3_NEAR_WORDS][Vcap]: each surface voxel's nearby set over lattice offsets
                              // (simulations with SC_NEARMASK), kept with nearStart/nearCount/nearList
    int* colAny;              // one word: some simulation rebuilds its contact lists this step (set by the
                              // staleness pass, cleared by the control pass)
    // tables
    const VoxMatDev* vmatTab;
    const LinkMatDev* lmatTab;
    const ExtDev* extTab;
    SimCtl* sims;
    double* hist;             // per sim 3*VX3_HIST: kinetic energy, total energy, max move histories
    const TokenDev* tokens;
    const int* simMats;       // per sim: global voxel-material then link-material table indices it uses
    int Vcap, Lcap, Ccap;     // Ccap: nbList capacity
    int hashTot;              // buckets in use (high-water mark, whole blocks)
    int cellHi;               // cellVox entries in use (high-water mark)
    int vHi, lHi, nSims;      // high-water marks: kernels cover [0, vHi) / [0, lHi) / [0, nSims)
    int solo;                 // >= 0: only this simulation takes the step (on the per-step kernels,
                              // whichever path it was admitted to); -1: every stream-path simulation;
                              // -2: every simulation, resident ones included
};

// Tables of a tiled resident launch (one workgroup per tile, TileDesc); unused by
// one-workgroup launches.
struct TileArgs {
    const TileDesc* tiles;   // per workgroup
    const int* vox;          // per tile slot: sim-local voxel; owned slots carry bit 31 when another tile reads them
    const int* link;         // per tile link: sim-local link; bit 31 = this tile writes it back
    const int* ends;         // per tile link: LDS slots of its ends, negative | positive << 16
    const int* xidx;         // per tile slot: its exchange record (written if owned, read if halo), -1 = none
    void* xch;               // exchange areas, in the state precision
    int* sync;               // per simulation its divergence word, then one arrival flag per
                             // workgroup from sync[flag0] (all zero at launch start)
    double* part;            // per simulation: centre-of-mass partial sums
    long long spin;          // wall-clock ticks a tile waits at a barrier before failing the simulation
    int repair;              // 1: run only the simulations whose last launch diverged (divergence word
                             // = 1 + the launch step), up to that step's link pass
    int flag0;               // index in sync of workgroup 0's arrival flag
};

}  // namespace vx3
#endif
