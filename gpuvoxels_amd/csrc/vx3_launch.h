// vx3_launch.h -- host entry points of the kernel translation units.
#ifndef VX3_LAUNCH_H
#define VX3_LAUNCH_H
#include <hip/hip_runtime.h>
#include "vx3_batch.h"

namespace vx3 {
struct StepOpts {
    bool contacts = false;     // some simulation has collisions: run the collision passes
    bool extras = false;       // some simulation has cilia or a force field: run its voxel pass
    bool prestep = false;      // some simulation has cilia, a force field or detach: run the step-start pass
    bool signals = false;      // some simulation has signals: run the signal-delivery pass
    bool attach = false;       // some simulation attaches sticky voxels: run the attach pass and conditions
    bool topo = false;         // some simulation has collisions and topology changes: run the nearby-list passes
    int block = 256;   // workgroup size of the element passes
};
// Kernel passes of one step, in launch order; ev (optional) receives a start/stop
// event pair per pass: ev[2*k], ev[2*k+1].  PASS_POISSON no longer runs on its own
// (the voxel pass computes the Poisson strain); its index keeps the ABI's numbering.
enum { PASS_POISSON = 0, PASS_LINKS, PASS_CONTACTS, PASS_VOXELS, PASS_CONTROL, PASS_COUNT };
void launchStep(const DevBatch<float>& b, hipStream_t s, const StepOpts& o, hipEvent_t* ev = nullptr);
void launchStep(const DevBatch<double>& b, hipStream_t s, const StepOpts& o, hipEvent_t* ev = nullptr);
// Resident multi-step kernel shapes: (block, voxels/thread, links per axis/thread), LDS bytes.
// A sim fits a shape when V <= block*vpt, each axis has <= block*lpa links, and it uses at
// most RES_MAX_VMATS voxel and RES_MAX_LMATS link material variants and RES_MAX_PHASES
// distinct actuation phase offsets.
struct ResidentShapeInfo { int block, vpt, lpa; size_t lds, ldsCol; };   // ldsCol: the variant with collisions
enum { RES_MAX_VMATS = 8, RES_MAX_LMATS = 16, RES_MAX_PHASES = 64 };
int residentShapes_f32(ResidentShapeInfo* out, int cap);
int residentShapes_f64(ResidentShapeInfo* out, int cap);
inline int residentShapes(bool f64, ResidentShapeInfo* out, int cap) {
    return f64 ? residentShapes_f64(out, cap) : residentShapes_f32(out, cap);
}
// Launches the resident kernel of shape `shape` (with collisions or without) over the sims
// in simList (device array).
hipError_t launchResident(const DevBatch<float>& b, hipStream_t s, int shape, bool collisions, const int* simList, int nSims,
                          int nSteps, hipEvent_t* ev = nullptr);
hipError_t launchResident(const DevBatch<double>& b, hipStream_t s, int shape, bool collisions, const int* simList, int nSims,
                          int nSteps, hipEvent_t* ev = nullptr);
// Launches the tiled resident kernel of shape `shape` (no collisions) over the tiles
// ta.tiles[0, nTiles) after clearing ta.sync[0, nSync); every tile is resident at once.
hipError_t launchTiled(const DevBatch<float>& b, hipStream_t s, int shape, const TileArgs& ta, int nTiles, int nSync,
                       int nSteps, hipEvent_t* ev = nullptr);
hipError_t launchTiled(const DevBatch<double>& b, hipStream_t s, int shape, const TileArgs& ta, int nTiles, int nSync,
                       int nSteps, hipEvent_t* ev = nullptr);
hipError_t tiledOccupancy(const DevBatch<float>& b, int shape, int* perCU);
hipError_t tiledOccupancy(const DevBatch<double>& b, int shape, int* perCU);
void launchDt(const DevBatch<float>& b, hipStream_t s, int s0, int s1);
void launchDt(const DevBatch<double>& b, hipStream_t s, int s0, int s1);
// copies the link end loads of voxels [0, nv) from src (slot stride srcFs) to dst (slot
// stride dstFs), in the state precision
void launchCopyLoads(const DevBatch<float>& b, hipStream_t s, const void* src, size_t srcFs, void* dst, size_t dstFs, int nv);
void launchCopyLoads(const DevBatch<double>& b, hipStream_t s, const void* src, size_t srcFs, void* dst, size_t dstFs, int nv);
}  // namespace vx3
#endif
