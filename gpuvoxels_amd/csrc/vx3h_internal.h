// vx3h_internal.h -- declarations shared by the engine's translation units.
#ifndef VX3H_INTERNAL_H
#define VX3H_INTERNAL_H
#include <string>
#include <vector>
#include "../../include/vx3h.h"
namespace vx3 {
extern thread_local std::string g_err;   // last error of the calling thread (vx3h_last_error)
// vx3h_batch_poll for a caller that only needs the simulations that stopped: the full
// status (centre of mass, fitness, distances: positions read back) of each simulation
// not running, after one read of the control blocks for all of them
vx3h_status pollFinished(vx3h_batch* b, std::vector<vx3h_sim_status>& finished);
// vx3h_sim_admit in two parts: the host model and (for a simulation too large for one
// workgroup) its tile plan, built on the calling thread with nothing shared written
// (several may run at once, on different threads); then the admission itself.  A
// prepared simulation is for the batch it was prepared for; admitPrepared frees it.
struct PreparedSim;
vx3h_status prepareSim(const vx3h_batch* b, const vx3h_sim_desc* desc, PreparedSim** out);
vx3h_status admitPrepared(vx3h_batch* b, const vx3h_sim_desc* desc, PreparedSim* p, int* slot);
void freePrepared(PreparedSim* p);
// voxels and link slots a prepared simulation takes
void preparedSize(const PreparedSim* p, long long* nv, long long* nl);
// room for nv more voxels and nl more links past the batch arrays' high-water marks, in
// one growth (admissions that follow then grow nothing)
vx3h_status reserveBatch(vx3h_batch* b, long long nv, long long nl);
}
#endif
