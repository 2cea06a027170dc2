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
}
#endif
