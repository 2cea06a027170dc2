// vx3_manager.h -- C++ simulation manager over the C ABI (include/vx3h.h).
//
// The host-side counterpart of VX3_SimulationManager
// (Voxelyze3/src/VX3/VX3_SimulationManager.cuh:15-46): constructed from per-device
// lists of VXD files, a base VXA and the input directory; start() reads every VXD
// merged into the base, runs all simulations of a device as one batch until each
// meets its stop condition, reaches the step cap or diverges, and fills h_results
// sorted by fitness (descending, NaN last; VX3_SimulationResult.h:26-33).
#ifndef VX3_MANAGER_H
#define VX3_MANAGER_H

#include <cstdio>
#include <string>
#include <vector>

#include "../../include/vx3h.h"

namespace vx3 {

struct SimulationResult {              // VX3_SimulationResult (VX3_SimulationResult.h:4-24)
    std::string vxa_filename;          // the VXD's file name
    double x = 0, y = 0, z = 0;        // current center of mass
    double voxSize = 0;
    double currentTime = 0;
    int num_voxel = 0;
    int num_measured_voxel = 0;
    double fitness_score = 0;
    bool SavePositionOfAllVoxels = false;
    std::vector<int> voxel_mats;
    std::vector<double> voxel_init_pos;   // 3 per voxel
    std::vector<double> voxel_position;   // 3 per voxel
    double total_distance_of_all_voxels = 0;
    double initialCenterOfMass[3] = {0, 0, 0};
    double currentCenterOfMass[3] = {0, 0, 0};
    int numClosePairs = 0;
    int state = 0;                      // VX3H_SIM_*: done or diverged
    vx3h_sim_status status{};           // as polled
};

struct ManagerOptions {
    int precision = VX3H_PREC_F64;      // state precision of the device engine
    int path = VX3H_PATH_AUTO;
    int steps_per_poll = 2000;          // steps between host polls of the batch
    bool allow_unsupported = false;     // accepted for callers that set it: nothing is left unsimulated
    bool verbose = true;
    FILE* history_out = nullptr;        // playback stream of RecordHistory simulations [stdout]
    bool voxcad_stop = false;           // apply the VXA's VoxCAD stop condition and equilibrium mode
};

class SimulationManager {
public:
    SimulationManager(std::vector<std::vector<std::string>> sub_batches, std::string base, std::string input_dir,
                      int num_of_devices, ManagerOptions opt = ManagerOptions());
    // Runs every simulation; throws std::runtime_error with the engine's message on failure.
    void start();
    std::vector<SimulationResult> h_results;

private:
    std::vector<std::vector<std::string>> sub_batches;
    std::string base, input_dir;
    int num_of_devices;
    ManagerOptions opt;
    void runDevice(int device, const std::vector<std::string>& files);
    void runHistory(vx3h_batch* b, int slot, const vx3h_vxa_info& info, double dtFrac);
};

}  // namespace vx3

// The reference's class names (Voxelyze3/src/VX3/VX3_SimulationManager.cuh:15,
// VX3_SimulationResult.h:4): the same constructor, start() and h_results.
using VX3_SimulationManager = vx3::SimulationManager;
using VX3_SimulationResult = vx3::SimulationResult;

#endif
