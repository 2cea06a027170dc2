// vx3h_build.h -- host-side model builder: vx3h_sim_desc -> device tables.
//
// Produces what VX3_VoxelyzeKernel's host constructor copies to the device
// (Voxelyze3/src/VX3/VX3_VoxelyzeKernel.cu:27-105) after CVX_Sim::Import has
// built the CPU model (Voxelyze3/src/VXA/VX_Sim.cpp:59-153), but as flat
// structure-of-arrays tables instead of pointer-linked objects.
#ifndef VX3H_BUILD_H
#define VX3H_BUILD_H

#include <string>
#include <vector>

#include "../../include/vx3h.h"
#include "vx3_types.h"

namespace vx3 {

struct HostModel {
    int nv = 0, nl = 0;
    double voxSize = 0.001;
    // per voxel-material (desc order)
    std::vector<VoxMatDev> vmats;
    // per link-material (creation order, CVoxelyze::linkMats)
    std::vector<LinkMatDev> lmats;
    // voxels (list order)
    std::vector<int> ijk;         // 3*nv
    std::vector<int> vmat;        // nv
    std::vector<double> phase;    // nv
    std::vector<int> vflags;      // nv, VF_* incl. link mask
    std::vector<float> tfac0;     // nv thermal size factor at t = 0
    std::vector<int> vext;        // nv index into exts or -1
    std::vector<ExtDev> exts;
    // links (list order)
    std::vector<int> lneg, lpos, laxis, lmat;
    std::vector<int> voxelLinks;  // 6*nv link index per direction or -1
    // scalars
    float tInit = 0;              // temperature assigned by the importer
    double com0[3] = {0, 0, 0};
    double optDt = 0;             // VX3 recommended dt at the first step
    float dt = 0;
    int flags = 0;                // SC_*
    bool anyPoisson = false;
    double tempAmp = 0, tempPeriod = 0, dtFrac = 0.9, dtFixed = 0;
    int maxSteps = 0;
    bool collisions = false;
    double boundingRadius = 0.75, watchDistance = 1.0;
    std::vector<vx3h_token> stopTok, fitTok;
    int saveAll = 0;
    // VX3 extras: cilia (6 per voxel: base then shift force, voxel frame; empty when
    // cilia are off) with each voxel-material's Cilia scale and switch-on time, and
    // the force-field programs per axis
    std::vector<double> cilia;
    std::vector<double> vmCilia, vmCiliaOn;
    std::vector<vx3h_token> ffTok[3];
    // VX3 attach: per voxel-material the device link-material index of a link between two
    // of its voxels (-1 unless sticky), and the attach-condition programs
    std::vector<int> vmSelf;
    std::vector<vx3h_token> atTok[5];
    bool extras() const { return !cilia.empty() || !ffTok[0].empty() || !ffTok[1].empty() || !ffTok[2].empty(); }
};

// Returns false and fills err on an invalid description.
bool buildModel(const vx3h_sim_desc* d, HostModel& out, std::string& err);

// MathTree evaluation (queue evaluator of Voxelyze3/src/Utils/VX3_MathTree.h:50-192).
double evalMathTree(const vx3h_token* toks, int n, double x, double y, double z, double hit, double t, double angle,
                    double closeness, int numClosePairs, int numVoxel);

}  // namespace vx3
#endif
