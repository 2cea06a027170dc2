// vx3_kernel.h -- VX3_VoxelyzeKernel: one simulation, stepped one doTimeStep at a time.
//
// The reference's per-simulation class (Voxelyze3/src/VX3/VX3_VoxelyzeKernel.cuh:20-31)
// is built on the host from a CVX_Sim (VX3_VoxelyzeKernel.cu:27-105) and stepped by a
// device thread that calls doTimeStep(dt) until StopConditionMet()
// (VX3_SimulationManager.cu:62-69).  Here the same object owns a one-simulation batch
// of the C ABI: it is built from a parsed VXA/VXD (vx3h_vxa, this engine's counterpart
// of CVX_Sim) or from a vx3h_sim_desc, and doTimeStep is called from the host; each
// call is one step of the per-step kernels (vx3h_sim_step), the plain
// one-simulation-per-launch path.  Many simulations at once run faster through a
// batch (vx3h_batch_step), which the manager (vx3_manager.h) uses.
//
// Header only: link with libvx3h.so.
#ifndef VX3_KERNEL_H
#define VX3_KERNEL_H

#include <stdexcept>
#include <string>

#include "../../include/vx3h.h"

class VX3_VoxelyzeKernel {
public:
    explicit VX3_VoxelyzeKernel(const vx3h_vxa* sim, int device = 0, int precision = VX3H_PREC_F64) {
        init(vx3h_vxa_desc(sim), device, precision);
    }
    explicit VX3_VoxelyzeKernel(const vx3h_sim_desc* desc, int device = 0, int precision = VX3H_PREC_F64) {
        init(desc, device, precision);
    }
    ~VX3_VoxelyzeKernel() { cleanup(); }
    VX3_VoxelyzeKernel(const VX3_VoxelyzeKernel&) = delete;
    VX3_VoxelyzeKernel& operator=(const VX3_VoxelyzeKernel&) = delete;

    // VX3_VoxelyzeKernel::doTimeStep (VX3_VoxelyzeKernel.cu:237-359): dt < 0 steps by
    // DtFrac x the cached recommended step, dt == 0 only counts the step, dt > 0 steps by
    // dt; false when a link diverged.
    bool doTimeStep(float dt = -1.0f) {
        int ok = 0;
        check(vx3h_sim_step(b_, slot_, dt, &ok));
        return ok != 0;
    }
    // VX3_VoxelyzeKernel::StopConditionMet (VX3_VoxelyzeKernel.cu:162-182): the stop
    // program holds, or the description's step budget is spent
    bool StopConditionMet() { return status().state == VX3H_SIM_DONE; }
    // OptimalDt before DtFrac (VX3_VoxelyzeKernel.cu:184-217); 0 until the first step
    double recommendedTimeStep() { return status().recommended_dt; }
    // currentTime, CurStepCount, centers of mass, fitness_score (computeFitness), ... now
    const vx3h_sim_status& status() {
        int n = 0;
        check(vx3h_batch_poll(b_, &st_, 1, &n));
        return st_;
    }
    // cleanup() (VX3_VoxelyzeKernel.cu:107-119): frees the device memory now
    void cleanup() {
        if (b_) vx3h_batch_destroy(b_);
        b_ = nullptr;
    }
    vx3h_batch* batch() const { return b_; }
    int slot() const { return slot_; }

private:
    vx3h_batch* b_ = nullptr;
    int slot_ = -1;
    vx3h_sim_status st_{};
    static void check(vx3h_status s) {
        if (s != VX3H_OK) throw std::runtime_error(std::string("vx3h: ") + vx3h_last_error());
    }
    void init(const vx3h_sim_desc* desc, int device, int precision) {
        if (!desc) throw std::runtime_error("vx3h: no simulation description");
        vx3h_batch_cfg cfg = {};
        cfg.precision = precision;
        cfg.path = VX3H_PATH_STREAM;
        check(vx3h_batch_create(device, &cfg, &b_));
        vx3h_status s = vx3h_sim_admit(b_, desc, &slot_);
        if (s != VX3H_OK) {
            const std::string msg = vx3h_last_error();
            cleanup();
            throw std::runtime_error("vx3h: " + msg);
        }
    }
};

#endif
