// vx3h_tune.cpp -- autotuner of the batch configuration.
//
// The reference has no tuner: its launch shapes are fixed in the source
// (VX3_VoxelyzeKernel.cu:259-269, 566-623).  Here the engine has choices that do not
// change any result: which resident kernel shape a simulation runs in (the preferred
// workgroup size), how many steps one resident launch runs, the element passes'
// workgroup size, and, on request, the collision rebuild period (extra rebuilds
// beyond the staleness rule add only watched pairs that cannot touch in between).
// Every candidate runs the same simulations from their initial state in a fresh
// batch: warm-up steps, then timed steps between two synchronizations; the fastest
// in voxel-steps per second wins.
#include <algorithm>
#include <chrono>
#include <vector>

#include "../../include/vx3h.h"
#include "vx3h_internal.h"

using vx3::g_err;

namespace {

// voxel-steps done so far over every admitted simulation
vx3h_status progress(vx3h_batch* b, double* vs) {
    int n = 0;
    vx3h_status st = vx3h_batch_poll(b, nullptr, 0, &n);
    if (st != VX3H_OK) return st;
    std::vector<vx3h_sim_status> s(n > 0 ? n : 1);
    if ((st = vx3h_batch_poll(b, s.data(), n, &n)) != VX3H_OK) return st;
    double t = 0;
    for (int i = 0; i < n; i++) t += (double)s[i].num_voxel * s[i].steps;
    *vs = t;
    return VX3H_OK;
}

vx3h_status trial(int device, const vx3h_sim_desc* const* descs, int n, const vx3h_batch_cfg& cfg, int warmup, int steps,
                  double* rate) {
    vx3h_batch* b = nullptr;
    vx3h_status st = vx3h_batch_create(device, &cfg, &b);
    if (st != VX3H_OK) return st;
    int slot;
    for (int i = 0; i < n && st == VX3H_OK; i++) st = vx3h_sim_admit(b, descs[i], &slot);
    double v0 = 0, v1 = 0;
    if (st == VX3H_OK && warmup > 0) st = vx3h_batch_step(b, warmup);
    if (st == VX3H_OK) st = progress(b, &v0);   // synchronizes
    const auto t0 = std::chrono::steady_clock::now();
    if (st == VX3H_OK) st = vx3h_batch_step(b, steps);
    if (st == VX3H_OK) st = vx3h_batch_sync(b);
    const auto t1 = std::chrono::steady_clock::now();
    if (st == VX3H_OK) st = progress(b, &v1);
    vx3h_batch_destroy(b);
    if (st != VX3H_OK) return st;
    const double sec = std::chrono::duration<double>(t1 - t0).count();
    *rate = sec > 0 ? (v1 - v0) / sec : 0.0;
    return VX3H_OK;
}

}  // namespace

extern "C" vx3h_status vx3h_autotune(int device, const vx3h_sim_desc* const* descs, int n_descs, const vx3h_batch_cfg* base,
                                     int warmup_steps, int trial_steps, int tune_rebuild_interval, vx3h_batch_cfg* best,
                                     vx3h_tune_trial* trials, int cap, int* n_trials) {
    if (!descs || n_descs <= 0 || !base || !best || trial_steps <= 0 || warmup_steps < 0) {
        g_err = "bad argument";
        return VX3H_EINVAL;
    }
    bool collisions = false;
    for (int i = 0; i < n_descs; i++) {
        if (!descs[i]) { g_err = "null description"; return VX3H_EINVAL; }
        collisions |= descs[i]->collisions_enabled != 0;
    }
    // workgroup sizes: 0 (the engine's order: smallest voxel capacity first), each
    // resident shape's block size, and 128 for the element passes
    std::vector<int> blocks = {0, 128, 256, 512};
    if (base->precision == VX3H_PREC_F32) blocks.push_back(1024);
    const std::vector<int> spls = base->path == VX3H_PATH_STREAM ? std::vector<int>{base->steps_per_launch}
                                                                   : std::vector<int>{250, 500, 1000, 2000};
    std::vector<int> intervals = {base->collision_rebuild_interval};
    if (tune_rebuild_interval && collisions) intervals = {0, 4, 8, 16};

    int count = 0;
    double bestRate = -1;
    for (int bs : blocks)
        for (int spl : spls)
            for (int iv : intervals) {
                vx3h_batch_cfg c = *base;
                c.block_size = bs;
                c.steps_per_launch = spl;
                c.collision_rebuild_interval = iv;
                double rate = 0;
                vx3h_status st = trial(device, descs, n_descs, c, warmup_steps, trial_steps, &rate);
                if (st != VX3H_OK) return st;
                if (trials && count < cap) trials[count] = vx3h_tune_trial{bs, spl, iv, rate};
                count++;
                if (rate > bestRate) { bestRate = rate; *best = c; }
            }
    if (n_trials) *n_trials = count;
    return VX3H_OK;
}
