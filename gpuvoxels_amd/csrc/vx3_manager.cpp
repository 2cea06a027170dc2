// vx3_manager.cpp -- SimulationManager: VXA/VXD files -> batched device runs -> results.
//
// The reference starts one device thread per simulation and steps each with its own
// child launches (VX3_SimulationManager.cu:138-156,398-414).  Here all simulations of a
// device share one batch on one stream; the host steps the batch in chunks and polls
// states between chunks, which is where finished simulations are noticed.
#include "vx3_manager.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace vx3 {

namespace {

std::string readFile(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot read " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

std::string joinPath(const std::string& dir, const std::string& file) {
    if (dir.empty() || (!file.empty() && file[0] == '/')) return file;
    return dir.back() == '/' ? dir + file : dir + "/" + file;
}

std::string baseName(const std::string& p) {
    size_t k = p.find_last_of('/');
    return k == std::string::npos ? p : p.substr(k + 1);
}

void check(vx3h_status st, const std::string& what) {
    if (st != VX3H_OK) throw std::runtime_error(what + ": " + vx3h_last_error());
}

}  // namespace

SimulationManager::SimulationManager(std::vector<std::vector<std::string>> sb, std::string b, std::string in, int n,
                                     ManagerOptions o)
    : sub_batches(std::move(sb)), base(std::move(b)), input_dir(std::move(in)), num_of_devices(n), opt(o) {}

void SimulationManager::start() {
    for (int d = 0; d < num_of_devices && d < (int)sub_batches.size(); d++)
        if (!sub_batches[d].empty()) runDevice(d, sub_batches[d]);
    std::stable_sort(h_results.begin(), h_results.end(), [](const SimulationResult& a, const SimulationResult& b) {
        if (std::isnan(b.fitness_score)) return !std::isnan(a.fitness_score);
        if (std::isnan(a.fitness_score)) return false;
        return a.fitness_score > b.fitness_score;
    });
}

void SimulationManager::runDevice(int device, const std::vector<std::string>& files) {
    if (opt.verbose) printf("=== set device to %d for %zu simulations ===\n", device, files.size());
    const std::string baseText = readFile(base);
    vx3h_batch_cfg cfg{};
    cfg.precision = opt.precision;
    cfg.path = opt.path;
    vx3h_batch* b = nullptr;
    check(vx3h_batch_create(device, &cfg, &b), "batch creation");
    struct Job {
        std::string name;
        int slot;
        vx3h_vxa* v;
        vx3h_vxa_info info;
        vx3h_sim_desc desc;   // the loader's description, with the VoxCAD stop keys when asked for
    };
    auto describe = [&](vx3h_vxa* v, const vx3h_vxa_info& info) {
        vx3h_sim_desc d = *vx3h_vxa_desc(v);
        if (opt.voxcad_stop) {
            d.stop_condition_type = info.stop_condition_type;
            d.stop_condition_value = info.stop_condition_value;
            d.equilibrium_mode = info.equilibrium_mode;
        }
        return d;
    };
    std::vector<Job> jobs, history;
    try {
        for (const std::string& f : files) {
            const std::string vxd = readFile(joinPath(input_dir, f));
            vx3h_vxa* v = nullptr;
            check(vx3h_vxa_parse(baseText.c_str(), vxd.c_str(), &v), f);
            vx3h_vxa_info info;
            const char* raw = nullptr;
            check(vx3h_vxa_get_info(v, &info, &raw), f);
            if (raw && *raw) printf("%s\n", raw);   // RawPrint (VX3_SimulationManager.cu:319-321)
            if (info.record_step_size > 0) {   // recorded simulations run alone, after the batch
                history.push_back({baseName(f), -1, v, info, describe(v, info)});
                continue;
            }
            int slot = -1;
            const vx3h_sim_desc d = describe(v, info);
            vx3h_status st = vx3h_sim_admit(b, &d, &slot);
            if (st != VX3H_OK) { vx3h_vxa_free(v); check(st, f); }
            jobs.push_back({baseName(f), slot, v, info, d});
        }
        // run until every simulation is done or diverged
        for (;;) {
            int n = 0;
            check(vx3h_batch_poll(b, nullptr, 0, &n), "poll");
            std::vector<vx3h_sim_status> sts(std::max(n, 1));
            check(vx3h_batch_poll(b, sts.data(), n, &n), "poll");
            bool running = false;
            for (int i = 0; i < n; i++) if (sts[i].state == VX3H_SIM_RUNNING) running = true;
            if (!running) break;
            check(vx3h_batch_step(b, opt.steps_per_poll), "step");
        }
        auto collect = [&](vx3h_batch* bt, const Job& j) {
            int n = 0;
            check(vx3h_batch_poll(bt, nullptr, 0, &n), "poll");
            std::vector<vx3h_sim_status> sts(std::max(n, 1));
            check(vx3h_batch_poll(bt, sts.data(), n, &n), "poll");
            const vx3h_sim_status* s = nullptr;
            for (int i = 0; i < n; i++) if (sts[i].slot == j.slot) s = &sts[i];
            if (!s) throw std::runtime_error(j.name + ": no status");
            const vx3h_sim_desc* d = vx3h_vxa_desc(j.v);
            SimulationResult r;
            r.vxa_filename = j.name;
            r.status = *s;
            r.state = s->state;
            r.currentTime = s->current_time;
            r.fitness_score = s->fitness;
            r.x = s->current_com[0]; r.y = s->current_com[1]; r.z = s->current_com[2];
            for (int k = 0; k < 3; k++) { r.initialCenterOfMass[k] = s->initial_com[k]; r.currentCenterOfMass[k] = s->current_com[k]; }
            r.voxSize = d->voxel_size;
            r.num_voxel = s->num_voxel;
            r.num_measured_voxel = s->num_measured_voxel;
            r.total_distance_of_all_voxels = s->total_distance_of_all_voxels;
            r.numClosePairs = s->num_close_pairs;
            r.SavePositionOfAllVoxels = d->save_position_of_all_voxels != 0;
            if (r.SavePositionOfAllVoxels) {
                const int V = d->n_voxels;
                r.voxel_init_pos.resize(3 * V);
                r.voxel_position.resize(3 * V);
                r.voxel_mats.resize(V);
                for (int i = 0; i < V; i++) r.voxel_mats[i] = d->materials[d->material[i]].matid;
                vx3h_state_view view{};
                view.pos = r.voxel_position.data();
                view.init_pos = r.voxel_init_pos.data();   // d_initialPosition (VX3_SimulationManager.cu:439-446)
                check(vx3h_sim_read_state(bt, j.slot, &view), j.name);
            }
            if (opt.verbose)
                printf("%d) Simulation %s ends: Time: %f%s.\n", device, j.name.c_str(), r.currentTime,
                       r.state == VX3H_SIM_DIVERGED ? " (diverged)" : "");
            h_results.push_back(std::move(r));
        };
        for (const Job& j : jobs) collect(b, j);
        vx3h_batch_destroy(b);
        b = nullptr;
        for (Job& j : history) {
            check(vx3h_batch_create(device, &cfg, &b), "batch creation");
            check(vx3h_sim_admit(b, &j.desc, &j.slot), j.name);
            runHistory(b, j.slot, j.info, j.desc.dt_frac);
            collect(b, j);
            vx3h_batch_destroy(b);
            b = nullptr;
        }
    } catch (...) {
        for (Job& j : jobs) vx3h_vxa_free(j.v);
        for (Job& j : history) vx3h_vxa_free(j.v);
        if (b) vx3h_batch_destroy(b);
        throw;
    }
    for (Job& j : jobs) vx3h_vxa_free(j.v);
    for (Job& j : history) vx3h_vxa_free(j.v);
}

// The playback stream of one simulation (VX3_SimulationManager.cu:40-108): the
// settings header, the real_stepsize line, then one record after every step j (the
// loop index, steps done - 1) with j % real_stepsize == 0, up to the stop condition,
// a divergence or 1000000 steps.
void SimulationManager::runHistory(vx3h_batch* b, int slot, const vx3h_vxa_info& info, double dtFrac) {
    FILE* out = opt.history_out ? opt.history_out : stdout;
    std::string text;
    auto fetch = [&](auto fn, auto... args) {
        size_t len = 0;
        check(fn(b, slot, args..., nullptr, 0, &len), "history");
        text.assign(len, '\0');
        check(fn(b, slot, args..., &text[0], len, &len), "history");
        text.resize(strlen(text.c_str()));
        fputs(text.c_str(), out);
    };
    auto status = [&]() {
        int n = 0;
        check(vx3h_batch_poll(b, nullptr, 0, &n), "poll");
        std::vector<vx3h_sim_status> sts(std::max(n, 1));
        check(vx3h_batch_poll(b, sts.data(), n, &n), "poll");
        for (int i = 0; i < n; i++) if (sts[i].slot == slot) return sts[i];
        throw std::runtime_error("history: no status");
    };
    fetch(vx3h_sim_history_header);
    // the engine computes the recommended step on the device with the first step, so
    // the first step runs before the real_stepsize line is printed (the text is the same)
    check(vx3h_batch_step(b, 1), "step");
    vx3h_sim_status s = status();
    const int real_stepsize =
        s.recommended_dt > 0 ? int(info.record_step_size / (10000 * s.recommended_dt * dtFrac)) + 1 : 1;
    fprintf(out, "real_stepsize: %d ; recommendedTimeStep %f; d_v3->DtFrac %f . \n", real_stepsize, s.recommended_dt, dtFrac);
    const long long maxSteps = 1000000;
    long long done = s.steps;
    if (done == 1 && s.state != VX3H_SIM_DIVERGED) fetch(vx3h_sim_history_record, 0, info.record_voxel, info.record_link);
    while (s.state == VX3H_SIM_RUNNING && done >= 1 && done < maxSteps) {
        // the next recorded step count: 1 + a multiple of real_stepsize
        const long long next = std::min(maxSteps, done + real_stepsize - (done - 1) % real_stepsize);
        check(vx3h_batch_step(b, (int)(next - done)), "step");
        s = status();
        if (s.steps <= done) break;
        done = s.steps;
        if (s.state == VX3H_SIM_DIVERGED) break;
        if ((done - 1) % real_stepsize == 0) fetch(vx3h_sim_history_record, (int)(done - 1), info.record_voxel, info.record_link);
    }
    fflush(out);
}

}  // namespace vx3
