// vx3h_server.cpp -- continuous-batching request loop over one device batch.
//
// The reference runs a fixed set of VXD files per process: it reads them all, starts
// one device thread per simulation and waits for the slowest
// (VX3_SimulationManager.cu:138-156; vx3_node_worker.cu:63-96).  Here requests
// arrive at any time: each is deep-copied into a queue, admitted in arrival order
// once its footprint (vx3h_sim_footprint) fits the budget left by the running
// ones, stepped with everything else on the device, and retired when it finishes,
// freeing its ranges for the next admissions.  Admission, retirement and result
// reads happen between rounds of quantum_steps steps.  Each request's host model and
// tile plan are built by the server's preparation threads from its submission on, while
// the device runs the current round, so an admission only uploads.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <system_error>
#include <thread>
#include <vector>

#include "../../include/vx3h.h"
#include "vx3h_internal.h"

using vx3::g_err;

namespace {

// an owned copy of a vx3h_sim_desc and every array it points to
struct DescCopy {
    vx3h_sim_desc d;
    std::vector<vx3h_material_desc> mats;
    std::vector<int> ijk, mat;
    std::vector<double> phase;
    std::vector<vx3h_external> exts;
    std::vector<vx3h_token> stop, fit, ff[3], att[5];
    std::vector<double> baseCilia, shiftCilia, initVel;
    void take(const vx3h_sim_desc& s) {
        d = s;
        mats.assign(s.materials, s.materials + s.n_materials);
        ijk.assign(s.ijk, s.ijk + 3 * (size_t)s.n_voxels);
        mat.assign(s.material, s.material + s.n_voxels);
        if (s.phase) phase.assign(s.phase, s.phase + s.n_voxels);
        if (s.n_externals) exts.assign(s.externals, s.externals + s.n_externals);
        if (s.n_stop_tokens) stop.assign(s.stop_tokens, s.stop_tokens + s.n_stop_tokens);
        if (s.n_fitness_tokens) fit.assign(s.fitness_tokens, s.fitness_tokens + s.n_fitness_tokens);
        d.materials = mats.data();
        d.ijk = ijk.data();
        d.material = mat.data();
        d.phase = s.phase ? phase.data() : nullptr;
        d.externals = exts.empty() ? nullptr : exts.data();
        d.stop_tokens = stop.empty() ? nullptr : stop.data();
        d.fitness_tokens = fit.empty() ? nullptr : fit.data();
        if (s.base_cilia_force) baseCilia.assign(s.base_cilia_force, s.base_cilia_force + 3 * (size_t)s.n_voxels);
        if (s.shift_cilia_force) shiftCilia.assign(s.shift_cilia_force, s.shift_cilia_force + 3 * (size_t)s.n_voxels);
        d.base_cilia_force = s.base_cilia_force ? baseCilia.data() : nullptr;
        d.shift_cilia_force = s.shift_cilia_force ? shiftCilia.data() : nullptr;
        if (s.init_velocity) initVel.assign(s.init_velocity, s.init_velocity + 3 * (size_t)s.n_voxels);
        d.init_velocity = s.init_velocity ? initVel.data() : nullptr;
        for (int k = 0; k < 3; k++) {
            if (s.n_force_field_tokens[k] > 0)
                ff[k].assign(s.force_field_tokens[k], s.force_field_tokens[k] + s.n_force_field_tokens[k]);
            d.force_field_tokens[k] = ff[k].empty() ? nullptr : ff[k].data();
        }
        for (int k = 0; k < 5; k++) {
            if (s.n_attach_tokens[k] > 0) att[k].assign(s.attach_tokens[k], s.attach_tokens[k] + s.n_attach_tokens[k]);
            d.attach_tokens[k] = att[k].empty() ? nullptr : att[k].data();
        }
    }
};

struct Request {
    long long ticket;
    DescCopy desc;
    size_t bytes = 0;
    double tSubmit = 0, tAdmit = 0;
    int slot = -1;
    // host model and tile plan (vx3::prepareSim), made by a preparation thread or, if none
    // has started on it, by the round that admits the request; state guarded by the
    // server's mutex
    enum { PREP_WAITING, PREP_RUNNING, PREP_DONE } prepState = PREP_WAITING;
    bool requeued = false;              // put back by a round (no longer counted ahead)
    vx3::PreparedSim* prep = nullptr;   // null when done: the preparation failed (admission reports why)
};

}  // namespace

struct vx3h_server {
    vx3h_server_cfg cfg{};
    vx3h_batch* batch = nullptr;
    size_t budget = 0, used = 0;
    std::chrono::steady_clock::time_point t0;
    std::mutex mu;                      // guards queue, nextTicket, prepQueue, stopping and Request::prepState
    std::deque<Request*> queue;
    std::deque<Request*> prepQueue;     // submitted, not yet prepared, in submission order
    std::condition_variable prepWork;   // a request to prepare, or stopping
    std::condition_variable prepDone;   // a preparation finished
    std::vector<std::thread> prepThreads;
    bool stopping = false;
    int ahead = 0;                      // requests the threads took and no round has yet (host memory bound)
    long long nextTicket = 1;
    std::map<int, Request*> running;    // slot -> request
    std::vector<vx3h_server_result> done;
    double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
};

namespace {

// at most this many prepared requests wait for admission (each holds its host model)
constexpr int kPrepAhead = 256;

void prepWorker(vx3h_server* sv) {
    std::unique_lock<std::mutex> lk(sv->mu);
    for (;;) {
        sv->prepWork.wait(lk, [&] { return sv->stopping || (!sv->prepQueue.empty() && sv->ahead < kPrepAhead); });
        if (sv->stopping) return;
        Request* r = sv->prepQueue.front();
        sv->prepQueue.pop_front();
        sv->ahead++;
        r->prepState = Request::PREP_RUNNING;
        lk.unlock();
        vx3::PreparedSim* p = nullptr;
        if (vx3::prepareSim(sv->batch, &r->desc.d, &p) != VX3H_OK) p = nullptr;
        lk.lock();
        r->prep = p;
        r->prepState = Request::PREP_DONE;
        sv->prepDone.notify_all();
    }
}

}  // namespace

extern "C" double vx3h_server_clock(const vx3h_server* sv) { return sv ? sv->now() : 0.0; }

extern "C" vx3h_status vx3h_server_create(int device, const vx3h_server_cfg* cfg, vx3h_server** out) {
    if (!cfg || !out) { g_err = "null argument"; return VX3H_EINVAL; }
    *out = nullptr;
    vx3h_server* sv = new vx3h_server();
    sv->cfg = *cfg;
    if (sv->cfg.quantum_steps <= 0) sv->cfg.quantum_steps = 1000;
    vx3h_status st = vx3h_batch_create(device, &sv->cfg.batch, &sv->batch);
    if (st != VX3H_OK) { delete sv; return st; }
    sv->budget = sv->cfg.batch.max_device_bytes;
    sv->t0 = std::chrono::steady_clock::now();
    // preparation threads: host work only, one core left to the thread that drives the device
    const unsigned hw = std::max(2u, std::thread::hardware_concurrency());
    try {
        for (unsigned t = 0; t < std::min(15u, hw - 1); t++) sv->prepThreads.emplace_back(prepWorker, sv);
    } catch (const std::system_error&) {   // fewer threads (none: each round prepares its own admissions)
    }
    *out = sv;
    return VX3H_OK;
}

extern "C" vx3h_status vx3h_server_destroy(vx3h_server* sv) {
    if (!sv) return VX3H_OK;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->stopping = true;
    }
    sv->prepWork.notify_all();
    for (auto& th : sv->prepThreads) th.join();
    vx3h_batch_destroy(sv->batch);
    for (Request* r : sv->queue) { vx3::freePrepared(r->prep); delete r; }
    for (auto& kv : sv->running) delete kv.second;
    delete sv;
    return VX3H_OK;
}

extern "C" vx3h_status vx3h_server_submit(vx3h_server* sv, const vx3h_sim_desc* desc, int max_steps, long long* ticket) {
    if (!sv || !desc) { g_err = "null argument"; return VX3H_EINVAL; }
    Request* r = new Request();
    r->desc.take(*desc);
    if (max_steps > 0) r->desc.d.max_steps = max_steps;
    vx3h_status st = vx3h_sim_footprint(&r->desc.d, sv->cfg.batch.precision, &r->bytes);
    if (st != VX3H_OK) { delete r; return st; }
    if (sv->budget && r->bytes > sv->budget) { delete r; g_err = "request larger than the server's device budget"; return VX3H_ENOMEM; }
    std::lock_guard<std::mutex> lk(sv->mu);
    r->ticket = sv->nextTicket++;
    r->tSubmit = sv->now();
    sv->queue.push_back(r);
    if (!sv->prepThreads.empty()) {
        sv->prepQueue.push_back(r);
        sv->prepWork.notify_one();
    }
    if (ticket) *ticket = r->ticket;
    return VX3H_OK;
}

extern "C" vx3h_status vx3h_server_round(vx3h_server* sv, int* active) {
    if (!sv) { g_err = "null server"; return VX3H_EINVAL; }
    // admit in arrival order while the footprint fits, each with its preparation: those
    // no preparation thread has started are prepared here on several threads, then the
    // round waits for the ones in progress
    std::vector<Request*> take, mine;
    std::vector<vx3::PreparedSim*> prep;
    bool freed = false;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        size_t bytes = 0;
        while (!sv->queue.empty()) {
            Request* h = sv->queue.front();
            if (sv->cfg.max_sims > 0 && (int)(sv->running.size() + take.size()) >= sv->cfg.max_sims) break;
            if (sv->budget && sv->used + bytes + h->bytes > sv->budget) break;
            sv->queue.pop_front();
            take.push_back(h);
            bytes += h->bytes;
            if (h->prepState == Request::PREP_WAITING) {
                h->prepState = Request::PREP_RUNNING;
                mine.push_back(h);
            } else if (!h->requeued) {
                sv->ahead--;   // a preparation thread's: room for the next
                freed = true;
            }
        }
        if (!mine.empty())   // no longer for the preparation threads
            sv->prepQueue.erase(std::remove_if(sv->prepQueue.begin(), sv->prepQueue.end(),
                                               [](Request* r) { return r->prepState != Request::PREP_WAITING; }),
                                sv->prepQueue.end());
    }
    if (freed) sv->prepWork.notify_all();
    {
        const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
        const size_t nt = std::min<size_t>(mine.size(), std::min(16u, hw));
        std::atomic<size_t> next{0};
        auto work = [&]() {
            for (size_t i; (i = next.fetch_add(1)) < mine.size();)
                if (vx3::prepareSim(sv->batch, &mine[i]->desc.d, &mine[i]->prep) != VX3H_OK) mine[i]->prep = nullptr;
        };
        std::vector<std::thread> pool;
        try {
            for (size_t t = 1; t < nt; t++) pool.emplace_back(work);
        } catch (const std::system_error&) {   // this thread and those started do the rest
        }
        work();
        for (auto& th : pool) th.join();
        std::unique_lock<std::mutex> lk(sv->mu);
        for (Request* r : mine) r->prepState = Request::PREP_DONE;
        sv->prepDone.wait(lk, [&] {
            for (Request* r : take)
                if (r->prepState != Request::PREP_DONE) return false;
            return true;
        });
        for (Request* r : take) {
            prep.push_back(r->prep);
            r->prep = nullptr;
        }
    }
    if (take.size() > 1) {   // one growth of the batch arrays for the whole round
        long long nv = 0, nl = 0;
        for (vx3::PreparedSim* p : prep) {
            long long a, c;
            vx3::preparedSize(p, &a, &c);
            nv += a;
            nl += c;
        }
        vx3::reserveBatch(sv->batch, nv, nl);   // best effort: admission grows what is still short
    }
    for (size_t i = 0; i < take.size(); i++) {
        Request* r = take[i];
        int slot = -1;
        vx3h_status st = vx3::admitPrepared(sv->batch, &r->desc.d, prep[i], &slot);   // frees prep[i]
        prep[i] = nullptr;
        if (st == VX3H_ENOMEM && !sv->running.empty()) {   // the engine's own budget: wait for retirements
            std::lock_guard<std::mutex> lk(sv->mu);
            for (size_t j = take.size(); j-- > i;) {
                if (j > i) { take[j]->prep = prep[j]; prep[j] = nullptr; }   // kept for the next round
                else take[j]->prepState = Request::PREP_WAITING;             // its preparation went with the admission
                take[j]->requeued = true;
                sv->queue.push_front(take[j]);
            }
            break;
        }
        if (st != VX3H_OK) {
            // a request that cannot run finishes at once, with its error as a failed state
            vx3h_server_result res{};
            res.ticket = r->ticket;
            res.status.state = VX3H_SIM_FAILED;
            res.status.slot = -1;
            res.t_submit = r->tSubmit;
            res.t_admit = res.t_done = sv->now();
            sv->done.push_back(res);
            delete r;
            continue;
        }
        r->slot = slot;
        r->tAdmit = sv->now();
        sv->used += r->bytes;
        sv->running[slot] = r;
    }
    if (!sv->running.empty()) {
        vx3h_status st = vx3h_batch_step(sv->batch, sv->cfg.quantum_steps);
        if (st != VX3H_OK) return st;
        std::vector<vx3h_sim_status> sts;   // results only of the simulations that stopped
        if ((st = vx3::pollFinished(sv->batch, sts)) != VX3H_OK) return st;
        const double t = sv->now();
        for (size_t i = 0; i < sts.size(); i++) {
            auto it = sv->running.find(sts[i].slot);
            if (it == sv->running.end()) continue;
            Request* r = it->second;
            vx3h_server_result res{};
            res.ticket = r->ticket;
            res.status = sts[i];
            res.t_submit = r->tSubmit;
            res.t_admit = r->tAdmit;
            res.t_done = t;
            res.voxel_steps = (double)sts[i].num_voxel * sts[i].steps;
            sv->done.push_back(res);
            if ((st = vx3h_sim_retire(sv->batch, r->slot)) != VX3H_OK) return st;
            sv->used -= r->bytes;
            sv->running.erase(it);
            delete r;
        }
    }
    if (active) {
        std::lock_guard<std::mutex> lk(sv->mu);
        *active = (int)(sv->running.size() + sv->queue.size());
    }
    return VX3H_OK;
}

extern "C" vx3h_status vx3h_server_results(vx3h_server* sv, vx3h_server_result* out, int cap, int* n) {
    if (!sv || !n) { g_err = "null argument"; return VX3H_EINVAL; }
    if (!out || cap <= 0) { *n = (int)sv->done.size(); return VX3H_OK; }
    const int k = std::min(cap, (int)sv->done.size());
    for (int i = 0; i < k; i++) out[i] = sv->done[i];
    sv->done.erase(sv->done.begin(), sv->done.begin() + k);
    *n = k;
    return VX3H_OK;
}
