// Serving loop (include/rin/serve.h): FIFO admission by free slot and HBM
// budget, one context (stream + memory pool) per slot, so admitted jobs run
// on separate streams and overlap on the device; grids kept resident per slot
// and resolution, meshes handed back in host memory.  The cut kernels keep
// their per-tet state in LDS or device workspaces (at most a few KB of
// private memory a lane), so several queues can run them at once.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "rin.h"
#include "rin/serve.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Job {
    uint64_t id = 0;
    rin_job_desc desc{};
    std::vector<rin_func_desc> funcs;
    uint64_t est = 0;
    int32_t status = 1;
    int32_t slot = -1;
    double t_submit = 0, t_start = 0, t_end = 0;
    rin_sizes sizes{};
    std::vector<uint64_t> rec11, vstart, verts, tfstart, tf, func;
    std::vector<double> xyz;
    std::string error;
};

// device bytes one plain-form run needs at most (values, per-vertex and
// per-tet stage buffers, records and outputs of an active fraction of tets)
uint64_t estimate_bytes(int res, uint32_t F) {
    const uint64_t N = (uint64_t)res + 1, V = N * N * N, T = 5ull * res * res * res;
    return V * (8ull * F + 48) + T * (24 + 16) + T / 2 * 420 + (64ull << 20);
}

}  // namespace

struct rin_server {
    int device = 0;
    uint64_t budget = 0;
    Clock::time_point t0 = Clock::now();
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<Job>> queue;
    std::unordered_map<uint64_t, std::shared_ptr<Job>> jobs;
    std::vector<std::thread> workers;
    std::vector<rin_ctx*> ctx;
    uint64_t next_id = 1, in_use = 0, in_flight = 0, max_in_flight = 0;
    uint64_t submitted = 0, completed = 0, failed = 0;
    std::vector<double> latencies;
    uint64_t tets = 0;
    double first_submit = -1, last_end = 0;
    bool stop = false;
    std::string err;

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }

    void worker(int slot) {
        rin_ctx* c = ctx[slot];
        std::map<int, rin_grid*> grids;  // resident per resolution
        for (;;) {
            std::shared_ptr<Job> J;
            {
                std::unique_lock<std::mutex> lk(mu);
                // FIFO: the head job starts when its memory fits (or nothing runs)
                cv_work.wait(lk, [&] {
                    return stop || (!queue.empty() && rin_server_admits(in_flight, in_use, queue.front()->est, budget));
                });
                if (stop && queue.empty()) break;
                if (queue.empty()) continue;
                J = queue.front();
                queue.pop_front();
                in_use += J->est;
                ++in_flight;
                max_in_flight = std::max(max_in_flight, in_flight);
                J->slot = slot;
                J->t_start = now();
            }
            cv_work.notify_all();
            std::string error;
            const int st = run(c, grids, *J, error);
            {
                std::lock_guard<std::mutex> lk(mu);
                J->status = st;
                J->error = error;
                in_use -= J->est;
                --in_flight;
                J->t_end = now();
                if (J->status == RIN_OK) {
                    ++completed;
                    latencies.push_back(J->t_end - J->t_submit);
                    tets += J->sizes.n_tets;
                } else {
                    ++failed;
                }
                last_end = std::max(last_end, J->t_end);
            }
            cv_work.notify_all();
            cv_done.notify_all();
        }
        for (auto& g : grids) rin_grid_release(c, g.second);
    }

    // runs one job on slot context c; the mesh goes to J's host arrays (the
    // status is published by the caller, under the lock)
    int run(rin_ctx* c, std::map<int, rin_grid*>& grids, Job& J, std::string& error) {
        rin_values* v = nullptr;
        rin_result* r = nullptr;
        int st = RIN_OK;
        auto fail = [&](int code) {
            error = rin_last_error(c);
            return code;
        };
        rin_grid*& g = grids[J.desc.res];
        if (!g) {
            const double lo[3] = {-1, -1, -1}, hi[3] = {1, 1, 1};
            st = rin_grid_regular(c, J.desc.res, lo, hi, &g);
            if (st != RIN_OK) {
                g = nullptr;
                grids.erase(J.desc.res);
                return fail(st);
            }
        }
        if ((st = rin_values_eval(c, g, J.funcs.data(), (uint32_t)J.funcs.size(), &v)) != RIN_OK) return fail(st);
        // implicit arrangements on the batched form (its plan, graph and tables
        // stay with the slot's grid for the next job of the same size)
        const uint32_t flags =
            (J.desc.mode == 1 ? RIN_FLAG_MI : (RIN_FLAG_BATCHED | RIN_FLAG_FIXED_FORMS)) | J.desc.flags;
        st = rin_run(c, g, v, flags, &r);
        if (st == RIN_OK) st = rin_result_sizes(c, r, &J.sizes);
        if (st == RIN_OK) {
            const rin_sizes& s = J.sizes;
            J.rec11.resize(11 * s.n_verts + 1);
            J.xyz.resize(3 * s.n_verts + 1);
            J.vstart.resize(s.n_faces + 1);
            J.verts.resize(s.n_face_verts + 1);
            J.tfstart.resize(s.n_faces + 1);
            J.tf.resize(2 * s.n_tet_faces + 1);
            J.func.resize(2 * s.n_faces + 1);
            st = rin_result_verts(c, r, J.rec11.data(), J.xyz.data());
            if (st == RIN_OK)
                st = rin_result_faces(c, r, J.vstart.data(), J.verts.data(), J.tfstart.data(), J.tf.data(),
                                      J.func.data());
        }
        if (st != RIN_OK) fail(st);
        if (r) rin_result_release(c, r);
        if (v) rin_values_release(c, v);
        return st;
    }
};

namespace {

int bad(rin_server* s, int code, const std::string& m) {
    if (s) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->err = m;
    }
    return code;
}

std::shared_ptr<Job> find(rin_server* s, uint64_t id, bool finished_only = false) {
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return nullptr;
    if (finished_only && it->second->status != RIN_OK) return nullptr;
    return it->second;
}

}  // namespace

extern "C" {

int rin_server_admits(uint64_t in_flight, uint64_t in_use, uint64_t estimate, uint64_t budget) {
    return in_flight == 0 || in_use + estimate <= budget;
}

int rin_server_create(int32_t device, int32_t n_slots, uint64_t hbm_budget, rin_server** out) {
    if (!out || n_slots < 1 || n_slots > 16) return RIN_ERR_ARG;
    auto* s = new rin_server;
    s->device = device;
    for (int i = 0; i < n_slots; ++i) {
        rin_ctx* c = nullptr;
        const int st = rin_create(&c, device);
        if (st != RIN_OK) {
            for (auto* x : s->ctx) rin_destroy(x);
            delete s;
            return st;
        }
        s->ctx.push_back(c);
    }
    uint64_t fr = 0, tot = 0;
    if (rin_device_memory(s->ctx[0], &fr, &tot) != RIN_OK) fr = 0;
    s->budget = hbm_budget ? hbm_budget : fr / 10 * 8;
    for (int i = 0; i < n_slots; ++i) s->workers.emplace_back([s, i] { s->worker(i); });
    *out = s;
    return RIN_OK;
}

int rin_server_destroy(rin_server* s) {
    if (!s) return RIN_ERR_ARG;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto& w : s->workers) w.join();
    for (auto* c : s->ctx) rin_destroy(c);
    delete s;
    return RIN_OK;
}

const char* rin_server_last_error(const rin_server* s) { return s ? s->err.c_str() : "null server"; }

double rin_server_now(const rin_server* s) { return s ? s->now() : 0.0; }

int rin_server_submit(rin_server* s, const rin_job_desc* d, uint64_t* id) {
    if (!s || !d || !id || (d->n_funcs && !d->funcs)) return bad(s, RIN_ERR_ARG, "null argument");
    if (d->mode != 0 && d->mode != 1) return bad(s, RIN_ERR_ARG, "mode must be 0 or 1");
    if (d->res < 1 || d->n_funcs < 1 || d->n_funcs > 64) return bad(s, RIN_ERR_ARG, "bad grid or function count");
    auto J = std::make_shared<Job>();
    J->desc = *d;
    J->funcs.assign(d->funcs, d->funcs + d->n_funcs);
    J->desc.funcs = nullptr;
    J->est = estimate_bytes(d->res, d->n_funcs);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        J->id = s->next_id++;
        J->t_submit = s->now();
        if (s->first_submit < 0) s->first_submit = J->t_submit;
        ++s->submitted;
        s->jobs[J->id] = J;
        s->queue.push_back(J);
        *id = J->id;
    }
    s->cv_work.notify_all();
    return RIN_OK;
}

int rin_server_wait(rin_server* s, uint64_t id, double timeout_s, rin_job_info* info) {
    if (!s) return RIN_ERR_ARG;
    auto J = find(s, id);
    if (!J) return bad(s, RIN_ERR_ARG, "unknown job");
    std::unique_lock<std::mutex> lk(s->mu);
    const bool done = s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), [&] { return J->status != 1; });
    if (info) {
        info->id = J->id;
        info->status = J->status;
        info->slot = J->slot;
        info->t_submit = J->t_submit;
        info->t_start = J->t_start;
        info->t_end = J->t_end;
        info->hbm_estimate = J->est;
        info->sizes = J->sizes;
    }
    if (!done) {
        s->err = "timed out waiting for job " + std::to_string(id);
        return RIN_ERR_INTERNAL;
    }
    if (J->status != RIN_OK) s->err = J->error;
    return J->status;
}

int rin_server_job_verts(rin_server* s, uint64_t id, uint64_t* rec11, double* xyz) {
    auto J = s ? find(s, id, true) : nullptr;
    if (!J) return bad(s, RIN_ERR_ARG, "job unknown, failed or not finished");
    if (rec11) std::memcpy(rec11, J->rec11.data(), 8 * 11 * J->sizes.n_verts);
    if (xyz) std::memcpy(xyz, J->xyz.data(), 8 * 3 * J->sizes.n_verts);
    return RIN_OK;
}

int rin_server_job_faces(rin_server* s, uint64_t id, uint64_t* vstart, uint64_t* verts, uint64_t* tfstart,
                         uint64_t* tf, uint64_t* func) {
    auto J = s ? find(s, id, true) : nullptr;
    if (!J) return bad(s, RIN_ERR_ARG, "job unknown, failed or not finished");
    const rin_sizes& z = J->sizes;
    if (vstart) std::memcpy(vstart, J->vstart.data(), 8 * (z.n_faces + 1));
    if (verts) std::memcpy(verts, J->verts.data(), 8 * z.n_face_verts);
    if (tfstart) std::memcpy(tfstart, J->tfstart.data(), 8 * (z.n_faces + 1));
    if (tf) std::memcpy(tf, J->tf.data(), 8 * 2 * z.n_tet_faces);
    if (func) std::memcpy(func, J->func.data(), 8 * 2 * z.n_faces);
    return RIN_OK;
}

int rin_server_job_release(rin_server* s, uint64_t id) {
    if (!s) return RIN_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end() || it->second->status == 1) {
        s->err = "job unknown or not finished";
        return RIN_ERR_ARG;
    }
    s->jobs.erase(it);
    return RIN_OK;
}

int rin_server_stats_get(rin_server* s, rin_server_stats* o) {
    if (!s || !o) return RIN_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    std::memset(o, 0, sizeof(*o));
    o->submitted = s->submitted;
    o->completed = s->completed;
    o->failed = s->failed;
    o->max_in_flight = s->max_in_flight;
    o->tets = s->tets;
    o->first_submit = s->first_submit;
    o->last_end = s->last_end;
    o->hbm_budget = s->budget;
    if (!s->latencies.empty()) {
        std::vector<double> l = s->latencies;
        std::sort(l.begin(), l.end());
        auto q = [&](double p) {
            const size_t i = (size_t)(p * (double)(l.size() - 1) + 0.5);
            return l[std::min(i, l.size() - 1)];
        };
        o->p50_latency = q(0.50);
        o->p99_latency = q(0.99);
        o->max_latency = l.back();
    }
    return RIN_OK;
}

}  // extern "C"
