// Continuous-batching flow server (include/lgf_flow.h, lgf_server_*).
//
// No counterpart in the reference, which flows saved configurations one after
// another from a CLI (LGF/LGF.cpp:163-196).  Here flow jobs arrive at any
// time; one scheduler thread owns the GPU context and loops:
//   1. admit queued jobs, first come first served, while their device
//      footprint (2 link buffers + RK3 accumulator) fits the budget, less
//      the measurement scratch (line products: up to 1 GiB per group of
//      lattices, or one lattice's worth if that is more);
//   2. choose the jobs that step: every admitted job (LGF_SERVE_ALL), or
//      (LGF_SERVE_SHORTEST_FIRST) the admitted jobs with the least remaining
//      link-updates, in that order, until they hold min_sites lattice sites,
//      so short jobs finish first while each launch still fills the GPU; the
//      choice is revisited when a job arrives or finishes;
//      run k RK3 steps for those jobs with one set of launches per precision
//      (a batch over their lattices, whatever their shapes; k = steps to the
//      nearest measurement, completion or admission tick);
//   3. measure the jobs that are due, retire finished jobs and free their
//      memory, which lets the next queued jobs in.
// New jobs are admitted only at ticks of kTick global steps (immediately
// when the server is idle), so jobs with a measurement interval dividing
// kTick reach their measurements together: the due jobs of a batch are then
// measured by one sweep over the batch, and a round is never cut short by
// another job's measurement.  The host does not wait for the GPU between
// rounds; it waits at measurements, which return results.
// Every job's arithmetic is independent of the other jobs in the batch, and
// its reductions run over its own tiles in a fixed order, so a job's results
// are bitwise the same alone and in any mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/lgf_flow.h"
#include "kernels.hpp"

namespace {

constexpr int kTick = 10;  // admission points are this many global steps apart
constexpr int kMaxMeasurements = 1024;

struct Job {
    uint64_t id = 0;
    lgf_request req{};
    std::vector<double> links;  // copy of the caller's initial links (may be empty)
    size_t bytes = 0;
    size_t line_bytes = 0;  // line products of one measurement of this job
    lgf_field* field = nullptr;
    int done = 0;
    double t = 0.0;
    int status = LGF_OK;
    double submit_s = 0, start_s = 0, finish_s = 0;
    std::vector<double> tm, E, Q;
};

using JobP = std::shared_ptr<Job>;

}  // namespace

struct lgf_server {
    int device = 0;
    size_t budget = 0;
    int max_active = 0;
    int policy = LGF_SERVE_SHORTEST_FIRST;  // guarded by mu
    int64_t min_sites = int64_t(1) << 20;
    double age_rate = 0.0;
    lgf_ctx* ctx = nullptr;
    std::thread thread;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<JobP> pending;
    std::map<uint64_t, JobP> finished;
    uint64_t next_id = 1;
    bool stop = false;
    bool tuning = false;  // lgf_server_autotune is using ctx: admit nothing
    int fail_batches = 0; // lgf_server_inject_batch_failures: batch creations still to fail
    std::string err;
    std::chrono::steady_clock::time_point t0;
    // counters (guarded by mu)
    uint64_t jobs_done = 0, batch_steps = 0;
    double link_updates = 0.0, busy_s = 0.0;

    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    void set_err(const std::string& m) {
        std::lock_guard<std::mutex> lk(mu);
        err = m;
    }
    void run();
};

static int64_t volume_of(const int* s) { return (int64_t)s[0] * s[1] * s[2] * s[3]; }

// device scratch a measurement of these jobs can hold: the line products of
// a group of lattices (the context's default cap) or of the largest one
constexpr size_t kMeasureScratchCap = size_t(1) << 30;
static size_t measure_reserve(size_t largest_line_bytes) { return std::max(kMeasureScratchCap, largest_line_bytes); }

static void record(Job& j, double E, double Q) {
    if ((int)j.tm.size() < kMaxMeasurements) {
        j.tm.push_back(j.t);
        j.E.push_back(E);
        j.Q.push_back(Q);
    }
}

static void measure_job(lgf_server* S, Job& j) {
    double E = 0, Q = 0, Qa = 0;
    int rc = lgf_measure(j.field, j.req.measure_flags, &E, &Q, &Qa, nullptr);
    if (rc != LGF_OK && rc != LGF_E_NAN) {
        j.status = rc;
        S->set_err(lgf_ctx_last_error(S->ctx));
        return;
    }
    record(j, E, Q);
}

// Measure the due jobs.  When every stepping job of a precision is due with
// the same flags, one sweep over that precision's batch measures them all;
// otherwise each due job's sweep is enqueued first and the results are
// collected afterwards, so the stream does not wait for the host in between.
static void measure_due(lgf_server* S, const std::vector<JobP>& stepping, const std::vector<JobP>& due_now,
                        lgf_batch* const batch[2]) {
    std::vector<JobP> rest;
    for (int p = 0; p < 2; ++p) {
        std::vector<JobP> in_batch, due_p;
        for (auto& j : stepping)
            if (j->req.precision == p) in_batch.push_back(j);
        for (auto& j : due_now)
            if (j->req.precision == p) due_p.push_back(j);
        if (due_p.empty()) continue;
        bool whole = batch[p] && due_p.size() == in_batch.size();
        for (auto& j : due_p) whole = whole && j->req.measure_flags == due_p[0]->req.measure_flags;
        if (whole) {
            std::vector<double> obs(LGF_NUM_OBS * in_batch.size());
            int rc = lgf_batch_measure_observables(batch[p], due_p[0]->req.measure_flags, obs.data());
            if (rc == LGF_OK || rc == LGF_E_NAN) {
                for (size_t i = 0; i < in_batch.size(); ++i)
                    record(*in_batch[i], obs[LGF_NUM_OBS * i + LGF_OBS_E], obs[LGF_NUM_OBS * i + LGF_OBS_Q]);
                continue;
            }
            S->set_err(lgf_ctx_last_error(S->ctx));
            for (auto& j : in_batch) j->status = rc;
            continue;
        }
        rest.insert(rest.end(), due_p.begin(), due_p.end());
    }
    std::vector<JobP> started;
    for (auto& j : rest) {
        int rc = lgf_measure_start(j->field, j->req.measure_flags);
        if (rc == LGF_OK) {
            started.push_back(j);
        } else {
            j->status = rc;
            S->set_err(lgf_ctx_last_error(S->ctx));
        }
    }
    for (auto& j : started) {
        double obs[LGF_NUM_OBS];
        int rc = lgf_measure_finish(j->field, obs);
        if (rc == LGF_OK || rc == LGF_E_NAN) {
            record(*j, obs[LGF_OBS_E], obs[LGF_OBS_Q]);
        } else {
            j->status = rc;
            S->set_err(lgf_ctx_last_error(S->ctx));
        }
    }
}

static double remaining_work(const Job& j) { return 4.0 * (double)volume_of(j.req.shape) * (j.req.nsteps - j.done); }

// measurements a job records: t = 0, every multiple of measure_every up to
// nsteps, and the end when it is not such a multiple (measure_every = 0: t = 0
// and the end)
static int64_t measurement_count(int nsteps, int every) {
    if (nsteps == 0) return 1;
    if (every == 0) return 2;
    return 1 + nsteps / every + (nsteps % every ? 1 : 0);
}

static bool due(const Job& j) {
    const int every = j.req.measure_every;
    return j.done == j.req.nsteps || (every > 0 && j.done % every == 0);
}

void lgf_server::run() {
    std::vector<JobP> active;    // admitted and resident
    std::vector<JobP> stepping;  // the subset that steps in this round
    lgf_batch* batch[2] = {nullptr, nullptr};
    bool dirty = true;
    int64_t gstep = 0;  // global steps since the server was last idle
    while (true) {
        std::vector<JobP> admitted;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_work.wait(lk, [&] { return stop || (!pending.empty() && !tuning) || !active.empty(); });
            if (stop) break;
            if (active.empty()) gstep = 0;
            size_t used = 0, lines = 0;
            for (auto& j : active) {
                used += j->bytes;
                lines = std::max(lines, j->line_bytes);
            }
            while (gstep % kTick == 0 && !pending.empty() && !tuning) {
                JobP j = pending.front();
                if (used + j->bytes + measure_reserve(std::max(lines, j->line_bytes)) > budget) break;
                if (max_active > 0 && (int)(active.size() + admitted.size()) >= max_active) break;
                pending.pop_front();
                used += j->bytes;
                lines = std::max(lines, j->line_bytes);
                admitted.push_back(j);
            }
        }
        for (auto& j : admitted) {
            j->start_s = now();
            int rc = lgf_field_create(ctx, j->req.shape, j->req.precision, &j->field);
            if (rc == LGF_OK) {
                if (!j->links.empty())
                    rc = lgf_field_upload(j->field, j->links.data());
                else if (j->req.start == 1)
                    rc = lgf_field_set_random(j->field, j->req.seed);
                j->links.clear();
                j->links.shrink_to_fit();
            }
            if (rc == LGF_OK) {
                measure_job(this, *j);  // t = 0 (LGF/LGF.cpp:178-184)
                rc = j->status;
            }
            if (rc != LGF_OK || j->req.nsteps == 0) {
                j->status = rc;
                if (rc != LGF_OK) set_err(lgf_ctx_last_error(ctx));
                if (j->field) lgf_field_destroy(j->field);
                j->field = nullptr;
                j->finish_s = now();
                std::lock_guard<std::mutex> lk(mu);
                finished[j->id] = j;
                jobs_done++;
                cv_done.notify_all();
                continue;
            }
            active.push_back(j);
            dirty = true;
        }
        if (active.empty()) continue;
        const double tb = now();
        if (dirty) {
            int pol;
            int64_t msites;
            double arate;
            {
                std::lock_guard<std::mutex> lk(mu);
                pol = policy;
                msites = min_sites;
                arate = age_rate;
            }
            stepping = active;
            if (pol == LGF_SERVE_SHORTEST_FIRST) {
                // remaining link-updates, less a credit for the time waited
                const double tnow = now();
                auto key = [&](const JobP& j) { return remaining_work(*j) - arate * (tnow - j->submit_s); };
                std::stable_sort(stepping.begin(), stepping.end(),
                                 [&](const JobP& a, const JobP& b) { return key(a) < key(b); });
                int64_t v = 0;
                size_t n = 0;
                while (n < stepping.size() && (n == 0 || v < msites)) v += volume_of(stepping[n++]->req.shape);
                stepping.resize(n);
            }
            for (int p = 0; p < 2; ++p) {
                if (batch[p]) lgf_batch_destroy(batch[p]);
                batch[p] = nullptr;
                std::vector<lgf_field*> fs;
                for (auto& j : stepping)
                    if (j->req.precision == p) fs.push_back(j->field);
                if (fs.empty()) continue;
                int rc;
                {
                    std::lock_guard<std::mutex> lk(mu);
                    rc = fail_batches > 0 ? LGF_E_NOMEM : LGF_OK;
                    if (fail_batches > 0) fail_batches--;
                }
                if (rc == LGF_OK) {
                    rc = lgf_batch_create(ctx, fs.data(), (int)fs.size(), &batch[p]);
                    if (rc != LGF_OK) set_err(std::string("batch creation failed: ") + lgf_ctx_last_error(ctx));
                } else {
                    set_err("batch creation failed (injected fault)");
                }
                if (rc != LGF_OK) {
                    // these jobs cannot step: they fail with the status and are retired below
                    batch[p] = nullptr;
                    for (auto& j : stepping)
                        if (j->req.precision == p) j->status = rc;
                }
            }
            dirty = false;
        }
        // steps to the nearest event of any stepping job, and to the next tick
        int k = kTick - (int)(gstep % kTick);
        for (auto& j : stepping) {
            int to_end = j->req.nsteps - j->done;
            int every = j->req.measure_every;
            int to_meas = every > 0 ? every - j->done % every : to_end;
            k = std::min(k, std::min(to_end, to_meas));
        }
        k = std::max(k, 1);
        double lu = 0.0;
        for (int p = 0; p < 2; ++p) {
            if (!batch[p]) continue;
            std::vector<double> eps;
            for (auto& j : stepping)
                if (j->req.precision == p) {
                    eps.push_back(j->req.eps);
                    lu += 4.0 * (double)volume_of(j->req.shape) * k;
                }
            int rc = lgf_batch_flow_steps(batch[p], eps.data(), k, 1);
            if (rc != LGF_OK) {
                set_err(lgf_ctx_last_error(ctx));
                for (auto& j : stepping)
                    if (j->req.precision == p) j->status = rc;
            }
        }
        gstep += k;
        std::vector<JobP> done_now, due_now;
        for (auto& j : stepping) {
            if (j->status != LGF_OK) continue;  // failed: did not step, retired below
            for (int s = 0; s < k; ++s) j->t += j->req.eps;  // the reference's t += eps per step
            j->done += k;
            if (due(*j)) due_now.push_back(j);
        }
        measure_due(this, stepping, due_now, batch);
        for (auto& j : stepping)
            if (j->done >= j->req.nsteps || j->status != LGF_OK) done_now.push_back(j);
        std::vector<JobP> still;
        for (auto& j : active)
            if (std::find(done_now.begin(), done_now.end(), j) == done_now.end()) still.push_back(j);
        if (!done_now.empty()) {
            // batches hold the fields: release them before the fields
            for (int p = 0; p < 2; ++p) {
                if (batch[p]) lgf_batch_destroy(batch[p]);
                batch[p] = nullptr;
            }
            dirty = true;
            for (auto& j : done_now) {
                lgf_field_destroy(j->field);
                j->field = nullptr;
                j->finish_s = now();
                std::lock_guard<std::mutex> lk(mu);
                finished[j->id] = j;
                jobs_done++;
                cv_done.notify_all();
            }
        }
        active.swap(still);
        {
            std::lock_guard<std::mutex> lk(mu);
            batch_steps += k;
            link_updates += lu;
            busy_s += now() - tb;
        }
    }
    for (int p = 0; p < 2; ++p)
        if (batch[p]) lgf_batch_destroy(batch[p]);
    std::lock_guard<std::mutex> lk(mu);
    for (auto& j : active) {
        if (j->field) lgf_field_destroy(j->field);
        j->field = nullptr;
        j->status = LGF_E_STATE;
        finished[j->id] = j;
    }
    for (auto& j : pending) {
        j->status = LGF_E_STATE;
        finished[j->id] = j;
    }
    pending.clear();
    cv_done.notify_all();
}

extern "C" int lgf_server_create(int device, size_t budget_bytes, int max_active, lgf_server** out) {
    if (!out || max_active < 0) return LGF_E_ARG;
    *out = nullptr;
    lgf_ctx* ctx = nullptr;
    int rc = lgf_ctx_create(device, &ctx);
    if (rc) return rc;
    if (budget_bytes == 0) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
            lgf_ctx_destroy(ctx);
            return LGF_E_HIP;
        }
        budget_bytes = (size_t)(0.8 * (double)fr);
    }
    lgf_server* s = new lgf_server();
    s->device = device;
    s->budget = budget_bytes;
    s->max_active = max_active;
    s->ctx = ctx;
    s->t0 = std::chrono::steady_clock::now();
    s->thread = std::thread([s] {
        (void)hipSetDevice(s->device);
        s->run();
    });
    *out = s;
    return LGF_OK;
}

extern "C" int lgf_server_destroy(lgf_server* s) {
    if (!s) return LGF_E_ARG;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->thread.joinable()) s->thread.join();
    lgf_ctx_destroy(s->ctx);
    delete s;
    return LGF_OK;
}

extern "C" int lgf_server_submit(lgf_server* s, const lgf_request* req, uint64_t* id) {
    if (!s || !req || !id) return LGF_E_ARG;
    size_t bytes = 0;
    int rc = lgf_field_footprint(req->shape, req->precision, &bytes);
    if (rc) return rc;
    if (req->nsteps < 0 || req->measure_every < 0 || !(req->eps == req->eps)) return LGF_E_ARG;
    if (measurement_count(req->nsteps, req->measure_every) > kMaxMeasurements) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->err = "request records " + std::to_string(measurement_count(req->nsteps, req->measure_every)) +
                 " measurements, more than the " + std::to_string(kMaxMeasurements) + " a result holds";
        return LGF_E_ARG;
    }
    const bool fp64_lines = req->precision == LGF_FP64 || (req->measure_flags & LGF_MEASURE_FP64);
    // the line products of one measurement: whole 64-site tiles, as the
    // measurement allocates them
    const size_t tiles = (size_t(volume_of(req->shape)) + lgf::kTileSites - 1) / lgf::kTileSites;
    const size_t line_bytes =
        tiles * lgf::kTileSites * lgf::line_bytes_per_site(fp64_lines ? lgf::kPrecFP64 : lgf::kPrecFP32);
    if (bytes + measure_reserve(line_bytes) > s->budget) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->err = "request needs " + std::to_string(bytes) + " bytes of device memory and " +
                 std::to_string(measure_reserve(line_bytes)) + " of measurement scratch, budget is " +
                 std::to_string(s->budget);
        return LGF_E_ADMISSION;
    }
    auto j = std::make_shared<Job>();
    j->req = *req;
    j->bytes = bytes;
    j->line_bytes = line_bytes;
    if (req->links) j->links.assign(req->links, req->links + (size_t)volume_of(req->shape) * 72);
    j->req.links = nullptr;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop) return LGF_E_STATE;
        j->id = s->next_id++;
        j->submit_s = s->now();
        s->pending.push_back(j);
        *id = j->id;
    }
    s->cv_work.notify_all();
    return LGF_OK;
}

extern "C" int lgf_server_wait(lgf_server* s, uint64_t id, double timeout_s, lgf_result* out) {
    if (!s || !out) return LGF_E_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    auto ready = [&] { return s->finished.count(id) > 0; };
    if (timeout_s < 0) {
        s->cv_done.wait(lk, ready);
    } else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ready)) {
        return LGF_E_STATE;
    }
    JobP j = s->finished[id];
    s->finished.erase(id);
    lk.unlock();
    std::memset(out, 0, sizeof *out);
    out->id = j->id;
    out->status = j->status;
    out->nmeas = (int)j->tm.size();
    out->submit_s = j->submit_s;
    out->start_s = j->start_s;
    out->finish_s = j->finish_s;
    for (int i = 0; i < out->nmeas; ++i) {
        out->t[i] = j->tm[i];
        out->E[i] = j->E[i];
        out->Q[i] = j->Q[i];
    }
    return LGF_OK;
}

extern "C" int lgf_server_set_policy(lgf_server* s, int policy, int64_t min_sites, double age_rate) {
    if (!s || (policy != LGF_SERVE_ALL && policy != LGF_SERVE_SHORTEST_FIRST) || min_sites < 0 || !(age_rate >= 0))
        return LGF_E_ARG;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->policy = policy;
        s->min_sites = min_sites;
        s->age_rate = age_rate;
    }
    s->cv_work.notify_all();
    return LGF_OK;
}

// The server's thread touches its context only while jobs are resident, and
// a job counts as done only after its field and batches are released; so
// with every submitted job done, and admission held off by `tuning`, the
// context is the caller's until the search ends.
extern "C" int lgf_server_autotune(lgf_server* s, const int* shapes, int nshapes, int precision, int nsteps, int* best) {
    if (!s || !shapes || nshapes <= 0 || nsteps <= 0) return LGF_E_ARG;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop || s->tuning || s->next_id - 1 != s->jobs_done) {
            s->err = "autotune needs an idle server: every submitted job finished";
            return LGF_E_STATE;
        }
        s->tuning = true;
    }
    int rc = hipSetDevice(s->device) == hipSuccess ? LGF_OK : LGF_E_HIP;
    if (!rc) rc = lgf_autotune_batch(s->ctx, shapes, nshapes, precision, nsteps, best, nullptr);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->tuning = false;
        if (rc) s->err = lgf_ctx_last_error(s->ctx);
    }
    s->cv_work.notify_all();
    return rc;
}

extern "C" int lgf_server_inject_batch_failures(lgf_server* s, int count) {
    if (!s || count < 0) return LGF_E_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    s->fail_batches = count;
    return LGF_OK;
}

extern "C" int lgf_server_now(lgf_server* s, double* now_s) {
    if (!s || !now_s) return LGF_E_ARG;
    *now_s = s->now();
    return LGF_OK;
}

extern "C" int lgf_server_stats(lgf_server* s, uint64_t* jobs_done, uint64_t* batch_steps, double* link_updates,
                                double* busy_s) {
    if (!s) return LGF_E_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    if (jobs_done) *jobs_done = s->jobs_done;
    if (batch_steps) *batch_steps = s->batch_steps;
    if (link_updates) *link_updates = s->link_updates;
    if (busy_s) *busy_s = s->busy_s;
    return LGF_OK;
}

// copied under the lock into a per-thread buffer: the scheduler may overwrite err
extern "C" const char* lgf_server_last_error(const lgf_server* s) {
    static thread_local std::string copy;
    if (!s) return "null server";
    lgf_server* m = const_cast<lgf_server*>(s);
    std::lock_guard<std::mutex> lk(m->mu);
    copy = m->err;
    return copy.c_str();
}
