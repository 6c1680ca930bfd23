// Serving loop: requests queue as they arrive; a scheduler thread admits them
// (FIFO with a short look-ahead, so a small job can pass a large one that
// does not fit yet) whenever a slot is free and the job's device footprint
// fits the memory budget, starts each admitted job on its own stream, and
// records completion. With cobatch_tail, a job's stream runs only its
// sampling passes; jobs whose passes have finished queue for a tail thread,
// which runs the denoise tails of everything queued as one batch, one launch
// per stage over all of them (rpd_render_finish_batch). While a batch runs,
// the jobs finishing their passes queue for the next one, and the scheduler
// goes on admitting. Built on the public C ABI, plus the engine's
// thread-local error setter so a waiter sees the message of a failure that
// happened on the scheduler thread.
#include "../../include/rpd_engine.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace rpd_internal {
int set_error(int code, const char* msg);  // engine.hip: the calling thread's rpd_last_error
}

namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

struct Request {
    uint64_t id = 0;
    rpd_scene* scene = nullptr;
    rpd_render_config cfg{};
    rpd_camera cam{};
    uint32_t seed = 0;
    uint64_t bytes = 0;
    Clock::time_point t_submit, t_admit, t_done;
    rpd_job* job = nullptr;
    int status = 0;
    std::string error;  // rpd_last_error of the failure, taken on the scheduler thread
    bool waited = false;  // a caller is blocked in rpd_server_wait on it
    bool tail = false;    // started with its passes only: the denoise tail is the scheduler's
    enum State { QUEUED, RUNNING, DONE, CLAIMED } state = QUEUED;
};

constexpr int kLookAhead = 8;

double percentile(std::vector<double> v, double q) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const double pos = q * double(v.size() - 1);
    const size_t lo = size_t(pos), hi = std::min(lo + 1, v.size() - 1);
    return v[lo] + (v[hi] - v[lo]) * (pos - double(lo));
}

}  // namespace

struct rpd_server {
    rpd_context* ctx = nullptr;
    rpd_server_config cfg{};
    uint64_t budget = 0;
    std::mutex mu;
    std::condition_variable cv_sched, cv_done;
    std::map<uint64_t, std::unique_ptr<Request>> all;
    std::deque<Request*> queue;
    std::vector<Request*> running;
    std::deque<Request*> tail_q;   // passes finished, tail not yet started
    int tail_busy = 0;             // jobs in the tail batch being run
    std::condition_variable cv_tail;
    std::thread tail_thr;
    bool stop_tail = false;
    uint64_t next_id = 1, peak = 0;
    uint64_t over_budget = 0;  // requests admitted beyond the budget (see the scheduler)
    uint64_t submitted = 0, completed = 0, failed = 0, tail_batches = 0, tail_jobs = 0;
    int max_seen = 0;
    std::vector<double> lat, qwait;
    bool stop = false;
    std::thread sched;

    uint64_t held() const {
        uint64_t u = 0;
        rpd_ctx_memory(ctx, &u, nullptr, nullptr, nullptr);  // an atomic load; no device call
        return u;
    }

    // Creates and launches the request's job; called with the lock held,
    // which it drops around the engine call.
    void start(Request* r, std::unique_lock<std::mutex>& lk) {
        r->t_admit = Clock::now();
        qwait.push_back(ms_between(r->t_submit, r->t_admit));
        r->tail = cfg.cobatch_tail && r->cfg.use_statmc;  // a job without StatMC has no tail
        lk.unlock();
        const int st = r->tail ? rpd_job_submit_passes(ctx, r->scene, &r->cfg, &r->cam, r->seed, &r->job)
                               : rpd_job_submit(ctx, r->scene, &r->cfg, &r->cam, r->seed, &r->job);
        const std::string err = st ? rpd_last_error() : "";
        lk.lock();
        if (st) {
            r->error = err;
            r->status = st;
            r->job = nullptr;
            r->t_done = Clock::now();
            r->state = Request::DONE;
            ++failed;
            cv_done.notify_all();
        } else {
            r->state = Request::RUNNING;
            running.push_back(r);
            max_seen = std::max(max_seen, in_flight());
            peak = std::max(peak, held());
        }
    }

    // admitted and not done: sampling, or waiting for or in a tail batch
    int in_flight() const { return int(running.size() + tail_q.size()) + tail_busy; }

    void finish(Request* r, int status, const std::string& err) {
        r->status = status;
        if (status) r->error = err;
        r->t_done = Clock::now();
        r->state = Request::DONE;
        lat.push_back(ms_between(r->t_submit, r->t_done));
        if (status) ++failed; else ++completed;
        cv_done.notify_all();
    }

    void loop() {
        std::unique_lock<std::mutex> lk(mu);
        while (true) {
            // completions; jobs whose passes finished queue for the tail thread
            for (size_t k = 0; k < running.size();) {
                Request* r = running[k];
                int done = 0;
                rpd_job_poll(r->job, &done);
                if (!done) { ++k; continue; }
                const int st = rpd_job_wait(r->job);  // joins the job's worker; already finished
                running.erase(running.begin() + long(k));
                if (!st && r->tail) {
                    tail_q.push_back(r);
                    cv_tail.notify_one();
                } else {
                    finish(r, st, st ? rpd_last_error() : "");
                }
            }
            // admission, against the bytes the context's jobs hold now: a finished
            // job keeps its device memory until whoever claimed it releases it,
            // so finished but unreleased jobs count as well as running ones.
            int looked = 0;
            for (auto it = queue.begin(); it != queue.end() && looked < kLookAhead &&
                                          in_flight() < cfg.max_in_flight;) {
                ++looked;
                if (held() + (*it)->bytes > budget) { ++it; continue; }
                Request* r = *it;
                queue.erase(it);
                start(r, lk);
                it = queue.begin();  // the queue changed while unlocked
                looked = 0;
            }
            // Nothing runs and nothing queued fits: the memory is held by finished
            // jobs not yet released. A caller blocked on a queued request (say,
            // one that claims in submission order while later requests finished
            // first) would wait for ever, so that request is admitted over the
            // budget; so is the oldest one once the server is stopping.
            if (in_flight() == 0 && !queue.empty()) {
                auto it = std::find_if(queue.begin(), queue.end(), [](const Request* r) { return r->waited; });
                if (it == queue.end() && stop) it = queue.begin();
                if (it != queue.end()) {
                    Request* r = *it;
                    queue.erase(it);
                    if (held() + r->bytes > budget) ++over_budget;
                    start(r, lk);
                }
            }
            if (stop && queue.empty() && in_flight() == 0) break;
            cv_sched.wait_for(lk, std::chrono::microseconds(running.empty() ? 2000 : 100));
        }
    }

    // Runs everything in the tail queue as one batch, until stopped.
    void tail_loop() {
        std::unique_lock<std::mutex> lk(mu);
        while (true) {
            cv_tail.wait(lk, [this]() { return !tail_q.empty() || stop_tail; });
            if (tail_q.empty()) break;
            std::vector<Request*> batch(tail_q.begin(), tail_q.end());
            tail_q.clear();
            tail_busy = int(batch.size());
            std::vector<rpd_job*> jobs;
            for (Request* r : batch) jobs.push_back(r->job);
            lk.unlock();
            const int st = rpd_render_finish_batch(ctx, jobs.data(), int(jobs.size()), nullptr);
            const std::string err = st ? rpd_last_error() : "";
            lk.lock();
            tail_busy = 0;
            ++tail_batches;
            tail_jobs += batch.size();
            for (Request* r : batch) finish(r, st, err);
            cv_sched.notify_one();   // slots are free again
        }
    }
};

extern "C" {

int rpd_server_create(rpd_context* ctx, const rpd_server_config* cfg, rpd_server** out) {
    if (!ctx || !cfg || !out) return RPD_ERR_INVALID_ARG;
    if (cfg->max_in_flight < 1) return RPD_ERR_INVALID_ARG;
    auto s = std::make_unique<rpd_server>();
    s->ctx = ctx;
    s->cfg = *cfg;
    s->budget = cfg->memory_budget;
    uint64_t fr = 0, tot = 0;
    int st = rpd_ctx_memory(ctx, nullptr, nullptr, &fr, &tot);
    if (st) return st;
    if (!s->budget) s->budget = fr;
    // Released jobs' slabs are kept for the next admitted jobs in what the
    // budget leaves of the device, less a reserve for scenes and the context's
    // own buffers, and never less than the context's default; a job that finds
    // no room empties the cache and allocates again.
    const uint64_t reserve = uint64_t(16) << 30, floor_cap = uint64_t(32) << 30;
    const uint64_t spare = tot > s->budget + reserve ? tot - s->budget - reserve : 0;
    st = rpd_ctx_set_buffer_cache(ctx, std::max(spare, floor_cap), nullptr);
    if (!st) st = rpd_ctx_set_held_limit(ctx, s->budget);
    if (st) return st;
    st = rpd_ctx_set_memory_budget(ctx, 0);  // admission is the server's; the context does not refuse
    if (!st) st = rpd_ctx_set_graph_passes(ctx, cfg->graph_passes);
    if (st) return st;
    rpd_server* raw = s.get();
    if (cfg->cobatch_tail) s->tail_thr = std::thread([raw]() { raw->tail_loop(); });
    s->sched = std::thread([raw]() { raw->loop(); });
    *out = s.release();
    return RPD_OK;
}

int rpd_server_submit(rpd_server* s, rpd_scene* scene, const rpd_render_config* cfg, const rpd_camera* cam,
                      uint32_t seed, uint64_t* request_id) {
    if (!s || !scene || !cfg || !cam || !request_id) return RPD_ERR_INVALID_ARG;
    uint64_t bytes = 0;
    const int st = rpd_job_bytes(cfg, &bytes);  // validates the configuration too
    if (st) return st;
    if (bytes > s->budget) return RPD_ERR_OUT_OF_MEMORY;  // could never be admitted
    auto r = std::make_unique<Request>();
    r->scene = scene; r->cfg = *cfg; r->cam = *cam; r->seed = seed; r->bytes = bytes;
    r->t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop) return RPD_ERR_STATE;
        r->id = s->next_id++;
        *request_id = r->id;
        s->queue.push_back(r.get());
        s->all[r->id] = std::move(r);
        ++s->submitted;
    }
    s->cv_sched.notify_one();
    return RPD_OK;
}

int rpd_server_wait(rpd_server* s, uint64_t request_id, rpd_job** job, double* latency_ms) {
    if (!s) return RPD_ERR_INVALID_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->all.find(request_id);
    if (it == s->all.end()) return RPD_ERR_INVALID_ARG;
    Request* r = it->second.get();
    if (r->state == Request::CLAIMED) return RPD_ERR_STATE;
    if (r->state == Request::QUEUED) {
        r->waited = true;
        s->cv_sched.notify_one();
    }
    s->cv_done.wait(lk, [r]() { return r->state == Request::DONE; });
    if (latency_ms) *latency_ms = ms_between(r->t_submit, r->t_done);
    const int status = r->status;
    if (job) {
        *job = r->job;
        r->job = nullptr;
    } else if (r->job) {
        rpd_job_release(r->job);
        r->job = nullptr;
    }
    r->state = Request::CLAIMED;
    if (status) return rpd_internal::set_error(status, r->error.c_str());
    return status;
}

int rpd_server_stats_get(rpd_server* s, rpd_server_stats* out) {
    if (!s || !out) return RPD_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    out->submitted = s->submitted;
    out->completed = s->completed;
    out->failed = s->failed;
    out->max_in_flight_seen = s->max_seen;
    out->peak_bytes = s->peak;
    out->over_budget = s->over_budget;
    out->latency_p50_ms = percentile(s->lat, 0.50);
    out->latency_p99_ms = percentile(s->lat, 0.99);
    double sum = 0.0, mx = 0.0;
    for (double v : s->lat) { sum += v; mx = std::max(mx, v); }
    out->latency_mean_ms = s->lat.empty() ? 0.0 : sum / double(s->lat.size());
    out->latency_max_ms = mx;
    out->queue_p50_ms = percentile(s->qwait, 0.50);
    out->queue_p99_ms = percentile(s->qwait, 0.99);
    out->tail_batches = s->tail_batches;
    out->tail_jobs = s->tail_jobs;
    return RPD_OK;
}

int rpd_server_destroy(rpd_server* s) {
    if (!s) return RPD_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_sched.notify_all();
    if (s->sched.joinable()) s->sched.join();   // returns once nothing is queued or in flight
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop_tail = true;
    }
    s->cv_tail.notify_all();
    if (s->tail_thr.joinable()) s->tail_thr.join();
    for (auto& kv : s->all)
        if (kv.second->job) rpd_job_release(kv.second->job);
    delete s;
    return RPD_OK;
}

}  // extern "C"
