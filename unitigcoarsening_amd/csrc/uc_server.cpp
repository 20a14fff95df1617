// uc_server.cpp -- request loop over the C ABI: jobs of different sizes
// arrive at any time, are admitted by device-memory budget onto one of
// n_slots worker contexts (each with its own HIP stream and host thread),
// and run concurrently on the device.  No counterpart in the reference,
// which runs one job per process (driver.cpp:422-662).
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/uc_api.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Job {
    uint64_t id = 0;
    std::vector<uint32_t> packed;
    uint64_t n = 0;
    uint32_t k = 0, l = 0;
    uint64_t estimate = 0;
    int state = UC_JOB_QUEUED;
    int status = UC_OK;
    std::vector<char> text;
    Clock::time_point submitted, started, finished;
};

struct Worker {
    uc_ctx *ctx = nullptr;
    bool busy = false;
    uint64_t reserved = 0;  // bytes granted to the running job
};

}  // namespace

struct uc_server {
    uc_server_opts opts{};
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Job *> queue;
    std::unordered_map<uint64_t, Job *> jobs;
    std::vector<Worker> workers;
    std::vector<std::thread> threads;
    uint64_t next_id = 1, peak_held = 0, done = 0;
    bool stopping = false;

    // device bytes the other workers hold or have been granted
    uint64_t held_by_others(size_t self) const
    {
        uint64_t s = 0;
        for (size_t i = 0; i < workers.size(); i++)
            if (i != self) s += std::max(workers[i].reserved, uc_ctx_held_bytes(workers[i].ctx));
        return s;
    }

    // called with mu held by idle worker `self`: can the queue head run here?
    // A worker keeps its buffers from job to job (no reallocation while the
    // budget allows); they can grow by at most the job's estimate.  When the
    // budget is short, this worker's buffers go back first, then those of
    // idle workers.
    bool admit(size_t self, Job *j)
    {
        const uint64_t budget = opts.hbm_budget;
        if (!budget) return true;
        const uint64_t need = j->estimate;
        uc_ctx *mine = workers[self].ctx;
        if (held_by_others(self) + uc_ctx_held_bytes(mine) + need <= budget) return true;
        if (uc_ctx_held_bytes(mine)) uc_ctx_trim(mine);
        if (held_by_others(self) + need <= budget) return true;
        for (size_t i = 0; i < workers.size(); i++)
            if (i != self && !workers[i].busy && uc_ctx_held_bytes(workers[i].ctx)) uc_ctx_trim(workers[i].ctx);
        return held_by_others(self) + need <= budget;
    }

    void run(size_t self)
    {
        Worker &w = workers[self];
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            cv_work.wait(lk, [&] {
                return stopping || (!queue.empty() && !w.busy && admit(self, queue.front()));
            });
            if (queue.empty()) {
                if (stopping) return;
                continue;
            }
            Job *j = queue.front();
            queue.pop_front();
            w.busy = true;
            w.reserved = uc_ctx_held_bytes(w.ctx) + j->estimate;  // what admit() counted
            j->state = UC_JOB_RUNNING;
            j->started = Clock::now();
            lk.unlock();

            int rc = uc_compact(w.ctx, j->packed.data(), j->n, j->k, j->l, 0);
            std::vector<char> text;
            if (rc == UC_OK) {
                text.resize(uc_text_size(w.ctx));
                rc = uc_text_copy(w.ctx, text.data(), text.size());
            }

            lk.lock();
            j->text.swap(text);
            j->status = rc;
            j->state = rc == UC_OK ? UC_JOB_DONE : UC_JOB_FAILED;
            j->finished = Clock::now();
            std::vector<uint32_t>().swap(j->packed);
            uint64_t held = 0;
            for (auto &x : workers) held += uc_ctx_held_bytes(x.ctx);
            peak_held = std::max(peak_held, held);
            done++;
            w.busy = false;
            w.reserved = 0;
            cv_done.notify_all();
            cv_work.notify_all();
        }
    }
};

extern "C" {

int uc_server_create(const uc_server_opts *opts, uc_server **out)
{
    if (!opts || !out || opts->n_slots == 0 || opts->n_slots > 64) return UC_E_ARG;
    *out = nullptr;
    uc_server *s = new uc_server();
    s->opts = *opts;
    s->workers.resize(opts->n_slots);
    for (auto &w : s->workers) {
        int rc = uc_ctx_create(&opts->ctx_opts, &w.ctx);
        if (rc != UC_OK) {
            for (auto &x : s->workers)
                if (x.ctx) uc_ctx_destroy(x.ctx);
            delete s;
            return rc;
        }
    }
    for (size_t i = 0; i < s->workers.size(); i++) s->threads.emplace_back([s, i] { s->run(i); });
    *out = s;
    return UC_OK;
}

int uc_server_destroy(uc_server *s)
{
    if (!s) return UC_OK;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        s->cv_done.wait(lk, [&] {
            if (!s->queue.empty()) return false;
            for (auto &w : s->workers)
                if (w.busy) return false;
            return true;
        });
        s->stopping = true;
    }
    s->cv_work.notify_all();
    for (auto &t : s->threads) t.join();
    for (auto &w : s->workers) uc_ctx_destroy(w.ctx);
    for (auto &kv : s->jobs) delete kv.second;
    delete s;
    return UC_OK;
}

int uc_submit(uc_server *s, const uint32_t *packed, uint64_t n, uint32_t k, uint32_t l, uint64_t *job)
{
    if (!s || !job || (!packed && n) || k < 2 || k > 64) return UC_E_ARG;
    Job *j = new Job();
    const uint64_t words = n * ((k + 15) / 16);
    j->packed.assign(packed, packed + words);
    j->n = n;
    j->k = k;
    j->l = l;
    j->estimate = uc_job_bytes_estimate(n, k);
    j->submitted = Clock::now();
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->opts.hbm_budget && j->estimate > s->opts.hbm_budget) {
        j->state = UC_JOB_FAILED;
        j->status = UC_E_NOMEM;
        j->started = j->finished = j->submitted;
    } else {
        s->queue.push_back(j);
    }
    j->id = s->next_id++;
    s->jobs[j->id] = j;
    *job = j->id;
    s->cv_work.notify_all();
    return UC_OK;
}

int uc_poll(uc_server *s, uint64_t job, int *state)
{
    if (!s || !state) return UC_E_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(job);
    if (it == s->jobs.end()) return UC_E_ARG;
    *state = it->second->state;
    return UC_OK;
}

int uc_wait(uc_server *s, uint64_t job)
{
    if (!s) return UC_E_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->jobs.find(job);
    if (it == s->jobs.end()) return UC_E_ARG;
    Job *j = it->second;
    s->cv_done.wait(lk, [&] { return j->state == UC_JOB_DONE || j->state == UC_JOB_FAILED; });
    return j->status;
}

int uc_result(uc_server *s, uint64_t job, char *dst, uint64_t cap, uint64_t *size, double *queue_ms, double *run_ms)
{
    if (!s) return UC_E_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(job);
    if (it == s->jobs.end()) return UC_E_ARG;
    Job *j = it->second;
    if (j->state != UC_JOB_DONE && j->state != UC_JOB_FAILED) return UC_E_STATE;
    if (size) *size = j->text.size();
    if (queue_ms) *queue_ms = std::chrono::duration<double, std::milli>(j->started - j->submitted).count();
    if (run_ms) *run_ms = std::chrono::duration<double, std::milli>(j->finished - j->started).count();
    if (j->status != UC_OK) return j->status;
    if (dst && cap >= j->text.size()) std::memcpy(dst, j->text.data(), j->text.size());
    return UC_OK;
}

int uc_release(uc_server *s, uint64_t job)
{
    if (!s) return UC_E_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(job);
    if (it == s->jobs.end()) return UC_E_ARG;
    if (it->second->state != UC_JOB_DONE && it->second->state != UC_JOB_FAILED) return UC_E_STATE;
    delete it->second;
    s->jobs.erase(it);
    return UC_OK;
}

int uc_server_stats(uc_server *s, uint64_t *peak_held, uint64_t *jobs_done)
{
    if (!s) return UC_E_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    if (peak_held) *peak_held = s->peak_held;
    if (jobs_done) *jobs_done = s->done;
    return UC_OK;
}

}  // extern "C"
