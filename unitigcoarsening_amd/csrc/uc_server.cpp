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
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/uc_api.h"
#include "uc_internal.h"

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
    std::unique_ptr<char[]> text;  // (not zero-filled first: the copy writes every byte)
    uint64_t text_size = 0;
    Clock::time_point submitted, started, finished;
};

struct Worker {
    uc_ctx *ctx = nullptr;
    bool busy = false;      // running a job or giving its buffers back
    bool trim_req = false;  // asked, while idle, to give its buffers back
    uint64_t held = 0;      // device bytes its context held when it last went idle
    uint64_t reserved = 0;  // while busy: the most its context can hold meanwhile
};

}  // namespace

// Every field below is guarded by `mu` except the contexts themselves: a
// worker's context is used only by that worker's thread (creation and
// destruction aside), and what the others need to know of it -- the bytes it
// held when it went idle, the bound while it runs -- is published under `mu`.
// The true peak of device bytes over all workers comes from the meter the
// contexts update at every allocation and free.
struct uc_server {
    uc_server_opts opts{};
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Job *> queue;
    std::unordered_map<uint64_t, Job *> jobs;
    std::vector<Worker> workers;
    std::vector<std::thread> threads;
    uc_meter meter;
    uint64_t next_id = 1, done = 0;
    bool stopping = false;

    // device bytes the other workers hold, or may grow to while busy
    uint64_t others(size_t self) const
    {
        uint64_t s = 0;
        for (size_t i = 0; i < workers.size(); i++)
            if (i != self) s += workers[i].busy ? workers[i].reserved : workers[i].held;
        return s;
    }

    // With mu held, for idle worker `self` and the queue head: 1 = start it
    // keeping this worker's buffers (they grow by at most the job's
    // estimate), 2 = start it after this worker gives its buffers back, 0 =
    // not yet.  In the last case idle workers holding memory are asked to
    // give it back; admission is tried again when they have.
    int admit(size_t self, const Job *j)
    {
        const uint64_t budget = opts.hbm_budget;
        if (!budget) return 1;
        const uint64_t o = others(self);
        if (o + workers[self].held + j->estimate <= budget) return 1;
        if (o + j->estimate <= budget) return 2;
        bool asked = false;
        for (size_t i = 0; i < workers.size(); i++) {
            Worker &x = workers[i];
            if (i != self && !x.busy && x.held && !x.trim_req) {
                x.trim_req = true;
                asked = true;
            }
        }
        if (asked) cv_work.notify_all();
        return 0;
    }

    void run(size_t self)
    {
        Worker &w = workers[self];
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            int how = 0;
            cv_work.wait(lk, [&] {
                if (w.trim_req || (stopping && queue.empty())) return true;
                if (queue.empty()) return false;
                how = admit(self, queue.front());
                return how != 0;
            });
            if (w.trim_req) {
                // give the buffers back outside the lock (hipFree waits for the device)
                w.trim_req = false;
                if (w.held) {
                    w.busy = true;
                    w.reserved = w.held;
                    lk.unlock();
                    uc_ctx_trim(w.ctx);
                    lk.lock();
                    w.held = 0;
                    w.busy = false;
                    w.reserved = 0;
                    cv_work.notify_all();
                }
                continue;
            }
            if (queue.empty()) {
                if (stopping) return;
                continue;
            }
            Job *j = queue.front();
            queue.pop_front();
            w.busy = true;
            w.reserved = (how == 1 ? w.held : 0) + j->estimate;  // what admit() counted
            j->state = UC_JOB_RUNNING;
            j->started = Clock::now();
            lk.unlock();

            if (how == 2) uc_ctx_trim(w.ctx);
            int rc = uc_compact(w.ctx, j->packed.data(), j->n, j->k, j->l, 0);
            std::unique_ptr<char[]> text;
            uint64_t text_size = 0;
            if (rc == UC_OK) {
                text_size = uc_text_size(w.ctx);
                text.reset(new (std::nothrow) char[text_size ? text_size : 1]);
                rc = text ? uc_text_copy(w.ctx, text.get(), text_size) : UC_E_NOMEM;
            }
            const uint64_t held_now = uc_ctx_held_bytes(w.ctx);

            lk.lock();
            j->text.swap(text);
            j->text_size = rc == UC_OK ? text_size : 0;
            j->status = rc;
            j->state = rc == UC_OK ? UC_JOB_DONE : UC_JOB_FAILED;
            j->finished = Clock::now();
            std::vector<uint32_t>().swap(j->packed);
            done++;
            w.busy = false;
            w.reserved = 0;
            w.held = held_now;
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
        uc_ctx_set_meter(w.ctx, &s->meter);
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
    if (size) *size = j->text_size;
    if (queue_ms) *queue_ms = std::chrono::duration<double, std::milli>(j->started - j->submitted).count();
    if (run_ms) *run_ms = std::chrono::duration<double, std::milli>(j->finished - j->started).count();
    if (j->status != UC_OK) return j->status;
    if (dst && cap >= j->text_size && j->text_size) std::memcpy(dst, j->text.get(), j->text_size);
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
    if (peak_held) *peak_held = s->meter.peak.load();
    if (jobs_done) *jobs_done = s->done;
    return UC_OK;
}

}  // extern "C"
