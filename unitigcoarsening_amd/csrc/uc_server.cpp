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
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/uc_api.h"
#include "uc_internal.h"

namespace {

using Clock = std::chrono::steady_clock;

// A host buffer of a job's input or text: page-locked when it could be had.
struct HostBuf {
    char *p = nullptr;
    uint64_t cap = 0;
    bool pinned = false;
};

// Input and text buffers recycled from job to job.  Page-locked, so that the
// copies to and from the device are one DMA each, from and into memory that is
// already mapped (a fresh host buffer costs a page fault per 4 KiB on first
// touch: a 108 MB text took 8.3 ms that way, against 2-4 ms into page-locked
// memory, on MI355X; a submission's copy of its input slowed likewise).  Sizes
// are rounded up to an eighth of their power of two, so that the buffers of
// jobs of similar size are interchangeable; at most KEEP bytes stay free.
class HostPool {
public:
    static constexpr uint64_t KEEP = 8ull << 30;

    HostBuf get(uint64_t bytes)
    {
        const uint64_t cap = round_up(bytes);
        {
            std::lock_guard<std::mutex> lk(mu_);
            auto it = free_.lower_bound(cap);
            if (it != free_.end() && it->first <= 2 * cap) {
                HostBuf b{it->second, it->first, true};
                free_bytes_ -= it->first;
                free_.erase(it);
                return b;
            }
        }
        HostBuf b;
        b.p = static_cast<char *>(uc_host_pinned_alloc(cap));
        b.cap = cap;
        b.pinned = b.p != nullptr;
        if (!b.p) b.p = new (std::nothrow) char[cap];
        if (!b.p) b.cap = 0;
        return b;
    }

    void put(HostBuf &b)
    {
        if (b.p && !b.pinned) delete[] b.p;
        if (b.p && b.pinned) {
            std::unique_lock<std::mutex> lk(mu_);
            if (free_bytes_ + b.cap <= KEEP) {
                free_.emplace(b.cap, b.p);
                free_bytes_ += b.cap;
            } else {
                lk.unlock();
                uc_host_pinned_free(b.p);
            }
        }
        b = HostBuf{};
    }

    ~HostPool()
    {
        for (auto &kv : free_) uc_host_pinned_free(kv.second);
    }

private:
    static uint64_t round_up(uint64_t bytes)
    {
        if (bytes <= (64u << 10)) return 64u << 10;
        uint64_t p2 = 1;
        while (p2 <= bytes / 2) p2 <<= 1;  // largest power of two <= bytes
        const uint64_t step = p2 / 8;
        return (bytes + step - 1) / step * step;
    }

    std::mutex mu_;
    std::multimap<uint64_t, char *> free_;
    uint64_t free_bytes_ = 0;
};

struct Job {
    uint64_t id = 0;
    HostBuf in;  // the k-mers, from the server's pool; given back when the job has run
    uint64_t n = 0;
    uint32_t k = 0, l = 0;
    uint64_t estimate = 0;
    int state = UC_JOB_QUEUED;
    int status = UC_OK;
    HostBuf text;  // from the server's pool; given back when the job is released
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
// contexts update at every allocation and free.  The host buffer pool has a
// lock of its own (taken without `mu` held, or after it, never before it), and
// a job's input buffer belongs to the submitting thread until the job is
// queued, then to the worker that runs it.
struct uc_server {
    uc_server_opts opts{};
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Job *> queue;
    std::unordered_map<uint64_t, Job *> jobs;
    std::vector<Worker> workers;
    std::vector<std::thread> threads;
    uc_meter meter;
    HostPool bufs;
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
            int rc = uc_compact(w.ctx, reinterpret_cast<const uint32_t *>(j->in.p), j->n, j->k, j->l, 0);
            bufs.put(j->in);
            HostBuf text;
            uint64_t text_size = 0;
            if (rc == UC_OK) {
                text_size = uc_text_size(w.ctx);
                text = bufs.get(text_size);
                if (!text.p)
                    rc = UC_E_NOMEM;
                else
                    rc = text.pinned ? uc_text_copy_pinned(w.ctx, text.p, text_size)
                                     : uc_text_copy(w.ctx, text.p, text_size);
                if (rc != UC_OK) bufs.put(text);
            }
            const uint64_t held_now = uc_ctx_held_bytes(w.ctx);

            lk.lock();
            j->text = text;
            j->text_size = rc == UC_OK ? text_size : 0;
            j->status = rc;
            j->state = rc == UC_OK ? UC_JOB_DONE : UC_JOB_FAILED;
            j->finished = Clock::now();
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
        // the workers' copies out overlap their copies in when a kernel makes
        // them (the DMA engine took them all one after another)
        uc_ctx_set_kernel_copies(w.ctx, true);
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
    for (auto &kv : s->jobs) {
        s->bufs.put(kv.second->in);
        s->bufs.put(kv.second->text);
        delete kv.second;
    }
    delete s;
    return UC_OK;
}

int uc_submit(uc_server *s, const uint32_t *packed, uint64_t n, uint32_t k, uint32_t l, uint64_t *job)
{
    if (!s || !job || (!packed && n) || k < 2 || k > 64) return UC_E_ARG;
    const uint64_t words = n * ((k + 15) / 16);
    const uint64_t estimate = uc_job_bytes_estimate(n, k);
    const bool over = s->opts.hbm_budget && estimate > s->opts.hbm_budget;
    Job *j = new Job();
    if (!over) {
        // the caller's k-mers are copied before submission returns
        j->in = s->bufs.get(words * 4);
        if (!j->in.p) {
            delete j;
            return UC_E_NOMEM;
        }
        if (words) std::memcpy(j->in.p, packed, words * 4);
    }
    j->n = n;
    j->k = k;
    j->l = l;
    j->estimate = estimate;
    j->submitted = Clock::now();
    std::lock_guard<std::mutex> lk(s->mu);
    if (over) {
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
    if (dst && cap >= j->text_size && j->text_size) std::memcpy(dst, j->text.p, j->text_size);
    return UC_OK;
}

int uc_release(uc_server *s, uint64_t job)
{
    if (!s) return UC_E_ARG;
    Job *j = nullptr;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->jobs.find(job);
        if (it == s->jobs.end()) return UC_E_ARG;
        if (it->second->state != UC_JOB_DONE && it->second->state != UC_JOB_FAILED) return UC_E_STATE;
        j = it->second;
        s->jobs.erase(it);
    }
    s->bufs.put(j->text);  // (outside the server's lock: a release of page-locked memory may wait)
    delete j;
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
