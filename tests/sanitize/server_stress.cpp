// server_stress.cpp -- many client threads against one request loop
// (uc_server.cpp over the stub contexts): every job's text must be the one
// its own k-mers give, and the workers' meter peak must stay within the
// budget.  Run under -fsanitize=thread and -fsanitize=address,undefined.
#include <atomic>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "uc_api.h"

static std::string expected(const std::vector<uint32_t> &p, uint64_t n)
{
    uint64_t h = 1469598103934665603ull;
    for (uint32_t w : p) h = (h ^ w) * 1099511628211ull;
    char buf[64];
    std::snprintf(buf, sizeof(buf), "%llu:%llx\n", (unsigned long long)n, (unsigned long long)h);
    return buf;
}

int main(int argc, char **argv)
{
    const int clients = argc > 1 ? std::atoi(argv[1]) : 6;
    const int per_client = argc > 2 ? std::atoi(argv[2]) : 60;
    const uint64_t budget = 3 * uc_job_bytes_estimate(20000, 31);
    uc_server_opts o;
    std::memset(&o, 0, sizeof(o));
    o.n_slots = 4;
    o.hbm_budget = budget;
    uc_server *s = nullptr;
    if (uc_server_create(&o, &s) != UC_OK) return 2;
    std::atomic<int> bad{0};
    std::vector<std::thread> th;
    for (int t = 0; t < clients; t++)
        th.emplace_back([&, t] {
            std::mt19937_64 rng(1000 + t);
            for (int i = 0; i < per_client; i++) {
                const uint64_t n = 1 + rng() % 20000;
                std::vector<uint32_t> p(n * 2);
                for (auto &w : p) w = (uint32_t)rng();
                uint64_t job = 0;
                if (uc_submit(s, p.data(), n, 31, 11, &job) != UC_OK) {
                    bad++;
                    continue;
                }
                if (rng() % 3 == 0) {
                    int st = -1;
                    uc_poll(s, job, &st);
                }
                if (uc_wait(s, job) != UC_OK) bad++;
                uint64_t size = 0;
                double q = 0, r = 0;
                std::string text(64, '\0');
                if (uc_result(s, job, &text[0], text.size(), &size, &q, &r) != UC_OK) bad++;
                text.resize(size);
                if (text != expected(p, n)) bad++;
                if (uc_release(s, job) != UC_OK) bad++;
            }
        });
    // a job above the whole budget is refused at once
    {
        std::vector<uint32_t> p(2 * 100000, 7);
        uint64_t job = 0;
        uc_submit(s, p.data(), 100000, 31, 11, &job);
        if (uc_wait(s, job) != UC_E_NOMEM) bad++;
    }
    for (auto &x : th) x.join();
    uint64_t peak = 0, done = 0;
    uc_server_stats(s, &peak, &done);
    uc_server_destroy(s);
    std::printf("jobs %llu peak %llu budget %llu bad %d\n", (unsigned long long)done, (unsigned long long)peak,
                (unsigned long long)budget, bad.load());
    if (done != (uint64_t)clients * per_client) return 1;
    if (peak > budget || bad) return 1;
    return 0;
}
