// stub_ctx.cpp -- host-only stand-in for the compaction contexts, so that the
// request loop (uc_server.cpp) can run under ThreadSanitizer / AddressSanitizer
// on a machine without a GPU.  Each stub "job" holds an amount of host memory
// that grows with its size (standing in for device buffers, counted on the
// server's meter exactly like the real contexts count theirs), sleeps a little
// and returns a text that is a function of its k-mers alone.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "uc_api.h"
#include "uc_internal.h"

struct uc_ctx {
    std::vector<char> mem;  // the "device" buffers
    uc_meter *meter = nullptr;
    std::string text;
    bool has_text = false;
};

static void grow(uc_ctx *c, size_t bytes)
{
    if (c->mem.size() >= bytes) return;
    const size_t old = c->mem.size();
    if (c->meter) c->meter->sub(old);
    std::vector<char>().swap(c->mem);
    c->mem.assign(bytes, 1);
    if (c->meter) c->meter->add(bytes);
}

extern "C" {

void uc_default_opts(uc_opts *o) { std::memset(o, 0, sizeof(*o)); }

int uc_ctx_create(const uc_opts *, uc_ctx **out)
{
    *out = new uc_ctx();
    return UC_OK;
}

int uc_ctx_destroy(uc_ctx *c)
{
    if (c && c->meter) c->meter->sub(c->mem.size());
    delete c;
    return UC_OK;
}

uint64_t uc_job_bytes_estimate(uint64_t n, uint32_t) { return 64 * n + 4096; }

int uc_compact(uc_ctx *c, const uint32_t *packed, uint64_t n, uint32_t k, uint32_t l, int)
{
    c->has_text = false;
    if (k < 2 || l < 1) return UC_E_ARG;
    grow(c, 32 * n + 1024);  // at most the estimate
    uint64_t h = 1469598103934665603ull;
    const uint64_t words = n * ((k + 15) / 16);
    for (uint64_t i = 0; i < words; i++) h = (h ^ packed[i]) * 1099511628211ull;
    std::this_thread::sleep_for(std::chrono::microseconds(50 + (h % 400)));
    char buf[64];
    std::snprintf(buf, sizeof(buf), "%llu:%llx\n", (unsigned long long)n, (unsigned long long)h);
    c->text = buf;
    c->has_text = true;
    return UC_OK;
}

uint64_t uc_text_size(const uc_ctx *c) { return c->has_text ? c->text.size() : 0; }

int uc_text_copy(uc_ctx *c, char *dst, uint64_t cap)
{
    if (!c->has_text || cap < c->text.size()) return UC_E_STATE;
    std::memcpy(dst, c->text.data(), c->text.size());
    return UC_OK;
}

int uc_ctx_trim(uc_ctx *c)
{
    if (c->meter) c->meter->sub(c->mem.size());
    std::vector<char>().swap(c->mem);
    return UC_OK;
}

uint64_t uc_ctx_held_bytes(const uc_ctx *c) { return c->mem.size(); }

}  // extern "C"

void uc_ctx_set_meter(uc_ctx *c, uc_meter *m)
{
    if (c->meter) c->meter->sub(c->mem.size());
    c->meter = m;
    if (m) m->add(c->mem.size());
}

// host memory stands in for page-locked memory
void *uc_host_pinned_alloc(uint64_t bytes) { return std::malloc(bytes ? bytes : 1); }

void uc_host_pinned_free(void *p) { std::free(p); }

int uc_text_copy_pinned(uc_ctx *c, char *dst, uint64_t cap) { return uc_text_copy(c, dst, cap); }

void uc_ctx_set_kernel_copies(uc_ctx *, bool) {}
