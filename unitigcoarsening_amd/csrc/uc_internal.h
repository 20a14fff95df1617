// uc_internal.h -- library-internal hooks between the compaction contexts
// (uc_core.hip) and the request loop (uc_server.cpp).  Not part of the C ABI.
#pragma once
#include <stdint.h>

#include <atomic>

struct uc_ctx;

// Device bytes held by a group of contexts (a server's workers): the sum,
// updated at every device allocation and free of every member, and its
// running maximum -- the true peak of the group, not a sample.
struct uc_meter {
    std::atomic<uint64_t> held{0};
    std::atomic<uint64_t> peak{0};

    void add(uint64_t b)
    {
        const uint64_t v = held.fetch_add(b) + b;
        uint64_t p = peak.load();
        while (v > p && !peak.compare_exchange_weak(p, v)) {
        }
    }
    void sub(uint64_t b) { held.fetch_sub(b); }
};

// Attach (or detach, with nullptr) a meter to a context; the bytes the
// context already holds move onto it.  Call from the context's own thread.
void uc_ctx_set_meter(uc_ctx *c, uc_meter *m);

// Page-locked host memory (nullptr when it cannot be had), and its release.
void *uc_host_pinned_alloc(uint64_t bytes);
void uc_host_pinned_free(void *p);

// uc_text_copy into page-locked memory from uc_host_pinned_alloc: one
// direct device-to-host copy, no staging.
int uc_text_copy_pinned(uc_ctx *c, char *dst, uint64_t cap);

// Copies from the context's device memory into page-locked host memory (its
// small counter reads, uc_text_copy_pinned) by a copy kernel instead of the
// DMA engine, so that they overlap the copies of other contexts sharing the
// device.  Call from the context's own thread.
void uc_ctx_set_kernel_copies(uc_ctx *c, bool on);
