// uc_text.cpp -- host side of the input boundary: the reference's k-mer text
// format (driver.cpp:251-354) parsed into the reference packing
// (driver.cpp:227-249), split over threads at line boundaries.
#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/uc_api.h"

namespace {

struct CharMap {
    uint8_t v[256];
    CharMap()
    {
        // driver.cpp:228-232: a zero-initialised table, so every byte other
        // than A, C, G, T reads as A
        std::memset(v, 0, sizeof(v));
        v[(unsigned char)'C'] = 1;
        v[(unsigned char)'G'] = 2;
        v[(unsigned char)'T'] = 3;
    }
};
const CharMap kMap;

unsigned n_threads(uint64_t len)
{
    unsigned hw = std::thread::hardware_concurrency();
    if (const char *e = std::getenv("OMP_NUM_THREADS")) {
        int v = std::atoi(e);
        if (v > 0) hw = (unsigned)v;
    }
    if (hw == 0) hw = 1;
    const uint64_t per = 1ull << 22;  // at least 4 MiB of text per thread
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(hw, (len + per - 1) / per));
}

}  // namespace

extern "C" int uc_parse_kmer_text(const char *buf, uint64_t len, uint32_t k, uint32_t **packed, uint64_t *n_out)
{
    if (!packed || !n_out || (!buf && len) || k < 2 || k > 64) return UC_E_ARG;
    *packed = nullptr;
    *n_out = 0;
    // line 1: the k-mer count (sscanf "%u" / "%lu", driver.cpp:296-304).
    // The reference compares it with the k-mers read only after whole read
    // chunks (driver.cpp:343) and reads on past it otherwise, so every line
    // is taken whatever the count says.
    uint64_t p = 0;
    bool digit = false;
    while (p < len && (buf[p] == ' ' || buf[p] == '\t')) p++;
    while (p < len && buf[p] >= '0' && buf[p] <= '9') {
        digit = true;
        p++;
    }
    if (!digit) return UC_E_FORMAT;
    while (p < len && buf[p] != '\n') p++;
    if (p < len) p++;
    const uint64_t body = p;
    // newlines and carriage returns at the end of the text end it
    while (len > body && (buf[len - 1] == '\n' || buf[len - 1] == '\r')) len--;
    // chunk boundaries at line starts
    const unsigned T = n_threads(len - body);
    std::vector<uint64_t> start(T + 1, len);
    start[0] = body;
    for (unsigned t = 1; t < T; t++) {
        uint64_t q = body + (len - body) * t / T;
        q = std::max(q, start[t - 1]);
        while (q < len && buf[q - 1] != '\n') q++;
        start[t] = q;
    }
    // lines per chunk (a line is a maximal run ending at '\n' or at the end)
    std::vector<uint64_t> lines(T + 1, 0);
    {
        std::vector<std::thread> th;
        for (unsigned t = 0; t < T; t++)
            th.emplace_back([&, t] {
                uint64_t cnt = 0;
                for (uint64_t q = start[t]; q < start[t + 1];) {
                    const void *nl = std::memchr(buf + q, '\n', start[t + 1] - q);
                    cnt++;
                    if (!nl) break;
                    q = (uint64_t)((const char *)nl - buf) + 1;
                }
                lines[t + 1] = cnt;
            });
        for (auto &x : th) x.join();
    }
    for (unsigned t = 0; t < T; t++) lines[t + 1] += lines[t];
    const uint64_t n = lines[T];
    const uint32_t rw = (k + 15) / 16;
    uint32_t *out = (uint32_t *)std::calloc(n ? n * rw : 1, sizeof(uint32_t));
    if (!out) return UC_E_NOMEM;
    std::atomic<int> bad{0};
    {
        std::vector<std::thread> th;
        for (unsigned t = 0; t < T; t++)
            th.emplace_back([&, t] {
                uint64_t idx = lines[t];
                for (uint64_t q = start[t]; q < start[t + 1] && idx < n; idx++) {
                    const char *nl = (const char *)std::memchr(buf + q, '\n', start[t + 1] - q);
                    const uint64_t e = nl ? (uint64_t)(nl - buf) : start[t + 1];
                    uint64_t ll = e - q;
                    if (ll > 0 && buf[q + ll - 1] == '\r') ll--;
                    if (ll < k) {
                        bad.store(1);
                        return;
                    }
                    uint32_t *dst = out + idx * rw;
                    for (uint32_t x = 0; x < k; x++) dst[x / 16] |= (uint32_t)kMap.v[(unsigned char)buf[q + x]] << (2 * (x & 15));
                    q = e + 1;
                }
            });
        for (auto &x : th) x.join();
    }
    if (bad.load()) {
        std::free(out);
        return UC_E_FORMAT;
    }
    *packed = out;
    *n_out = n;
    return UC_OK;
}

extern "C" int uc_read_kmer_file(const char *path, uint32_t k, uint32_t **packed, uint64_t *n)
{
    if (!path) return UC_E_ARG;
    FILE *f = std::fopen(path, "rb");
    if (!f) return UC_E_IO;
    if (std::fseek(f, 0, SEEK_END) != 0) {
        std::fclose(f);
        return UC_E_IO;
    }
    long sz = std::ftell(f);
    if (sz < 0) {
        std::fclose(f);
        return UC_E_IO;
    }
    std::rewind(f);
    char *buf = (char *)std::malloc(sz ? (size_t)sz : 1);
    if (!buf) {
        std::fclose(f);
        return UC_E_NOMEM;
    }
    size_t got = std::fread(buf, 1, (size_t)sz, f);
    std::fclose(f);
    if (got != (size_t)sz) {
        std::free(buf);
        return UC_E_IO;
    }
    int rc = uc_parse_kmer_text(buf, (uint64_t)sz, k, packed, n);
    std::free(buf);
    return rc;
}

extern "C" void uc_free(void *p) { std::free(p); }

// ---- packed binary k-mer files: a 32-byte header (magic, k, words per
// k-mer, count, zero) and the reference packing, little-endian

extern "C" int uc_write_packed_file(const char *path, const uint32_t *packed, uint64_t n, uint32_t k)
{
    if (!path || (!packed && n) || k < 2 || k > 64) return UC_E_ARG;
    FILE *f = std::fopen(path, "wb");
    if (!f) return UC_E_IO;
    uc_packed_header h;
    std::memcpy(h.magic, UC_PACKED_MAGIC, 8);
    h.k = k;
    h.words = (k + 15) / 16;
    h.n = n;
    h.reserved = 0;
    bool ok = std::fwrite(&h, sizeof(h), 1, f) == 1;
    const uint64_t words = n * h.words;
    for (uint64_t off = 0; ok && off < words;) {
        const uint64_t step = std::min<uint64_t>(words - off, 1ull << 24);
        ok = std::fwrite(packed + off, 4, step, f) == step;
        off += step;
    }
    ok = (std::fclose(f) == 0) && ok;
    return ok ? UC_OK : UC_E_IO;
}

extern "C" int uc_read_packed_header(const char *path, uc_packed_header *h)
{
    if (!path || !h) return UC_E_ARG;
    FILE *f = std::fopen(path, "rb");
    if (!f) return UC_E_IO;
    const bool got = std::fread(h, sizeof(*h), 1, f) == 1;
    std::fclose(f);
    if (!got || std::memcmp(h->magic, UC_PACKED_MAGIC, 8) != 0) return UC_E_FORMAT;
    if (h->k < 2 || h->k > 64 || h->words != (h->k + 15) / 16) return UC_E_FORMAT;
    return UC_OK;
}
