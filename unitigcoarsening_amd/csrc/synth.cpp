// synth.cpp -- seeded synthetic k-mer sets for tests and benchmarks.
//
// Genome model: uniform ACGT; optional repeats made of copies of random
// segments (length U[seg_min, seg_max]) from random sources to random
// targets with a per-base substitution rate, until `repeat_frac` of the
// genome has been overwritten; optional error k-mers: a fraction of the
// true k-mers gets one extra copy with one uniform substitution.  The k-mer
// set is every distinct canonical k-mer (canonical = lexicographic minimum
// of the k-mer and its reverse complement, A<C<G<T), returned in
// lexicographic order in the reference packing: ceil(k/16) uint32 words per
// k-mer, character i in bits 2*(i mod 16) of word i/16.
//
// Random numbers: splitmix64 seeds xoshiro256**.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/uc_synth.h"

namespace {

typedef unsigned __int128 u128;

// UC_SYNTH_PROGRESS=1: one line on stderr per stage of a large generation
// (the sorts of a few billion k-mers take minutes)
void progress(const char *stage, uint64_t a = 0)
{
    static const bool on = [] {
        const char *e = std::getenv("UC_SYNTH_PROGRESS");
        return e && e[0] == '1';
    }();
    if (!on) return;
    std::fprintf(stderr, "synth: %s %llu (%.1f s)\n", stage, (unsigned long long)a, omp_get_wtime());
    std::fflush(stderr);
}

struct Rng {
    uint64_t s[4];
    static uint64_t splitmix(uint64_t &x)
    {
        uint64_t z = (x += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    explicit Rng(uint64_t seed)
    {
        for (auto &w : s) w = splitmix(seed);
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next()
    {
        uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    uint64_t below(uint64_t n) { return (uint64_t)(((u128)next() * n) >> 64); }
    double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
};

// lexicographic code of a k-mer: character 0 most significant
template <class K>
K lex_mask(uint32_t k)
{
    return (2 * k >= sizeof(K) * 8) ? ~K(0) : ((K(1) << (2 * k)) - 1);
}

template <class K>
void to_packed(K lex, uint32_t k, uint32_t *dst)
{
    uint32_t w = (k + 15) / 16;
    std::memset(dst, 0, w * sizeof(uint32_t));
    for (uint32_t i = 0; i < k; i++) {
        uint32_t c = (uint32_t)(lex >> (2 * (k - 1 - i))) & 3u;
        dst[i / 16] |= c << (2 * (i & 15));
    }
}

template <class K>
void parallel_sort_unique(std::vector<K> &v)
{
    // MSD split on the top byte, then std::sort per part in parallel
    const int parts = 256;
    size_t n = v.size();
    if (n < (1u << 16)) {
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
        return;
    }
    K maxv = 0;
    for (size_t i = 0; i < n; i++) maxv = std::max(maxv, v[i]);
    int shift = 0;
    while ((maxv >> shift) >= (K)parts) shift++;
    std::vector<size_t> cnt(parts + 1, 0);
    for (size_t i = 0; i < n; i++) cnt[(size_t)(v[i] >> shift) + 1]++;
    for (int p = 0; p < parts; p++) cnt[p + 1] += cnt[p];
    std::vector<K> tmp(n);
    std::vector<size_t> pos(cnt.begin(), cnt.end() - 1);
    for (size_t i = 0; i < n; i++) tmp[pos[(size_t)(v[i] >> shift)]++] = v[i];
    progress("sort: parts scattered", n);
    int done = 0;
#pragma omp parallel for schedule(dynamic, 1)
    for (int p = 0; p < parts; p++) {
        std::sort(tmp.begin() + cnt[p], tmp.begin() + cnt[p + 1]);
        int d;
#pragma omp atomic capture
        d = ++done;
        if (d % 32 == 0) progress("sort: parts sorted", (uint64_t)d);
    }
    tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
    v.swap(tmp);
    progress("sort: distinct", v.size());
}

template <class K>
int run(const uc_synth_params *p, uint32_t **out, uint64_t *n_out)
{
    const uint32_t k = p->k;
    const uint64_t G = p->genome_len;
    if (G < k) {
        *out = (uint32_t *)std::malloc(sizeof(uint32_t));
        *n_out = 0;
        return 0;
    }
    Rng rng(p->seed);
    std::vector<uint8_t> g(G);
    for (uint64_t i = 0; i < G; i++) g[i] = (uint8_t)(rng.next() >> 62);
    if (p->repeat_frac > 0 && p->seg_max >= p->seg_min && p->seg_min > 0 && p->seg_max < G) {
        uint64_t covered = 0, target = (uint64_t)(p->repeat_frac * (double)G);
        const bool families = p->copies_max > 0;
        while (covered < target) {
            uint64_t len = p->seg_min + rng.below(p->seg_max - p->seg_min + 1);
            uint64_t src = rng.below(G - len + 1);
            std::vector<uint8_t> seg(g.begin() + src, g.begin() + src + len);
            uint64_t copies = 1;
            double sub = p->repeat_sub;
            if (families) {
                const double lo = std::log((double)std::max(1u, p->copies_min));
                const double hi = std::log((double)std::max(p->copies_min, p->copies_max));
                copies = (uint64_t)std::llround(std::exp(lo + (hi - lo) * rng.uniform()));
                sub = p->div_min + (p->div_max - p->div_min) * rng.uniform();
            }
            for (uint64_t r = 0; r < copies && covered < target; r++) {
                const uint64_t dst = rng.below(G - len + 1);
                for (uint64_t i = 0; i < len; i++) {
                    uint8_t c = seg[i];
                    if (sub > 0 && rng.uniform() < sub) c = (uint8_t)((c + 1 + rng.below(3)) & 3);
                    g[dst + i] = c;
                }
                covered += len;
            }
        }
    }
    progress("genome with repeats", G);
    const K mask = lex_mask<K>(k);
    const uint64_t nk = G - k + 1;
    std::vector<K> kmers(nk);
    const int nt = omp_get_max_threads();
    const uint64_t chunk = (nk + nt - 1) / nt;
#pragma omp parallel num_threads(nt)
    {
        uint64_t t = (uint64_t)omp_get_thread_num();
        uint64_t b = t * chunk, e = std::min(nk, b + chunk);
        if (b < e) {
            K fwd = 0, rc = 0;
            for (uint64_t i = b; i < e + k - 1; i++) {
                uint32_t c = g[i];
                fwd = ((fwd << 2) | (K)c) & mask;
                rc = (rc >> 2) | ((K)(3 - c) << (2 * (k - 1)));
                if (i + 1 >= b + k) kmers[i + 1 - k] = fwd < rc ? fwd : rc;
            }
        }
    }
    parallel_sort_unique(kmers);
    if (p->error_frac > 0) {
        uint64_t n0 = kmers.size();
        uint64_t n_err = (uint64_t)(p->error_frac * (double)n0);
        Rng er(p->seed ^ 0xA5A5A5A5DEADBEEFull);
        kmers.reserve(n0 + n_err);
        for (uint64_t e = 0; e < n_err; e++) {
            K x = kmers[er.below(n0)];
            uint32_t pos = (uint32_t)er.below(k);
            uint32_t sh = 2 * (k - 1 - pos);
            uint32_t c = (uint32_t)(x >> sh) & 3u;
            uint32_t nc = (c + 1 + (uint32_t)er.below(3)) & 3u;
            x = (x & ~((K)3 << sh)) | ((K)nc << sh);
            K rc = 0;
            for (uint32_t i = 0; i < k; i++) rc = (rc << 2) | (K)(3 - ((uint32_t)(x >> (2 * i)) & 3u));
            kmers.push_back(x < rc ? x : rc);
        }
        parallel_sort_unique(kmers);
    }
    progress("k-mers", kmers.size());
    const uint32_t w = (k + 15) / 16;
    const uint64_t n = kmers.size();
    uint32_t *buf = (uint32_t *)std::malloc((n ? n : 1) * w * sizeof(uint32_t));
    if (!buf) return -2;
#pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < n; i++) to_packed(kmers[i], k, buf + i * w);
    *out = buf;
    *n_out = n;
    return 0;
}

}  // namespace

extern "C" int uc_synth_kmers(const uc_synth_params *p, uint32_t **packed, uint64_t *n)
{
    if (!p || !packed || !n || p->k < 2 || p->k > 64) return -1;
    if (p->k <= 32) return run<uint64_t>(p, packed, n);
    return run<u128>(p, packed, n);
}

extern "C" void uc_synth_free(void *ptr) { std::free(ptr); }

// Reference text format (driver.cpp:251-256): count line, one k-mer per line.
extern "C" int uc_synth_format_text(const uint32_t *packed, uint64_t n, uint32_t k, char *dst, uint64_t cap,
                                    uint64_t *len)
{
    static const char ACGT[4] = {'A', 'C', 'G', 'T'};
    char head[32];
    int hl = snprintf(head, sizeof(head), "%llu\n", (unsigned long long)n);
    uint64_t need = (uint64_t)hl + n * (k + 1);
    *len = need;
    if (!dst) return 0;
    if (cap < need) return -1;
    std::memcpy(dst, head, hl);
    const uint32_t w = (k + 15) / 16;
#pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < n; i++) {
        char *o = dst + hl + i * (k + 1);
        const uint32_t *x = packed + i * w;
        for (uint32_t c = 0; c < k; c++) o[c] = ACGT[(x[c / 16] >> (2 * (c & 15))) & 3u];
        o[k] = '\n';
    }
    return 0;
}
