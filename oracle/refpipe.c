/*
 * refpipe.c -- serial restatement of the reference compaction pipeline,
 * phase by phase (test infrastructure and timed CPU baseline only; the
 * product never links it).
 *
 * Every function follows the reference file:line cited above it.  The
 * parallel Kokkos kernels become plain loops; where the reference relies on
 * atomics the loop runs in index order, which is one of the schedules the
 * reference admits.  Ids are 32-bit like the default (neither LARGE nor
 * HUGE) build, defs.h:38-41.  Defined for k = 15 (mod 16) only, like the
 * reference's hard-coded prefix mask (assembler.h:10,28,50,68).
 *
 * Departures, each because the reference is undefined there:
 *  - a self-looping k-mer returns ORC_E_SELFLOOP (heuristics.h:119-147 spins
 *    forever when hn(u) == u);
 *  - edge_count gets n+1 cells so the v == n write of the off-by-one test
 *    at assembler.h:183 stays in bounds;
 *  - output partition offsets that driver.cpp:567-587 leaves unset (no row
 *    boundary crosses them) are computed as "first row whose midpoint falls
 *    in the partition", which is what the kernel intends.
 */
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "uc_oracle.h"
#include "oracle_internal.h"

typedef uint32_t ord_t;
#define ORD_MAX UINT32_MAX
static const uint32_t PREF_MASK = 268435455u; /* 0x0FFFFFFF, assembler.h:10 */

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* ------------------------------------------------------------------ */
/* growable arrays                                                    */

typedef struct {
    ord_t *v;
    uint64_t n, cap;
} vec_t;

static int vec_push(vec_t *a, ord_t x)
{
    if (a->n == a->cap) {
        uint64_t c = a->cap ? 2 * a->cap : 64;
        ord_t *p = (ord_t *)realloc(a->v, c * sizeof(ord_t));
        if (!p) return ORC_E_NOMEM;
        a->v = p;
        a->cap = c;
    }
    a->v[a->n++] = x;
    return ORC_OK;
}

/* compressed rows: rows[0..nrows], entries[rows[nrows]] (a StaticCrsGraph) */
typedef struct {
    uint64_t nrows;
    uint64_t *rows;
    ord_t *ent;
} crs_t;

static void crs_free(crs_t *g)
{
    free(g->rows);
    free(g->ent);
    g->rows = NULL;
    g->ent = NULL;
    g->nrows = 0;
}

static int crs_alloc(crs_t *g, uint64_t nrows, uint64_t nent)
{
    g->nrows = nrows;
    g->rows = (uint64_t *)calloc(nrows + 1, sizeof(uint64_t));
    g->ent = (ord_t *)malloc((nent ? nent : 1) * sizeof(ord_t));
    return (g->rows && g->ent) ? ORC_OK : ORC_E_NOMEM;
}

typedef struct {
    char *buf;
    uint64_t len, cap;
} sbuf_t;

static int sb_reserve(sbuf_t *b, uint64_t extra)
{
    if (b->len + extra <= b->cap) return ORC_OK;
    uint64_t c = b->cap ? b->cap : 4096;
    while (c < b->len + extra) c *= 2;
    char *p = (char *)realloc(b->buf, c);
    if (!p) return ORC_E_NOMEM;
    b->buf = p;
    b->cap = c;
    return ORC_OK;
}

static const char ACGT[4] = {'A', 'C', 'G', 'T'};

/* lmin_buckets.h:48-56 get_char2 */
static inline uint32_t get_char2(const uint32_t *src, uint64_t offset, uint64_t char_id)
{
    return (src[offset + char_id / 16] >> (2 * (char_id % 16))) & 3u;
}

/* lmin_buckets.h:58-65 write_char (the atomic_or becomes a plain or) */
static inline void write_char(uint32_t *dst, uint64_t offset, uint64_t char_id, uint32_t val)
{
    dst[offset + char_id / 16] |= val << (2 * (char_id % 16));
}

/* ------------------------------------------------------------------ */
/* l-mer order                                                        */

static uint64_t splitmix64(uint64_t *s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* driver.cpp:356-402 generate_permutation.  The Kokkos XorShift64 pool is
 * replaced by a seeded splitmix64 stream; any random source yields a random
 * permutation, which is all the bucketing needs. */
int orc_generate_permutation(uint32_t l, uint64_t seed, uint32_t *perm)
{
    if (l < 1 || l > 15) return ORC_E_ARG;
    uint64_t n = 1ull << (2 * l), t = 2 * n;
    uint64_t *randoms = (uint64_t *)malloc(n * sizeof(uint64_t));
    ord_t *slots = (ord_t *)malloc(t * sizeof(ord_t));
    if (!randoms || !slots) {
        free(randoms);
        free(slots);
        return ORC_E_NOMEM;
    }
    uint64_t s = seed;
    for (uint64_t i = 0; i < n; i++) randoms[i] = splitmix64(&s);
    for (uint64_t i = 0; i < t; i++) slots[i] = ORD_MAX;
    uint64_t bsize = UINT64_MAX / t;
    for (uint64_t i = 0; i < n; i++) {
        uint64_t b = randoms[i] / bsize;
        for (;; b++) {
            if (b >= t) b -= t;
            if (slots[b] == ORD_MAX) {
                slots[b] = (ord_t)i;
                break;
            }
        }
    }
    uint64_t j = 0;
    for (uint64_t i = 0; i < t; i++)
        if (slots[i] != ORD_MAX) perm[j++] = slots[i];
    free(randoms);
    free(slots);
    return ORC_OK;
}

/* lmin_buckets.h:67-94 get_double_lmin: the l-mer code keeps the first
 * character most significant; l1 takes l-mers inside the (k-1)-prefix,
 * l2 those inside the (k-1)-suffix. */
void orc_double_lmin(const uint32_t *kmer, uint32_t k, uint32_t l, const uint32_t *order, uint32_t *pref_min,
                     uint32_t *suf_min)
{
    const ord_t lmer_mask = (ord_t)((1ull << (2 * l)) - 1);
    ord_t lmer = 0, l1 = ORD_MAX, l2 = ORD_MAX;
    for (uint32_t i = 0; i < k; i++) {
        lmer = ((lmer << 2) ^ get_char2(kmer, 0, i)) & lmer_mask;
        if (i + 1 >= l) {
            ord_t lx = order[lmer];
            if (i + 1 < k && l1 > lx) l1 = lx;
            if (i + 1 >= l + 1 && l2 > lx) l2 = lx;
        }
    }
    *pref_min = l1;
    *suf_min = l2;
}

/* ------------------------------------------------------------------ */
/* bucketing: lmin_buckets.h:288-433 find_l_minimizer on one chunk plus   */
/* driver.cpp:137-225 collect_buckets (a single chunk needs no id shift). */

void orc_buckets_free(orc_buckets *b)
{
    free(b->bucket_rows);
    free(b->kmer_src);
    free(b->cross_rows);
    free(b->cross_lend);
    free(b->cross_borrow);
    memset(b, 0, sizeof(*b));
}

static int bucketize_nt(const uint32_t *packed, uint64_t n, uint32_t k, uint32_t l, uint32_t n_buckets,
                        const uint32_t *order, orc_buckets *out, int nt);

int orc_bucketize(const uint32_t *packed, uint64_t n, uint32_t k, uint32_t l, uint32_t n_buckets,
                  const uint32_t *order, orc_buckets *out)
{
    return bucketize_nt(packed, n, k, l, n_buckets, order, out, 1);
}

/* nt threads compute the minimizers (the Kokkos parallel_for of "calc
 * lmins"); counting and partitioning stay in k-mer order, so the layout is
 * the serial one whatever nt is */
static int bucketize_nt(const uint32_t *packed, uint64_t n, uint32_t k, uint32_t l, uint32_t n_buckets,
                        const uint32_t *order, orc_buckets *out, int nt)
{
    memset(out, 0, sizeof(*out));
    if (k < 2 || k > 64 || l < 1 || l > 15 || l >= k || n_buckets < 1 || (n_buckets & (n_buckets - 1)))
        return ORC_E_ARG;
    const uint64_t B = n_buckets, bmask = B - 1, nl = 1ull << (2 * l);
    const uint32_t w = orc_words_per_kmer(k);
    ord_t *out_l = (ord_t *)malloc((n ? n : 1) * sizeof(ord_t));
    ord_t *in_l = (ord_t *)malloc((n ? n : 1) * sizeof(ord_t));
    uint64_t *small_cnt = (uint64_t *)calloc(nl, sizeof(uint64_t));
    uint64_t *small_off = (uint64_t *)calloc(nl + 1, sizeof(uint64_t));
    uint64_t *cross_cnt = (uint64_t *)calloc(B * B, sizeof(uint64_t));
    out->bucket_rows = (uint64_t *)calloc(B + 1, sizeof(uint64_t));
    out->cross_rows = (uint64_t *)calloc(B * B + 1, sizeof(uint64_t));
    int rc = ORC_E_NOMEM;
    if (!out_l || !in_l || !small_cnt || !small_off || !cross_cnt || !out->bucket_rows || !out->cross_rows) goto fail;
    out->n_buckets = n_buckets;
    /* "calc lmins" (:299) */
#pragma omp parallel for num_threads(nt) schedule(static)
    for (uint64_t i = 0; i < n; i++) orc_double_lmin(packed + i * w, k, l, order, &out_l[i], &in_l[i]);
    /* "count lmins" (:306-318) */
    for (uint64_t i = 0; i < n; i++) {
        uint64_t bo = out_l[i] & bmask, bi = in_l[i] & bmask;
        small_cnt[out_l[i]]++;
        if (bo != bi) {
            small_cnt[in_l[i]]++;
            cross_cnt[bo * B + bi]++;
            cross_cnt[bi * B + bo]++;
        }
    }
    /* small buckets of one large bucket are laid out by order value (:325-333) */
    uint64_t off = 0;
    for (uint64_t i = 0; i < B; i++) {
        for (uint64_t j = i; j < nl; j += B) {
            small_off[j] = off;
            off += small_cnt[j];
        }
        out->bucket_rows[i + 1] = off;
    }
    small_off[nl] = off;
    memset(small_cnt, 0, nl * sizeof(uint64_t));
    out->m = off;
    /* "prefix sum crosscut" (:343-349) */
    for (uint64_t i = 0; i < B * B; i++) {
        out->cross_rows[i + 1] = out->cross_rows[i] + cross_cnt[i];
        cross_cnt[i] = 0;
    }
    uint64_t total_cross = out->cross_rows[B * B];
    out->kmer_src = (uint64_t *)malloc((off ? off : 1) * sizeof(uint64_t));
    out->cross_lend = (uint32_t *)malloc((total_cross ? total_cross : 1) * sizeof(uint32_t));
    out->cross_borrow = (uint32_t *)malloc((total_cross ? total_cross : 1) * sizeof(uint32_t));
    if (!out->kmer_src || !out->cross_lend || !out->cross_borrow) goto fail;
    for (uint64_t i = 0; i < total_cross; i++) out->cross_lend[i] = out->cross_borrow[i] = ORD_MAX;
    /* "partition by lmins" (:357-383): lend entries carry the home-bucket
     * local id, borrow entries the local id of the duplicate, both at the
     * same offset of the (out,in) and (in,out) cross buckets. */
    for (uint64_t i = 0; i < n; i++) {
        ord_t lo = out_l[i];
        uint64_t insert = small_off[lo] + small_cnt[lo]++;
        out->kmer_src[insert] = i;
        uint64_t bo = lo & bmask, bi = in_l[i] & bmask;
        if (bi != bo) {
            ord_t local_id = (ord_t)(insert - out->bucket_rows[bo]);
            ord_t wl = in_l[i];
            insert = small_off[wl] + small_cnt[wl]++;
            out->kmer_src[insert] = i;
            uint64_t ob = bo * B + bi, ib = bi * B + bo, less = ib < ob ? ib : ob;
            uint64_t o = cross_cnt[less]++;
            out->cross_lend[out->cross_rows[ob] + o] = local_id;
            out->cross_borrow[out->cross_rows[ib] + o] = (ord_t)(insert - out->bucket_rows[bi]);
        }
    }
    rc = ORC_OK;
fail:
    free(out_l);
    free(in_l);
    free(small_cnt);
    free(small_off);
    free(cross_cnt);
    if (rc) orc_buckets_free(out);
    return rc;
}

/* ------------------------------------------------------------------ */
/* edge finding: assembler.h                                          */

/* assembler.h:7-22 fnv_pref: FNV-1a over the bytes of the (k-1)-prefix
 * words, little-endian, last word masked.  The byte goes through a signed
 * `char`, so bytes >= 0x80 are xor-ed in sign-extended. */
static uint32_t fnv_pref(const uint32_t *kmer, uint32_t comp)
{
    uint32_t hash = 2166136261u;
    for (uint32_t i = 0; i < comp; i++) {
        uint32_t x = kmer[i];
        if (i + 1 == comp) x &= PREF_MASK;
        for (int b = 0; b < 4; b++) {
            int8_t y = (int8_t)((x >> (8 * b)) & 255);
            hash ^= (uint32_t)(int32_t)y;
            hash *= 16777619u;
        }
    }
    return hash;
}

/* assembler.h:24-45 fnv_suf: same over the (k-1)-suffix, word i being
 * (w_i >> 2) | (w_{i+1} & 3) << 30 and the last word masked. */
static uint32_t fnv_suf(const uint32_t *kmer, uint32_t comp)
{
    uint32_t hash = 2166136261u;
    for (uint32_t i = 0; i < comp; i++) {
        uint32_t x = kmer[i] >> 2;
        if (i + 1 == comp)
            x &= PREF_MASK;
        else
            x |= (kmer[i + 1] & 3u) << 30;
        for (int b = 0; b < 4; b++) {
            int8_t y = (int8_t)((x >> (8 * b)) & 255);
            hash ^= (uint32_t)(int32_t)y;
            hash *= 16777619u;
        }
    }
    return hash;
}

uint32_t orc_fnv_pref(const uint32_t *kmer, uint32_t words) { return fnv_pref(kmer, words); }
uint32_t orc_fnv_suf(const uint32_t *kmer, uint32_t words) { return fnv_suf(kmer, words); }

/* assembler.h:47-63 cmp_pref */
static int cmp_pref(const uint32_t *a, const uint32_t *b, uint32_t comp)
{
    for (uint32_t i = 0; i < comp; i++) {
        if (i + 1 == comp) {
            if ((a[i] & PREF_MASK) != (b[i] & PREF_MASK)) return 0;
        } else if (a[i] != b[i]) {
            return 0;
        }
    }
    return 1;
}

/* assembler.h:65-85 cmp_suf: suffix of s1 against prefix of s2 */
static int cmp_suf(const uint32_t *s1, const uint32_t *s2, uint32_t comp)
{
    for (uint32_t i = 0; i < comp; i++) {
        uint32_t suf = s1[i] >> 2;
        if (i + 1 == comp) {
            if ((suf & PREF_MASK) != (s2[i] & PREF_MASK)) return 0;
        } else {
            suf |= (s1[i + 1] & 3u) << 30;
            if (suf != s2[i]) return 0;
        }
    }
    return 1;
}

/* assembler.h:113-118 init_hashmap: power of two >= 2 * largest bucket */
static uint64_t hashmap_size(uint64_t max_size)
{
    uint64_t s = 1;
    while (s < 2 * max_size) s <<= 1;
    return s;
}

/* assembler.h:120-150 generate_hashmap.  Slot values: < n owner, >= n a
 * prefix shared by several k-mers ("nullified"), ORD_MAX empty. */
static void generate_hashmap(ord_t *map, uint64_t map_size, const uint32_t *kmers, uint32_t comp, ord_t n)
{
    for (uint64_t i = 0; i < map_size; i++) map[i] = ORD_MAX;
    uint64_t cast = map_size - 1;
    for (ord_t i = 0; i < n; i++) {
        uint64_t h = fnv_pref(kmers + (uint64_t)comp * i, comp) & cast;
        if (map[h] == ORD_MAX) {
            map[h] = i;
            continue;
        }
        for (;;) {
            ord_t written = map[h];
            int nullify = 0;
            if (written >= n)
                written -= n;
            else
                nullify = 1;
            if (cmp_pref(kmers + (uint64_t)comp * i, kmers + (uint64_t)comp * written, comp)) {
                if (nullify) map[h] = i + n;
                break;
            }
            h = (h + 1) & cast;
            if (map[h] == ORD_MAX) {
                map[h] = i;
                break;
            }
        }
    }
}

/* assembler.h:87-111 find_vtx_from_edge */
static ord_t find_vtx_from_edge(const uint32_t *kmers, const ord_t *map, uint64_t map_size, uint64_t offset,
                                uint32_t comp, ord_t size)
{
    uint64_t cast = map_size - 1;
    uint64_t h = fnv_suf(kmers + offset, comp) & cast;
    while (map[h] != ORD_MAX) {
        ord_t addr = map[h];
        int nullified = 0;
        if (addr >= size) {
            addr -= size;
            nullified = 1;
        }
        if (cmp_suf(kmers + offset, kmers + (uint64_t)addr * comp, comp)) return nullified ? addr + size : addr;
        h = (h + 1) & cast;
    }
    return ORD_MAX;
}

typedef struct {
    ord_t *in, *out;
    uint64_t n;
} crosses_t;

/* assembler.h:164-214 assemble_pruned_graph */
static int assemble_pruned_graph(const uint32_t *kmers, uint32_t comp, ord_t n, const ord_t *map, uint64_t map_size,
                                 const ord_t *lends, const ord_t *borrows, uint64_t n_cross, ord_t *g,
                                 crosses_t *c)
{
    ord_t *in = (ord_t *)malloc(((uint64_t)n + 1) * sizeof(ord_t));
    ord_t *edge_count = (ord_t *)calloc((uint64_t)n + 1, sizeof(ord_t));
    c->in = (ord_t *)malloc((n_cross ? n_cross : 1) * sizeof(ord_t));
    c->out = (ord_t *)malloc((n_cross ? n_cross : 1) * sizeof(ord_t));
    c->n = n_cross;
    if (!in || !edge_count || !c->in || !c->out) {
        free(in);
        free(edge_count);
        return ORC_E_NOMEM;
    }
    for (ord_t i = 0; i < n; i++) in[i] = find_vtx_from_edge(kmers, map, map_size, (uint64_t)i * comp, comp, n);
    for (uint64_t i = 0; i < n_cross; i++)
        if (lends[i] != ORD_MAX) in[lends[i]] = ORD_MAX; /* "mark lends" */
    for (ord_t i = 0; i < n; i++) {
        ord_t v = in[i];
        if (v != ORD_MAX) {
            if (v > n) v -= n; /* the reference's `>` (should be `>=`) */
            edge_count[v]++;
        }
    }
    for (ord_t i = 0; i < n; i++) {
        ord_t v = in[i];
        if (v < n && edge_count[v] == 1) g[i] = v;
    }
    for (uint64_t i = 0; i < n_cross; i++) {
        c->out[i] = lends[i];
        c->in[i] = borrows[i];
    }
    free(in);
    free(edge_count);
    return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* HEC matching: heuristics.h:105-256                                 */

typedef struct {
    ord_t *ent;
    ord_t nc, n;
} interp_t;

/* heuristics.h:205-256 coarsen_HEC with parallel_map_construct
 * (:106-203) run in vertex order.  In that order every CAS on a free
 * vertex succeeds, so a vertex pairs with its hn if hn is free and joins
 * hn's aggregate otherwise; the "skipped mappings" pass finds nothing. */
static int coarsen_hec(const ord_t *g, ord_t n, interp_t *out)
{
    ord_t *hn = (ord_t *)malloc(((uint64_t)n + 1) * sizeof(ord_t));
    ord_t *vc = (ord_t *)malloc(((uint64_t)n + 1) * sizeof(ord_t));
    if (!hn || !vc) {
        free(hn);
        free(vc);
        return ORC_E_NOMEM;
    }
    for (ord_t i = 0; i < n; i++) vc[i] = hn[i] = ORD_MAX;
    for (ord_t i = 0; i < n; i++) /* "edge choose": predecessor */
        if (g[i] != ORD_MAX) hn[g[i]] = i;
    for (ord_t i = 0; i < n; i++) { /* "edge choose": else successor, else null aggregate */
        if (hn[i] == ORD_MAX) {
            if (g[i] != ORD_MAX)
                hn[i] = g[i];
            else
                vc[i] = 0;
        }
    }
    for (ord_t u = 0; u < n; u++) { /* "compute mappings" */
        ord_t v = hn[u];
        if (v == ORD_MAX) continue;
        if (v == u) {
            free(hn);
            free(vc);
            return ORC_E_SELFLOOP;
        }
        if (vc[u] != ORD_MAX) continue;
        if (vc[v] == ORD_MAX) {
            vc[u] = vc[v] = u + 1;
        } else {
            vc[u] = vc[v]; /* join v's aggregate (vc[v] <= n) */
        }
    }
    ord_t cnt = 0; /* "assign aggregates" */
    for (ord_t u = 0; u < n; u++) {
        if (vc[u] - 1 == u) {
            vc[u] = ++cnt;
        } else if (vc[u] > 0) {
            vc[u] = vc[u] - 1 + n;
        }
    }
    for (ord_t u = 0; u < n; u++) /* "propagate aggregates" */
        if (vc[u] >= n) vc[u] = vc[vc[u] - n];
    free(hn);
    out->ent = vc;
    out->nc = cnt + 1;
    out->n = n;
    return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* coarsening: coarseners.h, compact_graph.h                          */

/* coarseners.h:68-92 coarsen_de_bruijn_graph */
static ord_t *coarsen_de_bruijn_graph(const ord_t *g, const interp_t *interp)
{
    ord_t nc = interp->nc - 1;
    ord_t *e = (ord_t *)malloc(((uint64_t)nc + 1) * sizeof(ord_t));
    if (!e) return NULL;
    for (ord_t i = 0; i < nc; i++) e[i] = ORD_MAX;
    for (ord_t i = 0; i < interp->n; i++) {
        ord_t u = interp->ent[i];
        if (u > 0 && u != ORD_MAX) {
            ord_t f = g[i];
            if (f != ORD_MAX) {
                ord_t v = interp->ent[f];
                if (u != v) e[u - 1] = v - 1;
            }
        }
    }
    return e;
}

/* coarseners.h:125-146 transpose_null: vertices mapped to null_id, in order */
static int transpose_null(const interp_t *interp, ord_t null_id, vec_t *out)
{
    memset(out, 0, sizeof(*out));
    for (ord_t i = 0; i < interp->n; i++)
        if (interp->ent[i] == null_id && vec_push(out, i)) return ORC_E_NOMEM;
    return ORC_OK;
}

/* coarseners.h:149-235 transpose_non_null + transpose_and_sort: members of
 * every aggregate, then ordered along the path by a backward search from
 * the aggregate's exit vertex. */
static int transpose_and_sort(const interp_t *interp, const ord_t *g, crs_t *out)
{
    ord_t nc = interp->nc - 1;
    uint64_t *cnt = (uint64_t *)calloc((uint64_t)nc + 1, sizeof(uint64_t));
    if (!cnt) return ORC_E_NOMEM;
    uint64_t total = 0;
    for (ord_t i = 0; i < interp->n; i++)
        if (interp->ent[i] != 0) {
            cnt[interp->ent[i] - 1]++;
            total++;
        }
    if (crs_alloc(out, nc, total)) {
        free(cnt);
        return ORC_E_NOMEM;
    }
    for (ord_t r = 0; r < nc; r++) {
        out->rows[r + 1] = out->rows[r] + cnt[r];
        cnt[r] = 0;
    }
    for (ord_t i = 0; i < interp->n; i++) {
        ord_t v = interp->ent[i];
        if (v != 0) out->ent[out->rows[v - 1] + cnt[v - 1]++] = i;
    }
    free(cnt);
    for (ord_t r = 0; r < nc; r++) {
        ord_t vtx_id = r + 1;
        uint64_t start = out->rows[r], end = out->rows[r + 1];
        ord_t end_vertex = ORD_MAX;
        for (uint64_t x = start; x < end; x++) {
            ord_t u = out->ent[x];
            if (g[u] == ORD_MAX || interp->ent[g[u]] != vtx_id) {
                out->ent[x] = out->ent[end - 1];
                out->ent[end - 1] = u;
                end_vertex = u;
                break;
            }
        }
        end--;
        while (end > start) {
            for (uint64_t x = start; x < end; x++) {
                ord_t u = out->ent[x];
                if (g[u] == end_vertex) {
                    out->ent[x] = out->ent[end - 1];
                    out->ent[end - 1] = u;
                    end_vertex = u;
                    break;
                }
            }
            end--;
        }
    }
    return ORC_OK;
}

/* coarseners.h:237-246 collect_outputs_first: one single-vertex row per
 * vertex mapped to null_id */
static int collect_outputs_first(const interp_t *interp, ord_t null_id, crs_t *out)
{
    vec_t v;
    if (transpose_null(interp, null_id, &v)) return ORC_E_NOMEM;
    out->nrows = v.n;
    out->ent = v.v ? v.v : (ord_t *)malloc(sizeof(ord_t));
    out->rows = (uint64_t *)malloc((v.n + 1) * sizeof(uint64_t));
    if (!out->rows || !out->ent) return ORC_E_NOMEM;
    for (uint64_t i = 0; i <= v.n; i++) out->rows[i] = i;
    return ORC_OK;
}

/* compact_graph.h:23-52 collect_outputs: the fine-id rows of glue_old
 * selected by nulls */
static int collect_outputs(const crs_t *old, const vec_t *nulls, crs_t *out)
{
    uint64_t total = 0;
    for (uint64_t i = 0; i < nulls->n; i++) total += old->rows[nulls->v[i] + 1] - old->rows[nulls->v[i]];
    if (crs_alloc(out, nulls->n, total)) return ORC_E_NOMEM;
    for (uint64_t i = 0; i < nulls->n; i++) {
        ord_t u = nulls->v[i];
        uint64_t s = old->rows[u], e = old->rows[u + 1];
        out->rows[i + 1] = out->rows[i] + (e - s);
        memcpy(out->ent + out->rows[i], old->ent + s, (e - s) * sizeof(ord_t));
    }
    return ORC_OK;
}

/* compact_graph.h:55-112 collect_unitigs: compose glue_old through the
 * aggregates of the next level */
static int collect_unitigs(const crs_t *old, const crs_t *action, crs_t *out)
{
    uint64_t total = 0;
    for (uint64_t u = 0; u < action->nrows; u++)
        for (uint64_t i = action->rows[u]; i < action->rows[u + 1]; i++)
            total += old->rows[action->ent[i] + 1] - old->rows[action->ent[i]];
    if (crs_alloc(out, action->nrows, total)) return ORC_E_NOMEM;
    uint64_t w = 0;
    for (uint64_t u = 0; u < action->nrows; u++) {
        for (uint64_t i = action->rows[u]; i < action->rows[u + 1]; i++) {
            ord_t f = action->ent[i];
            uint64_t s = old->rows[f], e = old->rows[f + 1];
            memcpy(out->ent + w, old->ent + s, (e - s) * sizeof(ord_t));
            w += e - s;
        }
        out->rows[u + 1] = w;
    }
    return ORC_OK;
}

/* coarseners.h:270-347 coarsen_crosses: crossing k-mers whose path has
 * finished (null aggregate) are diverted to the ORD_MAX aggregate and
 * returned; the others follow their aggregate to the next level. */
static int coarsen_crosses(crosses_t *c, vec_t *rem, interp_t *interp, vec_t *cross_glues)
{
    vec_t keep;
    memset(&keep, 0, sizeof(keep));
    memset(cross_glues, 0, sizeof(*cross_glues));
    for (uint64_t x = 0; x < rem->n; x++) {
        ord_t i = rem->v[x];
        ord_t f = c->in[i] != ORD_MAX ? c->in[i] : c->out[i];
        if (f == ORD_MAX) continue;
        if (interp->ent[f] == 0) {
            if (vec_push(cross_glues, i)) return ORC_E_NOMEM;
        } else if (vec_push(&keep, i)) {
            return ORC_E_NOMEM;
        }
    }
    free(rem->v);
    *rem = keep;
    for (uint64_t x = 0; x < rem->n; x++) { /* "coarsen ins" (both passes) */
        ord_t i = rem->v[x];
        if (c->in[i] != ORD_MAX) c->in[i] = interp->ent[c->in[i]] - 1;
        if (c->out[i] != ORD_MAX) c->out[i] = interp->ent[c->out[i]] - 1;
    }
    for (uint64_t x = 0; x < cross_glues->n; x++) { /* "modify interp" */
        ord_t i = cross_glues->v[x];
        ord_t f = c->in[i] != ORD_MAX ? c->in[i] : c->out[i];
        interp->ent[f] = ORD_MAX;
    }
    return ORC_OK;
}

/* coarseners.h:248-268 relabel_crosses: finished crossing paths get dense
 * ids offset + rank, written back into the crossing lists */
static ord_t relabel_crosses(crosses_t *c, const vec_t *cross_glues, interp_t *interp, ord_t offset)
{
    ord_t r = 0;
    for (ord_t i = 0; i < interp->n; i++)
        if (interp->ent[i] == ORD_MAX) interp->ent[i] = r++;
    for (uint64_t x = 0; x < cross_glues->n; x++) {
        ord_t i = cross_glues->v[x];
        if (c->in[i] != ORD_MAX)
            c->in[i] = offset + interp->ent[c->in[i]];
        else if (c->out[i] != ORD_MAX)
            c->out[i] = offset + interp->ent[c->out[i]];
    }
    return offset + r;
}

/* coarseners.h:361-386 append_graph / collapse_list */
typedef struct {
    crs_t *g;
    uint64_t n, cap;
} crs_list_t;

static int list_push(crs_list_t *l, crs_t g)
{
    if (l->n == l->cap) {
        uint64_t c = l->cap ? 2 * l->cap : 8;
        crs_t *p = (crs_t *)realloc(l->g, c * sizeof(crs_t));
        if (!p) return ORC_E_NOMEM;
        l->g = p;
        l->cap = c;
    }
    l->g[l->n++] = g;
    return ORC_OK;
}

static int collapse_list(crs_list_t *l, crs_t *out)
{
    uint64_t rows = 0, ent = 0;
    for (uint64_t i = 0; i < l->n; i++) {
        rows += l->g[i].nrows;
        ent += l->g[i].rows[l->g[i].nrows];
    }
    if (crs_alloc(out, rows, ent)) return ORC_E_NOMEM;
    uint64_t r = 0, e = 0;
    for (uint64_t i = 0; i < l->n; i++) {
        crs_t *g = &l->g[i];
        for (uint64_t j = 0; j < g->nrows; j++) out->rows[r + j + 1] = g->rows[j + 1] + e;
        memcpy(out->ent + e, g->ent, g->rows[g->nrows] * sizeof(ord_t));
        r += g->nrows;
        e += g->rows[g->nrows];
        crs_free(g);
    }
    free(l->g);
    memset(l, 0, sizeof(*l));
    return ORC_OK;
}

/* coarseners.h:437-485 coarsen_de_bruijn_full_cycle (with crossing
 * tracking) and :388-435 coarsen_de_bruijn_full_cycle_final (c == NULL).
 * Consumes cur. */
static int coarsen_full_cycle(ord_t *cur, ord_t n, crosses_t *c, ord_t *cross_offset, crs_t *glue_out,
                              crs_t *cross_out)
{
    crs_list_t glue_list = {0, 0, 0}, cross_list = {0, 0, 0};
    crs_t glue_last = {0, 0, 0};
    vec_t rem;
    memset(&rem, 0, sizeof(rem));
    if (c)
        for (uint64_t i = 0; i < c->n; i++)
            if (vec_push(&rem, (ord_t)i)) return ORC_E_NOMEM;
    int first = 1, rc = ORC_OK;
    while (n > 0) {
        interp_t interp;
        crs_t glue;
        vec_t cross_glues;
        memset(&cross_glues, 0, sizeof(cross_glues));
        if ((rc = coarsen_hec(cur, n, &interp))) break;
        if ((rc = transpose_and_sort(&interp, cur, &glue))) break;
        if (c && (rc = coarsen_crosses(c, &rem, &interp, &cross_glues))) break;
        if (first) {
            crs_t o;
            if ((rc = collect_outputs_first(&interp, 0, &o)) || (rc = list_push(&glue_list, o))) break;
            if (c && ((rc = collect_outputs_first(&interp, ORD_MAX, &o)) || (rc = list_push(&cross_list, o)))) break;
            glue_last = glue;
        } else {
            vec_t nulls;
            crs_t o, nxt;
            if ((rc = transpose_null(&interp, 0, &nulls))) break;
            rc = collect_outputs(&glue_last, &nulls, &o);
            free(nulls.v);
            if (rc || (rc = list_push(&glue_list, o))) break;
            if (c) {
                vec_t xs;
                if ((rc = transpose_null(&interp, ORD_MAX, &xs))) break;
                rc = collect_outputs(&glue_last, &xs, &o);
                free(xs.v);
                if (rc || (rc = list_push(&cross_list, o))) break;
            }
            if ((rc = collect_unitigs(&glue_last, &glue, &nxt))) break;
            crs_free(&glue_last);
            crs_free(&glue);
            glue_last = nxt;
        }
        if (c) *cross_offset = relabel_crosses(c, &cross_glues, &interp, *cross_offset);
        free(cross_glues.v);
        first = 0;
        ord_t *next = coarsen_de_bruijn_graph(cur, &interp);
        ord_t nn = interp.nc - 1;
        free(interp.ent);
        free(cur);
        cur = next;
        n = nn;
        if (!cur) {
            rc = ORC_E_NOMEM;
            break;
        }
    }
    free(cur);
    free(rem.v);
    crs_free(&glue_last);
    if (rc) return rc;
    if ((rc = collapse_list(&glue_list, glue_out))) return rc;
    if (c) return collapse_list(&cross_list, cross_out);
    return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* output                                                             */

/* compact.h:161-210 write_intra_bucket_outputs: first k-mer in full, then
 * the last character of each later k-mer, then '\n' */
static int write_intra(sbuf_t *sb, const uint32_t *kmers, uint32_t k, uint32_t comp, const crs_t *glue,
                       uint64_t *lines)
{
    for (uint64_t i = 0; i < glue->nrows; i++) {
        uint64_t s = glue->rows[i], e = glue->rows[i + 1];
        if (sb_reserve(sb, (e - s) + k)) return ORC_E_NOMEM;
        for (uint64_t j = s; j < e; j++) {
            ord_t u = glue->ent[j];
            if (j > s) {
                sb->buf[sb->len++] = ACGT[get_char2(kmers, (uint64_t)u * comp, k - 1)];
            } else {
                for (uint32_t x = 0; x < k; x++) sb->buf[sb->len++] = ACGT[get_char2(kmers, (uint64_t)u * comp, x)];
            }
        }
        sb->buf[sb->len++] = '\n';
        (*lines)++;
    }
    return ORC_OK;
}

/* lmin_buckets.h:121-126 minitigs */
typedef struct {
    ord_t *lengths;
    uint64_t *rows; /* in 16-character words */
    uint32_t *chars;
    uint64_t size;
} minitigs_t;

/* lmin_buckets.h:172-214 generate_minitigs: 2-bit spelling of each crossing
 * path, first character of every k-mer but the last, which is spelled in
 * full; rows padded to whole 16-character words */
static int generate_minitigs(const crs_t *glues, const uint32_t *kmers, uint32_t k, uint32_t comp, minitigs_t *m)
{
    uint64_t n = glues->nrows;
    m->size = n;
    m->lengths = (ord_t *)malloc((n ? n : 1) * sizeof(ord_t));
    m->rows = (uint64_t *)calloc(n + 1, sizeof(uint64_t));
    if (!m->lengths || !m->rows) return ORC_E_NOMEM;
    for (uint64_t i = 0; i < n; i++) {
        m->lengths[i] = (ord_t)(k - 1 + glues->rows[i + 1] - glues->rows[i]);
        m->rows[i + 1] = m->rows[i] + (m->lengths[i] + 15) / 16;
    }
    m->chars = (uint32_t *)calloc(m->rows[n] ? m->rows[n] : 1, sizeof(uint32_t));
    if (!m->chars) return ORC_E_NOMEM;
    for (uint64_t i = 0; i < n; i++) {
        uint64_t wt = m->rows[i];
        for (uint64_t j = glues->rows[i]; j < glues->rows[i + 1]; j++) {
            ord_t u = glues->ent[j];
            uint64_t local = j - glues->rows[i];
            if (j + 1 < glues->rows[i + 1]) {
                write_char(m->chars, wt, local, get_char2(kmers, (uint64_t)u * comp, 0));
            } else {
                for (uint32_t x = 0; x < k; x++) write_char(m->chars, wt, local++, get_char2(kmers, (uint64_t)u * comp, x));
            }
        }
    }
    return ORC_OK;
}

/* compact.h:59-112 write_unitigs4 for one final row: every minitig but the
 * last drops its last k characters (the k-mer it shares with the next
 * one); '\n' ends the line */
static int write_stitched_row(sbuf_t *sb, const minitigs_t *all, const ord_t *ent, uint64_t s, uint64_t e, uint32_t k)
{
    uint64_t size = 1;
    for (uint64_t j = s; j < e; j++) size += all->lengths[ent[j]] - (j + 1 < e ? k : 0);
    if (sb_reserve(sb, size)) return ORC_E_NOMEM;
    for (uint64_t j = s; j < e; j++) {
        ord_t u = ent[j];
        uint64_t wsize = all->lengths[u] - (j + 1 < e ? k : 0);
        for (uint64_t x = 0; x < wsize; x++) sb->buf[sb->len++] = ACGT[get_char2(all->chars, all->rows[u], x)];
    }
    sb->buf[sb->len++] = '\n';
    return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* driver.cpp:422-662 main, from bucketing to the last append          */

int orc_refpipe(const uint32_t *packed, uint64_t n, uint32_t k, uint32_t l, uint32_t n_buckets,
                const uint32_t *order, char **text, uint64_t *text_len, orc_refpipe_stats *st)
{
    return orc_refpipe_par(packed, n, k, l, n_buckets, order, 1, text, text_len, st);
}

/* The reference's Kokkos OpenMP build runs every bucket's hashmap, graph,
 * coarsening and output on all host threads, bucket after bucket
 * (driver.cpp:479-521).  Here the threads take whole buckets instead
 * (buckets are independent until the cross graph): each bucket numbers its
 * finished crossing paths from 0 and writes its own text, and a prefix over
 * the buckets then gives the ids the serial loop's running offset gives, so
 * the text is the same bytes for every thread count. */
int orc_refpipe_par(const uint32_t *packed, uint64_t n, uint32_t k, uint32_t l, uint32_t n_buckets,
                    const uint32_t *order, int nthreads, char **text, uint64_t *text_len, orc_refpipe_stats *st)
{
    if (k % 16 != 15 || k > 63 || l < 1 || l >= k) return ORC_E_ARG;
    const int nt = nthreads > 0 ? nthreads : 1;
    orc_refpipe_stats stats;
    memset(&stats, 0, sizeof(stats));
    stats.threads = (uint64_t)nt;
    const uint32_t comp = orc_words_per_kmer(k);
    const uint64_t B = n_buckets;
    double t0 = now_s(), t = t0;
    orc_buckets bk;
    int rc = bucketize_nt(packed, n, k, l, n_buckets, order, &bk, nt);
    if (rc) return rc;
    stats.bucketing = now_s() - t;
    stats.n_kmers = n;
    stats.m_placements = bk.m;
    stats.n_cross = bk.cross_rows[B * B];

    sbuf_t sb = {0, 0, 0};
    uint64_t largest = 0;
    for (uint64_t i = 0; i < B; i++)
        if (bk.bucket_rows[i + 1] - bk.bucket_rows[i] > largest) largest = bk.bucket_rows[i + 1] - bk.bucket_rows[i];
    const uint64_t map_size = hashmap_size(largest);
    crosses_t *cross_list = (crosses_t *)calloc(B, sizeof(crosses_t));
    minitigs_t *mini = (minitigs_t *)calloc(B, sizeof(minitigs_t));
    sbuf_t *btext = (sbuf_t *)calloc(B, sizeof(sbuf_t));
    ord_t *bcross = (ord_t *)calloc(B + 1, sizeof(ord_t)); /* finished crossing paths per bucket */
    uint64_t *blines = (uint64_t *)calloc(B, sizeof(uint64_t));
    double tph[4] = {0, 0, 0, 0}; /* thread-seconds: assemble, coarsen, write, minitigs */
    ord_t cross_offset = 0;
    if (!cross_list || !mini || !btext || !bcross || !blines) {
        rc = ORC_E_NOMEM;
        goto out;
    }
    t = now_s();
#pragma omp parallel num_threads(nt) reduction(+ : tph[:4])
    {
        /* the hashmap and the bucket's k-mers: per thread, sized for the
         * largest bucket as driver.cpp:454-466 sizes them once */
        ord_t *map = (ord_t *)malloc(map_size * sizeof(ord_t));
        uint32_t *kb = (uint32_t *)malloc((largest ? largest : 1) * comp * sizeof(uint32_t));
        int trc = (!map || !kb) ? ORC_E_NOMEM : ORC_OK;
#pragma omp for schedule(dynamic, 1)
        for (uint64_t b = 0; b < B; b++) {
            if (trc) continue;
            /* driver.cpp:479-521: the per-bucket loop */
            uint64_t s = bk.bucket_rows[b], cnt = bk.bucket_rows[b + 1] - s;
            for (uint64_t i = 0; i < cnt; i++)
                memcpy(kb + i * comp, packed + bk.kmer_src[s + i] * comp, comp * sizeof(uint32_t));
            uint64_t cs = bk.cross_rows[B * b], ce = bk.cross_rows[B * (b + 1)];
            double tt = now_s();
            generate_hashmap(map, map_size, kb, comp, (ord_t)cnt);
            ord_t *g = (ord_t *)malloc((cnt ? cnt : 1) * sizeof(ord_t));
            if (!g) { trc = ORC_E_NOMEM; continue; }
            for (uint64_t i = 0; i < cnt; i++) g[i] = ORD_MAX;
            trc = assemble_pruned_graph(kb, comp, (ord_t)cnt, map, map_size, bk.cross_lend + cs, bk.cross_borrow + cs,
                                        ce - cs, g, &cross_list[b]);
            tph[0] += now_s() - tt;
            if (trc) { free(g); continue; }
            tt = now_s();
            crs_t glue, cross;
            ord_t local = 0;
            trc = coarsen_full_cycle(g, (ord_t)cnt, &cross_list[b], &local, &glue, &cross);
            tph[1] += now_s() - tt;
            if (trc) continue;
            bcross[b] = local;
            tt = now_s();
            trc = write_intra(&btext[b], kb, k, comp, &glue, &blines[b]);
            tph[2] += now_s() - tt;
            tt = now_s();
            if (!trc) trc = generate_minitigs(&cross, kb, k, comp, &mini[b]);
            tph[3] += now_s() - tt;
            crs_free(&glue);
            crs_free(&cross);
        }
        free(map);
        free(kb);
        if (trc) {
#pragma omp critical
            rc = trc;
        }
    }
    stats.buckets_wall = now_s() - t;
    stats.assemble = tph[0];
    stats.coarsen = tph[1];
    stats.write_intra = tph[2];
    stats.minitigs = tph[3];
    if (rc) goto out;
    /* crossing ids: bucket b's follow every earlier bucket's; the texts go
     * out bucket by bucket */
    for (uint64_t b = 0; b < B && !rc; b++) {
        const ord_t off = cross_offset;
        if (off)
            for (uint64_t i = 0; i < cross_list[b].n; i++) {
                if (cross_list[b].in[i] != ORD_MAX) cross_list[b].in[i] += off;
                if (cross_list[b].out[i] != ORD_MAX) cross_list[b].out[i] += off;
            }
        cross_offset += bcross[b];
        stats.n_lines += blines[b];
        if (btext[b].len) {
            if (sb_reserve(&sb, btext[b].len)) rc = ORC_E_NOMEM;
            else {
                memcpy(sb.buf + sb.len, btext[b].buf, btext[b].len);
                sb.len += btext[b].len;
            }
        }
    }
    if (rc) goto out;
    t = now_s();
    {
        /* driver.cpp:527-556 small graph from the aligned lend/borrow lists */
        ord_t *small_g = (ord_t *)malloc(((uint64_t)cross_offset + 1) * sizeof(ord_t));
        if (!small_g) { rc = ORC_E_NOMEM; goto out; }
        for (ord_t i = 0; i < cross_offset; i++) small_g[i] = ORD_MAX;
        for (uint64_t i = 0; i < B; i++)
            for (uint64_t j = 0; j < B; j++) {
                if (i == j) continue;
                uint64_t ob = bk.cross_rows[B * i + j] - bk.cross_rows[B * i];
                uint64_t ib = bk.cross_rows[B * j + i] - bk.cross_rows[B * j];
                uint64_t sz = bk.cross_rows[B * i + j + 1] - bk.cross_rows[B * i + j];
                for (uint64_t x = 0; x < sz; x++) {
                    ord_t u = cross_list[i].out[ob + x], v = cross_list[j].in[ib + x];
                    if (u != ORD_MAX && v != ORD_MAX) small_g[u] = v;
                }
            }
        stats.cross_offset = cross_offset;
        /* coarseners.h:388-435 on the small graph */
        crs_t res, unused;
        ord_t dummy = 0;
        memset(&unused, 0, sizeof(unused));
        rc = coarsen_full_cycle(small_g, cross_offset, NULL, &dummy, &res, &unused);
        if (rc) goto out;
        /* global cross id -> (bucket, row) of its minitig: the buckets'
         * cross rows were numbered consecutively bucket after bucket */
        minitigs_t all;
        memset(&all, 0, sizeof(all));
        all.size = cross_offset;
        all.lengths = (ord_t *)malloc(((uint64_t)cross_offset + 1) * sizeof(ord_t));
        all.rows = (uint64_t *)calloc((uint64_t)cross_offset + 1, sizeof(uint64_t));
        uint64_t words = 0;
        for (uint64_t b = 0; b < B; b++) words += mini[b].rows[mini[b].size];
        all.chars = (uint32_t *)malloc((words ? words : 1) * sizeof(uint32_t));
        if (!all.lengths || !all.rows || !all.chars) { rc = ORC_E_NOMEM; goto out; }
        uint64_t r = 0, wo = 0;
        for (uint64_t b = 0; b < B; b++) {
            for (uint64_t i = 0; i < mini[b].size; i++, r++) {
                all.lengths[r] = mini[b].lengths[i];
                all.rows[r + 1] = all.rows[r] + (mini[b].rows[i + 1] - mini[b].rows[i]);
            }
            memcpy(all.chars + wo, mini[b].chars, mini[b].rows[mini[b].size] * sizeof(uint32_t));
            wo += mini[b].rows[mini[b].size];
        }
        /* driver.cpp:561-639: rows go to output partitions by midpoint and
         * are written partition by partition, i.e. in row order */
        uint64_t E = res.rows[res.nrows], ps = (E + B) / B;
        uint64_t p = 0;
        for (uint64_t row = 0; row < res.nrows && !rc; row++) {
            uint64_t mid = (res.rows[row] + res.rows[row + 1]) / 2;
            while (p < mid / ps) p++; /* partition p starts at this row */
            rc = write_stitched_row(&sb, &all, res.ent, res.rows[row], res.rows[row + 1], k);
            stats.n_lines++;
        }
        free(all.lengths);
        free(all.rows);
        free(all.chars);
        crs_free(&res);
    }
    stats.cross_finish = now_s() - t;
out:
    if (cross_list)
        for (uint64_t b = 0; b < B; b++) {
            free(cross_list[b].in);
            free(cross_list[b].out);
        }
    if (mini)
        for (uint64_t b = 0; b < B; b++) {
            free(mini[b].lengths);
            free(mini[b].rows);
            free(mini[b].chars);
        }
    if (btext)
        for (uint64_t b = 0; b < B; b++) free(btext[b].buf);
    free(btext);
    free(bcross);
    free(blines);
    free(cross_list);
    free(mini);
    orc_buckets_free(&bk);
    stats.total = now_s() - t0;
    if (st) *st = stats;
    if (rc) {
        free(sb.buf);
        return rc;
    }
    if (!sb.buf) sb.buf = (char *)malloc(1);
    *text = sb.buf;
    *text_len = sb.len;
    return ORC_OK;
}
