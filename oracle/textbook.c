/*
 * textbook.c -- closed-form oracle: maximal non-branching paths of the
 * directed node-centric de Bruijn graph of a k-mer set (test infrastructure
 * only; the product never links it).
 *
 * Edge rule (what assembler.h:120-197 computes once its buckets are put back
 * together, lmin_buckets.h:357-383): u -> v is kept iff suffix_{k-1}(u) ==
 * prefix_{k-1}(v), exactly one k-mer has prefix suffix(u) (out-degree 1,
 * the non-"nullified" slot of assembler.h:130-148) and exactly one k-mer has
 * suffix prefix(v) (in-degree 1, assembler.h:180-197).  There is no reverse
 * complement anywhere.
 *
 * Output (compact.h:161-210, lmin_buckets.h:172-204): each unitig is its
 * first k-mer followed by the last character of every later k-mer, then
 * '\n'.  Unitigs are emitted walking from every head (no kept in-edge) in
 * input order; vertices left over lie on isolated cycles, which are spelled
 * starting at their lexicographically smallest k-mer (the reference leaves
 * the order of a cycle undefined, coarseners.h:191-230).
 */
#include <stdlib.h>
#include <string.h>

#include "uc_oracle.h"
#include "oracle_internal.h"

typedef struct {
    orc_u128 key;
    uint64_t owner;    /* last k-mer whose prefix is key */
    uint32_t n_pref;   /* k-mers with prefix == key */
    uint32_t n_suf;    /* k-mers with suffix == key */
    uint32_t used;
} slot_t;

typedef struct {
    slot_t *s;
    uint64_t mask;
} table_t;

static uint64_t mix64(uint64_t x)
{
    x ^= x >> 30;
    x *= 0xBF58476D1CE4E5B9ull;
    x ^= x >> 27;
    x *= 0x94D049BB133111EBull;
    x ^= x >> 31;
    return x;
}

static uint64_t hash128(orc_u128 k)
{
    return mix64((uint64_t)k ^ mix64((uint64_t)(k >> 64) + 0x9E3779B97F4A7C15ull));
}

static slot_t *table_get(table_t *t, orc_u128 key, int create)
{
    uint64_t h = hash128(key) & t->mask;
    for (;;) {
        slot_t *s = &t->s[h];
        if (!s->used) {
            if (!create) return NULL;
            s->used = 1;
            s->key = key;
            s->owner = UINT64_MAX;
            return s;
        }
        if (s->key == key) return s;
        h = (h + 1) & t->mask;
    }
}

static int table_init(table_t *t, uint64_t entries)
{
    uint64_t size = 1;
    while (size < 2 * entries + 2) size <<= 1;
    t->s = (slot_t *)calloc(size, sizeof(slot_t));
    t->mask = size - 1;
    return t->s ? ORC_OK : ORC_E_NOMEM;
}

/* succ[i] of the pruned graph; also rejects duplicated k-mers. */
static int build_succ(const uint32_t *packed, uint64_t n, uint32_t k, uint64_t *succ)
{
    if (k < 2 || k > 64) return ORC_E_ARG;
    uint32_t w = orc_words_per_kmer(k);
    orc_u128 kmask = orc_low_mask(2 * k), pmask = orc_low_mask(2 * (k - 1));
    orc_u128 *val = (orc_u128 *)malloc((n ? n : 1) * sizeof(orc_u128));
    if (!val) return ORC_E_NOMEM;
    for (uint64_t i = 0; i < n; i++) val[i] = orc_kmer_value(packed + i * w, w) & kmask;

    table_t dup, t;
    int rc = table_init(&dup, n);
    if (rc) { free(val); return rc; }
    for (uint64_t i = 0; i < n; i++) {
        slot_t *s = table_get(&dup, val[i], 1);
        if (s->n_pref++) { free(dup.s); free(val); return ORC_E_DUP; }
    }
    free(dup.s);

    rc = table_init(&t, 2 * n);
    if (rc) { free(val); return rc; }
    for (uint64_t i = 0; i < n; i++) {
        slot_t *p = table_get(&t, val[i] & pmask, 1);
        p->n_pref++;
        p->owner = i;
        slot_t *s = table_get(&t, val[i] >> 2, 1);
        s->n_suf++;
    }
    for (uint64_t u = 0; u < n; u++) {
        succ[u] = UINT64_MAX;
        slot_t *p = table_get(&t, val[u] >> 2, 0);
        if (!p || p->n_pref != 1) continue;             /* out-degree of u != 1 */
        uint64_t v = p->owner;
        slot_t *q = table_get(&t, val[v] & pmask, 0);
        if (q->n_suf != 1) continue;                      /* in-degree of v != 1 */
        succ[u] = v;
    }
    free(t.s);
    free(val);
    return ORC_OK;
}

int orc_textbook_succ(const uint32_t *packed, uint64_t n, uint32_t k, uint64_t *succ)
{
    return build_succ(packed, n, k, succ);
}

/* lexicographic key: character 0 most significant */
static orc_u128 lex_key(const uint32_t *kmer, uint32_t k)
{
    orc_u128 v = orc_kmer_value(kmer, orc_words_per_kmer(k)), r = 0;
    for (uint32_t i = 0; i < k; i++) r = (r << 2) | orc_char_at(v, i);
    return r;
}

typedef struct {
    char *buf;
    uint64_t len, cap;
} sbuf_t;

static int sb_put(sbuf_t *b, char c)
{
    if (b->len == b->cap) {
        uint64_t nc = b->cap ? 2 * b->cap : 4096;
        char *nb = (char *)realloc(b->buf, nc);
        if (!nb) return ORC_E_NOMEM;
        b->buf = nb;
        b->cap = nc;
    }
    b->buf[b->len++] = c;
    return ORC_OK;
}

static const char ACGT[4] = {'A', 'C', 'G', 'T'};

static int spell(sbuf_t *b, const uint32_t *packed, uint32_t w, uint32_t k, uint64_t start, const uint64_t *succ,
                 uint8_t *seen, int cycle)
{
    uint64_t v = start;
    int first = 1;
    while (v != UINT64_MAX && !(cycle && !first && v == start)) {
        orc_u128 x = orc_kmer_value(packed + v * w, w);
        if (first) {
            for (uint32_t i = 0; i < k; i++)
                if (sb_put(b, ACGT[orc_char_at(x, i)])) return ORC_E_NOMEM;
        } else if (sb_put(b, ACGT[orc_char_at(x, k - 1)])) {
            return ORC_E_NOMEM;
        }
        seen[v] = 1;
        first = 0;
        v = succ[v];
    }
    return sb_put(b, '\n');
}

int orc_textbook(const uint32_t *packed, uint64_t n, uint32_t k, char **text, uint64_t *text_len,
                 uint64_t *n_unitigs, uint64_t *n_cycles)
{
    uint32_t w = orc_words_per_kmer(k);
    uint64_t *succ = (uint64_t *)malloc((n ? n : 1) * sizeof(uint64_t));
    uint8_t *has_pred = (uint8_t *)calloc(n ? n : 1, 1);
    uint8_t *seen = (uint8_t *)calloc(n ? n : 1, 1);
    sbuf_t b = {0, 0, 0};
    uint64_t units = 0, cycles = 0;
    int rc = (succ && has_pred && seen) ? build_succ(packed, n, k, succ) : ORC_E_NOMEM;
    if (rc) goto done;
    for (uint64_t u = 0; u < n; u++)
        if (succ[u] != UINT64_MAX) has_pred[succ[u]] = 1;
    for (uint64_t u = 0; u < n && !rc; u++) {
        if (has_pred[u]) continue;
        rc = spell(&b, packed, w, k, u, succ, seen, 0);
        units++;
    }
    for (uint64_t u = 0; u < n && !rc; u++) {
        if (seen[u]) continue;
        /* isolated cycle through u: start at its smallest k-mer */
        uint64_t best = u;
        orc_u128 best_key = lex_key(packed + u * w, k);
        for (uint64_t v = succ[u]; v != u; v = succ[v]) {
            orc_u128 kv = lex_key(packed + v * w, k);
            if (kv < best_key) { best_key = kv; best = v; }
        }
        rc = spell(&b, packed, w, k, best, succ, seen, 1);
        units++;
        cycles++;
    }
done:
    free(succ);
    free(has_pred);
    free(seen);
    if (rc) { free(b.buf); return rc; }
    if (!b.buf) b.buf = (char *)malloc(1);
    *text = b.buf;
    *text_len = b.len;
    if (n_unitigs) *n_unitigs = units;
    if (n_cycles) *n_cycles = cycles;
    return ORC_OK;
}
