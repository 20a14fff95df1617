/*
 * common.c -- shared helpers of the CPU oracle (test infrastructure only).
 */
#include <stdlib.h>
#include <string.h>

#include "uc_oracle.h"
#include "oracle_internal.h"

uint32_t orc_words_per_kmer(uint32_t k) { return (k + 15) / 16; }

void orc_free(void *p) { free(p); }

/* driver.cpp:228-232: A,C,G,T -> 0..3, every other byte -> 0. */
static uint32_t char_code(unsigned char c)
{
    switch (c) {
    case 'C': return 1;
    case 'G': return 2;
    case 'T': return 3;
    default: return 0;
    }
}

/* driver.cpp:233-246: character x of k-mer j goes to bits 2*(x & 15) of word
 * x/16; the padding up to a multiple of 16 characters stays zero. */
int orc_pack(const char *chars, uint64_t n, uint32_t k, uint32_t *packed)
{
    if (k < 2 || k > 64) return ORC_E_ARG;
    uint32_t w = orc_words_per_kmer(k);
    for (uint64_t j = 0; j < n; j++) {
        uint32_t *dst = packed + j * w;
        memset(dst, 0, w * sizeof(uint32_t));
        for (uint32_t x = 0; x < k; x++)
            dst[x / 16] |= char_code((unsigned char)chars[j * k + x]) << (2 * (x & 15));
    }
    return ORC_OK;
}

/* The k-mer text rule of the library's parsers (uc_parse_kmer_text,
 * uc_load_text; the reference's load_kmers,
 * driver.cpp:251-354, on every file where that loop ends; see
 * orc_ref_load_kmers for the loop itself): the first line holds the count
 * (sscanf "%u"/"%lu"), which the reference checks only after whole read
 * chunks (driver.cpp:343), so every later line contributes its first k bytes
 * whatever the count says; the rest of a line (a KMC count, say) is skipped;
 * newlines and carriage returns at the end of the text end it; any other
 * line shorter than k (after one '\r' before its end) is a format error. */
int orc_parse_text(const char *buf, size_t len, uint32_t k, uint32_t **packed, uint64_t *n_out)
{
    if (k < 2 || k > 64) return ORC_E_ARG;
    size_t p = 0;
    int have_digit = 0;
    while (p < len && (buf[p] == ' ' || buf[p] == '\t')) p++;
    while (p < len && buf[p] >= '0' && buf[p] <= '9') {
        have_digit = 1;
        p++;
    }
    if (!have_digit) return ORC_E_FORMAT;
    while (p < len && buf[p] != '\n') p++;
    if (p < len) p++;
    size_t end = len;
    while (end > p && (buf[end - 1] == '\n' || buf[end - 1] == '\r')) end--;
    uint64_t lines = 0;
    for (size_t q = p; q < end; q++) lines += buf[q] == '\n';
    if (end > p) lines++;
    uint32_t w = orc_words_per_kmer(k);
    uint32_t *out = (uint32_t *)calloc(lines ? lines * w : 1, sizeof(uint32_t));
    if (!out) return ORC_E_NOMEM;
    uint64_t got = 0;
    while (got < lines) {
        size_t e = p;
        while (e < end && buf[e] != '\n') e++;
        size_t line_len = e - p;
        if (line_len > 0 && buf[p + line_len - 1] == '\r') line_len--;
        if (line_len < k) {
            free(out);
            return ORC_E_FORMAT;
        }
        uint32_t *dst = out + got * w;
        for (uint32_t x = 0; x < k; x++)
            dst[x / 16] |= char_code((unsigned char)buf[p + x]) << (2 * (x & 15));
        got++;
        p = e + 1;
    }
    *packed = out;
    *n_out = got;
    return ORC_OK;
}

/* driver.cpp:257-354 load_kmers itself, with its 64-chunk windowing:
 * chunk_size = sz / 64 (:268); each pass reads the chunk at `offset`,
 * shortened at the end of the file (:283-289); the count comes from the
 * chunk at offset 0 (:291-300); inside a chunk every '\n' is followed by a
 * k-mer when its k bytes fit in the chunk (:306-322, no check of the line's
 * length); `offset` moves to the end of the last k-mer taken (:330); after
 * a chunk the loop stops when total_read == n (:343), else at the end of the
 * file.  A pass that takes no k-mer leaves `offset` where it was, so the
 * reference would repeat it forever: ORC_E_HANG (a file ending in a newline
 * whose count differs from its lines, or a file under about 64 k-mer lines). */
int orc_ref_load_kmers(const char *buf, size_t sz, uint32_t k, uint32_t **packed, uint64_t *n_out)
{
    if (k < 2 || k > 64) return ORC_E_ARG;
    const size_t chunk_size = sz / 64;
    size_t offset = 0, cap = 1024, used = 0;
    uint64_t n = 0, total_read = 0;
    char *chars = (char *)malloc(cap);
    if (!chars) return ORC_E_NOMEM;
    int rc = ORC_OK;
    while (offset < sz) {
        size_t cs = chunk_size;
        if (offset + cs > sz) cs = sz - offset;
        const char *s = buf + offset, *f = s;
        if (offset == 0) {
            /* sscanf("%u"): leading white space, then digits */
            size_t q = 0;
            uint64_t v = 0;
            while (q < sz && (s[q] == ' ' || s[q] == '\t' || s[q] == '\n' || s[q] == '\r')) q++;
            while (q < sz && s[q] >= '0' && s[q] <= '9') v = v * 10 + (uint64_t)(s[q++] - '0');
            n = v;
        }
        size_t last_read = 0;
        while ((size_t)(f - s) < cs) {
            while ((size_t)(f - s) < cs && *f != '\n') f++;
            f++;
            if ((size_t)(f + k - s) > cs) break;
            last_read = (size_t)(f + k - s);
            if (used + k > cap) {
                cap = 2 * cap + k;
                char *grown = (char *)realloc(chars, cap);
                if (!grown) {
                    free(chars);
                    return ORC_E_NOMEM;
                }
                chars = grown;
            }
            memcpy(chars + used, f, k);
            used += k;
            total_read++;
            f += k;
        }
        offset += last_read;
        if (total_read == n) break;
        if (last_read == 0) {
            rc = ORC_E_HANG;
            break;
        }
    }
    if (rc != ORC_OK) {
        free(chars);
        return rc;
    }
    uint32_t *out = (uint32_t *)calloc(total_read ? total_read * orc_words_per_kmer(k) : 1, sizeof(uint32_t));
    if (!out) {
        free(chars);
        return ORC_E_NOMEM;
    }
    orc_pack(chars, total_read, k, out);
    free(chars);
    *packed = out;
    *n_out = total_read;
    return ORC_OK;
}

/* Seeded bijection on 2l-bit codes: add the seed's low word, multiply by an
 * odd constant, xorshift right, multiply by a second odd constant, xorshift
 * right, every step modulo 2^(2l).  Every step is invertible on [0, 4^l), so
 * the composition is a permutation.  The constants are odd and below 2^24 so
 * a GPU may use 24-bit multiplies when 2l <= 24 and get the same residues. */
uint32_t orc_lmer_order(uint32_t code, uint32_t l, uint64_t seed)
{
    uint32_t bits = 2 * l;
    uint32_t mask = bits >= 32 ? 0xFFFFFFFFu : ((1u << bits) - 1u);
    uint32_t h1 = (bits + 1) / 2, h2 = bits / 3 + 1;
    uint32_t x = ((code + ((uint32_t)seed & mask)) * ORC_MIX_C1) & mask;
    x ^= x >> h1;
    x = (x * ORC_MIX_C2) & mask;
    x ^= x >> h2;
    return x;
}
