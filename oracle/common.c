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

/* driver.cpp:251-354 load_kmers, restated without the 64-chunk windowing
 * (whose tiny-file defect is not reproduced): the first line is the count
 * (sscanf "%u"/"%lu"), each later line contributes its first k bytes, the
 * rest of the line (a KMC count, say) is skipped, reading stops once
 * `count` k-mers were taken or the text ends. */
int orc_parse_text(const char *buf, size_t len, uint32_t k, uint32_t **packed, uint64_t *n_out)
{
    if (k < 2 || k > 64) return ORC_E_ARG;
    size_t p = 0;
    uint64_t count = 0;
    int have_digit = 0;
    while (p < len && (buf[p] == ' ' || buf[p] == '\t')) p++;
    while (p < len && buf[p] >= '0' && buf[p] <= '9') {
        count = count * 10 + (uint64_t)(buf[p] - '0');
        have_digit = 1;
        p++;
    }
    if (!have_digit) return ORC_E_FORMAT;
    while (p < len && buf[p] != '\n') p++;
    if (p < len) p++;
    uint32_t w = orc_words_per_kmer(k);
    uint32_t *out = (uint32_t *)calloc(count ? count * w : 1, sizeof(uint32_t));
    if (!out) return ORC_E_NOMEM;
    uint64_t got = 0;
    while (got < count && p < len) {
        size_t e = p;
        while (e < len && buf[e] != '\n') e++;
        size_t line_len = e - p;
        if (line_len > 0 && buf[p + line_len - 1] == '\r') line_len--;
        if (line_len == 0 && e >= len) break; /* trailing newline at EOF */
        if (line_len < k) {
            free(out);
            return ORC_E_FORMAT;
        }
        uint32_t *dst = out + got * w;
        for (uint32_t x = 0; x < k; x++)
            dst[x / 16] |= char_code((unsigned char)buf[p + x]) << (2 * (x & 15));
        got++;
        p = (e < len) ? e + 1 : e;
    }
    *packed = out;
    *n_out = got;
    return ORC_OK;
}

/* Seeded bijection on 2l-bit codes: xor with the seed, then two rounds of
 * (multiply by an odd constant modulo 2^(2l), xorshift right), then xor with
 * the seed's high word.  Every step is invertible on [0, 4^l), so the
 * composition is a permutation.  The constants are odd and below 2^24 so a
 * GPU may use 24-bit multiplies when 2l <= 24 and get the same residues. */
uint32_t orc_lmer_order(uint32_t code, uint32_t l, uint64_t seed)
{
    uint32_t bits = 2 * l;
    uint32_t mask = bits >= 32 ? 0xFFFFFFFFu : ((1u << bits) - 1u);
    uint32_t h1 = (bits + 1) / 2, h2 = bits / 3 + 1;
    uint32_t x = (code ^ (uint32_t)seed) & mask;
    x = (x * ORC_MIX_C1) & mask;
    x ^= x >> h1;
    x = (x * ORC_MIX_C2) & mask;
    x ^= x >> h2;
    x ^= (uint32_t)(seed >> 32) & mask;
    return x;
}
