/*
 * uc_synth.h -- seeded synthetic k-mer sets (workload generator for tests
 * and benchmarks; host code, no GPU).  Not part of the compaction path and
 * with no counterpart in the reference, whose inputs come from a k-mer
 * counter such as KMC (README.md:3-4).
 */
#ifndef UC_SYNTH_H
#define UC_SYNTH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    uint64_t seed;
    uint64_t genome_len;  /* bases of the uniform random genome */
    uint32_t k;           /* 2..64 */
    double repeat_frac;   /* fraction of bases overwritten by repeat copies (0 = none) */
    uint64_t seg_min;     /* repeat segment length range */
    uint64_t seg_max;
    double repeat_sub;    /* per-base substitution rate inside a copy */
    double error_frac;    /* fraction of true k-mers given one substituted extra copy */
    /* repeat families (copies_max > 0): each family copies one source
     * segment c times, c log-uniform in [copies_min, copies_max], with a
     * per-family substitution rate uniform in [div_min, div_max] (replaces
     * repeat_sub); copies_max = 0 keeps one copy per segment */
    uint32_t copies_min, copies_max;
    double div_min, div_max;
} uc_synth_params;

/* Distinct canonical k-mers, lexicographically sorted, reference packing
 * (ceil(k/16) uint32 words per k-mer).  *packed is malloc'd; free with
 * uc_synth_free.  Returns 0 or a negative error. */
int uc_synth_kmers(const uc_synth_params *p, uint32_t **packed, uint64_t *n);
void uc_synth_free(void *ptr);

/* Reference text format: "<n>\n" then one k-mer per line.  With dst NULL
 * only *len is set. */
int uc_synth_format_text(const uint32_t *packed, uint64_t n, uint32_t k, char *dst, uint64_t cap, uint64_t *len);

#ifdef __cplusplus
}
#endif
#endif
