/*
 * uc_oracle.h -- CPU oracle for the unitig compaction path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library, its
 * C ABI, the CLI) links, loads or calls this code.  Only tests/, the smoke
 * check in __graft_entry__.py and the cpu_baseline leg of bench.py use it,
 * and only as the checker / the timed CPU baseline.
 *
 * Two independent restatements live here:
 *
 *  1. textbook.c  -- the closed form.  Maximal non-branching paths of the
 *     directed node-centric de Bruijn graph of the k-mer set, with the edge
 *     rule of assembler.h:120-197 (u->v iff suffix(u)==prefix(v), u has
 *     exactly one successor and v exactly one predecessor).  Serial hash map
 *     + walk from heads; cycles start at their lexicographically smallest
 *     k-mer.  Works for any 2 <= k <= 64.
 *
 *  2. refpipe.c   -- a serial restatement of the reference pipeline, phase by
 *     phase (text load, 2-bit packing, l-mer order permutation, double
 *     L-minimizers, 64-style bucketing with lend/borrow duplicates, FNV
 *     hashmap edge finding, HEC coarsening levels with crossing tracking,
 *     cross graph, final coarsening, output partitioning, minitig stitching).
 *     Each function cites the reference file:line it follows.  Like the
 *     reference it is defined only for k = 15 (mod 16).
 *
 * K-mers are passed in the reference's packing (driver.cpp:233-246):
 * ceil(k/16) uint32 words per k-mer, character i stored in bits
 * 2*(i mod 16) of word i/16, A=0 C=1 G=2 T=3, unused high bits zero.
 *
 * Status codes: 0 ok, negative on error (see ORC_E_*).
 */
#ifndef UC_ORACLE_H
#define UC_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_OK 0
#define ORC_E_ARG (-1)      /* bad argument (k, l, n_buckets ...) */
#define ORC_E_NOMEM (-2)    /* allocation failed */
#define ORC_E_FORMAT (-3)   /* malformed text input */
#define ORC_E_SELFLOOP (-4) /* input holds a self-looping k-mer: the reference's HEC never terminates */
#define ORC_E_DUP (-5)      /* input holds the same k-mer twice */
#define ORC_E_HANG (-6)     /* the reference's read loop would never end on this text */

/* ---- shared helpers --------------------------------------------------- */

/* words per k-mer in the reference packing */
uint32_t orc_words_per_kmer(uint32_t k);

/* Text input in the reference format (driver.cpp:251-354): line 1 holds the
 * k-mer count, each later line yields its first k characters.  Bytes other
 * than A/C/G/T map to 0 ('A') like the zero-initialised 256-entry table of
 * driver.cpp:228-232.  Every line is read whatever the count says (the
 * reference only compares the count after whole read chunks); newlines and
 * carriage returns at the end of the text end it.  Other lines shorter than
 * k characters are rejected (ORC_E_FORMAT).  *packed is malloc'd. */
int orc_parse_text(const char *buf, size_t len, uint32_t k, uint32_t **packed, uint64_t *n);

/* driver.cpp:257-354 load_kmers restated with its 64 read chunks, the count
 * test after each chunk and no line-length check: the k-mers the reference
 * reads from this text, or ORC_E_HANG where its loop would not end. */
int orc_ref_load_kmers(const char *buf, size_t sz, uint32_t k, uint32_t **packed, uint64_t *n);

/* driver.cpp:227-249 compact_kmers: n*k ASCII characters -> packed words. */
int orc_pack(const char *chars, uint64_t n, uint32_t k, uint32_t *packed);

/* Seeded bijection of l-mer codes (0 <= code < 4^l) onto orders.  The same
 * function is computed inside the HIP bucketing kernel, so CPU and GPU
 * bucket identically for the same seed. */
uint32_t orc_lmer_order(uint32_t code, uint32_t l, uint64_t seed);

/* ---- 1. textbook closed form ---------------------------------------- */

/* Unitig text (one line per unitig, "ACGT...\n") of the k-mer set.
 * *text is malloc'd (free with orc_free).  Duplicated k-mers -> ORC_E_DUP.
 * *n_cycles counts unitigs that are isolated cycles. */
int orc_textbook(const uint32_t *packed, uint64_t n, uint32_t k, char **text, uint64_t *text_len,
                 uint64_t *n_unitigs, uint64_t *n_cycles);

/* Degree summary used by tests: out[i] = successor index of k-mer i in the
 * pruned graph or UINT64_MAX. */
int orc_textbook_succ(const uint32_t *packed, uint64_t n, uint32_t k, uint64_t *succ);

/* ---- 2. restatement of the reference pipeline ------------------------ */

/* driver.cpp:356-402 generate_permutation: random keys, insertion into 2*4^l
 * slots with linear probing, compaction scan.  perm has 4^l entries. */
int orc_generate_permutation(uint32_t l, uint64_t seed, uint32_t *perm);

/* lmin_buckets.h:67-94 get_double_lmin on one k-mer. */
void orc_double_lmin(const uint32_t *kmer, uint32_t k, uint32_t l, const uint32_t *order,
                     uint32_t *pref_min, uint32_t *suf_min);

/* Bucketing summary (lmin_buckets.h:288-433 + driver.cpp:137-225).
 * bucket_rows: n_buckets+1 offsets; kmer_src: input index of every
 * placement in bucket order (m entries); cross_rows: n_buckets^2+1 offsets
 * of the cross lists; cross_lend / cross_borrow: bucket-local ids or
 * UINT32_MAX.  All arrays malloc'd and owned by the struct. */
typedef struct {
    uint32_t n_buckets;
    uint64_t m;
    uint64_t *bucket_rows;
    uint64_t *kmer_src;
    uint64_t *cross_rows;
    uint32_t *cross_lend;
    uint32_t *cross_borrow;
} orc_buckets;

int orc_bucketize(const uint32_t *packed, uint64_t n, uint32_t k, uint32_t l, uint32_t n_buckets,
                  const uint32_t *order, orc_buckets *out);
void orc_buckets_free(orc_buckets *b);

/* Per-phase wall times (seconds) of the last orc_refpipe call, at the
 * reference's own phase boundaries (driver.cpp:443,506,510,514,519,525,611). */
typedef struct {
    double bucketing, assemble, coarsen, write_intra, minitigs, cross_finish, total;
    uint64_t n_kmers, m_placements, n_cross, cross_offset, n_lines;
    /* orc_refpipe_par: the per-bucket loop's wall time (assemble .. minitigs
     * are then thread-seconds summed over the threads) and the thread count */
    double buckets_wall;
    uint64_t threads;
} orc_refpipe_stats;

/* Whole reference pipeline on packed k-mers: returns the unitig text in the
 * order the serial restatement writes it (intra-bucket unitigs bucket by
 * bucket, then the stitched cross-bucket unitigs partition by partition).
 * order: 4^l l-mer orders (e.g. from orc_generate_permutation). */
int orc_refpipe(const uint32_t *packed, uint64_t n, uint32_t k, uint32_t l, uint32_t n_buckets,
                const uint32_t *order, char **text, uint64_t *text_len, orc_refpipe_stats *stats);

/* The same pipeline on nthreads OpenMP threads: minimizers in parallel,
 * whole buckets per thread; the same text bytes as orc_refpipe. */
int orc_refpipe_par(const uint32_t *packed, uint64_t n, uint32_t k, uint32_t l, uint32_t n_buckets,
                    const uint32_t *order, int nthreads, char **text, uint64_t *text_len, orc_refpipe_stats *stats);

/* Restated pieces exposed for unit tests (assembler.h:7-45). */
uint32_t orc_fnv_pref(const uint32_t *kmer, uint32_t words);
uint32_t orc_fnv_suf(const uint32_t *kmer, uint32_t words);

void orc_free(void *p);

#ifdef __cplusplus
}
#endif
#endif
