/*
 * uc_api.h -- C ABI of the MI355X unitig compactor (libuc.so).
 *
 * The reference (euphoricpoptarts/UnitigCoarsening) has no FFI: its
 * boundary is the `main <k-mer file> k l <output file>` process
 * (driver.cpp:422-432) and, inside it, free functions over Kokkos views.
 * Each entry point below names the reference functions it replaces; the
 * CLI in csrc/uc_main.cpp keeps the process boundary itself.
 *
 * Conventions
 *  - Every function returns an int status: UC_OK (0) or a negative UC_E_*
 *    code (the "nonzero = error" convention of CHECK_RETSTAT,
 *    driver.cpp:15-23).  Nothing here calls exit().  uc_last_error() gives
 *    a message for the last failure on a context.
 *  - K-mers use the reference packing (driver.cpp:233-246): ceil(k/16)
 *    uint32 words per k-mer, character i in bits 2*(i mod 16) of word i/16,
 *    A=0 C=1 G=2 T=3, bits above the k-th character zero.
 *  - One host thread per context.  A context owns its device buffers; a
 *    result stays valid until the next job on the same context.
 *  - Slot ids are 32-bit: a job may hold at most 2^32-2 k-mer placements
 *    (k-mers plus crossing duplicates).
 */
#ifndef UC_API_H
#define UC_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define UC_OK 0
#define UC_E_ARG (-1)      /* invalid argument */
#define UC_E_NOMEM (-2)    /* device or host allocation failed / over budget */
#define UC_E_HIP (-3)      /* HIP runtime error */
#define UC_E_STATE (-4)    /* call out of order (e.g. uc_find_edges before uc_bucketize) */
#define UC_E_IO (-5)       /* file could not be opened / read / written */
#define UC_E_FORMAT (-6)   /* malformed k-mer text */
#define UC_E_LIMIT (-7)    /* job exceeds a compiled limit (32-bit slot ids, bucket count) */

#define UC_ALL_BUCKETS 0xFFFFFFFFu

/* bucket modes of the edge-finding phase */
#define UC_MODE_BATCHED 0 /* one launch per phase over every bucket (and every set) */
#define UC_MODE_PLAIN 1   /* one launch per bucket, one shared table sized for the largest
                             bucket, like the reference's per-bucket loop (driver.cpp:479-521) */

typedef struct uc_ctx uc_ctx;

typedef struct {
    int device;            /* HIP device ordinal */
    uint64_t seed;         /* seed of the l-mer order bijection and of the splitter sample */
    uint32_t n_buckets;    /* minimizer buckets per set, power of two, 1..1024
                              (the reference hard-codes 64, lmin_buckets.h:35) */
    uint32_t mode;         /* UC_MODE_BATCHED or UC_MODE_PLAIN */
    uint32_t load_pct;     /* hashmap load factor in percent (10..90; the reference uses <= 50,
                              assembler.h:113-118); 0 = default 50 */
    uint32_t sample_log2;  /* splitter sampling rate 2^-sample_log2 of the run heads; 0 = default 9 */
    uint64_t hbm_budget;   /* bytes this context may allocate on the device; 0 = no limit */
    uint32_t flags;        /* UC_FLAG_* */
    uint32_t fine_slots;   /* target slots per fine bucket (64..4096); 0 = default (512 for k <= 32, else 384; a
                              set of the batched form that wants more than 1024 fine buckets per coarse bucket
                              gets the fewest that average at most 1700 slots) */
    uint32_t group_slots;  /* most slots one LDS edge-finding workgroup takes (256..2048, power of two);
                              0 = default 2048 */
} uc_opts;

#define UC_FLAG_PROFILE 1u /* time every kernel launch with HIP events (uc_kernel_stat) */
#define UC_FLAG_XCD_MAP 2u /* give each XCD a contiguous range of bucket tiles */
#define UC_FLAG_KEEP_SRC 4u /* keep slot -> input k-mer ids for uc_bucket_copy */
#define UC_FLAG_DIRECT_SCATTER 8u /* both bucketing scatters write straight from registers instead of
                                     staging each tile in LDS sorted by bucket */
#define UC_FLAG_EAGER 16u /* launch the kernels of a contraction pass one by one instead of replaying
                             them as a captured HIP graph, and read its counters before emission (small
                             jobs otherwise read them once, after emission) */
#define UC_FLAG_COUNT_PROBES 32u /* count the hash-table probes of the edge phase (uc_edge_probes) */
#define UC_FLAG_LOW_MEMORY 64u /* free every buffer at the phase boundary after its last use, the loaded input
                                  (uc_load*) after the first scatter included: the device holds about half the
                                  bytes, each job pays a few hipFree/hipMalloc, the input is consumed (load again
                                  before the next uc_compact_loaded), and of the intermediate results only the
                                  text and the counts (uc_*_info_get) stay readable after a job */

void uc_default_opts(uc_opts *o);
int uc_ctx_create(const uc_opts *opts, uc_ctx **out);
int uc_ctx_destroy(uc_ctx *ctx);
const char *uc_last_error(const uc_ctx *ctx);
const char *uc_strerror(int status);

/* ---- whole job -------------------------------------------------------- */

/* Replaces driver.cpp:433-662 from the device copy of the packed k-mers to
 * the unitig text: bucketing, edge finding, contraction, emission.
 * `packed` holds n k-mers in the reference packing; `on_device` says
 * whether it points to device memory (no copy) or host memory (one H2D
 * copy).  The unitig text stays on the device (uc_text_*). */
int uc_compact(uc_ctx *ctx, const uint32_t *packed, uint64_t n, uint32_t k, uint32_t l, int on_device);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (tight serving loops, harnesses that forbid hipMalloc): uc_compact
 * on n k-mers already in device memory (`packed`, reference packing), with
 * every intermediate buffer carved from a static scratch array of
 * UC_SCRATCH_BYTES that the library holds on each device until the process
 * ends, and the unitig text written to the caller's device buffer `text`
 * (text_cap bytes).  Nothing is allocated: no hipMalloc, hipFree, graph
 * instantiation or event creation (profiling and probe counting are off for
 * the call; the contraction pass runs eagerly).
 *  - *text_size gets the text's size once contraction knows it, also when
 *    the call then fails with UC_E_NOMEM because text_cap is smaller.
 *  - A job too large for the scratch fails with UC_E_LIMIT: at once when n
 *    alone rules it out, else when the first buffer that does not fit is
 *    carved (before any kernel writes to it).  About 1.5 M 31-mers fit.
 *  - Every call on one device shares that device's scratch, which keeps the
 *    last call's contents: calls that overlap on one device (two contexts,
 *    two threads) race and must be serialised by the caller.
 *  - The context's own buffers and results are left as they were, except
 *    that nothing of the last job stays readable (uc_text_size is 0). */
#define UC_SCRATCH_BYTES (256ull << 20)
int uc_compact_noalloc(uc_ctx *ctx, const uint32_t *packed, uint64_t n, uint32_t k, uint32_t l, char *text,
                       uint64_t text_cap, uint64_t *text_size);

/* Whole job on the k-mers of the last uc_load (already resident in device
 * memory): bucketing, edges, contraction, emission, timed per phase. */
int uc_compact_loaded(uc_ctx *ctx, uint32_t l);

/* Batched form: n_sets independent k-mer sets concatenated in `packed`
 * (set s holds k-mers [set_offsets[s], set_offsets[s+1])); every phase
 * runs once over all (set, bucket) pairs.  Set s's unitigs are text bytes
 * [uc_set_text_offset(s), uc_set_text_offset(s+1)). */
int uc_compact_sets(uc_ctx *ctx, const uint32_t *packed, const uint64_t *set_offsets, uint32_t n_sets, uint32_t k,
                    uint32_t l, int on_device);

/* ---- phases (the drop-in pieces) -------------------------------------- */

/* Replaces driver.cpp:356-402 generate_permutation: writes the seeded
 * bijection of l-mer codes onto orders for every code < 4^l into an array
 * of 4^l uint32 (the table the bucketing kernel otherwise computes inline),
 * in device memory (on_device != 0) or host memory. */
int uc_lmer_order_table(uc_ctx *ctx, uint32_t l, uint32_t *table, int on_device);

/* Load packed k-mers (one set or several) for the phase-by-phase calls. */
int uc_load(uc_ctx *ctx, const uint32_t *packed, const uint64_t *set_offsets, uint32_t n_sets, uint32_t k,
            int on_device);

/* Replaces lmin_buckets.h:288-433 find_l_minimizer + driver.cpp:137-225
 * collect_buckets: prefix/suffix l-minimizers, bucket CSR with crossing
 * k-mers duplicated into their suffix bucket, lend/borrow pairs.
 * order_table: NULL for the inline seeded bijection, else a table of 4^l
 * orders (e.g. from uc_lmer_order_table, or any permutation such as the
 * reference's lmin_bucket_map) in device memory (table_on_device != 0) or
 * host memory (copied once). */
int uc_bucketize(uc_ctx *ctx, uint32_t l, const uint32_t *order_table, int table_on_device);

typedef struct {
    uint64_t n_kmers;        /* k-mers over all sets */
    uint64_t m_placements;   /* slots of the bucket CSR: k-mers + crossing duplicates */
    uint64_t n_cross;        /* crossing k-mers (lend/borrow pairs) */
    uint32_t n_buckets;      /* buckets per set */
    uint32_t n_sets;
    uint64_t largest_bucket;
    uint64_t n_fine_buckets;  /* fine buckets (coarse buckets split by further minimizer-order bits) */
    uint64_t largest_fine_bucket;
    uint32_t bucket_tile;    /* k-mers per bucketing tile */
    uint32_t fine_tile;      /* slots per fine-sort tile (0: no fine sort) */
    uint32_t lds_limit;      /* LDS bytes a workgroup may use on the device */
    uint32_t n_oversized;    /* fine buckets above the LDS group size: edge finding uses HBM tables */
    uint64_t oversized_slots; /* slots in those fine buckets */
    /* how the edge phase takes the oversized fine buckets (0 until it was planned):
       alone in a 4096-slot LDS area, split into LDS hash slices, or in HBM tables */
    uint64_t n_big_lds;
    uint64_t n_slice_tasks;
    uint64_t n_hbm_tables;
} uc_bucket_info;

int uc_bucket_info_get(uc_ctx *ctx, uc_bucket_info *info);

/* Host copies of the bucketing result for checking: bucket row offsets
 * (n_sets*n_buckets+1), slot -> input k-mer index (m; needs
 * UC_FLAG_KEEP_SRC), slot flags (m;
 * bit0 = home copy of a crossing k-mer, bit1 = borrowed duplicate) and the
 * pairs (home slot, borrowed slot) (n_cross each).  Any pointer may be NULL. */
int uc_bucket_copy(uc_ctx *ctx, uint64_t *bucket_rows, uint32_t *slot_src, uint8_t *slot_flags, uint32_t *pair_home,
                   uint32_t *pair_borrow);

/* Replaces assembler.h:113-150 init_hashmap/generate_hashmap and
 * :152-214 init_assembler/assemble_pruned_graph for one bucket (global
 * bucket id set*n_buckets+b) or UC_ALL_BUCKETS: open-addressing table of
 * (k-1)-prefixes, suffix lookups, in-degree counts, successor per slot
 * (kept only when out-degree and in-degree are both 1). */
int uc_find_edges(uc_ctx *ctx, uint32_t bucket);

/* Host copy of the successor array (m entries, UINT32_MAX = none), slot ids. */
int uc_succ_copy(uc_ctx *ctx, uint32_t *succ);
/* Hash-table slot inspections (insert and lookup probes, over every table
 * kind) of the last edge phase, with UC_FLAG_COUNT_PROBES (UC_E_STATE
 * without).  No counterpart in the reference: the hashmap probes/s figure
 * of the measurement. */
int uc_edge_probes(uc_ctx *ctx, uint64_t *probes);

/* Replaces coarseners.h:388-485 coarsen_de_bruijn_full_cycle(_final),
 * heuristics.h:205-256 coarsen_HEC, compact_graph.h and the cross-graph
 * build of driver.cpp:527-556: splices every borrowed duplicate back onto
 * its home copy, then contracts the non-branching paths (splitter
 * election, segment walks, segment-chain pointer jumping; isolated cycles
 * are cut at their lexicographically smallest k-mer). */
int uc_coarsen(uc_ctx *ctx);

typedef struct {
    uint64_t n_unitigs;
    uint64_t n_cycles;       /* unitigs that were isolated cycles */
    uint64_t text_bytes;     /* sum over unitigs of (length + k - 1 + 1) */
    uint64_t n_splitters;
    uint32_t chain_rounds;   /* pointer-jumping rounds over segment chains */
    uint64_t n_runs;         /* runs: chains of edges inside one LDS group (every k-mer is in one) */
} uc_coarsen_info;

int uc_coarsen_info_get(uc_ctx *ctx, uc_coarsen_info *info);

/* Replaces compact.h:161-210 write_intra_bucket_outputs,
 * lmin_buckets.h:172-283 generate_minitigs/partition_for_output,
 * driver.cpp:561-639 and compact.h:59-112 write_unitigs4: the unitig text
 * (one line per unitig) written on the device. */
int uc_emit(uc_ctx *ctx);

/* ---- results ---------------------------------------------------------- */

uint64_t uc_text_size(const uc_ctx *ctx);
const char *uc_text_device_ptr(const uc_ctx *ctx);
int uc_text_copy(uc_ctx *ctx, char *dst, uint64_t cap);
uint64_t uc_set_text_offset(const uc_ctx *ctx, uint32_t set);

/* Appends the text to a file (std::ofstream::app semantics, compact.h:7-37).
 * The copy out of HBM runs in 64 MiB pinned chunks, each overlapping the
 * write of the one before. */
int uc_text_append_file(uc_ctx *ctx, const char *path);

/* ---- text input (driver.cpp:251-354) ---------------------------------- */

/* Parses the reference's k-mer text (count line, then first k characters
 * of every later line, whatever the count says; bytes other than ACGT map
 * to A) into the reference packing on the host.  *packed is malloc'd (free
 * with uc_free). */
int uc_parse_kmer_text(const char *buf, uint64_t len, uint32_t k, uint32_t **packed, uint64_t *n);
int uc_read_kmer_file(const char *path, uint32_t k, uint32_t **packed, uint64_t *n);
void uc_free(void *p);

/* Replaces driver.cpp:257-354 load_kmers + driver.cpp:227-249 compact_kmers
 * on the device: the same text format and rules as uc_parse_kmer_text
 * (the count line must hold a number, and every later line is a k-mer
 * whatever it says: the reference compares the count with the k-mers read
 * only after whole read chunks, driver.cpp:310-343, and reads on past it; a
 * line shorter than k after dropping one '\r' before its end is
 * UC_E_FORMAT), parsed and packed by kernels and loaded as
 * one k-mer set, ready for uc_bucketize or uc_compact_loaded.  text is a
 * host buffer, or a device buffer when on_device is nonzero. */
int uc_load_text(uc_ctx *ctx, const char *text, uint64_t len, uint32_t k, int on_device);

/* uc_load_text of a file, read through 64 MiB pinned chunks, each read
 * overlapping the copy of the one before.  UC_E_IO if it cannot be opened
 * (driver.cpp:262-265). */
int uc_load_file(uc_ctx *ctx, const char *path, uint32_t k);

/* ---- packed binary k-mer files -----------------------------------------
 * A 32-byte header, then n * ceil(k/16) uint32 words in the reference
 * packing (the comp_vt buffer of driver.cpp:227-249, little-endian).  At
 * 3.1e9 63-mers this is 50 GB where the text is about 200 GB. */

#define UC_PACKED_MAGIC "UCKMERS1"

typedef struct {
    char magic[8];      /* UC_PACKED_MAGIC, no terminator */
    uint32_t k;
    uint32_t words;     /* ceil(k / 16) */
    uint64_t n;
    uint64_t reserved;  /* 0 */
} uc_packed_header;

int uc_write_packed_file(const char *path, const uint32_t *packed, uint64_t n, uint32_t k);
/* UC_E_IO if it cannot be read, UC_E_FORMAT if the header is not one. */
int uc_read_packed_header(const char *path, uc_packed_header *h);
/* Streams a packed file into HBM through 64 MiB pinned chunks (read of
 * chunk i+1 overlapping the copy of chunk i) and loads it as one k-mer
 * set; *k gets its k. */
int uc_load_packed_file(uc_ctx *ctx, const char *path, uint32_t *k);

/* Device address of the loaded k-mers in the 64-bit layout (ceil(k/32)
 * uint64 words per k-mer, char i at bits 2(i%32) of word i/32; for k in
 * 17-32 and 49-64 these are the bytes of the reference packing), valid until
 * the next load.  NULL when nothing is loaded. */
const uint64_t *uc_loaded_device_ptr(const uc_ctx *ctx);

/* The loaded k-mers in the reference packing (ceil(k/16) uint32 per
 * k-mer); *n gets their count.  packed may be NULL to ask for the count. */
int uc_loaded_copy(uc_ctx *ctx, uint32_t *packed, uint64_t cap, uint64_t *n);

/* ---- timing ----------------------------------------------------------- */

typedef struct {
    float bucketize_ms, edges_ms, coarsen_ms, emit_ms, total_ms;
    uint64_t device_bytes; /* device memory held by the context */
    uint64_t device_peak_bytes; /* most device memory the context held at once during the job */
} uc_timing;

/* Per-phase device times (HIP events on the context's stream) of the last
 * uc_compact / uc_compact_sets. */
int uc_timing_get(uc_ctx *ctx, uc_timing *t);

/* With UC_FLAG_PROFILE: accumulated device time and launch count of kernel
 * number idx (0..) since the last uc_kernel_stats_reset, on the context's
 * stream.  Returns UC_E_ARG past the last kernel. */
int uc_kernel_stat(uc_ctx *ctx, uint32_t idx, char *name, uint32_t name_cap, double *total_ms, uint64_t *launches);
int uc_kernel_stats_reset(uc_ctx *ctx);
/* With UC_FLAG_PROFILE: time only the launches of the named kernel (NULL or
 * "" for every kernel again).  Two events per launch of one kernel disturb a
 * timed run far less than two around every launch. */
int uc_kernel_stats_select(uc_ctx *ctx, const char *name);

/* Blocks until the context's stream is idle. */
int uc_sync(uc_ctx *ctx);

/* No counterpart in the reference.  Most probes of one insert into an
 * edge-finding slice table (fine buckets of 4097..65535 slots are split
 * into hash slices of about 2048 keys, each in an 8192-entry LDS table):
 * 1..8192, default 8192.  An insert that runs out of probes marks the slice
 * as overflowed, and the job's edge phase then runs again with every sliced
 * fine bucket on an HBM table (same result, one more edge phase).  A small
 * bound forces that route, which tests use. */
int uc_ctx_set_slice_probes(uc_ctx *ctx, uint32_t probes);

/* ---- autotuner ---------------------------------------------------------
 * Times whole compactions of the loaded job (uc_load / uc_load_text / ...)
 * under candidate settings, by coordinate descent over n_buckets,
 * fine_slots (bucket-to-workgroup tiling), group_slots (LDS table size per
 * workgroup), sample_log2 (contraction) and load_pct (hashmap load factor
 * of the HBM tables), and keeps the fastest in the context for every later
 * job.  The unitig set never depends on these settings; the line order
 * follows n_buckets and fine_slots.  Each candidate runs `reps` times
 * (0 = 2) after one warm-up; the minimum counts.  A fine_slots of 0 in the
 * result is the library's default choice (uc_opts.fine_slots). */
typedef struct {
    uint32_t n_buckets, fine_slots, group_slots, sample_log2, load_pct;
    uint32_t trials;       /* candidate settings timed */
    float default_ms;      /* the context's settings before tuning */
    float best_ms;
} uc_tune_result;

#define UC_TUNE_KEEP_BUCKETS 1u /* leave n_buckets as it is (a workload that fixes it) */

int uc_autotune(uc_ctx *ctx, uint32_t l, uint32_t reps, uint32_t tune_flags, uc_tune_result *out);

/* Frees every device buffer the context holds (they grow back on demand). */
int uc_ctx_trim(uc_ctx *ctx);

/* Device bytes a context holds right now. */
uint64_t uc_ctx_held_bytes(const uc_ctx *ctx);

/* Upper bound of the device bytes one job of n k-mers needs, whatever its
 * crossing rate (at most 2n placements). */
uint64_t uc_job_bytes_estimate(uint64_t n, uint32_t k);

/* UC_OK when a job of n k-mers, x of them crossing, fits the 32-bit slot ids
 * (n + x <= 2^32 - 2 placements); UC_E_LIMIT when it does not (the job then
 * fails with UC_E_LIMIT after bucket counting, before any slot is written);
 * UC_E_STATE when x > n.  Host only; replaces nothing in the reference,
 * whose `HUGE` build widens its ids instead (defs.h:30-33). */
int uc_check_placements(uint64_t n, uint64_t x);

/* ---- serving --------------------------------------------------------------
 * No counterpart in the reference, which runs one job per process.  A server
 * owns n_slots worker contexts, each with its own HIP stream and host
 * thread.  Jobs are queued in arrival order; a job starts when a worker is
 * free and its estimated footprint fits in hbm_budget together with the
 * memory the other workers hold (idle workers give their buffers back when
 * that is what makes a job fit).  Jobs on different workers run
 * concurrently on the device.  A job larger than the whole budget fails
 * with UC_E_NOMEM. */

typedef struct uc_server uc_server;

typedef struct {
    uc_opts ctx_opts;     /* options of every worker context */
    uint32_t n_slots;     /* concurrent jobs */
    uint64_t hbm_budget;  /* bytes all workers may hold together; 0 = no limit */
} uc_server_opts;

#define UC_JOB_QUEUED 0
#define UC_JOB_RUNNING 1
#define UC_JOB_DONE 2
#define UC_JOB_FAILED 3

int uc_server_create(const uc_server_opts *opts, uc_server **out);
/* Waits for queued and running jobs, then frees everything. */
int uc_server_destroy(uc_server *srv);
/* Copies n packed k-mers (host memory) and queues the job. */
int uc_submit(uc_server *srv, const uint32_t *packed, uint64_t n, uint32_t k, uint32_t l, uint64_t *job);
/* Non-blocking state of a job (UC_JOB_*). */
int uc_poll(uc_server *srv, uint64_t job, int *state);
/* Blocks until the job is done or failed; returns the job's status. */
int uc_wait(uc_server *srv, uint64_t job);
/* Result of a finished job: text size (always set) and text (when dst has
 * room), time queued and time running in milliseconds. */
int uc_result(uc_server *srv, uint64_t job, char *dst, uint64_t cap, uint64_t *size, double *queue_ms,
              double *run_ms);
/* Forgets a finished job and its text. */
int uc_release(uc_server *srv, uint64_t job);
/* Peak device bytes the workers held together (every allocation and free
 * of every worker context is counted as it happens), and jobs completed. */
int uc_server_stats(uc_server *srv, uint64_t *peak_held, uint64_t *jobs_done);

#ifdef __cplusplus
}
#endif
#endif
