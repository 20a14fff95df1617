// uc_main.cpp -- command-line drop-in for the reference's `main`, `large` and
// `huge` binaries (driver.cpp:422-662):
//
//     uc_main <k-mer file> k l <output file>
//
// Same arguments, same input format (count line, then one k-mer per line;
// the first k characters of each line are used), same output (one unitig per
// line, appended to the output file), same usage and I/O failures
// (driver.cpp:424-428, :262-265; compact.h:12-15).  Phase times go to stdout.
// The text is parsed on the device (uc_load_file), a packed binary k-mer
// file (uc_write_packed_file) is streamed in as it is, and the unitigs are
// appended through pinned chunks (uc_text_append_file).
// Build options the reference fixes at compile time come from the
// environment: UC_BUCKETS (64), UC_MODE (batched|plain), UC_SEED, UC_DEVICE.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/uc_api.h"

static uint64_t env_u64(const char *name, uint64_t dflt)
{
    const char *v = std::getenv(name);
    return v && *v ? std::strtoull(v, nullptr, 0) : dflt;
}

int main(int argc, char **argv)
{
    if (argc != 5) {
        printf("You input %d args\n", argc);
        fprintf(stderr, "Usage: %s <k-mer file> k l <output file>\n", argv[0]);
        return EXIT_FAILURE;
    }
    const char *in_path = argv[1];
    const uint32_t k = (uint32_t)atoi(argv[2]);
    const uint32_t l = (uint32_t)atoi(argv[3]);
    const char *out_path = argv[4];
    auto t0 = std::chrono::steady_clock::now();
    auto secs = [](std::chrono::steady_clock::time_point a) {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - a).count();
    };

    // the reference's open check comes first (driver.cpp:262-265)
    if (FILE *f = std::fopen(in_path, "rb")) {
        std::fclose(f);
    } else {
        printf("Error: Could not open input file. Exiting ...\n");
        exit(1);
    }

    uc_opts o;
    uc_default_opts(&o);
    o.n_buckets = (uint32_t)env_u64("UC_BUCKETS", 64);
    o.seed = env_u64("UC_SEED", o.seed);
    o.device = (int)env_u64("UC_DEVICE", 0);
    const char *mode = std::getenv("UC_MODE");
    if (mode && std::strcmp(mode, "plain") == 0) o.mode = UC_MODE_PLAIN;
    uc_ctx *ctx = nullptr;
    int rc = uc_ctx_create(&o, &ctx);
    if (rc != UC_OK) {
        fprintf(stderr, "uc_main: cannot create context: %s\n", uc_strerror(rc));
        return EXIT_FAILURE;
    }
    // a packed binary k-mer file streams straight into HBM; text goes to HBM
    // through pinned chunks and is parsed and packed on the device
    uc_packed_header ph;
    if (uc_read_packed_header(in_path, &ph) == UC_OK) {
        if (ph.k != k) {
            fprintf(stderr, "uc_main: %s holds %u-mers, not %u-mers\n", in_path, ph.k, k);
            uc_ctx_destroy(ctx);
            return EXIT_FAILURE;
        }
        uint32_t fk = 0;
        rc = uc_load_packed_file(ctx, in_path, &fk);
    } else {
        rc = uc_load_file(ctx, in_path, k);
    }
    if (rc == UC_E_IO) {
        printf("Error: Could not open input file. Exiting ...\n");
        exit(1);
    }
    uint64_t n = 0;
    if (rc == UC_OK) rc = uc_loaded_copy(ctx, nullptr, 0, &n);
    if (rc != UC_OK) {
        fprintf(stderr, "uc_main: %s reading %s: %s\n", uc_strerror(rc), in_path, uc_last_error(ctx));
        uc_ctx_destroy(ctx);
        return EXIT_FAILURE;
    }
    printf("Read %llu k-mers in %.3fs\n", (unsigned long long)n, secs(t0));
    rc = uc_compact_loaded(ctx, l);
    if (rc != UC_OK) {
        fprintf(stderr, "uc_main: %s: %s\n", uc_strerror(rc), uc_last_error(ctx));
        uc_ctx_destroy(ctx);
        return EXIT_FAILURE;
    }
    uc_timing t;
    uc_timing_get(ctx, &t);
    uc_coarsen_info ci;
    uc_coarsen_info_get(ctx, &ci);
    printf("Bucketed by l-minimizers in %.3fs\n", t.bucketize_ms / 1e3);
    printf("Found edges in %.3fs\n", t.edges_ms / 1e3);
    printf("Contracted paths in %.3fs\n", t.coarsen_ms / 1e3);
    printf("Spelled unitigs in %.3fs\n", t.emit_ms / 1e3);
    printf("Unitigs: %llu (isolated cycles: %llu), output bytes: %llu\n", (unsigned long long)ci.n_unitigs,
           (unsigned long long)ci.n_cycles, (unsigned long long)ci.text_bytes);
    auto tw = std::chrono::steady_clock::now();
    rc = uc_text_append_file(ctx, out_path);
    if (rc == UC_E_IO) {
        printf("Error: Could not open input file. Exiting ...\n");
        exit(1);
    }
    if (rc != UC_OK) {
        fprintf(stderr, "uc_main: %s: %s\n", uc_strerror(rc), uc_last_error(ctx));
        uc_ctx_destroy(ctx);
        return EXIT_FAILURE;
    }
    printf("Wrote output in %.3fs\n", secs(tw));
    printf("Total time: %.3fs\n", secs(t0));
    uc_ctx_destroy(ctx);
    return 0;
}
