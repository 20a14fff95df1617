// uc_kernels.h -- device-side building blocks of the compaction path (gfx950).
//
// Layouts (all in HBM, one context at a time):
//   k-mers      uint64 x W per k-mer, W = 1 for k <= 32, 2 for k <= 64;
//               character i in bits 2i of the 64*W-bit little-endian value
//               (the reference packing read as 64-bit words).
//   bucket CSR  slots [boff[g], boff[g+1]) hold bucket g = set*B + b:
//               every k-mer in its prefix-minimizer bucket, crossing
//               k-mers once more in their suffix-minimizer bucket.
//   tables      per bucket, 2^tbits[g] uint64 slots at toff[g]:
//               bits 0..31 owner (bucket-local slot), 32..59 key tag,
//               60..62 number of suffix hits, 63 prefix shared.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace uc {

constexpr uint32_t NONE = 0xFFFFFFFFu;
constexpr uint64_t EMPTY = ~0ull;
constexpr uint64_t SHARED_BIT = 1ull << 63;
constexpr uint32_t TAG_MASK = 0x0FFFFFFFu;
constexpr int HIT_SHIFT = 60;

// slot flags
constexpr uint8_t F_LENT = 1;    // home copy of a crossing k-mer: its out-edge lives elsewhere
constexpr uint8_t F_BORROW = 2;  // duplicate in the suffix bucket: carries the out-edge, dies at the splice
constexpr uint8_t F_CYCLE = 4;   // on an isolated cycle (set during contraction)
constexpr uint8_t F_VISIT = 8;   // reached from a path head (cycle detection only)
constexpr uint8_t F_RUNHEAD = 16;  // first k-mer of a run (a chain of edges inside one LDS group)

constexpr int BLOCK = 256;
constexpr int WAVES = BLOCK / 64;
constexpr int EDGE_TILE = 1024;               // slots per edge-finding tile
constexpr uint32_t MAX_BUCKETS = 1024;        // coarse buckets per set
constexpr uint32_t MAX_SUB = 1024;            // fine buckets per coarse bucket
constexpr uint32_t LDS_GROUP = 2048;          // slots per LDS edge-finding group
constexpr uint32_t BIG_GROUP = 4096;          // a lone fine bucket up to this size still runs in LDS
constexpr uint32_t SLICE_MAX = 65535;         // larger fine buckets up to this size run in LDS hash slices
constexpr uint32_t SLICE_KEYS = 2048;         // prefix keys per slice (slices per bucket: size / this)
constexpr uint32_t SLICE_TAB = 8192;          // LDS table entries of a slice
constexpr uint32_t DEFAULT_SAMPLE_LOG2 = 9;   // splitter sampling 2^-9 of the run heads

struct BTile {              // bucketing tile (one workgroup)
    uint64_t kbegin;        // first k-mer (pass 1) or slot (pass 2), global index
    uint64_t hbase;         // first histogram entry of the tile's set / coarse bucket
    uint32_t count;         // items in the tile
    uint32_t set;           // set (pass 1) or coarse bucket (pass 2)
    uint32_t ts;            // tiles of the set / coarse bucket
    uint32_t tlocal;        // tile index inside it
    uint32_t fmask;         // pass 1: fine-bucket mask B*S-1 of the set; pass 2: S
    uint32_t pad;
};

struct EGroup {             // edge-finding group: consecutive fine buckets of one
    uint32_t begin, end;    // coarse bucket, staged in LDS by one workgroup
};

struct SliceTask {          // one hash slice of one large fine bucket
    uint32_t begin, end;    // the fine bucket's slots
    uint32_t p, P;          // slice p of P: the keys whose hash falls in it
};

struct ETile {              // edge-finding tile: slots of one bucket
    uint32_t bucket, begin, end, pad;
};

struct LminParams {
    uint32_t k, l, B;
    uint32_t lmask, h1, h2, s0;  // order bijection: mask, xorshift amounts, seed word
    uint32_t w1, w2, k1;         // 2l - h1, 2l - h2, s0 * MIX_C1 (mod 2^32)
    const uint32_t *table;           // optional order table (4^l entries)
};

}  // namespace uc
