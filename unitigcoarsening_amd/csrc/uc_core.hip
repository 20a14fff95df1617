// uc_core.hip -- MI355X (gfx950) unitig compaction: kernels and the C ABI.
//
// Pipeline of one job (reference phases in brackets):
//   1. bucketing   [lmin_buckets.h:288-433, driver.cpp:137-225]
//      k_lmin_count: prefix/suffix l-minimizers per k-mer, per-tile bucket
//      histogram in LDS; device scan; k_lmin_scatter: stable wavefront-
//      ballot ranking, LDS-staged coalesced writes, crossing k-mers
//      duplicated into their suffix bucket, (home, borrow) pair list;
//      k_sub_count / k_sub_scatter: the same counting sort of each coarse
//      bucket by the next bits of the minimizer order (fine buckets).
//   2. edge finding [assembler.h:113-214]
//      k_edge_lds: one workgroup per group of fine buckets, k-mers staged
//      in LDS, insert / lookup / resolve against an LDS table; edges inside
//      the group are chained into runs by pointer jumping in LDS.  Fine
//      buckets too large for LDS use HBM tables (k_edge_insert / lookup /
//      resolve).
//   3. contraction [coarseners.h:68-485, heuristics.h:105-256,
//      compact_graph.h, driver.cpp:527-587]
//      k_splice re-attaches each borrowed duplicate's out-edge to its home
//      copy; splitters (path heads + a hashed 2^-s sample of run heads)
//      walk their segment run by run (k_seg_count); segment chains are
//      ranked by pointer jumping (k_chain_*).  Isolated cycles are cut at
//      their lexicographically smallest k-mer and the contraction repeats.
//   4. emission    [compact.h:59-210, lmin_buckets.h:172-283]
//      k_head_emit / k_rand_emit write each unitig: its first k-mer, then
//      the characters of its runs, which the edge kernel left in path order.
//   5. text input  [driver.cpp:227-354]: k_txt_count / k_txt_pack.
#include <hip/hip_runtime.h>

#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include "../../include/uc_api.h"
#include "uc_internal.h"
#include "uc_kernels.h"

namespace uc {

// ---------------------------------------------------------------------------
// device helpers

template <int W>
struct Kmer {
    uint64_t w[W];
};

__device__ __forceinline__ uint32_t fmix32(uint32_t h)
{
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    h *= 0xC2B2AE35u;
    h ^= h >> 16;
    return h;
}

template <int W>
__device__ __forceinline__ Kmer<W> load_kmer(const uint64_t *__restrict__ a, uint64_t i, uint64_t top_mask)
{
    Kmer<W> x;
    if constexpr (W == 1) {
        x.w[0] = a[i] & top_mask;
    } else {
        const ulonglong2 v = reinterpret_cast<const ulonglong2 *>(a)[i];
        x.w[0] = v.x;
        x.w[1] = v.y & top_mask;
    }
    return x;
}

template <int W>
__device__ __forceinline__ void store_kmer(uint64_t *__restrict__ a, uint64_t i, const Kmer<W> &x)
{
    if constexpr (W == 1) {
        a[i] = x.w[0];
    } else {
        reinterpret_cast<ulonglong2 *>(a)[i] = make_ulonglong2(x.w[0], x.w[1]);
    }
}

// (k-1)-prefix: characters 0..k-2 = the low 2(k-1) bits
template <int W>
__device__ __forceinline__ Kmer<W> prefix_key(const Kmer<W> &x, uint64_t pm0, uint64_t pm1)
{
    Kmer<W> r;
    r.w[0] = x.w[0] & pm0;
    if constexpr (W == 2) r.w[1] = x.w[1] & pm1;
    return r;
}

// (k-1)-suffix: characters 1..k-1 = the value shifted right by one character
template <int W>
__device__ __forceinline__ Kmer<W> suffix_key(const Kmer<W> &x)
{
    Kmer<W> r;
    if constexpr (W == 1) {
        r.w[0] = x.w[0] >> 2;
    } else {
        r.w[0] = (x.w[0] >> 2) | (x.w[1] << 62);
        r.w[1] = x.w[1] >> 2;
    }
    return r;
}

template <int W>
__device__ __forceinline__ bool key_eq(const Kmer<W> &a, const Kmer<W> &b)
{
    if constexpr (W == 1)
        return a.w[0] == b.w[0];
    else
        return a.w[0] == b.w[0] && a.w[1] == b.w[1];
}

// table position hash and an independent 28-bit tag
template <int W>
__device__ __forceinline__ void key_hash(const Kmer<W> &x, uint32_t &h, uint32_t &tag)
{
    const uint32_t a = (uint32_t)x.w[0], b = (uint32_t)(x.w[0] >> 32);
    // the position: one finalizer over the words folded by odd multiplies
    // (a finalizer per word measured 2% slower in edge finding on two-word
    // k-mers); the tag keeps its own chain
    if constexpr (W == 1) {
        h = fmix32(a ^ (b * 0x9E3779B9u));
        tag = fmix32(b ^ fmix32(a ^ 0x7FEB352Du)) & TAG_MASK;
    } else {
        const uint32_t c = (uint32_t)x.w[1], d = (uint32_t)(x.w[1] >> 32);
        h = fmix32(a ^ (b * 0x9E3779B9u) ^ fmix32(c ^ (d * 0x85EBCA77u)));
        tag = fmix32(d ^ fmix32(c ^ fmix32(b ^ fmix32(a ^ 0x7FEB352Du)))) & TAG_MASK;
    }
}

// lexicographic key: reverse the order of the 2-bit characters
__device__ __forceinline__ uint64_t rev_chars64(uint64_t x)
{
    uint64_t y = __builtin_bitreverse64(x);
    return ((y >> 1) & 0x5555555555555555ull) | ((y & 0x5555555555555555ull) << 1);
}

template <int W>
__device__ __forceinline__ bool lex_less(const Kmer<W> &a, const Kmer<W> &b)
{
    // both values sit in the low 2k bits; reversing all 64*W bits puts
    // character 0 in the top bits of both, so numeric order is lexicographic
    if constexpr (W == 1) {
        return rev_chars64(a.w[0]) < rev_chars64(b.w[0]);
    } else {
        const uint64_t ah = rev_chars64(a.w[0]), bh = rev_chars64(b.w[0]);
        if (ah != bh) return ah < bh;
        return rev_chars64(a.w[1]) < rev_chars64(b.w[1]);
    }
}

// How the order of an l-mer code is found: a caller's table, or the seeded
// bijection computed with 24-bit multiplies (2l <= 24: the full-rate
// v_mul_u32_u24, same residues modulo 2^(2l)) or 32-bit ones.  The mode is a
// template argument so that the hot loop has no branch on it (with a runtime
// branch the compiler merged both forms into the quarter-rate 32-bit one).
enum { LM_TABLE = 0, LM_MUL24 = 1, LM_MUL32 = 2 };

__host__ __device__ constexpr int lmin_mode(uint32_t l, bool table)
{
    return table ? LM_TABLE : (2 * l <= 24 ? LM_MUL24 : LM_MUL32);
}

// seeded bijection on 2l-bit l-mer codes (same function as the oracle's
// orc_lmer_order): add the seed, multiply by an odd constant, xorshift,
// multiply by another, xorshift, all modulo 2^(2l).  Only the low 2l bits of
// `code` count: a product's low bits depend only on its operands' low bits,
// so the rolling code may carry older characters above them.  Per round: one
// (24-bit) multiply(-add), one bit-field extract for the shifted term and one
// three-input op for mask and xor.
constexpr uint32_t MIX_C1 = 0x3779B1u, MIX_C2 = 0xEBCA6Bu;

template <int MODE>
__device__ __forceinline__ uint32_t lmer_order_m(uint32_t code, const LminParams &p)
{
    if constexpr (MODE == LM_TABLE) {
        return p.table[code & p.lmask];
    } else {
        uint32_t t;
        if constexpr (MODE == LM_MUL24)
            t = __umul24(code, MIX_C1) + p.k1;  // (code + s0) * C1
        else
            t = code * MIX_C1 + p.k1;
        uint32_t x = (t & p.lmask) ^ __builtin_amdgcn_ubfe(t, p.h1, p.w1);
        if constexpr (MODE == LM_MUL24)
            t = __umul24(x, MIX_C2);
        else
            t = x * MIX_C2;
        return (t & p.lmask) ^ __builtin_amdgcn_ubfe(t, p.h2, p.w2);
    }
}

__device__ __forceinline__ uint32_t lmer_order(uint32_t code, const LminParams &p)
{
    if (p.table) return lmer_order_m<LM_TABLE>(code, p);
    return 2 * p.l <= 24 ? lmer_order_m<LM_MUL24>(code, p) : lmer_order_m<LM_MUL32>(code, p);
}

// lmin_buckets.h:67-94: minimum order over the l-mers of the (k-1)-prefix
// (o1) and of the (k-1)-suffix (o2); the l-mer code keeps its first
// character most significant.  Both share the middle l-mers: o1 = min(first,
// middle), o2 = min(middle, last).  Characters are read 16 to a 32-bit word;
// the loop over a word's characters keeps a run-time (uniform) trip count, so
// the code stays small whatever k is.
template <int W, int MODE>
__device__ __forceinline__ void double_lmin(const Kmer<W> &x, const LminParams &p, uint32_t &o1, uint32_t &o2)
{
    const uint32_t k = p.k, l = p.l;
    uint32_t lmer = 0, mid = NONE;
    // characters [i0, i1) into the rolling code; hash (into mid) when HASH
    auto scan = [&](uint32_t i0, uint32_t i1, auto hash) {
#pragma unroll
        for (int q = 0; q < 2 * W; q++) {
            const uint32_t b = 16u * q;
            const uint32_t lo = i0 > b ? i0 : b, hi = i1 < b + 16u ? i1 : b + 16u;
            const uint32_t w = (uint32_t)(x.w[q / 2] >> (32 * (q % 2)));
#pragma unroll 1
            for (uint32_t i = lo; i < hi; i++) {  // uniform trip count
                // characters older than the l-mer stay above its 2l bits
                lmer = (lmer << 2) | __builtin_amdgcn_ubfe(w, 2 * (i - b), 2);
                if constexpr (decltype(hash)::value) mid = min(mid, lmer_order_m<MODE>(lmer, p));
            }
        }
    };
    scan(0, l, std::false_type{});
    const uint32_t first = lmer_order_m<MODE>(lmer, p);
    scan(l, k - 1, std::true_type{});
    scan(k - 1, k, std::false_type{});
    const uint32_t last = lmer_order_m<MODE>(lmer, p);
    o1 = min(first, mid);
    o2 = min(mid, last);
}

// 2-bit code -> 'A','C','G','T' (bytes 0x41, 0x43, 0x47, 0x54)
__device__ __forceinline__ char base_char(uint32_t c) { return (char)((0x54474341u >> (8 * c)) & 0xFFu); }

__device__ __forceinline__ bool sampled(uint32_t v, uint32_t salt, uint32_t shift)
{
    return shift != 0 && (fmix32(v ^ salt) >> (32 - shift)) == 0;
}

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup i runs on
// XCD i % 8).  Renumbering gives XCD x the contiguous tile range
// [x*per, (x+1)*per), so the buckets one XCD works on stay in its own L2.
constexpr uint32_t N_XCD = 8;

__device__ __forceinline__ uint32_t xcd_tile(uint32_t n_tiles)
{
    if (!(n_tiles & 0x80000000u)) return blockIdx.x < n_tiles ? blockIdx.x : NONE;  // natural order
    n_tiles &= 0x7FFFFFFFu;
    const uint32_t per = (n_tiles + N_XCD - 1) / N_XCD;
    const uint32_t t = (blockIdx.x % N_XCD) * per + blockIdx.x / N_XCD;
    return t < n_tiles ? t : NONE;
}

// ---------------------------------------------------------------------------
// generic kernels

template <typename T>
__global__ void k_fill(T *__restrict__ a, uint64_t n, T v)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        a[i] = v;
}

// A copy from device memory into page-locked host memory, the host side
// through its device mapping.  The copies of a server's workers went through
// the DMA engine one after another, inputs in and texts out, and a worker's
// small counter read waited behind another worker's text copy; a kernel's
// copies out overlap the engine's copies in and each other.  16-byte words
// when both ends allow them, else 4-byte words, else bytes (the branch is
// uniform).
__global__ __launch_bounds__(BLOCK) void k_host_copy(const char *__restrict__ src, char *__restrict__ dst,
                                                     uint64_t bytes)
{
    const uint64_t t = (uint64_t)blockIdx.x * BLOCK + threadIdx.x, nt = (uint64_t)gridDim.x * BLOCK;
    const uintptr_t al = reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(dst);
    uint64_t done = 0;
    if (al % 16 == 0) {
        const uint64_t nw = bytes / 16;
        for (uint64_t i = t; i < nw; i += nt)
            reinterpret_cast<uint4 *>(dst)[i] = reinterpret_cast<const uint4 *>(src)[i];
        done = nw * 16;
    } else if (al % 4 == 0) {
        const uint64_t nw = bytes / 4;
        for (uint64_t i = t; i < nw; i += nt)
            reinterpret_cast<uint32_t *>(dst)[i] = reinterpret_cast<const uint32_t *>(src)[i];
        done = nw * 4;
    }
    for (uint64_t i = done + t; i < bytes; i += nt) dst[i] = src[i];
}

// reference packing with an odd number of 32-bit words -> 64-bit words
__global__ void k_widen(const uint32_t *__restrict__ in, uint64_t n, uint32_t rw, uint32_t W, uint64_t *__restrict__ out)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        for (uint32_t j = 0; j < W; j++) {
            uint64_t lo = 2 * j < rw ? in[i * rw + 2 * j] : 0;
            uint64_t hi = 2 * j + 1 < rw ? in[i * rw + 2 * j + 1] : 0;
            out[i * W + j] = lo | (hi << 32);
        }
    }
}

// ---- exclusive scan (reduce, scan of block sums, scan) --------------------

constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = BLOCK * SCAN_ITEMS;

template <typename T>
__device__ __forceinline__ T wave_incl_scan(T v)
{
    const int lane = __lane_id();
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        T u = __shfl_up(v, o, 64);
        if (lane >= o) v += u;
    }
    return v;
}

// sum over a workgroup of NW waves (every thread gets it; sh holds NW words)
template <typename T, int NW>
__device__ __forceinline__ T block_reduce_n(T v, T *sh)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int w = threadIdx.x / 64;
    if (__lane_id() == 0) sh[w] = v;
    __syncthreads();
    T r = 0;
    for (int i = 0; i < NW; i++) r += sh[i];
    __syncthreads();
    return r;
}

template <typename T>
__device__ __forceinline__ T block_reduce(T v, T *sh)
{
    return block_reduce_n<T, WAVES>(v, sh);
}

template <typename T>
__global__ __launch_bounds__(BLOCK) void k_scan_reduce(const T *__restrict__ in, uint64_t n, T *__restrict__ partial)
{
    __shared__ T sh[WAVES];
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
    T s = 0;
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; j++) {
        uint64_t i = base + (uint64_t)j * BLOCK + threadIdx.x;
        if (i < n) s += in[i];
    }
    s = block_reduce(s, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// exclusive scan of one tile per block; `offset` holds per-block exclusive
// offsets (or is null); the block owning the last element writes the total
// (and a copy of it at `total2`, when given)
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_scan_tile(const T *in, T *out, uint64_t n, const T *__restrict__ offset,
                                                     T *__restrict__ total, T *__restrict__ total2)
{
    __shared__ T tile[SCAN_TILE];
    __shared__ T wsum[WAVES];
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; j++) {
        uint64_t i = base + (uint64_t)j * BLOCK + threadIdx.x;
        tile[j * BLOCK + threadIdx.x] = i < n ? in[i] : T(0);
    }
    __syncthreads();
    T loc[SCAN_ITEMS];
    T run = 0;
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; j++) {
        loc[j] = run;
        run += tile[threadIdx.x * SCAN_ITEMS + j];
    }
    T incl = wave_incl_scan(run);
    const int w = threadIdx.x / 64;
    if (__lane_id() == 63) wsum[w] = incl;
    __syncthreads();
    T wpre = 0, btot = 0;
    for (int i = 0; i < WAVES; i++) {
        if (i < w) wpre += wsum[i];
        btot += wsum[i];
    }
    const T boff = offset ? offset[blockIdx.x] : T(0);
    const T excl = boff + wpre + incl - run;
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; j++) tile[threadIdx.x * SCAN_ITEMS + j] = excl + loc[j];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; j++) {
        uint64_t i = base + (uint64_t)j * BLOCK + threadIdx.x;
        if (i < n) out[i] = tile[j * BLOCK + threadIdx.x];
    }
    if (threadIdx.x == 0 && base + SCAN_TILE >= n) {
        if (total) *total = boff + btot;
        if (total2) *total2 = boff + btot;
    }
}

// ---------------------------------------------------------------------------
// 1. bucketing

// Fine codes of the prefix and suffix minimizers, one pair per k-mer: two
// 16-bit halves of one word when every fine index fits 16 bits (C32), else
// a uint2.
template <bool C32>
__device__ __forceinline__ void put_codes(void *__restrict__ bkt, uint64_t i, uint32_t f1, uint32_t f2)
{
    if constexpr (C32)
        static_cast<uint32_t *>(bkt)[i] = f1 | (f2 << 16);
    else
        static_cast<uint2 *>(bkt)[i] = make_uint2(f1, f2);
}

template <bool C32>
__device__ __forceinline__ uint2 get_codes(const void *__restrict__ bkt, uint64_t i)
{
    if constexpr (C32) {
        const uint32_t v = static_cast<const uint32_t *>(bkt)[i];
        return make_uint2(v & 0xFFFFu, v >> 16);
    } else {
        return static_cast<const uint2 *>(bkt)[i];
    }
}

template <int W, int IPT, int MODE, bool C32>
__global__ __launch_bounds__(BLOCK) void k_lmin_count(const uint64_t *__restrict__ kin, const BTile *__restrict__ tiles,
                                                      LminParams p, uint64_t top_mask, void *__restrict__ bkt,
                                                      uint32_t *__restrict__ hist, uint32_t *__restrict__ xhist,
                                                      uint32_t n_tiles)
{
    extern __shared__ uint32_t cnt[];  // B buckets + 1 crossing counter
    // (XCD-contiguous tiles: neighbouring tiles write neighbouring words of
    // every bucket's histogram row, which one L2 then merges into whole lines)
    const uint32_t ti = xcd_tile(n_tiles);
    if (ti == NONE) return;
    const BTile t = tiles[ti];
    // fine bucket = order & (B*S - 1); its low log2(B) bits are the coarse
    // bucket.  A k-mer crosses when its two fine buckets differ (both copies
    // may then share a coarse bucket).
    for (uint32_t i = threadIdx.x; i <= p.B; i += BLOCK) cnt[i] = 0;
    __syncthreads();
    const uint32_t bmask = p.B - 1;
    // all of this thread's k-mers in flight at once
    Kmer<W> xs[IPT];
#pragma unroll
    for (int j = 0; j < IPT; j++) {
        const uint32_t idx = (uint32_t)j * BLOCK + threadIdx.x;
        if (idx < t.count) xs[j] = load_kmer<W>(kin, t.kbegin + idx, top_mask);
    }
#pragma unroll
    for (int j = 0; j < IPT; j++) {
        const uint32_t idx = (uint32_t)j * BLOCK + threadIdx.x;
        if (idx >= t.count) continue;
        uint32_t o1, o2;
        double_lmin<W, MODE>(xs[j], p, o1, o2);
        const uint32_t f1 = o1 & t.fmask, f2 = o2 & t.fmask;
        put_codes<C32>(bkt, t.kbegin + idx, f1, f2);
        atomicAdd(&cnt[f1 & bmask], 1u);
        if (f1 != f2) {
            atomicAdd(&cnt[f2 & bmask], 1u);
            atomicAdd(&cnt[p.B], 1u);
        }
    }
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < p.B; b += BLOCK) hist[t.hbase + (uint64_t)b * t.ts + t.tlocal] = cnt[b];
    if (threadIdx.x == 0) xhist[ti] = cnt[p.B];
}

// lanes whose home (mp) or borrow (ms) placement goes to bucket v
__device__ __forceinline__ void match_bucket(uint32_t v, uint32_t nbits, uint64_t V, uint64_t C,
                                             const uint64_t *BP, const uint64_t *BS, uint64_t &mp, uint64_t &ms)
{
    mp = V;
    ms = C;
#pragma unroll
    for (int j = 0; j < 10; j++) {
        if ((uint32_t)j < nbits) {
            if ((v >> j) & 1u) {
                mp &= BP[j];
                ms &= BS[j];
            } else {
                mp &= ~BP[j];
                ms &= ~BS[j];
            }
        }
    }
}

// block_excl_scan_n for PER = 2 or 4 values per thread at an aligned array:
// each thread's values move as one 8- or 16-byte word and stay in registers
// between the two passes (PER consecutive 4-byte words per lane put the lanes
// of a wave on only 32 / PER banks)
template <int NT, int PER>
__device__ uint32_t block_excl_scan_vec(uint32_t *a, uint32_t n, uint32_t *wsum)
{
    constexpr int NW = NT / 64;
    static_assert(PER == 2 || PER == 4, "vector scan of 2 or 4 values per thread");
    using V = typename std::conditional<PER == 4, uint4, uint2>::type;
    const uint32_t b0 = threadIdx.x * PER;
    const bool whole = b0 + PER <= n;
    uint32_t v[PER];
    if (whole) {
        const V x = *reinterpret_cast<const V *>(a + b0);
        v[0] = x.x;
        v[1] = x.y;
        if constexpr (PER == 4) {
            v[2] = reinterpret_cast<const uint4 &>(x).z;
            v[3] = reinterpret_cast<const uint4 &>(x).w;
        }
    } else {
#pragma unroll
        for (int i = 0; i < PER; i++) v[i] = b0 + i < n ? a[b0 + i] : 0u;
    }
    uint32_t loc = 0;
#pragma unroll
    for (int i = 0; i < PER; i++) loc += v[i];
    const uint32_t incl = wave_incl_scan(loc);
    const int w = threadIdx.x / 64;
    if (__lane_id() == 63) wsum[w] = incl;
    __syncthreads();
    uint32_t pre = incl - loc, tot = 0;
    for (int q = 0; q < NW; q++) {
        if (q < w) pre += wsum[q];
        tot += wsum[q];
    }
#pragma unroll
    for (int i = 0; i < PER; i++) {
        const uint32_t x = v[i];
        v[i] = pre;
        pre += x;
    }
    if (whole) {
        V y;
        y.x = v[0];
        y.y = v[1];
        if constexpr (PER == 4) {
            reinterpret_cast<uint4 &>(y).z = v[2];
            reinterpret_cast<uint4 &>(y).w = v[3];
        }
        *reinterpret_cast<V *>(a + b0) = y;
    } else {
#pragma unroll
        for (int i = 0; i < PER; i++)
            if (b0 + i < n) a[b0 + i] = v[i];
    }
    __syncthreads();
    return tot;
}

// exclusive scan of a[0..n) in place by a block of NT threads (wsum holds
// NT / 64 words); returns the total
template <int NT>
__device__ uint32_t block_excl_scan_n(uint32_t *a, uint32_t n, uint32_t *wsum)
{
    constexpr int NW = NT / 64;
    const uint32_t per = (n + NT - 1) / NT;
    // (per is uniform over the block, so every thread takes the same branch)
    if (per == 4 && !(reinterpret_cast<uintptr_t>(a) & 15)) return block_excl_scan_vec<NT, 4>(a, n, wsum);
    if (per == 2 && !(reinterpret_cast<uintptr_t>(a) & 7)) return block_excl_scan_vec<NT, 2>(a, n, wsum);
    const uint32_t b0 = threadIdx.x * per, b1 = min(n, b0 + per);
    uint32_t loc = 0;
    for (uint32_t i = b0; i < b1; i++) loc += a[i];
    const uint32_t incl = wave_incl_scan(loc);
    const int w = threadIdx.x / 64;
    if (__lane_id() == 63) wsum[w] = incl;
    __syncthreads();
    uint32_t pre = incl - loc, tot = 0;
    for (int q = 0; q < NW; q++) {
        if (q < w) pre += wsum[q];
        tot += wsum[q];
    }
    for (uint32_t i = b0; i < b1; i++) {
        const uint32_t v = a[i];
        a[i] = pre;
        pre += v;
    }
    __syncthreads();
    return tot;
}

__device__ __forceinline__ uint32_t block_excl_scan(uint32_t *a, uint32_t n, uint32_t *wsum)
{
    return block_excl_scan_n<BLOCK>(a, n, wsum);
}

// Stable scatter into the bucket CSR: inside a tile, placements keep the
// order of their k-mers (home copy before borrowed copy when both land in
// one bucket); per wave the rank among placements bound for the same bucket
// comes from log2(B) ballots of the bucket bits.  Every thread's codes and
// k-mers are loaded up front, so all of them are in flight at once.  With
// STAGE, the tile's placements are gathered in LDS in bucket order and leave
// as runs of consecutive slots; a tile with more placements than the stage
// holds (more than a quarter of its k-mers crossing) stores them directly.
// LDS of one pass-1 tile: the tile's start and global offset of every bucket,
// then a 16-bit placement count per bucket (and crossing count) for each of
// the tile's ng = tile / 64 (item, wave) groups
__host__ __device__ constexpr size_t lmin_lds_bytes(uint32_t B, uint32_t ng)
{
    return 2 * (size_t)B * 4 + ((size_t)(B + 1) * ng * 2 + 3) / 4 * 4;
}

// placements one pass-1 tile may stage: its k-mers plus a quarter
__host__ __device__ constexpr uint32_t lmin_stage_cap(uint32_t tile) { return tile + tile / 4; }

__host__ __device__ constexpr size_t lmin_stage_off(uint32_t B, uint32_t ng)
{
    return (lmin_lds_bytes(B, ng) + 15) & ~(size_t)15;
}

template <int W>
__host__ __device__ constexpr size_t lmin_stage_bytes(uint32_t B, uint32_t tile, bool keep)
{
    return lmin_stage_off(B, tile / 64) + (size_t)lmin_stage_cap(tile) * (8 * W + 2 + 2 + (keep ? 4 : 0));
}

// A pass-1 slot's fine index (below MAX_SUB) and its crossing flag share one
// 16-bit word, which is all the fine sort reads besides the k-mer
constexpr uint32_t SUB_MASK = 0x3FFFu;
constexpr uint32_t SUB_FLAG_SHIFT = 14;
static_assert(MAX_SUB - 1 <= SUB_MASK, "fine indices fit below the flag bits");
static_assert((F_LENT | F_BORROW) < 4, "crossing flags fit two bits");

// Pass 1 writes, per placement, the k-mer and either the final flag byte
// (flags != nullptr: no fine sort follows) or the fine index with the flag
// bits (sub != nullptr).
template <int W, int IPT, bool STAGE, int NT, bool C32>
__global__ __launch_bounds__(NT) void k_lmin_scatter(const uint64_t *__restrict__ kin, const BTile *__restrict__ tiles,
                                                        uint32_t B, uint32_t nbits, uint64_t top_mask,
                                                        const void *__restrict__ bkt,
                                                        const uint32_t *__restrict__ hscan,
                                                        const uint32_t *__restrict__ xscan, uint64_t *__restrict__ kout,
                                                        uint8_t *__restrict__ flags, uint32_t *__restrict__ src,
                                                        uint32_t *__restrict__ pair_home,
                                                        uint32_t *__restrict__ pair_borrow,
                                                        uint16_t *__restrict__ sub, uint32_t n_tiles)
{
    constexpr int NW = NT / 64;
    constexpr int NG = IPT * NW;  // (item, wave) groups, in the tile's order of k-mers
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ uint32_t wsum[NW];
    const bool keep = src != nullptr;
    uint32_t *tstart = reinterpret_cast<uint32_t *>(lds);  // bucket starts inside the tile
    uint32_t *delta = tstart + B;                          // slot of the tile's first placement of each bucket - start
    // placements per group and bucket (column B: crossing k-mers), then their
    // exclusive prefixes down each column
    uint16_t *gcnt = reinterpret_cast<uint16_t *>(delta + B);
    const uint32_t ti = xcd_tile(n_tiles);
    if (ti == NONE) return;
    const BTile t = tiles[ti];
    // all of this thread's codes and k-mers in flight at once
    uint2 codes[IPT];
    Kmer<W> xs[IPT];
#pragma unroll
    for (int j = 0; j < IPT; j++) {
        const uint32_t idx = (uint32_t)j * NT + threadIdx.x;
        codes[j] = make_uint2(0u, 0u);
        if (idx < t.count) {
            codes[j] = get_codes<C32>(bkt, t.kbegin + idx);
            xs[j] = load_kmer<W>(kin, t.kbegin + idx, top_mask);
        }
    }
    {
        uint32_t *g32 = reinterpret_cast<uint32_t *>(gcnt);
        for (uint32_t i = threadIdx.x; i < ((B + 1) * NG + 1) / 2; i += NT) g32[i] = 0;
    }
    // this tile's count per bucket is the next histogram entry minus its own
    for (uint32_t v = threadIdx.x; v < B; v += NT) {
        const uint64_t hi = t.hbase + (uint64_t)v * t.ts + t.tlocal;
        const uint32_t g0 = hscan[hi];
        tstart[v] = hscan[hi + 1] - g0;
        delta[v] = g0;
    }
    __syncthreads();
    block_excl_scan_n<NT>(tstart, B, wsum);
    for (uint32_t v = threadIdx.x; v < B; v += NT) delta[v] -= tstart[v];
    const int lane = __lane_id(), w = threadIdx.x / 64;
    const uint64_t lt = (1ull << lane) - 1ull, me = 1ull << lane;
    // 1. every item's rank among its wave's placements of the same bucket
    // (homes before borrows within a lane), and each wave's count per bucket
    uint32_t rh[IPT], rb[IPT], rc[IPT];
#pragma unroll
    for (int j = 0; j < IPT; j++) {
        rh[j] = rb[j] = rc[j] = 0;
        if ((uint32_t)j * NT >= t.count) break;  // uniform over the block
        const uint32_t idx = (uint32_t)j * NT + threadIdx.x;
        const bool valid = idx < t.count;
        uint32_t bp = 0, bs = 0;
        bool cross = false;
        if (valid) {
            bp = codes[j].x & (B - 1);
            bs = codes[j].y & (B - 1);
            cross = codes[j].x != codes[j].y;
        }
        const uint64_t V = __ballot(valid), C = __ballot(cross);
        uint64_t BP[10], BS[10];
#pragma unroll
        for (int q = 0; q < 10; q++) {
            BP[q] = (uint32_t)q < nbits ? (uint64_t)__ballot(valid && ((bp >> q) & 1u)) : 0ull;
            BS[q] = (uint32_t)q < nbits ? (uint64_t)__ballot(cross && ((bs >> q) & 1u)) : 0ull;
        }
        // placements of lower lanes in the same bucket: homes (hp*) and borrows (hs*)
        uint16_t *gc = gcnt + (size_t)(j * NW + w) * (B + 1);
        if (valid) {
            uint64_t hp = 0, hs = 0;
            match_bucket(bp, nbits, V, C, BP, BS, hp, hs);
            const uint64_t M = hp | hs;
            if ((M & (0ull - M)) == me) gc[bp] = (uint16_t)(__popcll(hp) + __popcll(hs));
            rh[j] = (uint32_t)(__popcll(hp & lt) + __popcll(hs & lt));
        }
        if (cross) {
            uint64_t hp = 0, hs = 0;
            match_bucket(bs, nbits, V, C, BP, BS, hp, hs);
            const uint64_t M = hp | hs;
            if ((M & (0ull - M)) == me) gc[bs] = (uint16_t)(__popcll(hp) + __popcll(hs));
            rb[j] = (uint32_t)(__popcll(hp & lt) + __popcll(hs & lt)) + (bp == bs ? 1u : 0u);
            rc[j] = (uint32_t)__popcll(C & lt);
        }
        if (C && (C & (0ull - C)) == me) gc[B] = (uint16_t)__popcll(C);
    }
    __syncthreads();
    // 2. each column's exclusive prefix over the groups in tile order
    // (at most 2 * tile placements per bucket: 16 bits hold them)
    for (uint32_t v = threadIdx.x; v <= B; v += NT) {
        uint32_t acc = 0;
#pragma unroll 8
        for (int g = 0; g < NG; g++) {
            const uint32_t c = gcnt[(size_t)g * (B + 1) + v];
            gcnt[(size_t)g * (B + 1) + v] = (uint16_t)acc;
            acc += c;
        }
    }
    __syncthreads();
    // 3. placements: k-mers, fine index with flags, bucket (and sources)
    const uint32_t xbase = xscan[ti];
    constexpr uint32_t CAP = lmin_stage_cap(IPT * NT);
    Kmer<W> *skm = reinterpret_cast<Kmer<W> *>(lds + lmin_stage_off(B, NG));
    uint16_t *ssub = reinterpret_cast<uint16_t *>(skm + CAP);
    uint16_t *sbk = ssub + CAP;
    uint32_t *ssrc = reinterpret_cast<uint32_t *>(sbk + CAP);  // 4-byte aligned: CAP is even
    const uint32_t placed = t.count + (xscan[ti + 1] - xbase);
    const bool staged = STAGE && placed <= CAP;  // uniform over the block
#pragma unroll
    for (int j = 0; j < IPT; j++) {
        const uint32_t idx = (uint32_t)j * NT + threadIdx.x;
        if (idx >= t.count) continue;
        const uint64_t g = t.kbegin + idx;
        const uint2 c = codes[j];
        const uint32_t fp = c.x >> nbits, fs = c.y >> nbits;  // fine indices inside the coarse buckets
        const uint32_t bp = c.x & (B - 1), bs = c.y & (B - 1);
        const bool cross = c.x != c.y;
        const uint16_t *gc = gcnt + (size_t)(j * NW + w) * (B + 1);
        const Kmer<W> x = xs[j];
        const uint32_t lh = tstart[bp] + gc[bp] + rh[j];
        const uint32_t wh = fp | ((cross ? (uint32_t)F_LENT : 0u) << SUB_FLAG_SHIFT);
        if (staged) {
            skm[lh] = x;
            ssub[lh] = (uint16_t)wh;
            sbk[lh] = (uint16_t)bp;
            if (keep) ssrc[lh] = (uint32_t)g;
        } else {
            const uint32_t ns = delta[bp] + lh;
            store_kmer<W>(kout, ns, x);
            if (flags) flags[ns] = (uint8_t)(wh >> SUB_FLAG_SHIFT);
            if (sub) sub[ns] = (uint16_t)wh;
            if (keep) src[ns] = (uint32_t)g;
        }
        if (cross) {
            const uint32_t lb = tstart[bs] + gc[bs] + rb[j];
            const uint32_t wb = fs | ((uint32_t)F_BORROW << SUB_FLAG_SHIFT);
            if (staged) {
                skm[lb] = x;
                ssub[lb] = (uint16_t)wb;
                sbk[lb] = (uint16_t)bs;
                if (keep) ssrc[lb] = (uint32_t)g;
            } else {
                const uint32_t ns = delta[bs] + lb;
                store_kmer<W>(kout, ns, x);
                if (flags) flags[ns] = F_BORROW;
                if (sub) sub[ns] = (uint16_t)wb;
                if (keep) src[ns] = (uint32_t)g;
            }
            const uint32_t px = gc[B] + rc[j];
            pair_home[xbase + px] = delta[bp] + lh;
            pair_borrow[xbase + px] = delta[bs] + lb;
        }
    }
    // staged: slot li of the tile's bucket-ordered placements goes to
    // delta[bucket] + li, so consecutive li of one bucket are consecutive slots
    if (staged) {
        __syncthreads();
        for (uint32_t li = threadIdx.x; li < placed; li += NT) {
            const uint32_t ns = delta[sbk[li]] + li;
            store_kmer<W>(kout, ns, skm[li]);
            const uint32_t v = ssub[li];
            if (flags) flags[ns] = (uint8_t)(v >> SUB_FLAG_SHIFT);
            if (sub) sub[ns] = (uint16_t)v;
            if (keep) src[ns] = ssrc[li];
        }
    }
}

// pass-1 tiles from the per-set plan: set s owns tiles [tfirst[s],
// tfirst[s + 1]) (at least one, possibly empty, so that every bucket row has
// a start in the histogram) and histogram entries from B * tfirst[s]
__global__ void k_plan1_tiles(const uint64_t *__restrict__ set_off, const uint64_t *__restrict__ tfirst,
                              const uint32_t *__restrict__ set_sub, uint32_t n_sets, uint32_t B, uint32_t tile,
                              uint64_t T, BTile *__restrict__ tiles)
{
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T) return;
    uint32_t lo = 0, hi = n_sets;  // the set with tfirst[lo] <= t < tfirst[lo + 1]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (tfirst[mid] <= t) lo = mid;
        else hi = mid;
    }
    const uint64_t ns = set_off[lo + 1] - set_off[lo], tl = t - tfirst[lo], skip = tl * tile;
    BTile bt;
    bt.kbegin = set_off[lo] + skip;
    bt.hbase = (uint64_t)B * tfirst[lo];
    bt.count = (uint32_t)(ns > skip ? (ns - skip < tile ? ns - skip : tile) : 0);
    bt.set = lo;
    bt.ts = (uint32_t)(tfirst[lo + 1] - tfirst[lo]);
    bt.tlocal = (uint32_t)tl;
    bt.fmask = B * set_sub[lo] - 1;
    bt.pad = 0;
    tiles[t] = bt;
}

// bucket row starts: boff[set*B + b] = scanned histogram at the set's first tile
// (and the same rows into boff2 when given: the fine rows of a job without
// a fine sort).  The slot count is *mdev when given (the host has not read
// it), else m.
__global__ void k_gather_boff(const uint32_t *__restrict__ hscan, const uint64_t *__restrict__ set_hbase,
                              const uint32_t *__restrict__ set_ts, uint32_t n_sets, uint32_t B, uint32_t m,
                              const uint32_t *__restrict__ mdev, uint32_t *__restrict__ boff,
                              uint32_t *__restrict__ boff2)
{
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t G = n_sets * B;
    uint32_t v;
    if (g < G) {
        const uint32_t s = g / B, b = g % B;
        v = hscan[set_hbase[s] + (uint64_t)b * set_ts[s]];
    } else if (g == G) {
        v = mdev ? *mdev : m;
    } else {
        return;
    }
    boff[g] = v;
    if (boff2) boff2[g] = v;
}

// ---- fine buckets: a stable counting sort of every coarse bucket's slots by
// their fine index (the next bits of the same minimizer order)

// Pass-2 tiles are planned on the device from the coarse rows: coarse
// bucket g (S_g fine buckets, ts_g = max(1, ceil(rows / tile)) tiles) owns
// tiles [tfirst_g, tfirst_g + ts_g) and histogram entries [hbase_g, hbase_g +
// ts_g * S_g), tile-major (entry t * S_g + v) so that every tile writes its
// counts contiguously.  Tiles past the last one are empty (fmask 0).

// one workgroup: tiles and histogram entries per coarse bucket, prefixes
constexpr int PLAN_NT = 1024;  // one workgroup walks all coarse buckets: the widest one

__global__ __launch_bounds__(PLAN_NT) void k_plan2_prefix(const uint32_t *__restrict__ boff,
                                                        const uint32_t *__restrict__ cb_first, uint32_t G, uint32_t tile,
                                                        uint32_t *__restrict__ cb_ts, uint32_t *__restrict__ cb_tfirst,
                                                        uint64_t *__restrict__ cb_hbase)
{
    __shared__ uint64_t wsum[PLAN_NT / 64], hsum[PLAN_NT / 64];
    __shared__ uint64_t tcarry, hcarry;
    if (threadIdx.x == 0) tcarry = hcarry = 0;
    __syncthreads();
    const int w = threadIdx.x / 64;
    for (uint32_t g0 = 0; g0 < G; g0 += PLAN_NT) {
        const uint32_t g = g0 + threadIdx.x;
        uint64_t ts = 0, hs = 0;
        if (g < G) {
            const uint32_t rows = boff[g + 1] - boff[g];
            ts = rows ? (rows + tile - 1) / tile : 1u;
            hs = ts * (cb_first[g + 1] - cb_first[g]);
        }
        const uint64_t ti = wave_incl_scan(ts), hi = wave_incl_scan(hs);
        if (__lane_id() == 63) {
            wsum[w] = ti;
            hsum[w] = hi;
        }
        __syncthreads();
        uint64_t tp = tcarry + ti - ts, hp = hcarry + hi - hs, tt = 0, ht = 0;
        for (int q = 0; q < PLAN_NT / 64; q++) {
            if (q < w) {
                tp += wsum[q];
                hp += hsum[q];
            }
            tt += wsum[q];
            ht += hsum[q];
        }
        if (g < G) {
            cb_ts[g] = (uint32_t)ts;
            cb_tfirst[g] = (uint32_t)tp;
            cb_hbase[g] = hp;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            tcarry += tt;
            hcarry += ht;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        cb_tfirst[G] = (uint32_t)tcarry;
        cb_hbase[G] = hcarry;
    }
}

// one thread per tile slot up to the upper bound
__global__ void k_plan2_tiles(const uint32_t *__restrict__ boff, const uint32_t *__restrict__ cb_first,
                              const uint32_t *__restrict__ cb_ts, const uint32_t *__restrict__ cb_tfirst,
                              const uint64_t *__restrict__ cb_hbase, uint32_t G, uint32_t tile, uint32_t n_tiles,
                              BTile *__restrict__ tiles)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_tiles) return;
    BTile bt;
    bt.pad = 0;
    if (i >= cb_tfirst[G]) {
        bt.kbegin = 0;
        bt.hbase = 0;
        bt.count = bt.set = bt.ts = bt.tlocal = bt.fmask = 0;
        tiles[i] = bt;
        return;
    }
    uint32_t lo = 0, hi = G;  // last g with cb_tfirst[g] <= i
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (cb_tfirst[mid] <= i) lo = mid; else hi = mid;
    }
    const uint32_t g = lo, t = i - cb_tfirst[g], rows = boff[g + 1] - boff[g];
    bt.kbegin = (uint64_t)boff[g] + (uint64_t)t * tile;
    bt.hbase = cb_hbase[g];
    bt.count = rows > t * tile ? min(tile, rows - t * tile) : 0u;
    bt.set = g;
    bt.ts = cb_ts[g];
    bt.tlocal = t;
    bt.fmask = cb_first[g + 1] - cb_first[g];
    tiles[i] = bt;
}

// counts per fine bucket of one planned fine-sort tile (up to SUB_TILE
// slots); every thread has all of its loads in flight before its first LDS add
constexpr int SUB_COUNT_NT = 512;

__global__ __launch_bounds__(SUB_COUNT_NT) void k_sub_count(const uint16_t *__restrict__ sub,
                                                            const BTile *__restrict__ tiles,
                                                            uint32_t *__restrict__ hist)
{
    constexpr int PER = 8;
    extern __shared__ uint32_t cnt[];
    const BTile t = tiles[blockIdx.x];
    const uint32_t S = t.fmask;
    if (S == 0) return;  // past the last tile
    for (uint32_t i = threadIdx.x; i < S; i += SUB_COUNT_NT) cnt[i] = 0;
    __syncthreads();
    for (uint32_t i0 = 0; i0 < t.count; i0 += PER * SUB_COUNT_NT) {
        uint32_t v[PER];
#pragma unroll
        for (int j = 0; j < PER; j++) {
            const uint32_t i = i0 + (uint32_t)j * SUB_COUNT_NT + threadIdx.x;
            v[j] = i < t.count ? (uint32_t)sub[t.kbegin + i] & SUB_MASK : NONE;
        }
#pragma unroll
        for (int j = 0; j < PER; j++)
            if (v[j] != NONE) atomicAdd(&cnt[v[j]], 1u);
    }
    __syncthreads();
    for (uint32_t v = threadIdx.x; v < S; v += SUB_COUNT_NT) hist[t.hbase + (uint64_t)t.tlocal * S + v] = cnt[v];
}

// per coarse bucket: each fine bucket's counts over the tiles become
// exclusive prefixes (in place), and the column totals go to colsum.  A block
// takes 64 columns (fine buckets) of one coarse bucket, lane = column; its
// waves split the tiles, so that G * S / 64 blocks share the work.
__global__ __launch_bounds__(BLOCK) void k_sub_colscan(uint32_t *__restrict__ hist, const uint64_t *__restrict__ cb_hbase,
                                                       const uint32_t *__restrict__ cb_ts,
                                                       const uint32_t *__restrict__ cb_first, uint32_t cpb,
                                                       uint32_t *__restrict__ colsum)
{
    __shared__ uint32_t part[WAVES][64];
    const uint32_t g = blockIdx.x / cpb, cb = blockIdx.x % cpb;
    const uint32_t f0 = cb_first[g], S = cb_first[g + 1] - f0;
    if (cb * 64 >= S) return;  // uniform over the block
    const uint32_t ts = cb_ts[g];
    const uint64_t hb = cb_hbase[g];
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x / 64;
    const uint32_t v = cb * 64 + lane;
    const bool on = v < S;
    const uint32_t per = (ts + WAVES - 1) / WAVES, t0 = min(ts, w * per), t1 = min(ts, t0 + per);
    uint32_t sum = 0;
    if (on) {
#pragma unroll 8
        for (uint32_t t = t0; t < t1; t++) sum += hist[hb + (uint64_t)t * S + v];
    }
    part[w][lane] = sum;
    __syncthreads();
    uint32_t run = 0;
    for (uint32_t q = 0; q < w; q++) run += part[q][lane];
    if (on) {
#pragma unroll 8
        for (uint32_t t = t0; t < t1; t++) {
            const uint32_t c = hist[hb + (uint64_t)t * S + v];
            hist[hb + (uint64_t)t * S + v] = run;
            run += c;
        }
        if (w == WAVES - 1) colsum[f0 + v] = run;
    }
}

// fine rows from the column totals: fboff[first_g + v] = boff[g] + totals of
// fine buckets < v of coarse bucket g
__global__ __launch_bounds__(BLOCK) void k_sub_rows(const uint32_t *__restrict__ colsum,
                                                    const uint32_t *__restrict__ cb_first,
                                                    const uint32_t *__restrict__ boff, uint32_t G, uint32_t m,
                                                    const uint32_t *__restrict__ mdev, uint32_t *__restrict__ fboff)
{
    __shared__ uint32_t tot[MAX_SUB];
    __shared__ uint32_t wsum[WAVES];
    const uint32_t g = blockIdx.x;
    const uint32_t f0 = cb_first[g], S = cb_first[g + 1] - f0;
    for (uint32_t v = threadIdx.x; v < S; v += BLOCK) tot[v] = colsum[f0 + v];
    __syncthreads();
    block_excl_scan(tot, S, wsum);
    for (uint32_t v = threadIdx.x; v < S; v += BLOCK) fboff[f0 + v] = boff[g] + tot[v];
    if (g == G - 1 && threadIdx.x == 0) fboff[cb_first[G]] = mdev ? *mdev : m;
}

// Pass-2 scatter with LDS staging.  A planned tile (up to ROUNDS rounds of
// IPT * NT slots of one coarse bucket) is sorted round by round; each round is
// ranked per wave: each wave owns a chunk of consecutive slots, counts its
// slots per fine bucket (ballot groups, one LDS add per group), the counts are
// scanned once per round, and a second sweep gives every slot its stable
// position inside the round, sorted by fine bucket.  The slots are staged
// there in LDS and written out in that order, so consecutive lanes store
// consecutive addresses of one fine bucket's run.  The running base of every
// fine bucket carries from round to round, so the histogram holds one row per
// planned tile (ROUNDS times fewer than one per round), and the next round's
// slots are loaded while the current one is ranked.
__host__ __device__ constexpr size_t sub_book_bytes(uint32_t S, uint32_t nw = WAVES)
{
    return (((size_t)nw * S + 3 * (size_t)S) * 4 + 15) / 16 * 16;
}

template <int W>
__host__ __device__ constexpr size_t sub_lds_bytes(uint32_t tile, uint32_t S, bool src, bool stage, uint32_t nw = WAVES)
{
    return sub_book_bytes(S, nw) + (stage ? (size_t)tile * 8 * W + (src ? (size_t)tile * 4 : 0) + (size_t)tile * 2 : 0);
}

constexpr uint32_t SUB_ROUNDS = 4;

template <int W, bool STAGE, int IPT, int NT>
__global__ __launch_bounds__(NT) void k_sub_scatter(const BTile *__restrict__ tiles, const uint16_t *__restrict__ sub,
                                                       const uint32_t *__restrict__ hscan,
                                                       const uint32_t *__restrict__ fboff,
                                                       const uint32_t *__restrict__ cb_first, const uint64_t *__restrict__ kin,
                                                       const uint32_t *__restrict__ sin, uint64_t *__restrict__ kout,
                                                       uint8_t *__restrict__ fout, uint32_t *__restrict__ sout,
                                                       uint32_t *__restrict__ perm, uint32_t n_tiles)
{
    constexpr int NW = NT / 64;
    constexpr uint32_t TILE = (uint32_t)IPT * NT, CHUNK = TILE / NW;  // slots per round / per wave
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ uint32_t wsum[NW];
    const uint32_t ti = xcd_tile(n_tiles);
    if (ti == NONE) return;
    const BTile t = tiles[ti];
    const uint32_t S = t.fmask;
    if (S == 0) return;  // past the last tile
    const uint32_t nbits = (uint32_t)__builtin_ctz(S);
    const bool src = sin != nullptr;
    uint32_t *wc = reinterpret_cast<uint32_t *>(lds);  // NW x S
    uint32_t *tstart = wc + NW * S;
    uint32_t *delta = tstart + S;
    uint32_t *gbase = delta + S;
    // staging (STAGE): slots sorted by fine bucket
    unsigned char *st = lds + sub_book_bytes(S, NW);
    Kmer<W> *skm = reinterpret_cast<Kmer<W> *>(st);
    uint32_t *ssrc = reinterpret_cast<uint32_t *>(st + (size_t)TILE * 8 * W);
    uint16_t *ssub = reinterpret_cast<uint16_t *>(st + (size_t)TILE * 8 * W + (src ? (size_t)TILE * 4 : 0));
    const int lane = __lane_id(), w = threadIdx.x / 64;
    const uint64_t lt = (1ull << lane) - 1ull, me = 1ull << lane;
    const uint64_t zero[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    const uint32_t c0 = (uint32_t)w * CHUNK;
    uint32_t *mywc = wc + (uint32_t)w * S;
    // the tile's base in every fine bucket
    for (uint32_t v = threadIdx.x; v < S; v += NT)
        gbase[v] = fboff[cb_first[t.set] + v] + hscan[t.hbase + (uint64_t)t.tlocal * S + v];
    // this lane's slots of a round: c0 + 64 j + lane; fine index | flags << 14
    uint32_t vf[IPT], nvf[IPT];
    Kmer<W> xs[IPT], nxs[IPT];
    uint32_t sr[IPT], nsr[IPT];
    auto load_round = [&](uint32_t r0, uint32_t *V, Kmer<W> *X, uint32_t *R) {
#pragma unroll
        for (int j = 0; j < IPT; j++) {
            const uint32_t idx = r0 + c0 + 64u * j + (uint32_t)lane;
            V[j] = NONE;
            R[j] = 0;
            if (idx < t.count) {
                const uint64_t s = t.kbegin + idx;
                V[j] = (uint32_t)sub[s];
                X[j] = load_kmer<W>(kin, s, ~0ull);
                if (src) R[j] = sin[s];
            }
        }
    };
    load_round(0, vf, xs, sr);
    for (uint32_t r0 = 0; r0 < t.count; r0 += TILE) {  // uniform over the block
        if (r0 + TILE < t.count) load_round(r0 + TILE, nvf, nxs, nsr);
        for (uint32_t i = threadIdx.x; i < NW * S; i += NT) wc[i] = 0;
        __syncthreads();
        // 1. per-wave counts
#pragma unroll
        for (int j = 0; j < IPT; j++)
            if (vf[j] != NONE) atomicAdd(&mywc[vf[j] & SUB_MASK], 1u);
        __syncthreads();
        // 2. round starts per fine bucket; per-wave bases; global minus local
        // offsets; the bases move past this round's slots
        for (uint32_t v = threadIdx.x; v < S; v += NT) {
            uint32_t tot = 0;
#pragma unroll
            for (int q = 0; q < NW; q++) tot += wc[q * S + v];
            tstart[v] = tot;
        }
        __syncthreads();
        block_excl_scan_n<NT>(tstart, S, wsum);
        for (uint32_t v = threadIdx.x; v < S; v += NT) {
            const uint32_t t0 = tstart[v];
            uint32_t b = t0;
#pragma unroll
            for (int q = 0; q < NW; q++) {
                const uint32_t cq = wc[q * S + v];
                wc[q * S + v] = b;
                b += cq;
            }
            delta[v] = gbase[v] - t0;
            gbase[v] += b - t0;
        }
        __syncthreads();
        // 3. stable positions: per pass, ballot groups of equal fine index
#pragma unroll
        for (int j = 0; j < IPT; j++) {
            const bool valid = vf[j] != NONE;
            const uint64_t V = __ballot(valid);
            if (!V) break;  // uniform over the wave
            const uint32_t v = vf[j] & SUB_MASK;
            uint64_t BP[10];
#pragma unroll
            for (int q = 0; q < 10; q++) BP[q] = (uint32_t)q < nbits ? (uint64_t)__ballot(valid && ((v >> q) & 1u)) : 0ull;
            uint64_t M = 0, none;
            match_bucket(v, nbits, V, 0ull, BP, zero, M, none);
            uint32_t old = 0;
            if (valid && (M & (0ull - M)) == me) old = atomicAdd(&mywc[v], (uint32_t)__popcll(M));
            const int leader = valid ? __builtin_ctzll(M) : lane;
            const uint32_t base = __shfl(old, leader, 64);
            if (valid) {
                const uint64_t s = t.kbegin + r0 + c0 + 64u * j + (uint32_t)lane;
                const uint32_t li = base + (uint32_t)__popcll(M & lt);
                const uint32_t ns = delta[v] + li;
                if constexpr (STAGE) {
                    skm[li] = xs[j];
                    ssub[li] = (uint16_t)vf[j];
                    if (src) ssrc[li] = sr[j];
                } else {
                    store_kmer<W>(kout, ns, xs[j]);
                    fout[ns] = (uint8_t)(vf[j] >> SUB_FLAG_SHIFT);
                    if (src) sout[ns] = sr[j];
                }
                perm[s] = ns;
            }
        }
        // 4. write out in sorted order
        if constexpr (STAGE) {
            __syncthreads();
            const uint32_t cnt = min(TILE, t.count - r0);
            for (uint32_t li = threadIdx.x; li < cnt; li += NT) {
                const uint32_t v = ssub[li];
                const uint32_t ns = delta[v & SUB_MASK] + li;
                store_kmer<W>(kout, ns, skm[li]);
                fout[ns] = (uint8_t)(v >> SUB_FLAG_SHIFT);
                if (src) sout[ns] = ssrc[li];
            }
        }
        __syncthreads();  // the next round reuses the counts and the stage
#pragma unroll
        for (int j = 0; j < IPT; j++) {
            vf[j] = nvf[j];
            xs[j] = nxs[j];
            sr[j] = nsr[j];
        }
    }
}

// (the pair count is *xdev when given: the host has not read it, and x is
// then its bound)
__global__ void k_pair_remap(uint32_t *__restrict__ pair_home, uint32_t *__restrict__ pair_borrow, uint32_t x,
                             const uint32_t *__restrict__ xdev, const uint32_t *__restrict__ perm)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= x || (xdev && i >= *xdev)) return;
    pair_home[i] = perm[pair_home[i]];
    pair_borrow[i] = perm[pair_borrow[i]];
}

// ---------------------------------------------------------------------------
// 2. edge finding
//
// One open-addressing table per bucket over the (k-1)-prefixes of the
// bucket's home copies; 16-byte entries:
//   key  k <= 32: the (k-1)-prefix itself (< 2^62, so never EMPTY);
//        k > 32:  claim word (tag << 32 | owner), the key read back from the
//        owner's k-mer on a tag match
//   own  bucket-local owner slot (30 bits) | owner's last character << 30
//   cnt  hits of suffix lookups (bits 0..15) | CNT_SHARED (prefix of more
//        than one k-mer)
// k_edge_insert claims prefixes, k_edge_lookup counts suffix hits and
// remembers the hit position, k_edge_resolve keeps u -> v when the prefix
// is unshared and was hit exactly once, writing link[u] = v
// and v's has-predecessor bit.

struct TEntry {
    uint64_t key;
    uint32_t own;
    uint32_t cnt;
};
constexpr uint32_t CNT_SHARED = 0x80000000u;
constexpr uint32_t OWN_MASK = 0x3FFFFFFFu;
constexpr uint32_t NO_LINK = NONE;

// Runs: maximal chains of edges u -> v with u and v in one LDS group and u
// not a borrowed duplicate (such edges never leave a fine bucket).  Every
// live k-mer belongs to one run.  A run's record sits at its first k-mer's
// slot (next run, length, where its characters are); rch holds the last
// character of every run member in path order
// (run-major inside the group's own slot range).  A lent k-mer ends its run
// and has no out-edge of its own: its link holds NONE in the low half and its
// run's first k-mer in the high half, so that the splice can continue the run
// through the duplicate's edge.  (Every other link is the successor in the
// low half; nothing reads a link's high half but the splice.)
struct RunOut {
    uint64_t *rec;
    char *rch;
    uint8_t *flags;
};

// run record (8 bytes): next run's first k-mer | (length - 1) << 32 (12 bits,
// a run never leaves its LDS group of at most 4096 slots) | (position of its
// characters in rch minus its own slot) << 44 (13 bits, signed: both lie in
// the group) | its first REC_INLINE characters as 2-bit codes << 57.  Runs of
// at most REC_INLINE k-mers (most: runs end where the minimizer changes) are
// spelled from the record alone, so following a unitig into another bucket
// reads one random line per run instead of two.
constexpr uint32_t REC_INLINE = 3;
__host__ __device__ __forceinline__ uint64_t make_rec(uint32_t next, uint32_t len, uint32_t slot, uint32_t pos,
                                                      uint32_t chars)
{
    return (uint64_t)next | ((uint64_t)((len - 1) & 0xFFFu) << 32) | ((uint64_t)((pos - slot) & 0x1FFFu) << 44) |
           ((uint64_t)(chars & 0x3Fu) << 57);
}
__device__ __forceinline__ uint32_t rec_chars(uint64_t r) { return (uint32_t)(r >> 57) & 0x3Fu; }
__device__ __forceinline__ uint32_t rec_next(uint64_t r) { return (uint32_t)r; }
__device__ __forceinline__ uint32_t rec_len(uint64_t r) { return (uint32_t)((r >> 32) & 0xFFFu) + 1u; }
__device__ __forceinline__ uint32_t rec_pos(uint64_t r, uint32_t slot)
{
    return slot + (uint32_t)(((int32_t)((uint32_t)(r >> 44) & 0x1FFFu) << 19) >> 19);
}

__device__ __forceinline__ uint32_t fastrange(uint32_t h, uint32_t cap)
{
    return (uint32_t)(((uint64_t)h * cap) >> 32);
}

template <int W>
__device__ __forceinline__ uint32_t char_at(const Kmer<W> &x, uint32_t i)
{
    // a select rather than an indexed read keeps x in registers
    const uint64_t v = W == 1 || i < 32 ? x.w[0] : x.w[W - 1];
    return (uint32_t)(v >> (2 * (i % 32))) & 3u;
}

template <int W>
__device__ __forceinline__ uint64_t claim_word(const Kmer<W> &key, uint32_t tag, uint32_t local)
{
    if constexpr (W == 1) return key.w[0];
    return ((uint64_t)tag << 32) | local;
}

// does occupant `cur` hold `key`?
template <int W>
__device__ __forceinline__ bool same_key(uint64_t cur, const Kmer<W> &key, uint32_t tag,
                                         const uint64_t *__restrict__ kmer, uint32_t base, uint64_t pm0, uint64_t pm1)
{
    if constexpr (W == 1) {
        return cur == key.w[0];
    } else {
        if ((uint32_t)(cur >> 32) != tag) return false;
        const Kmer<W> o = prefix_key<W>(load_kmer<W>(kmer, base + (uint32_t)cur, ~0ull), pm0, pm1);
        return key_eq<W>(o, key);
    }
}

__device__ __forceinline__ bool has_pred(const uint32_t *__restrict__ hp, uint32_t v)
{
    return (hp[v >> 5] >> (v & 31)) & 1u;
}

__global__ void k_table_clear(TEntry *__restrict__ t, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        reinterpret_cast<ulonglong2 *>(t)[i] = make_ulonglong2(EMPTY, 0ull);
}

// Hash-table slot inspections (probes) of the edge phase, counted per
// thread when `probes` is given (UC_FLAG_COUNT_PROBES) and added with one
// atomic per workgroup, spread over PROBE_SLOTS counters so that workgroups
// on different XCDs rarely meet on one (every thread of the workgroup calls
// this; red holds one word per wave)
constexpr uint32_t PROBE_SLOTS = 256;
template <int NT>
__device__ __forceinline__ void add_probes(unsigned long long *probes, uint32_t np, uint32_t *red)
{
    for (int o = 32; o > 0; o >>= 1) np += __shfl_xor(np, o, 64);
    if (__lane_id() == 0) red[threadIdx.x / 64] = np;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t t = 0;
        for (int w = 0; w < NT / 64; w++) t += red[w];
        atomicAdd(probes + blockIdx.x % PROBE_SLOTS, (unsigned long long)t);
    }
}

template <int W, bool CNT>
__global__ __launch_bounds__(BLOCK) void k_edge_insert(const uint64_t *__restrict__ kmer, const uint8_t *__restrict__ flags,
                                                       const ETile *__restrict__ tiles, const uint32_t *__restrict__ boff,
                                                       const uint64_t *__restrict__ toff,
                                                       const uint32_t *__restrict__ tcap, TEntry *table, uint32_t k,
                                                       uint64_t pm0, uint64_t pm1, uint32_t n_tiles,
                                                       unsigned long long *__restrict__ probes)
{
    __shared__ uint32_t pred_[BLOCK / 64];
    uint32_t np = 0;
    const uint32_t ti = xcd_tile(n_tiles);
    if (ti == NONE) return;
    const ETile et = tiles[ti];
    const uint32_t base = boff[et.bucket];
    TEntry *tb = table + toff[et.bucket];
    const uint32_t cap = tcap[et.bucket];
    for (uint32_t s = et.begin + threadIdx.x; s < et.end; s += BLOCK) {
        if (flags[s] & F_BORROW) continue;
        const Kmer<W> x = load_kmer<W>(kmer, s, ~0ull);
        const Kmer<W> key = prefix_key<W>(x, pm0, pm1);
        uint32_t h, tag;
        key_hash<W>(key, h, tag);
        const uint32_t local = s - base;
        const uint64_t cw = claim_word<W>(key, tag, local);
        uint32_t pos = fastrange(h, cap);
        for (uint32_t probe = 0; probe < cap; probe++) {
            if (CNT) np++;
            // claim first: a compare-and-swap returns the occupant without
            // pulling the line through a separate load
            const uint64_t cur =
                atomicCAS((unsigned long long *)&tb[pos].key, (unsigned long long)EMPTY, (unsigned long long)cw);
            if (cur == EMPTY) {
                tb[pos].own = local | (char_at<W>(x, k - 1) << 30);
                break;
            }
            if (same_key<W>(cur, key, tag, kmer, base, pm0, pm1)) {
                atomicOr(&tb[pos].cnt, CNT_SHARED);
                break;
            }
            pos = pos + 1 == cap ? 0 : pos + 1;
        }
    }
    if (CNT) add_probes<BLOCK>(probes, np, pred_);
}

template <int W, bool CNT>
__global__ __launch_bounds__(BLOCK) void k_edge_lookup(const uint64_t *__restrict__ kmer, const uint8_t *__restrict__ flags,
                                                       const ETile *__restrict__ tiles, const uint32_t *__restrict__ boff,
                                                       const uint64_t *__restrict__ toff,
                                                       const uint32_t *__restrict__ tcap, TEntry *table, uint64_t pm0,
                                                       uint64_t pm1, uint32_t *__restrict__ hit, uint32_t n_tiles,
                                                       unsigned long long *__restrict__ probes)
{
    __shared__ uint32_t pred_[BLOCK / 64];
    uint32_t np = 0;
    const uint32_t ti = xcd_tile(n_tiles);
    if (ti == NONE) return;
    const ETile et = tiles[ti];
    const uint32_t base = boff[et.bucket];
    TEntry *tb = table + toff[et.bucket];
    const uint32_t cap = tcap[et.bucket];
    for (uint32_t s = et.begin + threadIdx.x; s < et.end; s += BLOCK) {
        uint32_t res = NONE;
        if (!(flags[s] & F_LENT)) {
            const Kmer<W> key = suffix_key<W>(load_kmer<W>(kmer, s, ~0ull));
            uint32_t h, tag;
            key_hash<W>(key, h, tag);
            uint32_t pos = fastrange(h, cap);
            for (uint32_t probe = 0; probe < cap; probe++) {
                if (CNT) np++;
                const uint64_t cur = tb[pos].key;
                if (cur == EMPTY) break;
                if (same_key<W>(cur, key, tag, kmer, base, pm0, pm1)) {
                    atomicAdd(&tb[pos].cnt, 1u);
                    res = pos;
                    break;
                }
                pos = pos + 1 == cap ? 0 : pos + 1;
            }
        }
        hit[s] = res;
    }
    if (CNT) add_probes<BLOCK>(probes, np, pred_);
}

__global__ __launch_bounds__(BLOCK) void k_edge_resolve(const ETile *__restrict__ tiles, const uint32_t *__restrict__ boff,
                                                        const uint64_t *__restrict__ toff, const TEntry *table,
                                                        const uint32_t *__restrict__ hit, uint32_t *__restrict__ link,
                                                        uint32_t *__restrict__ hasp, const uint64_t *__restrict__ kmer,
                                                        uint32_t W, uint32_t k, RunOut ro, uint32_t n_tiles)
{
    const uint32_t ti = xcd_tile(n_tiles);
    if (ti == NONE) return;
    const ETile et = tiles[ti];
    const uint32_t base = boff[et.bucket];
    const TEntry *tb = table + toff[et.bucket];
    for (uint32_t s = et.begin + threadIdx.x; s < et.end; s += BLOCK) {
        const uint32_t h = hit[s];
        uint32_t l = NO_LINK;
        if (h != NONE) {
            const uint2 e = *reinterpret_cast<const uint2 *>(&tb[h].own);
            if (e.y == 1u) {  // unshared prefix, exactly one suffix hit
                const uint32_t v = base + (e.x & OWN_MASK);
                l = v;
                atomicOr(&hasp[v >> 5], 1u << (v & 31));
            }
        }
        // an oversized fine bucket: every k-mer is a run of its own (a lent
        // one's link holds its run head, itself, until the splice)
        const uint8_t f = ro.flags[s];
        link[s] = (f & F_LENT) ? s : l;
        if (!(f & F_BORROW)) {
            ro.flags[s] = (uint8_t)(f | F_RUNHEAD);
            const uint64_t word = kmer[(uint64_t)s * W + (k - 1) / 32];
            const uint32_t c = (uint32_t)(word >> (2 * ((k - 1) & 31))) & 3u;
            ro.rec[s] = make_rec((uint32_t)l, 1u, s, s, c);
            ro.rch[s] = base_char(c);
        }
    }
}

// LDS groups, planned on the device: one workgroup per coarse bucket; its
// thread j packs the fine buckets 8j .. 8j+7 in order into groups of at most
// gmax slots (a group closes when the next fine bucket would overflow it).
// k_plan_count counts each coarse bucket's groups (and lists the fine buckets
// above gmax); after a scan of the counts k_plan_write lays every group out
// in one dense list, so that the edge kernel's workgroups are all busy
// (workgroups are dealt round-robin to the XCDs: runs of empty entries
// between the coarse buckets had left some XCDs idle).
constexpr uint32_t PLAN_RUN = 8;  // fine buckets per planning thread
static_assert(MAX_SUB <= PLAN_RUN * BLOCK, "one run of fine buckets per planning thread");

__device__ __forceinline__ uint32_t pack_groups(const uint32_t *__restrict__ fboff, uint32_t f0, uint32_t v0,
                                                uint32_t v1, uint32_t gmax, EGroup *out, uint32_t *__restrict__ over,
                                                uint32_t *__restrict__ n_over)
{
    uint32_t ng = 0, gb = 0, ge = 0;
    bool open = false;
    for (uint32_t v = v0; v < v1; v++) {
        const uint32_t r0 = fboff[f0 + v], r1 = fboff[f0 + v + 1];
        if (r1 - r0 > gmax) {
            if (over) over[atomicAdd(n_over, 1u)] = f0 + v;
            if (open && out) out[ng] = EGroup{gb, ge};
            ng += open;
            open = false;
            continue;
        }
        if (open && r1 - gb <= gmax) {
            ge = r1;
        } else if (r1 == r0) {
            continue;  // an empty fine bucket opens no group
        } else {
            if (open && out) out[ng] = EGroup{gb, ge};
            ng += open;
            gb = r0;
            ge = r1;
            open = true;
        }
    }
    if (open && out) out[ng] = EGroup{gb, ge};
    return ng + open;
}

// per coarse bucket: its group count (block sum of the threads' counts)
__global__ __launch_bounds__(BLOCK) void k_plan_count(const uint32_t *__restrict__ fboff,
                                                      const uint32_t *__restrict__ cb_first, uint32_t gmax,
                                                      uint32_t *__restrict__ gcount, uint32_t *__restrict__ over,
                                                      uint32_t *__restrict__ n_over)
{
    __shared__ uint32_t red[WAVES];
    const uint32_t g = blockIdx.x, f0 = cb_first[g], S = cb_first[g + 1] - f0;
    // S <= MAX_SUB <= PLAN_RUN * BLOCK: one run of fine buckets per thread
    const uint32_t v0 = PLAN_RUN * threadIdx.x, v1 = v0 + PLAN_RUN < S ? v0 + PLAN_RUN : S;
    const uint32_t ng = v0 < S ? pack_groups(fboff, f0, v0, v1, gmax, nullptr, over, n_over) : 0u;
    const uint32_t tot = block_reduce(ng, red);
    if (threadIdx.x == 0) gcount[g] = tot;
}

// per coarse bucket: its groups from its scanned first entry
// (and, when given, a copy of two counters beside the offsets, for the same
// host read)
__global__ __launch_bounds__(BLOCK) void k_plan_write(const uint32_t *__restrict__ fboff,
                                                      const uint32_t *__restrict__ cb_first, uint32_t gmax,
                                                      const uint32_t *__restrict__ gfirst,
                                                      EGroup *__restrict__ groups, const uint32_t *__restrict__ tot,
                                                      uint32_t *__restrict__ tot_out)
{
    __shared__ uint32_t wsum[WAVES];
    if (tot && blockIdx.x == 0 && threadIdx.x < 2) tot_out[threadIdx.x] = tot[threadIdx.x];
    const uint32_t g = blockIdx.x, f0 = cb_first[g], S = cb_first[g + 1] - f0;
    const uint32_t v0 = PLAN_RUN * threadIdx.x, v1 = v0 + PLAN_RUN < S ? v0 + PLAN_RUN : S;
    const uint32_t ng = v0 < S ? pack_groups(fboff, f0, v0, v1, gmax, nullptr, nullptr, nullptr) : 0u;
    const uint32_t incl = wave_incl_scan(ng);
    const int w = threadIdx.x / 64;
    if (__lane_id() == 63) wsum[w] = incl;
    __syncthreads();
    uint32_t base = gfirst[g] + incl - ng;
    for (int q = 0; q < w; q++) base += wsum[q];
    if (ng) pack_groups(fboff, f0, v0, v1, gmax, groups + base, nullptr, nullptr);
}

// ---- the LDS form: one workgroup stages a group of consecutive fine buckets
// (k-mers) in LDS and runs insert, lookup and resolve against an LDS table;
// each thread keeps its own slots' flags and last characters in registers.
// Keys of different fine buckets never match (equal (k-1)-mers have
// equal minimizers), so a group shares one table.  HBM sees only the
// sequential k-mer/flag reads and link writes plus one bitmap word per 32
// slots.
//   LDS layout (nb = launch capacity, a power of two):
//     km   Kmer<W>[nb]   staged k-mers
//     tab  uint32[2nb]   owner (16 bits) | LDS_SHARED | hits << 24, or LDS_EMPTY
//     hitp uint16[nb]    table position of each slot's suffix hit
//     hb   uint32[nb/32 + 2]  has-predecessor bits, aligned to global words

constexpr uint32_t LDS_EMPTY = 0xFFFFFFFFu;
constexpr uint32_t LDS_SHARED = 1u << 16;
constexpr uint16_t LDS_NOHIT = 0xFFFFu;

template <int W>
__host__ __device__ constexpr size_t edge_lds_bytes(uint32_t nb)
{
    return (size_t)nb * (8 * W + 8 + 2) + ((size_t)nb / 32 + 2) * 4;
}

template <int W, int EB, int GMAX, bool CNT>
__global__ __launch_bounds__(EB) void k_edge_lds(const uint64_t *__restrict__ kmer, const EGroup *__restrict__ groups,
                                                    uint32_t nb, uint32_t k, uint64_t pm0, uint64_t pm1,
                                                    uint32_t *__restrict__ link, uint32_t *__restrict__ hasp, RunOut ro,
                                                    unsigned long long *__restrict__ probes)
{
    constexpr int EPT = GMAX / EB;  // slots per thread (a group holds at most GMAX)
    __shared__ uint32_t pred_[EB / 64];
    uint32_t np = 0;
    const uint8_t *flags = ro.flags;
    __shared__ uint32_t wsum[EB / 64];
    extern __shared__ __align__(16) unsigned char lds[];
    Kmer<W> *km = reinterpret_cast<Kmer<W> *>(lds);
    uint32_t *tab = reinterpret_cast<uint32_t *>(lds + (size_t)nb * 8 * W);
    uint16_t *hitp = reinterpret_cast<uint16_t *>(lds + (size_t)nb * (8 * W + 8));
    uint32_t *hb = reinterpret_cast<uint32_t *>(lds + (size_t)nb * (8 * W + 10));  // nb is a multiple of 64
    const EGroup gr = groups[blockIdx.x];
    const uint32_t n = gr.end - gr.begin;
    if (n == 0) return;  // an unused group entry
    uint32_t tcap = 64;
    while (tcap < 2 * n) tcap <<= 1;
    const uint32_t tmask = tcap - 1;
    const uint32_t off0 = gr.begin & 31;
    const uint32_t nw = (off0 + n + 31) / 32;
    // the thread's k-mers stay in registers through insert and lookup; every
    // load is in flight before the first LDS store
    static_assert(EPT <= 8, "one flag byte per item in a 64-bit word");
    Kmer<W> xs[EPT];
    uint64_t fs = 0;  // flag byte of item e at bits 8e
#pragma unroll
    for (int e = 0; e < EPT; e++) {
        const uint32_t i = threadIdx.x + (uint32_t)e * EB;
        if (i < n) {
            xs[e] = load_kmer<W>(kmer, gr.begin + i, ~0ull);
            fs |= (uint64_t)flags[gr.begin + i] << (8 * e);
        }
    }
    uint32_t lc = 0;  // last character of item e at bits 2e
#pragma unroll
    for (int e = 0; e < EPT; e++) {
        const uint32_t i = threadIdx.x + (uint32_t)e * EB;
        if (i < n) {
            km[i] = xs[e];
            lc |= char_at<W>(xs[e], k - 1) << (2 * e);
        }
    }
    for (uint32_t i = threadIdx.x; i < tcap; i += EB) tab[i] = LDS_EMPTY;
    for (uint32_t i = threadIdx.x; i < nw; i += EB) hb[i] = 0;
    // home positions of all items first (independent work), then the probes
    uint32_t hpos[EPT];
#pragma unroll
    for (int e = 0; e < EPT; e++) {
        const uint32_t i = threadIdx.x + (uint32_t)e * EB;
        hpos[e] = NONE;
        if (i < n && !((fs >> (8 * e)) & F_BORROW)) {
            uint32_t h, tag;
            key_hash<W>(prefix_key<W>(xs[e], pm0, pm1), h, tag);
            hpos[e] = h & tmask;
        }
    }
    __syncthreads();
    // prefixes of home copies
#pragma unroll
    for (int e = 0; e < EPT; e++) {
        if (hpos[e] == NONE) continue;
        const uint32_t i = threadIdx.x + (uint32_t)e * EB;
        const Kmer<W> key = prefix_key<W>(xs[e], pm0, pm1);
        uint32_t pos = hpos[e];
        for (uint32_t probe = 0; probe < tcap; probe++) {
            if (CNT) np++;
            const uint32_t cur = atomicCAS(&tab[pos], LDS_EMPTY, i);
            if (cur == LDS_EMPTY) break;
            if (key_eq<W>(prefix_key<W>(km[cur & 0xFFFFu], pm0, pm1), key)) {
                atomicOr(&tab[pos], LDS_SHARED);
                break;
            }
            pos = (pos + 1) & tmask;
        }
    }
    // lookup positions (suffixes of every copy that carries its out-edge here)
#pragma unroll
    for (int e = 0; e < EPT; e++) {
        const uint32_t i = threadIdx.x + (uint32_t)e * EB;
        hpos[e] = NONE;
        if (i < n && !((fs >> (8 * e)) & F_LENT)) {
            uint32_t h, tag;
            key_hash<W>(suffix_key<W>(xs[e]), h, tag);
            hpos[e] = h & tmask;
        }
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < EPT; e++) {
        const uint32_t i = threadIdx.x + (uint32_t)e * EB;
        if (i >= n) continue;
        uint16_t hp = LDS_NOHIT;
        if (hpos[e] != NONE) {
            const Kmer<W> key = suffix_key<W>(xs[e]);
            uint32_t pos = hpos[e];
            for (uint32_t probe = 0; probe < tcap; probe++) {
                if (CNT) np++;
                const uint32_t en = tab[pos];
                if (en == LDS_EMPTY) break;
                if (key_eq<W>(prefix_key<W>(km[en & 0xFFFFu], pm0, pm1), key)) {
                    atomicAdd(&tab[pos], 1u << 24);
                    hp = (uint16_t)pos;
                    break;
                }
                pos = (pos + 1) & tmask;
            }
        }
        hitp[i] = hp;
    }
    __syncthreads();
    // keep u -> v when v's prefix is unshared and was hit exactly once; the
    // edges out of k-mers that are not borrowed chain the runs
    uint16_t tgt[EPT];
#pragma unroll
    for (int e = 0; e < EPT; e++) {
        const uint32_t i = threadIdx.x + (uint32_t)e * EB;
        tgt[e] = LDS_NOHIT;
        if (i >= n) continue;
        const uint16_t hp = hitp[i];
        uint32_t l = NO_LINK;
        if (hp != LDS_NOHIT) {
            const uint32_t en = tab[hp];
            if (!(en & LDS_SHARED) && (en >> 24) == 1u) {
                const uint32_t v = en & 0xFFFFu;
                l = gr.begin + v;
                atomicOr(&hb[(off0 + v) >> 5], 1u << ((off0 + v) & 31));
                if (!((fs >> (8 * e)) & F_BORROW)) tgt[e] = (uint16_t)v;
            }
        }
        link[gr.begin + i] = l;
    }
    __syncthreads();
    // boundary words may be shared with neighbouring groups
    for (uint32_t j = threadIdx.x; j < nw; j += EB) {
        const uint32_t bits = hb[j];
        if (bits) atomicOr(&hasp[(gr.begin >> 5) + j], bits);
    }
    // runs: run predecessor of every slot (hitp is free now; the flags and
    // last characters stay in registers, so the k-mer staging area is free)
    uint16_t *predl = hitp;
    for (uint32_t i = threadIdx.x; i < n; i += EB) predl[i] = LDS_NOHIT;
    __syncthreads();
#pragma unroll
    for (int e = 0; e < EPT; e++)
        if (tgt[e] != LDS_NOHIT) predl[tgt[e]] = (uint16_t)(threadIdx.x + (uint32_t)e * EB);
    // run walks inside LDS: every run head follows its run (a run ends where
    // the minimizer changes, so runs are short).  tab is free: pd = run head |
    // distance << 16 in its first half, successors in its second half; sc
    // (over the k-mers) takes the run lengths, then their run-major starts.
    uint32_t *pd = tab;
    uint16_t *succ = reinterpret_cast<uint16_t *>(tab + nb);
    uint32_t *sc = reinterpret_cast<uint32_t *>(km);
    // every slot's last character (2-bit code), for the run records: in the
    // k-mer area past sc (8 W nb bytes hold 4 nb + nb)
    uint8_t *cc = reinterpret_cast<uint8_t *>(km) + 4 * (size_t)nb;
#pragma unroll
    for (int e = 0; e < EPT; e++) {
        const uint32_t i = threadIdx.x + (uint32_t)e * EB;
        if (i < n) {
            succ[i] = tgt[e];
            pd[i] = LDS_EMPTY;
            sc[i] = 0;
            cc[i] = (uint8_t)((lc >> (2 * e)) & 3u);
        }
    }
    __syncthreads();
    uint16_t rl[EPT];
#pragma unroll
    for (int e = 0; e < EPT; e++) {
        const uint32_t i = threadIdx.x + (uint32_t)e * EB;
        rl[e] = 0;
        if (i >= n || ((fs >> (8 * e)) & F_BORROW) || predl[i] != LDS_NOHIT) continue;
        uint32_t v = i, d = 0;
        for (;;) {
            pd[v] = i | (d << 16);
            const uint16_t nx = succ[v];
            if (nx == LDS_NOHIT || d + 1 >= n) break;
            v = nx;
            d++;
        }
        rl[e] = (uint16_t)(d + 1);
        sc[i] = d + 1;
    }
    __syncthreads();
    // no run head reaches it: the slot lies on a cycle inside the group; such
    // k-mers are runs of their own
#pragma unroll
    for (int e = 0; e < EPT; e++) {
        const uint32_t i = threadIdx.x + (uint32_t)e * EB;
        if (i < n && !((fs >> (8 * e)) & F_BORROW) && pd[i] == LDS_EMPTY) sc[i] = 1;
    }
    __syncthreads();
    block_excl_scan_n<EB>(sc, n, wsum);
#pragma unroll
    for (int e = 0; e < EPT; e++) {
        const uint32_t i = threadIdx.x + (uint32_t)e * EB;
        if (i >= n) continue;
        const uint32_t g = gr.begin + i;
        const uint8_t f = (uint8_t)((fs >> (8 * e)) & 0x1Fu & ~F_RUNHEAD);
        if (f & F_BORROW) {
            ro.flags[g] = f;
            continue;
        }
        const uint32_t x = pd[i];
        const bool cyc = x == LDS_EMPTY;
        const uint32_t a = cyc ? i : (x & 0xFFFFu), d = cyc ? 0u : (x >> 16);
        ro.rch[gr.begin + sc[a] + d] = base_char((lc >> (2 * e)) & 3u);
        ro.flags[g] = (uint8_t)(f | (a == i ? F_RUNHEAD : 0));
        if (a == i) {
            const uint32_t len = cyc ? 1u : (uint32_t)rl[e];
            uint32_t ch = cc[i];
            if (len >= 2) {
                const uint32_t v1 = succ[i];
                ch |= (uint32_t)cc[v1] << 2;
                if (len >= 3) ch |= (uint32_t)cc[succ[v1]] << 4;
            }
            ro.rec[g] = make_rec(cyc ? gr.begin + tgt[e] : NONE, len, g, gr.begin + sc[i], ch);
        }
        if (f & F_LENT) link[g] = gr.begin + a;  // its run head, until the splice
    }
    if (CNT) add_probes<EB>(probes, np, pred_);
}

// ---- hash slices: a fine bucket above BIG_GROUP (and at most SLICE_MAX)
// slots is split by the hash of the (k-1)-mer key into P slices of about
// SLICE_KEYS prefix keys; one workgroup per slice streams the whole bucket
// (from L2) three times: insert the prefixes of its slice into an LDS table,
// count the suffix hits of its slice, and resolve.  Equal keys share a
// slice, so every edge is found in exactly one workgroup.  Table entry: owner
// (16 bits) | LDS_SHARED | 7-bit tag << 17 | hits << 24.  Every k-mer is a
// run of its own, as in the HBM-table form.
constexpr uint32_t SLICE_TAGS = 0x7Fu << 17;

template <int W, bool CNT>
__device__ __forceinline__ bool slice_probe(const uint32_t *tab, const uint64_t *__restrict__ kmer, uint32_t begin,
                                            const Kmer<W> &key, uint32_t h, uint32_t tg, uint64_t pm0, uint64_t pm1,
                                            uint32_t &pos, uint32_t &np)
{
    pos = h & (SLICE_TAB - 1);
    for (uint32_t probe = 0; probe < SLICE_TAB; probe++) {
        if (CNT) np++;
        const uint32_t e = tab[pos];
        if (e == LDS_EMPTY) return false;
        if ((e & SLICE_TAGS) == tg &&
            key_eq<W>(prefix_key<W>(load_kmer<W>(kmer, begin + (e & 0xFFFFu), ~0ull), pm0, pm1), key))
            return true;
        pos = (pos + 1) & (SLICE_TAB - 1);
    }
    return false;
}

template <int W, bool CNT>
__global__ __launch_bounds__(1024) void k_edge_slice(const uint64_t *__restrict__ kmer,
                                                     const SliceTask *__restrict__ tasks, uint32_t k, uint64_t pm0,
                                                     uint64_t pm1, uint32_t *__restrict__ link,
                                                     uint32_t *__restrict__ hasp, RunOut ro, uint32_t *__restrict__ err,
                                                     unsigned long long *__restrict__ probes, uint32_t plimit)
{
    constexpr uint32_t NT = 1024;
    __shared__ uint32_t tab[SLICE_TAB];
    __shared__ uint32_t pred_[NT / 64];
    uint32_t np = 0;
    const SliceTask t = tasks[blockIdx.x];
    const uint32_t n = t.end - t.begin;
    for (uint32_t i = threadIdx.x; i < SLICE_TAB; i += NT) tab[i] = LDS_EMPTY;
    __syncthreads();
    // 1. prefixes of home copies in this slice
    for (uint32_t i = threadIdx.x; i < n; i += NT) {
        const uint32_t s = t.begin + i;
        if (ro.flags[s] & F_BORROW) continue;
        const Kmer<W> key = prefix_key<W>(load_kmer<W>(kmer, s, ~0ull), pm0, pm1);
        uint32_t h, tag;
        key_hash<W>(key, h, tag);
        if (fastrange(h, t.P) != t.p) continue;
        const uint32_t tg = (tag & 0x7Fu) << 17;
        uint32_t pos = h & (SLICE_TAB - 1), probe = 0;
        for (; probe < plimit; probe++) {
            if (CNT) np++;
            const uint32_t cur = atomicCAS(&tab[pos], LDS_EMPTY, i | tg);
            if (cur == LDS_EMPTY) break;
            if ((cur & SLICE_TAGS) == tg &&
                key_eq<W>(prefix_key<W>(load_kmer<W>(kmer, t.begin + (cur & 0xFFFFu), ~0ull), pm0, pm1), key)) {
                atomicOr(&tab[pos], LDS_SHARED);
                break;
            }
            pos = (pos + 1) & (SLICE_TAB - 1);
        }
        // table full (not expected at SLICE_KEYS per slice): the host runs
        // the edge phase again with these fine buckets on HBM tables
        if (probe == plimit) atomicOr(err, 1u);
    }
    __syncthreads();
    // 2. suffix hits of the copies that carry their out-edge here, this slice
    for (uint32_t i = threadIdx.x; i < n; i += NT) {
        const uint32_t s = t.begin + i;
        if (ro.flags[s] & F_LENT) continue;
        const Kmer<W> key = suffix_key<W>(load_kmer<W>(kmer, s, ~0ull));
        uint32_t h, tag, pos;
        key_hash<W>(key, h, tag);
        if (fastrange(h, t.P) != t.p) continue;
        if (slice_probe<W, CNT>(tab, kmer, t.begin, key, h, (tag & 0x7Fu) << 17, pm0, pm1, pos, np))
            atomicAdd(&tab[pos], 1u << 24);
    }
    __syncthreads();
    // 3. links and singleton runs: a lent k-mer belongs to its prefix's
    // slice, every other k-mer to its suffix's
    for (uint32_t i = threadIdx.x; i < n; i += NT) {
        const uint32_t s = t.begin + i;
        const uint8_t f = ro.flags[s];
        const Kmer<W> x = load_kmer<W>(kmer, s, ~0ull);
        uint32_t h, tag;
        if (f & F_LENT) {
            key_hash<W>(prefix_key<W>(x, pm0, pm1), h, tag);
            if (fastrange(h, t.P) != t.p) continue;
            link[s] = s;  // its own run head, until the splice
            ro.flags[s] = (uint8_t)(f | F_RUNHEAD);
            ro.rec[s] = make_rec(NONE, 1u, s, s, char_at<W>(x, k - 1));
            ro.rch[s] = base_char(char_at<W>(x, k - 1));
            continue;
        }
        const Kmer<W> key = suffix_key<W>(x);
        key_hash<W>(key, h, tag);
        if (fastrange(h, t.P) != t.p) continue;
        uint32_t l = NO_LINK;
        uint32_t pos;
        if (slice_probe<W, CNT>(tab, kmer, t.begin, key, h, (tag & 0x7Fu) << 17, pm0, pm1, pos, np)) {
            const uint32_t e = tab[pos];
            if (!(e & LDS_SHARED) && (e >> 24) == 1u) {  // unshared prefix, exactly one suffix hit
                const uint32_t v = t.begin + (e & 0xFFFFu);
                l = v;
                atomicOr(&hasp[v >> 5], 1u << (v & 31));
            }
        }
        link[s] = l;
        if (!(f & F_BORROW)) {
            ro.flags[s] = (uint8_t)(f | F_RUNHEAD);
            ro.rec[s] = make_rec((uint32_t)l, 1u, s, s, char_at<W>(x, k - 1));
            ro.rch[s] = base_char(char_at<W>(x, k - 1));
        }
    }
    if (CNT) add_probes<NT>(probes, np, pred_);
}

// ---------------------------------------------------------------------------
// 3. contraction
//
// After the splice every live k-mer has at most one successor and one
// predecessor, so the kept edges form disjoint simple paths (and, rarely,
// isolated cycles), each a chain of runs.  Splitters are the path heads plus
// a hashed 2^-s sample of the other run heads; a segment is a splitter and
// the runs up to the next splitter.  Segment lengths come from one walk per
// splitter (one step per run: its length and the next run), and chains of
// segments are ranked by pointer jumping over the sampled splitters only.
// Emission then walks each segment again, copying its runs' characters.

constexpr int CT_TILE = SCAN_TILE;  // slots per contraction tile
static_assert(CT_TILE <= 2048, "a tile's list entries hold slot offsets in 11 bits");

// the borrowed duplicate's out-edge moves to the home copy; the target keeps
// its has-predecessor bit
__global__ void k_splice(const uint32_t *__restrict__ pair_home, const uint32_t *__restrict__ pair_borrow, uint32_t x,
                         uint32_t *__restrict__ link, uint64_t *__restrict__ rec)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= x) return;
    const uint32_t h = pair_home[i];
    const uint32_t l = link[pair_borrow[i]];
    const uint32_t head = link[h];  // the lent k-mer's run head
    link[h] = l;
    // a lent k-mer ends its run: the run continues through the duplicate's edge
    reinterpret_cast<uint32_t *>(rec)[2ull * head] = l;  // the record's next-run field
}

// list index for every flagged lane: one atomic per wave
__device__ __forceinline__ uint32_t wave_append(bool flag, uint32_t *counter)
{
    const uint64_t bal = __ballot(flag);
    if (!bal) return NONE;
    const int leader = __builtin_ctzll(bal);
    uint32_t base = 0;
    if (__lane_id() == (uint32_t)leader) base = atomicAdd(counter, (uint32_t)__popcll(bal));
    base = __shfl(base, leader, 64);
    return base + (uint32_t)__popcll(bal & ((1ull << __lane_id()) - 1ull));
}

// counters in the context's small buffer
// (uint32 indices: one 128-byte line each, so that atomics on different
// counters never share a line; C_TOT holds two uint64 totals)
enum {
    C_LIST = 0, C_ERR = 32, C_ACTIVE = 64, C_NONCONV = 96, C_CYC = 128, C_CUT = 160, C_TOT = 192,
    C_SLICE = 256, C_WORDS = 288
};

// chain rounds launched before the host looks at the active counters; a
// chain of up to 2^(SPEC_ROUNDS - 1) sampled segments converges within them
constexpr uint32_t SPEC_ROUNDS = 2;

// A speculative contraction pass whose result emission may use, from its
// counters: no error flag, every chain at its head after the speculative
// rounds, no chain without a head (an isolated cycle), every one of the n
// k-mers reached.  Emission after a pass the host has not checked yet runs
// only when this holds: otherwise the unitig lengths and chain ancestors
// are not all written, and the kernels write nothing (the host then runs
// the pass again in its checked form).
__device__ __forceinline__ bool pass_settled(const uint32_t *__restrict__ cnt, uint32_t n)
{
    const uint64_t tot = *reinterpret_cast<const uint64_t *>(cnt + C_TOT + 2);
    return !cnt[C_ERR] && !cnt[C_SLICE] && (!cnt[C_LIST] || !cnt[C_ACTIVE + SPEC_ROUNDS - 1]) && !cnt[C_NONCONV] &&
           (uint32_t)tot == n;
}

__device__ __forceinline__ bool live_head(const uint8_t *__restrict__ flags, const uint32_t *__restrict__ hasp,
                                          uint32_t s)
{
    return !(flags[s] & F_BORROW) && !has_pred(hasp, s);
}

// heads whose first run has a next run, per tile: at mlist[tile start + i],
// i < mcnt[tile], (slot - tile start) | several segments << 11 | segment
// length << 32.  Every other head's unitig is its first run.
__device__ __forceinline__ uint64_t mlist_entry(uint32_t off, bool multi, uint32_t len)
{
    return (uint64_t)off | (multi ? 1ull << 11 : 0ull) | ((uint64_t)len << 32);
}

// one walk per splitter: segment length, next splitter; per tile the sum
// of (1 << 32 | length) over heads whose unitig is a single segment (the
// other heads add their count here and their length in k_chain_tail).
// Segment records are kept where the chain kernels read them (segments that
// end at or start from a sampled splitter); heads of several runs go to the
// tile's list, written contiguously
template <int NT>
__global__ __launch_bounds__(NT) void k_seg_count(const uint64_t *__restrict__ rec, const uint32_t *__restrict__ hasp,
                                                     const uint8_t *__restrict__ flags, uint32_t m, uint32_t salt,
                                                     uint32_t shift, uint64_t *__restrict__ seg,
                                                     uint32_t *__restrict__ prevseg, uint32_t *__restrict__ list,
                                                     uint64_t *__restrict__ mlist, uint32_t *__restrict__ mcnt,
                                                     uint32_t *__restrict__ cnt, uint64_t *__restrict__ tile_sum,
                                                     uint64_t *__restrict__ tile_cnt, uint32_t n_tiles)
{
    constexpr int ITEMS = CT_TILE / NT, NW = NT / 64;  // slots per thread; waves
    __shared__ uint64_t red[NW];
    const uint32_t ti = xcd_tile(n_tiles);
    if (ti == NONE) return;
    const uint64_t base = (uint64_t)ti * CT_TILE;
    // flags and has-predecessor words of this thread's slots, all in flight
    // at once (the flag array has slack past its end, so no load is guarded)
    uint8_t f[ITEMS];
    uint32_t hw[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
        const uint64_t s = base + (uint64_t)j * NT + threadIdx.x;
        f[j] = s < m ? flags[s] : (uint8_t)F_BORROW;
        hw[j] = hasp[(s < m ? s : 0) >> 5];
    }
    // splitters; sampled ones get list slots with one atomic per workgroup
    bool hd[ITEMS], rd[ITEMS];
    uint32_t nr = 0, nruns = 0;
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
        const uint32_t s = (uint32_t)(base + (uint64_t)j * NT + threadIdx.x);
        const bool cand = (f[j] & (F_BORROW | F_RUNHEAD)) == F_RUNHEAD;
        hd[j] = cand && !((hw[j] >> (s & 31)) & 1u);
        rd[j] = cand && !hd[j] && sampled(s, salt, shift);
        nruns += cand;
        nr += rd[j];
    }
    // every splitter's first run record in flight at once
    uint64_t first[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; j++)
        first[j] = (hd[j] || rd[j]) ? rec[base + (uint64_t)j * NT + threadIdx.x] : 0ull;
    uint32_t nm = 0;
#pragma unroll
    for (int j = 0; j < ITEMS; j++) nm += hd[j] && rec_next(first[j]) != NONE;
    // one scan ranks both: sampled splitters (low half), list heads (high)
    __shared__ uint64_t wtot[NW];
    __shared__ uint32_t bbase;
    const uint64_t both = ((uint64_t)nm << 32) | nr;
    const uint64_t incl = wave_incl_scan(both);
    const int w = threadIdx.x / 64;
    if (__lane_id() == 63) wtot[w] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t t = 0;
        for (int q = 0; q < NW; q++) t += wtot[q];
        bbase = (uint32_t)t ? atomicAdd(cnt + C_LIST, (uint32_t)t) : 0u;
        mcnt[ti] = (uint32_t)(t >> 32);
    }
    __syncthreads();
    uint64_t pre = incl - both;
    for (int q = 0; q < w; q++) pre += wtot[q];
    uint32_t li = bbase + (uint32_t)pre;
    uint64_t mi = base + (pre >> 32);
    uint64_t acc = 0;
    uint32_t nsplit = 0;
    for (int j = 0; j < ITEMS; j++) {
        if (!(hd[j] || rd[j])) continue;
        const uint32_t s = (uint32_t)(base + (uint64_t)j * NT + threadIdx.x);
        nsplit++;
        uint32_t r = rec_len(first[j]), v = rec_next(first[j]), steps = 1;
        while (v != NONE && !sampled(v, salt, shift)) {
            const uint64_t rr = rec[v];
            r += rec_len(rr);
            v = rec_next(rr);
            // in-degree <= 1 bounds every walk by m runs; the cap only guards
            // the GPU against corrupted run records
            if (++steps > m) {
                atomicOr(cnt + C_ERR, 1u);
                v = NONE;
                break;
            }
        }
        if (rd[j] || v != NONE) seg[s] = (uint64_t)r | ((uint64_t)v << 32);
        if (hd[j] && rec_next(first[j]) != NONE) mlist[mi++] = mlist_entry(s - (uint32_t)base, v != NONE, r);
        if (v != NONE) prevseg[v] = s;
        if (hd[j]) {
            acc += 1ull << 32;
            if (v == NONE) acc += r;
        }
        if (rd[j]) list[li++] = s;
    }
    // per-tile sums, no global atomics: (heads << 32 | head-segment lengths)
    // and (splitters << 32 | runs)
    acc = block_reduce_n<uint64_t, NW>(acc, red);
    const uint64_t sr = block_reduce_n<uint64_t, NW>(((uint64_t)nsplit << 32) | nruns, red);
    if (threadIdx.x == 0) {
        tile_sum[ti] = acc;
        tile_cnt[ti] = sr;
    }
}

// the ranked chain entries of the sampled splitters, from one chain buffer
// to the other
__global__ void k_chain_move(const uint32_t *__restrict__ list, uint32_t n, const uint64_t *__restrict__ from,
                             uint64_t *__restrict__ to)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t r = list[i];
        to[r] = from[r];
    }
}

// The chain kernels run over the sampled splitters, whose count the device
// holds (*np): a fixed grid strides over them, so no host read of the count
// stands between the segment walks and the chain rounds.

// sampled splitters: (preceding segment, its length)
__global__ void k_chain_init(const uint32_t *__restrict__ list, const uint32_t *__restrict__ np,
                             const uint32_t *__restrict__ prevseg, const uint64_t *__restrict__ seg,
                             uint64_t *__restrict__ chain)
{
    const uint32_t n = *np;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t r = list[i], a = prevseg[r];
        chain[r] = (uint64_t)a | ((uint64_t)(uint32_t)seg[a] << 32);
    }
}

// one synchronous pointer-jumping round: (ancestor, distance) pairs; a
// round after convergence copies every pair through unchanged.  *active
// ends non-zero when some chain has not reached its head.
__global__ void k_chain_round(const uint32_t *__restrict__ list, const uint32_t *__restrict__ np,
                              const uint32_t *__restrict__ hasp, const uint64_t *__restrict__ cin,
                              uint64_t *__restrict__ cout, uint32_t *__restrict__ active)
{
    const uint32_t n = *np;
    bool act = false;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t r = list[i];
        const uint64_t c = cin[r];
        const uint32_t a = (uint32_t)c;
        if (!has_pred(hasp, a)) {
            cout[r] = c;
        } else {
            const uint64_t c2 = cin[a];
            const uint32_t a2 = (uint32_t)c2;
            cout[r] = (uint64_t)a2 | ((uint64_t)((uint32_t)(c >> 32) + (uint32_t)(c2 >> 32)) << 32);
            act |= has_pred(hasp, a2);
        }
    }
    const uint64_t bal = __ballot(act);
    if (__lane_id() == 0 && bal) atomicAdd(active, (uint32_t)__popcll(bal));
}

// the tail segment of a multi-segment unitig gives the unitig length
__global__ void k_chain_tail(const uint32_t *__restrict__ list, const uint32_t *__restrict__ np,
                             const uint32_t *__restrict__ hasp, const uint64_t *__restrict__ chain,
                             const uint64_t *__restrict__ seg, uint32_t *__restrict__ ulen,
                             uint64_t *__restrict__ tile_sum, uint32_t *__restrict__ nonconv)
{
    const uint32_t n = *np;
    bool bad = false;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t r = list[i];
        const uint64_t c = chain[r];
        const uint32_t a = (uint32_t)c;
        if (has_pred(hasp, a)) {
            bad = true;  // no head behind this chain: an isolated cycle
        } else if ((uint32_t)(seg[r] >> 32) == NONE) {
            const uint32_t len = (uint32_t)(c >> 32) + (uint32_t)seg[r];
            ulen[a] = len;
            atomicAdd((unsigned long long *)&tile_sum[a / CT_TILE], (unsigned long long)len);
        }
    }
    const uint64_t bal = __ballot(bad);
    if (__lane_id() == 0 && bal) atomicAdd(nonconv, (uint32_t)__popcll(bal));
}

// ---- isolated cycles (only when some k-mers are not reached from a head) --

// marks every run of a segment that hangs off a head
__global__ __launch_bounds__(BLOCK) void k_cover(const uint64_t *__restrict__ rec, const uint32_t *__restrict__ hasp,
                                                 uint8_t *__restrict__ flags, const uint64_t *__restrict__ chain,
                                                 uint32_t m, uint32_t salt, uint32_t shift)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= m || (flags[s] & (F_BORROW | F_RUNHEAD)) != F_RUNHEAD) return;
    if (has_pred(hasp, s)) {
        if (!sampled(s, salt, shift)) return;
        if (has_pred(hasp, (uint32_t)chain[s])) return;
    }
    uint32_t v = s, steps = 0;
    do {
        flags[v] |= F_VISIT;
        v = rec_next(rec[v]);
    } while (v != NONE && !sampled(v, salt, shift) && ++steps <= m);
}

// runs that no head reaches lie on cycles; every member becomes a run of its
// own (its character stays where it is in rch)
__global__ void k_cycle_init(const uint32_t *__restrict__ link, uint8_t *__restrict__ flags, uint32_t m, RunOut ro,
                             uint64_t *__restrict__ lab, uint32_t *__restrict__ list, uint32_t *__restrict__ n_list)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    // (members converted by their run's head carry F_CYCLE and are skipped)
    if (s >= m || (flags[s] & (F_BORROW | F_RUNHEAD | F_VISIT | F_CYCLE)) != F_RUNHEAD) return;
    const uint64_t r = ro.rec[s];
    const uint32_t len = rec_len(r), pos = rec_pos(r, s);
    uint32_t u = s;
    for (uint32_t j = 0; j < len && u != NONE; j++) {
        const uint32_t l = link[u];
        flags[u] |= F_CYCLE | F_RUNHEAD;
        // the member's character: from the record, or its byte in rch ('A' 'C' 'G' 'T' -> 0..3)
        const uint32_t ch = len <= REC_INLINE ? (rec_chars(r) >> (2 * j)) & 3u
                                              : (((uint32_t)ro.rch[pos + j] >> 1) ^ ((uint32_t)ro.rch[pos + j] >> 2)) & 3u;
        ro.rec[u] = make_rec((uint32_t)l, 1u, u, pos + j, ch);
        lab[u] = (uint64_t)u | ((uint64_t)(uint32_t)l << 32);
        list[atomicAdd(n_list, 1u)] = u;
        u = (uint32_t)l;
    }
}

// min-label pointer jumping around each cycle: (smallest k-mer so far, jump)
template <int W>
__global__ void k_cycle_round(const uint32_t *__restrict__ list, uint32_t n, const uint64_t *__restrict__ kmer,
                              const uint64_t *__restrict__ lin, uint64_t *__restrict__ lout)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t v = list[i];
    const uint64_t c = lin[v];
    const uint32_t j = (uint32_t)(c >> 32), l1 = (uint32_t)c;
    const uint64_t c2 = lin[j];
    const uint32_t l2 = (uint32_t)c2, j2 = (uint32_t)(c2 >> 32);
    uint32_t best = l1;
    if (l2 != l1 && lex_less<W>(load_kmer<W>(kmer, l2, ~0ull), load_kmer<W>(kmer, l1, ~0ull))) best = l2;
    lout[v] = (uint64_t)best | ((uint64_t)j2 << 32);
}

// predecessors of cycle vertices (the main path keeps only a bitmap)
__global__ void k_cycle_pred(const uint32_t *__restrict__ list, uint32_t n, const uint32_t *__restrict__ link,
                             uint32_t *__restrict__ pred)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t v = list[i];
    pred[(uint32_t)link[v]] = v;
}

// each cycle loses the edge into its smallest k-mer
__global__ void k_cycle_cut(const uint32_t *__restrict__ list, uint32_t n, const uint64_t *__restrict__ lab,
                            const uint32_t *__restrict__ pred, uint32_t *__restrict__ link,
                            uint64_t *__restrict__ rec, uint32_t *__restrict__ hasp, uint8_t *__restrict__ flags,
                            uint32_t *__restrict__ n_cut)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t v = list[i];
    flags[v] &= (uint8_t)~F_CYCLE;
    if ((uint32_t)lab[v] == v) {
        link[pred[v]] = NO_LINK;
        reinterpret_cast<uint32_t *>(rec)[2ull * pred[v]] = NONE;  // the record's next-run field
        atomicAnd(&hasp[v >> 5], ~(1u << (v & 31)));
        atomicAdd(n_cut, 1u);
    }
}

__global__ void k_clear_flag(uint8_t *__restrict__ flags, uint32_t m, uint8_t bit)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (uint64_t)gridDim.x * blockDim.x)
        flags[i] &= (uint8_t)~bit;
}

// ---------------------------------------------------------------------------
// 4. emission
//
// Unitig u (in order of head slot) occupies text bytes [len_<u + u*k,
// ... + len_u + k): its first k-mer in full, the last character of every
// later k-mer, '\n'.

// unitig length of a list head (entry e at tile start t0): its segment's,
// or the chain's when the unitig has several segments
__device__ __forceinline__ uint32_t mlist_len(uint64_t e, const uint32_t *__restrict__ ulen, uint64_t t0)
{
    return (e >> 11) & 1u ? ulen[t0 + (e & 0x7FFu)] : (uint32_t)(e >> 32);
}

// LDS writes of one wave visible to the whole wave (no block barrier)
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// four characters from four 2-bit codes (the low byte of b, first code in
// the low bits), first character in the low byte
__device__ __forceinline__ uint32_t chars4(uint32_t b)
{
    // code q to byte q (b << 6q puts bits 2q..2q+1 at 8q..8q+1), then one
    // byte permute picks byte `code` of "ACGT" for each
    b &= 0xFFu;  // (the caller's higher bits are later characters)
    const uint32_t sel = (b | (b << 6) | (b << 12) | (b << 18)) & 0x03030303u;
    return __builtin_amdgcn_perm(0u, 0x54474341u, sel);
}

// characters i .. i + 3 of x (i + 3 < 32 W)
template <int W>
__device__ __forceinline__ uint32_t kmer_chars4(const Kmer<W> &x, uint32_t i)
{
    const uint32_t sh = 2 * (i % 32);
    if constexpr (W == 1) {
        return chars4((uint32_t)(x.w[0] >> sh));
    } else {
        const bool hi = i >= 32;
        uint64_t v = (hi ? x.w[1] : x.w[0]) >> sh;
        if (!hi && sh > 56) v |= x.w[1] << (64 - sh);
        return chars4((uint32_t)v);
    }
}

// the k characters of x at o, whose address is `at` modulo 4: single bytes
// up to the first 4-byte boundary and after the last, 4-byte words between
// (other unitigs own the neighbouring bytes)
template <int W>
__device__ __forceinline__ void put_kmer(char *o, uint64_t at, const Kmer<W> &x, uint32_t k)
{
    uint32_t lead = (4u - (uint32_t)(at & 3u)) & 3u;
    if (lead > k) lead = k;
    uint32_t i = 0;
    for (; i < lead; i++) o[i] = base_char(char_at<W>(x, i));
    for (; i + 4 <= k; i += 4) *reinterpret_cast<uint32_t *>(o + i) = kmer_chars4<W>(x, i);
    for (; i < k; i++) o[i] = base_char(char_at<W>(x, i));
}

// byte i (0..15) of a 16-byte line
__device__ __forceinline__ char line_byte(const uint4 &v, uint32_t i)
{
    const uint32_t q = i >> 2;
    const uint32_t d = q == 0 ? v.x : q == 1 ? v.y : q == 2 ? v.z : v.w;
    return (char)(d >> (8 * (i & 3u)));
}

// rch[pos, pos + len) to o, reading whole aligned 16-byte lines (rch is
// padded to a line past its last character)
__device__ __forceinline__ void copy_chars(char *o, const char *__restrict__ rch, uint32_t pos, uint32_t len)
{
    for (uint32_t i = 0; i < len;) {
        const uint32_t q = (pos + i) & ~15u, b0 = pos + i - q;
        const uint4 v = *reinterpret_cast<const uint4 *>(rch + q);
        const uint32_t nb = 16 - b0 < len - i ? 16 - b0 : len - i;
        for (uint32_t t = 0; t < nb; t++) o[i + t] = line_byte(v, b0 + t);
        i += nb;
    }
}

// one head's segment at o (address `at` modulo 4): first k-mer x, the later
// characters of its run (record r), the characters of every later run up to
// the next sampled splitter, '\n' when the unitig ends there
// characters from .. len - 1 of the run with record r at slot s: from the
// record when the run is short, else from rch
__device__ __forceinline__ void put_run(char *o, uint64_t r, const char *__restrict__ rch, uint32_t s, uint32_t from,
                                        uint32_t len)
{
    if (len <= REC_INLINE) {
        const uint32_t c = rec_chars(r);
        for (uint32_t j = from; j < len; j++) o[j - from] = base_char((c >> (2 * j)) & 3u);
    } else {
        copy_chars(o, rch, rec_pos(r, s) + from, len - from);
    }
}

// a head's first run at o (address `at` modulo 4): its first k-mer x, the
// later characters of its run (record r), '\n' when the run is the whole
// unitig; the runs after it are written after every head's first run
template <int W>
__device__ __forceinline__ void put_first(char *o, uint64_t at, const Kmer<W> &x, uint64_t r,
                                          const char *__restrict__ rch, uint32_t s, uint32_t k)
{
    put_kmer<W>(o, at, x, k);
    const uint32_t len = rec_len(r);
    put_run(o + k, r, rch, s, 1, len);  // the head's own character is the k-th above
    if (rec_next(r) == NONE) o[k + len - 1] = '\n';
}

// LDS staging bytes per wave in k_head_emit (a range that does not fit is
// written directly): 2 KiB for one-word k-mers, whose 256-thread groups
// would lose residency with more (measured on MI355X: 2 KiB 0.95 ms, 3 KiB
// 1.07 ms, unstaged 1.05 ms on the 50 Mbp job), 4 KiB for two-word k-mers
// (2 KiB: 8.8 against 8.4 ms on the 300 Mbp instance)
template <int W>
constexpr uint32_t emit_stage() { return W == 1 ? 2048 : 4096; }

// LDS index of slot i of an emission tile (one pad word after every 8)
__device__ __forceinline__ uint32_t tix(uint32_t i) { return i + (i >> 3); }

// per tile: exclusive prefix of (1 << 32 | length) over heads -> text
// offsets.  Every head writes its first run (put_first).  The 64 slots a
// wave takes at a time own one contiguous text range; when the range fits
// the wave's LDS stage, the heads write into LDS and the range goes out as
// aligned 16-byte stores, else each head writes its bytes to HBM itself.  A
// thread's next item (record and k-mer) is loaded while it writes the
// current one.  Then the heads of several runs (the tile's list) walk their
// later runs up to the first sampled splitter, many walks in flight at once.
template <int W, bool STAGE, int NT>
__global__ __launch_bounds__(NT) void k_head_emit(const uint64_t *__restrict__ kmer, const uint64_t *__restrict__ rec,
                                                     const char *__restrict__ rch, const uint32_t *__restrict__ hasp,
                                                     const uint8_t *__restrict__ flags,
                                                     const uint64_t *__restrict__ mlist, const uint32_t *__restrict__ mcnt,
                                                     const uint32_t *__restrict__ ulen,
                                                     const uint64_t *__restrict__ tile_off, uint32_t m, uint32_t k,
                                                     uint32_t salt, uint32_t shift, uint64_t *__restrict__ htoff,
                                                     char *__restrict__ text, uint32_t n_tiles,
                                                     const uint32_t *__restrict__ gate, uint32_t n_kmers)
{
    constexpr uint32_t STG = emit_stage<W>();
    if (gate && !pass_settled(gate, n_kmers)) return;  // (uniform over the grid)
    constexpr int NW = NT / 64, CTI = CT_TILE / NT;  // waves; items per thread
    // prefixes in slot order, then the tile's end; slot i at tix(i), one pad
    // word after every 8, so that a thread's 8 consecutive slots (the scan
    // below) sit 72 bytes from the next thread's: 4-way LDS bank access, as
    // for any 64 lanes of 8-byte words, where 64 bytes put every lane on one
    // of two bank pairs
    __shared__ uint64_t tile[CT_TILE + CT_TILE / 8 + 1];
    __shared__ uint64_t wsum[NW];
    __shared__ __align__(16) char stage[STAGE ? NW : 1][STAGE ? STG : 16];
    const uint32_t ti = xcd_tile(n_tiles);
    if (ti == NONE) return;
    const uint64_t t0 = (uint64_t)ti * CT_TILE;
    uint32_t hd = 0;  // per item: live head
    // the tile's list of heads of several runs, its entries, and every
    // item's flags, predecessor bit and run record, all in flight at once
    const uint32_t nml = mcnt[ti];
    uint32_t fh[CTI];
    uint64_t hr[CTI];
#pragma unroll
    for (int j = 0; j < CTI; j++) {
        const uint64_t s = t0 + (uint64_t)j * NT + threadIdx.x;
        fh[j] = s < m ? (uint32_t)flags[s] | (((hasp[s >> 5] >> (s & 31)) & 1u) << 8) : (uint32_t)F_BORROW;
        hr[j] = s < m ? rec[s] : 0ull;
    }
    uint64_t ml[CTI];
#pragma unroll
    for (int j = 0; j < CTI; j++) {
        const uint32_t e = (uint32_t)j * NT + threadIdx.x;
        ml[j] = e < nml ? mlist[t0 + e] : 0ull;
    }
    // a live head (not borrowed, no predecessor) whose run is its unitig has
    // its length now; the list fills in the others
#pragma unroll
    for (int j = 0; j < CTI; j++) {
        uint64_t v = 0;
        if (!(fh[j] & (F_BORROW | 0x100u))) {
            hd |= 1u << j;
            if (rec_next(hr[j]) == NONE) v = (1ull << 32) | rec_len(hr[j]);
        }
        tile[tix(j * NT + threadIdx.x)] = v;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < CTI; j++) {
        const uint32_t e = (uint32_t)j * NT + threadIdx.x;
        if (e < nml) tile[tix(ml[j] & 0x7FFu)] = (1ull << 32) | mlist_len(ml[j], ulen, t0);
    }
    __syncthreads();
    uint64_t loc[CTI];
    uint64_t run = 0;
#pragma unroll
    for (int j = 0; j < CTI; j++) {
        loc[j] = run;
        run += tile[tix(threadIdx.x * CTI + j)];
    }
    const uint64_t incl = wave_incl_scan(run);
    const int w = threadIdx.x / 64;
    if (__lane_id() == 63) wsum[w] = incl;
    __syncthreads();
    uint64_t wpre = 0;
    for (int i = 0; i < w; i++) wpre += wsum[i];
    const uint64_t excl = tile_off[ti] + wpre + incl - run;
#pragma unroll
    for (int j = 0; j < CTI; j++) tile[tix(threadIdx.x * CTI + j)] = excl + loc[j];
    if (threadIdx.x == NT - 1) tile[tix(CT_TILE)] = excl + run;
    __syncthreads();
    char *stg = stage[STAGE ? w : 0];
    const uint32_t lane = __lane_id();
    const uint32_t s0 = (uint32_t)t0 + threadIdx.x;
    Kmer<W> nx{};
    uint64_t nr = 0;
    if (hd & 1u) {
        nx = load_kmer<W>(kmer, s0, ~0ull);
        nr = rec[s0];
    }
#pragma unroll 1
    for (uint32_t j = 0; j < (uint32_t)CTI; j++) {
        const uint32_t s = s0 + j * NT;
        const uint64_t r = nr;
        const Kmer<W> x = nx;
        if ((hd >> (j + 1)) & 1u) {
            nx = load_kmer<W>(kmer, s + NT, ~0ull);
            nr = rec[s + NT];
        }
        const bool h = (hd >> j) & 1u;
        const uint64_t pre = tile[tix(j * NT + threadIdx.x)];
        const uint64_t off = (pre & 0xFFFFFFFFull) + (pre >> 32) * k;
        // the wave's text range [r0, r1): from its first slot to the slot after its last
        const uint32_t li = j * NT + (uint32_t)w * 64;
        const uint64_t p0 = tile[tix(li)], p1 = tile[tix(li + 64)];
        const uint64_t r0 = (p0 & 0xFFFFFFFFull) + (p0 >> 32) * k, r1 = (p1 & 0xFFFFFFFFull) + (p1 >> 32) * k;
        const uint64_t a = r0 & ~15ull;  // staged from r0's 16-byte line
        if (STAGE && r1 - a <= STG) {  // uniform over the wave
            // (the bytes of runs after a head's first run are written by later
            // kernels over what the stage holds there)
            if (h) put_first<W>(stg + (off - a), off, x, r, rch, s, k);
            wave_sync();
            // whole lines [f, e) as 16-byte stores; the partial first and last
            // lines byte by byte, one lane per byte
            const uint64_t f = (r0 + 15) >> 4, e = r1 >> 4;
            for (uint64_t c = f + lane; c < e; c += 64)
                *reinterpret_cast<uint4 *>(text + 16 * c) = *reinterpret_cast<const uint4 *>(stg + (16 * c - a));
            const uint64_t hb = lane < 16 ? r0 + lane : 16 * e + (lane - 16);
            const uint64_t hlim = lane < 16 ? (16 * f < r1 ? 16 * f : r1) : r1;
            if (lane < 32 && hb < hlim && (lane < 16 || e >= f)) text[hb] = stg[hb - a];
            wave_sync();
        } else if (h) {
            put_first<W>(text + off, off, x, r, rch, s, k);
        }
    }
    // the later runs of the heads of several runs (the tile's list), over
    // what the waves' stores above left there: the whole tile's text range
    // belongs to this group, so one barrier orders the two
    __syncthreads();
#pragma unroll 1
    for (int j = 0; j < CTI; j++) {
        const uint32_t e = (uint32_t)j * NT + threadIdx.x;
        if (e >= nml) break;
        const uint32_t ls = (uint32_t)(ml[j] & 0x7FFu);
        const uint32_t s = (uint32_t)t0 + ls;
        const uint64_t pre = tile[tix(ls)];
        const uint64_t off = (pre & 0xFFFFFFFFull) + (pre >> 32) * k;
        uint64_t r = rec[s];
        uint64_t p = off + k + rec_len(r) - 1;
        uint32_t v = rec_next(r), steps = 0;
        while (v != NONE && !sampled(v, salt, shift) && ++steps < m) {
            r = rec[v];
            const uint32_t len = rec_len(r);
            put_run(text + p, r, rch, v, 0, len);
            p += len;
            v = rec_next(r);
        }
        if (v == NONE)
            text[p] = '\n';
        else
            htoff[s] = off;  // a head of several segments: k_rand_emit places the later ones from it
    }
}

// the segment of sampled splitter r0, placed by its ranked chain entry
__device__ __forceinline__ void rand_segment(const uint64_t *__restrict__ rec, const char *__restrict__ rch, uint32_t r0,
                                             const uint64_t *__restrict__ chain, const uint64_t *__restrict__ htoff,
                                             uint32_t k, uint32_t salt, uint32_t shift, uint32_t n_slots,
                                             char *__restrict__ text)
{
    const uint64_t c = chain[r0];
    const uint32_t a = (uint32_t)c, d = (uint32_t)(c >> 32);
    uint64_t p = htoff[a] + (k - 1) + d;
    uint32_t v = r0, steps = 0;
    do {
        const uint64_t r = rec[v];
        put_run(text + p, r, rch, v, 0, rec_len(r));
        p += rec_len(r);
        v = rec_next(r);
    } while (v != NONE && !sampled(v, salt, shift) && ++steps < n_slots);
    if (v == NONE) text[p] = '\n';
}

// segments of sampled splitters: text positions from the ranked chain (the
// list's length *np from the device; the grid strides over it)
__global__ void k_rand_emit(const uint64_t *__restrict__ rec, const char *__restrict__ rch,
                            const uint32_t *__restrict__ list, const uint32_t *__restrict__ np,
                            const uint64_t *__restrict__ chain, const uint64_t *__restrict__ htoff, uint32_t k,
                            uint32_t salt, uint32_t shift, uint32_t n_slots, char *__restrict__ text,
                            const uint32_t *__restrict__ gate, uint32_t n_kmers)
{
    if (gate && !pass_settled(gate, n_kmers)) return;
    const uint32_t n = *np;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        rand_segment(rec, rch, list[i], chain, htoff, k, salt, shift, n_slots, text);
}

// text offset of each set's first unitig: the prefix at the set's first slot
// (one workgroup per set: the heads between the start of the set's first
// slot's contraction tile and that slot)
__global__ __launch_bounds__(BLOCK) void k_set_offsets(const uint32_t *__restrict__ boff, uint32_t n_sets,
                                                       uint32_t B, const uint32_t *__restrict__ hasp,
                                                       const uint8_t *__restrict__ flags,
                                                       const uint64_t *__restrict__ rec,
                                                       const uint64_t *__restrict__ mlist,
                                                       const uint32_t *__restrict__ mcnt,
                                                       const uint32_t *__restrict__ ulen,
                                                       const uint64_t *__restrict__ tile_off, uint32_t m,
                                                       const uint64_t *__restrict__ total, uint64_t *__restrict__ out,
                                                       const uint32_t *__restrict__ gate, uint32_t n_kmers)
{
    __shared__ uint64_t red[WAVES];
    if (gate && !pass_settled(gate, n_kmers)) return;
    const uint32_t s = blockIdx.x;
    const uint32_t sb = s < n_sets ? boff[(uint64_t)s * B] : m;
    if (sb >= m) {
        if (threadIdx.x == 0) out[s] = *total;  // (unitigs << 32 | k-mers) of the whole job
        return;
    }
    const uint32_t t = sb / CT_TILE;
    const uint64_t t0 = (uint64_t)t * CT_TILE;
    uint64_t acc = 0;
    for (uint32_t x = (uint32_t)t0 + threadIdx.x; x < sb; x += BLOCK)
        if (live_head(flags, hasp, x)) {
            const uint64_t r = rec[x];
            acc += (1ull << 32) | (rec_next(r) == NONE ? rec_len(r) : 0u);
        }
    // heads of several runs: their lengths from the tile's list
    for (uint32_t e = threadIdx.x; e < mcnt[t]; e += BLOCK) {
        const uint64_t v = mlist[t0 + e];
        if (t0 + (v & 0x7FFu) < sb) acc += mlist_len(v, ulen, t0);
    }
    acc = block_reduce(acc, red);
    if (threadIdx.x == 0) out[s] = tile_off[t] + acc;
}

// successor slot of every slot (for checking); before the splice a lent
// k-mer's link holds its run head and it has no successor yet
__global__ void k_link_succ(const uint32_t *__restrict__ link, const uint8_t *__restrict__ flags, uint32_t m,
                            bool lent_heads, uint32_t *__restrict__ succ)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (uint64_t)gridDim.x * blockDim.x)
        succ[i] = lent_heads && (flags[i] & F_LENT) ? NONE : link[i];
}

// ---------------------------------------------------------------------------
// 5. k-mer text on the device [driver.cpp:257-354 load_kmers,
//    driver.cpp:227-249 compact_kmers]
//
// The text sits in HBM with at least TXT_PAD zero bytes after it.  Tiles of
// TXT_TILE bytes, 16 per thread: k_txt_count counts the newlines of the body
// per tile; after a scan, k_txt_pack finds every newline again with its
// global index j and packs line j + 1 (the line that starts after it)
// straight into the 64-bit k-mer layout.  Line 0 starts at the body.

constexpr uint32_t TXT_TILE = BLOCK * 16;
constexpr uint32_t TXT_PAD = TXT_TILE + 256;

// bit b set iff byte b of the 16 at `pos` is '\n' and lo <= pos + b < hi
__device__ __forceinline__ uint32_t nl_mask16(uint4 v, uint64_t pos, uint64_t lo, uint64_t hi)
{
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
    uint32_t m = 0;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const uint32_t t = w[j] ^ 0x0A0A0A0Au;  // zero bytes where '\n'
        const uint32_t z = ~(((t & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | t | 0x7F7F7F7Fu);  // 0x80 per zero byte
        const uint32_t nib = ((z >> 7) & 1u) | ((z >> 14) & 2u) | ((z >> 21) & 4u) | ((z >> 28) & 8u);
        m |= nib << (4 * j);
    }
    if (pos < lo) m &= lo - pos >= 16 ? 0u : (0xFFFFu << (uint32_t)(lo - pos)) & 0xFFFFu;
    if (pos + 16 > hi) m &= hi <= pos ? 0u : (0xFFFFu >> (uint32_t)(pos + 16 - hi));
    return m;
}

__global__ __launch_bounds__(BLOCK) void k_txt_count(const char *__restrict__ txt, uint64_t body, uint64_t len,
                                                     uint64_t tile0, uint32_t *__restrict__ cnt)
{
    __shared__ uint32_t red[WAVES];
    const uint64_t pos = (tile0 + blockIdx.x) * TXT_TILE + (uint64_t)threadIdx.x * 16;
    const uint4 v = *reinterpret_cast<const uint4 *>(txt + pos);
    const uint32_t c = (uint32_t)__popc(nl_mask16(v, pos, body, len));
    const uint32_t t = block_reduce(c, red);
    if (threadIdx.x == 0) cnt[blockIdx.x] = t;
}

__device__ __forceinline__ uint32_t txt_code(char ch)
{
    // driver.cpp:228-232: A 0, C 1, G 2, T 3, every other byte 0
    return ch == 'C' ? 1u : ch == 'G' ? 2u : ch == 'T' ? 3u : 0u;
}

// line `idx` starting at `start`: its first k characters, or an error when
// the line (without one '\r' before its end) is shorter than k
template <int W>
__device__ __forceinline__ void pack_line(const char *__restrict__ txt, uint64_t len, uint64_t start, uint64_t idx,
                                          uint32_t k, uint64_t *__restrict__ out, uint32_t *__restrict__ err)
{
    uint64_t w[2] = {0, 0};
    bool bad = false;
    for (uint32_t i = 0; i < k; i++) {
        const uint64_t p = start + i;
        const char ch = p < len ? txt[p] : '\n';
        if (ch == '\n') {
            bad = true;
            break;
        }
        w[i >> 5] |= (uint64_t)txt_code(ch) << (2 * (i & 31));
    }
    if (!bad) {
        const char nx = start + k < len ? txt[start + k] : '\n';
        if (nx == '\n' && txt[start + k - 1] == '\r') bad = true;
    }
    if (bad) atomicOr(err, 1u);
#pragma unroll
    for (int j = 0; j < W; j++) out[idx * W + j] = w[j];
}

template <int W>
__global__ __launch_bounds__(BLOCK) void k_txt_pack(const char *__restrict__ txt, uint64_t body, uint64_t len,
                                                    uint64_t tile0, const uint32_t *__restrict__ base, uint64_t n,
                                                    uint32_t k, uint64_t *__restrict__ out, uint32_t *__restrict__ err)
{
    __shared__ uint32_t wsum[WAVES];
    const uint64_t pos = (tile0 + blockIdx.x) * TXT_TILE + (uint64_t)threadIdx.x * 16;
    const uint4 v = *reinterpret_cast<const uint4 *>(txt + pos);
    uint32_t m = nl_mask16(v, pos, body, len);
    const uint32_t c = (uint32_t)__popc(m);
    const uint32_t incl = wave_incl_scan(c);
    const int w = threadIdx.x / 64;
    if (__lane_id() == 63) wsum[w] = incl;
    __syncthreads();
    uint64_t j = base[blockIdx.x] + incl - c;
    for (int q = 0; q < w; q++) j += wsum[q];
    if (blockIdx.x == 0 && threadIdx.x == 0 && n > 0) pack_line<W>(txt, len, body, 0, k, out, err);
    while (m) {
        const uint32_t b = (uint32_t)__builtin_ctz(m);
        m &= m - 1;
        if (j + 1 < n) pack_line<W>(txt, len, pos + b + 1, j + 1, k, out, err);
        j++;
    }
}

// ---------------------------------------------------------------------------
// host side

// grid of whole XCD rounds covering n_tiles tiles (see xcd_tile)
static uint32_t xcd_grid(uint64_t n_tiles)
{
    return (uint32_t)(N_XCD * ((n_tiles + N_XCD - 1) / N_XCD));
}

static uint32_t grid_for(uint64_t n, uint32_t block = BLOCK)
{
    uint64_t g = (n + block - 1) / block;
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(g, 0x7FFFFFFFull));
}

struct DBuf {
    void *p = nullptr;
    size_t cap = 0;
};

// Scratch of uc_compact_noalloc: one array per device, reserved with the
// code object and held until the process ends.  Host code never passes the
// array's name to a kernel or a copy: the name is a host-side stand-in, and
// the device address comes from hipGetSymbolAddress (scratch_ptr below).
static __device__ __align__(256) unsigned char scratch[UC_SCRATCH_BYTES];

// bytes at the end of the scratch that no buffer is carved from: the few
// kernels that read a whole aligned line past a buffer's end stay inside
constexpr size_t SCRATCH_TAIL = 64 << 10;

// UC_FLAG_LOW_MEMORY: one device block per context, kept from job to job.
// Buffers are carved from it first-fit in address order (256-byte aligned)
// and given back at their last use, so phases reuse each other's memory and
// a job after the first calls no hipMalloc (a fresh allocation of tens of GB
// costs seconds).  Free ranges coalesce.
struct Pool {
    char *base = nullptr;
    size_t cap = 0, used = 0;
    std::map<size_t, size_t> free_;  // offset -> bytes, disjoint, not adjacent
};

}  // namespace uc

using namespace uc;

constexpr size_t RBUF_BYTES = 64u << 10;  // bytes of a context's pinned read buffer

struct uc_ctx {
    uc_opts opts;
    hipStream_t stream = nullptr;
    std::string err;
    size_t held = 0, held_peak = 0;  // device bytes held now / most held at once since the job began
    uc_meter *meter = nullptr;  // a server's account of its workers' device bytes

    // job
    uint32_t k = 0, l = 0, W = 1, B = 64, n_sets = 0;
    uint64_t n = 0, m = 0, x = 0, largest = 0;
    uint64_t top_mask = ~0ull, pm0 = 0, pm1 = 0;
    std::vector<uint64_t> set_off;   // n_sets + 1
    std::vector<uint32_t> boff_h;    // n_sets*B + 1 coarse bucket rows
    std::vector<uint32_t> set_sub;   // fine buckets per coarse bucket, per set
    std::vector<uint32_t> cb_first;  // first fine bucket of every coarse bucket
    std::vector<uint32_t> fboff_h;   // fine bucket rows
    bool loaded = false, bucketed = false, edged = false, coarsened = false, emitted = false;
    bool spliced = false;  // lent k-mers' links hold successors, no longer run heads
    bool edges_ready = false;
    std::vector<uint64_t> et_toff, et_first;
    std::vector<uint32_t> et_cap;
    std::vector<uint8_t> bucket_done;
    std::vector<size_t> grp_first, fb_first, bg_first, sl_first;  // per coarse bucket: first LDS group / HBM-table
                                                                 // bucket / big fine bucket / slice task
    uint32_t lds_nb = 64;                     // LDS capacity of the edge launch
    bool rows_valid = false;                  // boff_h / fboff_h mirror the device rows
    uint32_t lds_limit = 65536;               // dynamic LDS a workgroup may use on this device
    uint32_t tile1 = 0, tile2 = 0;            // k-mers per bucketing tile, slots per fine-sort tile
    uint64_t n_fallback = 0, n_groups = 0, n_big = 0;
    uint64_t table_slots = 0, n_done = 0;
    uint64_t n_unitigs = 0, n_cycles = 0, text_bytes = 0, n_splitters = 0, n_runs = 0;
    uint32_t chain_rounds = 0;
    std::vector<uint64_t> set_text;  // n_sets + 1
    uc_timing timing{};

    // device buffers
    const uint64_t *kin = nullptr;  // k-mers in 64-bit layout (the caller's buffer or `in`)
    DBuf otable, in, stage, bkt, tiles, hist, xhist, kmer2, flags2, src2, sub, perm, tiles2, hist2, cbhb, cbts,
        cbfirst, fboff, egroups, fbbase, scan_tmp, kmer, flags, src, pair_h, pair_b, boff, etiles,
        toff, tcap, table, hit, link, hasp, pred, seg, prevseg, chainA, chainB, cycB, list, clist, ulen, tsum,
        htoff, setoff, text, small, txt, txtcnt, rec, rch, tcnt, cbtf, over, colsum, bgroups, gfirst, stasks,
        mlist, mcnt, probes;
    // pinned host chunks of the file reader and writer
    // pinned bounce buffer of the small device reads (counters, offsets):
    // a copy into pinned memory completes sooner than one into pageable
    // memory, which HIP stages (measured on MI355X: 14.8 against 22.0 us per
    // read, launch and wait included)
    char *rbuf = nullptr;
    char *rbuf_dev = nullptr;  // its device mapping
    // copies into page-locked host memory by k_host_copy (a server's workers)
    bool kernel_copies = false;
    // a slice table of this bucketing overflowed: its sliced fine buckets
    // take HBM tables (set when the edge phase runs again for that)
    bool no_slices = false;
    uint32_t slice_probes = SLICE_TAB;  // most probes of one slice-table insert
    // a contraction pass whose counters the host has not read yet (the text
    // buffer is sized for the most text the job can have meanwhile)
    bool deferred = false;
    // a low-memory emission wrote head text offsets over the head lists
    bool lists_spent = false;
    // whole small jobs: bucketing leaves the placement totals on the device
    // (lazy_sizes), and m and x hold their bounds until the edge plan's read
    // (sizes_pending)
    bool lazy_sizes = false, sizes_pending = false;
    char *pin[2] = {nullptr, nullptr};
    hipEvent_t pin_ev[2] = {nullptr, nullptr};
    bool pin_used[2] = {false, false};
    // pinned pieces of large copies from and to pageable host memory
    char *spin[2] = {nullptr, nullptr};
    hipEvent_t spin_ev[2] = {nullptr, nullptr};
    bool spin_used[2] = {false, false};
    // the captured launches of a contraction pass, replayed while the key
    // (sizes, parameters, buffer addresses) stays the same
    hipGraphExec_t cgraph = nullptr;
    std::vector<uint64_t> cgraph_key;
    bool capturing = false;
    // contraction state
    uint64_t *chain = nullptr;  // ranked chain entries (chainA or chainB)
    uint32_t n_list = 0, salt = 0, shift = 5;
    hipEvent_t ev[6] = {};

    // per-kernel timing (UC_FLAG_PROFILE)
    struct KStat {
        std::string name;
        double ms = 0;
        uint64_t launches = 0;
    };
    std::vector<KStat> kstats;
    std::string prof_only;  // time only this kernel ("" for all)
    std::vector<std::pair<size_t, std::pair<hipEvent_t, hipEvent_t>>> pending;
    std::vector<hipEvent_t> evpool;

    // uc_compact_noalloc: the device address of this device's scratch
    // array; during such a call every buffer is carved from it (arena) and
    // the text goes to the caller's buffer (text_fixed)
    unsigned char *scratch_dev = nullptr;
    unsigned char *arena = nullptr;
    size_t arena_cap = 0, arena_used = 0;
    bool text_fixed = false;

    // UC_FLAG_LOW_MEMORY
    Pool pool;
    size_t pool_target = 0;  // bytes the pool should have for the current job (an upper bound)
};

// every device buffer of a context
static std::vector<DBuf *> all_bufs(uc_ctx *c)
{
    return {&c->kmer2,  &c->flags2,  &c->src2,   &c->sub,     &c->perm,    &c->tiles2, &c->hist2,
            &c->cbhb,   &c->cbts,    &c->cbfirst, &c->fboff,  &c->egroups, &c->fbbase, &c->otable,
            &c->in,     &c->stage,   &c->bkt,    &c->tiles,   &c->hist,    &c->xhist,  &c->scan_tmp,
            &c->kmer,   &c->flags,   &c->src,    &c->pair_h,  &c->pair_b,  &c->boff,
            &c->etiles,  &c->toff,   &c->tcap,    &c->table,   &c->hit,    &c->link,
            &c->hasp,   &c->pred,    &c->seg,    &c->prevseg, &c->chainA, &c->chainB,
            &c->cycB,   &c->list,    &c->clist,  &c->ulen,    &c->tsum,    &c->htoff,  &c->setoff,
            &c->text,   &c->small,   &c->txt,    &c->txtcnt,  &c->rec,     &c->rch,
            &c->tcnt,   &c->cbtf,    &c->over,   &c->colsum, &c->bgroups, &c->gfirst, &c->stasks,
            &c->mlist,  &c->mcnt,    &c->probes};
}

static void free_pinned(uc_ctx *c)
{
    for (int i = 0; i < 2; i++) {
        if (c->pin[i]) (void)hipHostFree(c->pin[i]);
        if (c->pin_ev[i]) (void)hipEventDestroy(c->pin_ev[i]);
        c->pin[i] = nullptr;
        c->pin_ev[i] = nullptr;
        c->pin_used[i] = false;
        if (c->spin[i]) (void)hipHostFree(c->spin[i]);
        if (c->spin_ev[i]) (void)hipEventDestroy(c->spin_ev[i]);
        c->spin[i] = nullptr;
        c->spin_ev[i] = nullptr;
        c->spin_used[i] = false;
    }
}

// tile-count argument of the tiled kernels: the top bit asks for the
// XCD-contiguous renumbering of xcd_tile
static uint32_t tile_arg(const uc_ctx *c, uint64_t n_tiles)
{
    return (uint32_t)n_tiles | ((c->opts.flags & UC_FLAG_XCD_MAP) ? 0x80000000u : 0u);
}

// XCD-contiguous tile order whatever the flags: the scatters, whose
// neighbouring tiles write neighbouring bytes of every bucket (one L2 then
// merges the partial lines; measured 1.9 -> 1.3 GB and 1.4 -> 0.7 GB
// written by the two passes on 48 M k-mers)
static uint32_t tile_arg_xcd(uint64_t n_tiles)
{
    return (uint32_t)n_tiles | 0x80000000u;
}

namespace {

int fail(uc_ctx *c, int code, const std::string &msg)
{
    if (c) c->err = msg;
    return code;
}

#define HIPCHK(ctx, expr)                                                                         \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            return fail(ctx, UC_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));       \
    } while (0)

#define CHK(expr)                    \
    do {                             \
        int rc_ = (expr);            \
        if (rc_ != UC_OK) return rc_; \
    } while (0)

// buffers a low-memory context carves from its pool: all but those of the
// text and table loads (their sizes come before the job's), which are
// allocated and freed on their own
static bool pooled(const uc_ctx *c, const DBuf &b)
{
    return (c->opts.flags & UC_FLAG_LOW_MEMORY) && !c->arena && &b != &c->stage && &b != &c->txt &&
           &b != &c->txtcnt && &b != &c->otable;
}

// Pool placement: buffers that live until the job's text is written are
// carved upward from the bottom, every other one downward from the top, so
// that when emission asks for the text the space above the survivors is one
// free range.
static bool keeps(const uc_ctx *c, const DBuf &b)
{
    for (const DBuf *k : {&c->kmer, &c->flags, &c->hasp, &c->rec, &c->rch, &c->mlist, &c->mcnt, &c->ulen,
                          &c->chainA, &c->list, &c->tsum, &c->text, &c->small, &c->boff, &c->fboff, &c->cbfirst,
                          &c->setoff, &c->probes, &c->htoff})
        if (k == &b) return true;
    return false;
}

static bool in_pool(const uc_ctx *c, const void *p)
{
    return c->pool.base && p >= c->pool.base && p < c->pool.base + c->pool.cap;
}

static size_t round256(size_t b) { return (b + 255) & ~(size_t)255; }

void note_peak(uc_ctx *c)
{
    c->held_peak = std::max(c->held_peak, c->held - c->pool.cap + c->pool.used);
}

// the pool, sized for the current job once per context (again when a larger
// job would get a larger block); everything free on the device but a margin
// at most, within hbm_budget
int pool_reserve(uc_ctx *c)
{
    Pool &P_ = c->pool;
    size_t fr = 0, tot = 0;
    HIPCHK(c, hipMemGetInfo(&fr, &tot));
    const size_t margin = std::max<size_t>(2ull << 30, tot / 64);
    auto room = [&](size_t extra) {
        size_t r = fr + extra > margin ? fr + extra - margin : 0;
        r = std::min(r, std::max<size_t>(c->pool_target, 64ull << 20));
        if (c->opts.hbm_budget) {
            const size_t other = c->held - P_.cap;
            r = std::min(r, c->opts.hbm_budget > other ? c->opts.hbm_budget - other : 0);
        }
        return r & ~(size_t)255;
    };
    if (P_.base) {
        // an idle pool grows when a larger job would get at least an eighth more
        if (P_.used || room(P_.cap) <= P_.cap + P_.cap / 8) return UC_OK;
        HIPCHK(c, hipFree(P_.base));
        c->held -= P_.cap;
        if (c->meter) c->meter->sub(P_.cap);
        P_ = Pool{};
        HIPCHK(c, hipMemGetInfo(&fr, &tot));
    }
    const size_t cap = room(0);
    if (cap < SCRATCH_TAIL + 4096)
        return fail(c, UC_E_NOMEM, "no device memory for the low-memory pool (" + std::to_string(fr) + " free)");
    void *p = nullptr;
    const hipError_t e = hipMalloc(&p, cap);
    if (e != hipSuccess)
        return fail(c, UC_E_NOMEM, "hipMalloc of the " + std::to_string(cap) + "-byte pool: " + hipGetErrorString(e));
    P_.base = static_cast<char *>(p);
    P_.cap = cap;
    P_.used = 0;
    P_.free_.clear();
    P_.free_[0] = cap - SCRATCH_TAIL;  // a tail no buffer gets: reads of a line past an end stay inside
    c->held += cap;
    if (c->meter) c->meter->add(cap);
    return UC_OK;
}

void pool_release(uc_ctx *c, DBuf &b)
{
    Pool &P_ = c->pool;
    size_t off = (size_t)(static_cast<char *>(b.p) - P_.base), len = round256(b.cap);
    P_.used -= len;
    auto nx = P_.free_.lower_bound(off);
    if (nx != P_.free_.end() && off + len == nx->first) {  // merge with the next range
        len += nx->second;
        nx = P_.free_.erase(nx);
    }
    if (nx != P_.free_.begin()) {
        auto pv = std::prev(nx);
        if (pv->first + pv->second == off) {  // and with the previous one
            pv->second += len;
            b.p = nullptr;
            b.cap = 0;
            return;
        }
    }
    P_.free_[off] = len;
    b.p = nullptr;
    b.cap = 0;
}

int pool_carve(uc_ctx *c, DBuf &b, size_t bytes)
{
    CHK(pool_reserve(c));
    Pool &P_ = c->pool;
    const size_t len = round256(bytes);
    auto take = [&](size_t off) {
        b.p = P_.base + off;
        b.cap = bytes;
        P_.used += len;
        note_peak(c);
        return UC_OK;
    };
    if (keeps(c, b)) {  // lowest range that fits, its low end
        for (auto it = P_.free_.begin(); it != P_.free_.end(); ++it) {
            if (it->second < len) continue;
            const size_t off = it->first, rest = it->second - len;
            P_.free_.erase(it);
            if (rest) P_.free_[off + len] = rest;
            return take(off);
        }
    } else {  // highest range that fits, its high end
        for (auto it = P_.free_.rbegin(); it != P_.free_.rend(); ++it) {
            if (it->second < len) continue;
            it->second -= len;
            const size_t off = it->first + it->second;
            if (it->second == 0) P_.free_.erase(std::next(it).base());
            return take(off);
        }
    }
    return fail(c, UC_E_NOMEM, "the low-memory pool of " + std::to_string(P_.cap) + " bytes has no free range of " +
                                   std::to_string(len) + " (" + std::to_string(P_.used) + " in use)");
}

int ensure(uc_ctx *c, DBuf &b, size_t bytes)
{
    if (bytes == 0) bytes = 8;
    if (b.cap >= bytes) return UC_OK;
    if (c->arena) {
        // uc_compact_noalloc: carve from the scratch, never allocate.  A
        // buffer asked for again with more bytes gets a new piece; contents
        // are never carried over (nor are they across hipMalloc below).
        const size_t off = (c->arena_used + 255) & ~(size_t)255;
        if (off > c->arena_cap || bytes > c->arena_cap - off)
            return fail(c, UC_E_LIMIT, "job needs more than the " + std::to_string(c->arena_cap) +
                                           " bytes of device scratch uc_compact_noalloc may use");
        b.p = c->arena + off;
        b.cap = bytes;
        c->arena_used = off + bytes;
        return UC_OK;
    }
    if (pooled(c, b)) {
        if (b.p) pool_release(c, b);  // (contents are not carried over)
        return pool_carve(c, b, bytes);
    }
    if (b.p) {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        HIPCHK(c, hipFree(b.p));
        c->held -= b.cap;
        if (c->meter) c->meter->sub(b.cap);
        b.p = nullptr;
        b.cap = 0;
    }
    if (c->opts.hbm_budget && c->held + bytes > c->opts.hbm_budget)
        return fail(c, UC_E_NOMEM, "device budget exceeded: need " + std::to_string(c->held + bytes) + " bytes");
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess) {
        b.p = nullptr;
        return fail(c, UC_E_NOMEM, std::string("hipMalloc of ") + std::to_string(bytes) + " bytes with " +
                                       std::to_string(c->held) + " held: " + hipGetErrorString(e));
    }
    b.cap = bytes;
    c->held += bytes;
    note_peak(c);
    if (c->meter) c->meter->add(bytes);
    return UC_OK;
}

// UC_FLAG_LOW_MEMORY: give a buffer back once the job is past its last use.
// A pool range is free for the next carve at once: work queued on the stream
// after this point is ordered after every kernel that used it.  Other
// buffers go back with hipFree, which waits for the device.
int drop(uc_ctx *c, DBuf &b)
{
    if (!(c->opts.flags & UC_FLAG_LOW_MEMORY) || c->arena || !b.p) return UC_OK;
    if (in_pool(c, b.p)) {
        pool_release(c, b);
        return UC_OK;
    }
    HIPCHK(c, hipFree(b.p));
    c->held -= b.cap;
    if (c->meter) c->meter->sub(b.cap);
    b.p = nullptr;
    b.cap = 0;
    return UC_OK;
}

template <typename T>
T *P(DBuf &b)
{
    return reinterpret_cast<T *>(b.p);
}

static RunOut run_out(uc_ctx *c)
{
    return RunOut{P<uint64_t>(c->rec), P<char>(c->rch), P<uint8_t>(c->flags)};
}

// Times one kernel launch with a pair of events on the context's stream
// when UC_FLAG_PROFILE is set; a no-op otherwise.
struct Prof {
    uc_ctx *c;
    size_t id = 0;
    hipEvent_t a = nullptr, b = nullptr;
    Prof(uc_ctx *ctx, const char *name) : c(ctx)
    {
        if (!(c->opts.flags & UC_FLAG_PROFILE) || c->capturing || c->arena ||
            (!c->prof_only.empty() && c->prof_only != name)) {
            c = nullptr;
            return;
        }
        for (id = 0; id < c->kstats.size() && c->kstats[id].name != name; id++) {
        }
        if (id == c->kstats.size()) c->kstats.push_back({name, 0.0, 0});
        a = take();
        b = take();
        if (!a || !b) {
            c = nullptr;
            return;
        }
        (void)hipEventRecord(a, c->stream);
    }
    hipEvent_t take()
    {
        hipEvent_t e = nullptr;
        if (!c->evpool.empty()) {
            e = c->evpool.back();
            c->evpool.pop_back();
        } else if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) {
            // timing only: no cache writeback and invalidation at every kernel boundary
            e = nullptr;
        }
        return e;
    }
    ~Prof()
    {
        if (!c) return;
        (void)hipEventRecord(b, c->stream);
        c->pending.push_back({id, {a, b}});
    }
};

// folds finished launch timings into kstats
int prof_flush(uc_ctx *c)
{
    if (c->pending.empty()) return UC_OK;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (auto &p : c->pending) {
        float ms = 0;
        HIPCHK(c, hipEventElapsedTime(&ms, p.second.first, p.second.second));
        c->kstats[p.first].ms += ms;
        c->kstats[p.first].launches++;
        c->evpool.push_back(p.second.first);
        c->evpool.push_back(p.second.second);
    }
    c->pending.clear();
    return UC_OK;
}

#define KLAUNCH(c, name, kernel, cfg, ...)                                  \
    do {                                                                    \
        Prof pf_(c, name);                                                  \
        hipLaunchKernelGGL(kernel, UC_CFG cfg, (c)->stream, __VA_ARGS__);   \
    } while (0);                                                            \
    HIPCHK(c, hipGetLastError())
#define UC_CFG(g, b, s) dim3(g), dim3(b), (s)

template <typename T>
int fill(uc_ctx *c, T *p, uint64_t n, T v)
{
    if (n == 0) return UC_OK;
    KLAUNCH(c, "k_fill", k_fill<T>, (std::min<uint32_t>(grid_for(n), 65536), BLOCK, 0), p, n, v);
    return UC_OK;
}

// dynamic LDS above the default 64 KiB must be asked for per kernel
template <typename K>
int allow_lds(uc_ctx *c, K kernel, size_t bytes)
{
    if (bytes <= 65536) return UC_OK;
    if (bytes > c->lds_limit) return fail(c, UC_E_LIMIT, "kernel needs more LDS than the device has");
    HIPCHK(c, hipFuncSetAttribute(reinterpret_cast<const void *>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)bytes));
    return UC_OK;
}

// exclusive scan of n values (in == out allowed); *d_total gets the sum.
// Block sums of every recursion level live in `tmp`.
template <typename T>
size_t scan_scratch(uint64_t n)
{
    size_t bytes = 0;
    for (uint64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE; nb > 1; nb = (nb + SCAN_TILE - 1) / SCAN_TILE)
        bytes += ((nb + 1) * sizeof(T) + 255) / 256 * 256;
    return bytes;
}

template <typename T>
int scan_level(uc_ctx *c, const T *in, T *out, uint64_t n, T *d_total, T *d_total2, char *tmp)
{
    const uint64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
    if (nb <= 1) {
        KLAUNCH(c, "k_scan_tile", k_scan_tile<T>, (1, BLOCK, 0), in, out, n, nullptr, d_total, d_total2);
        return UC_OK;
    }
    T *partial = reinterpret_cast<T *>(tmp);
    char *next = tmp + ((nb + 1) * sizeof(T) + 255) / 256 * 256;
    KLAUNCH(c, "k_scan_reduce", k_scan_reduce<T>, ((uint32_t)nb, BLOCK, 0), in, n, partial);
    CHK(scan_level<T>(c, partial, partial, nb, nullptr, nullptr, next));
    KLAUNCH(c, "k_scan_tile", k_scan_tile<T>, ((uint32_t)nb, BLOCK, 0), in, out, n, partial, d_total, d_total2);
    return UC_OK;
}

// exclusive scan of n values; the total goes to d_total and, when given, to
// d_total2 as well (a counter word the host reads)
template <typename T>
int scan_excl(uc_ctx *c, const T *in, T *out, uint64_t n, T *d_total, DBuf &tmp, T *d_total2 = nullptr)
{
    if (n == 0) {
        if (d_total) CHK(fill<T>(c, d_total, 1, T(0)));
        if (d_total2) CHK(fill<T>(c, d_total2, 1, T(0)));
        return UC_OK;
    }
    CHK(ensure(c, tmp, scan_scratch<T>(n) + 256));
    return scan_level<T>(c, in, out, n, d_total, d_total2, reinterpret_cast<char *>(tmp.p));
}

int select_dev(uc_ctx *c)
{
    HIPCHK(c, hipSetDevice(c->opts.device));
    return UC_OK;
}

// k_host_copy on the context's stream; `dst` or `src` is the device mapping
// of page-locked host memory
constexpr uint32_t HOST_COPY_GRID = 32;

int host_copy(uc_ctx *c, void *dst, const void *src, uint64_t bytes)
{
    if (bytes == 0) return UC_OK;
    const uint32_t g = (uint32_t)std::min<uint64_t>(HOST_COPY_GRID, (bytes + 16ull * BLOCK - 1) / (16ull * BLOCK));
    KLAUNCH(c, "k_host_copy", k_host_copy, (g, BLOCK, 0), static_cast<const char *>(src), static_cast<char *>(dst),
            bytes);
    return UC_OK;
}

// the device mapping of page-locked host memory from hipHostMalloc, or
// nullptr
void *host_mapping(const void *p)
{
    void *d = nullptr;
    if (hipHostGetDevicePointer(&d, const_cast<void *>(p), 0) != hipSuccess) d = nullptr;
    (void)hipGetLastError();
    return d;
}

int d2h(uc_ctx *c, void *dst, const void *src, size_t bytes)
{
    if (bytes == 0) return UC_OK;
    if (c->rbuf && bytes <= RBUF_BYTES) {
        if (c->kernel_copies && c->rbuf_dev)
            CHK(host_copy(c, c->rbuf_dev, src, bytes));
        else
            HIPCHK(c, hipMemcpyAsync(c->rbuf, src, bytes, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        std::memcpy(dst, c->rbuf, bytes);
        return UC_OK;
    }
    HIPCHK(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return UC_OK;
}

// host mirrors of the coarse and fine bucket rows, fetched when asked for
int host_rows(uc_ctx *c)
{
    if (c->rows_valid) return UC_OK;
    const uint64_t G = (uint64_t)c->n_sets * c->B, F = c->cb_first.empty() ? 0 : c->cb_first[G];
    c->boff_h.resize(G + 1);
    c->fboff_h.resize(F + 1);
    CHK(d2h(c, c->boff_h.data(), c->boff.p, (G + 1) * 4));
    CHK(d2h(c, c->fboff_h.data(), c->fboff.p, (F + 1) * 4));
    c->largest = 0;
    for (uint64_t g = 0; g < G; g++) c->largest = std::max<uint64_t>(c->largest, c->boff_h[g + 1] - c->boff_h[g]);
    c->rows_valid = true;
    return UC_OK;
}

int h2d(uc_ctx *c, void *dst, const void *src, size_t bytes)
{
    if (bytes == 0) return UC_OK;
    HIPCHK(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    return UC_OK;
}

// UC_FLAG_LOW_MEMORY: a new load gives back every buffer of the last job
// first (its results are no longer readable once a new job is loaded)
int drop_all(uc_ctx *c)
{
    if (!(c->opts.flags & UC_FLAG_LOW_MEMORY) || c->arena) return UC_OK;
    // (a captured contraction pass stays: its key holds the buffer addresses,
    // which the pool hands out again in the same order)
    for (DBuf *b : all_bufs(c)) CHK(drop(c, *b));
    c->kin = nullptr;
    return UC_OK;
}

void reset_job(uc_ctx *c)
{
    c->loaded = c->bucketed = c->edged = c->coarsened = c->emitted = false;
    c->edges_ready = c->no_slices = c->deferred = c->lists_spent = false;
    c->lazy_sizes = c->sizes_pending = false;
    c->n_unitigs = c->n_cycles = c->text_bytes = c->n_splitters = 0;
    c->chain_rounds = 0;
    c->m = c->x = c->largest = 0;
    c->rows_valid = false;
}

// ---- load -----------------------------------------------------------------

// k, word layout, masks and set layout of a job
int begin_job(uc_ctx *c, const uint64_t *set_offsets, uint32_t n_sets, uint32_t k)
{
    if (k < 2 || k > 64) return fail(c, UC_E_ARG, "k must be in [2, 64]");
    if (n_sets == 0) return fail(c, UC_E_ARG, "n_sets must be >= 1");
    c->k = k;
    c->W = k <= 32 ? 1 : 2;
    c->n_sets = n_sets;
    c->set_off.assign(set_offsets, set_offsets + n_sets + 1);
    for (uint32_t s = 0; s < n_sets; s++)
        if (c->set_off[s + 1] < c->set_off[s]) return fail(c, UC_E_ARG, "set offsets must be non-decreasing");
    if (c->set_off[0] != 0) return fail(c, UC_E_ARG, "set offsets must start at 0");
    c->n = c->set_off[n_sets];
    if (c->n >= 0xFFFFFFFFull) return fail(c, UC_E_LIMIT, "more than 2^32-2 k-mers");
    c->pool_target = uc_job_bytes_estimate(c->n, k);
    const uint32_t W = c->W;
    const uint32_t top_bits = 2 * k - 64 * (W - 1);
    c->top_mask = top_bits >= 64 ? ~0ull : ((1ull << top_bits) - 1);
    const uint32_t pbits = 2 * (k - 1);
    if (pbits >= 64) {
        c->pm0 = ~0ull;
        c->pm1 = pbits - 64 >= 64 ? ~0ull : ((1ull << (pbits - 64)) - 1);
    } else {
        c->pm0 = (1ull << pbits) - 1;
        c->pm1 = 0;
    }
    return UC_OK;
}

constexpr uint64_t PIN_CHUNK = 64ull << 20;  // bytes per pinned host chunk

int pinned(uc_ctx *c)
{
    for (int i = 0; i < 2; i++) {
        if (!c->pin[i]) {
            HIPCHK(c, hipHostMalloc((void **)&c->pin[i], PIN_CHUNK, hipHostMallocDefault));
            c->pin_used[i] = false;
        }
        if (!c->pin_ev[i]) HIPCHK(c, hipEventCreateWithFlags(&c->pin_ev[i], hipEventDisableTiming));
    }
    return UC_OK;
}

// Large copies from the device to pageable host memory go through two pinned
// pieces of STAGE_STEP bytes: the DMA of one piece overlaps the host copy of
// the other (25 MB of text: 1.2 ms against 1.8 ms for hipMemcpy from the
// device into pageable memory on MI355X).  Small ones take hipMemcpy
// directly.  Host-to-device copies measured no faster this way.
constexpr uint64_t STAGE_STEP = 4ull << 20;
constexpr uint64_t STAGE_MIN = 1ull << 20;

int stage_pinned(uc_ctx *c)
{
    for (int i = 0; i < 2; i++) {
        if (!c->spin[i]) {
            HIPCHK(c, hipHostMalloc((void **)&c->spin[i], STAGE_STEP, hipHostMallocDefault));
            c->spin_used[i] = false;
        }
        if (!c->spin_ev[i]) HIPCHK(c, hipEventCreateWithFlags(&c->spin_ev[i], hipEventDisableTiming));
    }
    return UC_OK;
}

int staged_d2h(uc_ctx *c, void *dst, const void *src, uint64_t bytes)
{
    if (bytes < STAGE_MIN) return d2h(c, dst, src, bytes);
    CHK(stage_pinned(c));
    const uint64_t np = (bytes + STAGE_STEP - 1) / STAGE_STEP;
    auto issue = [&](uint64_t i) -> int {
        const int b = (int)(i & 1);
        const uint64_t off = i * STAGE_STEP, step = std::min<uint64_t>(STAGE_STEP, bytes - off);
        if (c->spin_used[b]) HIPCHK(c, hipEventSynchronize(c->spin_ev[b]));
        HIPCHK(c, hipMemcpyAsync(c->spin[b], static_cast<const char *>(src) + off, step, hipMemcpyDeviceToHost,
                                 c->stream));
        HIPCHK(c, hipEventRecord(c->spin_ev[b], c->stream));
        c->spin_used[b] = true;
        return UC_OK;
    };
    CHK(issue(0));
    for (uint64_t i = 0; i < np; i++) {
        if (i + 1 < np) CHK(issue(i + 1));  // into the other piece, copied out already
        const int b = (int)(i & 1);
        HIPCHK(c, hipEventSynchronize(c->spin_ev[b]));
        const uint64_t off = i * STAGE_STEP;
        std::memcpy(static_cast<char *>(dst) + off, c->spin[b], std::min<uint64_t>(STAGE_STEP, bytes - off));
    }
    return UC_OK;
}

int do_load(uc_ctx *c, const uint32_t *packed, const uint64_t *set_offsets, uint32_t n_sets, uint32_t k, int on_device)
{
    CHK(drop_all(c));
    reset_job(c);
    CHK(begin_job(c, set_offsets, n_sets, k));
    const uint32_t rw = (k + 15) / 16;
    const uint32_t W = c->W;
    const size_t in_bytes = (size_t)c->n * rw * 4;
    if (rw == 2 * W) {
        // the reference packing with an even word count is the 64-bit layout
        if (on_device) {
            c->kin = reinterpret_cast<const uint64_t *>(packed);
        } else {
            CHK(ensure(c, c->in, in_bytes));
            CHK(h2d(c, c->in.p, packed, in_bytes));
            c->kin = P<uint64_t>(c->in);
        }
    } else {
        const uint32_t *src = packed;
        if (!on_device) {
            CHK(ensure(c, c->stage, in_bytes));
            CHK(h2d(c, c->stage.p, packed, in_bytes));
            src = P<uint32_t>(c->stage);
        }
        CHK(ensure(c, c->in, (size_t)c->n * W * 8));
        if (c->n) {
            KLAUNCH(c, "k_widen", k_widen, (std::min<uint32_t>(grid_for(c->n), 65536), BLOCK, 0), src, c->n, rw, W,
                                                                                        P<uint64_t>(c->in));
        }
        c->kin = P<uint64_t>(c->in);
    }
    c->loaded = true;
    return UC_OK;
}

// ---- k-mer text on the device ----------------------------------------------

// parse the text held in c->txt (len bytes, TXT_PAD zero bytes after it)
template <int W>
int parse_text(uc_ctx *c, uint64_t len, uint32_t k)
{
    const char *txt = P<char>(c->txt);
    // line 1: the k-mer count (driver.cpp:296-304), read in 4 KiB steps
    std::vector<char> h;
    uint64_t got = 0, nl = len;
    while (got < len && nl == len) {
        const uint64_t step = std::min<uint64_t>(len - got, 4096);
        h.resize(got + step);
        CHK(d2h(c, h.data() + got, txt + got, step));
        const void *q = std::memchr(h.data() + got, '\n', step);
        if (q) nl = (uint64_t)((const char *)q - h.data());
        got += step;
    }
    // the count is read, but every line is taken whatever it says: the
    // reference compares it with the k-mers read only after whole read
    // chunks (driver.cpp:343) and reads on past it otherwise
    uint64_t p = 0;
    bool digit = false;
    while (p < got && (h[p] == ' ' || h[p] == '\t')) p++;
    while (p < got && h[p] >= '0' && h[p] <= '9') {
        digit = true;
        p++;
    }
    if (!digit) return fail(c, UC_E_FORMAT, "no k-mer count on the first line");
    const uint64_t body = nl < len ? nl + 1 : len;
    // newlines and carriage returns at the end of the text end it
    uint64_t end = len;
    while (end > body) {
        const uint64_t step = std::min<uint64_t>(end - body, 4096);
        std::vector<char> tail(step);
        CHK(d2h(c, tail.data(), txt + end - step, step));
        uint64_t i = step;
        while (i > 0 && (tail[i - 1] == '\n' || tail[i - 1] == '\r')) i--;
        end -= step - i;
        if (i > 0) break;
    }
    uint64_t n = 0, nt = 0, tile0 = 0;
    if (body < end) {
        tile0 = body / TXT_TILE;
        nt = (end + TXT_TILE - 1) / TXT_TILE - tile0;
        CHK(ensure(c, c->txtcnt, (nt + 2) * 4));
        uint32_t *cnt = P<uint32_t>(c->txtcnt);
        KLAUNCH(c, "k_txt_count", k_txt_count, ((uint32_t)nt, BLOCK, 0), txt, body, end, tile0, cnt);
        CHK(scan_excl<uint32_t>(c, cnt, cnt, nt, cnt + nt, c->scan_tmp));
        uint32_t total = 0;
        CHK(d2h(c, &total, cnt + nt, 4));
        n = (uint64_t)total + 1;  // the last line ends at `end`, not at a newline
    }
    const uint64_t offs[2] = {0, n};
    CHK(begin_job(c, offs, 1, k));
    CHK(ensure(c, c->in, std::max<uint64_t>(n, 1) * W * 8));
    if (n) {
        uint32_t *cnt = P<uint32_t>(c->txtcnt);
        CHK(fill<uint32_t>(c, cnt + nt + 1, 1, 0u));
        KLAUNCH(c, "k_txt_pack", k_txt_pack<W>, ((uint32_t)nt, BLOCK, 0), txt, body, end, tile0, cnt, n, k,
                P<uint64_t>(c->in), cnt + nt + 1);
        uint32_t err = 0;
        CHK(d2h(c, &err, cnt + nt + 1, 4));
        if (err) return fail(c, UC_E_FORMAT, "a k-mer line is shorter than k");
    }
    c->kin = P<uint64_t>(c->in);
    c->loaded = true;
    return UC_OK;
}

int parse_loaded_text(uc_ctx *c, uint64_t len, uint32_t k)
{
    if (k < 2 || k > 64) return fail(c, UC_E_ARG, "k must be in [2, 64]");
    CHK(fill<uint8_t>(c, P<uint8_t>(c->txt) + len, TXT_PAD, (uint8_t)0));
    CHK(k <= 32 ? parse_text<1>(c, len, k) : parse_text<2>(c, len, k));
    // a low-memory context gives the text back once it is packed
    CHK(drop(c, c->txt));
    return drop(c, c->txtcnt);
}

int do_load_text(uc_ctx *c, const char *text, uint64_t len, uint32_t k, int on_device)
{
    CHK(drop_all(c));
    reset_job(c);
    CHK(ensure(c, c->txt, len + TXT_PAD));
    if (len) {
        if (on_device)
            HIPCHK(c, hipMemcpyAsync(c->txt.p, text, len, hipMemcpyDeviceToDevice, c->stream));
        else
            CHK(h2d(c, c->txt.p, text, len));
    }
    return parse_loaded_text(c, len, k);
}

// `len` bytes of an open file, from its current position, into device
// memory.  READ_THREADS readers each take every READ_THREADS-th piece of
// PIN_CHUNK / READ_THREADS bytes, read it with pread into their own part of
// one of the two pinned chunks, and copy it to the device on the context's
// stream; a reader reuses its part of a chunk once the copy out of it is done.
// (From the page cache of MI355X's host one reader took 22-26 GB/s, four
// 64-72 GB/s, above the 56 GB/s of the copies.)  UC_E_FORMAT when the file ends
// early.  Returns with the copies done.
constexpr int READ_THREADS = 4;

int stream_file(uc_ctx *c, FILE *f, uint64_t len, char *dst, const char *path)
{
    CHK(pinned(c));
    for (int b = 0; b < 2; b++)
        if (c->pin_used[b]) HIPCHK(c, hipEventSynchronize(c->pin_ev[b]));
    const long pos = std::ftell(f);
    if (pos < 0) return fail(c, UC_E_IO, std::string("could not read ") + path);
    const int fd = fileno(f);
    const uint64_t piece = PIN_CHUNK / READ_THREADS, np = (len + piece - 1) / piece;
    const int T = (int)std::min<uint64_t>(READ_THREADS, std::max<uint64_t>(np, 1));
    std::vector<hipEvent_t> ev(2 * (size_t)T, nullptr);
    int rc = UC_OK;
    for (auto &e : ev)
        if (rc == UC_OK && hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) rc = UC_E_HIP;
    std::atomic<int> err{rc};
    const int dev = c->opts.device;
    auto reader = [&](int t) {
        if (hipSetDevice(dev) != hipSuccess) {
            err = UC_E_HIP;
            return;
        }
        bool used[2] = {false, false};
        for (uint64_t i = (uint64_t)t, r = 0; i < np && err.load() == UC_OK; i += (uint64_t)T, r++) {
            const int b = (int)(r & 1);
            char *buf = c->pin[b] + (uint64_t)t * piece;
            if (used[b] && hipEventSynchronize(ev[2 * t + b]) != hipSuccess) {
                err = UC_E_HIP;
                break;
            }
            const uint64_t off = i * piece, step = std::min<uint64_t>(piece, len - off);
            uint64_t got = 0;
            while (got < step) {
                const ssize_t n = pread(fd, buf + got, step - got, (off_t)(pos + off + got));
                if (n < 0 && errno == EINTR) continue;
                if (n <= 0) {
                    int want = UC_OK;
                    err.compare_exchange_strong(want, n == 0 ? UC_E_FORMAT : UC_E_IO);
                    break;
                }
                got += (uint64_t)n;
            }
            if (got < step) break;
            if (hipMemcpyAsync(dst + off, buf, step, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
                hipEventRecord(ev[2 * t + b], c->stream) != hipSuccess) {
                err = UC_E_HIP;
                break;
            }
            used[b] = true;
        }
    };
    if (rc == UC_OK) {
        // a reader whose thread cannot be had runs on this one, after the first
        std::vector<std::thread> th;
        std::vector<int> here{0};
        for (int t = 1; t < T; t++) {
            try {
                th.emplace_back(reader, t);
            } catch (...) {
                here.push_back(t);
            }
        }
        for (int t : here) reader(t);
        for (auto &x : th) x.join();
        rc = err.load();
    }
    const bool synced = hipStreamSynchronize(c->stream) == hipSuccess;
    for (auto e : ev)
        if (e) (void)hipEventDestroy(e);
    c->pin_used[0] = c->pin_used[1] = false;
    if (rc == UC_E_FORMAT) return fail(c, rc, std::string("short read of ") + path);
    if (rc == UC_E_IO) return fail(c, rc, std::string("could not read ") + path);
    if (rc != UC_OK || !synced) return fail(c, UC_E_HIP, "copying the file to the device");
    return UC_OK;
}

int do_load_file(uc_ctx *c, const char *path, uint32_t k)
{
    CHK(drop_all(c));
    reset_job(c);
    FILE *f = std::fopen(path, "rb");
    if (!f) return fail(c, UC_E_IO, std::string("could not open input file ") + path);
    uint64_t len = 0;
    if (std::fseek(f, 0, SEEK_END) == 0) {
        const long sz = std::ftell(f);
        if (sz >= 0) len = (uint64_t)sz;
    }
    std::rewind(f);
    int rc = ensure(c, c->txt, len + TXT_PAD);
    if (rc == UC_OK) rc = stream_file(c, f, len, P<char>(c->txt), path);
    std::fclose(f);
    if (rc != UC_OK) return rc;
    return parse_loaded_text(c, len, k);
}

int do_load_packed_file(uc_ctx *c, const char *path, uint32_t *k_out)
{
    CHK(drop_all(c));
    reset_job(c);
    uc_packed_header h;
    const int hr = uc_read_packed_header(path, &h);
    if (hr == UC_E_IO) return fail(c, UC_E_IO, std::string("could not open input file ") + path);
    if (hr != UC_OK) return fail(c, hr, std::string("not a packed k-mer file: ") + path);
    const uint64_t offs[2] = {0, h.n};
    CHK(begin_job(c, offs, 1, h.k));
    const uint32_t rw = h.words, W = c->W;
    const uint64_t bytes = h.n * rw * 4;
    // even word counts are the 64-bit layout already; odd ones are widened
    DBuf &dst = rw == 2 * W ? c->in : c->stage;
    CHK(ensure(c, dst, std::max<uint64_t>(bytes, 8)));
    FILE *f = std::fopen(path, "rb");
    if (!f) return fail(c, UC_E_IO, std::string("could not open input file ") + path);
    int rc = std::fseek(f, (long)sizeof(uc_packed_header), SEEK_SET) == 0 ? UC_OK : fail(c, UC_E_IO, "seek");
    if (rc == UC_OK) rc = stream_file(c, f, bytes, P<char>(dst), path);
    std::fclose(f);
    if (rc != UC_OK) return rc;
    if (rw != 2 * W) {
        CHK(ensure(c, c->in, std::max<uint64_t>(h.n, 1) * W * 8));
        if (h.n)
            KLAUNCH(c, "k_widen", k_widen, (std::min<uint32_t>(grid_for(h.n), 65536), BLOCK, 0), P<uint32_t>(c->stage),
                    h.n, rw, W, P<uint64_t>(c->in));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->kin = P<uint64_t>(c->in);
    c->loaded = true;
    if (k_out) *k_out = h.k;
    return UC_OK;
}

// the text appended to a file: the copy of chunk i+1 overlaps the write of
// chunk i.  The file's blocks for the text are reserved first without
// changing its size (fallocate, FALLOC_FL_KEEP_SIZE; skipped where the file
// system has no such call): appending 0.68 GB into the page cache then took
// 13.1-13.9 GB/s on MI355X's host, against 11.3-11.5 GB/s without.
int append_text(uc_ctx *c, const char *path)
{
    const int fd = ::open(path, O_WRONLY | O_CREAT | O_APPEND | O_CLOEXEC, 0666);
    if (fd < 0) return fail(c, UC_E_IO, std::string("could not open output file ") + path);
    int rc = pinned(c);
    const uint64_t N = c->text_bytes, nch = (N + PIN_CHUNK - 1) / PIN_CHUNK;
    struct stat st;
    if (rc == UC_OK && N && ::fstat(fd, &st) == 0 && S_ISREG(st.st_mode))
        (void)::fallocate(fd, FALLOC_FL_KEEP_SIZE, st.st_size, (off_t)N);
    auto issue = [&](uint64_t i) -> int {
        const int b = (int)(i & 1);
        const uint64_t off = i * PIN_CHUNK, step = std::min<uint64_t>(PIN_CHUNK, N - off);
        HIPCHK(c, hipMemcpyAsync(c->pin[b], P<char>(c->text) + off, step, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipEventRecord(c->pin_ev[b], c->stream));
        c->pin_used[b] = true;
        return UC_OK;
    };
    if (rc == UC_OK && nch) rc = issue(0);
    for (uint64_t i = 0; rc == UC_OK && i < nch; i++) {
        if (i + 1 < nch) rc = issue(i + 1);
        if (rc != UC_OK) break;
        const int b = (int)(i & 1);
        if (hipEventSynchronize(c->pin_ev[b]) != hipSuccess) {
            rc = fail(c, UC_E_HIP, "hipEventSynchronize");
            break;
        }
        const uint64_t step = std::min<uint64_t>(PIN_CHUNK, N - i * PIN_CHUNK);
        for (uint64_t put = 0; put < step;) {
            const ssize_t w = ::write(fd, c->pin[b] + put, step - put);
            if (w < 0 && errno == EINTR) continue;
            if (w <= 0) {
                rc = fail(c, UC_E_IO, std::string("could not write ") + path);
                break;
            }
            put += (uint64_t)w;
        }
    }
    if (::close(fd) != 0 && rc == UC_OK) rc = fail(c, UC_E_IO, std::string("could not write ") + path);
    (void)hipStreamSynchronize(c->stream);
    return rc;
}

// ---- bucketing ---------------------------------------------------------------

// Every k-mer gets one home slot and every crossing k-mer one more, so the
// placements are m = n + x exactly.  The 32-bit histogram scan wraps when
// that reaches 2^32; the count is therefore taken from n and x (x <= n, a
// 32-bit tile scan of crossing flags) and checked before any buffer is sized
// or any slot is written.
int check_placements(uc_ctx *c, uint64_t n, uint64_t x)
{
    const int rc = uc_check_placements(n, x);
    if (rc == UC_E_LIMIT)
        return fail(c, rc, "more than 2^32-2 placements (" + std::to_string(n + x) + "): slot ids are 32-bit");
    if (rc != UC_OK) return fail(c, rc, "crossing count above the k-mer count");
    return UC_OK;
}

constexpr uint64_t FINE_CAPPED_MAX = 1700;

template <int W>
int run_bucketize(uc_ctx *c, uint32_t l, const uint32_t *table)
{
    const uint32_t B = c->B, S = c->n_sets;
    const uint32_t W_ = W;
    // fine buckets per coarse bucket, per set: about FINE_TARGET slots each
    const uint64_t fine_target = c->opts.fine_slots ? c->opts.fine_slots : (W_ == 1 ? 512 : 384);
    c->set_sub.assign(S, 1);
    bool any_fine = false;
    for (uint32_t s = 0; s < S; s++) {
        const uint64_t ns = c->set_off[s + 1] - c->set_off[s];
        const uint64_t want = (ns * 112 / 100 + (uint64_t)B * fine_target - 1) / ((uint64_t)B * fine_target);
        uint32_t sub = 1;
        while (sub < want && sub < MAX_SUB) sub <<= 1;
        // A set that wants more fine buckets than MAX_SUB gets the fewest that
        // still average at most FINE_CAPPED_MAX slots: the fine sort then writes
        // longer pieces of each fine bucket per round, and the buckets still
        // fit LDS groups (50 Mbp job: 512 instead of 1024 fine buckets per
        // coarse bucket, 3.91 against 3.96 ms per step on MI355X).  The plain
        // form keeps the count (its per-bucket launches measured 5.81 against
        // 5.31 ms with the fewer, fuller fine buckets).
        if (!c->opts.fine_slots && want > MAX_SUB && c->opts.mode != UC_MODE_PLAIN)
            while (sub > 1 && ns * 112 / 100 <= (uint64_t)B * (sub / 2) * FINE_CAPPED_MAX) sub >>= 1;
        c->set_sub[s] = sub;
        any_fine |= sub > 1;
    }
    const bool keep_src = (c->opts.flags & UC_FLAG_KEEP_SRC) != 0;
    c->tile2 = 0;
    // k-mers per bucketing tile
    constexpr int IPT1 = W == 1 ? 8 : 4;
    const uint32_t tile1 = IPT1 * BLOCK;
    c->tile1 = tile1;
    // tiles: every set gets at least one (possibly empty) tile so each bucket
    // row has a start in the histogram; the tiles are filled on the device
    std::vector<uint64_t> set_plan(3 * (uint64_t)S + 2);  // set offsets, tile prefix, then the sets' fine counts
    std::vector<uint64_t> set_hbase(S);
    std::vector<uint32_t> set_ts(S);
    uint64_t T = 0;
    for (uint32_t s = 0; s < S; s++) {
        const uint64_t ns = c->set_off[s + 1] - c->set_off[s];
        const uint32_t ts = (uint32_t)std::max<uint64_t>(1, (ns + tile1 - 1) / tile1);
        set_plan[s] = c->set_off[s];
        set_plan[S + 1 + s] = T;
        set_hbase[s] = (uint64_t)B * T;
        set_ts[s] = ts;
        T += ts;
    }
    set_plan[S] = c->set_off[S];
    set_plan[2 * (uint64_t)S + 1] = T;
    uint32_t *sub_h = reinterpret_cast<uint32_t *>(set_plan.data() + 2 * (uint64_t)S + 2);
    for (uint32_t s = 0; s < S; s++) sub_h[s] = c->set_sub[s];
    const uint64_t hb = (uint64_t)B * T;
    const uint64_t G = (uint64_t)S * B;
    // every coarse bucket's first fine bucket (from the sets' fine counts)
    c->cb_first.assign(G + 1, 0);
    for (uint64_t g = 0; g < G; g++) c->cb_first[g + 1] = c->cb_first[g] + c->set_sub[g / B];
    const uint64_t F = c->cb_first[G];
    // the host's plan in one upload, in 8-byte words: the coarse buckets'
    // first fine buckets (G + 1 u32), the set plan, the sets' histogram bases
    // (u64) and tile counts (u32); cbfirst lives on to emission
    const uint64_t w_cbf = (G + 2) / 2, w_plan = set_plan.size(), w_hb = S, w_ts = (S + 1) / 2;
    std::vector<uint64_t> up(w_cbf + w_plan + w_hb + w_ts, 0);
    std::memcpy(up.data(), c->cb_first.data(), (G + 1) * 4);
    std::memcpy(up.data() + w_cbf, set_plan.data(), w_plan * 8);
    std::memcpy(up.data() + w_cbf + w_plan, set_hbase.data(), S * 8);
    std::memcpy(up.data() + w_cbf + w_plan + w_hb, set_ts.data(), S * 4);
    CHK(ensure(c, c->cbfirst, up.size() * 8));
    CHK(h2d(c, c->cbfirst.p, up.data(), up.size() * 8));
    const uint64_t *sp = P<uint64_t>(c->cbfirst) + w_cbf;
    const uint64_t *sethb = sp + w_plan;
    const uint32_t *setts = reinterpret_cast<const uint32_t *>(sethb + w_hb);
    CHK(ensure(c, c->tiles, T * sizeof(BTile)));
    KLAUNCH(c, "k_plan1_tiles", k_plan1_tiles, (grid_for(T), BLOCK, 0), sp, sp + S + 1,
            reinterpret_cast<const uint32_t *>(sp + 2 * (uint64_t)S + 2), S, B, tile1, T, P<BTile>(c->tiles));
    CHK(ensure(c, c->hist, (hb + 1) * 4));
    CHK(ensure(c, c->xhist, (T + 1) * 4));

    LminParams p;
    p.k = c->k;
    p.l = l;
    p.B = B;
    const uint32_t bits = 2 * l;
    p.lmask = bits >= 32 ? 0xFFFFFFFFu : ((1u << bits) - 1u);
    p.h1 = (bits + 1) / 2;
    p.h2 = bits / 3 + 1;
    p.s0 = (uint32_t)c->opts.seed & p.lmask;
    p.w1 = bits - p.h1;
    p.w2 = bits - p.h2;
    p.k1 = p.s0 * MIX_C1;
    p.table = table;

    const uint64_t *kin = c->kin;
    // both fine codes of a k-mer in one word when every fine index fits 16 bits
    uint32_t smax = 1;
    for (uint32_t sb : c->set_sub) smax = std::max(smax, sb);
    const bool c32 = (uint64_t)B * smax <= 65536;
    CHK(ensure(c, c->bkt, std::max<uint64_t>(c->n, 1) * (c32 ? 4 : 8)));
#define UC_CNT(MODE, C32)                                                                                          \
    KLAUNCH(c, "k_lmin_count", (k_lmin_count<W, IPT1, MODE, C32>), (xcd_grid(T), BLOCK, (B + 1) * 4), kin,        \
            P<BTile>(c->tiles), p, c->top_mask, c->bkt.p, P<uint32_t>(c->hist), P<uint32_t>(c->xhist), tile_arg_xcd(T))
    switch (lmin_mode(l, table != nullptr)) {
    case LM_TABLE:
        if (c32) { UC_CNT(LM_TABLE, true); } else { UC_CNT(LM_TABLE, false); }
        break;
    case LM_MUL24:
        if (c32) { UC_CNT(LM_MUL24, true); } else { UC_CNT(LM_MUL24, false); }
        break;
    default:
        if (c32) { UC_CNT(LM_MUL32, true); } else { UC_CNT(LM_MUL32, false); }
        break;
    }
#undef UC_CNT
    // totals at hist[hb] and xhist[T], and side by side in the counters for one read
    CHK(ensure(c, c->small, C_WORDS * 4));
    uint32_t *cnt = P<uint32_t>(c->small);
    CHK(scan_excl<uint32_t>(c, P<uint32_t>(c->hist), P<uint32_t>(c->hist), hb, P<uint32_t>(c->hist) + hb, c->scan_tmp,
                            cnt + C_TOT));
    CHK(scan_excl<uint32_t>(c, P<uint32_t>(c->xhist), P<uint32_t>(c->xhist), T, P<uint32_t>(c->xhist) + T,
                            c->scan_tmp, cnt + C_TOT + 1));
    // A whole small job (lazy_sizes) does not read the totals here: the
    // buffers and grids below take the bounds x <= n and m <= 2n, the kernels
    // that need m or x read them on the device, and the host reads them with
    // the edge plan.
    const bool lazy = c->lazy_sizes;
    const uint32_t *mdev = lazy ? cnt + C_TOT : nullptr;
    if (lazy) {
        c->x = c->n;
        c->m = 2 * c->n;
        c->sizes_pending = true;
    } else {
        uint32_t tot[2];
        CHK(d2h(c, tot, cnt + C_TOT, 8));
        CHK(check_placements(c, c->n, tot[1]));
        c->x = tot[1];
        c->m = c->n + c->x;
        c->sizes_pending = false;
    }
    // pass 1 writes the final k-mers and flags when no set is split into
    // fine buckets, else the k-mers and the fine index with the flag bits
    DBuf &k1 = any_fine ? c->kmer2 : c->kmer, &s1 = any_fine ? c->src2 : c->src;
    CHK(ensure(c, k1, std::max<uint64_t>(c->m, 1) * W * 8));
    if (any_fine)
        CHK(ensure(c, c->sub, c->m * 2 + 8));
    else
        CHK(ensure(c, c->flags, c->m + 64));
    if (keep_src) CHK(ensure(c, s1, c->m * 4));
    CHK(ensure(c, c->pair_h, c->x * 4));
    CHK(ensure(c, c->pair_b, c->x * 4));
    CHK(ensure(c, c->boff, (G + 1) * 4));
    const uint32_t nbits = (uint32_t)__builtin_ctz(B);
    // 512 threads per 2048-k-mer tile (measured against 256 and 1024: the
    // staged form holds four tiles per CU, and 8 waves each hide more latency)
    constexpr int NT1 = 512;
    const bool stage1 = !(c->opts.flags & UC_FLAG_DIRECT_SCATTER) &&
                        lmin_stage_bytes<W>(B, tile1, keep_src) <= c->lds_limit;
#define UC_S1(ST, C32)                                                                                              \
    do {                                                                                                           \
        const size_t lds = ST ? lmin_stage_bytes<W>(B, tile1, keep_src) : lmin_lds_bytes(B, tile1 / 64);              \
        CHK(allow_lds(c, k_lmin_scatter<W, IPT1 * BLOCK / NT1, ST, NT1, C32>, lds));                                \
        KLAUNCH(c, "k_lmin_scatter", (k_lmin_scatter<W, IPT1 * BLOCK / NT1, ST, NT1, C32>), (xcd_grid(T), NT1, lds), \
                kin, P<BTile>(c->tiles), B, nbits, c->top_mask, c->bkt.p, P<uint32_t>(c->hist),                     \
                P<uint32_t>(c->xhist), P<uint64_t>(k1), any_fine ? nullptr : P<uint8_t>(c->flags),                  \
                keep_src ? P<uint32_t>(s1) : nullptr, P<uint32_t>(c->pair_h), P<uint32_t>(c->pair_b),               \
                any_fine ? P<uint16_t>(c->sub) : nullptr, tile_arg_xcd(T));                                         \
    } while (0)
    if (stage1) {
        if (c32) UC_S1(true, true);
        else UC_S1(true, false);
    } else {
        if (c32) UC_S1(false, true);
        else UC_S1(false, false);
    }
#undef UC_S1
    // the fine rows are the coarse rows when no set has a fine sort
    CHK(ensure(c, c->fboff, (F + 1) * 4));
    KLAUNCH(c, "k_gather_boff", k_gather_boff, (grid_for(G + 1), BLOCK, 0), P<uint32_t>(c->hist),
            sethb, setts, S, B, (uint32_t)c->m, mdev, P<uint32_t>(c->boff), any_fine ? nullptr : P<uint32_t>(c->fboff));
    c->rows_valid = false;
    if (c->opts.flags & UC_FLAG_LOW_MEMORY) {
        // the codes, the input and the tile histogram are read by pass 1 and
        // the row gather only; a loaded input is consumed
        CHK(drop(c, c->bkt));
        CHK(drop(c, c->hist));
        CHK(drop(c, c->xhist));
        if (c->kin == P<uint64_t>(c->in)) {
            CHK(drop(c, c->in));
            c->kin = nullptr;
            c->loaded = false;
        }
    }

    // fine buckets: stable sort of each coarse bucket's slots by fine index,
    // planned and run on the device (no host round trip)
    if (any_fine) {
        // slots per fine-sort round; LDS staging when the device has the room
        // (2048-slot rounds: half the LDS of 4096, so more tiles in flight per
        // CU; measured 0.82 -> 0.71 ms for the staged one-word scatter).  A
        // planned tile is SUB_ROUNDS rounds.
        constexpr int IPT2 = 8;
        // (1, 2 and 4 rounds measured 0.73, 0.71 and 0.68 ms for the count,
        // column scan and scatter together on the 50 Mbp job: the scatter takes
        // the same time whatever the rounds, the histogram work shrinks)
        const uint32_t round2 = IPT2 * BLOCK, tile2 = SUB_ROUNDS * round2;
        // (two-word k-mers measured faster unstaged here: 4.9 vs 5.5 ms on
        // 303 M slots; one-word ones staged: 0.82 vs 0.85 ms on 53 M)
        // threads per 2048-slot tile: 512 for the staged one-word form, 1024
        // for the direct two-word form (each measured best of 256/512/1024)
        constexpr int NT2 = W == 1 ? 512 : 1024;
        const bool stage2 = W == 1 && !(c->opts.flags & UC_FLAG_DIRECT_SCATTER) &&
                            sub_lds_bytes<W>(round2, smax, keep_src, true, NT2 / 64) <= c->lds_limit;
        c->tile2 = tile2;
        // tiles: at most one partial tile per coarse bucket beyond m / tile2
        const uint64_t T2 = G + (c->m + tile2 - 1) / tile2;
        const uint64_t h2 = (uint64_t)smax * T2;  // histogram entries, upper bound
        CHK(ensure(c, c->tiles2, T2 * sizeof(BTile)));
        CHK(ensure(c, c->hist2, (h2 + 1) * 4));
        CHK(ensure(c, c->cbhb, (G + 1) * 8));
        CHK(ensure(c, c->cbts, (G + 1) * 4));
        CHK(ensure(c, c->cbtf, (G + 1) * 4));
        KLAUNCH(c, "k_plan2_prefix", k_plan2_prefix, (1, PLAN_NT, 0), P<uint32_t>(c->boff), P<uint32_t>(c->cbfirst),
                (uint32_t)G, tile2, P<uint32_t>(c->cbts), P<uint32_t>(c->cbtf), P<uint64_t>(c->cbhb));
        KLAUNCH(c, "k_plan2_tiles", k_plan2_tiles, (grid_for(T2), BLOCK, 0), P<uint32_t>(c->boff),
                P<uint32_t>(c->cbfirst), P<uint32_t>(c->cbts), P<uint32_t>(c->cbtf), P<uint64_t>(c->cbhb), (uint32_t)G,
                tile2, (uint32_t)T2, P<BTile>(c->tiles2));
        KLAUNCH(c, "k_sub_count", k_sub_count, ((uint32_t)T2, SUB_COUNT_NT, smax * 4), P<uint16_t>(c->sub),
                P<BTile>(c->tiles2), P<uint32_t>(c->hist2));
        const uint32_t cpb = (smax + 63) / 64;  // column blocks per coarse bucket
        CHK(ensure(c, c->colsum, F * 4 + 4));
        KLAUNCH(c, "k_sub_colscan", k_sub_colscan, ((uint32_t)(G * cpb), BLOCK, 0), P<uint32_t>(c->hist2),
                P<uint64_t>(c->cbhb), P<uint32_t>(c->cbts), P<uint32_t>(c->cbfirst), cpb, P<uint32_t>(c->colsum));
        KLAUNCH(c, "k_sub_rows", k_sub_rows, ((uint32_t)G, BLOCK, 0), P<uint32_t>(c->colsum), P<uint32_t>(c->cbfirst),
                P<uint32_t>(c->boff), (uint32_t)G, (uint32_t)c->m, mdev, P<uint32_t>(c->fboff));
        CHK(ensure(c, c->kmer, std::max<uint64_t>(c->m, 1) * W * 8));
        CHK(ensure(c, c->flags, c->m + 64));
        if (keep_src) CHK(ensure(c, c->src, c->m * 4));
        CHK(ensure(c, c->perm, c->m * 4));
#define UC_S2(ST)                                                                                                   \
    do {                                                                                                           \
        const size_t lds2 = sub_lds_bytes<W>(round2, smax, keep_src, ST, NT2 / 64);                                 \
        CHK(allow_lds(c, k_sub_scatter<W, ST, IPT2 * BLOCK / NT2, NT2>, lds2));                                    \
        KLAUNCH(c, "k_sub_scatter", (k_sub_scatter<W, ST, IPT2 * BLOCK / NT2, NT2>), (xcd_grid(T2), NT2, lds2),      \
                P<BTile>(c->tiles2), P<uint16_t>(c->sub), P<uint32_t>(c->hist2), P<uint32_t>(c->fboff),             \
                P<uint32_t>(c->cbfirst), P<uint64_t>(c->kmer2), keep_src ? P<uint32_t>(c->src2) : nullptr,          \
                P<uint64_t>(c->kmer), P<uint8_t>(c->flags), keep_src ? P<uint32_t>(c->src) : nullptr,               \
                P<uint32_t>(c->perm), tile_arg_xcd(T2));                                                           \
    } while (0)
        if (stage2) UC_S2(true);
        else UC_S2(false);
#undef UC_S2
        if (c->x)
            KLAUNCH(c, "k_pair_remap", k_pair_remap, (grid_for(c->x), BLOCK, 0), P<uint32_t>(c->pair_h),
                    P<uint32_t>(c->pair_b), (uint32_t)c->x, lazy ? cnt + C_TOT + 1 : nullptr, P<uint32_t>(c->perm));
        // pass 2's input and the permutation are dead now
        CHK(drop(c, c->kmer2));
        CHK(drop(c, c->sub));
        CHK(drop(c, c->src2));
        CHK(drop(c, c->perm));
        CHK(drop(c, c->hist2));
    }
    c->l = l;
    c->bucketed = true;
    return UC_OK;
}

int do_bucketize(uc_ctx *c, uint32_t l, const uint32_t *table)
{
    if (!c->loaded) return fail(c, UC_E_STATE, "uc_bucketize before uc_load");
    if (l < 1 || l > 15 || l >= c->k) return fail(c, UC_E_ARG, "l must be in [1, min(15, k-1)]");
    c->edged = c->coarsened = c->emitted = false;
    c->edges_ready = c->no_slices = false;
    return c->W == 1 ? run_bucketize<1>(c, l, table) : run_bucketize<2>(c, l, table);
}

// ---- edges -----------------------------------------------------------------

// entries of a table holding `count` keys at load_pct percent
uint64_t table_cap(uint64_t count, uint32_t load_pct)
{
    return std::max<uint64_t>(64, (count * 100 + load_pct - 1) / load_pct);
}

int clear_table(uc_ctx *c, TEntry *t, uint64_t n)
{
    if (n == 0) return UC_OK;
    KLAUNCH(c, "k_table_clear", k_table_clear, (std::min<uint32_t>(grid_for(n), 65536), BLOCK, 0), t, n);
    return UC_OK;
}

// the edge phase's probe counters, when they are kept
static unsigned long long *probe_counter(uc_ctx *c)
{
    if (!(c->opts.flags & UC_FLAG_COUNT_PROBES)) return nullptr;
    return reinterpret_cast<unsigned long long *>(c->probes.p);
}

static int clear_probes(uc_ctx *c)
{
    if (!(c->opts.flags & UC_FLAG_COUNT_PROBES)) return UC_OK;
    CHK(ensure(c, c->probes, PROBE_SLOTS * 8));
    return fill<uint32_t>(c, P<uint32_t>(c->probes), 2 * PROBE_SLOTS, 0u);
}

template <int W, bool CNT>
int launch_global_edges_(uc_ctx *c, size_t first, size_t count)
{
    if (count == 0) return UC_OK;
    const ETile *tiles = P<ETile>(c->etiles) + first;
    KLAUNCH(c, "k_edge_insert", (k_edge_insert<W, CNT>), (xcd_grid(count), BLOCK, 0), P<uint64_t>(c->kmer),
            P<uint8_t>(c->flags), tiles, P<uint32_t>(c->fbbase), P<uint64_t>(c->toff), P<uint32_t>(c->tcap),
            P<TEntry>(c->table), c->k, c->pm0, c->pm1, tile_arg(c, count), probe_counter(c));
    KLAUNCH(c, "k_edge_lookup", (k_edge_lookup<W, CNT>), (xcd_grid(count), BLOCK, 0), P<uint64_t>(c->kmer),
            P<uint8_t>(c->flags), tiles, P<uint32_t>(c->fbbase), P<uint64_t>(c->toff), P<uint32_t>(c->tcap),
            P<TEntry>(c->table), c->pm0, c->pm1, P<uint32_t>(c->hit), tile_arg(c, count), probe_counter(c));
    KLAUNCH(c, "k_edge_resolve", k_edge_resolve, (xcd_grid(count), BLOCK, 0), tiles, P<uint32_t>(c->fbbase),
            P<uint64_t>(c->toff), P<TEntry>(c->table), P<uint32_t>(c->hit), P<uint32_t>(c->link),
            P<uint32_t>(c->hasp), P<uint64_t>(c->kmer), c->W, c->k, run_out(c), tile_arg(c, count));
    return UC_OK;
}

template <int W, bool CNT>
int launch_lds_edges_(uc_ctx *c, size_t first, size_t count)
{
    if (count == 0) return UC_OK;
    // threads per group: more waves per group hide the LDS and HBM latency
    // of its phases (measured: one-word k-mers 1.07 -> 0.81 ms at 512 threads,
    // two-word 11.4 -> 6.1 ms at 1024, against 256)
    constexpr int EB = W == 1 ? 512 : 1024;
    KLAUNCH(c, "k_edge_lds", (k_edge_lds<W, EB, LDS_GROUP, CNT>), ((uint32_t)count, EB, edge_lds_bytes<W>(c->lds_nb)),
            P<uint64_t>(c->kmer), P<EGroup>(c->egroups) + first, c->lds_nb, c->k, c->pm0, c->pm1,
            P<uint32_t>(c->link), P<uint32_t>(c->hasp), run_out(c), probe_counter(c));
    return UC_OK;
}

// fine buckets above BIG_GROUP and at most SLICE_MAX slots: one workgroup
// per hash slice
template <int W, bool CNT>
int launch_slice_edges_(uc_ctx *c, size_t first, size_t count)
{
    if (count == 0) return UC_OK;
    KLAUNCH(c, "k_edge_slice", (k_edge_slice<W, CNT>), ((uint32_t)count, 1024, 0), P<uint64_t>(c->kmer),
            P<SliceTask>(c->stasks) + first, c->k, c->pm0, c->pm1, P<uint32_t>(c->link), P<uint32_t>(c->hasp),
            run_out(c), P<uint32_t>(c->small) + C_SLICE, probe_counter(c), c->slice_probes);
    return UC_OK;
}

// single fine buckets above the group size but at most BIG_GROUP slots: one
// 1024-thread workgroup each, with the LDS area of BIG_GROUP slots
template <int W, bool CNT>
int launch_big_edges_(uc_ctx *c, size_t first, size_t count)
{
    if (count == 0) return UC_OK;
    const size_t lds = edge_lds_bytes<W>(BIG_GROUP);
    CHK(allow_lds(c, (k_edge_lds<W, 1024, BIG_GROUP, CNT>), lds));
    KLAUNCH(c, "k_edge_lds_big", (k_edge_lds<W, 1024, BIG_GROUP, CNT>), ((uint32_t)count, 1024, lds),
            P<uint64_t>(c->kmer), P<EGroup>(c->bgroups) + first, BIG_GROUP, c->k, c->pm0, c->pm1,
            P<uint32_t>(c->link), P<uint32_t>(c->hasp), run_out(c), probe_counter(c));
    return UC_OK;
}

// launch_global_edges with the probe-counting kernels when probes are counted
template <int W>
int launch_global_edges(uc_ctx *c, size_t first, size_t count)
{
    return probe_counter(c) ? launch_global_edges_<W, true>(c, first, count)
                            : launch_global_edges_<W, false>(c, first, count);
}

// launch_lds_edges with the probe-counting kernels when probes are counted
template <int W>
int launch_lds_edges(uc_ctx *c, size_t first, size_t count)
{
    return probe_counter(c) ? launch_lds_edges_<W, true>(c, first, count)
                            : launch_lds_edges_<W, false>(c, first, count);
}

// launch_slice_edges with the probe-counting kernels when probes are counted
template <int W>
int launch_slice_edges(uc_ctx *c, size_t first, size_t count)
{
    return probe_counter(c) ? launch_slice_edges_<W, true>(c, first, count)
                            : launch_slice_edges_<W, false>(c, first, count);
}

// launch_big_edges with the probe-counting kernels when probes are counted
template <int W>
int launch_big_edges(uc_ctx *c, size_t first, size_t count)
{
    return probe_counter(c) ? launch_big_edges_<W, true>(c, first, count)
                            : launch_big_edges_<W, false>(c, first, count);
}

// Work layout of the edge phase, once per bucketing result: consecutive
// fine buckets of one coarse bucket are packed into LDS groups of at most
// LDS_GROUP slots; a fine bucket larger than that gets a table of its own in
// HBM (the reference's layout: one open-addressing table per bucket,
// assembler.h:113-150).
int prepare_edges(uc_ctx *c)
{
    if (c->edges_ready) return UC_OK;
    const uint64_t G = (uint64_t)c->n_sets * c->B;
    const uint32_t load = c->opts.load_pct ? c->opts.load_pct : 50;
    const bool plain = c->opts.mode == UC_MODE_PLAIN;
    const uint32_t group_max = c->opts.group_slots ? c->opts.group_slots : LDS_GROUP;
    const uint64_t F = c->cb_first[G];
    // groups planned on the device in one dense list: coarse bucket g's
    // groups are entries [gfirst[g], gfirst[g + 1])
    CHK(ensure(c, c->egroups, std::max<uint64_t>(F, 1) * sizeof(EGroup)));
    CHK(ensure(c, c->over, F * 4 + 4));
    CHK(ensure(c, c->gfirst, (G + 4) * 4));
    CHK(ensure(c, c->small, C_WORDS * 4));
    uint32_t *cnt = P<uint32_t>(c->small);
    CHK(clear_probes(c));
    std::vector<uint32_t> gf(G + 1, 0);
    if (G) {
        // the oversized count is kept after the group offsets: one read for both
        uint32_t *gfd = P<uint32_t>(c->gfirst);
        CHK(fill<uint32_t>(c, gfd + G + 1, 1, 0u));
        KLAUNCH(c, "k_plan_count", k_plan_count, ((uint32_t)G, BLOCK, 0), P<uint32_t>(c->fboff),
                P<uint32_t>(c->cbfirst), group_max, gfd, P<uint32_t>(c->over), gfd + G + 1);
        CHK(scan_excl<uint32_t>(c, gfd, gfd, G, gfd + G, c->scan_tmp));
        // (sizes the bucketing left on the device come along: placements, crossings)
        const bool pend = c->sizes_pending;
        KLAUNCH(c, "k_plan_write", k_plan_write, ((uint32_t)G, BLOCK, 0), P<uint32_t>(c->fboff),
                P<uint32_t>(c->cbfirst), group_max, gfd, P<EGroup>(c->egroups), pend ? cnt + C_TOT : nullptr,
                gfd + G + 2);
        gf.resize(G + 4);
        CHK(d2h(c, gf.data(), gfd, (G + (pend ? 4 : 2)) * 4));
        if (pend) {
            CHK(check_placements(c, c->n, gf[G + 3]));
            c->x = gf[G + 3];
            c->m = c->n + c->x;
            c->sizes_pending = false;
        }
    }
    uint32_t n_over = 0;
    if (G) {
        n_over = gf[G + 1];
        gf.resize(G + 1);
    }
    c->lds_nb = group_max;
    c->grp_first.assign(gf.begin(), gf.end());
    // oversized fine buckets: their rows from the host mirror.  Those of at
    // most BIG_GROUP slots run alone in a larger LDS area; the rest get HBM
    // tables.
    std::vector<uint32_t> fb_base;
    std::vector<uint64_t> fb_cnt;
    std::vector<EGroup> big;
    std::vector<SliceTask> slices;
    c->fb_first.assign(G + 1, 0);
    c->bg_first.assign(G + 1, 0);
    c->sl_first.assign(G + 1, 0);
    if (n_over) {
        std::vector<uint32_t> ov(n_over);
        CHK(d2h(c, ov.data(), c->over.p, n_over * 4ull));
        std::sort(ov.begin(), ov.end());
        CHK(host_rows(c));
        const size_t big_lds = (size_t)BIG_GROUP * (8 * c->W + 10) + ((size_t)BIG_GROUP / 32 + 2) * 4;
        const bool big_ok = big_lds <= c->lds_limit;
        std::vector<uint32_t> ovh, ovb, ovs;  // HBM-table, big-LDS and sliced fine buckets
        std::vector<size_t> s_first;            // first slice task of each sliced fine bucket
        for (uint32_t f : ov) {
            const uint64_t sz = c->fboff_h[f + 1] - c->fboff_h[f];
            if (big_ok && sz <= BIG_GROUP) {
                ovb.push_back(f);
                big.push_back(EGroup{c->fboff_h[f], c->fboff_h[f + 1]});
            } else if (sz <= SLICE_MAX && !c->no_slices) {
                ovs.push_back(f);
                s_first.push_back(slices.size());
                const uint32_t P = (uint32_t)((sz + SLICE_KEYS - 1) / SLICE_KEYS);
                for (uint32_t q = 0; q < P; q++) slices.push_back(SliceTask{c->fboff_h[f], c->fboff_h[f + 1], q, P});
            } else {
                ovh.push_back(f);
                fb_base.push_back(c->fboff_h[f]);
                fb_cnt.push_back(sz);
            }
        }
        s_first.push_back(slices.size());
        for (uint64_t g = 0; g <= G; g++) {
            c->fb_first[g] = (size_t)(std::lower_bound(ovh.begin(), ovh.end(), c->cb_first[g]) - ovh.begin());
            c->bg_first[g] = (size_t)(std::lower_bound(ovb.begin(), ovb.end(), c->cb_first[g]) - ovb.begin());
            c->sl_first[g] = s_first[std::lower_bound(ovs.begin(), ovs.end(), c->cb_first[g]) - ovs.begin()];
        }
    }
    CHK(ensure(c, c->stasks, std::max<size_t>(slices.size(), 1) * sizeof(SliceTask)));
    if (!slices.empty()) CHK(h2d(c, c->stasks.p, slices.data(), slices.size() * sizeof(SliceTask)));
    CHK(fill<uint32_t>(c, cnt + C_SLICE, 1, 0u));
    CHK(ensure(c, c->bgroups, std::max<size_t>(big.size(), 1) * sizeof(EGroup)));
    if (!big.empty()) CHK(h2d(c, c->bgroups.p, big.data(), big.size() * sizeof(EGroup)));
    c->n_big = big.size();

    // HBM tables for oversized fine buckets; plain form shares one table
    // sized for the largest of them
    const size_t NF = fb_base.size();
    uint64_t fb_largest = 0;
    for (uint64_t x : fb_cnt) fb_largest = std::max(fb_largest, x);
    if (fb_largest > OWN_MASK) return fail(c, UC_E_LIMIT, "a bucket holds more than 2^30 slots");
    c->et_toff.assign(NF + 1, 0);
    c->et_cap.assign(NF, 0);
    uint64_t total = 0;
    const uint64_t shared_cap = table_cap(fb_largest, load);
    if (shared_cap > 0xFFFFFFFFull) return fail(c, UC_E_LIMIT, "bucket too large for one table");
    for (size_t f = 0; f < NF; f++) {
        c->et_cap[f] = (uint32_t)(plain ? shared_cap : table_cap(fb_cnt[f], load));
        c->et_toff[f] = plain ? 0 : total;
        if (!plain) total += c->et_cap[f];
    }
    if (plain && NF) total = shared_cap;
    c->table_slots = total;
    std::vector<ETile> et;
    c->et_first.assign(NF + 1, 0);
    for (size_t f = 0; f < NF; f++) {
        c->et_first[f] = et.size();
        for (uint64_t b = fb_base[f]; b < fb_base[f] + fb_cnt[f]; b += EDGE_TILE) {
            ETile t;
            t.bucket = (uint32_t)f;
            t.begin = (uint32_t)b;
            t.end = (uint32_t)std::min<uint64_t>(fb_base[f] + fb_cnt[f], b + EDGE_TILE);
            t.pad = 0;
            et.push_back(t);
        }
    }
    c->et_first[NF] = et.size();
    CHK(ensure(c, c->etiles, std::max<size_t>(et.size(), 1) * sizeof(ETile)));
    CHK(h2d(c, c->etiles.p, et.data(), et.size() * sizeof(ETile)));
    CHK(ensure(c, c->fbbase, NF * 4 + 8));
    CHK(h2d(c, c->fbbase.p, fb_base.data(), NF * 4));
    CHK(ensure(c, c->toff, (NF + 1) * 8));
    CHK(ensure(c, c->tcap, NF * 4 + 8));
    if (NF) {  // (read by the HBM-table kernels only)
        CHK(h2d(c, c->toff.p, c->et_toff.data(), (NF + 1) * 8));
        CHK(h2d(c, c->tcap.p, c->et_cap.data(), NF * 4));
    }
    CHK(ensure(c, c->table, total * sizeof(TEntry)));
    CHK(ensure(c, c->hit, c->m * 4));
    CHK(ensure(c, c->link, c->m * 4));
    CHK(ensure(c, c->hasp, (c->m + 31) / 32 * 4 + 8));
    CHK(ensure(c, c->rec, c->m * 8 + 16));
    CHK(ensure(c, c->rch, c->m + 16));
    // has-predecessor bits: every bucket sets only its own slots' bits
    CHK(fill<uint32_t>(c, P<uint32_t>(c->hasp), (c->m + 31) / 32, 0u));
    c->n_fallback = NF;
    c->n_groups = F;
    c->bucket_done.assign(G, 0);
    c->n_done = 0;
    c->edges_ready = true;
    return UC_OK;
}

// the oversized fine buckets [f0, f1) on the HBM tables
template <int W>
int run_fallback(uc_ctx *c, size_t f0, size_t f1)
{
    if (f0 == f1) return UC_OK;
    if (c->opts.mode != UC_MODE_PLAIN) {
        CHK(clear_table(c, P<TEntry>(c->table) + c->et_toff[f0], c->et_toff[f1 - 1] + c->et_cap[f1 - 1] - c->et_toff[f0]));
        return launch_global_edges<W>(c, c->et_first[f0], c->et_first[f1] - c->et_first[f0]);
    }
    for (size_t f = f0; f < f1; f++) {
        CHK(clear_table(c, P<TEntry>(c->table), c->et_cap[f]));
        CHK(launch_global_edges<W>(c, c->et_first[f], c->et_first[f + 1] - c->et_first[f]));
    }
    return UC_OK;
}

template <int W>
int run_edges(uc_ctx *c, uint32_t bucket)
{
    const bool fresh = !c->edges_ready;  // planning clears the bitmap and the probe counters itself
    CHK(prepare_edges(c));
    const uint64_t G = (uint64_t)c->n_sets * c->B;
    const bool plain = c->opts.mode == UC_MODE_PLAIN;
    if (bucket == UC_ALL_BUCKETS) {
        c->spliced = false;
        if (!fresh) {
            CHK(fill<uint32_t>(c, P<uint32_t>(c->hasp), (c->m + 31) / 32, 0u));
            CHK(clear_probes(c));
        }
        if (!plain) {
            CHK(launch_lds_edges<W>(c, 0, c->grp_first[G]));
            CHK(launch_big_edges<W>(c, 0, c->bg_first[G]));
            CHK(launch_slice_edges<W>(c, 0, c->sl_first[G]));
            CHK(run_fallback<W>(c, 0, c->fb_first[G]));
        } else {
            // the reference's per-bucket loop (driver.cpp:479-521)
            for (uint64_t g = 0; g < G; g++) {
                CHK(launch_lds_edges<W>(c, c->grp_first[g], c->grp_first[g + 1] - c->grp_first[g]));
                CHK(launch_big_edges<W>(c, c->bg_first[g], c->bg_first[g + 1] - c->bg_first[g]));
                CHK(launch_slice_edges<W>(c, c->sl_first[g], c->sl_first[g + 1] - c->sl_first[g]));
                CHK(run_fallback<W>(c, c->fb_first[g], c->fb_first[g + 1]));
            }
        }
        std::fill(c->bucket_done.begin(), c->bucket_done.end(), 1);
        c->n_done = G;
    } else {
        if (bucket >= G) return fail(c, UC_E_ARG, "bucket id out of range");
        // this bucket sets only its own slots' has-predecessor bits, which
        // start cleared from prepare_edges.  After a contraction (which
        // consumed the lent k-mers' run heads) every bucket is found again.
        if (c->spliced) {
            std::fill(c->bucket_done.begin(), c->bucket_done.end(), 0);
            c->n_done = 0;
            c->spliced = false;
        }
        CHK(launch_lds_edges<W>(c, c->grp_first[bucket], c->grp_first[bucket + 1] - c->grp_first[bucket]));
        CHK(launch_big_edges<W>(c, c->bg_first[bucket], c->bg_first[bucket + 1] - c->bg_first[bucket]));
        CHK(launch_slice_edges<W>(c, c->sl_first[bucket], c->sl_first[bucket + 1] - c->sl_first[bucket]));
        CHK(run_fallback<W>(c, c->fb_first[bucket], c->fb_first[bucket + 1]));
        if (!c->bucket_done[bucket]) {
            c->bucket_done[bucket] = 1;
            c->n_done++;
        }
    }
    c->edged = c->n_done == G;
    return UC_OK;
}

int do_edges(uc_ctx *c, uint32_t bucket)
{
    if (!c->bucketed) return fail(c, UC_E_STATE, "uc_find_edges before uc_bucketize");
    c->coarsened = c->emitted = false;
    CHK(c->W == 1 ? run_edges<1>(c, bucket) : run_edges<2>(c, bucket));
    if (c->edged) {  // every bucket done: the hash tables and their hit positions are dead
        CHK(drop(c, c->table));
        CHK(drop(c, c->hit));
    }
    return UC_OK;
}

// ---- contraction ----------------------------------------------------------

// one contraction pass: segment walks, chain ranking, unitig count/length
// per tile, scan of the tiles.  *live = k-mers reached from a head,
// *nonconv = sampled splitters whose chain has no head (on a cycle).
// internal status of a contraction pass whose edge phase had a slice table
// overflow: the edge phase runs again without slices (never returned by the
// C ABI)
constexpr int RETRY_SLICES = 1;
constexpr uint32_t CHAIN_BLOCKS = 1024;  // grid of the chain kernels (they stride)

// Segment walks, chain ranking and unitig totals.  The first attempt (spec)
// reads the device's counters once, at the end, after SPEC_ROUNDS chain
// rounds; if a chain is still open then, the pass runs again from the walks
// with a host-checked round loop (everything it writes is rewritten).
// kernels of a contraction pass, which the graph form captures
static const char *const PASS_KERNELS[] = {"k_fill",       "k_seg_count",   "k_scan_reduce", "k_scan_tile",
                                           "k_chain_init", "k_chain_round", "k_chain_tail"};

// the graph form unless asked for eager launches, or unless one of the
// pass's kernels is to be timed (captured launches carry no events)
static bool pass_graph(const uc_ctx *c)
{
    if ((c->opts.flags & UC_FLAG_EAGER) || c->arena) return false;  // instantiating a graph allocates
    if (!(c->opts.flags & UC_FLAG_PROFILE)) return true;
    if (c->prof_only.empty()) return false;
    for (const char *n : PASS_KERNELS)
        if (c->prof_only == n) return false;
    return true;
}

void drop_pass_graph(uc_ctx *c)
{
    if (c->cgraph) (void)hipGraphExecDestroy(c->cgraph);
    c->cgraph = nullptr;
    c->cgraph_key.clear();
}

// With `defer`, the speculative pass is only launched: its counters stay on
// the device for finish_deferred().
int contract_pass(uc_ctx *c, uint64_t *live, uint32_t *nonconv, bool spec = true, bool defer = false)
{
    const uint32_t m = (uint32_t)c->m;
    const uint64_t T = (c->m + CT_TILE - 1) / CT_TILE;
    uint32_t *cnt = P<uint32_t>(c->small);
    const uint32_t *np = cnt + C_LIST;
    uint64_t *ca = P<uint64_t>(c->chainA), *cb = P<uint64_t>(c->chainB);
    // walks, the splitter/run scan, then (spec) chain ranking and the
    // unitig-length scan; everything up to the counter read
    auto walks = [&]() -> int {
        CHK(fill<uint32_t>(c, cnt, C_SLICE, 0u));  // (the edge phase's slice flag stays)
        // 256 threads, 8 slots each: the walks of a thread's splitters
        // overlap (measured on MI355X, 50 Mbp job: 0.30 ms; 128 threads
        // 0.35, 512 threads 0.37, 1024 threads 0.74)
        KLAUNCH(c, "k_seg_count", k_seg_count<256>, (xcd_grid(T), 256, 0), P<uint64_t>(c->rec), P<uint32_t>(c->hasp),
                P<uint8_t>(c->flags), m, c->salt, c->shift, P<uint64_t>(c->seg), P<uint32_t>(c->prevseg),
                P<uint32_t>(c->list), P<uint64_t>(c->mlist), P<uint32_t>(c->mcnt), cnt, P<uint64_t>(c->tsum),
                P<uint64_t>(c->tcnt), tile_arg(c, T));
        CHK(scan_excl<uint64_t>(c, P<uint64_t>(c->tcnt), P<uint64_t>(c->tcnt), T,
                                reinterpret_cast<uint64_t *>(cnt + C_TOT), c->scan_tmp));
        return UC_OK;
    };
    auto totals = [&]() -> int {
        uint64_t *ttot = reinterpret_cast<uint64_t *>(cnt + C_TOT + 2);
        return scan_excl<uint64_t>(c, P<uint64_t>(c->tsum), P<uint64_t>(c->tsum), T, ttot, c->scan_tmp);
    };
    auto spec_pass = [&]() -> int {
        CHK(walks());
        uint64_t *a = ca, *b = cb;
        const uint32_t g = std::max<uint32_t>(1, std::min<uint32_t>(grid_for(m), CHAIN_BLOCKS));
        KLAUNCH(c, "k_chain_init", k_chain_init, (g, BLOCK, 0), P<uint32_t>(c->list), np, P<uint32_t>(c->prevseg),
                P<uint64_t>(c->seg), a);
        for (uint32_t r = 0; r < SPEC_ROUNDS; r++) {
            KLAUNCH(c, "k_chain_round", k_chain_round, (g, BLOCK, 0), P<uint32_t>(c->list), np,
                    P<uint32_t>(c->hasp), a, b, cnt + C_ACTIVE + r);
            std::swap(a, b);
        }
        KLAUNCH(c, "k_chain_tail", k_chain_tail, (g, BLOCK, 0), P<uint32_t>(c->list), np, P<uint32_t>(c->hasp), a,
                P<uint64_t>(c->seg), P<uint32_t>(c->ulen), P<uint64_t>(c->tsum), cnt + C_NONCONV);
        return totals();
    };
    uint32_t h[C_WORDS];
    uint32_t rounds = 0;
    auto read_counts = [&]() -> int {
        CHK(d2h(c, h, cnt, sizeof(h)));
        if (h[C_ERR]) return fail(c, UC_E_STATE, "successor array corrupted: a segment walk exceeded m steps");
        if (h[C_SLICE]) return c->no_slices ? fail(c, UC_E_STATE, "an edge-finding slice table overflowed")
                                            : RETRY_SLICES;
        return UC_OK;
    };
    if (spec && pass_graph(c)) {
        // no allocation or host wait may happen inside the capture
        CHK(ensure(c, c->scan_tmp, scan_scratch<uint64_t>(T) + 256));
        const std::vector<uint64_t> key = {
            m, c->salt, c->shift, tile_arg(c, T), (uint64_t)c->rec.p, (uint64_t)c->hasp.p, (uint64_t)c->flags.p,
            (uint64_t)c->seg.p, (uint64_t)c->prevseg.p, (uint64_t)c->list.p, (uint64_t)c->mlist.p,
            (uint64_t)c->mcnt.p, (uint64_t)c->small.p, (uint64_t)c->tsum.p, (uint64_t)c->tcnt.p,
            (uint64_t)c->chainA.p, (uint64_t)c->chainB.p, (uint64_t)c->ulen.p, (uint64_t)c->scan_tmp.p};
        if (!c->cgraph || key != c->cgraph_key) {
            drop_pass_graph(c);
            HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            c->capturing = true;
            const int rc = spec_pass();
            c->capturing = false;
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(c->stream, &g);
            if (rc != UC_OK || e != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                return rc != UC_OK ? rc : fail(c, UC_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
            }
            const hipError_t ei = hipGraphInstantiate(&c->cgraph, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ei != hipSuccess) {
                c->cgraph = nullptr;
                return fail(c, UC_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
            }
            c->cgraph_key = key;
        }
        HIPCHK(c, hipGraphLaunch(c->cgraph, c->stream));
    } else if (spec) {
        CHK(spec_pass());
    } else {
        CHK(walks());
        CHK(read_counts());
        const uint32_t nl = h[C_LIST];
        if (nl) {
            const uint32_t g = std::min<uint32_t>(grid_for(nl), CHAIN_BLOCKS);
            KLAUNCH(c, "k_chain_init", k_chain_init, (g, BLOCK, 0), P<uint32_t>(c->list), np,
                    P<uint32_t>(c->prevseg), P<uint64_t>(c->seg), ca);
            // a chain holds at most nl segments: ceil(log2(nl)) + 1 rounds
            // reach its head; chains still open after that run around a cycle
            uint32_t cap = 1;
            while ((1ull << cap) < nl) cap++;
            cap += 1;
            for (;;) {
                CHK(fill<uint32_t>(c, cnt + C_ACTIVE, 1, 0u));
                KLAUNCH(c, "k_chain_round", k_chain_round, (g, BLOCK, 0), P<uint32_t>(c->list), np,
                        P<uint32_t>(c->hasp), ca, cb, cnt + C_ACTIVE);
                std::swap(ca, cb);
                rounds++;
                uint32_t active;
                CHK(d2h(c, &active, cnt + C_ACTIVE, 4));
                if (active == 0 || rounds >= cap) break;
            }
            KLAUNCH(c, "k_chain_tail", k_chain_tail, (g, BLOCK, 0), P<uint32_t>(c->list), np,
                    P<uint32_t>(c->hasp), ca, P<uint64_t>(c->seg), P<uint32_t>(c->ulen), P<uint64_t>(c->tsum),
                    cnt + C_NONCONV);
        }
        CHK(totals());
    }
    if (spec && SPEC_ROUNDS % 2) std::swap(ca, cb);  // where the last speculative round wrote
    if (defer) {
        c->chain = ca;
        return UC_OK;
    }
    // the unitig totals land beside the other counters: one read for all
    CHK(read_counts());
    const uint32_t nl = h[C_LIST];
    if (spec) {
        // rounds to convergence: the first whose pairs all reached a head
        while (rounds < SPEC_ROUNDS && h[C_ACTIVE + rounds]) rounds++;
        if (nl && rounds == SPEC_ROUNDS) return contract_pass(c, live, nonconv, false);
        rounds = nl ? rounds + 1 : 0;
    }
    c->n_list = nl;
    uint64_t st;
    std::memcpy(&st, h + C_TOT, 8);
    c->n_splitters = st >> 32;
    c->n_runs = st & 0xFFFFFFFFull;
    c->chain = ca;
    c->chain_rounds += rounds;
    uint64_t tot;
    std::memcpy(&tot, h + C_TOT + 2, 8);
    c->n_unitigs = tot >> 32;
    *live = tot & 0xFFFFFFFFull;
    *nonconv = h[C_NONCONV];
    return UC_OK;
}

// isolated cycles: every k-mer no head reaches lies on one; each cycle is
// cut in front of its lexicographically smallest k-mer
template <int W>
int cut_cycles(uc_ctx *c, uint64_t *n_cut)
{
    const uint32_t m = (uint32_t)c->m;
    uint32_t *cnt = P<uint32_t>(c->small);
    KLAUNCH(c, "k_cover", k_cover, (grid_for(m), BLOCK, 0), P<uint64_t>(c->rec), P<uint32_t>(c->hasp),
            P<uint8_t>(c->flags), c->chain, m, c->salt, c->shift);
    CHK(fill<uint32_t>(c, cnt + C_CYC, 1, 0u));
    CHK(fill<uint32_t>(c, cnt + C_CUT, 1, 0u));
    uint64_t *la = P<uint64_t>(c->chainA) == c->chain ? P<uint64_t>(c->chainB) : P<uint64_t>(c->chainA);
    uint64_t *lb = P<uint64_t>(c->cycB);
    KLAUNCH(c, "k_cycle_init", k_cycle_init, (grid_for(m), BLOCK, 0), P<uint32_t>(c->link), P<uint8_t>(c->flags), m,
            run_out(c), la, P<uint32_t>(c->clist), cnt + C_CYC);
    uint32_t nl;
    CHK(d2h(c, &nl, cnt + C_CYC, 4));
    uint32_t rounds = 1;
    while ((1ull << (rounds - 1)) < nl) rounds++;
    for (uint32_t r = 0; r < rounds && nl; r++) {
        KLAUNCH(c, "k_cycle_round", k_cycle_round<W>, (grid_for(nl), BLOCK, 0), P<uint32_t>(c->clist), nl,
                P<uint64_t>(c->kmer), la, lb);
        std::swap(la, lb);
    }
    if (nl) {
        KLAUNCH(c, "k_cycle_pred", k_cycle_pred, (grid_for(nl), BLOCK, 0), P<uint32_t>(c->clist), nl,
                P<uint32_t>(c->link), P<uint32_t>(c->pred));
        KLAUNCH(c, "k_cycle_cut", k_cycle_cut, (grid_for(nl), BLOCK, 0), P<uint32_t>(c->clist), nl, la,
                P<uint32_t>(c->pred), P<uint32_t>(c->link), P<uint64_t>(c->rec), P<uint32_t>(c->hasp),
                P<uint8_t>(c->flags), cnt + C_CUT);
    }
    KLAUNCH(c, "k_clear_flag", k_clear_flag, (std::min<uint32_t>(grid_for(m), 65536), BLOCK, 0), P<uint8_t>(c->flags),
            m, F_VISIT);
    uint32_t cuts = 0;
    CHK(d2h(c, &cuts, cnt + C_CUT, 4));
    *n_cut = cuts;
    return UC_OK;
}

// every array a contraction pass writes or reads besides the edge result
int ensure_contraction(uc_ctx *c)
{
    const uint64_t T = (c->m + CT_TILE - 1) / CT_TILE;
    CHK(ensure(c, c->small, C_WORDS * 4));
    CHK(ensure(c, c->tcnt, (T + 1) * 8));
    CHK(ensure(c, c->seg, c->m * 8));
    CHK(ensure(c, c->mlist, c->m * 8));
    CHK(ensure(c, c->mcnt, (T + 1) * 4));
    CHK(ensure(c, c->prevseg, c->m * 4));
    CHK(ensure(c, c->ulen, c->m * 4));
    CHK(ensure(c, c->chainA, c->m * 8));
    CHK(ensure(c, c->chainB, c->m * 8));
    CHK(ensure(c, c->list, c->m * 4));
    CHK(ensure(c, c->tsum, (T + 1) * 8));
    return UC_OK;
}

// the edge phase again with every fine bucket that had a slice on an HBM
// table (a slice table overflowed), then contraction
int do_coarsen(uc_ctx *c, bool defer = false);
int redo_without_slices(uc_ctx *c, bool defer)
{
    c->no_slices = true;
    c->edges_ready = false;
    CHK(do_edges(c, UC_ALL_BUCKETS));
    return do_coarsen(c, defer);
}

// Segment walks, chain ranking and cycle cutting.  With `defer` (whole jobs
// that can hold the largest text, see can_defer()), the pass's counters are
// not read here: emission follows at once and finish_deferred() reads them.
int do_coarsen(uc_ctx *c, bool defer)
{
    if (!c->edged) return fail(c, UC_E_STATE, "uc_coarsen before uc_find_edges");
    c->emitted = false;
    c->deferred = c->lists_spent = false;
    const uint32_t m = (uint32_t)c->m;
    CHK(ensure_contraction(c));
    c->chain_rounds = 0;
    c->n_cycles = 0;
    c->n_list = 0;
    c->chain = P<uint64_t>(c->chainA);
    c->salt = (uint32_t)(c->opts.seed * 0x9E3779B97F4A7C15ull >> 32);
    c->shift = c->opts.sample_log2 ? c->opts.sample_log2 : DEFAULT_SAMPLE_LOG2;
    if (c->x && !c->spliced) {  // (once per edge result: it consumes the lent k-mers' run heads)
        KLAUNCH(c, "k_splice", k_splice, (grid_for(c->x), BLOCK, 0), P<uint32_t>(c->pair_h), P<uint32_t>(c->pair_b),
                (uint32_t)c->x, P<uint32_t>(c->link), P<uint64_t>(c->rec));
    }
    c->spliced = true;
    uint64_t live = 0;
    uint32_t nonconv = 0;
    if (defer) {
        CHK(contract_pass(c, &live, &nonconv, true, true));
        c->deferred = true;
        c->text_bytes = c->n * (c->k + 1);  // n isolated k-mers: the most text a job can have
        c->coarsened = true;
        return UC_OK;
    }
    if (m) {
        const int rc = contract_pass(c, &live, &nonconv);
        if (rc == RETRY_SLICES) return redo_without_slices(c, false);
        CHK(rc);
    }
    if (live != c->n || nonconv) {
        // cutting cycles reads the ranked chain and labels in the other chain
        // buffer; the pass's other arrays are rewritten by the next pass, so
        // a low-memory job gives them back meanwhile
        for (DBuf *b : {&c->seg, &c->mlist, &c->prevseg, &c->ulen, &c->list, &c->tsum, &c->tcnt, &c->mcnt})
            CHK(drop(c, *b));
        CHK(ensure(c, c->cycB, c->m * 8));
        CHK(ensure(c, c->clist, c->m * 4));
        CHK(ensure(c, c->pred, c->m * 4));
        uint64_t cut = 0;
        CHK(c->W == 1 ? cut_cycles<1>(c, &cut) : cut_cycles<2>(c, &cut));
        c->n_cycles = cut;
        for (DBuf *b : {&c->cycB, &c->clist, &c->pred}) CHK(drop(c, *b));
        CHK(ensure_contraction(c));
        CHK(contract_pass(c, &live, &nonconv));
        if (live != c->n || nonconv)
            return fail(c, UC_E_STATE, "contraction lost k-mers: " + std::to_string(live) + " of " +
                                           std::to_string(c->n) + " reached after cutting cycles");
    }
    c->text_bytes = live + c->n_unitigs * c->k;
    c->coarsened = true;
    if (c->opts.flags & UC_FLAG_LOW_MEMORY) {
        // emission reads the k-mers, flags, predecessor bits, run records and
        // characters, the head lists, unitig lengths, tile sums, the sampled
        // splitters and their ranked chain; everything else is dead.  The
        // ranked entries move to chainA, which the pool keeps low.
        if (c->chain != P<uint64_t>(c->chainA)) {
            if (c->n_list)
                KLAUNCH(c, "k_chain_move", k_chain_move, (std::min<uint32_t>(grid_for(c->n_list), 65536), BLOCK, 0),
                        P<uint32_t>(c->list), c->n_list, c->chain, P<uint64_t>(c->chainA));
            c->chain = P<uint64_t>(c->chainA);
        }
        // (small planning buffers included: left in the pool's top range
        // they would split the space the text needs)
        for (DBuf *b : all_bufs(c))
            if (!keeps(c, *b)) CHK(drop(c, *b));
        c->edged = false;  // the links are gone: contraction cannot run again on this edge result
    }
    return UC_OK;
}

int set_ranges(uc_ctx *c);
int finish_deferred(uc_ctx *c);

template <int W>
int run_emit(uc_ctx *c)
{
    const uint32_t m = (uint32_t)c->m;
    const uint64_t T = (c->m + CT_TILE - 1) / CT_TILE;
    if (c->text_fixed) {
        // the caller's buffer (uc_compact_noalloc); the kernels write bytes [0, text_bytes) only
        if (c->text.cap < c->text_bytes)
            return fail(c, UC_E_NOMEM, "the text buffer holds " + std::to_string(c->text.cap) + " bytes; the job's " +
                                           "text is " + std::to_string(c->text_bytes));
    } else {
        CHK(ensure(c, c->text, c->text_bytes + 1));
    }
    // Head text offsets of unitigs of several segments (written by
    // k_head_emit, read by k_rand_emit), one word per slot.  A low-memory job
    // of one set keeps them in the head lists' memory instead: a tile reads
    // all its own list entries (inside its own slot range) before its first
    // barrier and writes offsets only inside its own slot range after its
    // last, and nothing else reads the lists (k_set_offsets runs for several
    // sets only).  Emission cannot run again on the spent lists.
    if (c->lists_spent) return fail(c, UC_E_STATE, "uc_emit again after a low-memory emission (the head lists are spent)");
    const bool alias_htoff = (c->opts.flags & UC_FLAG_LOW_MEMORY) && c->n_sets == 1 && !c->deferred;
    if (alias_htoff)
        CHK(drop(c, c->htoff));
    else
        CHK(ensure(c, c->htoff, c->m * 8));
    uint64_t *htoff = alias_htoff ? P<uint64_t>(c->mlist) : P<uint64_t>(c->htoff);
    c->lists_spent = alias_htoff;
    uint32_t *cnt = P<uint32_t>(c->small);
    // a pass the host has not checked: the kernels check it themselves
    const uint32_t *gate = c->deferred ? cnt : nullptr;
    if (m) {
        // one-word k-mers from 256-thread groups, wider ones from 512-thread
        // groups, both staged in LDS (each measured faster than the other
        // choices on MI355X)
        constexpr int NT = W == 1 ? 256 : 512;
        KLAUNCH(c, "k_head_emit", (k_head_emit<W, true, NT>), (xcd_grid(T), NT, 0), P<uint64_t>(c->kmer),
                P<uint64_t>(c->rec), P<char>(c->rch), P<uint32_t>(c->hasp), P<uint8_t>(c->flags),
                P<uint64_t>(c->mlist), P<uint32_t>(c->mcnt), P<uint32_t>(c->ulen), P<uint64_t>(c->tsum), m, c->k,
                c->salt, c->shift, htoff, P<char>(c->text), tile_arg(c, T), gate, (uint32_t)c->n);
    }
    if (c->deferred || c->n_list) {
        // (the list's length is read by the kernel: a deferred pass has not
        // told the host yet)
        const uint32_t g = c->deferred ? std::max<uint32_t>(1, std::min<uint32_t>(grid_for(m), CHAIN_BLOCKS))
                                       : grid_for(c->n_list);
        KLAUNCH(c, "k_rand_emit", k_rand_emit, (g, BLOCK, 0), P<uint64_t>(c->rec), P<char>(c->rch),
                P<uint32_t>(c->list), cnt + C_LIST, c->chain, htoff, c->k, c->salt, c->shift, m, P<char>(c->text),
                gate, (uint32_t)c->n);
    }
    if (c->n_sets > 1) {
        // per-set text ranges: a set's unitigs start at its first bucket row
        CHK(ensure(c, c->setoff, (c->n_sets + 1) * 8));
        KLAUNCH(c, "k_set_offsets", k_set_offsets, (c->n_sets + 1, BLOCK, 0), P<uint32_t>(c->boff), c->n_sets,
                c->B, P<uint32_t>(c->hasp), P<uint8_t>(c->flags), P<uint64_t>(c->rec), P<uint64_t>(c->mlist),
                P<uint32_t>(c->mcnt), P<uint32_t>(c->ulen), P<uint64_t>(c->tsum), m,
                reinterpret_cast<const uint64_t *>(cnt + C_TOT + 2), P<uint64_t>(c->setoff), gate, (uint32_t)c->n);
    }
    return c->deferred ? finish_deferred(c) : set_ranges(c);
}

int do_emit(uc_ctx *c)
{
    if (!c->coarsened) return fail(c, UC_E_STATE, "uc_emit before uc_coarsen");
    return c->W == 1 ? run_emit<1>(c) : run_emit<2>(c);
}

// each set's text range (one set: the whole text), after emission
int set_ranges(uc_ctx *c)
{
    c->set_text.resize(c->n_sets + 1);
    if (c->n_sets == 1) {
        c->set_text[0] = 0;
        c->set_text[1] = c->text_bytes;
    } else {
        std::vector<uint64_t> so(c->n_sets + 1);
        CHK(d2h(c, so.data(), c->setoff.p, (c->n_sets + 1) * 8));
        for (uint32_t s = 0; s <= c->n_sets; s++) c->set_text[s] = (so[s] & 0xFFFFFFFFull) + (so[s] >> 32) * c->k;
    }
    c->emitted = true;
    return UC_OK;
}

// The end of a job whose contraction counters were left on the device: one
// read of them after emission.  A pass that did not settle (a chain still
// open after the speculative rounds, a cycle no head reaches) or whose edge
// phase had a slice table overflow runs again in the checked form, emission
// included; everything those write is rewritten.
int finish_deferred(uc_ctx *c)
{
    c->deferred = false;
    uint32_t h[C_WORDS];
    CHK(d2h(c, h, P<uint32_t>(c->small), sizeof(h)));
    if (h[C_ERR]) return fail(c, UC_E_STATE, "successor array corrupted: a segment walk exceeded m steps");
    const uint32_t nl = h[C_LIST];
    uint32_t rounds = 0;
    while (rounds < SPEC_ROUNDS && h[C_ACTIVE + rounds]) rounds++;
    uint64_t st, tot;
    std::memcpy(&st, h + C_TOT, 8);
    std::memcpy(&tot, h + C_TOT + 2, 8);
    const uint64_t live = tot & 0xFFFFFFFFull;
    if (h[C_SLICE] || (nl && rounds == SPEC_ROUNDS) || live != c->n || h[C_NONCONV]) {
        c->coarsened = false;
        if (h[C_SLICE]) {
            if (c->no_slices) return fail(c, UC_E_STATE, "an edge-finding slice table overflowed");
            CHK(redo_without_slices(c, false));
        } else {
            CHK(do_coarsen(c));
        }
        return do_emit(c);
    }
    c->n_list = nl;
    c->n_splitters = st >> 32;
    c->n_runs = st & 0xFFFFFFFFull;
    c->chain_rounds += nl ? rounds + 1 : 0;
    c->n_unitigs = tot >> 32;
    c->text_bytes = live + c->n_unitigs * c->k;
    return set_ranges(c);
}

// Whole jobs whose largest possible text (n isolated k-mers) fits in
// DEFER_TEXT_MAX bytes leave the contraction counters on the device until
// emission is done: one host read fewer per job.  Not for low-memory jobs
// (whose text buffer is sized exactly), the caller's text buffer, the static
// scratch or the eager form (which takes every step in the checked order).
constexpr uint64_t DEFER_TEXT_MAX = 256ull << 20;

// a whole job that may bucket with its placement totals left on the device:
// the jobs whose contraction counters are deferred too (their buffers are
// small enough to size for m <= 2n)
bool lazy_sizes_ok(const uc_ctx *c)
{
    const uint64_t bound = c->n * (c->k + 1) + 1;
    return c->n && !c->arena && !c->text_fixed && !(c->opts.flags & (UC_FLAG_LOW_MEMORY | UC_FLAG_EAGER)) &&
           bound <= DEFER_TEXT_MAX && (!c->opts.hbm_budget || c->held + 2 * bound <= c->opts.hbm_budget);
}

bool can_defer(const uc_ctx *c)
{
    const uint64_t bound = c->n * (c->k + 1) + 1;
    if (!c->m || c->arena || c->text_fixed || (c->opts.flags & (UC_FLAG_LOW_MEMORY | UC_FLAG_EAGER)) ||
        bound > DEFER_TEXT_MAX)
        return false;
    return !c->opts.hbm_budget || c->held + bound <= c->opts.hbm_budget;
}

// bucketing .. emission on the loaded k-mers, one event per phase boundary
int timed_phases(uc_ctx *c, uint32_t l)
{
    c->held_peak = c->held - c->pool.cap + c->pool.used;  // bytes in use, not the pool's reserve
    for (auto &e : c->ev)
        if (!e) HIPCHK(c, hipEventCreate(&e));
    HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
    c->lazy_sizes = lazy_sizes_ok(c);
    const int rb = do_bucketize(c, l, nullptr);
    c->lazy_sizes = false;
    CHK(rb);
    HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
    CHK(do_edges(c, UC_ALL_BUCKETS));
    HIPCHK(c, hipEventRecord(c->ev[2], c->stream));
    CHK(do_coarsen(c, can_defer(c)));
    HIPCHK(c, hipEventRecord(c->ev[3], c->stream));
    CHK(do_emit(c));
    HIPCHK(c, hipEventRecord(c->ev[4], c->stream));
    HIPCHK(c, hipEventSynchronize(c->ev[4]));
    float t[4];
    for (int i = 0; i < 4; i++) HIPCHK(c, hipEventElapsedTime(&t[i], c->ev[i], c->ev[i + 1]));
    c->timing.bucketize_ms = t[0];
    c->timing.edges_ms = t[1];
    c->timing.coarsen_ms = t[2];
    c->timing.emit_ms = t[3];
    HIPCHK(c, hipEventElapsedTime(&c->timing.total_ms, c->ev[0], c->ev[4]));
    c->timing.device_bytes = c->held;
    c->timing.device_peak_bytes = c->held_peak;
    return prof_flush(c);
}

int timed_job(uc_ctx *c, const uint32_t *packed, const uint64_t *offs, uint32_t n_sets, uint32_t k, uint32_t l,
              int on_device)
{
    CHK(do_load(c, packed, offs, n_sets, k, on_device));
    return timed_phases(c, l);
}

// Device address of the scratch array on `device`, which must be current:
// looked up once per device (every device holds its own copy of the array at
// its own address) and kept.  nullptr when the lookup failed.
constexpr int SCRATCH_DEVS = 64;

unsigned char *scratch_ptr(int device)
{
    static std::mutex mu;
    static unsigned char *ptr[SCRATCH_DEVS] = {};
    static bool done[SCRATCH_DEVS] = {};
    if (device < 0 || device >= SCRATCH_DEVS) return nullptr;
    std::lock_guard<std::mutex> g(mu);
    if (!done[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        ptr[device] = e == hipSuccess ? static_cast<unsigned char *>(p) : nullptr;
        done[device] = true;
    }
    return ptr[device];
}

}  // namespace

// ===========================================================================
// C ABI

extern "C" {

void uc_default_opts(uc_opts *o)
{
    std::memset(o, 0, sizeof(*o));
    o->device = 0;
    o->seed = 0x5EEDull;
    o->n_buckets = 64;
    o->mode = UC_MODE_BATCHED;
    o->load_pct = 50;
    o->sample_log2 = DEFAULT_SAMPLE_LOG2;
    o->hbm_budget = 0;
}

const char *uc_strerror(int s)
{
    switch (s) {
    case UC_OK: return "ok";
    case UC_E_ARG: return "invalid argument";
    case UC_E_NOMEM: return "out of memory";
    case UC_E_HIP: return "HIP runtime error";
    case UC_E_STATE: return "call out of order";
    case UC_E_IO: return "I/O error";
    case UC_E_FORMAT: return "malformed k-mer text";
    case UC_E_LIMIT: return "job exceeds a compiled limit";
    default: return "unknown status";
    }
}

int uc_ctx_create(const uc_opts *opts, uc_ctx **out)
{
    if (!out) return UC_E_ARG;
    *out = nullptr;
    uc_opts o;
    if (opts)
        o = *opts;
    else
        uc_default_opts(&o);
    if (o.n_buckets < 1 || o.n_buckets > MAX_BUCKETS || (o.n_buckets & (o.n_buckets - 1))) return UC_E_ARG;
    if (o.mode != UC_MODE_BATCHED && o.mode != UC_MODE_PLAIN) return UC_E_ARG;
    if (o.load_pct && (o.load_pct < 10 || o.load_pct > 90)) return UC_E_ARG;
    if (o.sample_log2 > 16) return UC_E_ARG;
    if (o.fine_slots && (o.fine_slots < 64 || o.fine_slots > 4096)) return UC_E_ARG;
    if (o.group_slots && (o.group_slots < 256 || o.group_slots > LDS_GROUP || (o.group_slots & (o.group_slots - 1))))
        return UC_E_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return UC_E_HIP;
    if (o.device < 0 || o.device >= ndev) return UC_E_ARG;
    if (hipSetDevice(o.device) != hipSuccess) return UC_E_HIP;
    uc_ctx *c = new uc_ctx();
    c->opts = o;
    c->B = o.n_buckets;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return UC_E_HIP;
    }
    int lds = 0;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, o.device) == hipSuccess && lds > 0)
        c->lds_limit = (uint32_t)lds;
    // set-up of uc_compact_noalloc, with the context's device current (a
    // failed lookup only disables that entry point)
    c->scratch_dev = scratch_ptr(o.device);
    // (without it the reads go to pageable memory)
    if (hipHostMalloc((void **)&c->rbuf, RBUF_BYTES, hipHostMallocDefault) != hipSuccess) c->rbuf = nullptr;
    *out = c;
    return UC_OK;
}

int uc_ctx_set_slice_probes(uc_ctx *c, uint32_t probes)
{
    if (!c || probes < 1 || probes > SLICE_TAB) return UC_E_ARG;
    c->slice_probes = probes;
    return UC_OK;
}

int uc_ctx_destroy(uc_ctx *c)
{
    if (!c) return UC_OK;
    (void)hipSetDevice(c->opts.device);
    (void)hipStreamSynchronize(c->stream);
    drop_pass_graph(c);
    for (DBuf *b : all_bufs(c))
        if (b->p && !in_pool(c, b->p)) (void)hipFree(b->p);
    if (c->pool.base) (void)hipFree(c->pool.base);
    if (c->meter) c->meter->sub(c->held);
    free_pinned(c);
    if (c->rbuf) (void)hipHostFree(c->rbuf);
    for (auto &e : c->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto &p : c->pending) {
        (void)hipEventDestroy(p.second.first);
        (void)hipEventDestroy(p.second.second);
    }
    for (auto e : c->evpool) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return UC_OK;
}

const char *uc_last_error(const uc_ctx *c) { return c ? c->err.c_str() : "no context"; }

int uc_compact(uc_ctx *c, const uint32_t *packed, uint64_t n, uint32_t k, uint32_t l, int on_device)
{
    if (!c || (!packed && n)) return UC_E_ARG;
    CHK(select_dev(c));
    const uint64_t offs[2] = {0, n};
    return timed_job(c, packed, offs, 1, k, l, on_device);
}

int uc_compact_noalloc(uc_ctx *c, const uint32_t *packed, uint64_t n, uint32_t k, uint32_t l, char *text,
                       uint64_t text_cap, uint64_t *text_size)
{
    if (!c || (!packed && n) || !text_size || (!text && text_cap)) return UC_E_ARG;
    *text_size = 0;
    if (k < 2 || k > 64) return fail(c, UC_E_ARG, "k must be in [2, 64]");
    CHK(select_dev(c));
    if (!c->scratch_dev) return fail(c, UC_E_HIP, "hipGetSymbolAddress found no scratch on this device");
    // every job keeps at least its bucketed k-mers twice plus a flag byte,
    // a link and a run record per slot: a job beyond that is refused before
    // anything runs; the carving refuses the rest as it meets them
    const uint64_t W = k <= 32 ? 1 : 2, floor_per_kmer = 16 * W + 13;
    const uint64_t usable = UC_SCRATCH_BYTES - SCRATCH_TAIL;
    if (n > usable / floor_per_kmer)
        return fail(c, UC_E_LIMIT, std::to_string(n) + " k-mers cannot fit the " + std::to_string(UC_SCRATCH_BYTES) +
                                       "-byte scratch of uc_compact_noalloc");
    HIPCHK(c, hipStreamSynchronize(c->stream));  // the context's own earlier work is done with its buffers
    // the context's own buffers step aside for the call and come back after it
    std::vector<DBuf *> bufs = all_bufs(c);
    std::vector<DBuf> saved;
    saved.reserve(bufs.size());
    for (DBuf *b : bufs) {
        saved.push_back(*b);
        *b = DBuf{};
    }
    const uint32_t flags = c->opts.flags;
    c->opts.flags = (flags & ~(UC_FLAG_PROFILE | UC_FLAG_COUNT_PROBES | UC_FLAG_KEEP_SRC)) | UC_FLAG_EAGER;
    c->arena = c->scratch_dev;
    c->arena_cap = usable;
    c->arena_used = 0;
    c->text = DBuf{text, text_cap};
    c->text_fixed = true;
    const uint64_t offs[2] = {0, n};
    int rc = do_load(c, packed, offs, 1, k, 1);
    if (rc == UC_OK) rc = do_bucketize(c, l, nullptr);
    if (rc == UC_OK) rc = do_edges(c, UC_ALL_BUCKETS);
    if (rc == UC_OK) rc = do_coarsen(c);
    if (rc == UC_OK) {
        *text_size = c->text_bytes;
        rc = do_emit(c);
    }
    const hipError_t se = hipStreamSynchronize(c->stream);
    if (rc == UC_OK && se != hipSuccess) rc = fail(c, UC_E_HIP, std::string("hipStreamSynchronize: ") + hipGetErrorString(se));
    for (size_t i = 0; i < bufs.size(); i++) *bufs[i] = saved[i];
    c->arena = nullptr;
    c->arena_cap = c->arena_used = 0;
    c->text_fixed = false;
    c->opts.flags = flags;
    // nothing of this job stays readable through the context: its buffers were the scratch
    reset_job(c);
    c->kin = nullptr;
    return rc;
}

int uc_compact_loaded(uc_ctx *c, uint32_t l)
{
    if (!c) return UC_E_ARG;
    if (!c->loaded) return fail(c, UC_E_STATE, "uc_compact_loaded before uc_load");
    CHK(select_dev(c));
    return timed_phases(c, l);
}

int uc_kernel_stat(uc_ctx *c, uint32_t idx, char *name, uint32_t name_cap, double *total_ms, uint64_t *launches)
{
    if (!c) return UC_E_ARG;
    CHK(prof_flush(c));
    if (idx >= c->kstats.size()) return UC_E_ARG;
    const auto &k = c->kstats[idx];
    if (name && name_cap) {
        std::strncpy(name, k.name.c_str(), name_cap - 1);
        name[name_cap - 1] = 0;
    }
    if (total_ms) *total_ms = k.ms;
    if (launches) *launches = k.launches;
    return UC_OK;
}

int uc_kernel_stats_select(uc_ctx *c, const char *name)
{
    if (!c) return UC_E_ARG;
    CHK(prof_flush(c));
    c->prof_only = name ? name : "";
    return UC_OK;
}

int uc_kernel_stats_reset(uc_ctx *c)
{
    if (!c) return UC_E_ARG;
    CHK(prof_flush(c));
    for (auto &k : c->kstats) {
        k.ms = 0;
        k.launches = 0;
    }
    return UC_OK;
}

int uc_compact_sets(uc_ctx *c, const uint32_t *packed, const uint64_t *set_offsets, uint32_t n_sets, uint32_t k,
                    uint32_t l, int on_device)
{
    if (!c || !set_offsets || n_sets == 0) return UC_E_ARG;
    if ((uint64_t)n_sets * c->B > 0x7FFFFFFFull) return fail(c, UC_E_LIMIT, "too many (set, bucket) pairs");
    CHK(select_dev(c));
    return timed_job(c, packed, set_offsets, n_sets, k, l, on_device);
}


int uc_load(uc_ctx *c, const uint32_t *packed, const uint64_t *set_offsets, uint32_t n_sets, uint32_t k, int on_device)
{
    if (!c || !set_offsets) return UC_E_ARG;
    CHK(select_dev(c));
    return do_load(c, packed, set_offsets, n_sets, k, on_device);
}

int uc_bucketize(uc_ctx *c, uint32_t l, const uint32_t *order_table, int table_on_device)
{
    if (!c) return UC_E_ARG;
    CHK(select_dev(c));
    const uint32_t *t = order_table;
    if (order_table && !table_on_device) {
        if (l < 1 || l > 15) return fail(c, UC_E_ARG, "l must be in [1, 15]");
        const size_t bytes = (size_t)4 << (2 * l);
        CHK(ensure(c, c->otable, bytes));
        CHK(h2d(c, c->otable.p, order_table, bytes));
        t = P<uint32_t>(c->otable);
    }
    return do_bucketize(c, l, t);
}

int uc_bucket_info_get(uc_ctx *c, uc_bucket_info *info)
{
    if (!c || !info) return UC_E_ARG;
    if (!c->bucketed) return fail(c, UC_E_STATE, "no bucketing result");
    CHK(select_dev(c));
    CHK(host_rows(c));
    info->n_kmers = c->n;
    info->m_placements = c->m;
    info->n_cross = c->x;
    info->n_buckets = c->B;
    info->n_sets = c->n_sets;
    info->largest_bucket = c->largest;
    info->n_fine_buckets = c->fboff_h.empty() ? 0 : c->fboff_h.size() - 1;
    info->bucket_tile = c->tile1;
    info->fine_tile = c->tile2;
    info->lds_limit = c->lds_limit;
    info->largest_fine_bucket = 0;
    info->n_oversized = 0;
    info->oversized_slots = 0;
    const uint64_t G = (uint64_t)c->n_sets * c->B;
    info->n_big_lds = c->edges_ready ? c->bg_first[G] : 0;
    info->n_slice_tasks = c->edges_ready ? c->sl_first[G] : 0;
    info->n_hbm_tables = c->edges_ready ? c->fb_first[G] : 0;
    const uint64_t group_max = c->opts.group_slots ? c->opts.group_slots : LDS_GROUP;
    for (size_t f = 0; f + 1 < c->fboff_h.size(); f++) {
        const uint64_t sz = c->fboff_h[f + 1] - c->fboff_h[f];
        info->largest_fine_bucket = std::max<uint64_t>(info->largest_fine_bucket, sz);
        if (sz > group_max) {
            info->n_oversized++;
            info->oversized_slots += sz;
        }
    }
    return UC_OK;
}

int uc_bucket_copy(uc_ctx *c, uint64_t *bucket_rows, uint32_t *slot_src, uint8_t *slot_flags, uint32_t *pair_home,
                   uint32_t *pair_borrow)
{
    if (!c) return UC_E_ARG;
    if (!c->bucketed) return fail(c, UC_E_STATE, "no bucketing result");
    CHK(select_dev(c));
    CHK(host_rows(c));
    if (bucket_rows)
        for (size_t i = 0; i < c->boff_h.size(); i++) bucket_rows[i] = c->boff_h[i];
    if (slot_src) {
        if (!(c->opts.flags & UC_FLAG_KEEP_SRC)) return fail(c, UC_E_STATE, "slot ids need UC_FLAG_KEEP_SRC");
        CHK(d2h(c, slot_src, c->src.p, c->m * 4));
    }
    if (slot_flags) {
        CHK(d2h(c, slot_flags, c->flags.p, c->m));
        for (uint64_t i = 0; i < c->m; i++) slot_flags[i] &= (F_LENT | F_BORROW);
    }
    if ((pair_home || pair_borrow) && c->x && !(c->pair_h.p && c->pair_b.p))
        return fail(c, UC_E_STATE, "the pairs were freed (UC_FLAG_LOW_MEMORY)");
    if (pair_home) CHK(d2h(c, pair_home, c->pair_h.p, c->x * 4));
    if (pair_borrow) CHK(d2h(c, pair_borrow, c->pair_b.p, c->x * 4));
    return UC_OK;
}

int uc_find_edges(uc_ctx *c, uint32_t bucket)
{
    if (!c) return UC_E_ARG;
    CHK(select_dev(c));
    return do_edges(c, bucket);
}

int uc_edge_probes(uc_ctx *c, uint64_t *probes)
{
    if (!c || !probes) return UC_E_ARG;
    // (after a job, also one whose edge result UC_FLAG_LOW_MEMORY has freed: the counters stay)
    if (!c->edged && !c->coarsened) return fail(c, UC_E_STATE, "no edge result");
    if (!(c->opts.flags & UC_FLAG_COUNT_PROBES)) return fail(c, UC_E_STATE, "probes are counted with UC_FLAG_COUNT_PROBES");
    if (!c->probes.p) return fail(c, UC_E_STATE, "no probe counters");
    CHK(select_dev(c));
    uint64_t v[PROBE_SLOTS];
    CHK(d2h(c, v, c->probes.p, sizeof(v)));
    *probes = 0;
    for (uint64_t x : v) *probes += x;
    return UC_OK;
}

int uc_succ_copy(uc_ctx *c, uint32_t *succ)
{
    if (!c || !succ) return UC_E_ARG;
    if (!c->edged) return fail(c, UC_E_STATE, "no edge result");
    CHK(select_dev(c));
    CHK(ensure(c, c->hit, c->m * 4));  // (UC_FLAG_LOW_MEMORY frees it after the edge phase)
    if (c->m) KLAUNCH(c, "k_link_succ", k_link_succ, (std::min<uint32_t>(grid_for(c->m), 65536), BLOCK, 0),
                      P<uint32_t>(c->link), P<uint8_t>(c->flags), (uint32_t)c->m, !c->spliced, P<uint32_t>(c->hit));
    return d2h(c, succ, c->hit.p, c->m * 4);
}

int uc_coarsen(uc_ctx *c)
{
    if (!c) return UC_E_ARG;
    CHK(select_dev(c));
    return do_coarsen(c);
}

int uc_coarsen_info_get(uc_ctx *c, uc_coarsen_info *info)
{
    if (!c || !info) return UC_E_ARG;
    if (!c->coarsened) return fail(c, UC_E_STATE, "no contraction result");
    info->n_unitigs = c->n_unitigs;
    info->n_cycles = c->n_cycles;
    info->text_bytes = c->text_bytes;
    info->n_splitters = c->n_splitters;
    info->n_runs = c->n_runs;
    info->chain_rounds = c->chain_rounds;
    return UC_OK;
}

int uc_emit(uc_ctx *c)
{
    if (!c) return UC_E_ARG;
    CHK(select_dev(c));
    return do_emit(c);
}

uint64_t uc_text_size(const uc_ctx *c) { return (c && c->emitted) ? c->text_bytes : 0; }

const char *uc_text_device_ptr(const uc_ctx *c) { return (c && c->emitted) ? (const char *)c->text.p : nullptr; }

int uc_text_copy(uc_ctx *c, char *dst, uint64_t cap)
{
    if (!c || (!dst && c->text_bytes)) return UC_E_ARG;
    if (!c->emitted) return fail(c, UC_E_STATE, "no text");
    if (cap < c->text_bytes) return fail(c, UC_E_ARG, "destination too small");
    CHK(select_dev(c));
    return staged_d2h(c, dst, c->text.p, c->text_bytes);
}

uint64_t uc_set_text_offset(const uc_ctx *c, uint32_t set)
{
    if (!c || !c->emitted || set > c->n_sets) return 0;
    return c->set_text[set];
}

int uc_text_append_file(uc_ctx *c, const char *path)
{
    if (!c || !path) return UC_E_ARG;
    if (!c->emitted) return fail(c, UC_E_STATE, "no text");
    CHK(select_dev(c));
    return append_text(c, path);
}

int uc_load_text(uc_ctx *c, const char *text, uint64_t len, uint32_t k, int on_device)
{
    if (!c || (!text && len)) return UC_E_ARG;
    CHK(select_dev(c));
    return do_load_text(c, text, len, k, on_device);
}

int uc_load_file(uc_ctx *c, const char *path, uint32_t k)
{
    if (!c || !path) return UC_E_ARG;
    if (k < 2 || k > 64) return fail(c, UC_E_ARG, "k must be in [2, 64]");
    CHK(select_dev(c));
    return do_load_file(c, path, k);
}

int uc_load_packed_file(uc_ctx *c, const char *path, uint32_t *k)
{
    if (!c || !path) return UC_E_ARG;
    CHK(select_dev(c));
    return do_load_packed_file(c, path, k);
}

const uint64_t *uc_loaded_device_ptr(const uc_ctx *c) { return (c && c->loaded) ? c->kin : nullptr; }

int uc_loaded_copy(uc_ctx *c, uint32_t *packed, uint64_t cap, uint64_t *n)
{
    if (!c) return UC_E_ARG;
    if (!c->loaded) return fail(c, UC_E_STATE, "nothing loaded");
    if (n) *n = c->n;
    if (!packed) return UC_OK;
    if (cap < c->n) return fail(c, UC_E_ARG, "destination too small");
    CHK(select_dev(c));
    const uint32_t W = c->W, rw = (c->k + 15) / 16;
    std::vector<uint64_t> h((size_t)c->n * W);
    CHK(d2h(c, h.data(), c->kin, h.size() * 8));
    for (uint64_t i = 0; i < c->n; i++)
        for (uint32_t j = 0; j < rw; j++) packed[i * rw + j] = (uint32_t)(h[i * W + j / 2] >> (32 * (j & 1)));
    return UC_OK;
}

int uc_timing_get(uc_ctx *c, uc_timing *t)
{
    if (!c || !t) return UC_E_ARG;
    *t = c->timing;
    return UC_OK;
}

int uc_sync(uc_ctx *c)
{
    if (!c) return UC_E_ARG;
    CHK(select_dev(c));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return UC_OK;
}

int uc_autotune(uc_ctx *c, uint32_t l, uint32_t reps, uint32_t tune_flags, uc_tune_result *out)
{
    if (!c || !out) return UC_E_ARG;
    if (!c->loaded) return fail(c, UC_E_STATE, "uc_autotune before a load");
    if (c->opts.flags & UC_FLAG_LOW_MEMORY)
        return fail(c, UC_E_ARG, "uc_autotune reruns the loaded job, which UC_FLAG_LOW_MEMORY consumes");
    CHK(select_dev(c));
    if (reps == 0) reps = 2;
    struct Cand {
        uint32_t B, fine, group, s, load;
    };
    // (fine_slots 0 is the library's own choice, the default target with the
    // capped rule for sets that want more than MAX_SUB fine buckets)
    Cand cur{c->B, c->opts.fine_slots,
             c->opts.group_slots ? c->opts.group_slots : LDS_GROUP, c->opts.sample_log2 ? c->opts.sample_log2 : DEFAULT_SAMPLE_LOG2,
             c->opts.load_pct ? c->opts.load_pct : 50};
    uint32_t trials = 0;
    auto apply = [&](const Cand &x) {
        c->B = x.B;
        c->opts.n_buckets = x.B;
        c->opts.fine_slots = x.fine;
        c->opts.group_slots = x.group;
        c->opts.sample_log2 = x.s;
        c->opts.load_pct = x.load;
    };
    auto time_of = [&](const Cand &x, float *ms) -> int {
        apply(x);
        trials++;
        CHK(timed_phases(c, l));
        float best = 1e30f;
        for (uint32_t r = 0; r < reps; r++) {
            CHK(timed_phases(c, l));
            best = std::min(best, c->timing.total_ms);
        }
        *ms = best;
        return UC_OK;
    };
    float best_ms;
    CHK(time_of(cur, &best_ms));
    const float default_ms = best_ms;
    const std::vector<uint32_t> Bs = {32, 64, 128, 256}, fines = {0, 256, 384, 512, 768, 1024},
                                groups = {1024, 2048}, ss = {4, 6, 8, 9, 10, 12}, loads = {35, 50, 70};
    for (int pass = 0; pass < 2; pass++) {
        bool moved = false;
        for (int dim = (tune_flags & UC_TUNE_KEEP_BUCKETS) ? 1 : 0; dim < 5; dim++) {
            const std::vector<uint32_t> &vals = dim == 0 ? Bs : dim == 1 ? fines : dim == 2 ? groups : dim == 3 ? ss : loads;
            for (uint32_t v : vals) {
                Cand x = cur;
                uint32_t &f = dim == 0 ? x.B : dim == 1 ? x.fine : dim == 2 ? x.group : dim == 3 ? x.s : x.load;
                if (f == v) continue;
                f = v;
                if (x.B > MAX_BUCKETS) continue;
                float ms;
                CHK(time_of(x, &ms));
                if (ms < best_ms * 0.99f) {
                    best_ms = ms;
                    cur = x;
                    moved = true;
                }
            }
        }
        if (!moved) break;
    }
    apply(cur);
    out->n_buckets = cur.B;
    out->fine_slots = cur.fine;
    out->group_slots = cur.group;
    out->sample_log2 = cur.s;
    out->load_pct = cur.load;
    out->trials = trials;
    out->default_ms = default_ms;
    out->best_ms = best_ms;
    return UC_OK;
}

int uc_ctx_trim(uc_ctx *c)
{
    if (!c) return UC_E_ARG;
    CHK(select_dev(c));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    drop_pass_graph(c);
    for (DBuf *b : all_bufs(c)) {
        if (b->p && !in_pool(c, b->p)) HIPCHK(c, hipFree(b->p));
        b->p = nullptr;
        b->cap = 0;
    }
    if (c->pool.base) HIPCHK(c, hipFree(c->pool.base));
    c->pool = Pool{};
    free_pinned(c);
    if (c->meter) c->meter->sub(c->held);
    c->held = 0;
    c->kin = nullptr;
    reset_job(c);
    return UC_OK;
}

uint64_t uc_ctx_held_bytes(const uc_ctx *c) { return c ? c->held : 0; }

}  // extern "C"

void uc_ctx_set_meter(uc_ctx *c, uc_meter *m)
{
    if (!c) return;
    if (c->meter) c->meter->sub(c->held);
    c->meter = m;
    if (m) m->add(c->held);
}

void *uc_host_pinned_alloc(uint64_t bytes)
{
    void *p = nullptr;
    // coherent: uc_text_copy_pinned may write it from a kernel
    if (hipHostMalloc(&p, bytes ? bytes : 1, hipHostMallocCoherent) != hipSuccess) return nullptr;
    return p;
}

void uc_host_pinned_free(void *p)
{
    if (p) (void)hipHostFree(p);
}

int uc_text_copy_pinned(uc_ctx *c, char *dst, uint64_t cap)
{
    if (!c || (!dst && c->text_bytes)) return UC_E_ARG;
    if (!c->emitted) return fail(c, UC_E_STATE, "no text");
    if (cap < c->text_bytes) return fail(c, UC_E_ARG, "destination too small");
    CHK(select_dev(c));
    if (!c->text_bytes) return UC_OK;
    void *d = c->kernel_copies ? host_mapping(dst) : nullptr;
    if (d)
        CHK(host_copy(c, d, c->text.p, c->text_bytes));
    else
        HIPCHK(c, hipMemcpyAsync(dst, c->text.p, c->text_bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return UC_OK;
}

void uc_ctx_set_kernel_copies(uc_ctx *c, bool on)
{
    if (!c || on == c->kernel_copies) return;
    (void)hipSetDevice(c->opts.device);
    (void)hipStreamSynchronize(c->stream);
    // the read buffer is written by k_host_copy now: coherent, so that the
    // kernel's stores reach host memory without a cache write-back
    if (c->rbuf) (void)hipHostFree(c->rbuf);
    c->rbuf = c->rbuf_dev = nullptr;
    if (hipHostMalloc((void **)&c->rbuf, RBUF_BYTES, on ? hipHostMallocCoherent : hipHostMallocDefault) != hipSuccess)
        c->rbuf = nullptr;
    if (on && c->rbuf) c->rbuf_dev = static_cast<char *>(host_mapping(c->rbuf));
    (void)hipGetLastError();
    c->kernel_copies = on;
}


extern "C" {

int uc_check_placements(uint64_t n, uint64_t x)
{
    if (x > n) return UC_E_STATE;
    if (n + x >= 0xFFFFFFFFull) return UC_E_LIMIT;
    return UC_OK;
}

uint64_t uc_job_bytes_estimate(uint64_t n, uint32_t k)
{
    // An upper bound whatever the crossing rate: every k-mer has one home
    // slot and at most one borrowed copy, so m = n + x <= 2n slots.
    // Per k-mer: the input copy, the bucket codes and the text of n isolated
    // k-mers (the most text a job can have).  Per slot: k-mers and flags
    // twice, fine index, permutation, slot sources (UC_FLAG_KEEP_SRC), the
    // fine-sort histogram (at most 4 bytes per fine bucket and 2048-slot
    // tile), hit, link, bitmap, run record and character, HBM hash tables
    // (at most two 16-byte entries per key), the contraction arrays
    // (segment, list entry, previous segment, unitig length, two chain
    // buffers, list, and the cycle buffers), head text offsets.
    const uint64_t W = k <= 32 ? 1 : 2;
    const uint64_t per_kmer = 8 * W + 8 + (k + 1);
    const uint64_t per_slot = (16 * W + 2) + 2 + 4 + 8 + 2 + (4 + 4 + 1 + 8 + 1) + 32 +
                              (8 + 8 + 4 + 4 + 8 + 8 + 4 + 8 + 4 + 4) + 8;
    return n * per_kmer + 2 * n * per_slot + (64ull << 20);
}

}  // extern "C"

namespace {
__global__ void k_order_table(uint32_t *t, uint64_t n, LminParams p)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        t[i] = lmer_order((uint32_t)i, p);
}
}  // namespace

extern "C" int uc_lmer_order_table(uc_ctx *c, uint32_t l, uint32_t *table, int on_device)
{
    if (!c || !table || l < 1 || l > 15) return UC_E_ARG;
    CHK(select_dev(c));
    LminParams p{};
    const uint32_t bits = 2 * l;
    p.lmask = (1u << bits) - 1u;
    p.h1 = (bits + 1) / 2;
    p.h2 = bits / 3 + 1;
    p.s0 = (uint32_t)c->opts.seed & p.lmask;
    p.w1 = bits - p.h1;
    p.w2 = bits - p.h2;
    p.k1 = p.s0 * MIX_C1;
    p.table = nullptr;
    const uint64_t n = 1ull << bits;
    uint32_t *dst = table;
    if (!on_device) {
        CHK(ensure(c, c->otable, n * 4));
        dst = P<uint32_t>(c->otable);
    }
    KLAUNCH(c, "k_order_table", k_order_table, (std::min<uint32_t>(grid_for(n), 65536), BLOCK, 0), dst, n, p);
    if (!on_device) return d2h(c, table, dst, n * 4);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return UC_OK;
}
