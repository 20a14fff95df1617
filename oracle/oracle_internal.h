/* oracle_internal.h -- helpers shared by the oracle's C files (test infrastructure only). */
#ifndef UC_ORACLE_INTERNAL_H
#define UC_ORACLE_INTERNAL_H

#include <stdint.h>

/* odd multipliers below 2^24 of the l-mer order bijection */
#define ORC_MIX_C1 0x3779B1u
#define ORC_MIX_C2 0xEBCA6Bu

typedef unsigned __int128 orc_u128;

/* k-mer as a 128-bit integer: character i at bits 2i (the reference packing
 * read as one little-endian integer). */
static inline orc_u128 orc_kmer_value(const uint32_t *kmer, uint32_t words)
{
    orc_u128 v = 0;
    for (uint32_t i = 0; i < words; i++) v |= (orc_u128)kmer[i] << (32 * i);
    return v;
}

static inline orc_u128 orc_low_mask(uint32_t bits)
{
    if (bits >= 128) return ~(orc_u128)0;
    return (((orc_u128)1) << bits) - 1;
}

static inline uint32_t orc_char_at(orc_u128 v, uint32_t i) { return (uint32_t)(v >> (2 * i)) & 3u; }

#endif
