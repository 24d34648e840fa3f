/*
 * Device helpers shared by the exact matchers (orb_match.hip), the bucketed
 * search of the learned matcher (orb_lindex.hip) and the forward pass's
 * centroid term (orb_mlp.hip).
 *
 * A (query, entry) pair is the packed key (distance << 23 | index): the
 * minimum key is the nearest entry with ties to the lowest index, and a
 * running top-2 is (min(k1, key), med3(k1, key, k2)).  The matrix-core
 * paths contract descriptor bits as FP4 codes, one per bit: the query's as
 * negated +/-1 codes and the entry's as 0/1 codes (dot = (entry set, query
 * clear) - (both set), distance = popcount(query) + that: the exact
 * matchers) or as +/-1 codes (256 - 2 distance, no per-query term: the
 * bucketed search and the centroid term).
 */
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace orbk
{

constexpr uint32_t kIdxBits = 23;
constexpr uint32_t kIdxMask = (1u << kIdxBits) - 1u;
constexpr uint32_t kNoKey = 0xffffffffu;

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v8i __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));

/* ---- FP4 operands.  A descriptor bit becomes one E2M1 code (4 bits):
   0x0 = 0, 0x2 = +1.0, 0xA = -1.0, so one 128-deep FP4 MFMA contracts 128
   bits.  Products lie in {-1, 0, 1} and a contraction sums at most 256 of
   them, so the FP32 accumulator holds the exact integer. */

/* bits 0..7 of x to bits 0, 4, ..., 28 (one per nibble) */
__device__ __forceinline__ uint32_t spread_nibbles(uint32_t x)
{
    x &= 0xffu;
    x = (x | (x << 12)) & 0x000f000fu;
    x = (x | (x << 6)) & 0x03030303u;
    x = (x | (x << 3)) & 0x11111111u;
    return x;
}

/* 8 bits -> 8 codes of 1 (bit set) / 0 (clear) */
__device__ __forceinline__ uint32_t fp4_b01_8(uint32_t x)
{
    return spread_nibbles(x) << 1;
}

/* 8 bits -> 8 codes of +1 (bit set) / -1 (clear) */
__device__ __forceinline__ uint32_t fp4_pm1_8(uint32_t x)
{
    return 0x22222222u | (spread_nibbles(~x) << 3);
}

/* 8 bits -> 8 codes of -1 (bit set) / +1 (clear) */
__device__ __forceinline__ uint32_t fp4_neg_pm1_8(uint32_t x)
{
    return 0x22222222u | (spread_nibbles(x) << 3);
}

/* one 32-bit descriptor word -> 32 codes (16 bytes): code i (bits 4i..4i+3
   of the 16 bytes) is bit i of w */
template <typename F>
__device__ __forceinline__ v4i fp4_word(uint32_t w, F code8)
{
    v4i r;
    r.x = (int)code8(w);
    r.y = (int)code8(w >> 8);
    r.z = (int)code8(w >> 16);
    r.w = (int)code8(w >> 24);
    return r;
}

/* FP4 codes of every byte value: entry b is fp4_b01_8(b) (LDS lookup
   tables are filled from this at kernel start) */
__device__ __forceinline__ void fill_code_table(uint32_t* table, int tid, int nthreads)
{
    for (int b = tid; b < 256; b += nthreads) table[b] = fp4_b01_8((uint32_t)b);
}

/* 32 codes of a descriptor word from a 256-entry byte table in LDS */
__device__ __forceinline__ v4i fp4_word_lut(uint32_t w, const uint32_t* table)
{
    v4i r;
    r.x = (int)table[w & 0xffu];
    r.y = (int)table[(w >> 8) & 0xffu];
    r.z = (int)table[(w >> 16) & 0xffu];
    r.w = (int)table[w >> 24];
    return r;
}

/* C + A x B, 16 x 16 x 128 on FP4 operands (unit scales), FP32 accumulate.
   Lane (g = lane / 16, col = lane % 16) gives A row col and B column col,
   elements 32g .. 32g + 31 of the 128 (the same positions in A and B). */
__device__ __forceinline__ v4f mfma_fp4(v4i a, v4i b, v4f c)
{
    const v8i a8 = {a.x, a.y, a.z, a.w, 0, 0, 0, 0};
    const v8i b8 = {b.x, b.y, b.z, b.w, 0, 0, 0, 0};
    return __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(a8, b8, c, 4, 4, 0, 127, 0, 127);
}

/* k2 = med3(k1, key, k2) (k1 <= k2 always holds) and k1 = min(k1, key).
   Written as min/max so the compiler selects v_med3_u32 and v_min_u32
   itself and places any wait states the MFMA results they read need. */
__device__ __forceinline__ void top2_push(uint32_t& k1, uint32_t& k2, uint32_t key)
{
    const uint32_t lo = min(k1, key), hi = max(k1, key);
    k2 = max(lo, min(hi, k2));
    k1 = lo;
}

/* key of a (+/-1 query, 0/1 entry) pair from the accumulated -dot:
   (popcount(query) - dot) << 23 | idx = (-dot << 23) + base_idx with
   base_idx = popcount(query) << 23 | idx (one v_lshl_add_u32) */
__device__ __forceinline__ uint32_t neg_dot_key(int negdot, uint32_t base_idx)
{
    return ((uint32_t)negdot << kIdxBits) + base_idx;
}

/* For the 0/1-entry matcher: the largest -dot whose key can still be at or
   below k2 (the distance must not exceed k2's; k2 = kNoKey admits every
   entry since -dot <= 256 - popcount(query)) */
__device__ __forceinline__ int neg_dot_bound(int qpop, uint32_t k2)
{
    return (int)(k2 >> kIdxBits) - qpop;
}

__device__ __forceinline__ int popc256(uint4 d0, uint4 d1)
{
    return __popc(d0.x) + __popc(d0.y) + __popc(d0.z) + __popc(d0.w) + __popc(d1.x) + __popc(d1.y) + __popc(d1.z) +
           __popc(d1.w);
}

} // namespace orbk
