/*
 * Device helpers shared by the exact matchers (orb_match.hip) and the
 * bucketed search of the learned matcher (orb_lindex.hip).
 *
 * A (query, entry) pair is the packed key (distance << 23 | index): the
 * minimum key is the nearest entry with ties to the lowest index, and a
 * running top-2 is (min(k1, key), med3(k1, key, k2)).  The INT8 MFMA paths
 * expand the query's bits to +/-1 bytes and the database entry's either to
 * 0/1 bytes (dot = (both set) - (entry set, query clear), distance =
 * popcount(query) - dot: fewer unpack instructions) or to +/-1 bytes
 * (distance = (256 - dot) / 2: no per-query term, fewer registers).  The
 * query operand is negated, so the accumulators hold -dot and a key is a
 * single shift-add.
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

/* k2 = med3(k1, key, k2) (k1 <= k2 always holds) and k1 = min(k1, key).
   Written as min/max so the compiler selects v_med3_u32 and v_min_u32
   itself and places any wait states the MFMA results they read need. */
__device__ __forceinline__ void top2_push(uint32_t& k1, uint32_t& k2, uint32_t key)
{
    const uint32_t lo = min(k1, key), hi = max(k1, key);
    k2 = max(lo, min(hi, k2));
    k1 = lo;
}

/* 16 bits -> 16 bytes of 1 (bit set) / 0 (bit clear) */
__device__ __forceinline__ v4i b01_16(uint32_t h)
{
    v4i r;
    r.x = (int)(((h & 15u) * 0x00204081u) & 0x01010101u);
    r.y = (int)((((h >> 4) & 15u) * 0x00204081u) & 0x01010101u);
    r.z = (int)((((h >> 8) & 15u) * 0x00204081u) & 0x01010101u);
    r.w = (int)((((h >> 12) & 15u) * 0x00204081u) & 0x01010101u);
    return r;
}

/* 4 bits -> 4 bytes of +1 (bit set) / -1 (bit clear): the multiply spreads
   bit i to bit 2 of byte i, so each selector byte is 4 (a byte of the first
   source, 0x01) or 0 (a byte of the second, 0xFF) */
__device__ __forceinline__ int pm1_nibble(uint32_t n)
{
    return (int)__builtin_amdgcn_perm(0x01010101u, 0xFFFFFFFFu, (n * 0x00810204u) & 0x04040404u);
}

/* 16 bits -> 16 +/-1 bytes */
__device__ __forceinline__ v4i pm1_16(uint32_t h)
{
    v4i r;
    r.x = pm1_nibble(h & 15u);
    r.y = pm1_nibble((h >> 4) & 15u);
    r.z = pm1_nibble((h >> 8) & 15u);
    r.w = pm1_nibble((h >> 12) & 15u);
    return r;
}

/* B operands of one query for the four 64-deep K steps of a 256-bit
   descriptor, negated: lane k-group g takes bits 64*ks + 16*g .. +15 and
   expands their complement (-1 where the bit is set, +1 where clear), so the
   MFMA accumulates -dot and a key is one shift-add of the accumulator */
__device__ __forceinline__ void query_operands_neg(uint4 d0, uint4 d1, int g, v4i* bq)
{
    const uint32_t w[8] = {d0.x, d0.y, d0.z, d0.w, d1.x, d1.y, d1.z, d1.w};
#pragma unroll
    for (int ks = 0; ks < 4; ks++) {
        const int hw = 4 * ks + g;
        bq[ks] = pm1_16(~(w[hw >> 1] >> (16 * (hw & 1))) & 0xffffu);
    }
}

/* key of a (+/-1 query, 0/1 entry) pair from the accumulated -dot:
   (popcount(query) - dot) << 23 | idx = (-dot << 23) + base_idx with
   base_idx = popcount(query) << 23 | idx (one v_lshl_add_u32) */
__device__ __forceinline__ uint32_t neg_dot_key(int negdot, uint32_t base_idx)
{
    return ((uint32_t)negdot << kIdxBits) + base_idx;
}

/* key of a (+/-1 query, +/-1 entry) pair from the accumulated -dot: the
   distance is (256 - dot) / 2 (dot is even), so the key is
   (-dot << 22) + (128 << 23) + idx modulo 2^32 */
__device__ __forceinline__ uint32_t pm1_neg_dot_key(int negdot, uint32_t idx)
{
    return ((uint32_t)negdot << (kIdxBits - 1)) + ((128u << kIdxBits) + idx);
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
