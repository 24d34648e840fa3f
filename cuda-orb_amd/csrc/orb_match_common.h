/*
 * Device helpers shared by the exact matchers (orb_match.hip) and the
 * bucketed search of the learned matcher (orb_lindex.hip).
 *
 * A (query, entry) pair is the packed key (distance << 23 | index): the
 * minimum key is the nearest entry with ties to the lowest index, and a
 * running top-2 is (min(k1, key), med3(k1, key, k2)).  For the INT8 MFMA path
 * the bits are expanded to +/-1 bytes and distance = (256 - dot) / 2.
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

__device__ __forceinline__ void top2_push(uint32_t& k1, uint32_t& k2, uint32_t key)
{
    k2 = min(k2, max(k1, key)); /* med3(k1, key, k2) since k1 <= k2 */
    k1 = min(k1, key);
}

/* 4 bits -> 4 bytes of +1 (bit set) / -1 (bit clear) */
__device__ __forceinline__ int pm1_nibble(uint32_t n)
{
    const uint32_t t = (n * 0x00204081u) & 0x01010101u;
    return (int)~(t * 0xFEu);
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
   descriptor: lane k-group g takes bits 64*ks + 16*g .. +15 */
__device__ __forceinline__ void query_operands(uint4 d0, uint4 d1, int g, v4i* bq)
{
    const uint32_t w[8] = {d0.x, d0.y, d0.z, d0.w, d1.x, d1.y, d1.z, d1.w};
#pragma unroll
    for (int ks = 0; ks < 4; ks++) {
        const int hw = 4 * ks + g;
        bq[ks] = pm1_16((w[hw >> 1] >> (16 * (hw & 1))) & 0xffffu);
    }
}

/* key of a +/-1 dot product: ((256 - dot) / 2) << 23 | idx = (256 - dot) << 22 | idx */
__device__ __forceinline__ uint32_t dot_key(int dot, uint32_t idx)
{
    return (uint32_t)((256u << 22) + idx) - ((uint32_t)dot << 22);
}

} // namespace orbk
