/*
 * Wavefront-level helpers shared by the HIP sources (device code only):
 * LDS visibility inside one wavefront, lane prefix counts, the XCD-aware
 * workgroup order, and an exact top-k selection over 1024 keys held 16 per
 * lane.
 */
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace orbk
{

/* LDS written by lanes of this wavefront is visible to its other lanes */
__device__ __forceinline__ void wave_sync_lds()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

/* Workgroup b of n runs logical unit xcd_remap(b): the hardware deals
   workgroups round-robin to the 8 XCDs, so this gives every XCD (and its L2)
   one contiguous range of units, and neighbouring units, which share source
   rows, land on the same L2. */
__device__ __forceinline__ int xcd_remap(int b, int n)
{
    const int q = n >> 3, r = n & 7, x = b & 7, k = b >> 3;
    return x * q + min(x, r) + k;
}

/* number of lanes below this one whose bit is set in m */
__device__ __forceinline__ int lanes_below(uint64_t m)
{
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

/* Sum over the 16 lanes of this lane's row (lanes 16r .. 16r+15), returned
   to every lane of the row: two quad swaps, then the half-row and row
   mirrors (DPP).  Every lane of the wavefront must be active. */
__device__ __forceinline__ int row_sum16(int v)
{
    v += __builtin_amdgcn_update_dpp(0, v, 0xb1, 0xf, 0xf, false);   /* quad_perm [1,0,3,2] */
    v += __builtin_amdgcn_update_dpp(0, v, 0x4e, 0xf, 0xf, false);   /* quad_perm [2,3,0,1] */
    v += __builtin_amdgcn_update_dpp(0, v, 0x141, 0xf, 0xf, false);  /* row_half_mirror */
    v += __builtin_amdgcn_update_dpp(0, v, 0x140, 0xf, 0xf, false);  /* row_mirror */
    return v;
}

/* Inclusive prefix sum over the 64 lanes (lane i gets v[0] + .. + v[i]):
   row shifts by 1, 2, 4, 8 within each 16-lane row, then the last lane of
   row 0 added to rows 1 and 3 and that of row 1 to rows 2 and 3 (DPP row
   broadcasts).  Every lane of the wavefront must be active. */
__device__ __forceinline__ int wave_incl_scan(int v)
{
    v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, true);   /* row_shr:1 */
    v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, true);   /* row_shr:2 */
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, true);   /* row_shr:4 */
    v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, true);   /* row_shr:8 */
    v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);  /* row_bcast:15 */
    v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);  /* row_bcast:31 */
    return v;
}

/* Sum over the 8 lanes of this lane's half row (lanes 8h .. 8h+7), returned
   to every lane of it: two quad swaps, then the half-row mirror (DPP).
   Every lane of the wavefront must be active. */
__device__ __forceinline__ int half_row_sum8(int v)
{
    v += __builtin_amdgcn_update_dpp(0, v, 0xb1, 0xf, 0xf, false);   /* quad_perm [1,0,3,2] */
    v += __builtin_amdgcn_update_dpp(0, v, 0x4e, 0xf, 0xf, false);   /* quad_perm [2,3,0,1] */
    v += __builtin_amdgcn_update_dpp(0, v, 0x141, 0xf, 0xf, false);  /* row_half_mirror */
    return v;
}

/* order-preserving key of a finite float (-0 and +0 share one key, above 0) */
__device__ __forceinline__ uint32_t float_order_key(float f)
{
    const uint32_t u = f == 0.f ? 0u : __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

/* The k-th largest of the wavefront's 64 NS keys (every key below 2^(HIBIT+1)),
   found bit by bit from the top: T grows while at least k keys are >= T.
   Keys 0 are never counted (every probe value is >= 1). */
template <int NS, int HIBIT>
__device__ __forceinline__ uint32_t wave_kth_key(const uint32_t* key, int k)
{
    uint32_t T = 0;
#pragma unroll
    for (int b = HIBIT; b >= 0; b--) {
        const uint32_t c = T | (1u << b);
        int cnt = 0;
#pragma unroll
        for (int s = 0; s < NS; s++) cnt += __popcll(__ballot(key[s] >= c));
        if (cnt >= k) T = c;
    }
    return T;
}

/* The wavefront's 64 values sorted in descending order across the lanes
   (lane 0 the largest): a bitonic network of 21 compare-exchange steps, the
   partner's value by a lane shuffle.  Every lane must be active. */
__device__ __forceinline__ uint32_t wave_sort_desc(uint32_t v)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int k = 2; k <= 64; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            const uint32_t o = (uint32_t)__shfl_xor((int)v, j, 64);
            /* blocks of k lanes alternate direction, the whole wave descending
               at k = 64; the lower lane of a pair keeps the larger value in a
               descending block */
            const bool keep_max = ((lane & j) == 0) == ((lane & k) == 0);
            v = keep_max ? max(v, o) : min(v, o);
        }
    }
    return v;
}

/* Top-k of up to 1024 entries, entry lane + 64 t holding key[t] (with PAIRS
   entry 128 (t / 2) + 2 lane + t % 2: a lane's slot pairs are adjacent
   entries) (key 0 marks an absent entry, every key below 2^(HIBIT+1); at least k entries present,
   1 <= k <= 1024): writes the k entry numbers of the largest keys, equal keys
   taken by ascending entry number, to out[0 .. k) -- with SORTED by
   descending key (equal keys by ascending entry number), otherwise in entry
   order.  cand is this wavefront's LDS buffer of CAP entries; with CAP below
   1024, more than CAP candidates (ties at the k-th lane maximum) are handled
   from the slots directly (k <= 64), and only k > 64 makes it return false
   having written nothing; otherwise it returns true.

   1. For k <= 64 the k-th largest lane maximum L bounds the answer from below
      (k lanes hold a key >= L); entries >= L are compacted into cand in entry
      order.  2. The exact k-th largest key T of those.  3. Entries above T
      (fewer than k): SORTED ranks them against each other, otherwise they
      keep entry order; entries equal to T take the remaining places in entry
      order. */
template <int HIBIT, bool SORTED, int CAP = 1024, bool PAIRS = false>
__device__ __forceinline__ bool wave_topk(const uint32_t (&key)[16], int k, uint2* cand, int* __restrict__ out)
{
    static_assert(CAP % 64 == 0 && CAP >= 64 && CAP <= 1024, "whole slots of 64 candidates");
    constexpr int NSL = CAP / 64;
    const int lane = threadIdx.x & 63;
    uint32_t L = 1;
    if (k <= 64) {
        uint32_t lm[1] = {0};
#pragma unroll
        for (int t = 0; t < 16; t++) lm[0] = max(lm[0], key[t]);
        L = wave_kth_key<1, HIBIT>(lm, k);
    }
    int nc = 0;
    if (PAIRS) {
        /* entry order: a lane's pair after the pairs of the lanes below it */
#pragma unroll
        for (int t = 0; t < 8; t++) {
            const bool p0 = key[2 * t] >= L, p1 = key[2 * t + 1] >= L;
            const uint64_t m0 = __ballot(p0), m1 = __ballot(p1);
            const int pos = nc + lanes_below(m0) + lanes_below(m1);
            const uint32_t id = (uint32_t)(128 * t + 2 * lane);
            if (p0 && pos < CAP) cand[pos] = make_uint2(key[2 * t], id);
            if (p1 && pos + (int)p0 < CAP) cand[pos + (int)p0] = make_uint2(key[2 * t + 1], id + 1u);
            nc += __popcll(m0) + __popcll(m1);
        }
    } else {
#pragma unroll
        for (int t = 0; t < 16; t++) {
            const bool p = key[t] >= L;
            const uint64_t m = __ballot(p);
            const int pos = nc + lanes_below(m);
            if (p && pos < CAP) cand[pos] = make_uint2(key[t], (uint32_t)(lane + 64 * t));
            nc += __popcll(m);
        }
    }
    if (nc > CAP) {
        if (k > 64) return false;
        /* more candidates than fit (ties at the bound): the exact k-th key
           over every slot, then the entries above it and the first ones
           equal to it (at most k <= 64, so they fit), each group in entry
           order, taken straight from the slots */
        const uint32_t T = wave_kth_key<16, HIBIT>(key, k);
        wave_sync_lds();
        int base = 0;
#pragma unroll
        for (int pass = 0; pass < 2; pass++) {
            auto take = [&](uint32_t kv, uint32_t id, int before) {
                const int pos = base + before;
                const bool p = pass == 0 ? kv > T : kv == T;
                if (p && pos < k) cand[pos] = make_uint2(kv, id);
            };
            if (PAIRS) {
#pragma unroll
                for (int t = 0; t < 8; t++) {
                    const bool p0 = pass == 0 ? key[2 * t] > T : key[2 * t] == T;
                    const bool p1 = pass == 0 ? key[2 * t + 1] > T : key[2 * t + 1] == T;
                    const uint64_t m0 = __ballot(p0), m1 = __ballot(p1);
                    const int before = lanes_below(m0) + lanes_below(m1);
                    const uint32_t id = (uint32_t)(128 * t + 2 * lane);
                    take(key[2 * t], id, before);
                    take(key[2 * t + 1], id + 1u, before + (int)p0);
                    base += __popcll(m0) + __popcll(m1);
                }
            } else {
#pragma unroll
                for (int t = 0; t < 16; t++) {
                    const bool p = pass == 0 ? key[t] > T : key[t] == T;
                    const uint64_t m = __ballot(p);
                    take(key[t], (uint32_t)(lane + 64 * t), lanes_below(m));
                    base += __popcll(m);
                }
            }
        }
        wave_sync_lds();
        if (SORTED) {
            uint32_t v = 0;
            if (lane < k) {
                const uint2 e = cand[lane];
                v = (e.x << 10) | (1023u - e.y);
            }
            v = wave_sort_desc(v);
            if (lane < k) out[lane] = (int)(1023u - (v & 1023u));
        } else if (lane < k) {
            out[lane] = (int)cand[lane].y;
        }
        return true;
    }
    wave_sync_lds();
    const int S = (nc + 63) >> 6;
    uint32_t ck[NSL];
    int ci[NSL];
#pragma unroll
    for (int s = 0; s < NSL; s++) {
        const int j = lane + 64 * s;
        const uint2 e = (s < S && j < nc) ? cand[j] : make_uint2(0u, 0u);
        ck[s] = e.x;
        ci[s] = (int)e.y;
    }
    /* slots at and above S hold key 0: only the occupied ones are counted */
    constexpr int N2 = NSL < 2 ? NSL : 2, N4 = NSL < 4 ? NSL : 4, N8 = NSL < 8 ? NSL : 8;
    const uint32_t T = S <= 1   ? wave_kth_key<1, HIBIT>(ck, k)
                       : S <= 2 ? wave_kth_key<N2, HIBIT>(ck, k)
                       : S <= 4 ? wave_kth_key<N4, HIBIT>(ck, k)
                       : S <= 8 ? wave_kth_key<N8, HIBIT>(ck, k)
                                : wave_kth_key<NSL, HIBIT>(ck, k);
    wave_sync_lds();
    if (SORTED && k <= 64) {
        /* the k entries (those above T, then the first ones equal to T, each
           group in entry order) one per lane, ranked by a sort of
           key << 10 | (1023 - entry): descending key, ties by ascending entry */
        static_assert(HIBIT <= 21, "key and entry number share 32 bits");
        int base = 0;
#pragma unroll
        for (int s = 0; s < NSL; s++) {
            if (s < S) {
                const bool p = ck[s] > T;
                const uint64_t m = __ballot(p);
                if (p) cand[base + lanes_below(m)] = make_uint2(ck[s], (uint32_t)ci[s]);
                base += __popcll(m);
            }
        }
#pragma unroll
        for (int s = 0; s < NSL; s++) {
            if (s < S && base < k) {
                const bool p = ck[s] == T;
                const uint64_t m = __ballot(p);
                const int rank = base + lanes_below(m);
                if (p && rank < k) cand[rank] = make_uint2(ck[s], (uint32_t)ci[s]);
                base += __popcll(m);
            }
        }
        wave_sync_lds();
        uint32_t v = 0;
        if (lane < k) {
            const uint2 e = cand[lane];
            v = (e.x << 10) | (1023u - e.y);
        }
        v = wave_sort_desc(v);
        if (lane < k) out[lane] = (int)(1023u - (v & 1023u));
        return true;
    }
    int g = 0;
#pragma unroll
    for (int s = 0; s < NSL; s++) {
        if (s < S) {
            const bool p = ck[s] > T;
            const uint64_t m = __ballot(p);
            if (p) {
                if (SORTED) cand[g + lanes_below(m)] = make_uint2(ck[s], (uint32_t)ci[s]);
                else out[g + lanes_below(m)] = ci[s];
            }
            g += __popcll(m);
        }
    }
    if (SORTED) wave_sync_lds();
    for (int i = lane; SORTED && i < g; i += 64) {
        const uint2 e = cand[i];
        int rank = 0;
#pragma unroll 4
        for (int j = 0; j < g; j++) {
            const uint2 o = cand[j];
            rank += (o.x > e.x) | ((o.x == e.x) & (j < i));
        }
        out[rank] = (int)e.y;
    }
    int base = g;
#pragma unroll
    for (int s = 0; s < NSL; s++) {
        if (s < S && base < k) {
            const bool p = ck[s] == T;
            const uint64_t m = __ballot(p);
            const int rank = base + lanes_below(m);
            if (p && rank < k) out[rank] = ci[s];
            base += __popcll(m);
        }
    }
    return true;
}

/* Inclusive prefix sum of a float over the 64 lanes (the pattern of
   wave_incl_scan, FP32 adds in its tree order).  Every lane must be active. */
__device__ __forceinline__ float wave_incl_scan_f(float v)
{
#define ORB_DPP_ADD(ctrl, rmask, bc)                                                                       \
    v = __fadd_rn(v, __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), ctrl, rmask, 0xf, bc)))
    ORB_DPP_ADD(0x111, 0xf, true);  /* row_shr:1 */
    ORB_DPP_ADD(0x112, 0xf, true);  /* row_shr:2 */
    ORB_DPP_ADD(0x114, 0xf, true);  /* row_shr:4 */
    ORB_DPP_ADD(0x118, 0xf, true);  /* row_shr:8 */
    ORB_DPP_ADD(0x142, 0xa, false); /* row_bcast:15 */
    ORB_DPP_ADD(0x143, 0xc, false); /* row_bcast:31 */
#undef ORB_DPP_ADD
    return v;
}

/* Probability-mass rule over the k best candidates, lane r holding the
   candidate of rank r (r < k <= 64) with weight w: the candidate of rank r is
   kept iff the weights of ranks 0 .. r-1 (the inclusive FP32 prefix sum in
   the wave_incl_scan_f order, minus its own weight) sum to less than
   `limit` -- rank 0 is always kept and the kept set is a prefix of the
   ranking.  Writes out[r] = its id if kept, else -1, for r < k.  Every lane
   must be active. */
__device__ __forceinline__ void wave_mass_keep(int id, float w, int k, float limit, int* __restrict__ out)
{
    const int lane = threadIdx.x & 63;
    const float before = __fsub_rn(wave_incl_scan_f(lane < k ? w : 0.f), lane < k ? w : 0.f);
    if (lane < k) out[lane] = before < limit ? id : -1;
}

} // namespace orbk
