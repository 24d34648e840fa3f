/*
 * Exact brute-force Hamming matchers for 256-bit descriptors (gfx950).
 *
 * The reference's gHammingMatch (orbd.cu:1471-1559) gives each query a
 * 32-thread block that re-reads the whole train set.  Here both matchers cut
 * the (query, database) plane into units of a query tile x a database chunk,
 * sized on the device from the real counts so a grid launched for the
 * capacities still spreads the actual work over every CU.  The popcount
 * matcher merges the chunks inside its one launch (device-scope atomic
 * words, see k_match_pop); the MFMA matcher, used for large dense problems,
 * writes per-chunk results that a merge pass combines.
 *
 * Each (query, entry) pair becomes a packed key (distance << 23 | index): the
 * minimum key is the nearest entry with ties to the lowest index, and the
 * running top-2 is (min(k1, key), med3(k1, key, k2)).
 *
 *   popcount: a wavefront holds 64 queries (one per lane) and streams its
 *     chunk with wave-uniform (scalar) loads: 8 x (xor, popcount) per pair.
 *   MFMA: the distance is an exact integer contraction,
 *     popcount(a ^ b) = popcount(a) - <a', b01> with a' the +/-1 expansion
 *     of the query's bits and b01 the 0/1 expansion of the entry's, computed
 *     by v_mfma_i32_16x16x64_i8 (INT8 in, INT32 accumulate) with -a' as the
 *     operand, so the accumulator holds -<a', b01> and a key is one shift-add.  A workgroup
 *     keeps 256 queries as +/-1 B operands in registers and unpacks each
 *     64-entry database tile into LDS once, the next tile's bytes loaded
 *     during the current one; a lane builds keys only for results that can
 *     still enter its top-2.
 */
#include <mutex>

#include "orb_internal.h"
#include "orb_match_common.h"

namespace orbk
{

constexpr int kGranule = 64;          /* database entries per chunk granule (popcount: 16 per wavefront) */
constexpr int kMinChunkGranules = 2;  /* at least 128 entries per chunk: bounds the merge work */
constexpr int kPopQ = 64;             /* popcount: queries per wavefront */
constexpr int kPopTile = 256;         /* popcount: database entries per LDS tile (one per thread) */
constexpr int kMfmaQ = 512;           /* MFMA: queries per workgroup */
constexpr int kMfmaRow = 128 + 16;    /* LDS row of one unpacked entry: 8 chunks of 32 FP4 codes, padded */
/* Split nq queries x ndb entries into qunits x chunks units for `slots`
   parallel workers (identical on host and device). */
struct Plan {
    int qunits, chunks, chunk;
};

__host__ __device__ inline Plan make_plan(int nq, int ndb, int qtile, int slots)
{
    Plan p;
    p.qunits = (nq + qtile - 1) / qtile;
    const int granules = (ndb + kGranule - 1) / kGranule;
    int c = p.qunits > 0 ? slots / p.qunits : 1;
    const int cmax = (granules + kMinChunkGranules - 1) / kMinChunkGranules;
    if (c > cmax) c = cmax;
    if (c < 1) c = 1;
    p.chunk = granules > 0 ? ((granules + c - 1) / c) * kGranule : kGranule;
    p.chunks = ndb > 0 ? (ndb + p.chunk - 1) / p.chunk : 0;
    return p;
}

/* ----------------------------------------------------------- popcount */

/* distance contribution of one 32-bit word of a ^ b: its set bits, or with
   H2 its set bit pairs (bits 2i, 2i+1), OpenCV's NORM_HAMMING2 */
template <bool H2>
__device__ __forceinline__ uint32_t word_dist(uint32_t x)
{
    if constexpr (H2) x = (x | (x >> 1)) & 0x55555555u;
    return __popc(x);
}

/* Outputs of query qi from its final top-2 keys: raw (idx, distance) pairs,
   and the reference's acceptance rule into the point (orbd.cu:1544-1558: both
   distances start at 256; a unique best below 64 is accepted). */
__device__ __forceinline__ void write_match(int qi, uint32_t k1, uint32_t k2, int* idx1, int* d1, int* idx2, int* d2,
                                            orb::OrbPoint* pts)
{
    const int a1 = k1 == kNoKey ? -1 : (int)(k1 & kIdxMask), b1 = k1 == kNoKey ? 257 : (int)(k1 >> kIdxBits);
    const int a2 = k2 == kNoKey ? -1 : (int)(k2 & kIdxMask), b2 = k2 == kNoKey ? 257 : (int)(k2 >> kIdxBits);
    if (idx1) idx1[qi] = a1;
    if (d1) d1[qi] = b1;
    if (idx2) idx2[qi] = a2;
    if (d2) d2[qi] = b2;
    if (pts) {
        const int s1 = min(b1, 256), s2 = min(b2, 256);
        orb::OrbPoint* p = pts + qi;
        if (s1 < s2 && s1 < 64) {
            p->match = a1;
            p->distance = s1;
        } else {
            p->match = -1;
            p->distance = -1;
        }
    }
}

/* A workgroup unit is 64 queries x one chunk; its 4 wavefronts scan a
   quarter of the chunk each and merge through LDS, so small problems still
   spread over many wavefronts without multiplying the merge work.

   The chunks of a query tile meet in one 64-bit word per query, (k1 << 32) |
   k2, merged by compare-and-swap at device scope; a counter per query tile
   tells the unit that finishes the tile's last chunk, which reads each word
   (resetting it to "no key"), writes the outputs and resets the counter, so
   the words and counters are clean for the next call and one launch does the
   whole match.  Every value that units exchange is a device-scope atomic, so
   no cache write-back is needed between them. */
template <bool H2>
__device__ __forceinline__ void match_pop_units(const uint8_t* __restrict__ q, int nq, const uint4* __restrict__ db,
                                                int ndb, int slots, int wg, int nwg,
                                                unsigned long long* __restrict__ state, int* __restrict__ done,
                                                int* __restrict__ idx1, int* __restrict__ d1,
                                                int* __restrict__ idx2, int* __restrict__ d2,
                                                orb::OrbPoint* __restrict__ pts)
{
    __shared__ uint2 s_k[3][64];
    __shared__ uint4 s_db[2][2 * kPopTile];
    const Plan P = make_plan(nq, ndb, kPopQ, slots);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (P.chunks == 0) {
        /* empty database: every query has no match */
        for (int qi = wg * 256 + threadIdx.x; qi < nq; qi += nwg * 256)
            write_match(qi, kNoKey, kNoKey, idx1, d1, idx2, d2, pts);
        return;
    }
    for (int u = wg; u < P.qunits * P.chunks; u += nwg) {
        const int qu = u % P.qunits, c = u / P.qunits;
        const int qi = qu * kPopQ + lane;
        uint4 qa = make_uint4(0, 0, 0, 0), qb = qa;
        if (qi < nq) {
            const uint4* qp = (const uint4*)(q + (long long)qi * 32);
            qa = qp[0];
            qb = qp[1];
        }
        const int j0 = min(ndb, c * P.chunk), j1 = min(ndb, j0 + P.chunk);
        uint32_t k1 = kNoKey, k2 = kNoKey;
        auto pair = [&](uint4 a, uint4 b, int j) {
            uint32_t d = word_dist<H2>(qa.x ^ a.x);
            d = word_dist<H2>(qa.y ^ a.y) + d;
            d = word_dist<H2>(qa.z ^ a.z) + d;
            d = word_dist<H2>(qa.w ^ a.w) + d;
            d = word_dist<H2>(qb.x ^ b.x) + d;
            d = word_dist<H2>(qb.y ^ b.y) + d;
            d = word_dist<H2>(qb.z ^ b.z) + d;
            d = word_dist<H2>(qb.w ^ b.w) + d;
            top2_push(k1, k2, (d << kIdxBits) | (uint32_t)j);
        };
        /* the chunk through LDS in tiles of 256 entries (one entry per thread,
           the next tile's loaded during the current one); wavefront wid takes
           the tile's entries wid, wid + 4, ..., every lane reading the same
           entry (a broadcast) */
        const int ntiles = (j1 - j0 + kPopTile - 1) / kPopTile;
        uint4 ea = make_uint4(0, 0, 0, 0), eb = ea;
        if (j0 + (int)threadIdx.x < j1) {
            ea = db[2 * (j0 + threadIdx.x)];
            eb = db[2 * (j0 + threadIdx.x) + 1];
        }
        for (int t = 0; t < ntiles; t++) {
            uint4* const tile = s_db[t & 1];
            tile[2 * threadIdx.x] = ea;
            tile[2 * threadIdx.x + 1] = eb;
            __syncthreads();
            const int jt = j0 + t * kPopTile, n = min(kPopTile, j1 - jt);
            if (t + 1 < ntiles && jt + kPopTile + (int)threadIdx.x < j1) {
                ea = db[2 * (jt + kPopTile + threadIdx.x)];
                eb = db[2 * (jt + kPopTile + threadIdx.x) + 1];
            }
            int e = wid;
            for (; e + 12 < n; e += 16) {
#pragma unroll
                for (int u = 0; u < 4; u++) pair(tile[2 * (e + 4 * u)], tile[2 * (e + 4 * u) + 1], jt + e + 4 * u);
            }
            for (; e < n; e += 4) pair(tile[2 * e], tile[2 * e + 1], jt + e);
        }
        if (wid > 0) s_k[wid - 1][lane] = make_uint2(k1, k2);
        __syncthreads();
        if (wid == 0) {
#pragma unroll
            for (int w = 0; w < 3; w++) {
                top2_push(k1, k2, s_k[w][lane].x);
                top2_push(k1, k2, s_k[w][lane].y);
            }
            if (P.chunks == 1) {
                if (qi < nq) write_match(qi, k1, k2, idx1, d1, idx2, d2, pts);
            } else {
                if (qi < nq) {
                    unsigned long long* wp = state + qi;
                    unsigned long long old = __hip_atomic_load(wp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    for (;;) {
                        const uint32_t o1 = (uint32_t)(old >> 32), o2 = (uint32_t)old;
                        uint32_t n1 = o1, n2 = o2;
                        top2_push(n1, n2, k1);
                        top2_push(n1, n2, k2);
                        if (n1 == o1 && n2 == o2) break;
                        if (__hip_atomic_compare_exchange_strong(wp, &old, ((unsigned long long)n1 << 32) | n2,
                                                                 __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                                 __HIP_MEMORY_SCOPE_AGENT))
                            break;
                    }
                }
                /* the merges above have completed before the counter moves */
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                int last = 0;
                if (lane == 0)
                    last = __hip_atomic_fetch_add(done + qu, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
                           P.chunks - 1;
                last = __shfl(last, 0, 64);
                if (last) {
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                    if (qi < nq) {
                        const unsigned long long v = __hip_atomic_exchange(state + qi, ~0ull, __ATOMIC_RELAXED,
                                                                           __HIP_MEMORY_SCOPE_AGENT);
                        write_match(qi, (uint32_t)(v >> 32), (uint32_t)v, idx1, d1, idx2, d2, pts);
                    }
                    if (lane == 0) __hip_atomic_store(done + qu, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
            }
        }
        __syncthreads();
    }
}

template <bool H2>
__global__ __launch_bounds__(256) void k_match_pop(const uint8_t* __restrict__ q, int nq_cap,
                                                   const int* __restrict__ d_nq, const uint4* __restrict__ db,
                                                   int ndb_cap, const int* __restrict__ d_ndb, int slots,
                                                   unsigned long long* __restrict__ state, int* __restrict__ done,
                                                   int* __restrict__ idx1, int* __restrict__ d1,
                                                   int* __restrict__ idx2, int* __restrict__ d2,
                                                   orb::OrbPoint* __restrict__ pts)
{
    const int nq = d_nq ? min(*d_nq, nq_cap) : nq_cap;
    const int ndb = d_ndb ? min(*d_ndb, ndb_cap) : ndb_cap;
    match_pop_units<H2>(q, nq, db, ndb, slots, blockIdx.x, gridDim.x, state, done, idx1, d1, idx2, d2, pts);
}

/* --------------------------------------------------------------- MFMA */

/* One MFMA unit of a workgroup of NW wavefronts: the 64*NW queries of tile
   qu (query qu*64*NW + wid*64 + qb*16 + col in lane col of k-group g)
   against database entries [j0, j1).  On return the top-2 keys of each query
   are merged over its 4 k-groups (every lane of a query holds them) and qidx
   holds the query indices.  Called by the whole workgroup (it synchronises
   through s_db, 2 tiles of kGranule * kMfmaRow bytes). */
template <int NW>
__device__ __forceinline__ void mfma_unit(const uint8_t* __restrict__ q, int nq, const uint8_t* __restrict__ db,
                                          int qu, int j0, int j1, int8_t* __restrict__ s_db,
                                          const uint32_t* __restrict__ s_code, uint32_t (&k1)[4],
                                          uint32_t (&k2)[4], int (&qidx)[4])
{
    constexpr int kThreads = 64 * NW, kWordsPerThread = kGranule * 8 / kThreads;
    static_assert(kWordsPerThread * kThreads == kGranule * 8, "whole descriptor words per thread");
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int g = lane >> 4, col = lane & 15;
    /* B operands: query (qb*16 + col) of this wave's 64, descriptor word
       4*ks + g (bits 128ks + 32g .. +31), negated (-1 where a bit is set) */
    v4i bq[4][2];
    /* bnd[qb]: a result can enter the top-2 only if its -dot is at most
       bnd = (k2 >> 23) - popcount(query), the current second distance
       as a -dot; "no key" admits every result until two are held.  The
       accumulators start at 256, so they hold 256 - dot >= 0, whose float
       bit patterns order like the values: the test is an unsigned compare
       with bbits[qb], the bits of 256 + bnd */
    int bnd[4];
    uint32_t bbits[4];
#pragma unroll
    for (int qb = 0; qb < 4; qb++) {
        const int qi = qu * kThreads + wid * 64 + qb * 16 + col;
        qidx[qb] = qi;
        uint4 d0 = make_uint4(0, 0, 0, 0), d1 = d0;
        if (qi < nq) {
            const uint4* qp = (const uint4*)(q + (long long)qi * 32);
            d0 = qp[0];
            d1 = qp[1];
        }
        const uint32_t w[8] = {d0.x, d0.y, d0.z, d0.w, d1.x, d1.y, d1.z, d1.w};
#pragma unroll
        for (int ks = 0; ks < 2; ks++) bq[qb][ks] = fp4_word(w[4 * ks + g], fp4_neg_pm1_8);
        bnd[qb] = neg_dot_bound(popc256(d0, d1), kNoKey);
        bbits[qb] = __float_as_uint((float)(256 + bnd[qb]));
        k1[qb] = k2[qb] = kNoKey;
    }
    const v4f c256 = {256.f, 256.f, 256.f, 256.f};
    /* unpack 64 entries (512 descriptor words) per tile: word tid + k*kThreads
       is word (tid + k*kThreads) % 8 of entry (tid + k*kThreads) / 8, row
       chunk of the same number, coded through the byte table.  Two LDS tile
       buffers: tile i+1 is unpacked into one while tile i is read from the
       other, so one barrier per tile separates them; the words of tile i+2
       are loaded meanwhile, through a buffer resource ending at entry j1
       (reads past it return 0) */
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc((void*)db, 0, j1 * 32, 0x00020000);
    auto fetch = [&](int j, uint32_t (&v)[kWordsPerThread]) {
#pragma unroll
        for (int k = 0; k < kWordsPerThread; k++) {
            const int wi = tid + k * kThreads;
            v[k] = __builtin_amdgcn_raw_buffer_load_b32(rs, (uint32_t)((j + (wi >> 3)) * 32 + (wi & 7) * 4), 0, 0);
        }
    };
    auto unpack = [&](const uint32_t (&v)[kWordsPerThread], int8_t* buf) {
#pragma unroll
        for (int k = 0; k < kWordsPerThread; k++) {
            const int wi = tid + k * kThreads;
            *(v4i*)(buf + (wi >> 3) * kMfmaRow + 16 * (wi & 7)) = fp4_word_lut(v[k], s_code);
        }
    };
    uint32_t nv[kWordsPerThread];
    __syncthreads();   /* the buffers' previous readers are done */
    fetch(j0, nv);
    unpack(nv, s_db);
    fetch(j0 + kGranule, nv);
    __syncthreads();
    int cur = 0;
    for (int jt = j0; jt < j1; jt += kGranule, cur ^= 1) {
        const int8_t* tile = s_db + cur * kGranule * kMfmaRow;
        if (jt + kGranule < j1) {
            unpack(nv, s_db + (cur ^ 1) * kGranule * kMfmaRow);
            fetch(jt + 2 * kGranule, nv);
        }
#pragma unroll
        for (int ab = 0; ab < 4; ab++) {
            /* A operands: entry (ab*16 + col) of the tile, words 4*ks + g */
            v4i a[2];
#pragma unroll
            for (int ks = 0; ks < 2; ks++) a[ks] = *(const v4i*)(tile + (ab * 16 + col) * kMfmaRow + 16 * (4 * ks + g));
            /* lane holds C[4g + r][col]: entries ab*16 + 4g + r, query qb*16 + col.
               The four query blocks' chains are interleaved and their results
               tested after all eight MFMAs */
            const int jb = jt + ab * 16 + 4 * g;
            v4f acc[4];
#pragma unroll
            for (int qb = 0; qb < 4; qb++) acc[qb] = mfma_fp4(a[0], bq[qb][0], c256);
#pragma unroll
            for (int qb = 0; qb < 4; qb++) acc[qb] = mfma_fp4(a[1], bq[qb][1], acc[qb]);
            /* a -dot above bnd means a distance above the second one, a key
               above k2, for all 4 entries: one test for the 4 query blocks
               (the minima are taken by every lane, outside any lane-dependent
               branch, so each MFMA runs with every lane enabled), and the
               rare lanes that pass update their top-2 */
            bool in[4];
            bool any = false;
#pragma unroll
            for (int qb = 0; qb < 4; qb++) {
                const uint32_t m = min(min(__float_as_uint(acc[qb][0]), __float_as_uint(acc[qb][1])),
                                       min(__float_as_uint(acc[qb][2]), __float_as_uint(acc[qb][3])));
                in[qb] = m <= bbits[qb];
                any |= in[qb];
            }
            if (any) {
#pragma unroll
                for (int qb = 0; qb < 4; qb++) {
                    if (in[qb]) {
                        /* (k2 >> 23) - bnd is the query's popcount */
                        const int qpop = (int)(k2[qb] >> 23) - bnd[qb];
                        const uint32_t base = ((uint32_t)qpop << 23) + (uint32_t)jb;
#pragma unroll
                        for (int r = 0; r < 4; r++)
                            if (jb + r < j1)
                                top2_push(k1[qb], k2[qb], neg_dot_key((int)acc[qb][r] - 256, base + (uint32_t)r));
                        bnd[qb] = neg_dot_bound(qpop, k2[qb]);
                        bbits[qb] = __float_as_uint((float)(256 + bnd[qb]));
                    }
                }
            }
        }
        __syncthreads();   /* tile read, next tile written */
    }
    /* merge the 4 k-groups that hold the same query */
#pragma unroll
    for (int qb = 0; qb < 4; qb++) {
#pragma unroll
        for (int m = 16; m <= 32; m <<= 1) {
            const uint32_t o1 = __shfl_xor(k1[qb], m, 64), o2 = __shfl_xor(k2[qb], m, 64);
            top2_push(k1[qb], k2[qb], o1);
            top2_push(k1[qb], k2[qb], o2);
        }
    }
}

/* dense matching: 512 queries per workgroup (8 wavefronts share each
   unpacked tile) */
constexpr int kMfmaWaves = 8;
static_assert(kMfmaQ == 64 * kMfmaWaves, "queries per MFMA workgroup");

__global__ __launch_bounds__(64 * kMfmaWaves) void k_match_mfma(const uint8_t* __restrict__ q, int nq_cap,
                                                                const int* __restrict__ d_nq,
                                                                const uint8_t* __restrict__ db, int ndb_cap,
                                                                const int* __restrict__ d_ndb, int slots,
                                                                uint2* __restrict__ partial)
{
    __shared__ __align__(16) int8_t s_db[2 * kGranule * kMfmaRow];
    __shared__ uint32_t s_code[256];
    fill_code_table(s_code, threadIdx.x, blockDim.x);   /* visible after mfma_unit's first barrier */
    const int nq = d_nq ? min(*d_nq, nq_cap) : nq_cap;
    const int ndb = d_ndb ? min(*d_ndb, ndb_cap) : ndb_cap;
    const Plan P = make_plan(nq, ndb, kMfmaQ, slots);
    const int g = (threadIdx.x & 63) >> 4;
    for (int u = blockIdx.x; u < P.qunits * P.chunks; u += slots) {
        const int qu = u % P.qunits, c = u / P.qunits;
        uint32_t k1[4], k2[4];
        int qidx[4];
        mfma_unit<kMfmaWaves>(q, nq, db, qu, c * P.chunk, min(ndb, c * P.chunk + P.chunk), s_db, s_code, k1, k2,
                              qidx);
#pragma unroll
        for (int qb = 0; qb < 4; qb++)
            if (g == 0 && qidx[qb] < nq) partial[(long long)c * nq + qidx[qb]] = make_uint2(k1[qb], k2[qb]);
    }
}

/* Batched reference-rule matching on the matrix cores: blockIdx.y is the
   pair and blockIdx.x one (query tile, database chunk) unit of it: kPairQ
   queries against kPairChunk entries.  Chunking bounds the launch by a
   chunk rather than by the largest pair's whole database.  The chunks of a
   tile meet as in the popcount matcher: one 64-bit word per query
   ((k1 << 32) | k2, merged by compare-and-swap at device scope) and one
   counter per tile; the unit that finishes the tile's last chunk reads the
   words (resetting them to "no key"), writes the matches and resets the
   counter, so the words and counters are clean for the next call.  A pair
   whose database fits one chunk writes its matches directly. */
constexpr int kPairWaves = 4, kPairQ = 64 * kPairWaves;
constexpr int kPairChunk = 16 * kGranule;

/* merge (k1, k2) into a query's merge word */
__device__ __forceinline__ void merge_word(unsigned long long* wp, uint32_t k1, uint32_t k2)
{
    unsigned long long old = __hip_atomic_load(wp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (;;) {
        const uint32_t o1 = (uint32_t)(old >> 32), o2 = (uint32_t)old;
        uint32_t n1 = o1, n2 = o2;
        top2_push(n1, n2, k1);
        top2_push(n1, n2, k2);
        if (n1 == o1 && n2 == o2) break;
        if (__hip_atomic_compare_exchange_strong(wp, &old, ((unsigned long long)n1 << 32) | n2, __ATOMIC_RELAXED,
                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            break;
    }
}

__global__ __launch_bounds__(64 * kPairWaves) void k_match_mfma_pairs(const MatchPair* __restrict__ pairs, int cap1,
                                                                      int cap2, int qcap,
                                                                      unsigned long long* __restrict__ state,
                                                                      int* __restrict__ done)
{
    __shared__ __align__(16) int8_t s_db[2 * kGranule * kMfmaRow];
    __shared__ uint32_t s_code[256];
    __shared__ int s_last;
    const MatchPair pr = pairs[blockIdx.y];
    const int nq = min(max(*pr.n1, 0), cap1), ndb = min(max(*pr.n2, 0), cap2);
    const int qtiles = (cap1 + kPairQ - 1) / kPairQ;
    const int qu = blockIdx.x % qtiles, c = blockIdx.x / qtiles;
    const int nch = ndb > 0 ? (ndb + kPairChunk - 1) / kPairChunk : 1;
    if (qu * kPairQ >= nq || c >= nch) return;
    if (ndb == 0) {
        const int qi = qu * kPairQ + threadIdx.x;
        if (qi < nq) write_match(qi, kNoKey, kNoKey, nullptr, nullptr, nullptr, nullptr, pr.pts);
        return;
    }
    fill_code_table(s_code, threadIdx.x, blockDim.x);   /* visible after mfma_unit's first barrier */
    uint32_t k1[4], k2[4];
    int qidx[4];
    mfma_unit<kPairWaves>(pr.q, nq, pr.db, qu, c * kPairChunk, min(ndb, (c + 1) * kPairChunk), s_db, s_code, k1, k2,
                          qidx);
    const int g = (threadIdx.x & 63) >> 4;
    if (nch == 1) {
#pragma unroll
        for (int qb = 0; qb < 4; qb++)
            if (g == 0 && qidx[qb] < nq)
                write_match(qidx[qb], k1[qb], k2[qb], nullptr, nullptr, nullptr, nullptr, pr.pts);
        return;
    }
    unsigned long long* words = state + (long long)blockIdx.y * qcap;
#pragma unroll
    for (int qb = 0; qb < 4; qb++)
        if (g == 0 && qidx[qb] < nq) merge_word(words + qidx[qb], k1[qb], k2[qb]);
    /* every merge of the workgroup has completed before the counter moves */
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __syncthreads();
    int* cnt = done + ((long long)blockIdx.y * qcap + qu * kPairQ) / 64;
    if (threadIdx.x == 0)
        s_last = __hip_atomic_fetch_add(cnt, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == nch - 1;
    __syncthreads();
    if (!s_last) return;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    const int qi = qu * kPairQ + threadIdx.x;
    if (qi < nq) {
        const unsigned long long v = __hip_atomic_exchange(words + qi, ~0ull, __ATOMIC_RELAXED,
                                                           __HIP_MEMORY_SCOPE_AGENT);
        write_match(qi, (uint32_t)(v >> 32), (uint32_t)v, nullptr, nullptr, nullptr, nullptr, pr.pts);
    }
    if (threadIdx.x == 0) __hip_atomic_store(cnt, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

/* ------------------------------------------------------------- merge */

/* merge the per-chunk top-2 keys of each query and write the outputs */
__global__ __launch_bounds__(256) void k_match_merge(int nq_cap, const int* __restrict__ d_nq, int ndb_cap,
                                                     const int* __restrict__ d_ndb, int qtile, int slots,
                                                     const uint2* __restrict__ partial, int* __restrict__ idx1,
                                                     int* __restrict__ d1, int* __restrict__ idx2,
                                                     int* __restrict__ d2, orb::OrbPoint* __restrict__ pts)
{
    const int nq = d_nq ? min(*d_nq, nq_cap) : nq_cap;
    const int ndb = d_ndb ? min(*d_ndb, ndb_cap) : ndb_cap;
    const Plan P = make_plan(nq, ndb, qtile, slots);
    for (int qi = blockIdx.x * blockDim.x + threadIdx.x; qi < nq; qi += gridDim.x * blockDim.x) {
        uint32_t k1 = kNoKey, k2 = kNoKey;
        int c = 0;
        for (; c + 4 <= P.chunks; c += 4) {
            uint2 v[4];
#pragma unroll
            for (int t = 0; t < 4; t++) v[t] = partial[(long long)(c + t) * nq + qi];
#pragma unroll
            for (int t = 0; t < 4; t++) {
                top2_push(k1, k2, v[t].x);
                top2_push(k1, k2, v[t].y);
            }
        }
        for (; c < P.chunks; c++) {
            const uint2 v = partial[(long long)c * nq + qi];
            top2_push(k1, k2, v.x);
            top2_push(k1, k2, v.y);
        }
        write_match(qi, k1, k2, idx1, d1, idx2, d2, pts);
    }
}

/* ---------------------------------------------------------- launchers */

constexpr int kPopBlocks = 2048;   /* workgroup slots of the popcount plan */
constexpr int kMfmaBlocks = 512;   /* workgroup slots (two 8-wave workgroups per CU): fewer, longer chunks
                                      keep more of each scan past the first entries, where most results
                                      still enter the top-2 (measured: 2.44 ms against 2.81 with 1024 on
                                      the 65536 x 262144 match) */

size_t match_partial_words(int nq, int ndb, int method, int* nchunks, int* chunk)
{
    (void)ndb;
    if (method != 1) {   /* the popcount matcher merges in place (match_state_words) */
        *nchunks = 0;
        *chunk = 0;
        return 0;
    }
    /* chunks * nq <= (slots / qunits + 1) * qunits * qtile <= (slots + qunits) * qtile */
    const int qtile = method == 1 ? kMfmaQ : kPopQ;
    const int slots = method == 1 ? kMfmaBlocks : kPopBlocks;
    const long long qunits = (nq + qtile - 1) / qtile;
    *nchunks = 0;
    *chunk = 0;
    long long entries = (slots + qunits) * (long long)qtile;
    if (entries < nq) entries = nq;
    return (size_t)entries * 2;
}

static void launch_merge(const MatchLaunch& m, int qtile, int slots, hipStream_t s, const StageEvents* ev)
{
    const int blocks = (m.nq + 255) / 256 < 1024 ? (m.nq + 255) / 256 : 1024;
    launch_stage(ev, kStMatchMerge, k_match_merge, dim3(blocks), dim3(256), s, m.nq, m.d_nq, m.ndb, m.d_ndb, qtile,
                 slots, (const uint2*)m.partial, m.idx1, m.d1, m.idx2, m.d2, m.pts);
}

hipError_t launch_match_popcount(const MatchLaunch& m, hipStream_t s, const StageEvents* ev)
{
    (void)hipGetLastError(); /* report only errors of these launches */
    if (m.nq <= 0) return hipSuccess;
    const int slots = kPopBlocks;
    /* enough workgroups for the units the capacities allow, at most kPopBlocks */
    const Plan P = make_plan(m.nq, m.ndb, kPopQ, slots);
    const long long units = (long long)P.qunits * (P.chunks > 0 ? P.chunks : 1);
    int blocks = (int)(units < kPopBlocks ? units : kPopBlocks);
    if (blocks < 1) blocks = 1;
    if (m.hamming2)
        launch_stage(ev, kStMatch, k_match_pop<true>, dim3(blocks), dim3(256), s, m.q, m.nq, m.d_nq,
                     (const uint4*)m.db, m.ndb, m.d_ndb, slots, m.state, m.done, m.idx1, m.d1, m.idx2, m.d2, m.pts);
    else
        launch_stage(ev, kStMatch, k_match_pop<false>, dim3(blocks), dim3(256), s, m.q, m.nq, m.d_nq,
                     (const uint4*)m.db, m.ndb, m.d_ndb, slots, m.state, m.done, m.idx1, m.d1, m.idx2, m.d2, m.pts);
    return hipGetLastError();
}

hipError_t launch_match_pairs(const MatchPair* d_pairs, int npairs, int cap1, int cap2, int qcap,
                              unsigned long long* state, int* done, hipStream_t s, const StageEvents* ev)
{
    (void)hipGetLastError(); /* report only errors of these launches */
    if (npairs <= 0 || cap1 <= 0) return hipSuccess;
    /* one workgroup per (query tile, database chunk) the capacities allow;
       units past a pair's counts return at once */
    const int qtiles = (cap1 + kPairQ - 1) / kPairQ;
    const int chunks = cap2 > 0 ? (cap2 + kPairChunk - 1) / kPairChunk : 1;
    launch_stage(ev, kStMatch, k_match_mfma_pairs, dim3(qtiles * chunks, npairs), dim3(kPairQ), s, d_pairs, cap1,
                 cap2, qcap, state, done);
    return hipGetLastError();
}

int match_pairs_qcap(int cap1)
{
    return (cap1 + kPairQ - 1) / kPairQ * kPairQ;
}

hipError_t launch_match_mfma(const MatchLaunch& m, hipStream_t s, const StageEvents* ev)
{
    (void)hipGetLastError(); /* report only errors of these launches */
    if (m.nq <= 0) return hipSuccess;
    const int slots = kMfmaBlocks;
    const Plan P = make_plan(m.nq, m.ndb, kMfmaQ, slots);
    const long long units = (long long)P.qunits * (P.chunks > 0 ? P.chunks : 1);
    int blocks = units < kMfmaBlocks ? (int)units : kMfmaBlocks;
    if (blocks < 1) blocks = 1;
    launch_stage(ev, kStMatch, k_match_mfma, dim3(blocks), dim3(kMfmaQ), s, m.q, m.nq, m.d_nq, m.db, m.ndb, m.d_ndb,
                 slots, (uint2*)m.partial);
    launch_merge(m, kMfmaQ, slots, s, ev);
    return hipGetLastError();
}

/* ------------------------------------------- allocation-free matching */

/* Scratch of orb_match_top2_noalloc, for callers that cannot allocate: one
   array per GPU for the life of the process, laid out as
     [merge words: 2 per query | tile counters: 1 per 64 queries | MFMA
      per-chunk top-2 keys],
   sized for kNoAllocMaxQueries queries.  In host code `scratch` names a
   host-side stand-in: kernels get the device address that
   match_scratch_setup obtains from hipGetSymbolAddress, once per device. */
constexpr size_t kScratchStateWords = 2 * (size_t)kNoAllocMaxQueries;
constexpr size_t kScratchDoneWords = (size_t)kNoAllocMaxQueries / kPopQ;
constexpr size_t kScratchPartialOff = (kScratchStateWords + kScratchDoneWords + 63) & ~(size_t)63;
/* MFMA partials: (slots + qunits) * qtile query entries of 2 words (match_partial_words) */
constexpr size_t kScratchPartialWords =
    2 * ((size_t)kMfmaBlocks + (size_t)kNoAllocMaxQueries / kMfmaQ) * (size_t)kMfmaQ;
constexpr size_t kScratchWords = kScratchPartialOff + kScratchPartialWords;
static_assert(kNoAllocMaxQueries % kMfmaQ == 0, "whole MFMA query tiles");

static __device__ __align__(256) uint32_t scratch[kScratchWords];

namespace
{
std::mutex g_scratch_mu;
uint32_t* g_scratch_dev[kMaxScratchDevices] = {};
}

size_t match_scratch_bytes()
{
    return kScratchWords * sizeof(uint32_t);
}

hipError_t match_scratch_setup(int device, uint32_t** ptr)
{
    *ptr = nullptr;
    if (device < 0 || device >= kMaxScratchDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_scratch_dev[device] = (uint32_t*)p;
    }
    *ptr = g_scratch_dev[device];
    return hipSuccess;
}

/* the merge words of nq queries back to "no key" and their tile counters to
   0 (the state k_match_pop expects; a kernel rather than a memset so the call
   can be recorded into a graph) */
__global__ __launch_bounds__(256) void k_scratch_reset(uint32_t* __restrict__ state, int nwords,
                                                       uint32_t* __restrict__ done, int ndone)
{
    for (int i = blockIdx.x * 256 + threadIdx.x; i < nwords; i += gridDim.x * 256) state[i] = kNoKey;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < ndone; i += gridDim.x * 256) done[i] = 0u;
}

hipError_t launch_match_noalloc(MatchLaunch m, int method, uint32_t* scr, hipStream_t s)
{
    (void)hipGetLastError(); /* report only errors of these launches */
    if (m.nq <= 0) return hipSuccess;
    if (m.nq > kNoAllocMaxQueries || !scr) return hipErrorInvalidValue;
    m.state = (unsigned long long*)scr;
    m.done = (int*)(scr + kScratchStateWords);
    m.partial = scr + kScratchPartialOff;
    if (method == 1) return launch_match_mfma(m, s, nullptr);
    const int nwords = 2 * m.nq, ndone = (m.nq + kPopQ - 1) / kPopQ;
    int blocks = (nwords + 255) / 256;
    if (blocks > 512) blocks = 512;
    hipLaunchKernelGGL(k_scratch_reset, dim3(blocks), dim3(256), 0, s, scr, nwords, (uint32_t*)m.done, ndone);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_match_popcount(m, s, nullptr);
}

} // namespace orbk
