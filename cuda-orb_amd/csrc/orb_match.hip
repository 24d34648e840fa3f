/*
 * Exact brute-force Hamming matchers for 256-bit descriptors (gfx950).
 *
 * The reference's gHammingMatch (orbd.cu:1471-1559) gives each query a
 * 32-thread block that re-reads the whole train set.  Here a workgroup holds
 * 256 queries (one per lane) in registers and streams a chunk of the
 * database through LDS; the database is split into chunks across workgroups
 * so even a few thousand queries fill the 256 CUs, and a merge pass combines
 * the per-chunk results.
 *
 * Each (query, entry) pair becomes a packed key (distance << 23 | index): the
 * minimum key is the nearest entry with ties to the lowest index, and the
 * running top-2 is (min(k1, key), med3(k1, key, k2)).
 */
#include "orb_internal.h"

namespace orbk
{

constexpr int kMatchQ = 256;       /* queries per workgroup */
constexpr int kMatchTile = 256;    /* database entries per LDS tile */
constexpr uint32_t kIdxBits = 23;
constexpr uint32_t kIdxMask = (1u << kIdxBits) - 1u;
constexpr uint32_t kNoKey = 0xffffffffu;

__device__ __forceinline__ void top2_push(uint32_t& k1, uint32_t& k2, uint32_t key)
{
    k2 = min(k2, max(k1, key)); /* med3(k1, key, k2) since k1 <= k2 */
    k1 = min(k1, key);
}

__global__ __launch_bounds__(256) void k_match_pop(const uint8_t* __restrict__ q, int nq_cap,
                                                   const int* __restrict__ d_nq, const uint8_t* __restrict__ db,
                                                   int ndb_cap, const int* __restrict__ d_ndb, int chunk,
                                                   uint32_t* __restrict__ partial)
{
    __shared__ uint4 s_db[kMatchTile][2];
    const int nq = d_nq ? min(*d_nq, nq_cap) : nq_cap;
    const int ndb = d_ndb ? min(*d_ndb, ndb_cap) : ndb_cap;
    const int qi = blockIdx.x * kMatchQ + threadIdx.x;
    if ((int)(blockIdx.x * kMatchQ) >= nq) return;
    const int j0 = blockIdx.y * chunk, j1 = min(ndb, j0 + chunk);
    uint4 qa = make_uint4(0, 0, 0, 0), qb = qa;
    if (qi < nq) {
        const uint4* qp = (const uint4*)(q + (long long)qi * 32);
        qa = qp[0];
        qb = qp[1];
    }
    uint32_t k1 = kNoKey, k2 = kNoKey;
    for (int jt = j0; jt < j1; jt += kMatchTile) {
        __syncthreads();
        const int j = jt + threadIdx.x;
        if (j < j1) {
            const uint4* dp = (const uint4*)(db + (long long)j * 32);
            s_db[threadIdx.x][0] = dp[0];
            s_db[threadIdx.x][1] = dp[1];
        }
        __syncthreads();
        const int nt = min(kMatchTile, j1 - jt);
        for (int t = 0; t < nt; t++) {
            const uint4 a = s_db[t][0], b = s_db[t][1];
            uint32_t d = __popc(qa.x ^ a.x);
            d = __popc(qa.y ^ a.y) + d;
            d = __popc(qa.z ^ a.z) + d;
            d = __popc(qa.w ^ a.w) + d;
            d = __popc(qb.x ^ b.x) + d;
            d = __popc(qb.y ^ b.y) + d;
            d = __popc(qb.z ^ b.z) + d;
            d = __popc(qb.w ^ b.w) + d;
            top2_push(k1, k2, (d << kIdxBits) | (uint32_t)(jt + t));
        }
    }
    if (qi < nq) {
        uint2* o = (uint2*)partial + (long long)blockIdx.y * nq_cap + qi;
        *o = make_uint2(k1, k2);
    }
}

/* merge the per-chunk top-2 keys of each query and write the outputs */
__global__ __launch_bounds__(256) void k_match_merge(int nq_cap, const int* __restrict__ d_nq, int nchunks,
                                                     int ndb_cap, const int* __restrict__ d_ndb, int chunk,
                                                     const uint32_t* __restrict__ partial, int* __restrict__ idx1,
                                                     int* __restrict__ d1, int* __restrict__ idx2,
                                                     int* __restrict__ d2, orb::OrbPoint* __restrict__ pts)
{
    const int nq = d_nq ? min(*d_nq, nq_cap) : nq_cap;
    const int ndb = d_ndb ? min(*d_ndb, ndb_cap) : ndb_cap;
    const int qi = blockIdx.x * blockDim.x + threadIdx.x;
    if (qi >= nq) return;
    uint32_t k1 = kNoKey, k2 = kNoKey;
    const int used = min(nchunks, (ndb + chunk - 1) / chunk);
    for (int c = 0; c < used; c++) {
        const uint2 v = ((const uint2*)partial)[(long long)c * nq_cap + qi];
        top2_push(k1, k2, v.x);
        top2_push(k1, k2, v.y);
    }
    const int a1 = k1 == kNoKey ? -1 : (int)(k1 & kIdxMask), b1 = k1 == kNoKey ? 257 : (int)(k1 >> kIdxBits);
    const int a2 = k2 == kNoKey ? -1 : (int)(k2 & kIdxMask), b2 = k2 == kNoKey ? 257 : (int)(k2 >> kIdxBits);
    if (idx1) idx1[qi] = a1;
    if (d1) d1[qi] = b1;
    if (idx2) idx2[qi] = a2;
    if (d2) d2[qi] = b2;
    if (pts) {
        /* orbd.cu:1544-1558: both distances start at 256; accept a unique best below 64 */
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

size_t match_partial_words(int nq, int ndb, int method, int* nchunks, int* chunk)
{
    (void)method;
    const int qblocks = (nq + kMatchQ - 1) / kMatchQ;
    const int tiles = (ndb + kMatchTile - 1) / kMatchTile;
    int nc = (2048 + qblocks - 1) / (qblocks > 0 ? qblocks : 1);
    if (nc > tiles) nc = tiles;
    if (nc < 1) nc = 1;
    int ch = ((tiles + nc - 1) / nc) * kMatchTile;
    if (ch < kMatchTile) ch = kMatchTile;
    nc = (ndb + ch - 1) / ch;
    if (nc < 1) nc = 1;
    *nchunks = nc;
    *chunk = ch;
    return (size_t)nc * (size_t)(nq > 0 ? nq : 1) * 2;
}

static void launch_merge(const MatchLaunch& m, hipStream_t s, const StageEvents* ev)
{
    ev_start(ev, kStMatchMerge, s);
    hipLaunchKernelGGL(k_match_merge, dim3((m.nq + 255) / 256), dim3(256), 0, s, m.nq, m.d_nq, m.nchunks, m.ndb,
                       m.d_ndb, m.chunk, m.partial, m.idx1, m.d1, m.idx2, m.d2, m.pts);
    ev_stop(ev, kStMatchMerge, s);
}

hipError_t launch_match_popcount(const MatchLaunch& m, hipStream_t s, const StageEvents* ev)
{
    if (m.nq <= 0) return hipSuccess;
    const dim3 grid((m.nq + kMatchQ - 1) / kMatchQ, m.nchunks);
    ev_start(ev, kStMatch, s);
    hipLaunchKernelGGL(k_match_pop, grid, dim3(256), 0, s, m.q, m.nq, m.d_nq, m.db, m.ndb, m.d_ndb, m.chunk,
                       m.partial);
    ev_stop(ev, kStMatch, s);
    launch_merge(m, s, ev);
    return hipGetLastError();
}

hipError_t launch_match_mfma(const MatchLaunch& m, hipStream_t s, const StageEvents* ev)
{
    (void)m;
    (void)s;
    (void)ev;
    return hipErrorNotSupported;
}

} // namespace orbk
