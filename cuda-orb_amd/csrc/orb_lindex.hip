/*
 * Bucketed database of the learned matcher (gfx950): a stable counting sort
 * of the database by bucket (entries of a bucket keep their index order),
 * the k-modes centroid update, per-bucket probe lists, the nearest-centroid
 * probe, and the bucketed exact search -- the INT8 MFMA contraction of
 * orb_match.hip applied to each (bucket, probing queries) block, with keys
 * carrying the sorted position (index order inside the bucket), turned into
 * database indices once per block, so ties resolve to the lowest index.
 */
#include <type_traits>

#include "orb_internal.h"
#include "orb_match_common.h"
#include "orb_mlp.h"
#include "orb_wave.h"

namespace orbk
{

constexpr int kLxRow = 128 + 16;  /* LDS row of one unpacked entry: 8 chunks of 32 FP4 codes, padded */
constexpr int kLxTile = 64;      /* entries per LDS tile */
constexpr int kLxQ = 256;        /* queries per unit */

__global__ void k_zero_int(int* p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ void k_fill_u32(uint32_t* p, int n, uint32_t v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

/* single workgroup: exclusive scan of nb counts (nb <= 4096) into offsets[0..nb]
   and of the unit counts ceil(count / unit) into units[0..nb] -- taken in the
   bucket order order[] when given, with unit_pos[u] the position of unit u;
   thread t holds positions 4t .. 4t+3 */
__global__ __launch_bounds__(1024) void k_scan_offsets(const int* __restrict__ count, int nb, int* __restrict__ offsets,
                                                       int* __restrict__ cursor, int unit, int* __restrict__ units,
                                                       const int* __restrict__ order, int* __restrict__ unit_pos)
{
    __shared__ int s_wc[16], s_wu[16];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    int c[4], u[4], sc = 0, su = 0;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int i = 4 * t + j;
        c[j] = i < nb ? count[i] : 0;
        const int cu = order ? (i < nb ? count[order[i]] : 0) : c[j];
        u[j] = unit > 0 ? (cu + unit - 1) / unit : 0;
        sc += c[j];
        su += u[j];
    }
    int xc = sc, xu = su;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int yc = __shfl_up(xc, o, 64), yu = __shfl_up(xu, o, 64);
        if (lane >= o) {
            xc += yc;
            xu += yu;
        }
    }
    if (lane == 63) {
        s_wc[w] = xc;
        s_wu[w] = xu;
    }
    __syncthreads();
    int ac = xc - sc, au = xu - su;
    for (int i = 0; i < w; i++) {
        ac += s_wc[i];
        au += s_wu[i];
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int i = 4 * t + j;
        if (i < nb) {
            offsets[i] = ac;
            if (cursor) cursor[i] = ac;
            if (units) units[i] = au;
            if (unit_pos)
                for (int v = 0; v < u[j]; v++) unit_pos[au + v] = i; /* unit -> its position */
        }
        ac += c[j];
        au += u[j];
    }
    if (t == 1023) {
        offsets[nb] = ac;
        if (units) units[nb] = au;
    }
}

/* Stable counting sort of the database by bucket (entries of a bucket keep
   their index order, so a position in the sorted copy orders like the index):
   per chunk of 256 entries the bucket histogram, stored bucket-major
   H[b * nchunks + c]; per bucket its total, then its chunk counts rebased to
   exclusive positions; each entry goes to its chunk's position for its bucket
   plus the number of earlier entries of the chunk in the same bucket. */
constexpr int kSortChunk = 256;

__global__ __launch_bounds__(256) void k_chunk_hist(const int* __restrict__ assign, int n, int nb, int nchunks,
                                                    int* __restrict__ H)
{
    __shared__ int h[1024];
    for (int b = threadIdx.x; b < nb; b += 256) h[b] = 0;
    __syncthreads();
    const int i = blockIdx.x * kSortChunk + threadIdx.x;
    if (i < n) atomicAdd(&h[assign[i]], 1);
    __syncthreads();
    for (int b = threadIdx.x; b < nb; b += 256) H[(long long)b * nchunks + blockIdx.x] = h[b];
}

/* one workgroup per bucket: rebase = 0 writes the row total to total[b];
   rebase = 1 turns the row into exclusive positions starting at offsets[b] */
__global__ __launch_bounds__(256) void k_chunk_rows(int* __restrict__ H, int nchunks, int rebase,
                                                    int* __restrict__ total, const int* __restrict__ offsets)
{
    __shared__ int s_w[4];
    int* row = H + (long long)blockIdx.x * nchunks;
    const int per = (nchunks + 255) / 256;
    const int c0 = min(nchunks, (int)threadIdx.x * per), c1 = min(nchunks, c0 + per);
    int sum = 0;
    for (int c = c0; c < c1; c++) sum += row[c];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int x = sum;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) s_w[w] = x;
    __syncthreads();
    if (!rebase) {
        if (threadIdx.x == 0) total[blockIdx.x] = s_w[0] + s_w[1] + s_w[2] + s_w[3];
        return;
    }
    int run = offsets[blockIdx.x] + x - sum;
    for (int i = 0; i < w; i++) run += s_w[i];
    for (int c = c0; c < c1; c++) {
        const int t = row[c];
        row[c] = run;
        run += t;
    }
}

__global__ __launch_bounds__(256) void k_chunk_scatter(const int* __restrict__ assign, int n, int nchunks,
                                                       const int* __restrict__ H, const uint8_t* __restrict__ db,
                                                       int* __restrict__ perm, uint8_t* __restrict__ db_sorted)
{
    __shared__ int sb[kSortChunk];
    const int i = blockIdx.x * kSortChunk + threadIdx.x;
    const int b = i < n ? assign[i] : -1;
    sb[threadIdx.x] = b;
    __syncthreads();
    if (i >= n) return;
    int r = 0;
    for (int t = 0; t < (int)threadIdx.x; t++) r += sb[t] == b;
    const int pos = H[(long long)b * nchunks + blockIdx.x] + r;
    perm[pos] = i;
    const uint4* src = (const uint4*)(db + (long long)i * 32);
    uint4* dst = (uint4*)(db_sorted + (long long)pos * 32);
    dst[0] = src[0];
    dst[1] = src[1];
}

/* one workgroup per bucket, one thread per bit: the majority bit of the
   bucket's entries (a tie or an empty bucket keeps the previous bit) */
__global__ __launch_bounds__(256) void k_kmodes_update(const uint8_t* __restrict__ db_sorted,
                                                       const int* __restrict__ offsets, uint8_t* __restrict__ centroids)
{
    const int b = blockIdx.x, k = threadIdx.x;
    const int o0 = offsets[b], o1 = offsets[b + 1];
    int cnt = 0;
    for (int j = o0; j < o1; j++) cnt += (db_sorted[(long long)j * 32 + (k >> 3)] >> (k & 7)) & 1;
    const int size = o1 - o0;
    const int old = (centroids[(long long)b * 32 + (k >> 3)] >> (k & 7)) & 1;
    const bool bit = 2 * cnt > size || (2 * cnt == size && old);
    const unsigned long long m = __ballot(bit);
    if ((k & 63) == 0) ((unsigned long long*)(centroids + (long long)b * 32))[k >> 6] = m;
}

constexpr int kPlChunk = 8192; /* probe-list entries per workgroup */

/* bucket histogram of a chunk of probe entries in LDS, then one global add per
   bucket (nb <= 1024) */
__global__ __launch_bounds__(256) void k_probe_count(const int* __restrict__ probes, int n, int nb,
                                                     int* __restrict__ count)
{
    __shared__ int h[1024];
    for (int b = threadIdx.x; b < nb; b += 256) h[b] = 0;
    __syncthreads();
    /* all of a thread's entries loaded first, then counted */
    const int i0 = blockIdx.x * kPlChunk + threadIdx.x, i1 = min(n, blockIdx.x * kPlChunk + kPlChunk);
    int v[kPlChunk / 256];
#pragma unroll
    for (int t = 0; t < kPlChunk / 256; t++) v[t] = i0 + 256 * t < i1 ? probes[i0 + 256 * t] : -1;
#pragma unroll
    for (int t = 0; t < kPlChunk / 256; t++)
        if (v[t] >= 0) atomicAdd(&h[v[t]], 1);
    __syncthreads();
    for (int b = threadIdx.x; b < nb; b += 256)
        if (h[b]) atomicAdd(&count[b], h[b]);
}

/* the same chunk again: one global range reserved per bucket, entries placed
   inside it through LDS cursors (i = query * nprobe + slot) */
__global__ __launch_bounds__(256) void k_probe_fill(const int* __restrict__ probes, int n, int nb,
                                                    int* __restrict__ cursor, int* __restrict__ lists)
{
    __shared__ int h[1024];
    for (int b = threadIdx.x; b < nb; b += 256) h[b] = 0;
    __syncthreads();
    const int i0 = blockIdx.x * kPlChunk + threadIdx.x, i1 = min(n, blockIdx.x * kPlChunk + kPlChunk);
    int v[kPlChunk / 256];
#pragma unroll
    for (int t = 0; t < kPlChunk / 256; t++) v[t] = i0 + 256 * t < i1 ? probes[i0 + 256 * t] : -1;
#pragma unroll
    for (int t = 0; t < kPlChunk / 256; t++)
        if (v[t] >= 0) atomicAdd(&h[v[t]], 1);
    __syncthreads();
    for (int b = threadIdx.x; b < nb; b += 256)
        if (h[b]) h[b] = atomicAdd(&cursor[b], h[b]);
    __syncthreads();
#pragma unroll
    for (int t = 0; t < kPlChunk / 256; t++)
        if (v[t] >= 0) lists[atomicAdd(&h[v[t]], 1)] = i0 + 256 * t;
}

/* Each unit: up to 256 list entries (query, slot) of one bucket x all of the
   bucket's entries.  A wave holds 64 queries as negated +/-1 FP4 B operands;
   each 64-entry tile of the bucket is unpacked into LDS once per workgroup as
   +/-1 FP4 codes (a 256-entry byte table in LDS), so a 16 x 16 block of
   -dot values is two 128-deep FP4 MFMAs, exact in FP32 (|-dot| <= 256).  Units are
   numbered along order[] (buckets by decreasing size) and handed out through
   the counter *next, so the largest come first and the workgroups finish
   together. */
__global__ __launch_bounds__(256, 4) void k_bucket_search(const uint8_t* __restrict__ q, int nprobe,
                                                          const int* __restrict__ lists,
                                                          const int* __restrict__ list_offsets,
                                                          const int* __restrict__ unit_offsets,
                                                          const int* __restrict__ unit_pos,
                                                          const int* __restrict__ order, int nb,
                                                          const uint8_t* __restrict__ db_sorted,
                                                          const int* __restrict__ perm,
                                                          const int* __restrict__ offsets, int* __restrict__ next,
                                                          uint32_t* __restrict__ qbound, uint2* __restrict__ partial)
{
    __shared__ __align__(16) int8_t s_db[kLxTile * kLxRow];
    __shared__ uint32_t s_code[256];
    __shared__ int s_u;
    const int total_units = unit_offsets[nb];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int g = lane >> 4, col = lane & 15;
    s_code[tid] = fp4_pm1_8((uint32_t)tid); /* +1 where a bit is set; read after the loop's first barrier */
    /* the accumulators start at 256, so they hold 256 - dot >= 0 */
    const v4f c256 = {256.f, 256.f, 256.f, 256.f};
    for (;;) {
        __syncthreads();
        if (tid == 0) s_u = atomicAdd(next, 1);
        __syncthreads();
        const int u = s_u;
        if (u >= total_units) break;
        const int lo = unit_pos[u]; /* position of unit u in the bucket order */
        const int b = order[lo];
        const int e0 = list_offsets[b] + (u - unit_offsets[lo]) * kLxQ, e1 = min(list_offsets[b + 1], e0 + kLxQ);
        v4i bq[4][2];
        /* A query's second distance found so far in any of its buckets
           (qbound) bounds its final second distance: entries farther than it
           cannot enter the final top-2 and are kept out by starting the
           top-2 at the phantom key (bound << 23 | kIdxMask), above every
           real key of that distance; phantoms are dropped at the end. */
        uint32_t k1[4], k2[4];
        /* acc = 256 - dot = 2 d for distance d, so a block's results can enter
           the top-2 only if its smallest acc is at most 2 (k2 >> 23): bb holds
           the float bits of that bound (non-negative floats order like their
           bits) */
        uint32_t bb[4];
#pragma unroll
        for (int qb = 0; qb < 4; qb++) {
            const int e = e0 + wid * 64 + qb * 16 + col;
            const int ent = e < e1 ? lists[e] : -1;
            uint4 d0 = make_uint4(0, 0, 0, 0), d1 = d0;
            uint32_t bd = 0xffffffffu;
            if (ent >= 0) {
                const int qi = ent / nprobe;
                const uint4* qp = (const uint4*)(q + (long long)qi * 32);
                d0 = qp[0];
                d1 = qp[1];
                bd = __hip_atomic_load(qbound + qi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            const uint32_t w[8] = {d0.x, d0.y, d0.z, d0.w, d1.x, d1.y, d1.z, d1.w};
#pragma unroll
            for (int ks = 0; ks < 2; ks++) bq[qb][ks] = fp4_word(w[4 * ks + g], fp4_neg_pm1_8);
            k1[qb] = k2[qb] = bd <= 256u ? (bd << kIdxBits) | kIdxMask : kNoKey;
            bb[qb] = __float_as_uint((float)(2u * (k2[qb] >> kIdxBits)));
        }
        const int j0 = offsets[b], j1 = offsets[b + 1];
        const bool wave_busy = e0 + wid * 64 < e1;
        /* thread -> entry e of a tile, bytes 8 qq .. 8 qq + 7; the next tile's
           bytes are loaded while this tile is searched */
        const int e = tid >> 2, qq = tid & 3;
        uint2 nv = make_uint2(0u, 0u);
        if (j0 + e < j1) nv = *(const uint2*)(db_sorted + (long long)(j0 + e) * 32 + qq * 8);
        /* keys carry the sorted position (index order inside the bucket); PART:
           the bucket's last tile, whose rows past j1 are skipped */
        auto search_tile = [&](int jt, auto part) {
            constexpr bool PART = decltype(part)::value;
#pragma unroll
            for (int ab = 0; ab < 4; ab++) {
                v4i a[2];
#pragma unroll
                for (int ks = 0; ks < 2; ks++) a[ks] = *(const v4i*)(s_db + (ab * 16 + col) * kLxRow + 16 * (4 * ks + g));
                const int jb = jt + ab * 16 + 4 * g;
#pragma unroll
                for (int qb = 0; qb < 4; qb++) {
                    v4f acc = mfma_fp4(a[0], bq[qb][0], c256);
                    acc = mfma_fp4(a[1], bq[qb][1], acc);
                    /* acc = 2 d: the key d << 23 | position is acc << 22 + position */
                    const uint32_t mb = min(min(__float_as_uint(acc[0]), __float_as_uint(acc[1])),
                                            min(__float_as_uint(acc[2]), __float_as_uint(acc[3])));
                    if (mb <= bb[qb]) {
#pragma unroll
                        for (int r = 0; r < 4; r++)
                            if (!PART || jb + r < j1)
                                top2_push(k1[qb], k2[qb], ((uint32_t)(int)acc[r] << (kIdxBits - 1)) + (uint32_t)(jb + r));
                        bb[qb] = __float_as_uint((float)(2u * (k2[qb] >> kIdxBits)));
                    }
                }
            }
        };
        for (int jt = j0; jt < j1; jt += kLxTile) {
            __syncthreads();
            {
                /* descriptor words 2 qq and 2 qq + 1 of entry e */
                *(v4i*)(s_db + e * kLxRow + 16 * (2 * qq)) = fp4_word_lut(nv.x, s_code);
                *(v4i*)(s_db + e * kLxRow + 16 * (2 * qq + 1)) = fp4_word_lut(nv.y, s_code);
                const int jn = jt + kLxTile + e;
                nv = make_uint2(0u, 0u);
                if (jn < j1) nv = *(const uint2*)(db_sorted + (long long)jn * 32 + qq * 8);
            }
            __syncthreads();
            /* a wavefront without queries (the tail of a bucket's last unit)
               only helps to stage the tiles */
            if (!wave_busy) continue;
            if (jt + kLxTile <= j1) search_tile(jt, std::false_type{});
            else search_tile(jt, std::true_type{});
        }
#pragma unroll
        for (int qb = 0; qb < 4; qb++) {
#pragma unroll
            for (int m = 16; m <= 32; m <<= 1) {
                const uint32_t o1 = __shfl_xor(k1[qb], m, 64), o2 = __shfl_xor(k2[qb], m, 64);
                top2_push(k1[qb], k2[qb], o1);
                top2_push(k1[qb], k2[qb], o2);
            }
            const int el = e0 + wid * 64 + qb * 16 + col;
            if (g == 0 && el < e1) {
                /* phantoms dropped, sorted positions back to database indices
                   (a real position is below kIdxMask: fewer than 2^23 - 1 entries) */
                const bool r1 = (k1[qb] & kIdxMask) != kIdxMask, r2 = (k2[qb] & kIdxMask) != kIdxMask;
                const uint32_t a1 = r1 ? (k1[qb] & ~kIdxMask) | (uint32_t)perm[k1[qb] & kIdxMask] : kNoKey;
                const uint32_t a2 = r2 ? (k2[qb] & ~kIdxMask) | (uint32_t)perm[k2[qb] & kIdxMask] : kNoKey;
                const int ent = lists[el];
                partial[ent] = make_uint2(a1, a2);
                if (r2) atomicMin(qbound + ent / nprobe, k2[qb] >> kIdxBits);
            }
        }
    }
}

/* low-confidence queries (max softmax < thresh) gathered for the exact search */
__global__ void k_fallback_compact(const uint8_t* __restrict__ q, int nq, const float* __restrict__ pmax, float thresh,
                                   int* __restrict__ fb_count, int* __restrict__ fb_index, int* __restrict__ fb_qidx,
                                   uint8_t* __restrict__ fb_q)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nq) return;
    if (!(pmax[i] < thresh)) {
        fb_index[i] = -1;
        return;
    }
    const int pos = atomicAdd(fb_count, 1);
    fb_index[i] = pos;
    fb_qidx[pos] = i;
    const uint4* s = (const uint4*)(q + (long long)i * 32);
    uint4* d = (uint4*)(fb_q + (long long)pos * 32);
    d[0] = s[0];
    d[1] = s[1];
}

/* top-2 of each query over its probed slots: 8 lanes per query, lane j takes
   slots j, j + 8, ... (contiguous 64-byte reads across the 8 lanes), then the
   lanes' top-2 meet by shuffles (the keys are distinct, so the order of the
   pushes does not matter) */
__global__ __launch_bounds__(256) void k_probe_merge(const uint2* __restrict__ partial, const int* __restrict__ probes,
                                                     int nq, int nprobe, const int* __restrict__ fb_index,
                                                     const int* __restrict__ ex_idx1, const int* __restrict__ ex_d1,
                                                     const int* __restrict__ ex_idx2, const int* __restrict__ ex_d2,
                                                     int* __restrict__ idx1, int* __restrict__ d1,
                                                     int* __restrict__ idx2, int* __restrict__ d2,
                                                     int* __restrict__ reset_count, int ncount,
                                                     uint32_t* __restrict__ reset_qbound)
{
    const int i = (blockIdx.x * blockDim.x + threadIdx.x) >> 3, j = threadIdx.x & 7;
    const bool live = i < nq;
    if (reset_count) {
        /* the probe-list counts and the search's bounds of this batch are no
           longer read: left clean for the next batch */
        if (blockIdx.x == 0)
            for (int t = threadIdx.x; t < ncount; t += blockDim.x) reset_count[t] = 0;
        if (live && j == 0) reset_qbound[i] = 0xffffffffu;
    }
    uint32_t k1 = kNoKey, k2 = kNoKey;
    if (live) {
        for (int s = j; s < nprobe; s += 8) {
            /* slots left out by the probability-mass rule (-1) were not searched */
            if (probes[(long long)i * nprobe + s] < 0) continue;
            const uint2 v = partial[(long long)i * nprobe + s];
            top2_push(k1, k2, v.x);
            top2_push(k1, k2, v.y);
        }
    }
#pragma unroll
    for (int o = 1; o < 8; o <<= 1) {
        const uint32_t o1 = (uint32_t)__shfl_xor((int)k1, o, 64), o2 = (uint32_t)__shfl_xor((int)k2, o, 64);
        top2_push(k1, k2, o1);
        top2_push(k1, k2, o2);
    }
    if (!live || j) return;
    const int f = fb_index ? fb_index[i] : -1;
    if (f >= 0) {
        if (idx1) idx1[i] = ex_idx1[f];
        if (d1) d1[i] = ex_d1[f];
        if (idx2) idx2[i] = ex_idx2[f];
        if (d2) d2[i] = ex_d2[f];
        return;
    }
    if (idx1) idx1[i] = k1 == kNoKey ? -1 : (int)(k1 & kIdxMask);
    if (d1) d1[i] = k1 == kNoKey ? 257 : (int)(k1 >> kIdxBits);
    if (idx2) idx2[i] = k2 == kNoKey ? -1 : (int)(k2 & kIdxMask);
    if (d2) d2[i] = k2 == kNoKey ? 257 : (int)(k2 >> kIdxBits);
}

/* gather rows idx[i] (or the first n) of q and, if given, their labels */
__global__ void k_gather(const uint8_t* __restrict__ q, const int* __restrict__ idx, const int* __restrict__ lab_src,
                         int n, uint8_t* __restrict__ qd, int* __restrict__ lab)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int r = idx ? idx[i] : i;
    const uint4* s4 = (const uint4*)(q + (long long)r * 32);
    uint4* d4 = (uint4*)(qd + (long long)i * 32);
    d4[0] = s4[0];
    d4[1] = s4[1];
    if (lab && lab_src) lab[i] = lab_src[r];
}

/* the nprobe nearest centroids of each query in Hamming distance (ties to the
   lower bucket id; by ascending distance with SORTED, else by bucket id) --
   the probe without the network.  With mass below 1 (k <= 64) only the
   nearest prefix whose mass reaches `mass` under the weights exp(-d / temp)
   is probed, by the rule of the network's probe (weights relative to the
   nearest centroid, wave_mass_keep), best first and -1 after.  The
   centroids are staged
   in LDS once per workgroup and each wavefront takes kCenQ queries in turn.
   CAP as for the network's probe: with CAP below 1024 a query whose
   candidates do not fit the LDS buffer is finished from its slots. */
constexpr int kCenQ = 8;
constexpr int kCenCap = 256;

template <bool SORTED, int CAP>
__global__ __launch_bounds__(256) void k_centroid_probe(const uint8_t* __restrict__ q, int nq,
                                                        const uint8_t* __restrict__ cen, int nb, int k, float mass,
                                                        float inv_temp, int* __restrict__ probes,
                                                        float* __restrict__ pmax)
{
    __shared__ uint4 s_c[1024 * 2];
    __shared__ uint2 s_cand[4][CAP];
    __shared__ int s_ids[4][64];
    for (int i = threadIdx.x; i < nb * 2; i += 256) s_c[i] = ((const uint4*)cen)[i];
    __syncthreads();
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int t = 0; t < kCenQ; t++) {
        const int r = (blockIdx.x * 4 + wv) * kCenQ + t;
        if (r >= nq) return;
        const uint4 qa = ((const uint4*)(q + (long long)r * 32))[0], qb = ((const uint4*)(q + (long long)r * 32))[1];
        /* key 0x10000 - distance: nearer centroids rank higher; 0 marks no centroid */
        uint32_t key[16];
#pragma unroll
        for (int u = 0; u < 16; u++) {
            const int c = lane + 64 * u;
            if (64 * u < nb) { /* nb is a multiple of 64 */
                const uint4 a = s_c[2 * c], b = s_c[2 * c + 1];
                const int d = __popc(qa.x ^ a.x) + __popc(qa.y ^ a.y) + __popc(qa.z ^ a.z) + __popc(qa.w ^ a.w) +
                              __popc(qb.x ^ b.x) + __popc(qb.y ^ b.y) + __popc(qb.z ^ b.z) + __popc(qb.w ^ b.w);
                key[u] = 0x10000u - (uint32_t)d;
            } else {
                key[u] = 0u;
            }
        }
        int* const out = probes + (long long)r * k;
        if (!(mass < 1.f)) {
            const bool ok = wave_topk<16, SORTED, CAP>(key, k, s_cand[wv], out);
            if (!ok && lane == 0) out[0] = -1;
        } else {
            uint32_t km = 0;
#pragma unroll
            for (int u = 0; u < 16; u++) km = max(km, key[u]);
#pragma unroll
            for (int o = 32; o >= 1; o >>= 1) km = max(km, (uint32_t)__shfl_xor((int)km, o, 64));
            float sw = 0.f;
#pragma unroll
            for (int u = 0; u < 16; u++) sw += 64 * u < nb ? __expf((float)((int)key[u] - (int)km) * inv_temp) : 0.f;
#pragma unroll
            for (int o = 32; o >= 1; o >>= 1) sw += __shfl_xor(sw, o, 64);
            if (!wave_topk<16, true, CAP>(key, k, s_cand[wv], s_ids[wv])) {
                if (lane == 0) out[0] = -1;
            } else {
                /* lane = rank */
                wave_sync_lds();
                const int c = lane < k ? s_ids[wv][lane] : 0;
                const uint4 a = s_c[2 * c], b = s_c[2 * c + 1];
                const int d = __popc(qa.x ^ a.x) + __popc(qa.y ^ a.y) + __popc(qa.z ^ a.z) + __popc(qa.w ^ a.w) +
                              __popc(qb.x ^ b.x) + __popc(qb.y ^ b.y) + __popc(qb.z ^ b.z) + __popc(qb.w ^ b.w);
                const uint32_t kc = 0x10000u - (uint32_t)d;
                wave_mass_keep(c, __expf((float)((int)kc - (int)km) * inv_temp), k, mass * sw, out);
            }
        }
        if (lane == 0) pmax[r] = 1.f;
        wave_sync_lds(); /* the next query reuses s_cand and s_ids */
    }
}

/* ---------------------------------------------------------------- launchers */

hipError_t launch_gather(const uint8_t* q, const int* idx, const int* lab_src, int n, uint8_t* qd, int* lab,
                         hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL(k_gather, dim3((n + 255) / 256), dim3(256), 0, s, q, idx, lab_src, n, qd, lab);
    return hipGetLastError();
}

hipError_t launch_centroid_probe(const uint8_t* q, int nq, const uint8_t* cen, int nb, int k, float mass, float temp,
                                 int* probes, float* pmax, bool sorted, hipStream_t s)
{
    if (nb <= 0 || nb > 1024 || nb % 64 || k < 1 || k > nb) return hipErrorInvalidValue;
    if (mass < 1.f && (k > 64 || !(mass > 0.f) || !(temp > 0.f))) return hipErrorInvalidValue;
    const float it = mass < 1.f ? 1.f / temp : 0.f;
    const dim3 grid((nq + 4 * kCenQ - 1) / (4 * kCenQ));
    if (nq <= 0) return hipGetLastError();
    const bool split = k <= 64 && nb > kCenCap;
#define CENTROID_PROBE(S, CAPV)                                                                          \
    hipLaunchKernelGGL((k_centroid_probe<S, CAPV>), grid, dim3(256), 0, s, q, nq, cen, nb, k, mass, it, probes, \
                       pmax)
    if (sorted && split) CENTROID_PROBE(true, kCenCap);
    else if (sorted) CENTROID_PROBE(true, 1024);
    else if (split) CENTROID_PROBE(false, kCenCap);
    else CENTROID_PROBE(false, 1024);
#undef CENTROID_PROBE
    return hipGetLastError();
}


static inline dim3 grid1(long long n) { return dim3((unsigned)((n + 255) / 256)); }

hipError_t launch_bucket_sort(const int* assign, int n, int nb, const uint8_t* db, int* H, int* hist, int* offsets,
                              int* perm, uint8_t* db_sorted, hipStream_t s)
{
    if (nb > 1024) return hipErrorInvalidValue;
    if (n <= 0) return hipSuccess;
    const int nchunks = (n + kSortChunk - 1) / kSortChunk;
    hipLaunchKernelGGL(k_chunk_hist, dim3(nchunks), dim3(256), 0, s, assign, n, nb, nchunks, H);
    hipLaunchKernelGGL(k_chunk_rows, dim3(nb), dim3(256), 0, s, H, nchunks, 0, hist, (const int*)nullptr);
    hipLaunchKernelGGL(k_scan_offsets, dim3(1), dim3(1024), 0, s, (const int*)hist, nb, offsets, (int*)nullptr, 0,
                       (int*)nullptr, (const int*)nullptr, (int*)nullptr);
    hipLaunchKernelGGL(k_chunk_rows, dim3(nb), dim3(256), 0, s, H, nchunks, 1, (int*)nullptr, (const int*)offsets);
    hipLaunchKernelGGL(k_chunk_scatter, dim3(nchunks), dim3(256), 0, s, assign, n, nchunks, (const int*)H, db, perm,
                       db_sorted);
    return hipGetLastError();
}

hipError_t launch_kmodes_update(const uint8_t* db_sorted, const int* offsets, int nb, uint8_t* centroids,
                                hipStream_t s)
{
    hipLaunchKernelGGL(k_kmodes_update, dim3(nb), dim3(256), 0, s, db_sorted, offsets, centroids);
    return hipGetLastError();
}

hipError_t launch_probe_lists(const int* probes, int nq, int nprobe, int nb, const int* order, int* list_count,
                              int* list_offsets, int* list_cursor, int* lists, int* unit_offsets, int* unit_pos,
                              bool zeroed, hipStream_t s)
{
    const long long n = (long long)nq * nprobe;
    if (!zeroed) hipLaunchKernelGGL(k_zero_int, grid1(nb + 1), dim3(256), 0, s, list_count, nb + 1);
    if (nb > 1024) return hipErrorInvalidValue;
    const dim3 chunks((unsigned)((n + kPlChunk - 1) / kPlChunk));
    if (n > 0) hipLaunchKernelGGL(k_probe_count, chunks, dim3(256), 0, s, probes, (int)n, nb, list_count);
    hipLaunchKernelGGL(k_scan_offsets, dim3(1), dim3(1024), 0, s, list_count, nb, list_offsets, list_cursor, kLxQ,
                       unit_offsets, order, unit_pos);
    if (n > 0) hipLaunchKernelGGL(k_probe_fill, chunks, dim3(256), 0, s, probes, (int)n, nb, list_cursor, lists);
    return hipGetLastError();
}

hipError_t launch_bucket_search(const uint8_t* q, int nprobe, const int* lists, const int* list_offsets,
                                const int* unit_offsets, const int* unit_pos, const int* order, int nb,
                                const uint8_t* db_sorted, const int* perm, const int* offsets, int* next,
                                uint32_t* qbound, int nq, uint2* partial, int ncu, bool filled, hipStream_t s)
{
    /* no bound yet for any query */
    if (!filled) hipLaunchKernelGGL(k_fill_u32, grid1(nq), dim3(256), 0, s, qbound, nq, 0xffffffffu);
    /* as many workgroups as stay resident: they take units until none are left */
    int per = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_bucket_search, 256, 0) != hipSuccess || per < 1) per = 1;
    const int blocks = per * (ncu > 0 ? ncu : 1);
    hipLaunchKernelGGL(k_bucket_search, dim3(blocks), dim3(256), 0, s, q, nprobe, lists, list_offsets, unit_offsets,
                       unit_pos, order, nb, db_sorted, perm, offsets, next, qbound, partial);
    return hipGetLastError();
}

hipError_t launch_fallback_compact(const uint8_t* q, int nq, const float* pmax, float thresh, int* fb_count,
                                   int* fb_index, int* fb_queries_idx, uint8_t* fb_q, hipStream_t s)
{
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(256), 0, s, fb_count, 1);
    if (nq > 0)
        hipLaunchKernelGGL(k_fallback_compact, grid1(nq), dim3(256), 0, s, q, nq, pmax, thresh, fb_count, fb_index,
                           fb_queries_idx, fb_q);
    return hipGetLastError();
}

hipError_t launch_probe_merge(const uint2* partial, const int* probes, int nq, int nprobe, const int* fb_index,
                              const int* ex_idx1, const int* ex_d1, const int* ex_idx2, const int* ex_d2, int* idx1,
                              int* d1, int* idx2, int* d2, int* reset_count, int ncount, uint32_t* reset_qbound,
                              hipStream_t s)
{
    if (nq > 0)
        hipLaunchKernelGGL(k_probe_merge, dim3((unsigned)((nq + 31) / 32)), dim3(256), 0, s, partial, probes, nq,
                           nprobe, fb_index, ex_idx1, ex_d1, ex_idx2, ex_d2, idx1, d1, idx2, d2, reset_count, ncount,
                           reset_qbound);
    return hipGetLastError();
}

} // namespace orbk
