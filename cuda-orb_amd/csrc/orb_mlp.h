/*
 * Internal interface of the learned bucket matcher kernels (orb_mlp.hip,
 * orb_lindex.hip) used by the host code in orb_lindex.cpp.
 */
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace orbk
{

/* C(m, n) = sum_k A(m, k) B(k, n) with A(m, k) = A[m*sam + k*sak],
   B(k, n) = B[k*sbk + n*sbn] (bf16 bits); M, N, K multiples of 64 (an error
   otherwise).
   Epilogue, in order: z += addin[m*ldc + n] if addin; z += cscale * (256 -
   2 Hamming(qbits row m, cbits row n)) if cbits (32-byte rows: the centroid
   term, exact in FP32); z += bias[n]; z *= gelu'(zmul[m*ldc + n]) if zmul;
   out32[m*ldc + n] = z; out16[m*ldc + n] = bf16(gelu ? gelu(z) : z).  With
   half (products of operands contiguous along K, no addin, zmul or gelu) z
   is first rounded to FP16, and out16 takes the FP16 bits instead. */
struct GemmArgs {
    const uint16_t* A;
    long long sam, sak;
    const uint16_t* B;
    long long sbk, sbn;
    int M, N, K;
    const float* addin;
    const float* bias;
    int gelu;
    const float* zmul;
    float* out32;
    uint16_t* out16;
    long long ldc;
    /* optional centroid term (forward products only): M rows of query bits,
       N rows of centroid bits, 32 bytes each */
    const uint8_t* qbits;
    const uint8_t* cbits;
    float cscale;
    /* products of operands contiguous along M and N (A^T B: the weight
       gradients) with only out32: the K range split into up to `splits`
       parts, ws holding splits x M x N floats of partial sums */
    float* ws;
    int splits;
    int half;
};
/* split count and workspace floats of such a product of M x N outputs over K */
int gemm_tn_splits(int M, int N, int K);

hipError_t launch_gemm(const GemmArgs& g, hipStream_t s);
/* dst (C x R) = src (R x C) transposed (bf16 bits); R and C multiples of 64 */
hipError_t launch_transpose16(const uint16_t* src, int R, int C, uint16_t* dst, hipStream_t s);
hipError_t launch_bits_pm1(const uint8_t* desc, int n, uint16_t* X, hipStream_t s);
hipError_t launch_softmax_ce(const float* Z, const int* labels, int n, int C, uint16_t* dZ, float* loss,
                             hipStream_t s);
/* acc[0] += the sum of loss[0 .. n) (fixed order) */
hipError_t launch_loss_sum(const float* loss, int n, float* acc, hipStream_t s);
hipError_t launch_colsum(const uint16_t* X, int n, int C, float* out, hipStream_t s);
hipError_t launch_adam(float* w, uint16_t* w16, const float* g, float* m, float* v, long long n, float lr, float b1,
                       float b2, float eps, int step, hipStream_t s);
hipError_t launch_to_bf16(const float* x, uint16_t* y, long long n, hipStream_t s);
/* top-k buckets per row of FP16 logits (ties to the lower id): best first when
   sorted, else in bucket-id order (the same set); with mass below 1 (k <= 64)
   the best-ranked prefix whose softmax mass reaches it, best first, -1 after */
hipError_t launch_probe(const uint16_t* Z, int n, int C, int k, float mass, int* probes, float* pmax, bool sorted,
                        hipStream_t s);
/* bucketed database (orb_lindex.hip) */
/* stable counting sort of n entries by bucket (assign): offsets[0..nb], perm
   (sorted position -> index, index order inside a bucket), db_sorted; H holds
   nb * ceil(n / 256) ints of scratch, hist nb ints */
hipError_t launch_bucket_sort(const int* assign, int n, int nb, const uint8_t* db, int* H, int* hist, int* offsets,
                              int* perm, uint8_t* db_sorted, hipStream_t s);
hipError_t launch_kmodes_update(const uint8_t* db_sorted, const int* offsets, int nb, uint8_t* centroids,
                                hipStream_t s);

/* probe lists: queries x nprobe (bucket ids) -> per-bucket lists of q*nprobe+slot;
   list_count has nb + 1 ints (the last is the search's unit counter), zeroed
   first unless `zeroed` (already zero: the previous query's merge reset it);
   unit_offsets counts search units along order[] (buckets by decreasing size)
   and unit_pos[u] (nq * nprobe / 256 + nb ints) is the order position of unit u */
hipError_t launch_probe_lists(const int* probes, int nq, int nprobe, int nb, const int* order, int* list_count,
                              int* list_offsets, int* list_cursor, int* lists, int* unit_offsets, int* unit_pos,
                              bool zeroed, hipStream_t s);
/* bucketed exact search: for every (query, probed bucket) the top-2 keys
   (distance << 23 | original index) over that bucket's entries, leaving out
   entries farther than the query's second distance in its other buckets
   (qbound: nq words, set to "no bound" first unless `filled`); units taken
   from the zeroed counter *next */
hipError_t launch_bucket_search(const uint8_t* q, int nprobe, const int* lists, const int* list_offsets,
                                const int* unit_offsets, const int* unit_pos, const int* order, int nb,
                                const uint8_t* db_sorted, const int* perm, const int* offsets, int* next,
                                uint32_t* qbound, int nq, uint2* partial, int ncu, bool filled, hipStream_t s);
/* per query: merge of the probed buckets, or the exact result when the query
   was sent to the fallback (fb_index >= 0); with reset_count, it also zeroes
   reset_count[0 .. ncount) and sets reset_qbound[0 .. nq) to "no bound",
   ready for the next query batch */
hipError_t launch_probe_merge(const uint2* partial, const int* probes, int nq, int nprobe, const int* fb_index,
                              const int* ex_idx1,
                              const int* ex_d1, const int* ex_idx2, const int* ex_d2, int* idx1, int* d1, int* idx2,
                              int* d2, int* reset_count, int ncount, uint32_t* reset_qbound, hipStream_t s);
/* rows idx[i] (or the first n) of q, with labels lab_src[idx[i]] when given */
hipError_t launch_gather(const uint8_t* q, const int* idx, const int* lab_src, int n, uint8_t* qd, int* lab,
                         hipStream_t s);
/* the k nearest centroids of each query (Hamming), ties to the lower id */
hipError_t launch_centroid_probe(const uint8_t* q, int nq, const uint8_t* cen, int nb, int k, float mass, float temp,
                                 int* probes, float* pmax, bool sorted, hipStream_t s);
/* compaction of the low-confidence queries for the exact fallback */
hipError_t launch_fallback_compact(const uint8_t* q, int nq, const float* pmax, float thresh, int* fb_count,
                                   int* fb_index, int* fb_queries_idx, uint8_t* fb_q, hipStream_t s);

} // namespace orbk
