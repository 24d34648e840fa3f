/*
 * Learned bucket matcher kernels (gfx950): a BF16 MLP 256 -> H -> H -> C
 * (GELU, FP32 accumulate) that maps a descriptor to database buckets, its
 * training step, and the bucketed exact search.  No counterpart in the
 * reference; exact search (orb_match.hip) stays the fallback and the judge
 * of recall.
 *
 *   k_gemm      C = A B on v_mfma_f32_16x16x32_bf16, 64 x 64 workgroup tiles,
 *               A and B addressed through strides, epilogue fused: + bias,
 *               GELU, or the GELU-derivative product of the backward pass,
 *               FP32 and/or BF16 stores
 *   k_gemm_fwd  the forward products of 1024 rows and more: 256 x 128 tiles
 *               in an XCD-aware order, 32-deep K steps in three LDS stages
 *               filled by buffer loads with the LDS bit, epilogue with the
 *               centroid term on FP4 matrix cores, bias, GELU, BF16, FP32 or
 *               FP16 (query logits) stores through an LDS stage
 *   k_gemm_nt128, k_gemm_nt
 *               products with both operands contiguous along K below 1024
 *               rows and the data gradients: 128 x 128 tiles, 128-deep K
 *               steps with buffer loads (K a multiple of 128) or 64-deep
 *               steps double-buffered in LDS, LDS-staged epilogue
 *   k_gemm_tn, k_splitk_sum
 *               weight gradients A^T B (both contiguous along M and N): the
 *               batch split over ~256 workgroups, parts summed in order
 *   k_transpose16  bf16 transposes (K-contiguous weight copies)
 *   k_bits_pm1  descriptor bits -> +/-1 BF16 inputs
 *   k_softmax_ce, k_loss_sum
 *               row softmax, cross-entropy and its gradient (scaled 1/n); the
 *               epoch loss summed on the device in a fixed order
 *   k_colsum    bias gradients
 *   k_adam      Adam on FP32 master weights, BF16 copies refreshed
 *   k_probe     softmax max and the exact top-k buckets of each query's FP16
 *               logits, or the probability-mass prefix of them
 */
#include "orb_internal.h"
#include "orb_match_common.h"
#include "orb_mlp.h"
#include "orb_wave.h"

namespace orbk
{

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f)
{
    /* round to nearest even */
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float bf2f(uint16_t h)
{
    return __uint_as_float((uint32_t)h << 16);
}

/* x rounded to FP16 (nearest even), and FP16 bits <-> float */
__device__ __forceinline__ float round_h(float x)
{
    return (float)(_Float16)x;
}

__device__ __forceinline__ uint32_t h_bits(float x)
{
    return (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)x);
}

__device__ __forceinline__ float h2f(uint32_t h)
{
    return (float)__builtin_bit_cast(_Float16, (uint16_t)h);
}

/* erf(x) in the form 1 - exp(-t P(t)), t = min(|x|, 4), sign restored: P is a
   degree-7 polynomial fitted for this kernel by weighted least squares on
   [0, 4] against the float64 erf (weights t exp(-t P), the sensitivity of erf
   to P; Chebyshev nodes).  Evaluated in float32 the maximum absolute error is
   1.1e-7 (erf(4) rounds to 1); one branchless pass of 8 FMAs and v_exp_f32. */
__device__ __forceinline__ float erf_fit(float x)
{
    const float t = fminf(fabsf(x), 4.0f);
    float p = 0x1.03f97p-15f;
    p = fmaf(p, t, -0x1.3ffa26p-12f);
    p = fmaf(p, t, 0x1.0b9928p-10f);
    p = fmaf(p, t, 0x1.232c32p-11f);
    p = fmaf(p, t, -0x1.411e5p-6f);
    p = fmaf(p, t, 0x1.a59628p-4f);
    p = fmaf(p, t, 0x1.45eff6p-1f);
    p = fmaf(p, t, 0x1.20dd88p+0f);
    const float e = __builtin_amdgcn_exp2f(-(t * p) * 1.44269504f);
    return copysignf(1.0f - e, x);
}

/* gelu(z) = z (1 + erf(z / sqrt 2)) / 2 for the BF16 activations: with
   u = min(|z|, 4 sqrt 2) and h = erfc(u / sqrt 2) / 2 = 2^(-u Q(u) - 1),
   gelu = z - z h (z >= 0) or z h (z < 0).  Q is a degree-5 polynomial fitted
   by weighted least squares on (0, 4 sqrt 2] to -log2(erfc(u / sqrt 2)) / u
   (weights u erfc(u / sqrt 2), the sensitivity of erfc to Q; Chebyshev
   nodes).  Evaluated in float32 the error is at most 7.3e-7 absolute and
   0.26 % relative where |gelu| > 1e-4, under half a BF16 step. */
__device__ __forceinline__ float gelu(float z)
{
    const float u = fminf(fabsf(z), 5.65685425f);
    float q = -0x1.4d320ep-16f;
    q = fmaf(q, u, 0x1.5b3d9ep-11f);
    q = fmaf(q, u, -0x1.fc8a3ap-8f);
    q = fmaf(q, u, 0x1.b1e27cp-5f);
    q = fmaf(q, u, 0x1.d61574p-2f);
    q = fmaf(q, u, 0x1.26afbep+0f);
    const float h = __builtin_amdgcn_exp2f(fmaf(-u, q, -1.0f));
    const float zh = z * h;
    return z >= 0.f ? z - zh : zh;
}

__device__ __forceinline__ float gelu_grad(float z)
{
    return 0.5f * (1.0f + erf_fit(z * 0.70710678118654752f)) + z * 0.39894228040143268f * __expf(-0.5f * z * z);
}

/* 8 staged values -> 8 output halves: FP16 bits (values already rounded)
   with HALF, else BF16 of gelu(v) (GELU) or v */
template <bool GELU, bool HALF>
__device__ __forceinline__ uint4 pack_out16_t(const float (&v)[8])
{
    uint32_t w[4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
        if constexpr (HALF) {
            w[t] = h_bits(v[2 * t]) | (h_bits(v[2 * t + 1]) << 16);
        } else {
            const float e0 = GELU ? gelu(v[2 * t]) : v[2 * t];
            const float e1 = GELU ? gelu(v[2 * t + 1]) : v[2 * t + 1];
            w[t] = (uint32_t)f2bf(e0) | ((uint32_t)f2bf(e1) << 16);
        }
    }
    return make_uint4(w[0], w[1], w[2], w[3]);
}

/* the same with the flags of g */
__device__ __forceinline__ uint4 pack_out16(const GemmArgs& g, const float (&v)[8])
{
    uint32_t w[4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
        if (g.half) {
            w[t] = h_bits(v[2 * t]) | (h_bits(v[2 * t + 1]) << 16);
        } else {
            const float e0 = g.gelu ? gelu(v[2 * t]) : v[2 * t];
            const float e1 = g.gelu ? gelu(v[2 * t + 1]) : v[2 * t + 1];
            w[t] = (uint32_t)f2bf(e0) | ((uint32_t)f2bf(e1) << 16);
        }
    }
    return make_uint4(w[0], w[1], w[2], w[3]);
}

/* ------------------------------------------------------------------ GEMM */

constexpr int kGT = 64;        /* workgroup tile (M and N) */
constexpr int kGK = 32;        /* K step */
constexpr int kGLd = kGK + 8;  /* LDS row (bf16 elements), padded */

/* Tile loader: a 64 x 32 block of X(i, k) = X[i*si + k*sk] into LDS as
   dst[i][k].  Contiguous-k blocks move 8 elements per thread; otherwise the
   block is read along i (contiguous when si == 1) and stored transposed. */
__device__ __forceinline__ void load_tile(uint16_t (*dst)[kGLd], const uint16_t* __restrict__ X, long long si,
                                          long long sk, int i0, int k0)
{
    const int tid = threadIdx.x;
    if (sk == 1) {
        const int i = tid >> 2, kk = (tid & 3) * 8;
        const uint4 v = *(const uint4*)(X + (long long)(i0 + i) * si + k0 + kk);
        *(uint4*)&dst[i][kk] = v;
    } else if (si == 1) {
        /* 32 rows of k, 64 contiguous i each: thread -> k = tid/8, i = 8*(tid%8) */
        const int kk = tid >> 3, i = (tid & 7) * 8;
        const uint4 v = *(const uint4*)(X + (long long)(k0 + kk) * sk + i0 + i);
        const uint16_t* e = (const uint16_t*)&v;
#pragma unroll
        for (int t = 0; t < 8; t++) dst[i + t][kk] = e[t];
    } else {
        for (int e = tid; e < kGT * kGK; e += 256) {
            const int i = e / kGK, kk = e % kGK;
            dst[i][kk] = X[(long long)(i0 + i) * si + (long long)(k0 + kk) * sk];
        }
    }
}

__global__ __launch_bounds__(256) void k_gemm(GemmArgs g)
{
    __shared__ __align__(16) uint16_t sA[kGT][kGLd];
    __shared__ __align__(16) uint16_t sB[kGT][kGLd];
    const int tiles_n = g.N / kGT;
    const int m0 = (blockIdx.x / tiles_n) * kGT, n0 = (blockIdx.x % tiles_n) * kGT;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int wm = (wid >> 1) * 32, wn = (wid & 1) * 32;
    const int row = lane & 15, grp = lane >> 4;
    f32x4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) acc[a][b] = (f32x4){0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < g.K; k0 += kGK) {
        __syncthreads();
        load_tile(sA, g.A, g.sam, g.sak, m0, k0);
        load_tile(sB, g.B, g.sbn, g.sbk, n0, k0);
        __syncthreads();
        bf16x8 av[2], bv[2];
#pragma unroll
        for (int a = 0; a < 2; a++) av[a] = *(const bf16x8*)&sA[wm + 16 * a + row][8 * grp];
#pragma unroll
        for (int b = 0; b < 2; b++) bv[b] = *(const bf16x8*)&sB[wn + 16 * b + row][8 * grp];
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int b = 0; b < 2; b++)
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[a], bv[b], acc[a][b], 0, 0, 0);
    }
    /* lane holds C[4*grp + r][row] of each 16 x 16 block */
#pragma unroll
    for (int a = 0; a < 2; a++) {
#pragma unroll
        for (int b = 0; b < 2; b++) {
            const int n = n0 + wn + 16 * b + row;
            const float bias = g.bias ? g.bias[n] : 0.f;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int m = m0 + wm + 16 * a + 4 * grp + r;
                const long long o = (long long)m * g.ldc + n;
                float z = acc[a][b][r];
                if (g.addin) z = z + g.addin[o];
                z = z + bias;
                if (g.zmul) z = z * gelu_grad(g.zmul[o]);
                if (g.out32) g.out32[o] = z;
                if (g.out16) g.out16[o] = f2bf(g.gelu ? gelu(z) : z);
            }
        }
    }
}

/* C = A^T B for operands contiguous along M and N (A(m, k) = A[k*sak + m],
   B(k, n) = B[k*sbk + n]): the weight-gradient products, K being the batch.
   128 x 128 tiles, four wavefronts of 64 x 64 (4 x 4 MFMA blocks); the K
   range is split into parts of `kper` (one workgroup per tile and part, so a
   product of a few tiles still fills the chip), each part's sums written to
   ws[part][m][n] and added up in part order by k_splitk_sum (the result does
   not depend on scheduling).  The MFMA wants 8 consecutive k per lane and
   the operands are contiguous along m: a thread loads 8 rows of k x 2
   columns (8 dwords, 256 contiguous bytes per row across a wavefront) and
   writes each column's 8 k values as one 16-byte LDS row segment; the next
   step's dwords are loaded during the current step's MFMAs.  Columns past M
   or N read zeros (buffer loads) and are not stored. */
constexpr int kTT = 128, kTK = 32, kTLd = kTK + 8; /* LDS row: 80 B */

__global__ __launch_bounds__(256, 2) void k_gemm_tn(GemmArgs g, int kper)
{
    __shared__ __align__(16) uint16_t sA[2][kTT * kTLd];
    __shared__ __align__(16) uint16_t sB[2][kTT * kTLd];
    const int tiles_n = (g.N + kTT - 1) / kTT, nt = ((g.M + kTT - 1) / kTT) * tiles_n;
    const int part = blockIdx.x / nt, t = blockIdx.x % nt;
    const int m0 = (t / tiles_n) * kTT, n0 = (t % tiles_n) * kTT;
    const int kb = part * kper, ke = min(g.K, kb + kper);
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int wm = (wid >> 1) * 64, wn = (wid & 1) * 64;
    const int row = lane & 15, grp = lane >> 4;
    /* loader: column pair cp (columns 2 cp, 2 cp + 1), k rows 8 ko .. 8 ko + 7 of a step */
    const int cp = tid & 63, ko = tid >> 6;
    const __amdgpu_buffer_rsrc_t rA =
        __builtin_amdgcn_make_buffer_rsrc((void*)g.A, 0, (int)min((long long)g.K * g.sak * 2, 0x7fffffffll), 0x00020000);
    const __amdgpu_buffer_rsrc_t rB =
        __builtin_amdgcn_make_buffer_rsrc((void*)g.B, 0, (int)min((long long)g.K * g.sbk * 2, 0x7fffffffll), 0x00020000);
    const uint32_t oob = 0x80000000u; /* columns past M or N */
    const uint32_t ca = m0 + 2 * cp < g.M ? (uint32_t)((m0 + 2 * cp) * 2) : oob;
    const uint32_t cb = n0 + 2 * cp < g.N ? (uint32_t)((n0 + 2 * cp) * 2) : oob;
    uint32_t xa[8], xb[8];
    auto gload = [&](int k0) {
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int k = k0 + 8 * ko + j;
            const bool in = k < ke;
            xa[j] = __builtin_amdgcn_raw_buffer_load_b32(rA, in ? ca : oob, (uint32_t)(k * g.sak * 2), 0);
            xb[j] = __builtin_amdgcn_raw_buffer_load_b32(rB, in ? cb : oob, (uint32_t)(k * g.sbk * 2), 0);
        }
    };
    auto lstore = [&](int buf) {
        /* column 2 cp: the low halves of the 8 dwords, column 2 cp + 1: the high halves */
        const uint4 alo = make_uint4(__builtin_amdgcn_perm(xa[1], xa[0], 0x05040100u), __builtin_amdgcn_perm(xa[3], xa[2], 0x05040100u),
                                     __builtin_amdgcn_perm(xa[5], xa[4], 0x05040100u), __builtin_amdgcn_perm(xa[7], xa[6], 0x05040100u));
        const uint4 ahi = make_uint4(__builtin_amdgcn_perm(xa[1], xa[0], 0x07060302u), __builtin_amdgcn_perm(xa[3], xa[2], 0x07060302u),
                                     __builtin_amdgcn_perm(xa[5], xa[4], 0x07060302u), __builtin_amdgcn_perm(xa[7], xa[6], 0x07060302u));
        const uint4 blo = make_uint4(__builtin_amdgcn_perm(xb[1], xb[0], 0x05040100u), __builtin_amdgcn_perm(xb[3], xb[2], 0x05040100u),
                                     __builtin_amdgcn_perm(xb[5], xb[4], 0x05040100u), __builtin_amdgcn_perm(xb[7], xb[6], 0x05040100u));
        const uint4 bhi = make_uint4(__builtin_amdgcn_perm(xb[1], xb[0], 0x07060302u), __builtin_amdgcn_perm(xb[3], xb[2], 0x07060302u),
                                     __builtin_amdgcn_perm(xb[5], xb[4], 0x07060302u), __builtin_amdgcn_perm(xb[7], xb[6], 0x07060302u));
        *(uint4*)&sA[buf][(2 * cp) * kTLd + 8 * ko] = alo;
        *(uint4*)&sA[buf][(2 * cp + 1) * kTLd + 8 * ko] = ahi;
        *(uint4*)&sB[buf][(2 * cp) * kTLd + 8 * ko] = blo;
        *(uint4*)&sB[buf][(2 * cp + 1) * kTLd + 8 * ko] = bhi;
    };
    f32x4 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) acc[a][b] = (f32x4){0.f, 0.f, 0.f, 0.f};
    const int nk = (ke - kb + kTK - 1) / kTK;
    if (nk > 0) {
        gload(kb);
        lstore(0);
    }
    __syncthreads();
    for (int kt = 0; kt < nk; kt++) {
        const int cur = kt & 1;
        if (kt + 1 < nk) gload(kb + (kt + 1) * kTK);
        __builtin_amdgcn_sched_barrier(0);
        bf16x8 av[4], bv[4];
#pragma unroll
        for (int a = 0; a < 4; a++) av[a] = *(const bf16x8*)&sA[cur][(wm + 16 * a + row) * kTLd + 8 * grp];
#pragma unroll
        for (int b = 0; b < 4; b++) bv[b] = *(const bf16x8*)&sB[cur][(wn + 16 * b + row) * kTLd + 8 * grp];
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
            for (int b = 0; b < 4; b++)
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[a], bv[b], acc[a][b], 0, 0, 0);
        __builtin_amdgcn_sched_barrier(0);
        if (kt + 1 < nk) lstore(cur ^ 1);
        __syncthreads();
    }
    /* lane holds C[4 grp + r][row] of each 16 x 16 block */
    float* const ws = g.ws + (long long)part * g.M * g.N;
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) {
            const int n = n0 + wn + 16 * b + row;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int m = m0 + wm + 16 * a + 4 * grp + r;
                if (m < g.M && n < g.N) ws[(long long)m * g.N + n] = acc[a][b][r];
            }
        }
}

/* out[m*ldc + n] = ws[0][m][n] + ws[1][m][n] + ... (parts in order), four
   outputs per thread */
__global__ __launch_bounds__(256) void k_splitk_sum(const float* __restrict__ ws, int parts, int M, int N, long long ldc,
                                                    float* __restrict__ out)
{
    const long long i = ((long long)blockIdx.x * blockDim.x + threadIdx.x) * 4;
    if (i >= (long long)M * N) return;
    float4 v = *(const float4*)(ws + i);
    for (int p = 1; p < parts; p++) {
        const float4 w = *(const float4*)(ws + (long long)p * M * N + i);
        v.x += w.x;
        v.y += w.y;
        v.z += w.z;
        v.w += w.w;
    }
    const long long m = i / N, n = i % N;
    *(float4*)(out + m * ldc + n) = v;
}

/* Hamming distance of two 32-byte descriptors */
__device__ __forceinline__ int ham256(const uint4& a0, const uint4& a1, const uint4& b0, const uint4& b1)
{
    return __popc(a0.x ^ b0.x) + __popc(a0.y ^ b0.y) + __popc(a0.z ^ b0.z) + __popc(a0.w ^ b0.w) +
           __popc(a1.x ^ b1.x) + __popc(a1.y ^ b1.y) + __popc(a1.z ^ b1.z) + __popc(a1.w ^ b1.w);
}

/* the centroid term of output (m, n): cscale * (256 - 2 d), exact in FP32 */
__device__ __forceinline__ float centroid_term(const GemmArgs& g, int d)
{
    return g.cscale * (float)(256 - 2 * d);
}

constexpr int kNT = 128; /* workgroup tile of the NT products (M and N) */

/* Epilogue of the NT products for one 128 x 128 tile: acc as the MFMA left it
   (lane holds C[4*grp + r][row] of each 16 x 16 block); st is LDS of at least
   128 x 132 floats, free once the K loop has ended on a barrier. */
__device__ __forceinline__ void nt_epilogue(const GemmArgs& g, f32x4 (&acc)[4][4], float* st, int m0, int n0)
{
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int wm = (wid >> 1) * 64, wn = (wid & 1) * 64;
    const int row = lane & 15, grp = lane >> 4;
    if (g.cbits) {
        /* acc += the centroid term: this lane's 4 columns (one per block
           column) against each of its 16 rows */
        uint4 c[4][2];
#pragma unroll
        for (int b = 0; b < 4; b++) {
            const uint4* cp = (const uint4*)(g.cbits + (long long)min(n0 + wn + 16 * b + row, g.N - 1) * 32);
            c[b][0] = cp[0];
            c[b][1] = cp[1];
        }
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const uint4* qp = (const uint4*)(g.qbits + (long long)min(m0 + wm + 16 * a + 4 * grp + r, g.M - 1) * 32);
                const uint4 q0 = qp[0], q1 = qp[1];
#pragma unroll
                for (int b = 0; b < 4; b++) acc[a][b][r] += centroid_term(g, ham256(q0, q1, c[b][0], c[b][1]));
            }
    }
    if (!g.addin && !g.zmul) {
        /* z = acc + bias staged as a 128 x 128 fp32 tile in LDS (the loop ended
           on a barrier), then written along rows: 16 B per lane */
        constexpr int kOLd = kNT + 4;
#pragma unroll
        for (int b = 0; b < 4; b++) {
            const int nl = wn + 16 * b + row;
            const float bias = g.bias && n0 + nl < g.N ? g.bias[n0 + nl] : 0.f;
#pragma unroll
            for (int a = 0; a < 4; a++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const float z = acc[a][b][r] + bias;
                    st[(wm + 16 * a + 4 * grp + r) * kOLd + nl] = g.half ? round_h(z) : z;
                }
        }
        __syncthreads();
        if (g.out32) {
#pragma unroll 4
            for (int i = 0; i < 16; i++) {
                const int c = tid + 256 * i, rl = c >> 5, cl = (c & 31) * 4;
                const int m = m0 + rl, n = n0 + cl;
                if (m < g.M && n < g.N)
                    *(float4*)(g.out32 + (long long)m * g.ldc + n) = *(const float4*)(st + rl * kOLd + cl);
            }
        }
        if (g.out16) {
#pragma unroll 2
            for (int i = 0; i < 8; i++) {
                const int c = tid + 256 * i, rl = c >> 4, cl = (c & 15) * 8;
                const int m = m0 + rl, n = n0 + cl;
                if (m < g.M && n < g.N) {
                    const float4 lo = *(const float4*)(st + rl * kOLd + cl);
                    const float4 hi = *(const float4*)(st + rl * kOLd + cl + 4);
                    const float v[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
                    *(uint4*)(g.out16 + (long long)m * g.ldc + n) = pack_out16(g, v);
                }
            }
        }
        return;
    }
    /* lane holds C[4*grp + r][row] of each 16 x 16 block */
#pragma unroll
    for (int a = 0; a < 4; a++) {
#pragma unroll
        for (int b = 0; b < 4; b++) {
            const int n = n0 + wn + 16 * b + row;
            if (n >= g.N) continue;
            const float bias = g.bias ? g.bias[n] : 0.f;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int m = m0 + wm + 16 * a + 4 * grp + r;
                if (m >= g.M) continue;
                const long long o = (long long)m * g.ldc + n;
                float z = acc[a][b][r];
                if (g.addin) z = z + g.addin[o];
                z = z + bias;
                if (g.zmul) z = z * gelu_grad(g.zmul[o]);
                if (g.out32) g.out32[o] = z;
                if (g.out16) g.out16[o] = f2bf(g.gelu ? gelu(z) : z);
            }
        }
    }
}

/* C = A B with both operands contiguous along K (A(m, k) = A[m*sam + k],
   B(k, n) = B[n*sbn + k]): 128 x 128 workgroup tiles, four wavefronts of
   64 x 64 (4 x 4 MFMA blocks), 64-deep K steps double-buffered in LDS with the
   next step's global loads issued before the current step's MFMAs.  Rows
   beyond M or N are read clamped and never stored.  Same epilogue as k_gemm. */
constexpr int kNK = 64;
constexpr int kNLd = kNK + 8; /* LDS row (bf16 elements): 144 B, conflict-free b128 reads */

__global__ __launch_bounds__(256, 2) void k_gemm_nt(GemmArgs g)
{
    /* A and B tiles (two buffers each); after the loop, the fp32 output tile */
    __shared__ __align__(16) uint16_t s_lds[4 * kNT * kNLd];
    uint16_t (*sA)[kNT * kNLd] = (uint16_t (*)[kNT * kNLd])s_lds;
    uint16_t (*sB)[kNT * kNLd] = (uint16_t (*)[kNT * kNLd])(s_lds + 2 * kNT * kNLd);
    /* tiles in M-major order, each XCD a contiguous range: the N tiles of an
       M block share its rows of A in one L2 */
    const int tiles_n = (g.N + kNT - 1) / kNT, u = xcd_remap(blockIdx.x, gridDim.x);
    const int m0 = (u / tiles_n) * kNT, n0 = (u % tiles_n) * kNT;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int wm = (wid >> 1) * 64, wn = (wid & 1) * 64;
    const int row = lane & 15, grp = lane >> 4;
    /* loader: thread -> tile rows lr + 32 j, k offset lk (8 elements) */
    const int lr = tid >> 3, lk = (tid & 7) * 8;
    const uint16_t* pa[4];
    const uint16_t* pb[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const long long mr = min(m0 + lr + 32 * j, g.M - 1), nr = min(n0 + lr + 32 * j, g.N - 1);
        pa[j] = g.A + mr * g.sam + lk;
        pb[j] = g.B + nr * g.sbn + lk;
    }
    const int sto = (lr * kNLd + lk) * 2; /* byte offset of this thread's LDS stores */
    char* const lA = (char*)&sA[0][0] + sto;
    char* const lB = (char*)&sB[0][0] + sto;
    constexpr int kBuf = kNT * kNLd * 2, kRow32 = 32 * kNLd * 2;
    const char* rA = (const char*)&sA[0][0] + ((wm + row) * kNLd + 8 * grp) * 2;
    const char* rB = (const char*)&sB[0][0] + ((wn + row) * kNLd + 8 * grp) * 2;
    const int nk = g.K / kNK;
    uint4 a0 = *(const uint4*)pa[0], a1 = *(const uint4*)pa[1], a2 = *(const uint4*)pa[2], a3 = *(const uint4*)pa[3];
    uint4 b0 = *(const uint4*)pb[0], b1 = *(const uint4*)pb[1], b2 = *(const uint4*)pb[2], b3 = *(const uint4*)pb[3];
    *(uint4*)(lA) = a0;
    *(uint4*)(lA + kRow32) = a1;
    *(uint4*)(lA + 2 * kRow32) = a2;
    *(uint4*)(lA + 3 * kRow32) = a3;
    *(uint4*)(lB) = b0;
    *(uint4*)(lB + kRow32) = b1;
    *(uint4*)(lB + 2 * kRow32) = b2;
    *(uint4*)(lB + 3 * kRow32) = b3;
    __syncthreads();
    f32x4 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) acc[a][b] = (f32x4){0.f, 0.f, 0.f, 0.f};
    for (int kt = 0; kt < nk; kt++) {
        const int cur = (kt & 1) * kBuf, nxt = kBuf - cur;
        /* the next step's tiles, in flight during this step's MFMAs (the last
           step reloads its own: harmless, never stored over live data) */
        const int k1 = min(kt + 1, nk - 1) * kNK;
        a0 = *(const uint4*)(pa[0] + k1);
        a1 = *(const uint4*)(pa[1] + k1);
        a2 = *(const uint4*)(pa[2] + k1);
        a3 = *(const uint4*)(pa[3] + k1);
        b0 = *(const uint4*)(pb[0] + k1);
        b1 = *(const uint4*)(pb[1] + k1);
        b2 = *(const uint4*)(pb[2] + k1);
        b3 = *(const uint4*)(pb[3] + k1);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int ks = 0; ks < 2; ks++) {
            bf16x8 av[4], bv[4];
#pragma unroll
            for (int a = 0; a < 4; a++) av[a] = *(const bf16x8*)(rA + cur + (16 * a * kNLd + 32 * ks) * 2);
#pragma unroll
            for (int b = 0; b < 4; b++) bv[b] = *(const bf16x8*)(rB + cur + (16 * b * kNLd + 32 * ks) * 2);
#pragma unroll
            for (int a = 0; a < 4; a++)
#pragma unroll
                for (int b = 0; b < 4; b++)
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[a], bv[b], acc[a][b], 0, 0, 0);
        }
        __builtin_amdgcn_sched_barrier(0);
        *(uint4*)(lA + nxt) = a0;
        *(uint4*)(lA + nxt + kRow32) = a1;
        *(uint4*)(lA + nxt + 2 * kRow32) = a2;
        *(uint4*)(lA + nxt + 3 * kRow32) = a3;
        *(uint4*)(lB + nxt) = b0;
        *(uint4*)(lB + nxt + kRow32) = b1;
        *(uint4*)(lB + nxt + 2 * kRow32) = b2;
        *(uint4*)(lB + nxt + 3 * kRow32) = b3;
        __syncthreads();
    }
    nt_epilogue(g, acc, (float*)s_lds, m0, n0);
}

/* The same product with 128-deep K steps (K a multiple of 128): one
   LDS tile of A and one of B (69.6 KB: two workgroups per CU), the next
   step's 16 B loads (buffer loads: rows beyond M or N read as zeros) in
   registers while the 64 MFMAs of the current step run, two barriers per
   step.  Twice the MFMA work per exposed load latency of the 64-deep form. */
constexpr int kNK2 = 128;
constexpr int kNLd2 = kNK2 + 8; /* LDS row (bf16 elements): 272 B, conflict-free b128 reads */

__global__ __launch_bounds__(256, 2) void k_gemm_nt128(GemmArgs g)
{
    __shared__ __align__(16) uint16_t s_lds[2 * kNT * kNLd2];
    /* tiles in M-major order, each XCD a contiguous range: the N tiles of an
       M block share its rows of A in one L2 */
    const int tiles_n = (g.N + kNT - 1) / kNT, u = xcd_remap(blockIdx.x, gridDim.x);
    const int m0 = (u / tiles_n) * kNT, n0 = (u % tiles_n) * kNT;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int wm = (wid >> 1) * 64, wn = (wid & 1) * 64;
    const int row = lane & 15, grp = lane >> 4;
    /* loader: thread -> tile rows lr + 16 j (j < 8), k offset lk (8 elements) */
    const int lr = tid >> 4, lk = (tid & 15) * 8;
    const __amdgpu_buffer_rsrc_t rA = __builtin_amdgcn_make_buffer_rsrc((void*)g.A, 0, (int)(g.M * g.sam * 2), 0x00020000);
    const __amdgpu_buffer_rsrc_t rB = __builtin_amdgcn_make_buffer_rsrc((void*)g.B, 0, (int)(g.N * g.sbn * 2), 0x00020000);
    const uint32_t vA = (uint32_t)((lr * g.sam + lk) * 2), vB = (uint32_t)((lr * g.sbn + lk) * 2);
    char* const lA = (char*)s_lds + (lr * kNLd2 + lk) * 2;
    char* const lB = lA + kNT * kNLd2 * 2;
    constexpr int kRow16 = 16 * kNLd2 * 2;
    const char* rdA = (const char*)s_lds + ((wm + row) * kNLd2 + 8 * grp) * 2;
    const char* rdB = (const char*)s_lds + (kNT * kNLd2 + (wn + row) * kNLd2 + 8 * grp) * 2;
    const int nk = g.K / kNK2;
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    u32x4 x[16];
    auto gload = [&](int kt) {
        const uint32_t kb = (uint32_t)(kt * kNK2 * 2);
#pragma unroll
        for (int j = 0; j < 8; j++) {
            x[j] = __builtin_amdgcn_raw_buffer_load_b128(rA, vA, (uint32_t)((m0 + 16 * j) * g.sam * 2) + kb, 0);
            x[8 + j] = __builtin_amdgcn_raw_buffer_load_b128(rB, vB, (uint32_t)((n0 + 16 * j) * g.sbn * 2) + kb, 0);
        }
    };
    auto lstore = [&]() {
#pragma unroll
        for (int j = 0; j < 8; j++) {
            *(u32x4*)(lA + j * kRow16) = x[j];
            *(u32x4*)(lB + j * kRow16) = x[8 + j];
        }
    };
    f32x4 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) acc[a][b] = (f32x4){0.f, 0.f, 0.f, 0.f};
    gload(0);
    lstore();
    __syncthreads();
    for (int kt = 0; kt < nk; kt++) {
        /* the next step in flight (the last step reloads itself: never stored) */
        gload(min(kt + 1, nk - 1));
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int ks = 0; ks < 4; ks++) {
            bf16x8 av[4], bv[4];
#pragma unroll
            for (int a = 0; a < 4; a++) av[a] = *(const bf16x8*)(rdA + (16 * a * kNLd2 + 32 * ks) * 2);
#pragma unroll
            for (int b = 0; b < 4; b++) bv[b] = *(const bf16x8*)(rdB + (16 * b * kNLd2 + 32 * ks) * 2);
#pragma unroll
            for (int a = 0; a < 4; a++)
#pragma unroll
                for (int b = 0; b < 4; b++)
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[a], bv[b], acc[a][b], 0, 0, 0);
        }
        __builtin_amdgcn_sched_barrier(0);
        __syncthreads();
        if (kt + 1 < nk) lstore();
        __syncthreads();
    }
    nt_epilogue(g, acc, (float*)s_lds, m0, n0);
}

/* The forward products (no addin / zmul; K a multiple of 32) on 256 x
   128 workgroup tiles (M x N): four wavefronts of 128 x 64 outputs (8 x 4
   MFMA blocks, 128 FP32 accumulators per lane), 32-deep K steps.  The tiles
   travel from memory straight into LDS (buffer loads with the LDS bit, 16
   bytes per lane, rows past M or N read as zeros) through three stages of
   24 KB, two steps ahead of the MFMAs, with no staging registers; two
   workgroups share a CU, so one's epilogue overlaps the other's MFMAs.  In
   a stage a 32-element row is 64 bytes whose four 16-byte chunks sit at
   slot chunk ^ ((row >> 2) & 3), so the 16 rows of a fragment read hit 16
   distinct bank groups.  The MFMA takes the weight rows as its first operand
   and the activation rows as its second, so a lane holds four consecutive
   output columns of one row: bias and GELU are applied in registers, each
   wavefront stages 16 rows at a time in LDS and writes them back as whole
   row segments (16 bytes per lane).  The K order of the sums is the 128 x
   128 kernels' (32-deep MFMA steps in K order), so the results are the
   same. */
constexpr int kFM = 256, kFN = 128;
constexpr int kFK = 32;
constexpr int kFStage = (kFM + kFN) * kFK * 2; /* bytes */
#define ORB_LDS_PTR(p) ((__attribute__((address_space(3))) void*)(p))

template <bool GELU, bool HALF>
__global__ __launch_bounds__(256, 2) void k_gemm_fwd(GemmArgs g)
{
    __shared__ __align__(16) char s_st[3 * kFStage];
    const int tiles_n = (g.N + kFN - 1) / kFN, u = xcd_remap(blockIdx.x, gridDim.x);
    const int m0 = (u / tiles_n) * kFM, n0 = (u % tiles_n) * kFN;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wid = __builtin_amdgcn_readfirstlane(tid >> 6); /* uniform: the DMA's LDS address goes to M0 */
    const int wm = (wid >> 1) * 128, wn = (wid & 1) * 64;
    const int row = lane & 15, grp = lane >> 4;
    const __amdgpu_buffer_rsrc_t rA = __builtin_amdgcn_make_buffer_rsrc((void*)g.A, 0, (int)(g.M * g.sam * 2), 0x00020000);
    const __amdgpu_buffer_rsrc_t rB = __builtin_amdgcn_make_buffer_rsrc((void*)g.B, 0, (int)(g.N * g.sbn * 2), 0x00020000);
    /* DMA: instruction j of wavefront wid fills the 16-row block q = wid + 4 j
       of a stage (blocks 0..15 the activation rows, 16..23 the weight rows);
       lane -> row lane >> 2 of the block, slot lane & 3, which holds chunk
       (lane & 3) ^ ((lane >> 4) & 3) of that row */
    const int drow = lane >> 2, dch = (lane & 3) ^ ((lane >> 4) & 3);
    const uint32_t vA = (uint32_t)((drow * g.sam + 8 * dch) * 2), vB = (uint32_t)((drow * g.sbn + 8 * dch) * 2);
    const int nk = g.K / kFK;
    auto issue = [&](char* st, int kt) {
        const uint32_t kb = (uint32_t)(kt * kFK * 2);
#pragma unroll
        for (int j = 0; j < 4; j++)
            __builtin_amdgcn_raw_ptr_buffer_load_lds(rA, ORB_LDS_PTR(st + (wid + 4 * j) * 1024), 16, vA,
                                                     (uint32_t)(m0 + 16 * (wid + 4 * j)) * (uint32_t)g.sam * 2 + kb,
                                                     0, 0);
#pragma unroll
        for (int j = 0; j < 2; j++)
            __builtin_amdgcn_raw_ptr_buffer_load_lds(rB, ORB_LDS_PTR(st + (16 + wid + 4 * j) * 1024), 16, vB,
                                                     (uint32_t)(n0 + 16 * (wid + 4 * j)) * (uint32_t)g.sbn * 2 + kb,
                                                     0, 0);
    };
    /* fragment reads: row `row` of a 16-row block, chunk grp at its slot */
    const int rdo = row * 64 + 16 * (grp ^ ((lane >> 2) & 3));
    f32x4 acc[4][8];
#pragma unroll
    for (int b = 0; b < 4; b++)
#pragma unroll
        for (int a = 0; a < 8; a++) acc[b][a] = (f32x4){0.f, 0.f, 0.f, 0.f};
    /* the activation fragments in two halves of four (fewer live registers) */
    auto mfma_step = [&](const char* st) {
        const char* const px = st + wm * 64 + rdo;
        const char* const pw = st + kFM * 64 + wn * 64 + rdo;
        bf16x8 wv[4], xv[4];
#pragma unroll
        for (int b = 0; b < 4; b++) wv[b] = *(const bf16x8*)(pw + 16 * b * 64);
#pragma unroll
        for (int h = 0; h < 2; h++) {
#pragma unroll
            for (int a = 0; a < 4; a++) xv[a] = *(const bf16x8*)(px + 16 * (4 * h + a) * 64);
#pragma unroll
            for (int a = 0; a < 4; a++)
#pragma unroll
                for (int b = 0; b < 4; b++)
                    acc[b][4 * h + a] =
                        __builtin_amdgcn_mfma_f32_16x16x32_bf16(wv[b], xv[a], acc[b][4 * h + a], 0, 0, 0);
        }
    };
    /* step kt: wait for its stage (this wavefront's six loads of it; those of
       step kt + 1 may stay in flight) and for the reads of the stage about to
       be refilled, barrier, refill that stage with step kt + 2, MFMAs */
    auto step = [&](char* cur, char* refill, int kt) {
        if (kt + 1 < nk) __builtin_amdgcn_s_waitcnt(0x0076); /* vmcnt(6) lgkmcnt(0) */
        else __builtin_amdgcn_s_waitcnt(0x0070);             /* vmcnt(0) lgkmcnt(0) */
        __builtin_amdgcn_s_barrier();
        if (kt + 2 < nk) issue(refill, kt + 2);
        mfma_step(cur);
    };
    issue(s_st, 0);
    if (nk > 1) issue(s_st + kFStage, 1);
    for (int kt = 0, c = 0; kt < nk; kt++, c = c == 2 ? 0 : c + 1)
        step(s_st + c * kFStage, s_st + (c == 0 ? 2 : c - 1) * kFStage, kt);
    __syncthreads(); /* every load landed and every read done: stage 0 becomes the epilogue's */
    /* epilogue: lane holds z[m][n .. n+3] of each block (m = the block's row
       `row`, n = 4 grp); per 16-row block of the wavefront's tile the values go
       through this wavefront's 16 x 64 stage (the K loop ended on a barrier) */
    if (g.cbits) {
        /* acc += the centroid term on the matrix cores: descriptor bits as FP4
           codes +1 (set) / -1 (clear), so a 256-deep FP4 contraction of a
           centroid row with a query row is 256 - 2 d exactly (the same value
           the popcount form gives; the same FMA then adds it).  Codes of every
           byte value from a table in the third stage's LDS (free after the
           loop); lane (grp, row) takes words grp and 4 + grp of its centroid
           (A operand, the output column) and of its query (B, the row). */
        uint32_t* const tab = (uint32_t*)(s_st + 2 * kFStage);
        tab[tid] = fp4_pm1_8((uint32_t)tid);
        __syncthreads();
        v4i cc[4][2];
#pragma unroll
        for (int b = 0; b < 4; b++) {
            const uint32_t* cp = (const uint32_t*)(g.cbits + (long long)min(n0 + wn + 16 * b + row, g.N - 1) * 32);
            cc[b][0] = fp4_word_lut(cp[grp], tab);
            cc[b][1] = fp4_word_lut(cp[4 + grp], tab);
        }
#pragma unroll
        for (int a = 0; a < 8; a++) {
            const uint32_t* qp = (const uint32_t*)(g.qbits + (long long)min(m0 + wm + 16 * a + row, g.M - 1) * 32);
            const v4i q0 = fp4_word_lut(qp[grp], tab), q1 = fp4_word_lut(qp[4 + grp], tab);
#pragma unroll
            for (int b = 0; b < 4; b++) {
                v4f t = mfma_fp4(cc[b][0], q0, (v4f){0.f, 0.f, 0.f, 0.f});
                t = mfma_fp4(cc[b][1], q1, t);
#pragma unroll
                for (int i = 0; i < 4; i++) acc[b][a][i] += g.cscale * t[i];
            }
        }
    }
    float* const st = (float*)s_st + wid * 16 * 68;
    float4 bias[4];
#pragma unroll
    for (int b = 0; b < 4; b++) {
        const int n = n0 + wn + 16 * b + 4 * grp;
        bias[b] = g.bias && n < g.N ? *(const float4*)(g.bias + n) : make_float4(0.f, 0.f, 0.f, 0.f);
    }
#pragma unroll
    for (int a = 0; a < 8; a++) {
#pragma unroll
        for (int b = 0; b < 4; b++) {
            float4 z = make_float4(acc[b][a][0] + bias[b].x, acc[b][a][1] + bias[b].y, acc[b][a][2] + bias[b].z,
                                   acc[b][a][3] + bias[b].w);
            /* FP32 stores of the rounded values; FP16 stores round in the pack */
            if constexpr (HALF)
                if (g.out32) z = make_float4(round_h(z.x), round_h(z.y), round_h(z.z), round_h(z.w));
            *(float4*)(st + row * 68 + 16 * b + 4 * grp) = z;
        }
        wave_sync_lds();
        if (g.out32) {
            /* 16 lanes per row, 4 rows per pass */
#pragma unroll
            for (int p = 0; p < 4; p++) {
                const int rl = 4 * p + (lane >> 4), cl = 4 * (lane & 15);
                const int m = m0 + wm + 16 * a + rl, n = n0 + wn + cl;
                if (m < g.M && n < g.N)
                    *(float4*)(g.out32 + (long long)m * g.ldc + n) = *(const float4*)(st + rl * 68 + cl);
            }
        }
        if (g.out16) {
            /* 8 lanes per row, 8 rows per pass */
#pragma unroll
            for (int p = 0; p < 2; p++) {
                const int rl = 8 * p + (lane >> 3), cl = 8 * (lane & 7);
                const int m = m0 + wm + 16 * a + rl, n = n0 + wn + cl;
                if (m < g.M && n < g.N) {
                    const float4 lo = *(const float4*)(st + rl * 68 + cl);
                    const float4 hi = *(const float4*)(st + rl * 68 + cl + 4);
                    const float v[8] = {lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
                    *(uint4*)(g.out16 + (long long)m * g.ldc + n) = pack_out16_t<GELU, HALF>(v);
                }
            }
        }
        wave_sync_lds(); /* the stage is rewritten for the next block */
    }
}

/* dst (C x R) = src (R x C) transposed, bf16 bits; R and C multiples of 64 */
__global__ __launch_bounds__(256) void k_transpose16(const uint16_t* __restrict__ src, int R, int C,
                                                     uint16_t* __restrict__ dst)
{
    __shared__ uint16_t t[64][66];
    const int tc = C / 64;
    const int r0 = (blockIdx.x / tc) * 64, c0 = (blockIdx.x % tc) * 64;
    for (int e = threadIdx.x; e < 64 * 64; e += 256) t[e >> 6][e & 63] = src[(long long)(r0 + (e >> 6)) * C + c0 + (e & 63)];
    __syncthreads();
    for (int e = threadIdx.x; e < 64 * 64; e += 256)
        dst[(long long)(c0 + (e >> 6)) * R + r0 + (e & 63)] = t[e & 63][e >> 6];
}

/* ------------------------------------------------------------ elementwise */

/* descriptor bits -> +/-1 bf16 (bit k of the descriptor -> column k): one
   thread per descriptor byte, 8 bf16 values in one 16-byte store */
__global__ void k_bits_pm1(const uint8_t* __restrict__ desc, int n, uint16_t* __restrict__ X)
{
    const uint32_t pos = 0x3F80u, neg = 0xBF80u; /* bf16 +1, -1 */
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)n * 32) return;
    const uint32_t b = desc[i];
    uint32_t w[4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const uint32_t lo = (b >> (2 * t)) & 1u ? pos : neg, hi = (b >> (2 * t + 1)) & 1u ? pos : neg;
        w[t] = lo | (hi << 16);
    }
    *(uint4*)(X + i * 8) = make_uint4(w[0], w[1], w[2], w[3]);
}

/* acc[0] += loss[0] + .. + loss[n-1]: one workgroup, a fixed summation order
   (strided per-thread sums, then a tree), so the reported loss of a training
   run does not depend on scheduling */
__global__ __launch_bounds__(1024) void k_loss_sum(const float* __restrict__ loss, int n, float* __restrict__ acc)
{
    __shared__ float s_part[1024];
    float t = 0.f;
    for (int i = threadIdx.x; i < n; i += 1024) t += loss[i];
    s_part[threadIdx.x] = t;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s_part[threadIdx.x] += s_part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) acc[0] += s_part[0];
}

/* one wavefront per row of C logits: loss and bf16 gradient (softmax - onehot) / n */
__global__ __launch_bounds__(256) void k_softmax_ce(const float* __restrict__ Z, const int* __restrict__ labels,
                                                    int n, int C, uint16_t* __restrict__ dZ, float* __restrict__ loss)
{
    const int r = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (r >= n) return;
    const float* z = Z + (long long)r * C;
    float m = -INFINITY;
    for (int c = lane; c < C; c += 64) m = fmaxf(m, z[c]);
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    float s = 0.f;
    for (int c = lane; c < C; c += 64) s += __expf(z[c] - m);
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o, 64);
    const int y = labels[r];
    const float inv = 1.f / s, scale = 1.f / (float)n;
    for (int c = lane; c < C; c += 64) {
        const float p = __expf(z[c] - m) * inv;
        dZ[(long long)r * C + c] = f2bf((p - (c == y ? 1.f : 0.f)) * scale);
    }
    if (lane == 0 && loss) loss[r] = -(z[y] - m - __logf(s));
}

/* Column sums of a bf16 (n x C) matrix into fp32 (bias gradients).  One
   1024-thread workgroup per 64 columns: thread (g, c) sums rows g, g+16, ...
   of column c (every row read as one 128-byte segment by a wavefront, eight
   rows in flight per thread), then the 16 partial sums of a column are added
   in a fixed order, so the result does not depend on scheduling. */
__global__ __launch_bounds__(1024) void k_colsum(const uint16_t* __restrict__ X, int n, int C, float* __restrict__ out)
{
    __shared__ float s_part[16][64];
    const int col = threadIdx.x & 63, g = threadIdx.x >> 6;
    const int c = blockIdx.x * 64 + col;
    float s = 0.f;
    if (c < C) {
        int r = g;
        for (; r + 7 * 16 < n; r += 8 * 16) {
            float v[8];
#pragma unroll
            for (int k = 0; k < 8; k++) v[k] = bf2f(X[(long long)(r + 16 * k) * C + c]);
#pragma unroll
            for (int k = 0; k < 8; k++) s += v[k];
        }
        for (; r < n; r += 16) s += bf2f(X[(long long)r * C + c]);
    }
    s_part[g][col] = s;
    __syncthreads();
    if (g == 0 && c < C) {
        float t = 0.f;
#pragma unroll
        for (int k = 0; k < 16; k++) t += s_part[k][col];
        out[c] = t;
    }
}

/* Adam on fp32 master weights; bf16 copies refreshed */
__global__ void k_adam(float* __restrict__ w, uint16_t* __restrict__ w16, const float* __restrict__ g,
                       float* __restrict__ m, float* __restrict__ v, long long n, float lr, float b1, float b2,
                       float eps, float c1, float c2)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
    w[i] = wi;
    if (w16) w16[i] = f2bf(wi);
}

/* fp32 -> bf16 copy */
__global__ void k_to_bf16(const float* __restrict__ x, uint16_t* __restrict__ y, long long n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = f2bf(x[i]);
}

/* candidates one wavefront of the first probe pass holds in LDS (2 KB) */
constexpr int kProbeCap = 256;

/* Max softmax probability (*pmax_r) and the top-k logits (out[0 .. k)) of
   the FP16 logit row z, one wavefront (ties to the lower bucket id; C <= 1024,
   a multiple of 64), by descending logit with SORTED, else by bucket id; cand
   holds CAP candidates (more are handled from the slots for k <= 64; false,
   nothing written, only for k > 64 with CAP below 1024, never launched).
   With mass below 1
   (k <= 64) only the best-ranked prefix whose softmax mass reaches `mass` is
   probed: the top-k are ranked (logit descending, ties to the lower id), the
   weights exp(z - max) of the ranks before a candidate are summed
   (wave_mass_keep), and the candidate is kept iff that sum is below mass
   times the row's sum of exp(z - max); the row then holds the kept buckets
   best first and -1 in the remaining slots, whatever SORTED says.  ids: 64
   ints of LDS for this wavefront. */
/* order-preserving 16-bit key of FP16 bits (-0 and +0 share one key, above
   0: every value has a key of at least 1) */
__device__ __forceinline__ uint32_t half_order_key(uint32_t h)
{
    return (h & 0x7fffu) == 0u ? 0x8000u : (h & 0x8000u) ? (~h & 0xffffu) : (h | 0x8000u);
}

template <bool SORTED, int CAP>
__device__ __forceinline__ bool probe_row(const uint16_t* __restrict__ z, int C, int k, float mass,
                                          int* __restrict__ out, float* __restrict__ pmax_r, uint2* cand, int* ids)
{
    const int lane = threadIdx.x & 63;
    /* one dword per lane and slot pair t: buckets 128 t + 2 lane and + 1 (C
       is even, so a pair is present whole or not at all) */
    const uint32_t* const z2 = (const uint32_t*)z;
    uint32_t hb[8];
#pragma unroll
    for (int t = 0; t < 8; t++) hb[t] = 128 * t + 2 * lane < C ? z2[lane + 64 * t] : 0u;
    float v[16];
#pragma unroll
    for (int t = 0; t < 8; t++) {
        const bool present = 128 * t + 2 * lane < C;
        v[2 * t] = present ? h2f(hb[t] & 0xffffu) : -INFINITY;
        v[2 * t + 1] = present ? h2f(hb[t] >> 16) : -INFINITY;
    }
    float m = -INFINITY;
#pragma unroll
    for (int t = 0; t < 16; t++) m = fmaxf(m, v[t]);
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    float s = 0.f;
#pragma unroll
    for (int t = 0; t < 16; t++) s += 128 * (t >> 1) + 2 * lane < C ? __expf(v[t] - m) : 0.f;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) *pmax_r = 1.f / s;
    uint32_t key[16];
#pragma unroll
    for (int t = 0; t < 8; t++) {
        const bool present = 128 * t + 2 * lane < C;
        key[2 * t] = present ? half_order_key(hb[t] & 0xffffu) : 0u;
        key[2 * t + 1] = present ? half_order_key(hb[t] >> 16) : 0u;
    }
    if (!(mass < 1.f)) return wave_topk<15, SORTED, CAP, true>(key, k, cand, out);
    /* the top k best first into LDS, then the rule over lane = rank */
    if (!wave_topk<15, true, CAP, true>(key, k, cand, ids)) return false;
    wave_sync_lds();
    const int id = lane < k ? ids[lane] : 0;
    wave_mass_keep(id, __expf(h2f(z[id]) - m), k, mass * s, out);
    wave_sync_lds(); /* ids are reused by the wavefront's next query */
    return true;
}

/* one wavefront per query.  With CAP below 1024 (LDS for CAP candidates per
   wavefront, which leaves room for more resident wavefronts) a query whose
   candidates do not fit is finished from its slots (probe_row). */
template <bool SORTED, int CAP>
__global__ __launch_bounds__(256) void k_probe(const uint16_t* __restrict__ Z, int n, int C, int k, float mass,
                                               int* __restrict__ probes, float* __restrict__ pmax)
{
    __shared__ uint2 s_cand[4][CAP];
    __shared__ int s_ids[4][64];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int r = blockIdx.x * 4 + wv;
    if (r >= n) return;
    if (!probe_row<SORTED, CAP>(Z + (long long)r * C, C, k, mass, probes + (long long)r * k, pmax + r, s_cand[wv],
                                s_ids[wv]) &&
        lane == 0)
        probes[(long long)r * k] = -1;
}

template <bool SORTED>
static void probe_launch(const uint16_t* Z, int n, int C, int k, float mass, int* probes, float* pmax, hipStream_t s)
{
    const dim3 grid((n + 3) / 4);
    if (k <= 64 && C > kProbeCap) {
        /* a query whose candidates overflow the LDS buffer is finished from
           its slots inside the kernel: no second pass */
        hipLaunchKernelGGL((k_probe<SORTED, kProbeCap>), grid, dim3(256), 0, s, Z, n, C, k, mass, probes, pmax);
    } else {
        hipLaunchKernelGGL((k_probe<SORTED, 1024>), grid, dim3(256), 0, s, Z, n, C, k, mass, probes, pmax);
    }
}

/* ---------------------------------------------------------------- launchers */

int gemm_tn_splits(int M, int N, int K)
{
    /* enough parts for about 256 workgroups, each part at least 128 deep */
    const int tiles = ((M + kTT - 1) / kTT) * ((N + kTT - 1) / kTT);
    int sp = (256 + tiles - 1) / tiles;
    sp = min(sp, max(1, K / 128));
    return max(sp, 1);
}

hipError_t launch_gemm(const GemmArgs& g, hipStream_t s)
{
    if (g.M <= 0 || g.N <= 0 || g.K <= 0 || g.M % kGT || g.N % kGT || g.K % kNK) return hipErrorInvalidValue;
    const bool nt = g.sak == 1 && g.sbk == 1 && g.sam % 8 == 0 && g.sbn % 8 == 0;
    if (g.cbits && (!g.qbits || g.addin || g.zmul || g.sak != 1 || g.sbk != 1)) return hipErrorInvalidValue;
    if (g.half && (g.addin || g.zmul || g.gelu || !nt)) return hipErrorInvalidValue;
    if (nt) {
        const int tiles = ((g.M + kNT - 1) / kNT) * ((g.N + kNT - 1) / kNT);
        /* 128-deep steps when both K segments allow them and the operands fit 32-bit buffer offsets */
        const bool deep = g.K % kNK2 == 0 && (long long)g.M * g.sam * 2 < (1ll << 31) &&
                          (long long)g.N * g.sbn * 2 < (1ll << 31);
        const bool big = !g.addin && !g.zmul && g.M >= 4 * kFM && g.N >= kFN && g.K % kFK == 0 &&
                         g.ldc % 8 == 0 && (long long)g.M * g.sam * 2 < (1ll << 31) &&
                         (long long)g.N * g.sbn * 2 < (1ll << 31);
        if (big) {
            const int tiles_big = ((g.M + kFM - 1) / kFM) * ((g.N + kFN - 1) / kFN);
            if (g.half) hipLaunchKernelGGL((k_gemm_fwd<false, true>), dim3(tiles_big), dim3(256), 0, s, g);
            else if (g.gelu) hipLaunchKernelGGL((k_gemm_fwd<true, false>), dim3(tiles_big), dim3(256), 0, s, g);
            else hipLaunchKernelGGL((k_gemm_fwd<false, false>), dim3(tiles_big), dim3(256), 0, s, g);
        } else if (deep) hipLaunchKernelGGL(k_gemm_nt128, dim3(tiles), dim3(256), 0, s, g);
        else hipLaunchKernelGGL(k_gemm_nt, dim3(tiles), dim3(256), 0, s, g);
    } else if (g.ws && g.sam == 1 && g.sbn == 1 && g.splits >= 1 && g.out32 && !g.out16 && !g.addin && !g.bias &&
               !g.zmul && !g.gelu && g.ldc % 4 == 0 && g.N % 4 == 0 && (long long)g.K * g.sak * 2 < (1ll << 30) &&
               (long long)g.K * g.sbk * 2 < (1ll << 30)) {
        const int tiles = ((g.M + kTT - 1) / kTT) * ((g.N + kTT - 1) / kTT);
        const int kper = ((g.K + g.splits - 1) / g.splits + kTK - 1) / kTK * kTK;
        const int parts = (g.K + kper - 1) / kper;
        hipLaunchKernelGGL(k_gemm_tn, dim3(tiles * parts), dim3(256), 0, s, g, kper);
        const long long n4 = (long long)g.M * g.N / 4;
        hipLaunchKernelGGL(k_splitk_sum, dim3((unsigned)((n4 + 255) / 256)), dim3(256), 0, s, (const float*)g.ws, parts,
                           g.M, g.N, g.ldc, g.out32);
    } else {
        hipLaunchKernelGGL(k_gemm, dim3((g.M / kGT) * (g.N / kGT)), dim3(256), 0, s, g);
    }
    return hipGetLastError();
}

hipError_t launch_transpose16(const uint16_t* src, int R, int C, uint16_t* dst, hipStream_t s)
{
    if (R <= 0 || C <= 0 || R % 64 || C % 64) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_transpose16, dim3((R / 64) * (C / 64)), dim3(256), 0, s, src, R, C, dst);
    return hipGetLastError();
}

hipError_t launch_bits_pm1(const uint8_t* desc, int n, uint16_t* X, hipStream_t s)
{
    const long long tot = (long long)n * 32;
    if (tot > 0)
        hipLaunchKernelGGL(k_bits_pm1, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, desc, n, X);
    return hipGetLastError();
}

hipError_t launch_softmax_ce(const float* Z, const int* labels, int n, int C, uint16_t* dZ, float* loss, hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL(k_softmax_ce, dim3((n + 3) / 4), dim3(256), 0, s, Z, labels, n, C, dZ, loss);
    return hipGetLastError();
}

hipError_t launch_loss_sum(const float* loss, int n, float* acc, hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL(k_loss_sum, dim3(1), dim3(1024), 0, s, loss, n, acc);
    return hipGetLastError();
}

hipError_t launch_colsum(const uint16_t* X, int n, int C, float* out, hipStream_t s)
{
    hipLaunchKernelGGL(k_colsum, dim3((C + 63) / 64), dim3(1024), 0, s, X, n, C, out);
    return hipGetLastError();
}

hipError_t launch_adam(float* w, uint16_t* w16, const float* g, float* m, float* v, long long n, float lr, float b1,
                       float b2, float eps, int step, hipStream_t s)
{
    const float c1 = 1.f / (1.f - powf(b1, (float)step)), c2 = 1.f / (1.f - powf(b2, (float)step));
    hipLaunchKernelGGL(k_adam, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, w, w16, g, m, v, n, lr, b1, b2,
                       eps, c1, c2);
    return hipGetLastError();
}

hipError_t launch_to_bf16(const float* x, uint16_t* y, long long n, hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, y, n);
    return hipGetLastError();
}

hipError_t launch_probe(const uint16_t* Z, int n, int C, int k, float mass, int* probes, float* pmax, bool sorted,
                        hipStream_t s)
{
    if (C <= 0 || C > 1024 || C % 64 || k < 1 || k > C || (mass < 1.f && (k > 64 || !(mass > 0.f))))
        return hipErrorInvalidValue;
    if (n > 0) {
        if (sorted) probe_launch<true>(Z, n, C, k, mass, probes, pmax, s);
        else probe_launch<false>(Z, n, C, k, mass, probes, pmax, s);
    }
    return hipGetLastError();
}

} // namespace orbk
