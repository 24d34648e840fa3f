/*
 * Learned bucket matcher kernels (gfx950): a BF16 MLP 256 -> H -> H -> C
 * (GELU, FP32 accumulate) that maps a descriptor to database buckets, its
 * training step, and the bucketed exact search.  No counterpart in the
 * reference; exact search (orb_match.hip) stays the fallback and the judge
 * of recall.
 *
 *   k_gemm      C = A B on v_mfma_f32_16x16x32_bf16, 64 x 64 workgroup tiles,
 *               A and B addressed through strides (so A^T B and A B^T need no
 *               transposed copies), epilogue fused: + bias, GELU, or the
 *               GELU-derivative product of the backward pass, FP32 and/or
 *               BF16 stores
 *   k_bits_pm1  descriptor bits -> +/-1 BF16 inputs
 *   k_softmax_ce  row softmax, cross-entropy and its gradient (scaled 1/n)
 *   k_colsum    bias gradients
 *   k_adam      Adam on FP32 master weights, BF16 copies refreshed
 *   k_probe     softmax max and the top-k buckets of each query
 */
#include "orb_internal.h"
#include "orb_mlp.h"

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

__device__ __forceinline__ float gelu(float z)
{
    return z * 0.5f * (1.0f + erff(z * 0.70710678118654752f));
}

__device__ __forceinline__ float gelu_grad(float z)
{
    return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * __expf(-0.5f * z * z);
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

/* ------------------------------------------------------------ elementwise */

/* descriptor bits -> +/-1 bf16 (bit k of the descriptor -> column k) */
__global__ void k_bits_pm1(const uint8_t* __restrict__ desc, int n, uint16_t pos, uint16_t neg,
                           uint16_t* __restrict__ X)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)n * 256) return;
    const long long r = i >> 8;
    const int k = (int)(i & 255);
    const int bit = (desc[r * 32 + (k >> 3)] >> (k & 7)) & 1;
    X[i] = bit ? pos : neg;
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

/* column sums of a bf16 (n x C) matrix into fp32 (bias gradients) */
__global__ void k_colsum(const uint16_t* __restrict__ X, int n, int C, float* __restrict__ out)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    float s = 0.f;
    for (int r = 0; r < n; r++) s += bf2f(X[(long long)r * C + c]);
    out[c] = s;
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

/* one wavefront per query: max softmax probability and the top-k logits
   (ties to the lower bucket id) */
__global__ __launch_bounds__(256) void k_probe(const float* __restrict__ Z, int n, int C, int k,
                                               int* __restrict__ probes, float* __restrict__ pmax)
{
    const int r = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (r >= n) return;
    const float* z = Z + (long long)r * C;
    constexpr int kPer = 16; /* C <= 1024 */
    float v[kPer];
#pragma unroll
    for (int t = 0; t < kPer; t++) {
        const int c = lane + 64 * t;
        v[t] = c < C ? z[c] : -INFINITY;
    }
    float m = -INFINITY;
#pragma unroll
    for (int t = 0; t < kPer; t++) m = fmaxf(m, v[t]);
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    float s = 0.f;
#pragma unroll
    for (int t = 0; t < kPer; t++) s += (lane + 64 * t) < C ? __expf(v[t] - m) : 0.f;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) pmax[r] = 1.f / s;
    for (int j = 0; j < k; j++) {
        /* argmax with the lowest index on ties */
        float bv = -INFINITY;
        int bi = 0x7fffffff;
#pragma unroll
        for (int t = 0; t < kPer; t++) {
            const int c = lane + 64 * t;
            if (c < C && (v[t] > bv || (v[t] == bv && c < bi))) { bv = v[t]; bi = c; }
        }
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) {
            const float ov = __shfl_xor(bv, o, 64);
            const int oi = __shfl_xor(bi, o, 64);
            if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
        }
        if (lane == 0) probes[(long long)r * k + j] = bi;
        if ((bi & 63) == lane) v[bi >> 6] = -INFINITY;
    }
}

/* ---------------------------------------------------------------- launchers */

hipError_t launch_gemm(const GemmArgs& g, hipStream_t s)
{
    hipLaunchKernelGGL(k_gemm, dim3((g.M / kGT) * (g.N / kGT)), dim3(256), 0, s, g);
    return hipGetLastError();
}

hipError_t launch_bits_pm(const uint8_t* desc, int n, uint16_t pos, uint16_t neg, uint16_t* X, hipStream_t s)
{
    const long long tot = (long long)n * 256;
    if (tot > 0)
        hipLaunchKernelGGL(k_bits_pm1, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, desc, n, pos, neg, X);
    return hipGetLastError();
}

hipError_t launch_bits_pm1(const uint8_t* desc, int n, uint16_t* X, hipStream_t s)
{
    return launch_bits_pm(desc, n, (uint16_t)0x3F80, (uint16_t)0xBF80, X, s);
}

hipError_t launch_softmax_ce(const float* Z, const int* labels, int n, int C, uint16_t* dZ, float* loss, hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL(k_softmax_ce, dim3((n + 3) / 4), dim3(256), 0, s, Z, labels, n, C, dZ, loss);
    return hipGetLastError();
}

hipError_t launch_colsum(const uint16_t* X, int n, int C, float* out, hipStream_t s)
{
    hipLaunchKernelGGL(k_colsum, dim3((C + 255) / 256), dim3(256), 0, s, X, n, C, out);
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

hipError_t launch_probe(const float* Z, int n, int C, int k, int* probes, float* pmax, hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL(k_probe, dim3((n + 3) / 4), dim3(256), 0, s, Z, n, C, k, probes, pmax);
    return hipGetLastError();
}

} // namespace orbk
