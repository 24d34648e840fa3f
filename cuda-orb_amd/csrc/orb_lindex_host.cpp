/*
 * Host side of the learned bucket matcher (include/orb_capi.h, orb_lindex_*):
 * k-modes partition of a descriptor database into buckets, training of the
 * BF16 MLP that maps a query descriptor to the bucket of its exact nearest
 * neighbour, and the probe-then-search query path with the exact matcher as
 * fallback.  No counterpart in the reference (it has only the brute-force
 * matcher, orbd.cu:1471-1559).
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "orb_capi.h"
#include "orb_internal.h"
#include "orb_mlp.h"

using namespace orbk;

/* internal helpers of orb_capi.cpp */
int orb__ctx_stream(orb_ctx* ctx, void** stream);
int orb__ctx_fail(orb_ctx* ctx, int code, const char* msg);
int orb__match_dcount_top2(orb_ctx* ctx, const uint8_t* q, int nq_cap, const int* d_nq, const uint8_t* db, int ndb,
                           int* idx1, int* d1, int* idx2, int* d2);

namespace
{

constexpr int kIn = 256;

struct Rng {
    uint64_t s;
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

template <typename T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    ~DBuf() { reset(); }
    void reset()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    bool ensure(size_t count)
    {
        if (count <= n) return true;
        reset();
        if (hipMalloc((void**)&p, sizeof(T) * (count ? count : 1)) != hipSuccess) {
            p = nullptr;
            return false;
        }
        n = count;
        return true;
    }
};

} // namespace

struct orb_lindex {
    orb_ctx* ctx = nullptr;
    orb_lindex_params prm{};
    int H = 0, C = 0;
    int ncu = 256; /* compute units of the device (sizes the bucket-search grid) */
    /* parameters: offsets of W1, b1, W2, b2, W3, b3 in one fp32 vector */
    size_t off[7] = {};
    DBuf<float> w, g, m, v;
    DBuf<uint16_t> w16; /* bf16 copies of W1, W2, W3 (same offsets minus biases is awkward: full copy) */
    DBuf<uint16_t> w16t; /* W1^T (H x 256), W2^T (H x H), W3^T (C x H): K-contiguous forward operands */
    bool wt_fresh = false; /* w16t matches w16 */
    int step = 0;
    /* probability-mass rule of the probes (1: every one of the nprobe buckets)
       and the distance temperature of its centroid-mode weights */
    float mass = 1.f, ctemp = 4.f;
    /* batch workspace */
    int cap_batch = 0;
    DBuf<uint16_t> X, A1, A2, dZ1, dZ2, dZ3;
    DBuf<float> Z1, Z2, Z3, loss, eloss;
    DBuf<uint16_t> Z3h; /* FP16 logits of the query path */
    DBuf<float> ws; /* split-K partial sums of the weight gradients */
    DBuf<int> labels, order;
    /* database */
    int ndb = 0;
    DBuf<uint8_t> db, db_sorted, centroids;
    uint16_t cscale16 = 0;   /* bf16 bits of centroid_scale (0: no centroid term) */
    DBuf<int> assign, hist, offsets, chunk_hist, perm;
    DBuf<int> border; /* bucket ids by decreasing size (ties by id): the search's unit order */
    /* query workspace */
    DBuf<int> probes, list_count, list_offsets, list_cursor, lists, unit_offsets, unit_pos, fb_count, fb_index, fb_qidx;
    DBuf<int> ex_idx1, ex_d1, ex_idx2, ex_d2, tmp_idx;
    DBuf<float> pmax;
    DBuf<uint8_t> fb_q;
    DBuf<uint2> partial;
    DBuf<uint32_t> qbound; /* per query: the smallest second distance found so far (bucket search) */
    /* list_count is zero and qbound[0 .. clean_rows) holds "no bound": the
       state each query batch's merge leaves for the next one (0: unknown) */
    int clean_rows = 0;
};

namespace
{

hipStream_t stream_of(orb_lindex* L)
{
    void* s = nullptr;
    orb__ctx_stream(L->ctx, &s);
    return (hipStream_t)s;
}

int lfail(orb_lindex* L, int code, const char* msg) { return orb__ctx_fail(L ? L->ctx : nullptr, code, msg); }

#define LTRY(L, call)                                                                   \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) {                                                         \
            std::string m_ = std::string(#call) + ": " + hipGetErrorString(e_);         \
            return lfail(L, e_ == hipErrorOutOfMemory ? ORB_E_NOMEM : ORB_E_HIP, m_.c_str()); \
        }                                                                               \
    } while (0)

#define LALLOC(L, buf, n) \
    do { if (!(buf).ensure(n)) return lfail(L, ORB_E_NOMEM, "device allocation failed"); } while (0)

size_t nparams(const orb_lindex* L) { return L->off[6]; }

/* float -> bf16 bits, round to nearest even */
uint16_t bf16_bits(float x)
{
    uint32_t u;
    memcpy(&u, &x, 4);
    return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

/* the float value of bf16 bits */
float bf16_value(uint16_t h)
{
    const uint32_t u = (uint32_t)h << 16;
    float x;
    memcpy(&x, &u, 4);
    return x;
}

int ensure_batch(orb_lindex* L, int B)
{
    if (B <= L->cap_batch) return ORB_OK;
    const size_t b = (size_t)B;
    LALLOC(L, L->X, b * kIn);
    LALLOC(L, L->A1, b * L->H);
    LALLOC(L, L->A2, b * L->H);
    LALLOC(L, L->dZ1, b * L->H);
    LALLOC(L, L->dZ2, b * L->H);
    LALLOC(L, L->dZ3, b * L->C);
    LALLOC(L, L->Z1, b * L->H);
    LALLOC(L, L->Z2, b * L->H);
    LALLOC(L, L->Z3, b * L->C);
    LALLOC(L, L->loss, b);
    LALLOC(L, L->labels, b);
    L->cap_batch = B;
    return ORB_OK;
}

GemmArgs gemm(const uint16_t* A, long long sam, long long sak, const uint16_t* B, long long sbk, long long sbn, int M,
              int N, int K)
{
    GemmArgs g;
    memset(&g, 0, sizeof(g));
    g.A = A;
    g.sam = sam;
    g.sak = sak;
    g.B = B;
    g.sbk = sbk;
    g.sbn = sbn;
    g.M = M;
    g.N = N;
    g.K = K;
    g.ldc = N;
    return g;
}

/* the K-contiguous weight copies the forward products read, refreshed after
   the weights change */
int ensure_wt(orb_lindex* L, hipStream_t s)
{
    if (L->wt_fresh) return ORB_OK;
    const int H = L->H, C = L->C;
    const uint16_t* w16 = L->w16.p;
    uint16_t* wt = L->w16t.p;
    const size_t t1 = (size_t)kIn * H, t2 = t1 + (size_t)H * H;
    LTRY(L, launch_transpose16(w16 + L->off[0], kIn, H, wt, s));
    LTRY(L, launch_transpose16(w16 + L->off[2], H, H, wt + t1, s));
    LTRY(L, launch_transpose16(w16 + L->off[4], H, C, wt + t2, s));
    L->wt_fresh = true;
    return ORB_OK;
}

/* Network forward of B rows of X (the +/-1 expansion of the descriptors
   qbits) into Z3 (logits); keep_z
   also stores the pre-activations Z1, Z2 the backward pass needs.  The
   centroid term is added to the last layer's products in their epilogue.
   Without keep_z (queries) the logits are rounded to FP16: into z3h as FP16
   bits if given, else as floats into z3 (L->Z3 if null);
   with keep_z (training) they stay FP32. */
int forward(orb_lindex* L, int B, hipStream_t s, bool keep_z, const uint8_t* qbits, float* z3 = nullptr,
            uint16_t* z3h = nullptr)
{
    const int H = L->H, C = L->C;
    const float* w = L->w.p;
    const uint16_t* wt = L->w16t.p;
    const size_t t1 = (size_t)kIn * H, t2 = t1 + (size_t)H * H;
    int rc = ensure_wt(L, s);
    if (rc != ORB_OK) return rc;
    uint16_t* const a1 = L->A1.p;
    uint16_t* const a2 = L->A2.p;
    GemmArgs g = gemm(L->X.p, kIn, 1, wt, 1, kIn, B, H, kIn);
    g.bias = w + L->off[1];
    g.gelu = 1;
    g.out32 = keep_z ? L->Z1.p : nullptr;
    g.out16 = a1;
    LTRY(L, launch_gemm(g, s));
    g = gemm(a1, H, 1, wt + t1, 1, H, B, H, H);
    g.bias = w + L->off[3];
    g.gelu = 1;
    g.out32 = keep_z ? L->Z2.p : nullptr;
    g.out16 = a2;
    LTRY(L, launch_gemm(g, s));
    g = gemm(a2, H, 1, wt + t2, 1, H, B, C, H);
    if (L->cscale16) {
        /* centroid term centroid_scale * <x, c_b> = centroid_scale * (256 - 2 Hamming) */
        if (!L->centroids.p) return lfail(L, ORB_E_STATE, "orb_lindex_build has not been called");
        g.qbits = qbits;
        g.cbits = L->centroids.p;
        g.cscale = bf16_value(L->cscale16);
    }
    g.bias = w + L->off[5];
    g.half = keep_z ? 0 : 1;
    if (z3h) g.out16 = z3h;
    else g.out32 = z3 ? z3 : L->Z3.p;
    LTRY(L, launch_gemm(g, s));
    return ORB_OK;
}


/* weight gradient out = A^T dZ (A: B x M, dZ: B x N, row-major), the batch
   split over the split-K workspace */
int wgrad(orb_lindex* L, const uint16_t* A, int M, const uint16_t* dZ, int N, int B, float* out, hipStream_t s)
{
    GemmArgs g = gemm(A, 1, M, dZ, N, 1, M, N, B);
    g.out32 = out;
    g.splits = gemm_tn_splits(M, N, B);
    LALLOC(L, L->ws, (size_t)g.splits * M * N);
    g.ws = L->ws.p;
    LTRY(L, launch_gemm(g, s));
    return ORB_OK;
}

/* backward into L->g after forward and labels; loss per row in L->loss */
int backward(orb_lindex* L, int B, hipStream_t s)
{
    const int H = L->H, C = L->C;
    const uint16_t* w16 = L->w16.p;
    float* gr = L->g.p;
    LTRY(L, launch_softmax_ce(L->Z3.p, L->labels.p, B, C, L->dZ3.p, L->loss.p, s));
    int rc = wgrad(L, L->A2.p, H, L->dZ3.p, C, B, gr + L->off[4], s); /* dW3 = A2^T dZ3 */
    if (rc != ORB_OK) return rc;
    LTRY(L, launch_colsum(L->dZ3.p, B, C, gr + L->off[5], s));
    GemmArgs g = gemm(L->dZ3.p, C, 1, w16 + L->off[4], 1, C, B, H, C); /* dZ2 = (dZ3 W3^T) gelu'(Z2) */
    g.zmul = L->Z2.p;
    g.out16 = L->dZ2.p;
    LTRY(L, launch_gemm(g, s));
    rc = wgrad(L, L->A1.p, H, L->dZ2.p, H, B, gr + L->off[2], s); /* dW2 = A1^T dZ2 */
    if (rc != ORB_OK) return rc;
    LTRY(L, launch_colsum(L->dZ2.p, B, H, gr + L->off[3], s));
    g = gemm(L->dZ2.p, H, 1, w16 + L->off[2], 1, H, B, H, H); /* dZ1 = (dZ2 W2^T) gelu'(Z1) */
    g.zmul = L->Z1.p;
    g.out16 = L->dZ1.p;
    LTRY(L, launch_gemm(g, s));
    rc = wgrad(L, L->X.p, kIn, L->dZ1.p, H, B, gr + L->off[0], s); /* dW1 = X^T dZ1 */
    if (rc != ORB_OK) return rc;
    LTRY(L, launch_colsum(L->dZ1.p, B, H, gr + L->off[1], s));
    return ORB_OK;
}

/* the nprobe buckets of each of nq queries (best first when sorted, else in
   bucket-id order) and their max softmax probability (1 without the network) */
int compute_probes(orb_lindex* L, const uint8_t* d_q, int nq, int mode, int* probes, float* pmax, bool sorted,
                   hipStream_t s)
{
    const int np = L->prm.nprobe, nb = L->C;
    if (mode == ORB_LINDEX_MLP) {
        /* network forward in chunks (FP16 logits), top-nprobe buckets and max softmax */
        const int chunk = 65536;
        int rc = ensure_batch(L, nq < chunk ? nq : chunk);
        if (rc != ORB_OK) return rc;
        LALLOC(L, L->Z3h, (size_t)(nq < chunk ? nq : chunk) * nb);
        for (int c0 = 0; c0 < nq; c0 += chunk) {
            const int B = nq - c0 < chunk ? nq - c0 : chunk;
            const uint8_t* qb = d_q + (size_t)c0 * 32;
            LTRY(L, launch_bits_pm1(qb, B, L->X.p, s));
            rc = forward(L, B, s, false, qb, nullptr, L->Z3h.p);
            if (rc != ORB_OK) return rc;
            LTRY(L, launch_probe(L->Z3h.p, B, nb, np, L->mass, probes + (size_t)c0 * np, pmax + c0, sorted, s));
        }
    } else {
        LTRY(L, launch_centroid_probe(d_q, nq, L->centroids.p, nb, np, L->mass, L->ctemp, probes, pmax, sorted, s));
    }
    return ORB_OK;
}

} // namespace

extern "C" {

int orb_lindex_create(orb_ctx* ctx, const orb_lindex_params* p, orb_lindex** out)
{
    if (!ctx || !p || !out) return orb__ctx_fail(ctx, ORB_E_ARG, "null argument");
    *out = nullptr;
    if (p->nbuckets < 64 || p->nbuckets > 1024 || p->nbuckets % 64)
        return orb__ctx_fail(ctx, ORB_E_ARG, "nbuckets must be a multiple of 64 in [64, 1024]");
    if (p->hidden < 64 || p->hidden > 4096 || p->hidden % 64)
        return orb__ctx_fail(ctx, ORB_E_ARG, "hidden must be a multiple of 64 in [64, 4096]");
    if (p->nprobe < 1 || p->nprobe > p->nbuckets) return orb__ctx_fail(ctx, ORB_E_ARG, "nprobe out of range");
    orb_lindex* L = new orb_lindex();
    L->ctx = ctx;
    {
        int dev = 0, ncu = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && ncu > 0)
            L->ncu = ncu;
    }
    L->prm = *p;
    L->H = p->hidden;
    L->C = p->nbuckets;
    const size_t sizes[6] = {(size_t)kIn * L->H, (size_t)L->H, (size_t)L->H * L->H, (size_t)L->H,
                             (size_t)L->H * L->C, (size_t)L->C};
    L->off[0] = 0;
    for (int i = 0; i < 6; i++) L->off[i + 1] = L->off[i] + sizes[i];
    const size_t np = nparams(L);
    if (!L->w.ensure(np) || !L->g.ensure(np) || !L->m.ensure(np) || !L->v.ensure(np) || !L->w16.ensure(np) ||
        !L->w16t.ensure((size_t)kIn * L->H + (size_t)L->H * L->H + (size_t)L->H * L->C)) {
        delete L;
        return orb__ctx_fail(ctx, ORB_E_NOMEM, "device allocation failed");
    }
    /* uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights, zero biases */
    std::vector<float> h(np, 0.f);
    Rng rng{p->seed * 0x2545F4914F6CDD1Dull + 11};
    const int fan_in[3] = {kIn, L->H, L->H};
    L->cscale16 = p->centroid_scale != 0.f ? bf16_bits(p->centroid_scale) : 0;
    for (int t = 0; t < 3; t++) {
        const double a = 1.0 / std::sqrt((double)fan_in[t]);
        for (size_t i = L->off[2 * t]; i < L->off[2 * t + 1]; i++) h[i] = (float)((2.0 * rng.uniform() - 1.0) * a);
    }
    /* with the centroid term the last layer starts at zero: the untrained
       network ranks buckets as nearest-centroid probing does */
    if (L->cscale16)
        for (size_t i = L->off[4]; i < L->off[5]; i++) h[i] = 0.f;
    int rc = orb_lindex_set_weights(L, h.data());
    if (rc != ORB_OK) {
        delete L;
        return rc;
    }
    *out = L;
    return ORB_OK;
}

int orb_lindex_destroy(orb_lindex* L)
{
    if (!L) return ORB_OK;
    (void)hipStreamSynchronize(stream_of(L));
    delete L;
    return ORB_OK;
}

int orb_lindex_num_params(const orb_lindex* L)
{
    return L ? (int)nparams(L) : 0;
}

int orb_lindex_nprobe(const orb_lindex* L)
{
    return L ? L->prm.nprobe : 0;
}

int orb_lindex_set_weights(orb_lindex* L, const float* h_w)
{
    if (!L || !h_w) return lfail(L, ORB_E_ARG, "null argument");
    hipStream_t s = stream_of(L);
    const size_t np = nparams(L);
    LTRY(L, hipMemcpyAsync(L->w.p, h_w, sizeof(float) * np, hipMemcpyHostToDevice, s));
    LTRY(L, hipMemsetAsync(L->m.p, 0, sizeof(float) * np, s));
    LTRY(L, hipMemsetAsync(L->v.p, 0, sizeof(float) * np, s));
    LTRY(L, launch_to_bf16(L->w.p, L->w16.p, (long long)np, s));
    L->wt_fresh = false;
    LTRY(L, hipStreamSynchronize(s));
    L->step = 0;
    return ORB_OK;
}

int orb_lindex_get_weights(orb_lindex* L, float* h_w)
{
    if (!L || !h_w) return lfail(L, ORB_E_ARG, "null argument");
    hipStream_t s = stream_of(L);
    LTRY(L, hipMemcpyAsync(h_w, L->w.p, sizeof(float) * nparams(L), hipMemcpyDeviceToHost, s));
    LTRY(L, hipStreamSynchronize(s));
    return ORB_OK;
}

int orb_lindex_build(orb_lindex* L, const uint8_t* d_db, int ndb, int iters)
{
    if (!L || !d_db || ndb <= 0) return lfail(L, ORB_E_ARG, "bad database");
    if (ndb >= (1 << 23)) return lfail(L, ORB_E_CAPACITY, "database above 2^23-1 entries");
    if (iters < 0) return lfail(L, ORB_E_ARG, "iters must be >= 0");
    hipStream_t s = stream_of(L);
    const int nb = L->C;
    LALLOC(L, L->db, (size_t)ndb * 32);
    LALLOC(L, L->db_sorted, (size_t)ndb * 32);
    LALLOC(L, L->assign, ndb);
    LALLOC(L, L->perm, ndb);
    LALLOC(L, L->hist, nb);
    LALLOC(L, L->offsets, nb + 1);
    LALLOC(L, L->chunk_hist, (size_t)nb * ((ndb + 255) / 256));
    const bool have_centroids = L->centroids.p != nullptr && L->ndb == -1;
    LALLOC(L, L->centroids, (size_t)nb * 32);
    LTRY(L, hipMemcpyAsync(L->db.p, d_db, (size_t)ndb * 32, hipMemcpyDeviceToDevice, s));
    if (!have_centroids) {
        /* initial centroids: nb distinct database rows, evenly spaced with a seeded offset */
        Rng rng{L->prm.seed * 0x9E3779B97F4A7C15ull + 5};
        std::vector<int> pick(nb);
        const double step = (double)ndb / nb;
        for (int b = 0; b < nb; b++) pick[b] = (int)((b + rng.uniform()) * step) % ndb;
        for (int b = 0; b < nb; b++)
            LTRY(L, hipMemcpyAsync(L->centroids.p + (size_t)b * 32, d_db + (size_t)pick[b] * 32, 32,
                                   hipMemcpyDeviceToDevice, s));
    }
    L->ndb = ndb;
    /* Lloyd iterations in Hamming space: assign with the exact matcher, then
       bucket-wise bit majority; a final assignment builds the bucketed copy */
    for (int it = 0; it <= iters; it++) {
        int rc = orb_match_top2(L->ctx, L->db.p, ndb, L->centroids.p, nb, ORB_MATCH_MFMA, L->assign.p, nullptr,
                                nullptr, nullptr);
        if (rc != ORB_OK) return rc;
        LTRY(L, launch_bucket_sort(L->assign.p, ndb, nb, L->db.p, L->chunk_hist.p, L->hist.p, L->offsets.p,
                                   L->perm.p, L->db_sorted.p, s));
        if (it < iters) LTRY(L, launch_kmodes_update(L->db_sorted.p, L->offsets.p, nb, L->centroids.p, s));
    }
    {
        std::vector<int> off(nb + 1), ord(nb);
        LTRY(L, hipMemcpyAsync(off.data(), L->offsets.p, sizeof(int) * (nb + 1), hipMemcpyDeviceToHost, s));
        LTRY(L, hipStreamSynchronize(s));
        for (int b = 0; b < nb; b++) ord[b] = b;
        std::stable_sort(ord.begin(), ord.end(),
                         [&](int a, int b) { return off[a + 1] - off[a] > off[b + 1] - off[b]; });
        LALLOC(L, L->border, nb);
        LTRY(L, hipMemcpyAsync(L->border.p, ord.data(), sizeof(int) * nb, hipMemcpyHostToDevice, s));
    }
    LTRY(L, hipStreamSynchronize(s));
    return ORB_OK;
}

int orb_lindex_bucket_info(orb_lindex* L, int* h_offsets, uint8_t* h_centroids, int* d_bucket_of)
{
    if (!L || L->ndb <= 0) return lfail(L, ORB_E_STATE, "orb_lindex_build has not been called");
    hipStream_t s = stream_of(L);
    if (h_offsets)
        LTRY(L, hipMemcpyAsync(h_offsets, L->offsets.p, sizeof(int) * (L->C + 1), hipMemcpyDeviceToHost, s));
    if (h_centroids)
        LTRY(L, hipMemcpyAsync(h_centroids, L->centroids.p, (size_t)L->C * 32, hipMemcpyDeviceToHost, s));
    if (d_bucket_of)
        LTRY(L, hipMemcpyAsync(d_bucket_of, L->assign.p, sizeof(int) * L->ndb, hipMemcpyDeviceToDevice, s));
    LTRY(L, hipStreamSynchronize(s));
    return ORB_OK;
}

int orb_lindex_grads(orb_lindex* L, const uint8_t* d_q, const int* d_labels, int n, float* h_grads, float* h_loss)
{
    if (!L || !d_q || !d_labels || n <= 0 || n % 64) return lfail(L, ORB_E_ARG, "n must be a positive multiple of 64");
    hipStream_t s = stream_of(L);
    int rc = ensure_batch(L, n);
    if (rc != ORB_OK) return rc;
    LALLOC(L, L->fb_q, (size_t)n * 32);
    LTRY(L, launch_gather(d_q, nullptr, d_labels, n, L->fb_q.p, L->labels.p, s));
    LTRY(L, launch_bits_pm1(L->fb_q.p, n, L->X.p, s));
    rc = forward(L, n, s, true, L->fb_q.p);
    if (rc != ORB_OK) return rc;
    rc = backward(L, n, s);
    if (rc != ORB_OK) return rc;
    if (h_grads) LTRY(L, hipMemcpyAsync(h_grads, L->g.p, sizeof(float) * nparams(L), hipMemcpyDeviceToHost, s));
    std::vector<float> lv(n);
    LTRY(L, hipMemcpyAsync(lv.data(), L->loss.p, sizeof(float) * n, hipMemcpyDeviceToHost, s));
    LTRY(L, hipStreamSynchronize(s));
    if (h_loss) {
        double t = 0;
        for (float x : lv) t += x;
        *h_loss = (float)(t / n);
    }
    return ORB_OK;
}

int orb_lindex_train(orb_lindex* L, const uint8_t* d_q, const int* d_target, int n, int epochs, int batch, float lr,
                     float* h_epoch_loss)
{
    if (!L || !d_q || !d_target || n < batch || batch <= 0 || batch % 64 || epochs < 1)
        return lfail(L, ORB_E_ARG, "bad training arguments (batch a multiple of 64, n >= batch)");
    if (L->ndb <= 0) return lfail(L, ORB_E_STATE, "orb_lindex_build has not been called");
    hipStream_t s = stream_of(L);
    int rc = ensure_batch(L, batch);
    if (rc != ORB_OK) return rc;
    LALLOC(L, L->order, n);
    LALLOC(L, L->fb_q, (size_t)batch * 32);
    LALLOC(L, L->tmp_idx, n);
    /* labels: bucket of each target database entry */
    {
        std::vector<int> tgt(n), asg(L->ndb);
        LTRY(L, hipMemcpyAsync(tgt.data(), d_target, sizeof(int) * n, hipMemcpyDeviceToHost, s));
        LTRY(L, hipMemcpyAsync(asg.data(), L->assign.p, sizeof(int) * L->ndb, hipMemcpyDeviceToHost, s));
        LTRY(L, hipStreamSynchronize(s));
        for (int i = 0; i < n; i++) {
            if (tgt[i] < 0 || tgt[i] >= L->ndb) return lfail(L, ORB_E_ARG, "target index outside the database");
            tgt[i] = asg[tgt[i]];
        }
        LTRY(L, hipMemcpyAsync(L->tmp_idx.p, tgt.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
        LTRY(L, hipStreamSynchronize(s));
    }
    Rng rng{L->prm.seed * 0xD1B54A32D192ED03ull + 3};
    std::vector<int> ord(n);
    const int nbatch = n / batch;
    /* cosine decay of the step size from lr to lr / 20 over the whole run */
    const double total_steps = (double)epochs * nbatch;
    long long t_step = 0;
    /* per epoch: the sum of every training row's loss, accumulated on the device */
    LALLOC(L, L->eloss, epochs);
    LTRY(L, hipMemsetAsync(L->eloss.p, 0, sizeof(float) * epochs, s));
    for (int e = 0; e < epochs; e++) {
        for (int i = 0; i < n; i++) ord[i] = i;
        for (int i = n - 1; i > 0; i--) {
            const int j = (int)(rng.next() % (uint64_t)(i + 1));
            const int t = ord[i];
            ord[i] = ord[j];
            ord[j] = t;
        }
        LTRY(L, hipMemcpyAsync(L->order.p, ord.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
        for (int bi = 0; bi < nbatch; bi++) {
            LTRY(L, launch_gather(d_q, L->order.p + (size_t)bi * batch, L->tmp_idx.p, batch, L->fb_q.p,
                                  L->labels.p, s));
            LTRY(L, launch_bits_pm1(L->fb_q.p, batch, L->X.p, s));
            rc = forward(L, batch, s, true, L->fb_q.p);
            if (rc != ORB_OK) return rc;
            rc = backward(L, batch, s);
            if (rc != ORB_OK) return rc;
            if (h_epoch_loss) LTRY(L, launch_loss_sum(L->loss.p, batch, L->eloss.p + e, s));
            L->step++;
            const double frac = (double)(t_step++) / total_steps;
            const float lr_t = (float)(lr * (0.05 + 0.95 * 0.5 * (1.0 + cos(3.14159265358979323846 * frac))));
            LTRY(L, launch_adam(L->w.p, L->w16.p, L->g.p, L->m.p, L->v.p, (long long)nparams(L), lr_t, 0.9f, 0.999f,
                                1e-8f, L->step, s));
            L->wt_fresh = false;
        }
    }
    if (h_epoch_loss) {
        std::vector<float> el(epochs);
        LTRY(L, hipMemcpyAsync(el.data(), L->eloss.p, sizeof(float) * epochs, hipMemcpyDeviceToHost, s));
        LTRY(L, hipStreamSynchronize(s));
        for (int e = 0; e < epochs; e++) h_epoch_loss[e] = (float)((double)el[e] / ((double)nbatch * batch));
    }
    LTRY(L, hipStreamSynchronize(s));
    return ORB_OK;
}

int orb_lindex_logits(orb_lindex* L, const uint8_t* d_q, int n, float* d_logits)
{
    if (!L || !d_q || !d_logits || n <= 0 || n % 64) return lfail(L, ORB_E_ARG, "n must be a positive multiple of 64");
    hipStream_t s = stream_of(L);
    const int chunk = 65536;
    int rc = ensure_batch(L, n < chunk ? n : chunk);
    if (rc != ORB_OK) return rc;
    for (int c0 = 0; c0 < n; c0 += chunk) {
        const int B = n - c0 < chunk ? n - c0 : chunk;
        LTRY(L, launch_bits_pm1(d_q + (size_t)c0 * 32, B, L->X.p, s));
        /* the product stores 16-byte row pieces: an unaligned buffer gets a copy */
        float* const dst = d_logits + (size_t)c0 * L->C;
        const bool direct = ((uintptr_t)dst & 15) == 0;
        rc = forward(L, B, s, false, d_q + (size_t)c0 * 32, direct ? dst : nullptr);
        if (rc != ORB_OK) return rc;
        if (!direct)
            LTRY(L, hipMemcpyAsync(dst, L->Z3.p, sizeof(float) * B * L->C, hipMemcpyDeviceToDevice, s));
    }
    return ORB_OK;
}

int orb_lindex_query(orb_lindex* L, const uint8_t* d_q, int nq, int mode, int* d_idx1, int* d_d1, int* d_idx2,
                     int* d_d2, int* h_nfallback)
{
    if (!L || !d_q || nq <= 0 || nq % 64) return lfail(L, ORB_E_ARG, "nq must be a positive multiple of 64");
    if (L->ndb <= 0) return lfail(L, ORB_E_STATE, "orb_lindex_build has not been called");
    if (mode != ORB_LINDEX_MLP && mode != ORB_LINDEX_CENTROIDS) return lfail(L, ORB_E_ARG, "unknown mode");
    hipStream_t s = stream_of(L);
    const int np = L->prm.nprobe, nb = L->C;
    const int* const count_before = L->list_count.p;
    const uint32_t* const bound_before = L->qbound.p;
    LALLOC(L, L->probes, (size_t)nq * np);
    LALLOC(L, L->pmax, nq);
    LALLOC(L, L->list_count, nb + 1);
    LALLOC(L, L->list_offsets, nb + 1);
    LALLOC(L, L->list_cursor, nb);
    LALLOC(L, L->unit_offsets, nb + 1);
    LALLOC(L, L->unit_pos, (size_t)nq * np / 256 + nb);
    LALLOC(L, L->lists, (size_t)nq * np);
    LALLOC(L, L->partial, (size_t)nq * np);
    LALLOC(L, L->qbound, nq);
    LALLOC(L, L->fb_count, 1);
    LALLOC(L, L->fb_index, nq);
    LALLOC(L, L->fb_qidx, nq);
    LALLOC(L, L->fb_q, (size_t)nq * 32);
    LALLOC(L, L->ex_idx1, nq);
    LALLOC(L, L->ex_d1, nq);
    LALLOC(L, L->ex_idx2, nq);
    LALLOC(L, L->ex_d2, nq);
    /* counts and bounds left clean by the previous batch's merge need no
       clearing launches; a new buffer or a batch that stopped early does */
    if (L->list_count.p != count_before || L->qbound.p != bound_before) L->clean_rows = 0;
    const bool clean = L->clean_rows >= nq;
    L->clean_rows = 0;
    /* the search needs the probed set only: its order is left as it comes */
    int rc = compute_probes(L, d_q, nq, mode, L->probes.p, L->pmax.p, false, s);
    if (rc != ORB_OK) return rc;
    LTRY(L, launch_probe_lists(L->probes.p, nq, np, nb, L->border.p, L->list_count.p, L->list_offsets.p,
                               L->list_cursor.p, L->lists.p, L->unit_offsets.p, L->unit_pos.p, clean, s));
    LTRY(L, launch_bucket_search(d_q, np, L->lists.p, L->list_offsets.p, L->unit_offsets.p, L->unit_pos.p, L->border.p, nb,
                                 L->db_sorted.p, L->perm.p, L->offsets.p, L->list_count.p + nb, L->qbound.p, nq,
                                 L->partial.p, L->ncu, clean, s));
    const float thresh = mode == ORB_LINDEX_MLP ? L->prm.fallback : -1.f;
    /* the max softmax probability is above 0: a threshold at or below 0 (and
       centroid mode) sends no query to the exact search, which is skipped */
    const bool may_fallback = thresh > 0.f;
    if (may_fallback) {
        LTRY(L, launch_fallback_compact(d_q, nq, L->pmax.p, thresh, L->fb_count.p, L->fb_index.p, L->fb_qidx.p,
                                        L->fb_q.p, s));
        /* exact search of the low-confidence queries (count read on the device) */
        rc = orb__match_dcount_top2(L->ctx, L->fb_q.p, nq, L->fb_count.p, L->db.p, L->ndb, L->ex_idx1.p, L->ex_d1.p,
                                    L->ex_idx2.p, L->ex_d2.p);
        if (rc != ORB_OK) return rc;
    }
    LTRY(L, launch_probe_merge(L->partial.p, L->probes.p, nq, np, may_fallback ? L->fb_index.p : nullptr, L->ex_idx1.p,
                               L->ex_d1.p, L->ex_idx2.p, L->ex_d2.p, d_idx1, d_d1, d_idx2, d_d2, L->list_count.p,
                               nb + 1, L->qbound.p, s));
    L->clean_rows = nq;
    if (h_nfallback) {
        if (may_fallback) {
            LTRY(L, hipMemcpyAsync(h_nfallback, L->fb_count.p, sizeof(int), hipMemcpyDeviceToHost, s));
            LTRY(L, hipStreamSynchronize(s));
        } else {
            *h_nfallback = 0;
        }
    }
    return ORB_OK;
}

/* The probed buckets of nq queries (nq x nprobe, best first, ties to the
   lower bucket id) and their max softmax probabilities (1 in centroid mode). */
int orb_lindex_probes(orb_lindex* L, const uint8_t* d_q, int nq, int mode, int* d_probes, float* d_pmax)
{
    if (!L || !d_q || !d_probes || !d_pmax || nq <= 0 || nq % 64)
        return lfail(L, ORB_E_ARG, "nq must be a positive multiple of 64");
    if (L->ndb <= 0) return lfail(L, ORB_E_STATE, "orb_lindex_build has not been called");
    if (mode != ORB_LINDEX_MLP && mode != ORB_LINDEX_CENTROIDS) return lfail(L, ORB_E_ARG, "unknown mode");
    return compute_probes(L, d_q, nq, mode, d_probes, d_pmax, true, stream_of(L));
}

int orb_lindex_set_probe(orb_lindex* L, int nprobe, float fallback)
{
    if (!L) return lfail(L, ORB_E_ARG, "null argument");
    if (nprobe < 1 || nprobe > L->C) return lfail(L, ORB_E_ARG, "nprobe out of range");
    if (L->mass < 1.f && nprobe > 64) return lfail(L, ORB_E_ARG, "nprobe above 64 with a probability-mass rule");
    L->prm.nprobe = nprobe;
    L->prm.fallback = fallback;
    return ORB_OK;
}

int orb_lindex_set_mass(orb_lindex* L, float mass, float centroid_temperature)
{
    if (!L) return lfail(L, ORB_E_ARG, "null argument");
    if (!(mass > 0.f) || !(mass <= 1.f)) return lfail(L, ORB_E_ARG, "mass must be in (0, 1]");
    if (!(centroid_temperature > 0.f) || !std::isfinite(centroid_temperature))
        return lfail(L, ORB_E_ARG, "centroid temperature must be positive");
    if (mass < 1.f && L->prm.nprobe > 64) return lfail(L, ORB_E_ARG, "a probability-mass rule needs nprobe <= 64");
    L->mass = mass;
    L->ctemp = centroid_temperature;
    return ORB_OK;
}

/* File: "ORBLIDX1", int32 version=1, nbuckets, hidden, nprobe, float fallback,
   then all fp32 parameters (W1 b1 W2 b2 W3 b3, row-major), then the
   nbuckets x 32 centroid bytes. */
int orb_lindex_save(orb_lindex* L, const char* path)
{
    if (!L || !path) return lfail(L, ORB_E_ARG, "null argument");
    if (L->ndb <= 0) return lfail(L, ORB_E_STATE, "nothing to save before orb_lindex_build");
    std::vector<float> w(nparams(L));
    std::vector<uint8_t> cen((size_t)L->C * 32);
    int rc = orb_lindex_get_weights(L, w.data());
    if (rc != ORB_OK) return rc;
    rc = orb_lindex_bucket_info(L, nullptr, cen.data(), nullptr);
    if (rc != ORB_OK) return rc;
    FILE* f = fopen(path, "wb");
    if (!f) return lfail(L, ORB_E_ARG, "cannot open the file for writing");
    /* version 2: after the fallback threshold, the bf16 bits of centroid_scale */
    const int32_t hdr[4] = {2, L->C, L->H, L->prm.nprobe};
    const uint32_t cs = L->cscale16;
    const bool ok = fwrite("ORBLIDX1", 1, 8, f) == 8 && fwrite(hdr, sizeof(hdr), 1, f) == 1 &&
                    fwrite(&L->prm.fallback, sizeof(float), 1, f) == 1 && fwrite(&cs, sizeof(cs), 1, f) == 1 &&
                    fwrite(w.data(), sizeof(float), w.size(), f) == w.size() &&
                    fwrite(cen.data(), 1, cen.size(), f) == cen.size();
    fclose(f);
    return ok ? ORB_OK : lfail(L, ORB_E_ARG, "write failed");
}

int orb_lindex_load(orb_lindex* L, const char* path)
{
    if (!L || !path) return lfail(L, ORB_E_ARG, "null argument");
    FILE* f = fopen(path, "rb");
    if (!f) return lfail(L, ORB_E_ARG, "cannot open the file");
    char magic[8];
    int32_t hdr[4];
    float fb;
    uint32_t cs = 0;
    bool ok = fread(magic, 1, 8, f) == 8 && memcmp(magic, "ORBLIDX1", 8) == 0 && fread(hdr, sizeof(hdr), 1, f) == 1 &&
              (hdr[0] == 1 || hdr[0] == 2) && hdr[1] == L->C && hdr[2] == L->H && hdr[3] >= 1 && hdr[3] <= L->C &&
              fread(&fb, sizeof(float), 1, f) == 1;
    if (ok && hdr[0] == 2) ok = fread(&cs, sizeof(cs), 1, f) == 1 && cs <= 0xFFFFu;
    std::vector<float> w(nparams(L));
    std::vector<uint8_t> cen((size_t)L->C * 32);
    ok = ok && fread(w.data(), sizeof(float), w.size(), f) == w.size() && fread(cen.data(), 1, cen.size(), f) == cen.size();
    fclose(f);
    if (!ok) return lfail(L, ORB_E_ARG, "not a matching learned-index file (magic, version or shape)");
    int rc = orb_lindex_set_weights(L, w.data());
    if (rc != ORB_OK) return rc;
    LALLOC(L, L->centroids, (size_t)L->C * 32);
    hipStream_t s = stream_of(L);
    LTRY(L, hipMemcpyAsync(L->centroids.p, cen.data(), cen.size(), hipMemcpyHostToDevice, s));
    LTRY(L, hipStreamSynchronize(s));
    L->prm.nprobe = hdr[3];
    L->prm.fallback = fb;
    if (hdr[3] > 64) L->mass = 1.f; /* the mass rule covers at most 64 probes */
    L->cscale16 = (uint16_t)cs;
    L->ndb = -1; /* centroids loaded: orb_lindex_build(db, n, 0) assigns the database to them */
    return ORB_OK;
}

} // extern "C"
