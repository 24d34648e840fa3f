/*
 * gfx950 kernels of the ORB detect / describe path.
 *
 * One launch per stage covers every (frame, level) of a frame set:
 *   k_pyramid   raw + 7x7-blurred levels from the input, LDS halo tiles
 *               (orbd.cu:514-535 pyramid, orbd.cu:1273-1338 blur)
 *   k_detect    FAST-9 segment test + Harris / FAST score on every level,
 *               64-bit atomic max into the level-0 key map (orbd.cu:672-756)
 *   k_nms       square cross-level NMS, wavefront-ballot compaction of the
 *               survivors of each 256-pixel row segment (orbd.cu:919-983)
 *   k_select    per-frame scan of the segment counts and, when more than
 *               max_pts survive, an exact radix select of the score cut
 *   k_emit      survivors -> OrbPoint in raster order
 *   k_describe  intensity-centroid angle + rBRIEF, one wavefront per keypoint
 *               (orbd.cu:1001-1036, 1341-1468)
 * Every floating-point step is the explicit IEEE operation (__f*_rn) the CPU
 * oracle performs, so results are bit-identical to it.
 */
#include "orb_internal.h"

namespace orbk
{

/* ------------------------------------------------------------ device math */

/* centroid angle, orbd.cu:986-998 */
__device__ __forceinline__ float dev_atan2(float y, float x)
{
    const float ax = fabsf(x), ay = fabsf(y);
    const float a = __fdiv_rn(fminf(ax, ay), fmaxf(ax, ay));
    const float s = __fmul_rn(a, a);
    float r = __fmaf_rn(__fmaf_rn(__fmaf_rn(-0.0464964749f, s, 0.15931422f), s, -0.327622764f),
                        __fmul_rn(s, a), a);
    r = (ay > ax) ? __fsub_rn(1.5707963267948966f, r) : r;
    r = (x < 0) ? (float)__dsub_rn(3.14159265358979323846, (double)r) : r;
    r = (y < 0) ? -r : r;
    return r;
}

/* sin/cos: Cody-Waite reduction by pi/2 (three-part split) and the Cephes
   sinf/cosf minimax polynomials on [-pi/4, pi/4]; replaces __sinf/__cosf of
   orbd.cu:1446-1447 with a definition the CPU reproduces exactly. */
__device__ __forceinline__ void dev_sincos(float a, float& sn, float& cs)
{
    const float j = rintf(__fmul_rn(a, 0.636619772f));
    float r = __fmaf_rn(-j, 1.5703125f, a);
    r = __fmaf_rn(-j, 4.837512969970703125e-4f, r);
    r = __fmaf_rn(-j, 7.54978995489188216e-8f, r);
    const int q = ((int)j) & 3;
    const float z = __fmul_rn(r, r);
    const float sp = __fmaf_rn(
        __fmaf_rn(__fmaf_rn(-1.9515295891e-4f, z, 8.3321608736e-3f), z, -1.6666654611e-1f),
        __fmul_rn(z, r), r);
    const float cp = __fmaf_rn(
        __fmaf_rn(__fmaf_rn(2.443315711809948e-5f, z, -1.388731625493765e-3f), z, 4.166664568298827e-2f),
        __fmul_rn(z, z), __fmaf_rn(-0.5f, z, 1.0f));
    switch (q) {
    case 0: sn = sp; cs = cp; break;
    case 1: sn = cp; cs = -sp; break;
    case 2: sn = -sp; cs = -cp; break;
    default: sn = -cp; cs = sp; break;
    }
}

/* ------------------------------------------------------------- utilities */

__device__ __forceinline__ Job find_job(const Job* __restrict__ jobs, int n, int b)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].first_block <= b) lo = mid;
        else hi = mid - 1;
    }
    return jobs[lo];
}

/* reflect-101 (cuda_utils.h:145-152), clamped so halo entries that no output
   uses stay inside the level */
__device__ __forceinline__ int reflect101(int i, int n)
{
    if (i < 0) i = -i;
    if (i >= n) i = n + n - 2 - i;
    return min(max(i, 0), n - 1);
}

__device__ __forceinline__ const int* level_tab(const FrameDev& F, const int* tabs, int l)
{
    return tabs + F.tab_off[l];
}

/* level-l pixel (x, y) computed from the input image */
__device__ __forceinline__ int src_value(const FrameDev& F, const int* __restrict__ tabs, int l, int x, int y)
{
    if (l == 0) return F.img[(long long)y * F.ip + x];
    if (F.ref_mode) return F.img[(long long)(y << l) * F.ip + (x << l)];
    const int* t = level_tab(F, tabs, l);
    const int tx = t[x], ty = t[F.lw[l] + y];
    const int x0 = tx >> 8, fx = tx & 255, y0 = ty >> 8, fy = ty & 255;
    const int x1 = min(x0 + 1, F.w0 - 1), y1 = min(y0 + 1, F.h0 - 1);
    const uint8_t* r0 = F.img + (long long)y0 * F.ip;
    const uint8_t* r1 = F.img + (long long)y1 * F.ip;
    const int top = (256 - fx) * r0[x0] + fx * r0[x1];
    const int bot = (256 - fx) * r1[x0] + fx * r1[x1];
    return ((256 - fy) * top + fy * bot + 32768) >> 16;
}

__device__ __forceinline__ int to_l0(const FrameDev& F, const int* __restrict__ tabs, int l, int v, bool is_y)
{
    if (l == 0) return v;
    if (F.ref_mode) return v << l;
    const int* t = level_tab(F, tabs, l);
    return is_y ? t[2 * F.lw[l] + F.lh[l] + v] : t[F.lw[l] + F.lh[l] + v];
}

/* level-0 coordinate -> level coordinate: the unique v with to_l0(v) == x0 */
__device__ __forceinline__ int from_l0(const FrameDev& F, const int* __restrict__ tabs, int l, int x0, bool is_y)
{
    if (l == 0) return x0;
    if (F.ref_mode) return x0 >> l;
    const int n = is_y ? F.lh[l] : F.lw[l];
    int v = __float2int_rn(__fmul_rn((float)x0, F.inv_sc[l]));
    v = min(max(v, 0), n - 1);
    const int m = to_l0(F, tabs, l, v, is_y);
    if (m < x0 && v + 1 < n) v++;
    else if (m > x0 && v > 0) v--;
    return v;
}

__device__ __forceinline__ int wave_sum(int v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

/* ---------------------------------------------------------------- pyramid */

__global__ __launch_bounds__(256) void k_pyramid(const FrameDev* __restrict__ frames, const Job* __restrict__ jobs,
                                                 int njobs, Tables T)
{
    __shared__ uint8_t s_raw[kPyrTH + 6][kPyrTW + 8];
    __shared__ float s_h[kPyrTH + 6][kPyrTW + 1];
    const Job jb = find_job(jobs, njobs, blockIdx.x);
    const FrameDev& F = frames[jb.frame];
    const int l = jb.level;
    const int t = blockIdx.x - jb.first_block;
    const int x0 = (t % jb.tiles_x) * kPyrTW, y0 = (t / jb.tiles_x) * kPyrTH;
    const int w = F.lw[l], h = F.lh[l], p = F.lp[l];
    const int tid = threadIdx.x;
    uint8_t* raw = F.raw + F.loff[l];
    uint8_t* blur = F.blur + F.loff[l];
    const bool inside = x0 < w;

    if (inside) {
        for (int i = tid; i < (kPyrTH + 6) * (kPyrTW + 6); i += 256) {
            const int r = i / (kPyrTW + 6), c = i - r * (kPyrTW + 6);
            const int yy = reflect101(y0 - 3 + r, h), xx = reflect101(x0 - 3 + c, w);
            s_raw[r][c] = (uint8_t)src_value(F, T.tabs, l, xx, yy);
        }
    }
    __syncthreads();
    for (int i = tid; i < kPyrTH * kPyrTW; i += 256) {
        const int r = i / kPyrTW, c = i - r * kPyrTW;
        const int y = y0 + r, x = x0 + c;
        if (y < h && x < p) {
            uint8_t v;
            if (x < w) {
                v = s_raw[r + 3][c + 3];
            } else {
                /* row padding: level 0 keeps the input's bytes, other levels 0 */
                v = (l == 0 && x < F.ip) ? F.img[(long long)y * F.ip + x] : (uint8_t)0;
                blur[(long long)y * p + x] = v;
            }
            raw[(long long)y * p + x] = v;
        }
    }
    if (!inside) return;
    for (int i = tid; i < (kPyrTH + 6) * kPyrTW; i += 256) {
        const int r = i / kPyrTW, c = i - r * kPyrTW;
        const uint8_t* s = &s_raw[r][c + 3];
        float v = __fmul_rn(T.k0, (float)s[0]);
        v = __fmaf_rn(T.k1, (float)((int)s[-1] + (int)s[1]), v);
        v = __fmaf_rn(T.k2, (float)((int)s[-2] + (int)s[2]), v);
        v = __fmaf_rn(T.k3, (float)((int)s[-3] + (int)s[3]), v);
        s_h[r][c] = v;
    }
    __syncthreads();
    for (int i = tid; i < kPyrTH * kPyrTW; i += 256) {
        const int r = i / kPyrTW, c = i - r * kPyrTW;
        const int y = y0 + r, x = x0 + c;
        if (y < h && x < w) {
            float v = __fmul_rn(T.k0, s_h[r + 3][c]);
            v = __fmaf_rn(T.k1, __fadd_rn(s_h[r + 2][c], s_h[r + 4][c]), v);
            v = __fmaf_rn(T.k2, __fadd_rn(s_h[r + 1][c], s_h[r + 5][c]), v);
            v = __fmaf_rn(T.k3, __fadd_rn(s_h[r][c], s_h[r + 6][c]), v);
            blur[(long long)y * p + x] = (uint8_t)__float2int_rn(v);
        }
    }
}

/* -------------------------------------------------------------- detection */

__constant__ int c_circle_dx[16] = {0, 1, 2, 3, 3, 3, 2, 1, 0, -1, -2, -3, -3, -3, -2, -1};
__constant__ int c_circle_dy[16] = {3, 3, 2, 1, 0, -1, -2, -3, -3, -3, -2, -1, 0, 1, 2, 3};

/* OpenCV cornerScore<16> as in orbd.cu:538-578 */
__device__ float fast_score(const int* ring, int v, int threshold)
{
    int d[25];
#pragma unroll
    for (int k = 0; k < 16; k++) d[k] = (short)(v - ring[k]);
#pragma unroll
    for (int k = 16; k < 25; k++) d[k] = d[k - 16];
    int a0 = threshold;
#pragma unroll
    for (int k = 0; k < 16; k += 2) {
        int a = min(d[k + 1], d[k + 2]);
        a = min(a, d[k + 3]);
        if (a <= a0) continue;
        a = min(a, d[k + 4]); a = min(a, d[k + 5]); a = min(a, d[k + 6]);
        a = min(a, d[k + 7]); a = min(a, d[k + 8]);
        a0 = max(a0, min(a, d[k]));
        a0 = max(a0, min(a, d[k + 9]));
    }
    int b0 = -a0;
#pragma unroll
    for (int k = 0; k < 16; k += 2) {
        int b = max(d[k + 1], d[k + 2]);
        b = max(b, d[k + 3]); b = max(b, d[k + 4]); b = max(b, d[k + 5]);
        if (b >= b0) continue;
        b = max(b, d[k + 6]); b = max(b, d[k + 7]); b = max(b, d[k + 8]);
        b0 = min(b0, max(b, d[k]));
        b0 = min(b0, max(b, d[k + 9]));
    }
    return (float)(-b0 - 1);
}

/* Harris response, orbd.cu:651-669: 32-bit wrapping products, then
   fma(-0.04, B, A) * (1/(4*7*255))^4 */
__device__ float harris_score(const uint8_t* __restrict__ c, int p, float scale)
{
    int win[9][9];
#pragma unroll
    for (int r = 0; r < 9; r++)
#pragma unroll
        for (int q = 0; q < 9; q++) win[r][q] = c[(long long)(r - 4) * p + (q - 4)];
    int dxx = 0, dyy = 0, dxy = 0;
#pragma unroll
    for (int i = 1; i < 8; i++) {
#pragma unroll
        for (int j = 1; j < 8; j++) {
            const int dx = (win[i][j + 1] - win[i][j - 1]) * 2 + (win[i - 1][j + 1] - win[i - 1][j - 1]) +
                           (win[i + 1][j + 1] - win[i + 1][j - 1]);
            const int dy = (win[i + 1][j] - win[i - 1][j]) * 2 + (win[i + 1][j - 1] - win[i - 1][j - 1]) +
                           (win[i + 1][j + 1] - win[i - 1][j + 1]);
            dxx += dx * dx;
            dyy += dy * dy;
            dxy += dx * dy;
        }
    }
    const int A = (int)((unsigned)dxx * (unsigned)dyy - (unsigned)dxy * (unsigned)dxy);
    const unsigned S = (unsigned)(dxx + dyy);
    const int B = (int)(S * S);
    return __fmul_rn(__fmaf_rn(-0.04f, __int2float_rn(B), __int2float_rn(A)), scale);
}

__global__ __launch_bounds__(256) void k_detect(const FrameDev* __restrict__ frames, const Job* __restrict__ jobs,
                                                int njobs, Tables T, const uint32_t* __restrict__ epochs)
{
    const Job jb = find_job(jobs, njobs, blockIdx.x);
    const FrameDev& F = frames[jb.frame];
    const int l = jb.level;
    const int t = blockIdx.x - jb.first_block;
    const int ix = 3 + (t % jb.tiles_x) * kDetTW + (threadIdx.x & (kDetTW - 1));
    const int iy = 3 + (t / jb.tiles_x) * kDetTH + (threadIdx.x / kDetTW);
    const int w = F.lw[l], h = F.lh[l], p = F.lp[l];
    if (ix >= w - 3 || iy >= h - 3) return;
    const uint8_t* c = F.raw + F.loff[l] + (long long)iy * p + ix;
    const int v = c[0];
    const int thr = T.threshold;
    int ring[16];
    unsigned dark = 0, bright = 0;
#pragma unroll
    for (int k = 0; k < 16; k++) {
        ring[k] = c[c_circle_dx[k] + c_circle_dy[k] * p];
        dark |= (unsigned)(ring[k] - v < -thr) << k;
        bright |= (unsigned)(ring[k] - v > thr) << k;
    }
    /* quick test of orbd.cu:687-700: every opposite pair has a darker
       (bit 1) / brighter (bit 2) member; the darker arc is tested first and
       exclusively (orbd.cu:703-740) */
    const bool d1 = ((dark | (dark >> 8)) & 0xffu) == 0xffu;
    const bool d2 = ((bright | (bright >> 8)) & 0xffu) == 0xffu;
    unsigned m = d1 ? dark : (d2 ? bright : 0u);
    if (m == 0) return;
    /* > 8 consecutive of the 25-entry wrapped ring == a circular run of 9 */
    unsigned r = m | (m << 16);
    unsigned run = r;
#pragma unroll
    for (int k = 1; k < 9; k++) run &= r >> k;
    if (run == 0) return;
    const float score = T.score_type == 0 ? harris_score(c, p, T.harris_scale) : fast_score(ring, v, thr);
    if (!(score > 0.f)) return;
    const int x0 = to_l0(F, T.tabs, l, ix, false), y0 = to_l0(F, T.tabs, l, iy, true);
    const unsigned long long key = ((unsigned long long)epochs[jb.frame] << kEpochShift) |
                                   ((unsigned long long)__float_as_uint(score) << 3) |
                                   (unsigned long long)(7 - l);
    atomicMax(F.keys + (long long)y0 * F.p0 + x0, key);
}

/* -------------------------------------------------------------------- NMS */

__global__ __launch_bounds__(256) void k_nms(const FrameDev* __restrict__ frames, const int* __restrict__ block_frame,
                                             const uint32_t* __restrict__ epochs, Candidate* __restrict__ cand,
                                             int* __restrict__ blk_count)
{
    __shared__ int s_w[4];
    const int f = block_frame[blockIdx.x];
    const FrameDev& F = frames[f];
    const int b = blockIdx.x - F.nms_first;
    const int row = b / F.nms_segs, seg = b - row * F.nms_segs;
    const int y = F.border + row;
    const int x = F.border + seg * kNmsSeg + threadIdx.x;
    const uint32_t epoch = epochs[f];
    bool surv = false;
    uint32_t sbits = 0, lev = 0;
    if (x < F.w0 - F.border) {
        const unsigned long long* kp = F.keys + (long long)y * F.p0 + x;
        const unsigned long long k = *kp;
        if ((uint32_t)(k >> kEpochShift) == epoch) {
            sbits = (uint32_t)(k >> 3) & 0x7fffffffu;
            lev = 7u - (uint32_t)(k & 7u);
            const int r = F.radius[lev];
            surv = true;
            for (int dy = -r; dy <= r && surv; dy++) {
                const unsigned long long* rp = kp + (long long)dy * F.p0;
                for (int dx = -r; dx <= r; dx++) {
                    const unsigned long long nk = rp[dx];
                    if ((uint32_t)(nk >> kEpochShift) != epoch || (dx == 0 && dy == 0)) continue;
                    const uint32_t ns = (uint32_t)(nk >> 3) & 0x7fffffffu;
                    if (ns > sbits || (ns == sbits && dx <= 0 && dy <= 0)) {
                        surv = false;
                        break;
                    }
                }
            }
        }
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const unsigned long long m = __ballot(surv);
    const int pre = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) s_w[wid] = __popcll(m);
    __syncthreads();
    int base = 0;
    for (int i = 0; i < wid; i++) base += s_w[i];
    if (surv) {
        Candidate c;
        c.xy = ((uint32_t)y << 16) | (uint32_t)x;
        c.score = sbits;
        c.level = lev;
        c.pad = 0;
        cand[(long long)blockIdx.x * (kNmsSeg / 2) + base + pre] = c;
    }
    if (threadIdx.x == 0) blk_count[blockIdx.x] = s_w[0] + s_w[1] + s_w[2] + s_w[3];
}

/* ---------------------------------------------------------------- select */

/* exclusive scan over the 1024 threads of a workgroup */
__device__ int block_scan_1024(int v, int* s_tmp, int& total)
{
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) s_tmp[wid] = x;
    __syncthreads();
    if (wid == 0) {
        int s = lane < 16 ? s_tmp[lane] : 0;
#pragma unroll
        for (int o = 1; o < 16; o <<= 1) {
            const int y = __shfl_up(s, o, 64);
            if (lane >= o) s += y;
        }
        if (lane < 16) s_tmp[16 + lane] = s;
    }
    __syncthreads();
    total = s_tmp[16 + 15];
    const int res = x - v + (wid > 0 ? s_tmp[16 + wid - 1] : 0);
    __syncthreads();
    return res;
}

__global__ __launch_bounds__(1024) void k_select(const FrameDev* __restrict__ frames, const Candidate* __restrict__ cand,
                                                 const int* __restrict__ blk_count, int* __restrict__ blk_out,
                                                 int* __restrict__ blk_eq, SelectState* __restrict__ sel,
                                                 uint32_t* __restrict__ epochs)
{
    __shared__ int s_tmp[32];
    __shared__ int s_hist[256];
    __shared__ int s_pick[2];
    const int f = blockIdx.x;
    const FrameDev& F = frames[f];
    const int nb = F.nms_rows * F.nms_segs, first = F.nms_first;
    const int tid = threadIdx.x;
    const int per = (nb + 1023) / 1024;
    const int b0 = min(nb, tid * per), b1 = min(nb, b0 + per);
    const int cap = F.cap;
    const uint32_t epoch = epochs[f];

    int sum = 0;
    for (int b = b0; b < b1; b++) sum += blk_count[first + b];
    int total;
    int run = block_scan_1024(sum, s_tmp, total);

    uint32_t thr = 0;
    int need = 0;
    if (total <= cap) {
        for (int b = b0; b < b1; b++) {
            blk_out[first + b] = run;
            blk_eq[first + b] = 0;
            run += blk_count[first + b];
        }
    } else {
        /* exact radix select of the cap-th largest score (8 bits per pass) */
        uint32_t prefix = 0, pmask = 0;
        int k = cap;
        for (int shift = 24; shift >= 0; shift -= 8) {
            for (int i = tid; i < 256; i += 1024) s_hist[i] = 0;
            __syncthreads();
            for (int b = b0; b < b1; b++) {
                const int n = blk_count[first + b];
                const Candidate* cb = cand + (long long)(first + b) * (kNmsSeg / 2);
                for (int i = 0; i < n; i++) {
                    const uint32_t s = cb[i].score;
                    if ((s & pmask) == prefix) atomicAdd(&s_hist[(s >> shift) & 255u], 1);
                }
            }
            __syncthreads();
            if (tid == 0) {
                int acc = 0, d = 255;
                for (; d > 0; d--) {
                    if (acc + s_hist[d] >= k) break;
                    acc += s_hist[d];
                }
                s_pick[0] = d;
                s_pick[1] = k - acc;
            }
            __syncthreads();
            prefix |= (uint32_t)s_pick[0] << shift;
            pmask |= 255u << shift;
            k = s_pick[1];
            __syncthreads();
        }
        thr = prefix;
        need = k;
        /* per-thread counts of > thr and == thr over its blocks */
        int gt_sum = 0, eq_sum = 0;
        for (int b = b0; b < b1; b++) {
            const int n = blk_count[first + b];
            const Candidate* cb = cand + (long long)(first + b) * (kNmsSeg / 2);
            for (int i = 0; i < n; i++) {
                gt_sum += cb[i].score > thr;
                eq_sum += cb[i].score == thr;
            }
        }
        int eq_tot;
        int eq_before = block_scan_1024(eq_sum, s_tmp, eq_tot);
        /* kept per thread: gt + min(eq, max(0, need - eq_before)) */
        int kept_thr = gt_sum + min(eq_sum, max(0, need - eq_before));
        int kept_tot;
        int out = block_scan_1024(kept_thr, s_tmp, kept_tot);
        for (int b = b0; b < b1; b++) {
            const int n = blk_count[first + b];
            const Candidate* cb = cand + (long long)(first + b) * (kNmsSeg / 2);
            int g = 0, e = 0;
            for (int i = 0; i < n; i++) {
                g += cb[i].score > thr;
                e += cb[i].score == thr;
            }
            blk_out[first + b] = out;
            blk_eq[first + b] = eq_before;
            out += g + min(e, max(0, need - eq_before));
            eq_before += e;
        }
        total = cap;
    }
    if (tid == 0) {
        sel[f].thr = thr;
        sel[f].need = need;
        sel[f].epoch = epoch;
        *F.count = total;
    }
    /* advance the epoch; on wrap-around clear this frame's key map */
    if (epoch + 1 >= kEpochLimit) {
        const long long n = (long long)F.h0 * F.p0;
        for (long long i = tid; i < n; i += 1024) F.keys[i] = 0ull;
        if (tid == 0) epochs[f] = 1;
    } else if (tid == 0) {
        epochs[f] = epoch + 1;
    }
}

/* ------------------------------------------------------------------ emit */

__device__ __forceinline__ int block_scan_256(int v, int* s_tmp)
{
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const unsigned long long m = __ballot(v != 0);
    const int pre = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) s_tmp[wid] = __popcll(m);
    __syncthreads();
    int base = 0;
    for (int i = 0; i < wid; i++) base += s_tmp[i];
    __syncthreads();
    return base + pre;
}

__global__ __launch_bounds__(256) void k_emit(const FrameDev* __restrict__ frames, const int* __restrict__ block_frame,
                                              const Candidate* __restrict__ cand, const int* __restrict__ blk_count,
                                              const int* __restrict__ blk_out, const int* __restrict__ blk_eq,
                                              const SelectState* __restrict__ sel)
{
    __shared__ int s_tmp[4];
    const int n = blk_count[blockIdx.x];
    if (n == 0) return;
    const int f = block_frame[blockIdx.x];
    const FrameDev& F = frames[f];
    const SelectState S = sel[f];
    const int i = threadIdx.x;
    const bool has = i < n;
    Candidate c = {0, 0, 0, 0};
    if (has) c = cand[(long long)blockIdx.x * (kNmsSeg / 2) + i];
    const bool is_eq = has && c.score == S.thr;
    const int eq_rank = block_scan_256(is_eq, s_tmp);
    const bool kept = has && (c.score > S.thr || (is_eq && blk_eq[blockIdx.x] + eq_rank < S.need));
    const int rank = block_scan_256(kept, s_tmp);
    if (kept) {
        orb::OrbPoint* o = F.pts + blk_out[blockIdx.x] + rank;
        o->x = (int)(c.xy & 0xffffu);
        o->y = (int)(c.xy >> 16);
        o->octave = (int)c.level;
        o->score = __uint_as_float(c.score);
        o->angle = 0.f;
        o->match = -1;
        o->distance = -1;
    }
}

/* --------------------------------------------------------- angle + describe */

__device__ __forceinline__ int sample(const uint8_t* __restrict__ c, int p, int2 pt, float sn, float cs)
{
    const float x = (float)pt.x, y = (float)pt.y;
    const int ix = __float2int_rn(__fmaf_rn(x, cs, -__fmul_rn(y, sn)));
    const int iy = __float2int_rn(__fmaf_rn(x, sn, __fmul_rn(y, cs)));
    return c[(long long)iy * p + ix];
}

__global__ __launch_bounds__(256) void k_describe(const FrameDev* __restrict__ frames, int blocks_per_frame, Tables T,
                                                  int compute_desc)
{
    __shared__ int2 s_pat[512];
    __shared__ int s_umax[32];
    __shared__ uint8_t s_code[kDescPerBlock][128];
    const int f = blockIdx.x / blocks_per_frame;
    const FrameDev& F = frames[f];
    for (int i = threadIdx.x; i < 512; i += 256) s_pat[i] = T.pattern[i];
    if (threadIdx.x < 17) s_umax[threadIdx.x] = T.umax[threadIdx.x];
    __syncthreads();
    const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int pi = (blockIdx.x - f * blocks_per_frame) * kDescPerBlock + wid;
    const int n = *F.count;
    if (pi >= n) return;
    orb::OrbPoint* P = F.pts + pi;
    const int l = P->octave;
    const int cx = from_l0(F, T.tabs, l, P->x, false), cy = from_l0(F, T.tabs, l, P->y, true);
    const int p = F.lp[l];
    const long long co = F.loff[l] + (long long)cy * p + cx;

    /* intensity centroid over the umax disc of the raw level (orbd.cu:1011-1032) */
    const uint8_t* craw = F.raw + co;
    const int hk = T.half_patch, side = 2 * hk + 1;
    int m10 = 0, m01 = 0;
    for (int i = lane; i < side * side; i += 64) {
        const int v = i / side - hk, u = i % side - hk;
        if (abs(u) <= s_umax[abs(v)]) {
            const int I = craw[(long long)v * p + u];
            m10 += u * I;
            m01 += v * I;
        }
    }
    m10 = wave_sum(m10);
    m01 = wave_sum(m01);
    const float ang = (m01 == 0 && m10 == 0) ? 0.f : dev_atan2(__int2float_rn(m01), __int2float_rn(m10));
    if (lane == 0) P->angle = ang;
    if (!compute_desc) return;

    float sn, cs;
    dev_sincos(ang, sn, cs);
    const uint8_t* cb = F.blur + co;
    unsigned long long* out = (unsigned long long*)(F.desc + (long long)pi * 32);
    if (T.wta_k == 2) {
        /* bit k of the descriptor compares points 2k and 2k+1 (orbd.cu:1351-1361) */
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int k = lane + 64 * j;
            const bool bit = sample(cb, p, s_pat[2 * k], sn, cs) < sample(cb, p, s_pat[2 * k + 1], sn, cs);
            const unsigned long long m = __ballot(bit);
            if (lane == 0) out[j] = m;
        }
    } else {
        /* 2-bit argmax codes, code q = 4*byte + slot (orbd.cu:1364-1434) */
        for (int j = 0; j < 2; j++) {
            const int q = lane + 64 * j, byte = q >> 2, slot = q & 3;
            int code;
            if (T.wta_k == 3) {
                const int2* pp = s_pat + 12 * byte;
                const int i0 = slot < 3 ? 3 * slot : 8, i1 = slot < 3 ? 3 * slot + 1 : 9,
                          i2 = slot < 3 ? 3 * slot + 2 : 11;
                const int t0 = sample(cb, p, pp[i0], sn, cs), t1 = sample(cb, p, pp[i1], sn, cs),
                          t2 = sample(cb, p, pp[i2], sn, cs);
                code = t2 > t1 ? (t2 > t0 ? 2 : 0) : (t1 > t0);
            } else {
                const int2* pp = s_pat + 16 * byte + 4 * slot;
                int t0 = sample(cb, p, pp[0], sn, cs), t1 = sample(cb, p, pp[1], sn, cs);
                int t2 = sample(cb, p, pp[2], sn, cs), t3 = sample(cb, p, pp[3], sn, cs);
                int u = 0, v = 2;
                if (t1 > t0) { t0 = t1; u = 1; }
                if (t3 > t2) { t2 = t3; v = 3; }
                code = t0 > t2 ? u : v;
            }
            s_code[wid][q] = (uint8_t)code;
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        if (lane < 32) {
            const uint8_t* sc = s_code[wid] + 4 * lane;
            F.desc[(long long)pi * 32 + lane] = (uint8_t)(sc[0] | (sc[1] << 2) | (sc[2] << 4) | (sc[3] << 6));
        }
    }
}

/* ------------------------------------------------------------ debug math */

__global__ void k_debug_sincos(const float* in, int n, float* s, float* c)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        float a, b;
        dev_sincos(in[i], a, b);
        s[i] = a;
        c[i] = b;
    }
}

__global__ void k_debug_atan2(const float* yx, int n, float* out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = dev_atan2(yx[2 * i], yx[2 * i + 1]);
}

/* ------------------------------------------------------------- launchers */

hipError_t launch_detect(const DetectLaunch& d, int compute_desc, hipStream_t s, const StageEvents* ev)
{
    if (d.pyr_blocks > 0) {
        ev_start(ev, kStPyramid, s);
        hipLaunchKernelGGL(k_pyramid, dim3(d.pyr_blocks), dim3(256), 0, s, d.frames, d.pyr_jobs, d.pyr_njobs, d.t);
        ev_stop(ev, kStPyramid, s);
    }
    if (d.det_blocks > 0) {
        ev_start(ev, kStDetect, s);
        hipLaunchKernelGGL(k_detect, dim3(d.det_blocks), dim3(256), 0, s, d.frames, d.det_jobs, d.det_njobs, d.t,
                           d.epochs);
        ev_stop(ev, kStDetect, s);
    }
    if (d.nms_blocks > 0) {
        ev_start(ev, kStNms, s);
        hipLaunchKernelGGL(k_nms, dim3(d.nms_blocks), dim3(256), 0, s, d.frames, d.nms_block_frame, d.epochs,
                           d.cand, d.blk_count);
        ev_stop(ev, kStNms, s);
    }
    ev_start(ev, kStSelect, s);
    hipLaunchKernelGGL(k_select, dim3(d.nframes), dim3(1024), 0, s, d.frames, d.cand, d.blk_count, d.blk_out,
                       d.blk_eq, d.sel, d.epochs);
    ev_stop(ev, kStSelect, s);
    if (d.nms_blocks > 0) {
        ev_start(ev, kStEmit, s);
        hipLaunchKernelGGL(k_emit, dim3(d.nms_blocks), dim3(256), 0, s, d.frames, d.nms_block_frame, d.cand,
                           d.blk_count, d.blk_out, d.blk_eq, d.sel);
        ev_stop(ev, kStEmit, s);
    }
    if (compute_desc && d.desc_blocks_per_frame > 0) {
        ev_start(ev, kStDescribe, s);
        hipLaunchKernelGGL(k_describe, dim3(d.desc_blocks_per_frame * d.nframes), dim3(256), 0, s, d.frames,
                           d.desc_blocks_per_frame, d.t, compute_desc);
        ev_stop(ev, kStDescribe, s);
    }
    return hipGetLastError();
}

hipError_t launch_debug_sincos(const float* in, int n, float* s, float* c, hipStream_t st)
{
    if (n > 0) hipLaunchKernelGGL(k_debug_sincos, dim3((n + 255) / 256), dim3(256), 0, st, in, n, s, c);
    return hipGetLastError();
}

hipError_t launch_debug_atan2(const float* yx, int n, float* out, hipStream_t st)
{
    if (n > 0) hipLaunchKernelGGL(k_debug_atan2, dim3((n + 255) / 256), dim3(256), 0, st, yx, n, out);
    return hipGetLastError();
}

} // namespace orbk
