/*
 * gfx950 kernels of the ORB detect / describe path.
 *
 * One launch per stage covers every (frame, level) of a frame set:
 *   k_pyramid   raw + 7x7-blurred levels from the input, one wavefront per
 *               strip walking its rows in registers, neighbours across lanes
 *               (orbd.cu:514-535 pyramid, orbd.cu:1273-1338 blur)
 *   k_detect    FAST-9 segment test + Harris / FAST score on every level,
 *               64-bit atomic max into the level-0 key map (orbd.cu:672-756)
 *   k_nms       square cross-level NMS of the pixels detection marked, the
 *               survivors of each 2048-pixel row segment in x order
 *               (orbd.cu:919-983)
 *   k_select    per-frame scan of the segment counts, keypoints out in
 *               raster order; with more than max_pts survivors, an exact
 *               radix select of the score cut
 *   k_describe  intensity-centroid angle + rBRIEF, one wavefront per keypoint
 *               (orbd.cu:1001-1036, 1341-1468)
 *   k_export    OpenCV cv::KeyPoint-layout rows of a keypoint set, honouring
 *               retain_topn
 * Every floating-point step is the explicit IEEE operation (__f*_rn) the CPU
 * oracle performs, so results are bit-identical to it.
 */
#include "orb_internal.h"
#include "orb_wave.h"

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

/* Pointers read from FrameDev are generic; the global-space view makes the
   compiler emit global_* instead of flat_* memory operations, which it can
   wait on per class (vmcnt) without draining LDS traffic as well. */
#if defined(__HIP_DEVICE_COMPILE__)
#define ORB_GLOBAL __attribute__((address_space(1)))
#else
#define ORB_GLOBAL /* the host pass only parses device code */
#endif
template <typename T>
__device__ __forceinline__ ORB_GLOBAL T* as_global(T* p)
{
    return (ORB_GLOBAL T*)p;
}

/* The job of unit b (wave-uniform): job_of holds the job index of every
   unit, so the lookup is two dependent scalar loads whatever the number of
   jobs (frames x levels x bands). */
__device__ __forceinline__ Job find_job(const Job* __restrict__ jobs, const int* __restrict__ job_of, int b)
{
    return jobs[__builtin_amdgcn_readfirstlane(job_of[b])];
}

/* Fixed-point bilinear source position of level index i along an axis of S
   input and D level pixels: floor(((2i+1)S - D) * 256 / (2D)) clamped to
   [0, (S-1)*256] (centre-aligned, 8 fractional bits).  The numerator is an
   integer below 2^53 and a non-integral quotient lies at least 1/(2D) from
   the next integer, far beyond the division's rounding, so the floor of the
   IEEE double quotient is the integer quotient. */
__device__ __forceinline__ int resize_pos(int i, int S, int D)
{
    const double num = (double)((long long)(2 * i + 1) * S - D) * 256.0;
    const double q = floor(__ddiv_rn(num, (double)(2 * D)));
    return (int)fmin(fmax(q, 0.0), (double)(S - 1) * 256.0);
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

/* level coordinate -> level-0 coordinate: rn(v * s^l), evaluated as the
   host tables are, floor(v * s^l + 0.5) in double without contraction */
__device__ __forceinline__ int to_l0(const FrameDev& F, const int* __restrict__ tabs, int l, int v, bool is_y)
{
    (void)tabs;
    (void)is_y;
    if (l == 0) return v;
    if (F.ref_mode) return v << l;
    return (int)floor(__dadd_rn(__dmul_rn((double)v, F.sc[l]), 0.5));
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

/* v of lane i - 1 (lane 0: 0) and of lane i + 1 (lane 63: 0): DPP
   wavefront shifts by one lane, a VALU move each (no LDS round trip) */
__device__ __forceinline__ uint32_t from_lane_below(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x138, 0xf, 0xf, true);   /* wave_shr:1 */
}
__device__ __forceinline__ uint32_t from_lane_above(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x130, 0xf, 0xf, true);   /* wave_shl:1 */
}

/* Loaded bytes of one level row for the 4 columns of a lane, kept apart from
   their use so the loads of later rows are in flight while a row is computed.
   MODE 0: level 0 from a dword-aligned input (dword loads + byte permute);
   MODE 1: byte gathers (level 0 of an unaligned input, or scale-2
   decimation); MODE 2: bilinear, one dword per column and source row;
   MODE 3: bilinear, one 8-byte window per source row holding the taps of all
   4 columns (every lane's taps within 8 bytes: scales up to 2); MODE 4:
   bilinear, two 8-byte windows per source row, for columns 0-1 and 2-3;
   MODE 5: level 0 where no lane of the wavefront reflects a column or owns
   row padding: the lane's own dword.  The batch stage is bound by the
   vector memory instructions per pixel, so a wavefront uses the mode with
   the fewest loads whose windows hold every lane's taps. */
template <int MODE>
struct PyrRow {
    uint32_t v[MODE == 0 ? 3 : (MODE == 3 ? 4 : (MODE == 5 ? 1 : 8))];
    int fy;
};

typedef unsigned short pyr_u16x2 __attribute__((ext_vector_type(2)));

/* Per-lane constants of a strip.  All memory goes through buffer resources:
   the row base is a wave-uniform scalar offset and the column a 32-bit
   vector offset, and an offset at or above the resource's size (kOob) reads
   0 and drops stores, which replaces branches around edge rows and lanes.
   kOob exceeds every level and image size (< 2^28 bytes), and a row offset
   and a column offset that are both kOob still sum below 2^32. */
constexpr uint32_t kOob = 0x40000000u;

typedef unsigned int v4u __attribute__((ext_vector_type(4)));
typedef unsigned int v2u __attribute__((ext_vector_type(2)));

struct PyrLane {
    __amdgpu_buffer_rsrc_t src;  /* the input image, h0 * ip bytes */
    int ip, l, h0;
    /* MODE 1: xa = source byte offset in a row, xb = the padding columns' own bytes.
       MODE 2: xa = offset of the source dword holding the column pair (x0, x1),
       xb = v_perm selector giving x0 | x1 << 16 from it, wx = (256 - fx) | fx << 16.
       MODE 3: xa[0] = offset of the lane's 8-byte window, xb = v_perm selectors
       of each column's pair in it, wx as MODE 2.  MODE 4: xa[0], xa[1] = the
       windows of columns 0-1 and 2-3, xb and wx as MODE 3 */
    uint32_t xa[4], xb[4], wx[4];
    uint32_t base4, sel, direct; /* MODE 0: reflect window start, byte selector, the lane's own dword */
    uint32_t padmask;            /* bytes of the lane's dword that lie in the row padding */
    uint32_t ipmask;             /* level 0: padding bytes that exist in the input row */
};

/* `live` false (rows past the walk): every load reads the out-of-range offset,
   which returns 0 without touching memory */
template <int MODE>
__device__ __forceinline__ void pyr_load(const PyrLane& P, int yr, int ty, bool live, PyrRow<MODE>& R)
{
    if constexpr (MODE == 3) {
        /* one (unaligned) 8-byte window per source row holds every tap */
        const int ya = ty >> 8, yb = min(ya + 1, P.h0 - 1);
        const uint32_t sa = live ? (uint32_t)(ya * P.ip) : kOob, sb = live ? (uint32_t)(yb * P.ip) : kOob;
        const auto a = __builtin_amdgcn_raw_buffer_load_b64(P.src, P.xa[0], sa, 0);
        const auto b = __builtin_amdgcn_raw_buffer_load_b64(P.src, P.xa[0], sb, 0);
        R.v[0] = a[0];
        R.v[1] = a[1];
        R.v[2] = b[0];
        R.v[3] = b[1];
        R.fy = ty & 255;
    } else if constexpr (MODE == 4) {
        const int ya = ty >> 8, yb = min(ya + 1, P.h0 - 1);
        const uint32_t sa = live ? (uint32_t)(ya * P.ip) : kOob, sb = live ? (uint32_t)(yb * P.ip) : kOob;
        const auto a0 = __builtin_amdgcn_raw_buffer_load_b64(P.src, P.xa[0], sa, 0);
        const auto a1 = __builtin_amdgcn_raw_buffer_load_b64(P.src, P.xa[1], sa, 0);
        const auto b0 = __builtin_amdgcn_raw_buffer_load_b64(P.src, P.xa[0], sb, 0);
        const auto b1 = __builtin_amdgcn_raw_buffer_load_b64(P.src, P.xa[1], sb, 0);
        R.v[0] = a0[0];
        R.v[1] = a0[1];
        R.v[2] = a1[0];
        R.v[3] = a1[1];
        R.v[4] = b0[0];
        R.v[5] = b0[1];
        R.v[6] = b1[0];
        R.v[7] = b1[1];
        R.fy = ty & 255;
    } else if constexpr (MODE == 2) {
        /* one (unaligned) dword per source row and column holds both taps */
        const int ya = ty >> 8, yb = min(ya + 1, P.h0 - 1);
        const uint32_t sa = live ? (uint32_t)(ya * P.ip) : kOob, sb = live ? (uint32_t)(yb * P.ip) : kOob;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            R.v[2 * k + 0] = __builtin_amdgcn_raw_buffer_load_b32(P.src, P.xa[k], sa, 0);
            R.v[2 * k + 1] = __builtin_amdgcn_raw_buffer_load_b32(P.src, P.xa[k], sb, 0);
        }
        R.fy = ty & 255;
    } else if constexpr (MODE == 1) {
        const uint32_t so = live ? (uint32_t)((yr << P.l) * P.ip) : kOob;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            R.v[k] = __builtin_amdgcn_raw_buffer_load_b8(P.src, P.xa[k], so, 0);
            R.v[4 + k] = __builtin_amdgcn_raw_buffer_load_b8(P.src, P.xb[k], so, 0);
        }
        R.fy = 0;
    } else if constexpr (MODE == 5) {
        const uint32_t so = live ? (uint32_t)(yr * P.ip) : kOob;
        R.v[0] = __builtin_amdgcn_raw_buffer_load_b32(P.src, P.direct, so, 0);
        R.fy = 0;
    } else {
        const uint32_t so = live ? (uint32_t)(yr * P.ip) : kOob;
        R.v[0] = __builtin_amdgcn_raw_buffer_load_b32(P.src, P.base4, so, 0);
        R.v[1] = __builtin_amdgcn_raw_buffer_load_b32(P.src, P.base4 + 4, so, 0);
        R.v[2] = __builtin_amdgcn_raw_buffer_load_b32(P.src, P.direct, so, 0);
        R.fy = 0;
    }
}

/* the 4 level pixels (reflect-101 columns) and the raw output dword (row
   padding: the input's bytes at level 0, 0 elsewhere) */
template <int MODE>
__device__ __forceinline__ void pyr_pixels(const PyrLane& P, const PyrRow<MODE>& R, uint32_t& hin, uint32_t& out)
{
    if constexpr (MODE == 2 || MODE == 3 || MODE == 4) {
        /* ((256-fy)((256-fx)a + fx b) + fy((256-fx)c + fx d) + 32768) >> 16 is
           exact integer arithmetic, so the rows are combined first, both
           columns at once in 16-bit lanes (each sum <= 255 * 256), then the
           columns by a dot product; the result (< 2^24) is byte 2 */
        const unsigned short wa = (unsigned short)(256 - R.fy), wb = (unsigned short)R.fy;
        uint32_t x[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            uint32_t ab, cd;
            if constexpr (MODE == 3) {
                ab = __builtin_amdgcn_perm(R.v[1], R.v[0], P.xb[k]);
                cd = __builtin_amdgcn_perm(R.v[3], R.v[2], P.xb[k]);
            } else if constexpr (MODE == 4) {
                const int wi = 2 * (k >> 1);
                ab = __builtin_amdgcn_perm(R.v[wi + 1], R.v[wi], P.xb[k]);
                cd = __builtin_amdgcn_perm(R.v[4 + wi + 1], R.v[4 + wi], P.xb[k]);
            } else {
                ab = __builtin_amdgcn_perm(R.v[2 * k], R.v[2 * k], P.xb[k]);
                cd = __builtin_amdgcn_perm(R.v[2 * k + 1], R.v[2 * k + 1], P.xb[k]);
            }
            const pyr_u16x2 col = __builtin_bit_cast(pyr_u16x2, ab) * wa + __builtin_bit_cast(pyr_u16x2, cd) * wb;
            x[k] = __builtin_amdgcn_udot2(col, __builtin_bit_cast(pyr_u16x2, P.wx[k]), 32768u, false);
        }
        hin = __builtin_amdgcn_perm(x[1], x[0], 0x0c0c0602u) | __builtin_amdgcn_perm(x[3], x[2], 0x06020c0cu);
        out = hin & ~P.padmask;
    } else if constexpr (MODE == 1) {
        hin = R.v[0] | (R.v[1] << 8) | (R.v[2] << 16) | (R.v[3] << 24);
        const uint32_t pad = R.v[4] | (R.v[5] << 8) | (R.v[6] << 16) | (R.v[7] << 24);
        out = (hin & ~P.padmask) | pad;
    } else if constexpr (MODE == 5) {
        hin = R.v[0];
        out = hin;
    } else {
        hin = __builtin_amdgcn_perm(R.v[1], R.v[0], P.sel);
        out = (hin & ~P.padmask) | (R.v[2] & P.padmask & P.ipmask);
    }
}

/* One wavefront builds a strip of kPyrStripW columns x TH rows of one level:
   lane i owns the 4 columns xs - 4 + 4i .. +3 (lanes 1..60 own the strip;
   lanes 0 and 61 only feed the horizontal taps of their neighbours).  The lane walks rows ys-3 .. ys+TH+2
   (reflect-101 at the level's top and bottom): per row it forms the 4 level
   pixels (copy of the input, scale-2 decimation, or 8-bit bilinear from the
   input), takes its neighbours' pixels across lanes, runs the horizontal
   7-tap pass and keeps the last 8 results in registers, from which the
   vertical pass emits the blurred row 3 above.  The loads of a row are issued
   three rows ahead.  The raw and blurred dwords of 4 rows are gathered in the
   wavefront's LDS rows and leave as one 16-byte store per lane and output
   (15 lanes per row), a quarter of the store instructions of dword stores;
   stores outside the strip's rows or columns go to kOob.
   Strips cover the whole pitch: row padding is written as the input's bytes
   at level 0 and as 0 elsewhere, in both the raw and the blurred level
   (orbd.cu:514-535, 1273-1338; blur out of place). */
template <int MODE, int TH>
__device__ __forceinline__ void pyr_walk(const PyrLane& P, const Tables& T, int ys, int h, int p,
                                         __amdgpu_buffer_rsrc_t raw, __amdgpu_buffer_rsrc_t blur, int tyv,
                                         int xs, uint32_t* __restrict__ s_rows)
{
    const int lane = threadIdx.x & 63;
    /* gather rows: lane i's dword of row q at s_rows[q * 64 + i - 1] (lanes
       1..60 at 0..59, the halo lanes in 60..63); the 16-byte store of lane m
       (< 60) takes chunk m % 15 of row m / 15 */
    uint32_t* s_raw = s_rows;
    uint32_t* s_blur = s_rows + 256;
    const int wslot = (lane - 1) & 63;
    const int gq = lane / 15, gc = lane - 15 * gq;
    const int gcol = xs + 16 * gc;
    const bool gok = lane < 60 && gcol < p;
    const int grd = gok ? 64 * gq + 4 * gc : 0;
    /* rows walked: the strip's rows plus 3 above and below, in steps of 8 */
    const int nwalk = min(TH, h - ys) + 6;
    const int row_end = min(ys + TH, h);
    auto row_ty = [&](int i) { return __builtin_amdgcn_readlane(tyv, min(i, nwalk - 1)); };
    auto row_y = [&](int i) { return reflect101(ys - 3 + min(i, nwalk - 1), h); };
    /* walked row i is loaded into slot i % 4, three rows before it is used;
       the loop is unrolled by 8, so every slot and ring entry is a fixed
       register set, and at most 3 rows of loads and stores (38 vector memory
       operations) are outstanding, inside what the wait counter can track */
    PyrRow<MODE> st[4];
#pragma unroll
    for (int r = 0; r < 3; r++) {
        /* in row order, as the waits in the loop assume */
        __builtin_amdgcn_sched_barrier(0);
        pyr_load<MODE>(P, row_y(r), row_ty(r), r < nwalk, st[r]);
    }
    float4 ring[8];
    uint32_t outring[8];
    for (int i0 = 0; i0 < nwalk; i0 += 8) {
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int i = i0 + j;
            /* keep the prefetch where it is written: the scheduler would sink it */
            __builtin_amdgcn_sched_barrier(0);
            pyr_load<MODE>(P, row_y(i + 3), row_ty(i + 3), i + 3 < nwalk, st[(j + 3) % 4]);
            __builtin_amdgcn_sched_barrier(0);
            uint32_t hin, rawout;
            pyr_pixels<MODE>(P, st[j % 4], hin, rawout);
            const uint32_t lf = from_lane_below(hin), rt = from_lane_above(hin);
            const int rr = ys - 3 + i;
            outring[j] = rawout;
            s_raw[64 * (j & 3) + wslot] = rawout;
            /* horizontal pass: q1..q10 = columns cx-3 .. cx+6, each byte
               converted to float once; sums of two integers <= 255 are exact
               in FP32, so they equal the converted integer sums.  The empty
               asm keeps the compiler from folding the float sums back into
               integer sums converted one by one (two instructions a sum) */
            float q[12];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                q[k] = (float)((lf >> (8 * k)) & 255);
                q[4 + k] = (float)((hin >> (8 * k)) & 255);
                q[8 + k] = (float)((rt >> (8 * k)) & 255);
            }
#pragma unroll
            for (int k = 1; k <= 10; k++) asm("" : "+v"(q[k]));
            float4 hv;
            float* hf = (float*)&hv;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                float v = __fmul_rn(T.k0, q[k + 4]);
                v = __fmaf_rn(T.k1, __fadd_rn(q[k + 3], q[k + 5]), v);
                v = __fmaf_rn(T.k2, __fadd_rn(q[k + 2], q[k + 6]), v);
                v = __fmaf_rn(T.k3, __fadd_rn(q[k + 1], q[k + 7]), v);
                hf[k] = v;
            }
            ring[j] = hv;
            /* vertical pass for row rr - 3 from ring rows rr-6 .. rr */
            const int o = rr - 3;
            const float* c0 = (const float*)&ring[(j + 2) % 8];
            const float* c1 = (const float*)&ring[(j + 3) % 8];
            const float* c2 = (const float*)&ring[(j + 4) % 8];
            const float* c3 = (const float*)&ring[(j + 5) % 8];
            const float* c4 = (const float*)&ring[(j + 6) % 8];
            const float* c5 = (const float*)&ring[(j + 7) % 8];
            const float* c6 = (const float*)&ring[j];
            /* 0 <= a < 255.5: a + 1.5 * 2^23 rounds a half to even into the
               low mantissa byte, which is the blurred pixel */
            uint32_t rb[4];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                float a = __fmul_rn(T.k0, c3[k]);
                a = __fmaf_rn(T.k1, __fadd_rn(c2[k], c4[k]), a);
                a = __fmaf_rn(T.k2, __fadd_rn(c1[k], c5[k]), a);
                a = __fmaf_rn(T.k3, __fadd_rn(c0[k], c6[k]), a);
                rb[k] = __float_as_uint(__fadd_rn(a, 12582912.0f));
            }
            uint32_t out = __builtin_amdgcn_perm(rb[1], rb[0], 0x0c0c0400u) | __builtin_amdgcn_perm(rb[3], rb[2], 0x04000c0cu);
            /* padding columns of the blurred row keep the raw row's bytes */
            out = (out & ~P.padmask) | (outring[(j + 5) % 8] & P.padmask);
            s_blur[64 * (j & 3) + wslot] = out;
            if ((j & 3) == 3) {
                /* raw rows rr - 3 .. rr and blurred rows o - 3 .. o */
                const v4u rv = *(const v4u*)&s_raw[grd];
                const v4u bv = *(const v4u*)&s_blur[grd];
                const int gr = rr - 3 + gq, gb = o - 3 + gq;
                const uint32_t ro = gok && gr >= ys && gr < row_end ? (uint32_t)(gr * p + gcol) : kOob;
                const uint32_t bo = gok && gb >= ys && gb < row_end ? (uint32_t)(gb * p + gcol) : kOob;
                __builtin_amdgcn_raw_buffer_store_b128(rv, raw, ro, 0, 0);
                __builtin_amdgcn_raw_buffer_store_b128(bv, blur, bo, 0, 0);
            }
        }
    }
}

template <int TH>
__global__ __launch_bounds__(256) void k_pyramid(const FrameDev* __restrict__ frames, const Job* __restrict__ jobs,
                                                 const int* __restrict__ job_of, int nunits, Tables T)
{
    static_assert(TH + 6 <= 64, "one lane per walked row holds its bilinear row entry");
    static_assert((TH + 6) % 8 == 0, "a full strip walks whole unrolled groups of 8 rows");
    /* per wavefront: 4 gathered rows of raw and of blurred dwords */
    __shared__ __align__(16) uint32_t s_gather[4 * 512];
    const int lane = threadIdx.x & 63;
    const int u = __builtin_amdgcn_readfirstlane(xcd_remap(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6));
    if (u >= nunits) return;
    const Job jb = find_job(jobs, job_of, u);
    const FrameDev& F = frames[jb.frame];
    const int l = jb.level;
    const int t = u - jb.first_block;
    const int xs = (t % jb.tiles_x) * kPyrStripW, ys = (jb.row0 + t / jb.tiles_x) * TH;
    const int w = F.lw[l], h = F.lh[l], p = F.lp[l];
    const int cx = xs - 4 + 4 * lane;
    PyrLane P;
    P.ip = F.ip;
    P.l = l;
    P.h0 = F.h0;
    P.src = __builtin_amdgcn_make_buffer_rsrc((void*)F.img, 0, F.h0 * F.ip, 0x00020000);
    const bool aligned = (F.ip & 3) == 0 && (((uintptr_t)F.img) & 3) == 0;
    int mode = l == 0 ? (aligned ? 0 : 1) : (F.ref_mode ? 1 : 2);
    P.padmask = 0;
    P.ipmask = 0;
    int xr[4], x0s[4], fxs[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int x = cx + k;
        xr[k] = reflect101(x, w);   /* the reflect-101 column feeds the horizontal taps */
        if (x >= w) P.padmask |= 0xffu << (8 * k);
        if (l == 0 && x >= w && x < F.ip) P.ipmask |= 0xffu << (8 * k);
        if (mode == 2) {
            /* the dword at min(x0, w0 - 4) ends inside the input row; at x0 = w0 - 1
               both taps are that pixel, which fx = 0 gives exactly */
            const int v = resize_pos(xr[k], F.w0, w);
            const int x0 = v >> 8, fx = x0 == F.w0 - 1 ? 0 : (v & 255);
            const int d = min(x0, F.w0 - 4), sh = x0 - d;
            x0s[k] = x0;
            fxs[k] = fx;
            P.xa[k] = (uint32_t)d;
            P.xb[k] = (uint32_t)sh | 0x0c00u | (uint32_t)(sh + 1) << 16 | 0x0c000000u;
            P.wx[k] = (uint32_t)(256 - fx) | (uint32_t)fx << 16;
        } else {
            P.xa[k] = (uint32_t)(xr[k] << (mode == 1 ? l : 0));
            /* MODE 1 padding bytes (level 0 only): the input byte at the column itself */
            P.xb[k] = (l == 0 && x >= w && x < F.ip) ? (uint32_t)x : kOob;
            P.wx[k] = 0;
        }
    }
    if (mode == 2 && F.w0 >= 8) {
        /* MODE 3 when every lane's taps lie in the 8 bytes from
           min(its lowest tap, w0 - 8) (inside the input row), else MODE 4
           when those of columns 0-1 and of 2-3 each lie in such a window; a
           second tap past a window only occurs at x0 = w0 - 1, where fx = 0
           ignores it */
        for (int nw = 1; nw <= 2 && mode == 2; nw++) {
            const int per = 4 / nw;
            bool fits = true;
            uint32_t sel[4], st[2];
#pragma unroll
            for (int wi = 0; wi < 2; wi++) {
                int lo = x0s[wi * 2];
#pragma unroll
                for (int k = 0; k < 4; k++)
                    if (k / per == wi) lo = min(lo, x0s[k]);
                st[wi] = (uint32_t)min(lo, F.w0 - 8);
            }
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int o = x0s[k] - (int)st[k / per];
                fits = fits && o <= 7 && (o <= 6 || fxs[k] == 0);
                sel[k] = (uint32_t)o | 0x0c00u | (uint32_t)min(o + 1, 7) << 16 | 0x0c000000u;
            }
            if (__ballot(!fits) == 0) {
                mode = nw == 1 ? 3 : 4;
                P.xa[0] = st[0];
                P.xa[1] = st[1];
#pragma unroll
                for (int k = 0; k < 4; k++) P.xb[k] = sel[k];
            }
        }
    }
    if (mode == 0) {
        const int lo = min(min(xr[0], xr[1]), min(xr[2], xr[3]));
        P.base4 = (uint32_t)(lo & ~3);
        P.sel = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) P.sel |= (uint32_t)(xr[k] - (lo & ~3)) << (8 * k);
        P.direct = cx >= 0 ? (uint32_t)cx : kOob;
        /* MODE 5: every lane's 4 columns are its own dword (no reflection,
           no row padding) */
        const bool own = xr[0] == cx && xr[3] == cx + 3 && P.padmask == 0;
        if (__ballot(!own) == 0) mode = 5;
    }
    /* lane i holds the bilinear row entry of walked row i */
    const int tyv = (mode >= 2 && mode <= 4) && lane < TH + 6 ? resize_pos(reflect101(ys - 3 + lane, h), F.h0, h) : 0;
    const int lbytes = h * p;
    __amdgpu_buffer_rsrc_t raw = __builtin_amdgcn_make_buffer_rsrc((void*)(F.raw + F.loff[l]), 0, lbytes, 0x00020000);
    __amdgpu_buffer_rsrc_t blur = __builtin_amdgcn_make_buffer_rsrc((void*)(F.blur + F.loff[l]), 0, lbytes, 0x00020000);
    uint32_t* s_rows = s_gather + 512 * (threadIdx.x >> 6);
    if (mode == 0) pyr_walk<0, TH>(P, T, ys, h, p, raw, blur, tyv, xs, s_rows);
    else if (mode == 1) pyr_walk<1, TH>(P, T, ys, h, p, raw, blur, tyv, xs, s_rows);
    else if (mode == 3) pyr_walk<3, TH>(P, T, ys, h, p, raw, blur, tyv, xs, s_rows);
    else if (mode == 4) pyr_walk<4, TH>(P, T, ys, h, p, raw, blur, tyv, xs, s_rows);
    else if (mode == 5) pyr_walk<5, TH>(P, T, ys, h, p, raw, blur, tyv, xs, s_rows);
    else pyr_walk<2, TH>(P, T, ys, h, p, raw, blur, tyv, xs, s_rows);
}

/* -------------------------------------------------------------- detection */


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

/* detection tile in LDS: rows y0-4 .. y0+TH+4, columns x0-4 .. x0+TW+4 of the
   flat pyramid (so reads past a level's edge see exactly the bytes the flat
   layout puts there) */
constexpr int kDetSW = kDetTW + 8, kDetSH = kDetTH + 8;

/* circle offsets (orbd.cu:466-469) as LDS offsets in the staged tile */
#define RING_OFF(dx, dy) ((dy) * kDetSW + (dx))
__device__ __forceinline__ int ring_off(int k)
{
    switch (k) {
    case 0: return RING_OFF(0, 3);   case 1: return RING_OFF(1, 3);   case 2: return RING_OFF(2, 2);
    case 3: return RING_OFF(3, 1);   case 4: return RING_OFF(3, 0);   case 5: return RING_OFF(3, -1);
    case 6: return RING_OFF(2, -2);  case 7: return RING_OFF(1, -3);  case 8: return RING_OFF(0, -3);
    case 9: return RING_OFF(-1, -3); case 10: return RING_OFF(-2, -2); case 11: return RING_OFF(-3, -1);
    case 12: return RING_OFF(-3, 0); case 13: return RING_OFF(-3, 1); case 14: return RING_OFF(-2, 2);
    default: return RING_OFF(-1, 3);
    }
}

/* Harris response of orbd.cu:651-669 from the sums of the Sobel products
   over the 7x7 window: 32-bit wrapping products, then
   fma(-0.04, B, A) * (1/(4*7*255))^4.  The sums are of 49 integer terms each
   (every |dx|, |dy| <= 1020), so their order does not change them. */
__device__ __forceinline__ float harris_from_sums(int dxx, int dyy, int dxy, float scale)
{
    const int A = (int)((unsigned)dxx * (unsigned)dyy - (unsigned)dxy * (unsigned)dxy);
    const unsigned S = (unsigned)(dxx + dyy);
    const int B = (int)(S * S);
    return __fmul_rn(__fmaf_rn(-0.04f, __int2float_rn(B), __int2float_rn(A)), scale);
}

/* exclusive scan of one int per thread over a workgroup of NT threads */
template <int NT>
__device__ __forceinline__ int block_scan(int v, int* s_tmp, int& total)
{
    constexpr int NW = NT / 64;
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
        int q = lane < NW ? s_tmp[lane] : 0;
#pragma unroll
        for (int o = 1; o < NW; o <<= 1) {
            const int y = __shfl_up(q, o, 64);
            if (lane >= o) q += y;
        }
        if (lane < NW) s_tmp[NW + lane] = q;
    }
    __syncthreads();
    total = s_tmp[2 * NW - 1];
    const int res = x - v + (wid > 0 ? s_tmp[NW + wid - 1] : 0);
    __syncthreads();
    return res;
}

/* Slot of this lane in a list that the wavefront appends to: one LDS atomic
   per wavefront, lanes ordered by lane id; -1 where pred is false. */
__device__ __forceinline__ int wave_append(bool pred, int* counter)
{
    const unsigned long long m = __ballot(pred);
    if (m == 0) return -1;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((long long)m) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(counter, __popcll(m));
    base = __shfl(base, leader, 64);
    const int before = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0));
    return pred ? base + before : -1;
}

/* FAST-9 + score of one detection tile, in three passes over the staged
   level bytes:
   1. every pixel: the compass test.  A contiguous 9-arc of the 16-ring
      contains two compass points (ring 0/4/8/12) that are 4 apart, so a
      corner has two such neighbouring compass points both darker than c-t
      or both brighter than c+t; the rest are rejected here.
   2. the compass survivors, compacted: the exact segment test of
      orbd.cu:685-740 (quick test on opposite pairs, >8 consecutive of the
      wrapped ring, darker arc first and exclusively).
   3. the corners, compacted: Harris (or FAST) score and the 64-bit max into
      the level-0 key map (orbd.cu:742-755). */
__global__ __launch_bounds__(256) void k_detect(const FrameDev* __restrict__ frames, const Job* __restrict__ jobs,
                                                const int* __restrict__ job_of, Tables T,
                                                const uint32_t* __restrict__ epochs)
{
    __shared__ __align__(16) uint8_t s_tile[kDetSH * kDetSW];
    __shared__ uint16_t s_cand[kDetTW * kDetTH];
    __shared__ uint16_t s_corner[kDetTW * kDetTH];
    __shared__ int s_n[2];
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    const Job jb = find_job(jobs, job_of, b);
    const FrameDev& F = frames[jb.frame];
    const int l = jb.level;
    const int t = b - jb.first_block;
    const int x0 = (t % jb.tiles_x) * kDetTW, y0 = (t / jb.tiles_x) * kDetTH;
    const int w = F.lw[l], h = F.lh[l], p = F.lp[l];
    const int tid = threadIdx.x;
    if (tid < 2) s_n[tid] = 0;
    /* stage kDetSH rows x kDetSW bytes as dwords (every row start is dword
       aligned) through a buffer resource at the tile's first byte, so every
       address is a 32-bit offset: all loads first, then the LDS writes.  The
       tile lies inside the flat pyramid and its guard rows. */
    {
        const uint8_t* base = F.raw + F.loff[l] + (long long)(y0 - 4) * p + (x0 - 4);
        const __amdgpu_buffer_rsrc_t rs =
            __builtin_amdgcn_make_buffer_rsrc((void*)base, 0, kDetSH * p, 0x00020000);
        constexpr int dw_per_row = kDetSW / 4, ndw = kDetSH * dw_per_row, per = (ndw + 255) / 256;
        uint32_t v[per];
#pragma unroll
        for (int k = 0; k < per; k++) {
            const int i = tid + 256 * k;
            const int r = i / dw_per_row, c = i - r * dw_per_row;
            if (i < ndw) v[k] = __builtin_amdgcn_raw_buffer_load_b32(rs, (uint32_t)(r * p + 4 * c), 0, 0);
        }
#pragma unroll
        for (int k = 0; k < per; k++) {
            const int i = tid + 256 * k;
            if (i < ndw) *(uint32_t*)&s_tile[4 * i] = v[k];
        }
    }
    __syncthreads();
    const int thr = T.threshold;
    /* pass 1: compass test on every pixel of the tile, 4 pixels (one LDS
       dword) at a time, then one workgroup-wide compaction.  With q0, q4, q8,
       q12 the compass points (ring (0,3), (3,0), (0,-3), (-3,0)), "two
       neighbouring compass points both darker than v-t" is
       max(min(q0,q8), min(q4,q12)) < v-t and "both brighter than v+t" is
       min(max(q0,q8), max(q4,q12)) > v+t; the even and the odd bytes are
       tested in the two 16-bit lanes of packed integer operations, where a
       difference's sign bit is the comparison. */
    int ncand;
    {
        constexpr int kG = kDetTW / 4, kStep = 256 / kG, kRows = kDetTH / kStep;
        static_assert(kRows * 4 <= 32, "a bit per pixel of the thread's rows");
        const int lx = 4 * (tid & (kG - 1));
        uint32_t colm = 0;
#pragma unroll
        for (int j = 0; j < 4; j++) colm |= (uint32_t)(x0 + lx + j >= 3 && x0 + lx + j < w - 3) << j;
        const pyr_u16x2 t2 = {(unsigned short)thr, (unsigned short)thr};
        uint32_t m = 0;
#pragma unroll
        for (int rr = 0; rr < kRows; rr++) {
            const int ly = tid / kG + kStep * rr;
            const int iy = y0 + ly;
            const uint8_t* c = &s_tile[(ly + 4) * kDetSW + lx + 4];
            const uint32_t C = *(const uint32_t*)c;
            const uint32_t Lw = *(const uint32_t*)(c - 4), Rw = *(const uint32_t*)(c + 4);
            const uint32_t U = *(const uint32_t*)(c - 3 * kDetSW), D = *(const uint32_t*)(c + 3 * kDetSW);
            const uint32_t L = __builtin_amdgcn_alignbyte(C, Lw, 1);  /* bytes at x - 3 */
            const uint32_t R = __builtin_amdgcn_alignbyte(Rw, C, 3);  /* bytes at x + 3 */
            uint32_t hit[2];
#pragma unroll
            for (int half = 0; half < 2; half++) {
                /* bytes 0,2 (half 0) or 1,3 (half 1) as two 16-bit lanes */
                const uint32_t sel = half ? 0x0c030c01u : 0x0c020c00u;
                const pyr_u16x2 v = __builtin_bit_cast(pyr_u16x2, __builtin_amdgcn_perm(C, C, sel));
                const pyr_u16x2 q0 = __builtin_bit_cast(pyr_u16x2, __builtin_amdgcn_perm(D, D, sel));
                const pyr_u16x2 q4 = __builtin_bit_cast(pyr_u16x2, __builtin_amdgcn_perm(R, R, sel));
                const pyr_u16x2 q8 = __builtin_bit_cast(pyr_u16x2, __builtin_amdgcn_perm(U, U, sel));
                const pyr_u16x2 q12 = __builtin_bit_cast(pyr_u16x2, __builtin_amdgcn_perm(L, L, sel));
                const pyr_u16x2 mn = __builtin_elementwise_max(__builtin_elementwise_min(q0, q8),
                                                               __builtin_elementwise_min(q4, q12));
                const pyr_u16x2 mx = __builtin_elementwise_min(__builtin_elementwise_max(q0, q8),
                                                               __builtin_elementwise_max(q4, q12));
                /* mn + t - v < 0 (darker), v + t - mx < 0 (brighter): 16-bit
                   differences in [-255, 510], sign bits 15 and 31 */
                const uint32_t dk = __builtin_bit_cast(uint32_t, (pyr_u16x2)(mn + t2 - v));
                const uint32_t br = __builtin_bit_cast(uint32_t, (pyr_u16x2)(v + t2 - mx));
                hit[half] = (dk | br) & 0x80008000u;
            }
            /* pixel j of the dword -> bit j */
            const uint32_t tb = (hit[0] >> 15) | (hit[1] >> 14);
            const uint32_t nib = (tb | (tb >> 14)) & 15u;
            m |= (iy >= 3 && iy < h - 3 ? nib & colm : 0u) << (4 * rr);
        }
        /* each wavefront reserves its candidates' slots with one LDS add
           (the order of the list does not matter) */
        const int cnt = __popc(m);
        const int incl = wave_incl_scan(cnt);
        int base = 0;
        if ((tid & 63) == 63) base = atomicAdd(&s_n[0], incl);
        int off = __builtin_amdgcn_readlane(base, 63) + incl - cnt;
        for (; m; m &= m - 1) {
            const int bit = __ffs(m) - 1;
            s_cand[off++] = (uint16_t)(((tid / kG + kStep * (bit >> 2)) << 8) | (lx + (bit & 3)));
        }
    }
    __syncthreads();
    ncand = s_n[0];
    /* pass 2: exact segment test on the candidates */
    for (int i0 = 0; i0 < ncand; i0 += 256) {
        const int i = i0 + tid;
        bool corner = false;
        int e = 0;
        if (i < ncand) {
            e = s_cand[i];
            const uint8_t* c = &s_tile[((e >> 8) + 4) * kDetSW + (e & 255) + 4];
            const int v = c[0];
            /* ring points k and k + 8 in the two 16-bit lanes of one word;
               q - (v - t) < 0 (darker) and (v + t) - q < 0 (brighter) are the
               sign bits of 16-bit differences in [-255, 510], moved to bits
               k and 16 + k, then the high half down to bits 8..15 */
            const pyr_u16x2 vm = {(unsigned short)(v - thr), (unsigned short)(v - thr)};
            const pyr_u16x2 vp = {(unsigned short)(v + thr), (unsigned short)(v + thr)};
            uint32_t dk = 0, bk = 0;
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const pyr_u16x2 q = {(unsigned short)c[ring_off(k)], (unsigned short)c[ring_off(k + 8)]};
                const uint32_t sd = __builtin_bit_cast(uint32_t, (pyr_u16x2)(q - vm));
                const uint32_t sb = __builtin_bit_cast(uint32_t, (pyr_u16x2)(vp - q));
                dk |= (sd >> (15 - k)) & (0x00010001u << k);
                bk |= (sb >> (15 - k)) & (0x00010001u << k);
            }
            const unsigned dark = (dk & 0xffu) | ((dk >> 8) & 0xff00u);
            const unsigned bright = (bk & 0xffu) | ((bk >> 8) & 0xff00u);
            /* quick test of orbd.cu:687-700, then >8 consecutive of the wrapped
               ring (a circular run of 9) on the darker arc first, exclusively
               (orbd.cu:703-740) */
            const bool d1 = ((dark | (dark >> 8)) & 0xffu) == 0xffu;
            const bool d2 = ((bright | (bright >> 8)) & 0xffu) == 0xffu;
            const unsigned m = d1 ? dark : (d2 ? bright : 0u);
            const unsigned r2 = m | (m << 16);
            unsigned run = r2;
#pragma unroll
            for (int k = 1; k < 9; k++) run &= r2 >> k;
            corner = run != 0;
        }
        const int slot = wave_append(corner, &s_n[1]);
        if (slot >= 0) s_corner[slot] = (uint16_t)e;
    }
    __syncthreads();
    /* pass 3: score the corners into the key map */
    const int n = s_n[1];
    if (n == 0) return;
    const unsigned long long ep = (unsigned long long)epochs[jb.frame] << kEpochShift;
    const auto emit = [&](int e, float score) {
        const int ix = x0 + (e & 255), iy = y0 + (e >> 8);
        const int X0 = to_l0(F, T.tabs, l, ix, false), Y0 = to_l0(F, T.tabs, l, iy, true);
        const unsigned long long key = ep | ((unsigned long long)__float_as_uint(score) << 3) |
                                       (unsigned long long)(7 - l);
        __hip_atomic_fetch_max(as_global(F.keys) + (long long)Y0 * F.p0 + X0, key, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_fetch_or(as_global(F.cbits) + (long long)Y0 * (F.p0 >> 5) + (X0 >> 5), 1u << (X0 & 31),
                              __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    };
    if (T.score_type == 0) {
        /* Harris, 8 lanes per corner and 8 corners per wavefront pass (a tile
           holds a few corners, so one lane per corner would leave most lanes
           idle through the 7x7 window): lane r < 7 of a corner takes window
           row r - 3, forms the column sums u + 2m + d and d - u of its 9
           columns once and the Sobel dx, dy of the row's 7 positions from
           them, then the 8 lanes sum the products (exact integer sums) */
        const int lane = tid & 63, wid = tid >> 6, wr = lane & 7;
        for (int g = 8 * wid; g < n; g += 32) {
            const int ci = g + (lane >> 3);
            int sxx = 0, syy = 0, sxy = 0, e = 0;
            if (ci < n) e = s_corner[ci];
            if (ci < n && wr < 7) {
                /* the 9 bytes of columns -4..4 of rows r-1, r, r+1, read as
                   the 3 aligned dwords around them and shifted into place */
                const int a = ((e >> 8) + 4 + wr - 3) * kDetSW + (e & 255);
                const int sh = 8 * (a & 3);
                uint32_t x[3][3];
#pragma unroll
                for (int rr = 0; rr < 3; rr++) {
                    const uint32_t* w = (const uint32_t*)&s_tile[(a & ~3) + (rr - 1) * kDetSW];
                    const uint32_t w0 = w[0], w1 = w[1], w2 = w[2];
                    x[rr][0] = (uint32_t)(((uint64_t)w1 << 32 | w0) >> sh);
                    x[rr][1] = (uint32_t)(((uint64_t)w2 << 32 | w1) >> sh);
                    x[rr][2] = w2 >> sh;
                }
                int hs[9], vs[9];
#pragma unroll
                for (int k = 0; k < 9; k++) {
                    const int u = (x[0][k >> 2] >> (8 * (k & 3))) & 255;
                    const int m = (x[1][k >> 2] >> (8 * (k & 3))) & 255;
                    const int d = (x[2][k >> 2] >> (8 * (k & 3))) & 255;
                    hs[k] = u + 2 * m + d;
                    vs[k] = d - u;
                }
#pragma unroll
                for (int x = 0; x < 7; x++) {
                    const int dx = hs[x + 2] - hs[x];
                    const int dy = vs[x] + 2 * vs[x + 1] + vs[x + 2];
                    sxx += __mul24(dx, dx);
                    syy += __mul24(dy, dy);
                    sxy += __mul24(dx, dy);
                }
            }
            sxx = half_row_sum8(sxx);
            syy = half_row_sum8(syy);
            sxy = half_row_sum8(sxy);
            if (ci < n && wr == 0) {
                const float score = harris_from_sums(sxx, syy, sxy, T.harris_scale);
                if (score > 0.f) emit(e, score);
            }
        }
    } else {
        for (int i = tid; i < n; i += 256) {
            const int e = s_corner[i];
            const uint8_t* c = &s_tile[((e >> 8) + 4) * kDetSW + (e & 255) + 4];
            int ring[16];
#pragma unroll
            for (int k = 0; k < 16; k++) ring[k] = c[ring_off(k)];
            const float score = fast_score(ring, c[0], thr);
            if (score > 0.f) emit(e, score);
        }
    }
}

/* -------------------------------------------------------------------- NMS */


/* NMS of orbd.cu:945-966: a key survives unless a live neighbour in its
   (2r+1)^2 square (r the radius of the key's own level) is greater, or equal
   at dx <= 0 && dy <= 0, with the reference's flat indexing (pitch p0).
   Detection marks every scored corner in a bitmap of the level-0 pixels, so
   the NMS only visits marked pixels: one wavefront per 2048-pixel row
   segment reads its 64 bitmap words (and clears them for the next call) and
   compacts the marked pixels in x order.  A segment holds a few marked
   pixels, so kNmsLanes lanes share one: lane `sub` of the group checks
   square rows dy = sub - r, sub - r + kNmsLanes, ..., the group's verdicts
   are combined by a ballot, and the survivors are written in x order.  One
   neighbourhood row at a time per lane keeps the wavefront small (many
   wavefronts in flight to cover the scattered key reads); 4 lanes per pixel
   measured best against 8 and 16 on both the 1080p pair and the 128-frame
   batch. */
constexpr int kNmsLanes = 4;

/* Row dy of the square around x: suppressed by a neighbour at |dx| <= r?
   (R >= r is the largest radius among the wavefront's pixels, so the R
   loads are unrolled and issued together) */
template <int R>
__device__ __forceinline__ bool nms_row_suppressed(const ORB_GLOBAL unsigned long long* __restrict__ row, int x,
                                                   int dy, uint32_t s, uint32_t epoch, int r)
{
    unsigned long long nk[2 * R + 1];
#pragma unroll
    for (int dx = -R; dx <= R; dx++) nk[dx + R] = row[x + dx];
    bool sup = false;
#pragma unroll
    for (int dx = -R; dx <= R; dx++) {
        const uint32_t ns = (uint32_t)(nk[dx + R] >> 3) & 0x7fffffffu;
        const bool live = (uint32_t)(nk[dx + R] >> kEpochShift) == epoch;
        const bool self = dx == 0 && dy == 0;
        sup |= !self && abs(dx) <= r && live && (ns > s || (ns == s && dx <= 0 && dy <= 0));
    }
    return sup;
}

/* the same for any radius (scale-2 octaves >= 3 have radii above 4) */
__device__ __forceinline__ bool nms_row_suppressed_any(const ORB_GLOBAL unsigned long long* __restrict__ row, int x,
                                                       int dy, uint32_t s, uint32_t epoch, int r)
{
    bool sup = false;
    for (int dx = -r; dx <= r; dx++) {
        if (dx == 0 && dy == 0) continue;
        const unsigned long long nk = row[x + dx];
        const uint32_t ns = (uint32_t)(nk >> 3) & 0x7fffffffu;
        sup |= (uint32_t)(nk >> kEpochShift) == epoch && (ns > s || (ns == s && dx <= 0 && dy <= 0));
    }
    return sup;
}

__global__ __launch_bounds__(256) void k_nms(const FrameDev* __restrict__ frames, const int* __restrict__ unit_frame,
                                             int nunits, const uint32_t* __restrict__ epochs,
                                             Candidate* __restrict__ cand, int* __restrict__ seg_count)
{
    __shared__ uint16_t s_x[4][kNmsSeg];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    /* consecutive rows on one XCD: their shared neighbour rows hit its L2 */
    const int u = xcd_remap(blockIdx.x, gridDim.x) * 4 + wid;
    if (u >= nunits) return;
    const int f = unit_frame[u];
    const FrameDev& F = frames[f];
    const int lu = u - F.nms_first;
    const int row = lu / F.nms_segs, c = lu - row * F.nms_segs;
    const int y = F.border + row;
    const int wpr = F.p0 >> 5;
    const int w = c * (kNmsSeg / 32) + lane;
    uint32_t bits = 0;
    if (w < wpr) {
        ORB_GLOBAL uint32_t* bw = as_global(F.cbits) + (long long)y * wpr + w;
        bits = *bw;
        if (bits) *bw = 0u;
    }
    /* pixels inside [border, w0 - border) only */
    const int x0 = w * 32;
    const int lo = max(F.border - x0, 0), hi = min(F.w0 - F.border - x0, 32);
    bits = hi <= lo ? 0u : bits & ((hi == 32 ? 0xffffffffu : ((1u << hi) - 1u)) & ~((1u << lo) - 1u));
    /* marked pixels of the segment, in x order, into LDS */
    const int nl = __popc(bits);
    int incl = nl;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(incl, o, 64);
        if (lane >= o) incl += t;
    }
    const int total = __shfl(incl, 63, 64);
    int pos = incl - nl;
    for (uint32_t b = bits; b; b &= b - 1) s_x[wid][pos++] = (uint16_t)(x0 + __ffs(b) - 1);
    wave_sync_lds();
    const uint32_t epoch = epochs[f];
    const ORB_GLOBAL unsigned long long* rowp = as_global(F.keys) + (long long)y * F.p0;
    ORB_GLOBAL Candidate* out = as_global(cand) + (long long)u * kNmsCap;
    const int g = lane / kNmsLanes, sub = lane % kNmsLanes;
    int nout = 0;
    for (int c0 = 0; c0 < total; c0 += 64) {
        /* one lane per marked pixel of the chunk: its key, score and radius
           (radius 0: no live key) */
        const int nchunk = min(64, total - c0);
        int xj = 0, rj = 0;
        uint32_t sj = 0, posj = 0;
        if (lane < nchunk) {
            xj = s_x[wid][c0 + lane];
            const unsigned long long key = rowp[xj];
            if ((uint32_t)(key >> kEpochShift) == epoch) {
                sj = (uint32_t)(key >> 3) & 0x7fffffffu;
                const uint32_t lv = 7u - (uint32_t)(key & 7u);
                rj = F.radius[lv];
                posj = (lv << 28) | ((uint32_t)y << 14) | (uint32_t)xj;
            }
        }
        int rw = rj;
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) rw = max(rw, __shfl_xor(rw, o, 64));
        /* groups of kNmsLanes lanes take the chunk's pixels 64 / kNmsLanes at a
           time, in x order */
        for (int q = 0; q < nchunk; q += 64 / kNmsLanes) {
            const int src = q + g;
            const int x = __shfl(xj, src, 64), r = __shfl(rj, src, 64);
            const uint32_t sc = (uint32_t)__shfl((int)sj, src, 64);
            const uint32_t pw = (uint32_t)__shfl((int)posj, src, 64);
            const bool mine = src < nchunk && r > 0;
            bool sup = false;
            if (mine) {
                for (int dy = sub - r; dy <= r; dy += kNmsLanes) {
                    const ORB_GLOBAL unsigned long long* rr = rowp + (long long)dy * F.p0;
                    switch (rw) {
                    case 1: sup |= nms_row_suppressed<1>(rr, x, dy, sc, epoch, r); break;
                    case 2: sup |= nms_row_suppressed<2>(rr, x, dy, sc, epoch, r); break;
                    case 3: sup |= nms_row_suppressed<3>(rr, x, dy, sc, epoch, r); break;
                    case 4: sup |= nms_row_suppressed<4>(rr, x, dy, sc, epoch, r); break;
                    default: sup |= nms_row_suppressed_any(rr, x, dy, sc, epoch, r);
                    }
                }
            }
            const unsigned long long m = __ballot(sup);
            const bool surv = mine && sub == 0 && ((m >> (kNmsLanes * g)) & ((1ull << kNmsLanes) - 1)) == 0;
            const unsigned long long ms = __ballot(surv);
            if (surv) {
                Candidate cd;
                cd.pos = pw;
                cd.score = sc;
                out[nout + lanes_below(ms)] = cd;
            }
            nout += __popcll(ms);
        }
    }
    if (lane == 0) seg_count[u] = nout;
}

/* ---------------------------------------------------------- select + emit */

__device__ __forceinline__ void emit_point(ORB_GLOBAL orb::OrbPoint* o, Candidate c)
{
    o->x = (int)(c.pos & 0x3fffu);
    o->y = (int)((c.pos >> 14) & 0x3fffu);
    o->octave = (int)(c.pos >> 28);
    o->score = __uint_as_float(c.score);
    o->angle = 0.f;
    o->match = -1;
    o->distance = -1;
}

/* segment offsets a select workgroup keeps in LDS (3840x2160 needs 4196) */
constexpr int kSelSegs = 6144;

/* One workgroup per frame: prefix sum of the NMS segment counts and the
   keypoints written in raster order.  With more than max_pts survivors, an
   exact radix select finds the score of the max_pts-th best; all higher
   scores and the first `need` equal ones in raster order are kept. */
__global__ __launch_bounds__(1024) void k_select(const FrameDev* __restrict__ frames, const Candidate* __restrict__ cand,
                                                 const int* __restrict__ blk_count, uint32_t* __restrict__ epochs)
{
    __shared__ int s_tmp[32];
    __shared__ int s_hist[256];
    __shared__ int s_pick[2];
    __shared__ int s_off[kSelSegs + 1];
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
    int out = block_scan<1024>(sum, s_tmp, total);

    if (total <= cap && nb <= kSelSegs) {
        /* every output slot in parallel: segment offsets in LDS, the slot's
           segment by binary search (the last offset at or below the slot) */
        for (int b = b0; b < b1; b++) {
            s_off[b] = out;
            out += blk_count[first + b];
        }
        if (tid == 0) s_off[nb] = total;
        __syncthreads();
        for (int k = tid; k < total; k += 1024) {
            int lo = 0, hi = nb - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (s_off[mid] <= k) lo = mid;
                else hi = mid - 1;
            }
            emit_point(as_global(F.pts) + k, cand[(long long)(first + lo) * kNmsCap + (k - s_off[lo])]);
        }
    } else if (total <= cap) {
        for (int b = b0; b < b1; b++) {
            const int n = blk_count[first + b];
            const Candidate* cb = cand + (long long)(first + b) * kNmsCap;
            for (int i = 0; i < n; i++) emit_point(as_global(F.pts) + out++, cb[i]);
        }
    } else {
        uint32_t prefix = 0, pmask = 0;
        int k = cap;
        for (int shift = 24; shift >= 0; shift -= 8) {
            for (int i = tid; i < 256; i += 1024) s_hist[i] = 0;
            __syncthreads();
            for (int b = b0; b < b1; b++) {
                const int n = blk_count[first + b];
                const Candidate* cb = cand + (long long)(first + b) * kNmsCap;
                for (int i = 0; i < n; i++) {
                    const uint32_t sv = cb[i].score;
                    if ((sv & pmask) == prefix) atomicAdd(&s_hist[(sv >> shift) & 255u], 1);
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
        const uint32_t thr = prefix;
        const int need = k;
        int gt = 0, eq = 0;
        for (int b = b0; b < b1; b++) {
            const int n = blk_count[first + b];
            const Candidate* cb = cand + (long long)(first + b) * kNmsCap;
            for (int i = 0; i < n; i++) {
                gt += cb[i].score > thr;
                eq += cb[i].score == thr;
            }
        }
        int tot;
        int eq_before = block_scan<1024>(eq, s_tmp, tot);
        out = block_scan<1024>(gt + min(eq, max(0, need - eq_before)), s_tmp, tot);
        for (int b = b0; b < b1; b++) {
            const int n = blk_count[first + b];
            const Candidate* cb = cand + (long long)(first + b) * kNmsCap;
            for (int i = 0; i < n; i++) {
                const Candidate c = cb[i];
                bool kept = c.score > thr;
                if (c.score == thr) kept = eq_before++ < need;
                if (kept) emit_point(as_global(F.pts) + out++, c);
            }
        }
        total = cap;
    }
    if (tid == 0) *as_global(F.count) = total;
    /* advance the epoch; on wrap-around clear this frame's key map */
    if (epoch + 1 >= kEpochLimit) {
        const long long n = (long long)F.h0 * F.p0;
        for (long long i = tid; i < n; i += 1024) as_global(F.keys)[i] = 0ull;
        if (tid == 0) epochs[f] = 1;
    } else if (tid == 0) {
        epochs[f] = epoch + 1;
    }
}

/* ------------------------------------------------------ OpenCV-layout export */

/* One workgroup converts a keypoint set (raster order, as detection writes it)
   to OpenCV's cv::KeyPoint layout: pt = level coordinate x layerScale,
   layerScale = (float)pow(scale, level), size = patch_size x layerScale, angle
   in degrees [0, 360), response = score, octave = level, class_id = -1.  With
   retain > 0 and more points than that, the `retain` best by score are kept
   (ties in raster order, the rule of the max_pts cut) by the same exact radix
   select as k_select; the kept points stay in raster order and their
   descriptors follow them.  Scores are positive, so their bits order them. */
__global__ __launch_bounds__(1024) void k_export(const orb::OrbPoint* __restrict__ pts,
                                                 const uint8_t* __restrict__ desc, const int* __restrict__ d_n,
                                                 int cap, int retain, ExportArgs A, orb_cv_keypoint* __restrict__ kp,
                                                 uint8_t* __restrict__ desc_out, int* __restrict__ d_nout)
{
    __shared__ int s_tmp[32];
    __shared__ int s_hist[256];
    __shared__ int s_pick[2];
    const int tid = threadIdx.x;
    const int n = d_n ? min(max(*d_n, 0), cap) : cap;
    const int per = (n + 1023) / 1024;
    const int i0 = min(n, tid * per), i1 = min(n, i0 + per);
    const bool cut = retain > 0 && n > retain;
    uint32_t thr = 0;
    int need = 0;
    if (cut) {
        uint32_t prefix = 0, pmask = 0;
        int k = retain;
        for (int shift = 24; shift >= 0; shift -= 8) {
            for (int i = tid; i < 256; i += 1024) s_hist[i] = 0;
            __syncthreads();
            for (int i = i0; i < i1; i++) {
                const uint32_t sv = __float_as_uint(pts[i].score);
                if ((sv & pmask) == prefix) atomicAdd(&s_hist[(sv >> shift) & 255u], 1);
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
    }
    auto kept = [&](uint32_t sv, int& eq_seen) {
        if (!cut || sv > thr) return true;
        return sv == thr && eq_seen++ < need;
    };
    /* equal-to-threshold points before this thread's range, then the output slots */
    int eq = 0, tot;
    if (cut)
        for (int i = i0; i < i1; i++) eq += __float_as_uint(pts[i].score) == thr;
    int eq_before = block_scan<1024>(eq, s_tmp, tot);
    int mine = 0;
    {
        int e = eq_before;
        for (int i = i0; i < i1; i++) mine += kept(__float_as_uint(pts[i].score), e);
    }
    int total;
    int out = block_scan<1024>(mine, s_tmp, total);
    for (int i = i0; i < i1; i++) {
        const orb::OrbPoint P = pts[i];
        if (!kept(__float_as_uint(P.score), eq_before)) continue;
        const int l = min(max(P.octave, 0), A.levels - 1);
        int lx, ly;
        if (A.ref_mode) {
            lx = P.x >> l;
            ly = P.y >> l;
        } else {
            lx = (int)floor(__dadd_rn(__ddiv_rn((double)P.x, A.sc[l]), 0.5));
            ly = (int)floor(__dadd_rn(__ddiv_rn((double)P.y, A.sc[l]), 0.5));
        }
        orb_cv_keypoint o;
        o.x = __fmul_rn((float)lx, A.lsc[l]);
        o.y = __fmul_rn((float)ly, A.lsc[l]);
        o.size = __fmul_rn(A.patch, A.lsc[l]);
        float deg = __fmul_rn(P.angle, 57.2957795130823208768f);
        if (deg < 0.f) deg = __fadd_rn(deg, 360.f);
        if (deg >= 360.f) deg = __fsub_rn(deg, 360.f);
        o.angle = deg;
        o.response = P.score;
        o.octave = l;
        o.class_id = -1;
        kp[out] = o;
        if (desc && desc_out) {
            const uint4* src = (const uint4*)(desc + (long long)i * 32);
            uint4* dst = (uint4*)(desc_out + (long long)out * 32);
            dst[0] = src[0];
            dst[1] = src[1];
        }
        out++;
    }
    if (tid == 0 && d_nout) *d_nout = total;
}

/* --------------------------------------------------------- angle + describe */

/* staged patches per wavefront: the raw 31-row disc region (36-byte rows) and
   the blurred 37x37 region of the rotated pattern (40-byte rows), each row the
   flat-layout bytes around the keypoint */
constexpr int kRawRow = 36, kBlurRow = 40;

__device__ __forceinline__ int sample_lds(const uint8_t* bp, int2 pt, float sn, float cs)
{
    const float x = (float)pt.x, y = (float)pt.y;
    const int ix = __float2int_rn(__fmaf_rn(x, cs, -__fmul_rn(y, sn)));
    const int iy = __float2int_rn(__fmaf_rn(x, sn, __fmul_rn(y, cs)));
    return bp[iy * kBlurRow + ix];
}

/* the same sample read from the flat blurred pyramid: offset c of the
   keypoint from level 0's start, bytes outside the pyramid read as 0 (the
   reference's flat addressing), for patterns that reach past the staged
   37 x 37 region */
__device__ __forceinline__ int sample_flat(const ORB_GLOBAL uint8_t* base, long long total, long long c, int p, int2 pt,
                                           float sn, float cs)
{
    const float x = (float)pt.x, y = (float)pt.y;
    const int ix = __float2int_rn(__fmaf_rn(x, cs, -__fmul_rn(y, sn)));
    const int iy = __float2int_rn(__fmaf_rn(x, sn, __fmul_rn(y, cs)));
    const long long i = c + (long long)iy * p + ix;
    return (i >= 0 && i < total) ? base[i] : 0;
}

/* PLAIN: WTA_K 2 with the patch-31 table (every sample inside the staged
   patch), the common case compiled on its own so it keeps fewer registers */
template <bool PLAIN>
__global__ __launch_bounds__(256) void k_describe(const FrameDev* __restrict__ frames, int blocks_per_frame, Tables T,
                                                  int compute_desc)
{
    __shared__ uint8_t s_code[kDescPerBlock][128];
    __shared__ __align__(16) uint8_t s_blur[kDescPerBlock][37 * kBlurRow];
    /* consecutive keypoints (raster neighbours, overlapping patches) on one XCD */
    const int bid = xcd_remap(blockIdx.x, gridDim.x);
    const int f = bid / blocks_per_frame;
    const int lb = bid - f * blocks_per_frame;
    const FrameDev& F = frames[f];
    const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int hk = T.half_patch;
    /* The angle's moments come straight from the loaded raw patch: dword
       i = lane + 64k of the 31 x 36-byte patch is row i / 9, bytes
       4 (i % 9) .. +3.  Per dword, fixed for every keypoint: the row's
       offset v from the centre, its umax bound (-1 outside the patch) and the
       column offset of its first byte (less the keypoint's alignment) */
    constexpr int kRawDw = kRawRow / 4, kRawN = 31 * kRawDw, kRawK = (kRawN + 63) / 64;
    int rrow[kRawK], rcol[kRawK], rv[kRawK], rum[kRawK];
#pragma unroll
    for (int k = 0; k < kRawK; k++) {
        const int i = lane + 64 * k, r = i / kRawDw, d = i - r * kRawDw;
        rrow[k] = r;
        rcol[k] = 4 * d;
        rv[k] = r - 15;
        rum[k] = (i < kRawN && abs(r - 15) <= hk) ? T.umax[abs(r - 15)] : -1;
    }
    /* the wavefronts work independently (no workgroup barrier): the count,
       the first point and this lane's pattern points (WTA_K 2) are all
       requested at once */
    int pi = lb * kDescPerBlock + wid;
    const int n = *as_global(F.count);
    int2 pat2[8];
    if (PLAIN || T.wta_k == 2) {
#pragma unroll
        for (int j = 0; j < 4; j++) {
            pat2[2 * j] = as_global(T.pattern)[2 * (lane + 64 * j)];
            pat2[2 * j + 1] = as_global(T.pattern)[2 * (lane + 64 * j) + 1];
        }
    }
    for (; pi < n; pi += blocks_per_frame * kDescPerBlock) {
        ORB_GLOBAL orb::OrbPoint* P = as_global(F.pts) + pi;
        const int px = __builtin_amdgcn_readfirstlane(P->x), py = __builtin_amdgcn_readfirstlane(P->y);
        const int l = __builtin_amdgcn_readfirstlane(P->octave);
        const int cx = from_l0(F, T.tabs, l, px, false), cy = from_l0(F, T.tabs, l, py, true);
        const int p = F.lp[l];
        const long long co = F.loff[l] + (long long)cy * p + cx;

        /* both patches with dword loads (every row shares the keypoint's
           alignment because the pitch is a multiple of 128) */
        const ORB_GLOBAL uint8_t* ra = as_global(F.raw) + co - 15 - 15LL * p;
        const int ro = (int)((uintptr_t)ra & 3);
        const ORB_GLOBAL uint8_t* rb = ra - ro;
        const ORB_GLOBAL uint8_t* ba = as_global(F.blur) + co - 18 - 18LL * p;
        const int bo = (int)((uintptr_t)ba & 3);
        const ORB_GLOBAL uint8_t* bb = ba - bo;
        constexpr int nblur = 37 * (kBlurRow / 4), pblur = (nblur + 63) / 64;
        uint32_t vr[kRawK], vb[pblur];
#pragma unroll
        for (int k = 0; k < kRawK; k++)
            vr[k] = rum[k] >= 0 ? *(const ORB_GLOBAL uint32_t*)(rb + (uint32_t)(rrow[k] * p + rcol[k])) : 0u;
#pragma unroll
        for (int k = 0; k < pblur; k++) {
            const int i = lane + 64 * k;
            const int r = i / (kBlurRow / 4), d = i - r * (kBlurRow / 4);
            if (compute_desc && i < nblur) vb[k] = *(const ORB_GLOBAL uint32_t*)(bb + (uint32_t)(r * p + 4 * d));
        }
#pragma unroll
        for (int k = 0; k < pblur; k++) {
            const int i = lane + 64 * k;
            if (compute_desc && i < nblur) *(uint32_t*)&s_blur[wid][4 * i] = vb[k];
        }

        /* intensity centroid over the umax disc of the raw level
           (orbd.cu:1011-1032): byte b of dword k is column
           u = rcol + b - ro - 15 of row v; integer sums, so their order does
           not change them */
        int m10 = 0, m01 = 0;
#pragma unroll
        for (int k = 0; k < kRawK; k++) {
            const int u0 = rcol[k] - ro - 15;
            int su = 0, sI = 0;
#pragma unroll
            for (int b = 0; b < 4; b++) {
                const int u = u0 + b;
                const int I = abs(u) <= rum[k] ? (int)((vr[k] >> (8 * b)) & 255u) : 0;
                su += u * I;
                sI += I;
            }
            m10 += su;
            m01 += rv[k] * sI;
        }
        m10 = wave_sum(m10);
        m01 = wave_sum(m01);
        const float ang = (m01 == 0 && m10 == 0) ? 0.f : dev_atan2(__int2float_rn(m01), __int2float_rn(m10));
        if (lane == 0) P->angle = ang;
        wave_sync_lds();
        if (compute_desc) {
            float sn, cs;
            dev_sincos(ang, sn, cs);
            const uint8_t* bp = &s_blur[wid][18 * kBlurRow + bo + 18];
            /* patch 31: every rotated point lies within 18 of the centre, in the
               staged patch; the random points of other patch sizes sit off
               centre (orbd.cu:359-369) and are read from the flat pyramid
               (patch 30 has the same half patch as 31 but random points) */
            const bool flat = !PLAIN && !T.pattern31;
            const ORB_GLOBAL uint8_t* fb = as_global(F.blur);
            auto smp = [&](int2 pt) { return flat ? sample_flat(fb, F.total, co, p, pt, sn, cs) : sample_lds(bp, pt, sn, cs); };
            if (PLAIN || T.wta_k == 2) {
                /* bit k of the descriptor compares points 2k and 2k+1 (orbd.cu:1351-1361) */
                ORB_GLOBAL unsigned long long* outd = (ORB_GLOBAL unsigned long long*)(as_global(F.desc) + (long long)pi * 32);
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const bool bit = smp(pat2[2 * j]) < smp(pat2[2 * j + 1]);
                    const unsigned long long m = __ballot(bit);
                    if (lane == 0) outd[j] = m;
                }
            } else {
                /* 2-bit argmax codes, code q = 4*byte + slot (orbd.cu:1364-1434) */
                for (int j = 0; j < 2; j++) {
                    const int q = lane + 64 * j, byte = q >> 2, slot = q & 3;
                    int code;
                    if (T.wta_k == 3) {
                        const ORB_GLOBAL int2* pp = as_global(T.pattern) + 12 * byte;
                        const int i0 = slot < 3 ? 3 * slot : 8, i1 = slot < 3 ? 3 * slot + 1 : 9,
                                  i2 = slot < 3 ? 3 * slot + 2 : 11;
                        const int t0 = smp(pp[i0]), t1 = smp(pp[i1]),
                                  t2 = smp(pp[i2]);
                        code = t2 > t1 ? (t2 > t0 ? 2 : 0) : (t1 > t0);
                    } else {
                        const ORB_GLOBAL int2* pp = as_global(T.pattern) + 16 * byte + 4 * slot;
                        int t0 = smp(pp[0]), t1 = smp(pp[1]);
                        int t2 = smp(pp[2]), t3 = smp(pp[3]);
                        int u = 0, v = 2;
                        if (t1 > t0) { t0 = t1; u = 1; }
                        if (t3 > t2) { t2 = t3; v = 3; }
                        code = t0 > t2 ? u : v;
                    }
                    s_code[wid][q] = (uint8_t)code;
                }
                wave_sync_lds();
                if (lane < 32) {
                    const uint8_t* sc = s_code[wid] + 4 * lane;
                    as_global(F.desc)[(long long)pi * 32 + lane] = (uint8_t)(sc[0] | (sc[1] << 2) | (sc[2] << 4) | (sc[3] << 6));
                }
            }
        }
        wave_sync_lds();
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

hipError_t launch_pyramid(const DetectLaunch& d, int variant, hipStream_t s, const StageEvents* ev)
{
    switch (variant) {
#define ORB_PYR_CASE(v)                                                                                         \
    case v:                                                                                                     \
        launch_stage(ev, kStPyramid, k_pyramid<kPyrTHs[v]>, dim3(d.pyr_blocks), dim3(256), s, d.frames, d.pyr_jobs, \
                     d.pyr_map, d.pyr_units, d.t);                                                              \
        break;
        ORB_PYR_CASE(0)
        ORB_PYR_CASE(1)
        ORB_PYR_CASE(2)
        ORB_PYR_CASE(3)
        ORB_PYR_CASE(4)
#undef ORB_PYR_CASE
    default: return hipErrorInvalidValue;
    }
    return hipSuccess;
}

hipError_t launch_detect(const DetectLaunch& d, int compute_desc, hipStream_t s, const StageEvents* ev)
{
    (void)hipGetLastError(); /* report only errors of these launches */
    if (d.pyr_blocks > 0) {
        hipError_t e = launch_pyramid(d, d.pyr_variant, s, ev);
        if (e != hipSuccess) return e;
    }
    if (d.det_blocks > 0)
        launch_stage(ev, kStDetect, k_detect, dim3(d.det_blocks), dim3(256), s, d.frames, d.det_jobs, d.det_map,
                     d.t, d.epochs);
    if (d.nms_blocks > 0)
        launch_stage(ev, kStNms, k_nms, dim3((d.nms_blocks + 3) / 4), dim3(256), s, d.frames, d.nms_block_frame,
                     d.nms_blocks, d.epochs, d.cand, d.blk_count);
    launch_stage(ev, kStSelect, k_select, dim3(d.nframes), dim3(1024), s, d.frames, d.cand, d.blk_count, d.epochs);
    if (compute_desc && d.desc_blocks_per_frame > 0) {
        if (d.t.wta_k == 2 && d.t.pattern31)
            launch_stage(ev, kStDescribe, k_describe<true>, dim3(d.desc_blocks_per_frame * d.nframes), dim3(256), s,
                         d.frames, d.desc_blocks_per_frame, d.t, compute_desc);
        else
            launch_stage(ev, kStDescribe, k_describe<false>, dim3(d.desc_blocks_per_frame * d.nframes), dim3(256), s,
                         d.frames, d.desc_blocks_per_frame, d.t, compute_desc);
    }
    return hipGetLastError();
}

hipError_t launch_export(const orb::OrbPoint* pts, const uint8_t* desc, const int* d_n, int cap, int retain,
                         const ExportArgs& A, orb_cv_keypoint* kp, uint8_t* desc_out, int* d_nout, hipStream_t s)
{
    (void)hipGetLastError();
    hipLaunchKernelGGL(k_export, dim3(1), dim3(1024), 0, s, pts, desc, d_n, cap, retain, A, kp, desc_out, d_nout);
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
