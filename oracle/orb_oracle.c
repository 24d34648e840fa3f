/*
 * CPU oracle for the ORB detect / describe / match path -- TEST INFRASTRUCTURE.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
 * file, as the checker.  It is a plain C restatement of the reference's
 * semantics (file:line citations are relative to the reference tree), built
 * with -ffp-contract=off so that every floating-point operation below is the
 * single IEEE operation written (fused multiply-adds are spelled fmaf).
 *
 * Semantics fixed here where the reference is nondeterministic or undefined:
 *  - flat pyramid addressing: levels are contiguous, row pitch alignUp(w,128),
 *    and a read outside [0, total) of the flat buffer returns 0
 *    (orb.cpp:151-175; the reference reads out of its allocation there);
 *  - Harris integer products wrap modulo 2^32 like the reference's int math
 *    (orbd.cu:653-668); the float tail is fmaf(-0.04f, (float)B, (float)A) * s;
 *  - the 7x7 blur is out of place (the reference blurs in place, racing
 *    across tiles, orbd.cu:1756) with reflect-101 borders and fmaf taps;
 *  - keypoints come out in raster order (the reference appends with
 *    atomicInc, orbd.cu:970); when more than max_pts survive NMS the max_pts
 *    with the highest score are kept, ties by raster order;
 *  - sin/cos are this file's own Cody-Waite + Cephes minimax polynomials
 *    (the reference uses __sinf/__cosf, orbd.cu:1446-1447);
 *  - an angle whose two moments are both 0 is 0 (the reference yields NaN).
 *
 * Parity pins for each function are the CPU tests in tests/test_oracle.py
 * (closed forms, brute force, invariants); see the function comments.
 */
#include "orb_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

static const int32_t k_pattern31[1024] = {
#include "pattern31.inc"
};

/* ------------------------------------------------------------------ tables */

/* orbd.cu:57-63: lut[255 + (n - c)] = 1 darker, 2 brighter, 0 similar. */
void orc_fast_lut(int t, uint8_t lut[512])
{
    for (int i = -255; i <= 255; i++)
        lut[i + 255] = (uint8_t)(i < -t ? 1 : (i > t ? 2 : 0));
    lut[511] = 0;
}

/* orbd.cu:66-92: OpenCV's symmetric disc half-widths for the centroid patch. */
void orc_umax(int patch_size, int* umax)
{
    int half = patch_size / 2;
    umax[half + 1] = 0;
    float v = half * sqrtf(2.f) / 2;
    int vmax = (int)floorf(v + 1);
    int vmin = (int)ceilf(v);
    for (int i = 0; i <= vmax; i++)
        umax[i] = (int)roundf(sqrtf((float)(half * half - i * i)));
    for (int i = half, v0 = 0; i >= vmin; --i) {
        while (umax[v0] == umax[v0 + 1])
            ++v0;
        umax[i] = v0;
        ++v0;
    }
}

/* orbd.cu:413-433: 7-tap Gaussian, sigma 2, normalised in float. */
void orc_gauss(float k[4])
{
    const float svar = -1.f / (2.f * 2.f * 2.f);
    float sum = 0.f;
    for (int i = 0; i <= 3; i++) {
        k[i] = expf((float)(i * i) * svar);
        sum = sum + (k[i] + (i == 0 ? 0.f : k[i]));
    }
    sum = 1.f / sum;
    for (int i = 0; i <= 3; i++)
        k[i] *= sum;
}

/* orbd.cu:436-441: (1 / (4 * 7 * 255))^4 in float. */
float orc_harris_scale(void)
{
    float s = 1.f / (4 * 7 * 255.f);
    return s * s * s * s;
}

/* glibc random(3) TYPE_3 generator (additive feedback r[i] = r[i-3] + r[i-31],
   seeded by the Park-Miller LCG, 310 discarded outputs), restated from its
   published description; the reference calls srand/rand at orbd.cu:380-392. */
static int32_t g_ring[31]; /* r[n-31 .. n-1], g_f indexes r[n-31] */
static int g_f;

void orc_glibc_srand(unsigned seed)
{
    int32_t r[344];
    if (seed == 0) seed = 1;
    r[0] = (int32_t)seed;
    for (int i = 1; i < 31; i++) {
        long hi = r[i - 1] / 127773, lo = r[i - 1] % 127773;
        long word = 16807 * lo - 2836 * hi;
        if (word < 0) word += 2147483647;
        r[i] = (int32_t)word;
    }
    for (int i = 31; i < 34; i++) r[i] = r[i - 31];
    for (int i = 34; i < 344; i++) r[i] = (int32_t)((uint32_t)r[i - 31] + (uint32_t)r[i - 3]);
    for (int i = 0; i < 31; i++) g_ring[i] = r[313 + i];
    g_f = 0; /* the next output is r[344] = r[313] + r[341] */
}

int orc_glibc_rand(void)
{
    /* r[n] = r[n-31] + r[n-3]; ring index g_f holds r[n-31] */
    int32_t rn31 = g_ring[g_f];
    int32_t rn3 = g_ring[(g_f + 28) % 31];
    int32_t v = (int32_t)((uint32_t)rn31 + (uint32_t)rn3);
    g_ring[g_f] = v;
    g_f = (g_f + 1) % 31;
    return (int)((uint32_t)v >> 1);
}

/* orbd.cu:95-410: the base set of 512 points is the table (patch_size 31) or,
   for another patch size p, 512 points drawn with srand(0x34985739), each
   coordinate rand() % p - -p / 2 (orbd.cu:359-369: the subtraction of a
   negated half makes every offset +p/2, so the points lie in
   [p/2, p + p/2 - 1], off the patch centre, a reference bug kept as it is).
   WTA_K 2 uses the base set as it is; WTA_K 3/4 resample it into 128 tuples
   of distinct points with srand(0x12345678). */
int orc_pattern_patch(int patch_size, int wta_k, int32_t* xy)
{
    if (wta_k < 2 || wta_k > 4 || patch_size < 3 || patch_size > 31) return -1;
    int32_t base[1024];
    if (patch_size == 31) {
        memcpy(base, k_pattern31, sizeof(base));
    } else {
        orc_glibc_srand(0x34985739u);
        for (int i = 0; i < 512; i++) {
            base[2 * i] = orc_glibc_rand() % patch_size - -patch_size / 2;
            base[2 * i + 1] = orc_glibc_rand() % patch_size - -patch_size / 2;
        }
    }
    if (wta_k == 2) {
        memcpy(xy, base, sizeof(base));
        return 0;
    }
    memset(xy, 0, 1024 * sizeof(int32_t));
    orc_glibc_srand(0x12345678u);
    const int ntuples = 128;
    for (int i = 0; i < ntuples; i++) {
        for (int k = 0; k < wta_k; k++) {
            for (;;) {
                int idx = orc_glibc_rand() % 512;
                int px_ = base[2 * idx], py_ = base[2 * idx + 1];
                int k1;
                for (k1 = 0; k1 < k; k1++) {
                    int q = wta_k * i + k1;
                    if (px_ == xy[2 * q] && py_ == xy[2 * q + 1]) break;
                }
                if (k1 == k) {
                    xy[2 * (wta_k * i + k)] = px_;
                    xy[2 * (wta_k * i + k) + 1] = py_;
                    break;
                }
            }
        }
    }
    return 0;
}

int orc_pattern(int wta_k, int32_t* xy)
{
    return orc_pattern_patch(31, wta_k, xy);
}

/* -------------------------------------------------------------------- math */

#define ORC_H_PI 1.5707963267948966f
#define ORC_M_PI 3.14159265358979323846

/* orbd.cu:986-998 (only IEEE ops: division, multiply, fma). */
float orc_fast_atan2(float y, float x)
{
    const float ax = fabsf(x), ay = fabsf(y);
    const float a = fminf(ax, ay) / fmaxf(ax, ay);
    const float s = a * a;
    float r = fmaf(fmaf(fmaf(-0.0464964749f, s, 0.15931422f), s, -0.327622764f), s * a, a);
    r = (ay > ax) ? ORC_H_PI - r : r;
    r = (x < 0) ? (float)(ORC_M_PI - (double)r) : r;
    r = (y < 0) ? -r : r;
    return r;
}

/* sin/cos: Cody-Waite reduction by pi/2 with a three-part split of pi/2, then
   the Cephes sinf/cosf minimax polynomials on [-pi/4, pi/4] (S. Moshier,
   Cephes Math Library).  Replaces __sinf/__cosf of orbd.cu:1446-1447. */
void orc_sincosf(float a, float* sn, float* cs)
{
    const float j = rintf(a * 0.636619772f);
    float r = fmaf(-j, 1.5703125f, a);
    r = fmaf(-j, 4.837512969970703125e-4f, r);
    r = fmaf(-j, 7.54978995489188216e-8f, r);
    const int q = ((int)j) & 3;
    const float z = r * r;
    const float sp = fmaf(fmaf(fmaf(-1.9515295891e-4f, z, 8.3321608736e-3f), z, -1.6666654611e-1f),
                          z * r, r);
    const float cp = fmaf(fmaf(fmaf(2.443315711809948e-5f, z, -1.388731625493765e-3f), z,
                               4.166664568298827e-2f),
                          z * z, fmaf(-0.5f, z, 1.0f));
    switch (q) {
    case 0: *sn = sp; *cs = cp; break;
    case 1: *sn = cp; *cs = -sp; break;
    case 2: *sn = -sp; *cs = -cp; break;
    default: *sn = -cp; *cs = sp; break;
    }
}

/* ---------------------------------------------------------------- geometry */

static int align_up(int a, int b) { return (a % b != 0) ? (a - a % b + b) : a; }

/* orb.cpp:145-191 for scale 2 (level count min(noctaves, log2(min(w,h)/80)+1, 5),
   w_l = w_{l-1} >> 1, pitch alignUp(w,128), levels contiguous).  For another
   scale s in (1,2): w_l = floor(w / s^l + 0.5), levels kept while the short
   side is >= 16, at most 8.  NMS radius: 2^l (orbd.cu:941) or max(1, rn(s^l)). */
int orc_geometry(int w, int h, int noctaves, double scale, orc_geom* g)
{
    memset(g, 0, sizeof(*g));
    if (w <= 0 || h <= 0 || noctaves < 1) return -1;
    if (!(scale > 1.0 && scale <= 2.0)) return -1;
    g->scale = scale;
    g->ref_mode = (scale == 2.0);
    int L;
    if (g->ref_mode) {
        int m = (w < h ? w : h) / 80;
        if (m < 1) return -1;
        L = (int)log2f((float)m) + 1;
        if (L > noctaves) L = noctaves;
        if (L > 5) L = 5; /* MAX_OCTAVE, orbd.cu:18 */
        for (int l = 0; l < L; l++) {
            g->w[l] = l == 0 ? w : g->w[l - 1] >> 1;
            g->h[l] = l == 0 ? h : g->h[l - 1] >> 1;
            g->sc[l] = (double)(1 << l);
            g->radius[l] = 1 << l;
        }
    } else {
        if ((w < h ? w : h) < 16) return -1;
        L = 0;
        double sc = 1.0;
        for (int l = 0; l < noctaves && l < ORC_MAX_LEVELS; l++) {
            if (l > 0) sc *= scale;
            int wl = (int)floor(w / sc + 0.5), hl = (int)floor(h / sc + 0.5);
            if ((wl < hl ? wl : hl) < 16) break;
            g->w[l] = wl;
            g->h[l] = hl;
            g->sc[l] = sc;
            int r = (int)floor(sc + 0.5);
            g->radius[l] = r < 1 ? 1 : r;
            L++;
        }
    }
    g->nlevels = L;
    long off = 0;
    for (int l = 0; l < L; l++) {
        g->p[l] = align_up(g->w[l], 128);
        g->off[l] = off;
        off += (long)g->h[l] * g->p[l];
    }
    g->total = off;
    return 0;
}

/* level-local coordinate -> level-0 coordinate (orbd.cu:745-746 for scale 2) */
int orc_level_to_l0(const orc_geom* g, int level, int v)
{
    if (g->ref_mode) return v << level;
    return (int)floor(v * g->sc[level] + 0.5);
}

/* inverse of orc_level_to_l0 on its image (orbd.cu:1009 for scale 2) */
int orc_l0_to_level(const orc_geom* g, int level, int v)
{
    if (g->ref_mode) return v >> level;
    return (int)floor(v / g->sc[level] + 0.5);
}

/* threads of the parallel loops below (1: the plain scalar oracle).  Every
   parallel loop writes disjoint outputs, so results do not depend on it. */
static int g_threads = 1;

void orc_set_threads(int n)
{
    g_threads = n < 1 ? 1 : n;
}

int orc_get_threads(void)
{
    return g_threads;
}

static inline int px(const uint8_t* buf, long total, long i)
{
    return (i >= 0 && i < total) ? buf[i] : 0;
}

/* ----------------------------------------------------------------- pyramid */

/* bilinear source table for one axis: level index -> (i0 << 8) | frac, fixed
   point with 8 fractional bits, centre-aligned ((i + 0.5) * S / D - 0.5). */
static int resize_tab(int i, int S, int D)
{
    long long num = ((long long)(2 * i + 1) * S - D) * 256;
    long long pos = num / (2LL * D);
    if (pos < 0) pos = 0;
    if (pos > (long long)(S - 1) * 256) pos = (long long)(S - 1) * 256;
    return (int)pos;
}

static int dealBorder(int i, int sz) /* cuda_utils.h:145-152 */
{
    if (i < 0) return -i;
    if (i >= sz) return sz + sz - 2 - i;
    return i;
}

/* orbd.cu:1273-1338 semantics, out of place: horizontal 7-tap pass on integer
   pair sums, vertical pass on the float rows, round half to even. */
void orc_blur_level(const orc_geom* g, int l, const uint8_t* raw, uint8_t* blur)
{
    float k[4];
    orc_gauss(k);
    const int w = g->w[l], h = g->h[l], p = g->p[l];
    const uint8_t* src = raw + g->off[l];
    uint8_t* dst = blur + g->off[l];
    float* rows = (float*)malloc(sizeof(float) * (size_t)w * h);
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
    for (int y = 0; y < h; y++) {
        const uint8_t* r = src + (long)y * p;
        for (int x = 0; x < w; x++) {
            float s = k[0] * (float)r[x];
            for (int i = 1; i <= 3; i++)
                s = fmaf(k[i], (float)((int)r[dealBorder(x - i, w)] + (int)r[dealBorder(x + i, w)]), s);
            rows[(long)y * w + x] = s;
        }
    }
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
    for (int y = 0; y < h; y++) {
        for (int x = 0; x < w; x++) {
            float s = k[0] * rows[(long)y * w + x];
            for (int i = 1; i <= 3; i++)
                s = fmaf(k[i], rows[(long)dealBorder(y - i, h) * w + x] + rows[(long)dealBorder(y + i, h) * w + x], s);
            dst[(long)y * p + x] = (uint8_t)(int)rintf(s);
        }
    }
    free(rows);
}

/* orbd.cu:1579-1600: level 0 is a verbatim copy of the pitched input (padding
   included); level l >= 1 is L0[y<<l][x<<l] for scale 2, or the fixed-point
   bilinear resize of L0 otherwise; padding of levels >= 1 is 0.  The blurred
   copy keeps level 0's padding bytes, as the in-place blur does. */
void orc_pyramid(const orc_geom* g, const uint8_t* img, uint8_t* raw, uint8_t* blur)
{
    const int w0 = g->w[0], h0 = g->h[0], p0 = g->p[0];
    memset(raw, 0, (size_t)g->total);
    memcpy(raw, img, (size_t)h0 * p0);
    for (int l = 1; l < g->nlevels; l++) {
        uint8_t* dst = raw + g->off[l];
        const int w = g->w[l], h = g->h[l], p = g->p[l];
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1)
        for (int y = 0; y < h; y++) {
            for (int x = 0; x < w; x++) {
                int v;
                if (g->ref_mode) {
                    v = img[(long)(y << l) * p0 + (x << l)];
                } else {
                    int ty = resize_tab(y, h0, h), tx = resize_tab(x, w0, w);
                    int y0 = ty >> 8, fy = ty & 255, x0 = tx >> 8, fx = tx & 255;
                    int y1 = y0 + 1 < h0 ? y0 + 1 : h0 - 1;
                    int x1 = x0 + 1 < w0 ? x0 + 1 : w0 - 1;
                    int a = img[(long)y0 * p0 + x0], b = img[(long)y0 * p0 + x1];
                    int c = img[(long)y1 * p0 + x0], d = img[(long)y1 * p0 + x1];
                    int top = (256 - fx) * a + fx * b, bot = (256 - fx) * c + fx * d;
                    v = ((256 - fy) * top + fy * bot + 32768) >> 16;
                }
                dst[(long)y * p + x] = (uint8_t)v;
            }
        }
    }
    memset(blur, 0, (size_t)g->total);
    memcpy(blur, img, (size_t)h0 * p0); /* level-0 padding bytes */
    for (int l = 0; l < g->nlevels; l++)
        orc_blur_level(g, l, raw, blur);
}

/* --------------------------------------------------------------- detection */

static const int k_circle[16][2] = {
    {0, 3}, {1, 3}, {2, 2}, {3, 1}, {3, 0}, {3, -1}, {2, -2}, {1, -3},
    {0, -3}, {-1, -3}, {-2, -2}, {-3, -1}, {-3, 0}, {-3, 1}, {-2, 2}, {-1, 3}}; /* orbd.cu:466-469 */

/* orbd.cu:685-740: quick rejection on opposite pairs, then >8 consecutive of
   the 25-entry wrapped ring strictly darker (tested first, exclusively) or
   strictly brighter. */
int orc_fast_corner(const uint8_t* raw, long total, long c, int pitch, const uint8_t* lut, int t)
{
    int ring[25];
    for (int k = 0; k < 16; k++)
        ring[k] = px(raw, total, c + k_circle[k][0] + (long)k_circle[k][1] * pitch);
    for (int k = 16; k < 25; k++) ring[k] = ring[k - 16];
    const int v = px(raw, total, c);
    const uint8_t* tab = lut + 255 - v;
    int d = tab[ring[0]] | tab[ring[8]];
    if (d == 0) return 0;
    d &= tab[ring[2]] | tab[ring[10]];
    d &= tab[ring[4]] | tab[ring[12]];
    d &= tab[ring[6]] | tab[ring[14]];
    if (d == 0) return 0;
    d &= tab[ring[1]] | tab[ring[9]];
    d &= tab[ring[3]] | tab[ring[11]];
    d &= tab[ring[5]] | tab[ring[13]];
    d &= tab[ring[7]] | tab[ring[15]];
    if (d & 1) {
        int vt = v - t, count = 0;
        for (int k = 0; k < 25; k++) {
            if (ring[k] < vt) { if (++count > 8) return 1; }
            else count = 0;
        }
    } else if (d & 2) {
        int vt = v + t, count = 0;
        for (int k = 0; k < 25; k++) {
            if (ring[k] > vt) { if (++count > 8) return 1; }
            else count = 0;
        }
    }
    return 0;
}

/* orbd.cu:651-669: Sobel gradients summed over the 7x7 window in 32-bit ints;
   the two products wrap modulo 2^32 as in the reference. */
float orc_harris(const uint8_t* raw, long total, long c, int p)
{
    int32_t dxx = 0, dyy = 0, dxy = 0;
    for (int i = -3; i <= 3; i++) {
        for (int j = -3; j <= 3; j++) {
            long q = c + (long)i * p + j;
            int dx = (px(raw, total, q + 1) - px(raw, total, q - 1)) * 2 +
                     (px(raw, total, q - p + 1) - px(raw, total, q - p - 1)) +
                     (px(raw, total, q + p + 1) - px(raw, total, q + p - 1));
            int dy = (px(raw, total, q + p) - px(raw, total, q - p)) * 2 +
                     (px(raw, total, q + p - 1) - px(raw, total, q - p - 1)) +
                     (px(raw, total, q + p + 1) - px(raw, total, q - p + 1));
            dxx += dx * dx;
            dyy += dy * dy;
            dxy += dx * dy;
        }
    }
    int32_t A = (int32_t)((uint32_t)dxx * (uint32_t)dyy - (uint32_t)dxy * (uint32_t)dxy);
    int32_t S = dxx + dyy;
    int32_t B = (int32_t)((uint32_t)S * (uint32_t)S);
    return fmaf(-0.04f, (float)B, (float)A) * orc_harris_scale();
}

/* orbd.cu:538-578 (OpenCV cornerScore<16>): largest threshold keeping the
   corner, returned as -b0 - 1. */
float orc_fast_score(const uint8_t* raw, long total, long c, int p, int threshold)
{
    int v = px(raw, total, c);
    int d[25];
    for (int k = 0; k < 16; k++)
        d[k] = (short)(v - px(raw, total, c + k_circle[k][0] + (long)k_circle[k][1] * p));
    for (int k = 16; k < 25; k++) d[k] = d[k - 16];
    int a0 = threshold;
    for (int k = 0; k < 16; k += 2) {
        int a = d[k + 1] < d[k + 2] ? d[k + 1] : d[k + 2];
        a = a < d[k + 3] ? a : d[k + 3];
        if (a <= a0) continue;
        for (int m = 4; m <= 8; m++) a = a < d[k + m] ? a : d[k + m];
        int t1 = a < d[k] ? a : d[k];
        a0 = a0 > t1 ? a0 : t1;
        int t2 = a < d[k + 9] ? a : d[k + 9];
        a0 = a0 > t2 ? a0 : t2;
    }
    int b0 = -a0;
    for (int k = 0; k < 16; k += 2) {
        int b = d[k + 1] > d[k + 2] ? d[k + 1] : d[k + 2];
        for (int m = 3; m <= 5; m++) b = b > d[k + m] ? b : d[k + m];
        if (b >= b0) continue;
        for (int m = 6; m <= 8; m++) b = b > d[k + m] ? b : d[k + m];
        int t1 = b > d[k] ? b : d[k];
        b0 = b0 < t1 ? b0 : t1;
        int t2 = b > d[k + 9] ? b : d[k + 9];
        b0 = b0 < t2 ? b0 : t2;
    }
    return (float)(-b0 - 1);
}

/* orbd.cu:672-756 + 1616-1634: per level (ascending) the segment test on
   [3,w-3) x [3,h-3); a corner's score replaces the level-0 map value at its
   level-0 coordinate iff strictly greater (map starts at 0, so the lowest level
   wins ties and non-positive scores are never recorded). */
void orc_detect_map(const orc_geom* g, const uint8_t* raw, int threshold, int score_type,
                    float* vmap, int* lmap)
{
    uint8_t lut[512];
    orc_fast_lut(threshold, lut);
    const int p0 = g->p[0];
    memset(vmap, 0, sizeof(float) * (size_t)g->h[0] * p0);
    memset(lmap, 0, sizeof(int) * (size_t)g->h[0] * p0);
    /* levels in ascending order (the lowest level wins ties); the rows of a
       level map to distinct level-0 rows, so they run in parallel */
    for (int l = 0; l < g->nlevels; l++) {
        const int w = g->w[l], h = g->h[l], p = g->p[l];
#pragma omp parallel for schedule(dynamic, 8) num_threads(g_threads) if (g_threads > 1)
        for (int iy = 3; iy < h - 3; iy++) {
            for (int ix = 3; ix < w - 3; ix++) {
                long c = g->off[l] + (long)iy * p + ix;
                if (!orc_fast_corner(raw, g->total, c, p, lut, threshold)) continue;
                float s = score_type == 0 ? orc_harris(raw, g->total, c, p)
                                          : orc_fast_score(raw, g->total, c, p, threshold);
                long idx = (long)orc_level_to_l0(g, l, iy) * p0 + orc_level_to_l0(g, l, ix);
                if (vmap[idx] < s) {
                    vmap[idx] = s;
                    lmap[idx] = l;
                }
            }
        }
    }
}

typedef struct { float s; long r; } orc_cand;

static int cand_cmp(const void* a, const void* b)
{
    const orc_cand* x = (const orc_cand*)a;
    const orc_cand* y = (const orc_cand*)b;
    if (x->s != y->s) return x->s > y->s ? -1 : 1;
    return x->r < y->r ? -1 : (x->r > y->r);
}

static int long_cmp(const void* a, const void* b)
{
    long x = *(const long*)a, y = *(const long*)b;
    return x < y ? -1 : (x > y);
}

/* orbd.cu:919-983: square (2r+1)^2 suppression, r = radius of the centre's
   level; a neighbour suppresses if it is positive and greater, or equal and at
   dx <= 0 && dy <= 0.  Map reads use flat indexing with pitch p0, 0 outside
   [0, h0*p0).  Output: raster order; top max_pts by (score desc, raster asc). */
int orc_nms(const orc_geom* g, const float* vmap, const int* lmap, int border, int max_pts,
            orc_point* pts, int* n_survivors)
{
    if (border < 3) border = 3;
    const int w0 = g->w[0], h0 = g->h[0], p0 = g->p[0];
    const long total = (long)h0 * p0;
    long cap = 1024, n = 0;
    long* surv = (long*)malloc(sizeof(long) * cap);
    for (int y = border; y < h0 - border; y++) {
        for (int x = border; x < w0 - border; x++) {
            long vi = (long)y * p0 + x;
            float v = vmap[vi];
            if (!(v > 0)) continue;
            int r = g->radius[lmap[vi]];
            int supp = 0;
            for (int dy = -r; dy <= r && !supp; dy++) {
                for (int dx = -r; dx <= r; dx++) {
                    if (dx == 0 && dy == 0) continue;
                    long ni = vi + (long)dy * p0 + dx;
                    float nv = (ni >= 0 && ni < total) ? vmap[ni] : 0.f;
                    if (nv > 0 && (nv > v || (nv == v && dx <= 0 && dy <= 0))) { supp = 1; break; }
                }
            }
            if (supp) continue;
            if (n == cap) { cap *= 2; surv = (long*)realloc(surv, sizeof(long) * cap); }
            surv[n++] = vi;
        }
    }
    if (n_survivors) *n_survivors = (int)n;
    long keep = n;
    if (n > max_pts) {
        orc_cand* c = (orc_cand*)malloc(sizeof(orc_cand) * n);
        for (long i = 0; i < n; i++) { c[i].s = vmap[surv[i]]; c[i].r = surv[i]; }
        qsort(c, n, sizeof(orc_cand), cand_cmp);
        for (long i = 0; i < max_pts; i++) surv[i] = c[i].r;
        free(c);
        keep = max_pts;
        qsort(surv, keep, sizeof(long), long_cmp);
    }
    for (long i = 0; i < keep; i++) {
        long vi = surv[i];
        pts[i].x = (int)(vi % p0);
        pts[i].y = (int)(vi / p0);
        pts[i].octave = lmap[vi];
        pts[i].score = vmap[vi];
        pts[i].angle = 0.f;
        pts[i].match = -1;
        pts[i].distance = -1;
    }
    free(surv);
    return (int)keep;
}

/* ------------------------------------------------------ angle and describe */

static long center_of(const orc_geom* g, const orc_point* pt)
{
    int l = pt->octave;
    int cx = orc_l0_to_level(g, l, pt->x), cy = orc_l0_to_level(g, l, pt->y);
    return g->off[l] + (long)cy * g->p[l] + cx;
}

/* orbd.cu:1001-1036: intensity centroid over the umax disc of the raw level. */
void orc_angles(const orc_geom* g, const uint8_t* raw, orc_point* pts, int n, int patch_size)
{
    int umax[64];
    orc_umax(patch_size, umax);
    const int half = patch_size / 2;
#pragma omp parallel for schedule(dynamic, 32) num_threads(g_threads) if (g_threads > 1)
    for (int i = 0; i < n; i++) {
        const long c = center_of(g, &pts[i]);
        const int p = g->p[pts[i].octave];
        int m01 = 0, m10 = 0;
        for (int u = -half; u <= half; ++u) m10 += u * px(raw, g->total, c + u);
        for (int v = 1; v <= half; ++v) {
            int vsum = 0;
            for (int u = -umax[v]; u <= umax[v]; ++u) {
                int vp = px(raw, g->total, c + u + (long)v * p);
                int vm = px(raw, g->total, c + u - (long)v * p);
                vsum += vp - vm;
                m10 += u * (vp + vm);
            }
            m01 += v * vsum;
        }
        pts[i].angle = (m01 == 0 && m10 == 0) ? 0.f : orc_fast_atan2((float)m01, (float)m10);
    }
}

/* orbd.cu:1341-1348: rotate a pattern point and read the blurred level;
   rotation ix = rn(fma(x, c, -(y*s))), iy = rn(fma(x, s, y*c)). */
static int sample(const orc_geom* g, const uint8_t* b, long c, int p, const int32_t* pat, int idx,
                  float sn, float cs)
{
    const float x = (float)pat[2 * idx], y = (float)pat[2 * idx + 1];
    const int ix = (int)rintf(fmaf(x, cs, -(y * sn)));
    const int iy = (int)rintf(fmaf(x, sn, y * cs));
    return px(b, g->total, c + (long)iy * p + ix);
}

/* orbd.cu:1351-1468: 32 bytes per keypoint; WTA_K 2 compares pairs (bit j of
   byte b from points 16b+2j, 16b+2j+1), WTA_K 3 and 4 write 2-bit argmax codes
   (including the reference's 8,9,11 fourth triple for WTA_K 3). */
void orc_describe_patch(const orc_geom* g, const uint8_t* blur, const orc_point* pts, int n, int patch_size,
                        int wta_k, uint8_t* desc)
{
    int32_t pat[1024];
    orc_pattern_patch(patch_size, wta_k, pat);
#pragma omp parallel for schedule(dynamic, 32) num_threads(g_threads) if (g_threads > 1)
    for (int i = 0; i < n; i++) {
        float sn, cs;
        orc_sincosf(pts[i].angle, &sn, &cs);
        const long c = center_of(g, &pts[i]);
        const int p = g->p[pts[i].octave];
        uint8_t* out = desc + 32L * i;
        for (int b = 0; b < 32; b++) {
            int val = 0;
            if (wta_k == 2) {
                const int32_t* pp = pat + 2 * 16 * b;
                for (int j = 0; j < 8; j++)
                    val |= (sample(g, blur, c, p, pp, 2 * j, sn, cs) <
                            sample(g, blur, c, p, pp, 2 * j + 1, sn, cs)) << j;
            } else if (wta_k == 3) {
                const int32_t* pp = pat + 2 * 12 * b;
                static const int tri[4][3] = {{0, 1, 2}, {3, 4, 5}, {6, 7, 8}, {8, 9, 11}};
                for (int q = 0; q < 4; q++) {
                    int t0 = sample(g, blur, c, p, pp, tri[q][0], sn, cs);
                    int t1 = sample(g, blur, c, p, pp, tri[q][1], sn, cs);
                    int t2 = sample(g, blur, c, p, pp, tri[q][2], sn, cs);
                    int k = t2 > t1 ? (t2 > t0 ? 2 : 0) : (t1 > t0);
                    val |= k << (2 * q);
                }
            } else {
                const int32_t* pp = pat + 2 * 16 * b;
                for (int q = 0; q < 4; q++) {
                    int t0 = sample(g, blur, c, p, pp, 4 * q + 0, sn, cs);
                    int t1 = sample(g, blur, c, p, pp, 4 * q + 1, sn, cs);
                    int t2 = sample(g, blur, c, p, pp, 4 * q + 2, sn, cs);
                    int t3 = sample(g, blur, c, p, pp, 4 * q + 3, sn, cs);
                    int u = 0, v = 2;
                    if (t1 > t0) { t0 = t1; u = 1; }
                    if (t3 > t2) { t2 = t3; v = 3; }
                    val |= (t0 > t2 ? u : v) << (2 * q);
                }
            }
            out[b] = (uint8_t)val;
        }
    }
}

void orc_describe(const orc_geom* g, const uint8_t* blur, const orc_point* pts, int n, int wta_k, uint8_t* desc)
{
    orc_describe_patch(g, blur, pts, n, 31, wta_k, desc);
}

/* ---------------------------------------------------------------- pipeline */

/* orb.cpp:57-101 + orb.cpp:194-201: pyramid, detect map, NMS and cap,
   orientation on the raw levels, blur, descriptors on the blurred levels. */
int orc_detect_and_compute(const orc_params* prm, const uint8_t* img, int w, int h,
                           orc_point* pts, uint8_t* desc, int* n_out, int* n_survivors)
{
    orc_geom g;
    if (orc_geometry(w, h, prm->noctaves, prm->scale, &g) != 0) return -1;
    uint8_t* raw = (uint8_t*)malloc((size_t)g.total);
    uint8_t* blur = (uint8_t*)malloc((size_t)g.total);
    float* vmap = (float*)malloc(sizeof(float) * (size_t)g.h[0] * g.p[0]);
    int* lmap = (int*)malloc(sizeof(int) * (size_t)g.h[0] * g.p[0]);
    orc_pyramid(&g, img, raw, blur);
    orc_detect_map(&g, raw, prm->fast_threshold, prm->score_type, vmap, lmap);
    int n = orc_nms(&g, vmap, lmap, prm->edge_threshold, prm->max_pts, pts, n_survivors);
    orc_angles(&g, raw, pts, n, prm->patch_size);
    orc_describe_patch(&g, blur, pts, n, prm->patch_size, prm->wta_k, desc);
    *n_out = n;
    free(raw); free(blur); free(vmap); free(lmap);
    return 0;
}

/* ----------------------------------------------------------------- matching */

static inline int ham(const uint8_t* a, const uint8_t* b)
{
    uint64_t x[4], y[4];
    memcpy(x, a, 32);
    memcpy(y, b, 32);
    return __builtin_popcountll(x[0] ^ y[0]) + __builtin_popcountll(x[1] ^ y[1]) +
           __builtin_popcountll(x[2] ^ y[2]) + __builtin_popcountll(x[3] ^ y[3]);
}

/* orbd.cu:1471-1559: best and second-best distance (both start at 256);
   accept iff d1 < d2 && d1 < 64, else match = distance = -1.  An accepted
   best is unique, so the reference's lane order never shows. */
void orc_match(const uint8_t* d1, int n1, const uint8_t* d2, int n2, int* match, int* dist)
{
#pragma omp parallel for schedule(dynamic, 16) num_threads(g_threads) if (g_threads > 1)
    for (int i = 0; i < n1; i++) {
        int b1 = 256, b2 = 256, i1 = 0;
        for (int j = 0; j < n2; j++) {
            int d = ham(d1 + 32L * i, d2 + 32L * j);
            if (d < b1) { b2 = b1; b1 = d; i1 = j; }
            else if (d < b2) b2 = d;
        }
        if (b1 < b2 && b1 < 64) { match[i] = i1; dist[i] = b1; }
        else { match[i] = -1; dist[i] = -1; }
    }
}

/* OpenCV's NORM_HAMMING2 on 256-bit rows: the number of the 128 bit pairs
   (bits 2i, 2i+1 of the little-endian row) in which a and b differ, counted
   pair by pair */
static inline int ham2(const uint8_t* a, const uint8_t* b)
{
    int d = 0;
    for (int i = 0; i < 32; i++) {
        const int x = a[i] ^ b[i];
        for (int k = 0; k < 8; k += 2) d += ((x >> k) & 3) != 0;
    }
    return d;
}

/* exact nearest and second nearest, ties broken by the lowest index:
   (d1, idx1) is the lexicographic minimum of (distance, index), (d2, idx2) the
   minimum over the other entries.  With no second entry idx2 = -1, d2 = 257.
   norm 0: Hamming distance; norm 1: NORM_HAMMING2 (ham2). */
void orc_match_top2_norm(const uint8_t* q, int nq, const uint8_t* db, int ndb, int nthreads, int norm,
                         int* idx1, int* d1, int* idx2, int* d2)
{
#ifdef _OPENMP
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
#else
    (void)nthreads;
#endif
    for (int i = 0; i < nq; i++) {
        int b1 = 257, b2 = 257, j1 = -1, j2 = -1;
        const uint8_t* qi = q + 32L * i;
        for (int j = 0; j < ndb; j++) {
            int d = norm ? ham2(qi, db + 32L * j) : ham(qi, db + 32L * j);
            if (d < b1) { b2 = b1; j2 = j1; b1 = d; j1 = j; }
            else if (d < b2) { b2 = d; j2 = j; }
        }
        idx1[i] = j1; d1[i] = b1; idx2[i] = j2; d2[i] = b2;
    }
}

void orc_match_top2(const uint8_t* q, int nq, const uint8_t* db, int ndb, int nthreads,
                    int* idx1, int* d1, int* idx2, int* d2)
{
    orc_match_top2_norm(q, nq, db, ndb, nthreads, 0, idx1, d1, idx2, d2);
}
