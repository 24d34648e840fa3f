/*
 * Host side of the C ABI (include/orb_capi.h): context, parameter tables,
 * pyramid geometry, frame-set memory and launch sequencing.
 *
 * The reference keeps parameters in module-global __constant__ symbols
 * (orbd.cu:32-41) and reallocates its scratch on every size change
 * (orb.cpp:145-191); here a context owns the tables and a cached frame set
 * (geometry, job tables and HBM buffers for a list of image sizes), so a
 * repeated call with the same sizes allocates nothing and syncs nothing.
 */
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "orb_capi.h"
#include "orb_internal.h"

using namespace orbk;

namespace
{

const int32_t k_pattern31[1024] = {
#include "orb_pattern31.inc"
};

std::string g_create_error;

int align_up(int a, int b) { return (a % b != 0) ? (a - a % b + b) : a; }

/* Pyramid geometry.  Scale 2 follows orb.cpp:145-191: levels
   min(noctaves, (int)log2f(min(w,h)/80) + 1) capped at MAX_OCTAVE (5),
   w_l = w_{l-1} >> 1, NMS radius 2^l.  Another scale s in (1,2):
   w_l = floor(w / s^l + 0.5) while the short side is >= 16 (at most 8 levels),
   NMS radius max(1, floor(s^l + 0.5)).  Pitch alignUp(w_l, 128), levels
   contiguous. */
struct Geom {
    int L = 0, ref = 1;
    double scale = 2.0;
    int w[kMaxLevels] = {}, h[kMaxLevels] = {}, p[kMaxLevels] = {}, radius[kMaxLevels] = {};
    long long off[kMaxLevels] = {};
    double sc[kMaxLevels] = {};
    long long total = 0;
};

bool compute_geom(int w, int h, int noctaves, double scale, Geom& g)
{
    g = Geom();
    if (w <= 0 || h <= 0 || noctaves < 1) return false;
    if (!(scale > 1.0 && scale <= 2.0)) return false;
    g.scale = scale;
    g.ref = scale == 2.0;
    int L;
    if (g.ref) {
        const int m = (w < h ? w : h) / 80;
        if (m < 1) return false;
        L = (int)log2f((float)m) + 1;
        if (L > noctaves) L = noctaves;
        if (L > kRefMaxLevels) L = kRefMaxLevels;
        for (int l = 0; l < L; l++) {
            g.w[l] = l == 0 ? w : g.w[l - 1] >> 1;
            g.h[l] = l == 0 ? h : g.h[l - 1] >> 1;
            g.sc[l] = (double)(1 << l);
            g.radius[l] = 1 << l;
        }
    } else {
        if ((w < h ? w : h) < 16) return false;
        L = 0;
        double sc = 1.0;
        for (int l = 0; l < noctaves && l < kMaxLevels; l++) {
            if (l > 0) sc *= scale;
            const int wl = (int)floor(w / sc + 0.5), hl = (int)floor(h / sc + 0.5);
            if ((wl < hl ? wl : hl) < 16) break;
            g.w[l] = wl;
            g.h[l] = hl;
            g.sc[l] = sc;
            const int r = (int)floor(sc + 0.5);
            g.radius[l] = r < 1 ? 1 : r;
            L++;
        }
    }
    g.L = L;
    long long off = 0;
    for (int l = 0; l < L; l++) {
        g.p[l] = align_up(g.w[l], 128);
        g.off[l] = off;
        off += (long long)g.h[l] * g.p[l];
    }
    g.total = off;
    return L > 0;
}

/* bilinear source position of level index i along an axis of S source and D
   level pixels: (i0 << 8) | frac, centre-aligned, 8 fractional bits */
int resize_tab(int i, int S, int D)
{
    long long num = ((long long)(2 * i + 1) * S - D) * 256;
    long long pos = num / (2LL * D);
    if (pos < 0) pos = 0;
    if (pos > (long long)(S - 1) * 256) pos = (long long)(S - 1) * 256;
    return (int)pos;
}

size_t round256(size_t v) { return (v + 255) & ~(size_t)255; }

struct FrameSet {
    int pins = 0;              /* graphs that recorded launches on this set's buffers */
    unsigned long long last_use = 0;
    std::vector<int> ws, hs;
    int noctaves = 0, border = 0, cap = 0;
    double scale = 0;
    int nframes = 0;
    std::vector<FrameDev> hframes;
    void* arena = nullptr;
    size_t arena_bytes = 0;
    FrameDev* d_frames = nullptr;
    Job* d_pyr[kPyrVariants] = {};
    Job* d_det = nullptr;
    int* d_pyr_map[kPyrVariants] = {};   /* job index of every pyramid unit */
    int* d_det_map = nullptr;            /* job index of every detection tile */
    int pyr_njobs[kPyrVariants] = {}, pyr_units[kPyrVariants] = {}, det_njobs = 0, det_blocks = 0;
    /* autotuner: measured ms per pyramid variant and the choice (-1: not tuned) */
    float pyr_ms[kPyrVariants] = {};
    int pyr_choice = -1;
    int* d_nms_frame = nullptr;
    int nms_blocks = 0;
    Candidate* d_cand = nullptr;
    int* d_blk_count = nullptr;
    uint32_t* d_epochs = nullptr;
    int* d_tabs = nullptr;
};

} // namespace

struct orb_ctx {
    int device = 0;
    int cu_count = 256;        /* compute units of the device (persistent grids) */
    hipStream_t stream = nullptr;
    bool inited = false;
    orb_params prm{};
    int2* d_pattern = nullptr;
    int pyr_forced = -1;       /* pyramid variant forced by orb_tune_set, or -1 */
    /* graph capture: calls are recorded on capture_stream instead of run */
    bool capturing = false;
    hipStream_t user_stream = nullptr;
    hipStream_t capture_stream = nullptr;
    float gauss[4] = {};
    float harris_scale = 0.f;
    int umax[32] = {};
    /* frame sets by frame sizes (and, for sets a graph recorded, buffer
       pointers); fs points at the one of the current call */
    std::vector<FrameSet*> sets;
    FrameSet* fs = nullptr;
    unsigned long long use_clock = 0;
    /* matcher workspace: [merge words: 2 per query | tile counters: 1 per 64
       queries | MFMA per-chunk partials]; the first two regions start as "no
       key" / 0 and every popcount match leaves them so */
    uint32_t* d_match_partial = nullptr;
    size_t match_partial_cap = 0;   /* words of the partial region */
    size_t match_state_q = 0;       /* queries the merge region holds */
    /* batched matching: the device pair table and the host copy it was uploaded from */
    MatchPair* d_pairs = nullptr;
    size_t pairs_cap = 0;
    std::vector<MatchPair> h_pairs;
    int partial_pins = 0;                  /* graphs that recorded the matcher workspace */
    std::vector<uint32_t*> retired_partial;/* replaced workspaces still named by a graph */
    std::vector<FrameSet*> captured_sets;  /* during capture: sets the recorded calls use */
    bool captured_partial = false;
    /* device address of the allocation-free matcher's scratch array on this
       context's device (null if its lookup failed: scratch_err says why) */
    uint32_t* scratch = nullptr;
    std::string scratch_err;
    std::string err;
    /* stage timing: events recorded around launches while prof_mask != 0 */
    unsigned prof_mask = 0;
    std::vector<hipEvent_t> ev_pool;
    struct Rec { int stage; hipEvent_t a, b; };
    std::vector<Rec> recs;
};

namespace
{

int fail(orb_ctx* c, int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    else g_create_error = buf;
    return code;
}

int hip_fail(orb_ctx* c, hipError_t e, const char* what)
{
    return fail(c, e == hipErrorOutOfMemory ? ORB_E_NOMEM : ORB_E_HIP, "%s: %s", what, hipGetErrorString(e));
}

#define HIP_TRY(ctx, call)                                   \
    do {                                                     \
        hipError_t e_ = (call);                              \
        if (e_ != hipSuccess) return hip_fail(ctx, e_, #call); \
    } while (0)

/* setUmax, orbd.cu:66-92 */
void make_umax(int patch_size, int* umax)
{
    const int half = patch_size / 2;
    umax[half + 1] = 0;
    const float v = half * sqrtf(2.f) / 2;
    const int vmax = (int)floorf(v + 1);
    const int vmin = (int)ceilf(v);
    for (int i = 0; i <= vmax; i++) umax[i] = (int)roundf(sqrtf((float)(half * half - i * i)));
    for (int i = half, v0 = 0; i >= vmin; --i) {
        while (umax[v0] == umax[v0 + 1]) ++v0;
        umax[i] = v0;
        ++v0;
    }
}

/* setGaussianKernel, orbd.cu:413-433 */
void make_gauss(float k[4])
{
    const float svar = -1.f / (2.f * 2.f * 2.f);
    float sum = 0.f;
    for (int i = 0; i <= 3; i++) {
        k[i] = expf((float)(i * i) * svar);
        sum = sum + (k[i] + (i == 0 ? 0.f : k[i]));
    }
    sum = 1.f / sum;
    for (int i = 0; i <= 3; i++) k[i] *= sum;
}

/* The additive-feedback generator of glibc's random(3) in its default
   TYPE_3 form, which the reference's srand/rand calls produce
   (orbd.cu:380-392): 31 words seeded by the Park-Miller step
   x -> 16807 x mod (2^31 - 1), then w[n] = w[n-31] + w[n-3] (mod 2^32), the
   first 310 of those discarded, output w[n] >> 1.  The state lives in the
   object, so callers' own rand() streams and other threads are untouched. */
class Type3Rand
{
public:
    /* valid for 0 < seed < 2^31 - 1, where the Park-Miller step is the plain
       modular product (the only seed used here is 0x12345678) */
    explicit Type3Rand(uint32_t seed)
    {
        uint64_t x = seed;
        hist[0] = (uint32_t)x;
        for (int i = 1; i < 31; i++) {
            x = (16807u * x) % 2147483647u;
            hist[i] = (uint32_t)x;
        }
        for (int i = 31; i < 34; i++) hist[i] = hist[i - 31];   /* w[31..33] copy w[0..2] */
        n = 34;
        for (int i = 0; i < 310; i++) step();
    }
    int next() { return (int)(step() >> 1); }

private:
    static constexpr int kLag = 34;   /* holds w[n-34 .. n-1] */
    uint32_t hist[kLag];
    long long n;
    uint32_t step()
    {
        const uint32_t v = hist[(n - 31) % kLag] + hist[(n - 3) % kLag];
        hist[n % kLag] = v;
        n++;
        return v;
    }
};

/* setPattern, orbd.cu:95-410.  The base set of 512 points is the table for
   patch 31; for another patch size p it is drawn as the reference draws it,
   srand(0x34985739) and each coordinate rand() % p - -p / 2 (orbd.cu:359-369:
   subtracting the negated half adds p/2, so the points sit off the patch
   centre in [p/2, p + p/2 - 1]; kept as the reference computes it).  WTA_K 2
   uses the base set; WTA_K 3/4 draw 128 tuples of distinct base points with
   srand(0x12345678). */
void make_pattern(int patch_size, int wta_k, int32_t* xy)
{
    int32_t base[1024];
    if (patch_size == 31) {
        memcpy(base, k_pattern31, sizeof(base));
    } else {
        Type3Rand pick(0x34985739u);
        for (int i = 0; i < 1024; i++) base[i] = pick.next() % patch_size - -patch_size / 2;
    }
    if (wta_k == 2) {
        memcpy(xy, base, sizeof(base));
        return;
    }
    memset(xy, 0, 1024 * sizeof(int32_t));
    Type3Rand rng(0x12345678u);
    for (int t = 0; t < 128; t++) {
        int2* tuple = (int2*)xy + wta_k * t;
        for (int k = 0; k < wta_k; k++) {
            bool fresh = false;
            while (!fresh) {
                const int idx = rng.next() % 512;
                const int2 cand = make_int2(base[2 * idx], base[2 * idx + 1]);
                fresh = true;
                for (int e = 0; e < k && fresh; e++) fresh = !(tuple[e].x == cand.x && tuple[e].y == cand.y);
                if (fresh) tuple[k] = cand;
            }
        }
    }
}

void free_frameset(FrameSet* fs)
{
    if (fs->arena) (void)hipFree(fs->arena);
    delete fs;
}

constexpr int kMaxFreeSets = 4;   /* unpinned frame sets kept for reuse */

struct CallFrames {
    int nf;
    const uint8_t* const* imgs;
    const int* pitches;
    orb::OrbPoint* const* pts;
    uint8_t* const* desc;
    int* d_counts;
    int compute_desc;
};

bool same_buffers(const FrameSet& fs, const CallFrames& cf)
{
    for (int f = 0; f < cf.nf; f++) {
        const FrameDev& F = fs.hframes[f];
        if (F.img != cf.imgs[f] || F.ip != cf.pitches[f] || F.pts != cf.pts[f] ||
            F.desc != (cf.compute_desc ? cf.desc[f] : nullptr) || F.count != cf.d_counts + f)
            return false;
    }
    return true;
}

/* Build geometry, job tables and buffers for a list of frame sizes, or
   reuse a set built before. */
int build_frameset(orb_ctx* c, const std::vector<int>& ws, const std::vector<int>& hs, const CallFrames& cf)
{
    const orb_params& P = c->prm;
    const int border = P.edge_threshold < 3 ? 3 : P.edge_threshold;
    /* a set with these sizes; one a graph recorded only with the same buffers */
    FrameSet* pick = nullptr;
    for (FrameSet* s : c->sets) {
        if (!(s->ws == ws && s->hs == hs && s->noctaves == P.noctaves && s->scale == P.scale_factor &&
              s->border == border && s->cap == P.max_pts))
            continue;
        const bool same = same_buffers(*s, cf);
        if (s->pins > 0 && !same) continue;
        if (!pick || (same && !same_buffers(*pick, cf))) pick = s;
    }
    if (pick) {
        c->fs = pick;
        pick->last_use = ++c->use_clock;
        return ORB_OK;
    }
    if (c->capturing)
        return fail(c, ORB_E_STATE, "new frame sizes or parameters during graph capture: make the same call once "
                                    "before orb_capture_begin");
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    /* keep at most kMaxFreeSets unpinned sets: drop the least recently used */
    for (;;) {
        int nfree = 0;
        FrameSet* lru = nullptr;
        for (FrameSet* s : c->sets)
            if (s->pins == 0) {
                nfree++;
                if (!lru || s->last_use < lru->last_use) lru = s;
            }
        if (nfree < kMaxFreeSets || !lru) break;
        c->sets.erase(std::find(c->sets.begin(), c->sets.end(), lru));
        if (c->fs == lru) c->fs = nullptr;
        free_frameset(lru);
    }
    FrameSet* nfs = new FrameSet();
    c->sets.push_back(nfs);
    c->fs = nfs;
    nfs->last_use = ++c->use_clock;
    FrameSet& fs = *nfs;
    const int nf = (int)ws.size();
    std::vector<Geom> geo(nf);
    for (int f = 0; f < nf; f++) {
        if (ws[f] > 16383 || hs[f] > 16383)
            return fail(c, ORB_E_CAPACITY, "frame %d: %dx%d exceeds 16383", f, ws[f], hs[f]);
        if (!compute_geom(ws[f], hs[f], P.noctaves, P.scale_factor, geo[f]))
            return fail(c, ORB_E_ARG, "frame %d: %dx%d too small for the pyramid (scale %g)", f, ws[f], hs[f],
                        P.scale_factor);
    }
    /* sizes of every sub-buffer */
    std::vector<size_t> pyr_off(nf), key_off(nf);
    std::vector<int> guard(nf), key_guard_rows(nf), tab_base(nf);
    size_t pyr_bytes = 0, key_bytes = 0;
    int tab_ints = 0;
    std::vector<Job> pyr_jobs[kPyrVariants], det_jobs;
    std::vector<int> nms_frame;
    int pyr_units[kPyrVariants] = {}, det_blocks = 0, nms_blocks = 0;
    fs.hframes.assign(nf, FrameDev());
    for (int f = 0; f < nf; f++) {
        const Geom& g = geo[f];
        FrameDev& F = fs.hframes[f];
        memset(&F, 0, sizeof(F));
        F.w0 = g.w[0];
        F.h0 = g.h[0];
        F.p0 = g.p[0];
        F.nlevels = g.L;
        F.ref_mode = g.ref;
        F.border = border;
        F.cap = P.max_pts;
        F.total = g.total;
        int rmax = 1;
        for (int l = 0; l < g.L; l++) {
            F.lw[l] = g.w[l];
            F.lh[l] = g.h[l];
            F.lp[l] = g.p[l];
            F.loff[l] = g.off[l];
            F.radius[l] = g.radius[l];
            F.inv_sc[l] = (float)(1.0 / g.sc[l]);
            F.sc[l] = g.sc[l];
            F.tab_off[l] = tab_ints;
            if (!g.ref && l > 0) tab_ints += 2 * (g.w[l] + g.h[l]);
            if (g.radius[l] > rmax) rmax = g.radius[l];

            if (g.w[l] > 6 && g.h[l] > 6) {
                /* tiles cover [0, w-3) x [0, h-3); pixels outside [3, w-3) x [3, h-3) are masked */
                const int dtx = (g.w[l] - 3 + kDetTW - 1) / kDetTW, dty = (g.h[l] - 3 + kDetTH - 1) / kDetTH;
                det_jobs.push_back(Job{f, l, dtx, det_blocks, 0});
                det_blocks += dtx * dty;
            }
        }
        /* pyramid strips cover the whole pitch (padding written too).  Units
           go in kPyrBands horizontal bands of the frame, each band holding the
           strips of every level over it, so the contiguous unit range an XCD
           runs reads one band of the input into its L2 */
        for (int v = 0; v < kPyrVariants; v++) {
            for (int band = 0; band < kPyrBands; band++) {
                for (int l = 0; l < g.L; l++) {
                    const int ptx = (g.p[l] + kPyrStripW - 1) / kPyrStripW;
                    const int nch = (g.h[l] + kPyrTHs[v] - 1) / kPyrTHs[v];
                    const int c0 = band * nch / kPyrBands, c1 = (band + 1) * nch / kPyrBands;
                    if (c1 <= c0) continue;
                    pyr_jobs[v].push_back(Job{f, l, ptx, pyr_units[v], c0});
                    pyr_units[v] += ptx * (c1 - c0);
                }
            }
        }
        /* flat-pyramid guard: covers the deepest stencil / patch overhang
           beyond either end (Harris +-4, describe +-18 rows/cols, and the
           detection tile staged up to kDetTH + 4 rows past a level's end) */
        guard[f] = (kDetTH + 20) * g.p[0];
        pyr_off[f] = pyr_bytes;
        pyr_bytes += round256((size_t)g.total + 2 * (size_t)guard[f]);
        key_guard_rows[f] = rmax + 1;
        F.rmax = rmax;
        key_off[f] = key_bytes;
        key_bytes += round256(sizeof(unsigned long long) * (size_t)(g.h[0] + 2 * key_guard_rows[f]) * g.p[0]);
        const int rows = g.h[0] - 2 * border, span = g.w[0] - 2 * border;
        F.nms_first = nms_blocks;
        F.nms_x0 = 0; /* segments start at x = 0; pixels outside the border are masked */
        F.nms_rows = rows > 0 && span > 0 ? rows : 0;
        F.nms_segs = rows > 0 && span > 0 ? (g.w[0] - border + kNmsSeg - 1) / kNmsSeg : 0;
        for (int b = 0; b < F.nms_rows * F.nms_segs; b++) nms_frame.push_back(f);
        nms_blocks += F.nms_rows * F.nms_segs;
    }
    /* one arena for everything, zeroed once */
    size_t off = 0;
    auto carve = [&](size_t bytes) { size_t o = off; off += round256(bytes ? bytes : 1); return o; };
    const size_t o_frames = carve(sizeof(FrameDev) * nf);
    size_t o_pyr_jobs[kPyrVariants];
    for (int v = 0; v < kPyrVariants; v++) o_pyr_jobs[v] = carve(sizeof(Job) * pyr_jobs[v].size());
    const size_t o_det_jobs = carve(sizeof(Job) * det_jobs.size());
    /* the job of every unit, so a unit finds its job in one lookup */
    std::vector<int> pyr_map[kPyrVariants], det_map((size_t)det_blocks);
    for (int v = 0; v < kPyrVariants; v++) {
        pyr_map[v].resize((size_t)pyr_units[v]);
        for (size_t j = 0; j < pyr_jobs[v].size(); j++) {
            const int end = j + 1 < pyr_jobs[v].size() ? pyr_jobs[v][j + 1].first_block : pyr_units[v];
            for (int u = pyr_jobs[v][j].first_block; u < end; u++) pyr_map[v][u] = (int)j;
        }
    }
    for (size_t j = 0; j < det_jobs.size(); j++) {
        const int end = j + 1 < det_jobs.size() ? det_jobs[j + 1].first_block : det_blocks;
        for (int u = det_jobs[j].first_block; u < end; u++) det_map[u] = (int)j;
    }
    size_t o_pyr_map[kPyrVariants];
    for (int v = 0; v < kPyrVariants; v++) o_pyr_map[v] = carve(sizeof(int) * pyr_map[v].size());
    const size_t o_det_map = carve(sizeof(int) * det_map.size());
    const size_t o_nms_frame = carve(sizeof(int) * nms_frame.size());
    const size_t o_cand = carve(sizeof(Candidate) * (size_t)nms_blocks * kNmsCap);
    const size_t o_cnt = carve(sizeof(int) * nms_blocks);
    const size_t o_epochs = carve(sizeof(uint32_t) * nf);
    const size_t o_tabs = carve(sizeof(int) * (size_t)tab_ints);
    const size_t o_pyr = carve(2 * pyr_bytes);
    const size_t o_keys = carve(key_bytes);
    std::vector<size_t> o_bits(nf);
    for (int f = 0; f < nf; f++) o_bits[f] = carve(sizeof(uint32_t) * (size_t)geo[f].h[0] * (geo[f].p[0] / 32));
    const size_t total = off;
    void* arena = nullptr;
    hipError_t e = hipMalloc(&arena, total);
    if (e != hipSuccess) return hip_fail(c, e, "hipMalloc(frame set)");
    fs.arena = arena;
    fs.arena_bytes = total;
    char* base = (char*)arena;
    HIP_TRY(c, hipMemsetAsync(arena, 0, total, c->stream));
    fs.d_frames = (FrameDev*)(base + o_frames);
    for (int v = 0; v < kPyrVariants; v++) fs.d_pyr[v] = (Job*)(base + o_pyr_jobs[v]);
    fs.d_det = (Job*)(base + o_det_jobs);
    for (int v = 0; v < kPyrVariants; v++) fs.d_pyr_map[v] = (int*)(base + o_pyr_map[v]);
    fs.d_det_map = (int*)(base + o_det_map);
    fs.d_nms_frame = (int*)(base + o_nms_frame);
    fs.d_cand = (Candidate*)(base + o_cand);
    fs.d_blk_count = (int*)(base + o_cnt);
    fs.d_epochs = (uint32_t*)(base + o_epochs);
    fs.d_tabs = (int*)(base + o_tabs);
    uint8_t* pyr = (uint8_t*)(base + o_pyr);
    unsigned long long* keys = (unsigned long long*)(base + o_keys);
    /* coordinate tables for scale != 2 */
    std::vector<int> tabs((size_t)tab_ints);
    for (int f = 0; f < nf; f++) {
        const Geom& g = geo[f];
        FrameDev& F = fs.hframes[f];
        F.raw = pyr + pyr_off[f] + guard[f];
        F.blur = pyr + pyr_bytes + pyr_off[f] + guard[f];
        F.keys = (unsigned long long*)((char*)keys + key_off[f]) + (size_t)key_guard_rows[f] * g.p[0];
        F.cbits = (uint32_t*)(base + o_bits[f]);
        if (g.ref) continue;
        for (int l = 1; l < g.L; l++) {
            int* t = tabs.data() + F.tab_off[l];
            for (int x = 0; x < g.w[l]; x++) t[x] = resize_tab(x, g.w[0], g.w[l]);
            for (int y = 0; y < g.h[l]; y++) t[g.w[l] + y] = resize_tab(y, g.h[0], g.h[l]);
            for (int x = 0; x < g.w[l]; x++) t[g.w[l] + g.h[l] + x] = (int)floor(x * g.sc[l] + 0.5);
            for (int y = 0; y < g.h[l]; y++) t[2 * g.w[l] + g.h[l] + y] = (int)floor(y * g.sc[l] + 0.5);
        }
    }
    std::vector<uint32_t> epochs(nf, 1u);
    for (int v = 0; v < kPyrVariants; v++)
        HIP_TRY(c, hipMemcpyAsync(fs.d_pyr[v], pyr_jobs[v].data(), sizeof(Job) * pyr_jobs[v].size(),
                                  hipMemcpyHostToDevice, c->stream));
    if (!det_jobs.empty())
        HIP_TRY(c, hipMemcpyAsync(fs.d_det, det_jobs.data(), sizeof(Job) * det_jobs.size(), hipMemcpyHostToDevice, c->stream));
    for (int v = 0; v < kPyrVariants; v++)
        if (!pyr_map[v].empty())
            HIP_TRY(c, hipMemcpyAsync(fs.d_pyr_map[v], pyr_map[v].data(), sizeof(int) * pyr_map[v].size(),
                                      hipMemcpyHostToDevice, c->stream));
    if (!det_map.empty())
        HIP_TRY(c, hipMemcpyAsync(fs.d_det_map, det_map.data(), sizeof(int) * det_map.size(), hipMemcpyHostToDevice,
                                  c->stream));
    if (!nms_frame.empty())
        HIP_TRY(c, hipMemcpyAsync(fs.d_nms_frame, nms_frame.data(), sizeof(int) * nms_frame.size(), hipMemcpyHostToDevice, c->stream));
    if (tab_ints)
        HIP_TRY(c, hipMemcpyAsync(fs.d_tabs, tabs.data(), sizeof(int) * tabs.size(), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipMemcpyAsync(fs.d_epochs, epochs.data(), sizeof(uint32_t) * nf, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    fs.ws = ws;
    fs.hs = hs;
    fs.noctaves = P.noctaves;
    fs.scale = P.scale_factor;
    fs.border = border;
    fs.cap = P.max_pts;
    fs.nframes = nf;
    for (int v = 0; v < kPyrVariants; v++) fs.pyr_njobs[v] = (int)pyr_jobs[v].size();
    for (int v = 0; v < kPyrVariants; v++) fs.pyr_units[v] = pyr_units[v];
    fs.det_njobs = (int)det_jobs.size();
    fs.det_blocks = det_blocks;
    fs.nms_blocks = nms_blocks;
    /* force the first per-call upload */
    for (auto& F : fs.hframes) F.img = nullptr;
    return ORB_OK;
}

/* Events for the stages of one call (null when timing is off). */
const StageEvents* take_events(orb_ctx* c, StageEvents& ev, unsigned stages)
{
    const unsigned m = c->capturing ? 0u : c->prof_mask & stages;
    if (!m) return nullptr;
    for (int s = 0; s < kNumStages; s++) {
        ev.on[s] = (m >> s) & 1u;
        if (!ev.on[s]) continue;
        for (int k = 0; k < 2; k++) {
            hipEvent_t e = nullptr;
            if (!c->ev_pool.empty()) {
                e = c->ev_pool.back();
                c->ev_pool.pop_back();
            } else if (hipEventCreate(&e) != hipSuccess) {
                ev.on[s] = false;
                break;
            }
            (k == 0 ? ev.start[s] : ev.stop[s]) = e;
        }
    }
    return &ev;
}

/* Keep the event pairs of the stages the call launched; the others go back
   to the pool unrecorded (they would otherwise report stale times). */
void commit_events(orb_ctx* c, const StageEvents& ev)
{
    for (int s = 0; s < kNumStages; s++) {
        if (!ev.on[s]) continue;
        if (ev.rec[s]) {
            c->recs.push_back({s, ev.start[s], ev.stop[s]});
        } else {
            c->ev_pool.push_back(ev.start[s]);
            c->ev_pool.push_back(ev.stop[s]);
        }
    }
}

Tables make_tables(const orb_ctx* c)
{
    Tables t;
    memset(&t, 0, sizeof(t));
    t.tabs = c->fs->d_tabs;
    t.pattern = c->d_pattern;
    t.k0 = c->gauss[0];
    t.k1 = c->gauss[1];
    t.k2 = c->gauss[2];
    t.k3 = c->gauss[3];
    t.harris_scale = c->harris_scale;
    t.threshold = c->prm.fast_threshold;
    t.score_type = c->prm.score_type;
    t.wta_k = c->prm.wta_k;
    t.half_patch = c->prm.patch_size / 2;
    t.pattern31 = c->prm.patch_size == 31;
    for (int i = 0; i < 17; i++) t.umax[i] = c->umax[i];
    return t;
}

void set_pyramid_variant(const FrameSet& fs, DetectLaunch& d, int v)
{
    d.pyr_variant = v;
    d.pyr_jobs = fs.d_pyr[v];
    d.pyr_map = fs.d_pyr_map[v];
    d.pyr_njobs = fs.pyr_njobs[v];
    d.pyr_units = fs.pyr_units[v];
    d.pyr_blocks = (fs.pyr_units[v] + 3) / 4;
}

/* Autotuner for the pyramid strip height: on the first call with a new list
   of frame sizes, every variant runs once untimed and kTuneReps times between
   events on the context's stream (the stage is idempotent: it rewrites the
   same levels from the same input); the fastest mean is kept for the frame
   set. */
int tune_pyramid(orb_ctx* c, DetectLaunch& d)
{
    constexpr int kTuneReps = 3;
    FrameSet& fs = *c->fs;
    hipEvent_t a = nullptr, b = nullptr;
    HIP_TRY(c, hipEventCreate(&a));
    hipError_t e = hipEventCreate(&b);
    if (e != hipSuccess) {
        (void)hipEventDestroy(a);
        return hip_fail(c, e, "hipEventCreate");
    }
    int best = 0;
    for (int v = 0; v < kPyrVariants && e == hipSuccess; v++) {
        set_pyramid_variant(fs, d, v);
        e = launch_pyramid(d, v, c->stream);
        if (e == hipSuccess) e = hipEventRecord(a, c->stream);
        for (int r = 0; r < kTuneReps && e == hipSuccess; r++) e = launch_pyramid(d, v, c->stream);
        if (e == hipSuccess) e = hipEventRecord(b, c->stream);
        if (e == hipSuccess) e = hipEventSynchronize(b);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
        fs.pyr_ms[v] = ms / kTuneReps;
        if (fs.pyr_ms[v] < fs.pyr_ms[best]) best = v;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (e != hipSuccess) return hip_fail(c, e, "pyramid autotune");
    fs.pyr_choice = best;
    return ORB_OK;
}

int run_frames(orb_ctx* c, int nf, const uint8_t* const* imgs, const int* ws, const int* hs, const int* pitches,
               orb::OrbPoint* const* pts, uint8_t* const* desc, int* d_counts, int compute_desc)
{
    if (!c->inited) return fail(c, ORB_E_STATE, "orb_init has not been called");
    if (nf <= 0) return fail(c, ORB_E_ARG, "nframes must be positive");
    std::vector<int> vw(ws, ws + nf), vh(hs, hs + nf);
    for (int f = 0; f < nf; f++) {
        if (!imgs[f] || !pts[f] || !d_counts) return fail(c, ORB_E_ARG, "frame %d: null pointer", f);
        if (compute_desc && !desc[f]) return fail(c, ORB_E_ARG, "frame %d: null descriptor buffer", f);
        if (pitches[f] < ws[f]) return fail(c, ORB_E_ARG, "frame %d: pitch %d < width %d", f, pitches[f], ws[f]);
    }
    const CallFrames cf{nf, imgs, pitches, pts, desc, d_counts, compute_desc};
    int rc = build_frameset(c, vw, vh, cf);
    if (rc != ORB_OK) return rc;
    FrameSet& fs = *c->fs;
    if (c->capturing && std::find(c->captured_sets.begin(), c->captured_sets.end(), c->fs) == c->captured_sets.end())
        c->captured_sets.push_back(c->fs);
    bool changed = false;
    for (int f = 0; f < nf; f++) {
        FrameDev& F = fs.hframes[f];
        if (F.img != imgs[f] || F.ip != pitches[f] || F.pts != pts[f] || F.desc != (compute_desc ? desc[f] : nullptr) ||
            F.count != d_counts + f) {
            F.img = imgs[f];
            F.ip = pitches[f];
            F.pts = pts[f];
            F.desc = compute_desc ? desc[f] : nullptr;
            F.count = d_counts + f;
            changed = true;
        }
    }
    if (changed) {
        if (c->capturing)
            return fail(c, ORB_E_STATE, "new image, point or descriptor pointers during graph capture: make the same "
                                        "call once before orb_capture_begin");
        HIP_TRY(c, hipMemcpyAsync(fs.d_frames, fs.hframes.data(), sizeof(FrameDev) * nf, hipMemcpyHostToDevice,
                                  c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
    }
    DetectLaunch d;
    memset(&d, 0, sizeof(d));
    d.frames = fs.d_frames;
    d.nframes = nf;

    d.det_jobs = fs.d_det;
    d.det_map = fs.d_det_map;
    d.det_njobs = fs.det_njobs;
    d.det_blocks = fs.det_blocks;
    d.nms_blocks = fs.nms_blocks;
    d.nms_block_frame = fs.d_nms_frame;
    d.cand = fs.d_cand;
    d.blk_count = fs.d_blk_count;
    d.epochs = fs.d_epochs;
    /* describe: a fixed number of workgroups per frame loop over the points */
    const int need_blocks = (c->prm.max_pts + kDescPerBlock - 1) / kDescPerBlock;
    d.desc_blocks_per_frame = compute_desc ? (need_blocks < kDescMaxBlocks ? need_blocks : kDescMaxBlocks) : 0;
    d.t = make_tables(c);
    int variant = c->pyr_forced;
    if (variant < 0) {
        if (fs.pyr_choice < 0) {
            if (c->capturing) return fail(c, ORB_E_STATE, "the autotuner has not run for these frame sizes");
            rc = tune_pyramid(c, d);
            if (rc != ORB_OK) return rc;
        }
        variant = fs.pyr_choice;
    }
    set_pyramid_variant(fs, d, variant);
    StageEvents ev;
    memset(&ev, 0, sizeof(ev));
    hipError_t e = launch_detect(d, compute_desc, c->stream, take_events(c, ev, 0x2fu));
    commit_events(c, ev);
    if (e != hipSuccess) return hip_fail(c, e, "detect launch");
    return ORB_OK;
}

size_t state_words(size_t nq)
{
    return (2 * nq + (nq + 63) / 64 + 63) & ~(size_t)63;
}

int ensure_partial(orb_ctx* c, size_t words, int nq)
{
    if (c->capturing) c->captured_partial = true;
    if (words <= c->match_partial_cap && (size_t)nq <= c->match_state_q) return ORB_OK;
    if (c->capturing)
        return fail(c, ORB_E_STATE, "matcher workspace grows during graph capture: make the same call once before "
                                    "orb_capture_begin");
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (c->d_match_partial) {
        if (c->partial_pins > 0) c->retired_partial.push_back(c->d_match_partial); /* a graph still names it */
        else (void)hipFree(c->d_match_partial);
    }
    const size_t pw = words > c->match_partial_cap ? words : c->match_partial_cap;
    const size_t q = (size_t)nq > c->match_state_q ? (size_t)nq : c->match_state_q;
    c->d_match_partial = nullptr;
    c->match_partial_cap = 0;
    c->match_state_q = 0;
    HIP_TRY(c, hipMalloc((void**)&c->d_match_partial, (state_words(q) + pw) * sizeof(uint32_t)));
    HIP_TRY(c, hipMemsetAsync(c->d_match_partial, 0xff, 8 * q, c->stream));
    HIP_TRY(c, hipMemsetAsync(c->d_match_partial + 2 * q, 0, 4 * ((q + 63) / 64), c->stream));
    c->match_partial_cap = pw;
    c->match_state_q = q;
    return ORB_OK;
}

/* noalloc: run on the device's scratch array instead of the context's
   workspace, with no allocation, no memset and no stage events */
int run_match(orb_ctx* c, const uint8_t* q, int nq, const int* d_nq, const uint8_t* db, int ndb, const int* d_ndb,
              int method, int* idx1, int* d1, int* idx2, int* d2, orb::OrbPoint* pts, bool noalloc = false)
{
    if (nq < 0 || ndb < 0) return fail(c, ORB_E_ARG, "negative descriptor count");
    if (ndb >= (1 << 23)) return fail(c, ORB_E_CAPACITY, "database of %d entries exceeds 2^23-1", ndb);
    if (noalloc && nq > kNoAllocMaxQueries)
        return fail(c, ORB_E_CAPACITY, "%d queries exceed the allocation-free limit of %d", nq, kNoAllocMaxQueries);
    if (nq == 0) return ORB_OK;
    if (!q || (ndb > 0 && !db)) return fail(c, ORB_E_ARG, "null descriptor pointer");
    if (method != ORB_MATCH_POPCOUNT && method != ORB_MATCH_MFMA && method != ORB_MATCH_HAMMING2)
        return fail(c, ORB_E_ARG, "unknown method %d", method);
    MatchLaunch m;
    memset(&m, 0, sizeof(m));
    if (noalloc) {
        if (!c->scratch)
            return fail(c, ORB_E_STATE, "no scratch array on device %d: %s", c->device, c->scratch_err.c_str());
        m.q = q;
        m.nq = nq;
        m.db = db;
        m.ndb = ndb;
        m.idx1 = idx1;
        m.d1 = d1;
        m.idx2 = idx2;
        m.d2 = d2;
        m.hamming2 = method == ORB_MATCH_HAMMING2;
        const hipError_t e = launch_match_noalloc(m, method, c->scratch, c->stream);
        if (e != hipSuccess) return hip_fail(c, e, "allocation-free match");
        return ORB_OK;
    }
    const size_t words = match_partial_words(nq, ndb, method, &m.nchunks, &m.chunk);
    int rc = ensure_partial(c, words, nq);
    if (rc != ORB_OK) return rc;
    m.q = q;
    m.nq = nq;
    m.d_nq = d_nq;
    m.db = db;
    m.ndb = ndb;
    m.d_ndb = d_ndb;
    m.state = (unsigned long long*)c->d_match_partial;
    m.done = (int*)(c->d_match_partial + 2 * c->match_state_q);
    m.partial = c->d_match_partial + state_words(c->match_state_q);
    m.idx1 = idx1;
    m.d1 = d1;
    m.idx2 = idx2;
    m.d2 = d2;
    m.pts = pts;
    m.hamming2 = method == ORB_MATCH_HAMMING2;
    StageEvents ev;
    memset(&ev, 0, sizeof(ev));
    const StageEvents* pev = take_events(c, ev, 0xc0u);
    hipError_t e = method == ORB_MATCH_MFMA ? launch_match_mfma(m, c->stream, pev)
                                            : launch_match_popcount(m, c->stream, pev);
    commit_events(c, ev);
    if (e != hipSuccess) return hip_fail(c, e, method == ORB_MATCH_MFMA ? "MFMA match" : "popcount match");
    return ORB_OK;
}

} // namespace

/* helpers for the learned-index host code (orb_lindex.cpp) */
int orb__ctx_stream(orb_ctx* c, void** stream)
{
    if (!c) return ORB_E_ARG;
    if (c->capturing) return fail(c, ORB_E_STATE, "learned-index calls cannot be recorded into a graph");
    *stream = (void*)c->stream;
    return ORB_OK;
}

int orb__ctx_fail(orb_ctx* c, int code, const char* msg)
{
    return fail(c, code, "%s", msg);
}

int orb__match_dcount_top2(orb_ctx* c, const uint8_t* q, int nq_cap, const int* d_nq, const uint8_t* db, int ndb,
                           int* idx1, int* d1, int* idx2, int* d2)
{
    return run_match(c, q, nq_cap, d_nq, db, ndb, nullptr, ORB_MATCH_MFMA, idx1, d1, idx2, d2, nullptr);
}

extern "C" {

int orb_ctx_create(orb_ctx** out, int device)
{
    if (!out) return fail(nullptr, ORB_E_ARG, "null context pointer");
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) return fail(nullptr, ORB_E_HIP, "no HIP device available");
    if (device < 0 || device >= n) return fail(nullptr, ORB_E_ARG, "device %d out of range [0,%d)", device, n);
    e = hipSetDevice(device);
    if (e != hipSuccess) return fail(nullptr, ORB_E_HIP, "hipSetDevice: %s", hipGetErrorString(e));
    orb_ctx* c = new orb_ctx();
    c->device = device;
    /* the allocation-free matcher's scratch address on this device, looked
       up once per device with the device current */
    e = match_scratch_setup(device, &c->scratch);
    if (e != hipSuccess) c->scratch_err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0)
        c->cu_count = cus;
    e = hipMalloc((void**)&c->d_pattern, 512 * sizeof(int2));
    if (e != hipSuccess) {
        delete c;
        return fail(nullptr, ORB_E_NOMEM, "hipMalloc(pattern): %s", hipGetErrorString(e));
    }
    *out = c;
    return ORB_OK;
}

int orb_ctx_destroy(orb_ctx* c)
{
    if (!c) return ORB_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (FrameSet* s : c->sets) free_frameset(s);
    c->sets.clear();
    for (uint32_t* p : c->retired_partial) (void)hipFree(p);
    for (auto& r : c->recs) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    for (auto e : c->ev_pool) (void)hipEventDestroy(e);
    if (c->d_pattern) (void)hipFree(c->d_pattern);
    if (c->capture_stream) (void)hipStreamDestroy(c->capture_stream);
    if (c->d_match_partial) (void)hipFree(c->d_match_partial);
    if (c->d_pairs) (void)hipFree(c->d_pairs);
    delete c;
    return ORB_OK;
}

int orb_set_stream(orb_ctx* c, void* stream)
{
    if (!c) return fail(nullptr, ORB_E_ARG, "null context");
    if (c->capturing) {
        if ((hipStream_t)stream == c->user_stream) return ORB_OK;
        return fail(c, ORB_E_STATE, "stream change during graph capture");
    }
    c->stream = (hipStream_t)stream;
    return ORB_OK;
}

struct orb_graph {
    orb_ctx* ctx = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<FrameSet*> sets;   /* pinned: their buffers stay as recorded */
    bool partial = false;          /* pins the matcher workspace */
};

int orb_capture_begin(orb_ctx* c)
{
    if (!c) return fail(nullptr, ORB_E_ARG, "null context");
    if (c->capturing) return fail(c, ORB_E_STATE, "capture already in progress");
    if (!c->capture_stream) HIP_TRY(c, hipStreamCreateWithFlags(&c->capture_stream, hipStreamNonBlocking));
    HIP_TRY(c, hipStreamBeginCapture(c->capture_stream, hipStreamCaptureModeThreadLocal));
    c->user_stream = c->stream;
    c->stream = c->capture_stream;
    c->capturing = true;
    c->captured_sets.clear();
    c->captured_partial = false;
    return ORB_OK;
}

int orb_capture_end(orb_ctx* c, orb_graph** out)
{
    if (!c || !out) return fail(c, ORB_E_ARG, "null argument");
    *out = nullptr;
    if (!c->capturing) return fail(c, ORB_E_STATE, "no capture in progress");
    c->capturing = false;
    c->stream = c->user_stream;
    hipGraph_t g = nullptr;
    HIP_TRY(c, hipStreamEndCapture(c->capture_stream, &g));
    hipGraphExec_t x = nullptr;
    hipError_t e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        return hip_fail(c, e, "hipGraphInstantiate");
    }
    orb_graph* gr = new orb_graph();
    gr->ctx = c;
    gr->graph = g;
    gr->exec = x;
    gr->sets = c->captured_sets;
    for (FrameSet* s : gr->sets) s->pins++;
    gr->partial = c->captured_partial;
    if (gr->partial) c->partial_pins++;
    c->captured_sets.clear();
    *out = gr;
    return ORB_OK;
}

int orb_graph_launch(orb_graph* g)
{
    if (!g) return fail(nullptr, ORB_E_ARG, "null graph");
    orb_ctx* c = g->ctx;
    if (c->capturing) return fail(c, ORB_E_STATE, "graph launch during capture");
    HIP_TRY(c, hipGraphLaunch(g->exec, c->stream));
    return ORB_OK;
}

int orb_graph_destroy(orb_graph* g)
{
    if (!g) return ORB_OK;
    orb_ctx* c = g->ctx;
    (void)hipStreamSynchronize(c->stream); /* replays in flight still use the pinned buffers */
    (void)hipGraphExecDestroy(g->exec);
    (void)hipGraphDestroy(g->graph);
    for (FrameSet* s : g->sets) s->pins--;
    if (g->partial && --c->partial_pins == 0) {
        for (uint32_t* p : c->retired_partial) (void)hipFree(p);
        c->retired_partial.clear();
    }
    delete g;
    return ORB_OK;
}

int orb_init(orb_ctx* c, const orb_params* p)
{
    if (!c || !p) return fail(c, ORB_E_ARG, "null argument");
    if (p->noctaves < 1) return fail(c, ORB_E_ARG, "noctaves must be >= 1");
    if (p->wta_k < 2 || p->wta_k > 4) return fail(c, ORB_E_ARG, "wta_k must be 2, 3 or 4");
    if (p->score_type != orb::HARRIS_SCORE && p->score_type != orb::FAST_SCORE)
        return fail(c, ORB_E_ARG, "unknown score type %d", p->score_type);
    /* above 31 the reference overflows its umax table (orbd.cu:37) */
    if (p->patch_size < 3 || p->patch_size > 31) return fail(c, ORB_E_ARG, "patch_size must be in [3, 31]");
    if (p->fast_threshold < 0 || p->fast_threshold > 255) return fail(c, ORB_E_ARG, "fast_threshold out of [0,255]");
    if (p->max_pts <= 0) return fail(c, ORB_E_ARG, "max_pts must be positive");
    if (p->max_pts > (1 << 22)) return fail(c, ORB_E_CAPACITY, "max_pts above 2^22");
    if (!(p->scale_factor > 1.0 && p->scale_factor <= 2.0)) return fail(c, ORB_E_ARG, "scale_factor must be in (1,2]");
    HIP_TRY(c, hipSetDevice(c->device));
    c->prm = *p;
    make_umax(p->patch_size, c->umax);
    make_gauss(c->gauss);
    {
        const float s = 1.f / (4 * 7 * 255.f); /* setScaleSqSq, orbd.cu:436-441 */
        c->harris_scale = s * s * s * s;
    }
    int32_t pat[1024];
    make_pattern(p->patch_size, p->wta_k, pat);
    HIP_TRY(c, hipMemcpyAsync(c->d_pattern, pat, sizeof(pat), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    c->inited = true;
    return ORB_OK;
}

int orb_num_levels(const orb_ctx* c, int w, int h)
{
    if (!c || !c->inited) return 0;
    Geom g;
    if (!compute_geom(w, h, c->prm.noctaves, c->prm.scale_factor, g)) return 0;
    return g.L;
}

int orb_detect_and_compute(orb_ctx* c, const uint8_t* img, int w, int h, int pitch, orb::OrbPoint* pts,
                           uint8_t* desc, int* d_count, int compute_desc)
{
    if (!c) return fail(nullptr, ORB_E_ARG, "null context");
    return run_frames(c, 1, &img, &w, &h, &pitch, &pts, &desc, d_count, compute_desc);
}

int orb_detect_and_compute_batch(orb_ctx* c, int nf, const uint8_t* const* imgs, const int* ws, const int* hs,
                                 const int* pitches, orb::OrbPoint* const* pts, uint8_t* const* desc, int* d_counts,
                                 int compute_desc)
{
    if (!c) return fail(nullptr, ORB_E_ARG, "null context");
    if (!imgs || !ws || !hs || !pitches || !pts || (compute_desc && !desc))
        return fail(c, ORB_E_ARG, "null frame array");
    return run_frames(c, nf, imgs, ws, hs, pitches, pts, desc, d_counts, compute_desc);
}

int orb_match(orb_ctx* c, orb::OrbPoint* pts1, const uint8_t* desc1, int n1, const uint8_t* desc2, int n2)
{
    if (!c) return fail(nullptr, ORB_E_ARG, "null context");
    if (n1 > 0 && !pts1) return fail(c, ORB_E_ARG, "null point array");
    return run_match(c, desc1, n1, nullptr, desc2, n2, nullptr, ORB_MATCH_POPCOUNT, nullptr, nullptr, nullptr,
                     nullptr, pts1);
}

int orb_match_dcount(orb_ctx* c, orb::OrbPoint* pts1, const uint8_t* desc1, const int* d_n1, int cap1,
                     const uint8_t* desc2, const int* d_n2, int cap2)
{
    if (!c) return fail(nullptr, ORB_E_ARG, "null context");
    if (!pts1 || !d_n1 || !d_n2) return fail(c, ORB_E_ARG, "null pointer");
    return run_match(c, desc1, cap1, d_n1, desc2, cap2, d_n2, ORB_MATCH_POPCOUNT, nullptr, nullptr, nullptr,
                     nullptr, pts1);
}

int orb_match_batch(orb_ctx* c, int npairs, orb::OrbPoint* const* pts1, const uint8_t* const* desc1, const int* d_n1,
                    int cap1, const uint8_t* const* desc2, const int* d_n2, int cap2)
{
    if (!c) return fail(nullptr, ORB_E_ARG, "null context");
    if (npairs < 0 || cap1 < 0 || cap2 < 0) return fail(c, ORB_E_ARG, "negative count or capacity");
    if (cap2 >= (1 << 23)) return fail(c, ORB_E_CAPACITY, "database capacity %d exceeds 2^23-1", cap2);
    if (npairs == 0 || cap1 == 0) return ORB_OK;
    if (c->capturing) return fail(c, ORB_E_STATE, "orb_match_batch cannot be recorded into a graph");
    if (!pts1 || !desc1 || !desc2 || !d_n1 || !d_n2) return fail(c, ORB_E_ARG, "null pointer");
    std::vector<MatchPair> hp((size_t)npairs);
    for (int i = 0; i < npairs; i++) {
        if (!pts1[i] || !desc1[i] || (cap2 > 0 && !desc2[i])) return fail(c, ORB_E_ARG, "pair %d: null pointer", i);
        hp[i] = MatchPair{desc1[i], desc2[i], d_n1 + i, d_n2 + i, pts1[i]};
    }
    const int qcap = match_pairs_qcap(cap1);
    if ((long long)npairs * qcap > (1LL << 30))
        return fail(c, ORB_E_CAPACITY, "%d pairs of %d queries exceed the batched matcher's workspace", npairs, cap1);
    int rc = ensure_partial(c, 0, npairs * qcap);
    if (rc != ORB_OK) return rc;
    const bool same = hp.size() == c->h_pairs.size() &&
                      memcmp(hp.data(), c->h_pairs.data(), sizeof(MatchPair) * hp.size()) == 0;
    if (!same) {
        if ((size_t)npairs > c->pairs_cap) {
            HIP_TRY(c, hipStreamSynchronize(c->stream));
            if (c->d_pairs) (void)hipFree(c->d_pairs);
            c->d_pairs = nullptr;
            c->pairs_cap = 0;
            HIP_TRY(c, hipMalloc((void**)&c->d_pairs, sizeof(MatchPair) * (size_t)npairs));
            c->pairs_cap = (size_t)npairs;
        }
        HIP_TRY(c, hipMemcpyAsync(c->d_pairs, hp.data(), sizeof(MatchPair) * hp.size(), hipMemcpyHostToDevice,
                                  c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        c->h_pairs.swap(hp);
    }
    StageEvents ev;
    memset(&ev, 0, sizeof(ev));
    const StageEvents* pev = take_events(c, ev, 0x40u);
    hipError_t e = launch_match_pairs(c->d_pairs, npairs, cap1, cap2, qcap, (unsigned long long*)c->d_match_partial,
                                      (int*)(c->d_match_partial + 2 * c->match_state_q), c->stream, pev);
    commit_events(c, ev);
    if (e != hipSuccess) return hip_fail(c, e, "batched match");
    return ORB_OK;
}

int orb_match_top2(orb_ctx* c, const uint8_t* q, int nq, const uint8_t* db, int ndb, int method, int* idx1, int* d1,
                   int* idx2, int* d2)
{
    if (!c) return fail(nullptr, ORB_E_ARG, "null context");
    return run_match(c, q, nq, nullptr, db, ndb, nullptr, method, idx1, d1, idx2, d2, nullptr);
}

int orb_match_top2_noalloc(orb_ctx* c, const uint8_t* q, int nq, const uint8_t* db, int ndb, int method, int* idx1,
                           int* d1, int* idx2, int* d2)
{
    if (!c) return fail(nullptr, ORB_E_ARG, "null context");
    return run_match(c, q, nq, nullptr, db, ndb, nullptr, method, idx1, d1, idx2, d2, nullptr, true);
}

int orb_export_keypoints(orb_ctx* c, const orb::OrbPoint* pts, const uint8_t* desc, const int* d_n, int cap,
                         orb_cv_keypoint* kp, uint8_t* desc_out, int* d_nout)
{
    if (!c) return fail(nullptr, ORB_E_ARG, "null context");
    if (!c->inited) return fail(c, ORB_E_STATE, "orb_init has not been called");
    if (cap < 0) return fail(c, ORB_E_ARG, "negative capacity");
    if (cap == 0) return ORB_OK;
    if (!pts || !kp) return fail(c, ORB_E_ARG, "null point pointer");
    ExportArgs A;
    memset(&A, 0, sizeof(A));
    const double sf = c->prm.scale_factor;
    A.ref_mode = sf == 2.0;
    A.levels = c->prm.noctaves < kMaxLevels ? c->prm.noctaves : kMaxLevels;
    A.patch = (float)c->prm.patch_size;
    double sc = 1.0;
    for (int l = 0; l < kMaxLevels; l++) {
        if (l > 0) sc *= sf;   /* the product detection uses for s^l */
        A.sc[l] = A.ref_mode ? (double)(1 << l) : sc;
        A.lsc[l] = (float)pow(sf, (double)l);
    }
    hipError_t e = launch_export(pts, desc, d_n, cap, c->prm.retain_topn, A, kp, desc_out, d_nout, c->stream);
    if (e != hipSuccess) return hip_fail(c, e, "export launch");
    return ORB_OK;
}

int orb_sync(orb_ctx* c)
{
    if (!c) return fail(nullptr, ORB_E_ARG, "null context");
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return ORB_OK;
}

const char* orb_last_error(const orb_ctx* c)
{
    return c ? c->err.c_str() : g_create_error.c_str();
}

size_t orb_workspace_bytes(const orb_ctx* c)
{
    if (!c) return 0;
    size_t b = c->d_match_partial ? (c->match_partial_cap + state_words(c->match_state_q)) * sizeof(uint32_t) : 0;
    for (const FrameSet* s : c->sets) b += s->arena_bytes;
    return b;
}

int orb_profile_enable(orb_ctx* c, unsigned stage_mask)
{
    if (!c) return fail(nullptr, ORB_E_ARG, "null context");
    c->prof_mask = stage_mask & ((1u << ORB_NUM_STAGES) - 1u) & ~(1u << 4); /* stage 4 is unused */
    return ORB_OK;
}

int orb_profile_read(orb_ctx* c, double* ms_sum, int* counts)
{
    if (!c || !ms_sum || !counts) return fail(c, ORB_E_ARG, "null argument");
    for (int s = 0; s < ORB_NUM_STAGES; s++) {
        ms_sum[s] = 0.0;
        counts[s] = 0;
    }
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    for (auto& r : c->recs) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
            ms_sum[r.stage] += ms;
            counts[r.stage] += 1;
        }
        c->ev_pool.push_back(r.a);
        c->ev_pool.push_back(r.b);
    }
    c->recs.clear();
    (void)hipGetLastError(); /* an unrecorded pair leaves an error that is not the caller's */
    return ORB_OK;
}

int orb_tune_set(orb_ctx* c, int stage, int variant)
{
    if (!c) return fail(nullptr, ORB_E_ARG, "null context");
    if (stage != ORB_TUNE_PYRAMID) return fail(c, ORB_E_ARG, "unknown tuning stage %d", stage);
    if (variant < -1 || variant >= kPyrVariants) return fail(c, ORB_E_ARG, "variant %d out of range", variant);
    c->pyr_forced = variant;
    return ORB_OK;
}

int orb_tune_get(orb_ctx* c, int stage, int* variant, int* nvariants, float* ms)
{
    if (!c || !variant || !nvariants) return fail(c, ORB_E_ARG, "null argument");
    if (stage != ORB_TUNE_PYRAMID) return fail(c, ORB_E_ARG, "unknown tuning stage %d", stage);
    *nvariants = kPyrVariants;
    *variant = c->pyr_forced >= 0 ? c->pyr_forced : (c->fs ? c->fs->pyr_choice : -1);
    if (ms)
        for (int v = 0; v < kPyrVariants; v++) ms[v] = c->fs ? c->fs->pyr_ms[v] : 0.f;
    return ORB_OK;
}

int orb_debug_pattern(int patch_size, int wta_k, int32_t* xy)
{
    if (!xy || wta_k < 2 || wta_k > 4 || patch_size < 3 || patch_size > 31)
        return fail(nullptr, ORB_E_ARG, "orb_debug_pattern: patch_size 3..31, wta_k 2..4 and a buffer");
    make_pattern(patch_size, wta_k, xy);
    return ORB_OK;
}

int orb_debug_sincos(orb_ctx* c, const float* in, int n, float* s, float* co)
{
    if (!c) return fail(nullptr, ORB_E_ARG, "null context");
    hipError_t e = launch_debug_sincos(in, n, s, co, c->stream);
    return e == hipSuccess ? ORB_OK : hip_fail(c, e, "debug sincos");
}

int orb_debug_atan2(orb_ctx* c, const float* yx, int n, float* out)
{
    if (!c) return fail(nullptr, ORB_E_ARG, "null context");
    hipError_t e = launch_debug_atan2(yx, n, out, c->stream);
    return e == hipSuccess ? ORB_OK : hip_fail(c, e, "debug atan2");
}

int orb_debug_pyramid(orb_ctx* c, int frame, int which, uint8_t* d_out, size_t cap, size_t* nbytes)
{
    if (!c || !nbytes) return fail(c, ORB_E_ARG, "null argument");
    if (!c->fs || frame < 0 || frame >= c->fs->nframes)
        return fail(c, ORB_E_STATE, "no frame %d in the last call", frame);
    if (which != 0 && which != 1) return fail(c, ORB_E_ARG, "which must be 0 (raw) or 1 (blurred)");
    const FrameDev& F = c->fs->hframes[frame];
    *nbytes = (size_t)F.total;
    if (!d_out) return ORB_OK;
    if (cap < (size_t)F.total) return fail(c, ORB_E_CAPACITY, "buffer of %zu bytes < %lld", cap, F.total);
    HIP_TRY(c, hipMemcpyAsync(d_out, which ? F.blur : F.raw, (size_t)F.total, hipMemcpyDeviceToDevice, c->stream));
    return ORB_OK;
}

} // extern "C"
