/*
 * Internal layout shared by the host orchestration (orb_capi.cpp) and the
 * gfx950 kernels (orb_kernels.hip, orb_match.hip).
 *
 * A "frame set" is a list of images processed by one launch per stage.  Each
 * frame owns, in HBM:
 *   raw   flat pyramid of unblurred levels: levels contiguous, row pitch
 *         alignUp(w_l,128), with GUARD zero bytes before level 0 and after the
 *         last level, so every stencil / patch read stays inside the allocation
 *         and reads 0 outside the pyramid (the reference's flat addressing,
 *         orb.cpp:151-175);
 *   blur  the same layout holding the 7x7 Gaussian of every level;
 *   keys  the level-0 score map, one uint64 per level-0 pixel (pitch p0) with
 *         guard rows for the NMS window: key = epoch<<34 | score_bits<<3 |
 *         (7 - level).  A max over keys is the reference's "strictly greater
 *         score wins, lowest level wins ties" (orbd.cu:750-754), and a key of
 *         an older epoch reads as empty, so the map is never cleared.
 */
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "orb_capi.h"
#include "orb_structures.h"

namespace orbk
{

constexpr int kMaxLevels = 8;
constexpr int kRefMaxLevels = 5;     /* MAX_OCTAVE, orbd.cu:18 */
constexpr int kEpochShift = 34;
constexpr uint32_t kEpochLimit = 1u << 30;

/* pyramid unit: one wavefront per strip of 240 columns (60 lanes x 4, a
   multiple of 16 so every 16-byte store of a row lies in one strip) x TH
   rows of one level; 4 units per 256-thread workgroup.  The strip heights
   are the autotuner's variants (TH + 6 walked rows fit one wavefront; taller
   strips, measured up to 122 rows, were slower at batch scale). */
constexpr int kPyrStripW = 240;
constexpr int kPyrBands = 8;         /* horizontal bands per frame in the unit order */
constexpr int kPyrVariants = 5;
constexpr int kPyrTHs[kPyrVariants] = {18, 26, 34, 42, 58};
/* detection tile: 128 x 32 pixels, 256 threads */
constexpr int kDetTW = 128, kDetTH = 32;
/* NMS unit: one wavefront per level-0 row segment of 2048 pixels (64 words
   of the corner bitmap); at most every other pixel of a row survives, so
   kNmsCap candidates fit a segment */
constexpr int kNmsSeg = 2048;
constexpr int kNmsCap = kNmsSeg / 2;
/* describe: one wavefront per keypoint, 4 per workgroup, at most
   kDescMaxBlocks workgroups per frame looping over the points */
constexpr int kDescPerBlock = 4;
constexpr int kDescMaxBlocks = 512;

struct FrameDev {
    /* per call */
    const uint8_t* img;
    int ip;                    /* input row pitch */
    orb::OrbPoint* pts;
    uint8_t* desc;
    int* count;
    /* geometry */
    int w0, h0, p0;
    int nlevels, ref_mode;
    int border, cap;
    int lw[kMaxLevels], lh[kMaxLevels], lp[kMaxLevels], radius[kMaxLevels];
    long long loff[kMaxLevels];
    float inv_sc[kMaxLevels];  /* 1/s^l, level-0 -> level coordinate (scale != 2) */
    double sc[kMaxLevels];     /* s^l, level -> level-0 coordinate (scale != 2) */
    int tab_off[kMaxLevels];   /* into the int table: xsrc[w], ysrc[h], x2l0[w], y2l0[h] */
    long long total;
    uint8_t* raw;              /* level-0 start of the raw flat pyramid */
    uint8_t* blur;             /* level-0 start of the blurred flat pyramid */
    unsigned long long* keys;  /* row 0 of the key map */
    uint32_t* cbits;           /* corner bitmap: bit x of row y set where a scored corner landed
                                  (p0/32 words per row; the NMS clears the words it reads) */
    /* NMS units: rows [border, h0-border), nms_segs segments of kNmsSeg pixels per
       row from x = 0, from unit nms_first (also the index of the unit's segment) */
    int nms_first, nms_rows, nms_segs, nms_x0;
    int rmax;                  /* largest NMS radius of the frame's levels */
};

/* A contiguous range of work units covering one (frame, level) region:
   tiles_x units per row of tiles, tile rows from row0. */
struct Job {
    int frame, level, tiles_x, first_block, row0;
};

/* NMS survivor: pos = level << 28 | y << 14 | x (level-0 coordinates), score = float bits */
struct Candidate {
    uint32_t pos;
    uint32_t score;
};

struct Tables {
    const int* tabs;          /* resize / coordinate tables */
    const int2* pattern;      /* 512 sampling points */
    float k0, k1, k2, k3;     /* Gaussian taps */
    float harris_scale;
    int threshold, score_type, wta_k, half_patch;
    int pattern31;            /* the sampling points are the patch-31 table (patch size 31) */
    int umax[17];
};

struct DetectLaunch {
    const FrameDev* frames;
    int nframes;
    /* jobs sorted by first_block, and the job index of every unit (pyramid:
       one wavefront per unit; detection: one workgroup per unit) */
    const Job* pyr_jobs; const int* pyr_map; int pyr_njobs; int pyr_blocks; int pyr_units; int pyr_variant;
    const Job* det_jobs; const int* det_map; int det_njobs; int det_blocks;
    int nms_blocks;             /* NMS units (row segments) */
    const int* nms_block_frame; /* frame of each NMS unit */
    Candidate* cand;            /* kNmsCap per NMS unit */
    int* blk_count;             /* survivors per NMS unit */
    uint32_t* epochs;           /* per frame */
    int max_blocks_per_frame;
    int desc_blocks_per_frame;
    Tables t;
};

/* Stage ids for event timing (orb_profile_read). */
enum Stage { kStPyramid = 0, kStDetect, kStNms, kStSelect, kStUnused4, kStDescribe, kStMatch, kStMatchMerge,
             kNumStages };

/* Optional start/stop events per stage, recorded around each launch. */
struct StageEvents {
    hipEvent_t start[kNumStages];
    hipEvent_t stop[kNumStages];
    bool on[kNumStages];
    mutable bool rec[kNumStages];   /* set by launch_stage: the stage was launched */
};

/* orb_kernels.hip */
hipError_t launch_detect(const DetectLaunch& d, int compute_desc, hipStream_t s, const StageEvents* ev);
/* the pyramid stage alone with strip-height variant `variant` (autotuner) */
hipError_t launch_pyramid(const DetectLaunch& d, int variant, hipStream_t s, const StageEvents* ev = nullptr);
/* OpenCV-layout export: per-level scales of the context's geometry */
struct ExportArgs {
    double sc[kMaxLevels];  /* s^l: level -> level-0 coordinate, as detection rounds it */
    float lsc[kMaxLevels];  /* OpenCV layerScale: (float)pow(scale, l) */
    float patch;            /* patch size */
    int ref_mode, levels;
};
hipError_t launch_export(const orb::OrbPoint* pts, const uint8_t* desc, const int* d_n, int cap, int retain,
                         const ExportArgs& A, orb_cv_keypoint* kp, uint8_t* desc_out, int* d_nout, hipStream_t s);
hipError_t launch_debug_sincos(const float* in, int n, float* s, float* c, hipStream_t st);
hipError_t launch_debug_atan2(const float* yx, int n, float* out, hipStream_t st);

/* orb_match.hip */
struct MatchLaunch {
    const uint8_t* q; int nq; const int* d_nq;
    const uint8_t* db; int ndb; const int* d_ndb;
    uint32_t* partial; int nchunks; int chunk;   /* MFMA: per-chunk top-2 of each query */
    unsigned long long* state; int* done;       /* popcount: merge word per query, counter per 64-query tile */
    int* idx1; int* d1; int* idx2; int* d2;   /* top-2 outputs (may be null) */
    orb::OrbPoint* pts;                       /* reference-rule output (may be null) */
    bool hamming2;                            /* popcount matcher: count differing bit pairs */
};
hipError_t launch_match_popcount(const MatchLaunch& m, hipStream_t s, const StageEvents* ev);
/* one pair of a batched reference-rule match: queries q (count *n1), database db (count *n2) */
struct MatchPair {
    const uint8_t* q;
    const uint8_t* db;
    const int* n1;
    const int* n2;
    orb::OrbPoint* pts;
};
/* npairs reference-rule matches in one launch (FP4 MFMA); state/done hold
   qcap = match_pairs_qcap(cap1) queries per pair (merge words all ones,
   counters zero, as the launch leaves them) */
hipError_t launch_match_pairs(const MatchPair* d_pairs, int npairs, int cap1, int cap2, int qcap,
                              unsigned long long* state, int* done, hipStream_t s, const StageEvents* ev);
int match_pairs_qcap(int cap1);
hipError_t launch_match_mfma(const MatchLaunch& m, hipStream_t s, const StageEvents* ev);

/* Allocation-free matching (orb_match_top2_noalloc): a file-scope device
   array in orb_match.hip holds the matcher workspace for up to
   kNoAllocMaxQueries queries.  match_scratch_setup (device current) looks up
   that array's address on `device` once and returns it; later calls return
   the cached address.  launch_match_noalloc runs orb_match_top2's kernels on
   it (method 0/2 popcount, 1 MFMA), resetting what must start cleared with a
   kernel, and allocates nothing. */
constexpr int kNoAllocMaxQueries = ORB_NOALLOC_MAX_QUERIES;
constexpr int kMaxScratchDevices = 64;
size_t match_scratch_bytes();
hipError_t match_scratch_setup(int device, uint32_t** ptr);
hipError_t launch_match_noalloc(MatchLaunch m, int method, uint32_t* scratch, hipStream_t s);

/* Launch one kernel of stage st.  When the stage is timed, its events go with
   the dispatch itself (hipExtLaunchKernel marks the kernel's own start and
   end), so timing adds no packets to the stream and no bubble between
   kernels. */
template <typename F, typename... Args>
inline void launch_stage(const StageEvents* ev, int st, F kernel, dim3 grid, dim3 block, hipStream_t s,
                         Args... args)
{
    if (ev && ev->on[st]) {
        hipExtLaunchKernelGGL(kernel, grid, block, 0, s, ev->start[st], ev->stop[st], 0, args...);
        ev->rec[st] = true;
    } else {
        hipLaunchKernelGGL(kernel, grid, block, 0, s, args...);
    }
}
size_t match_partial_words(int nq, int ndb, int method, int* nchunks, int* chunk);

} // namespace orbk
