/*
 * C ABI of the MI355X ORB library (liborb_amd.so).
 *
 * Plain pointers and sizes only.  Device pointers are HIP device memory on the
 * context's device; every call is asynchronous on the context's stream unless
 * it says otherwise.  Every entry point returns ORB_OK or a negative ORB_E*
 * code; orb_last_error() gives the message.  A context is not thread-safe:
 * calls on one context must be serialised by the caller.
 *
 * Each entry point names the reference interface it replaces (file:line in
 * Accustomer/CUDA-ORB).  The reference keeps its tables in module-global
 * __constant__ memory (orbd.cu:32-41); here they live in the context.
 */
#pragma once

#include <stddef.h>
#include <stdint.h>

#include "orb_structures.h"

#ifdef __cplusplus
extern "C" {
#define ORB_POINT orb::OrbPoint
#else
#define ORB_POINT struct OrbPoint
#endif

#define ORB_OK 0
#define ORB_E_ARG -1      /* invalid argument (null pointer, size, parameter) */
#define ORB_E_STATE -2    /* orb_init not called, or wrong call order */
#define ORB_E_HIP -3      /* HIP runtime error */
#define ORB_E_NOMEM -4    /* device allocation failed */
#define ORB_E_CAPACITY -5 /* a size exceeds a compiled-in or configured limit */

#define ORB_MAX_LEVELS 8

/* Detector parameters: the arguments of Orbor::init (orb.h:18-19) plus the
   pyramid scale factor (2.0 = the reference's nearest-neighbour octaves). */
typedef struct {
    int noctaves;         /* pyramid levels requested (orb.h:36) */
    int edge_threshold;   /* border without keypoints (orb.h:40) */
    int wta_k;            /* 2, 3 or 4 (orb.h:42) */
    int score_type;       /* HARRIS_SCORE or FAST_SCORE (orb.h:47) */
    int patch_size;       /* 3..31 (orb.h:49); 31 uses the reference's table, others its random points */
    int fast_threshold;   /* FAST intensity threshold (orb.h:51) */
    int retain_topn;      /* unused by detection, as in the reference (orb.cpp:41); orb_export_keypoints honours it */
    int max_pts;          /* keypoints kept per image (orb.h:55) */
    double scale_factor;  /* 2.0: reference octaves; (1,2): bilinear levels */
} orb_params;

typedef struct orb_ctx orb_ctx;

/* Create a context on HIP device `device`, using the device's null stream.
   Replaces the implicit global state of orbd.cu:32-41 and Orbor() (orb.cpp:14). */
int orb_ctx_create(orb_ctx** ctx, int device);

/* Free a context and its device memory (Orbor::~Orbor, orb.cpp:19-29). */
int orb_ctx_destroy(orb_ctx* ctx);

/* Stream for all later asynchronous calls (hipStream_t as void*; NULL = null
   stream).  No counterpart in the reference, which only uses stream 0. */
int orb_set_stream(orb_ctx* ctx, void* stream);

/* Upload parameters and tables: replaces setMaxNumPoints, setFastThresholdLUT,
   setUmax, setPattern, setGaussianKernel, setScaleSqSq (orbd.h:10-30,
   orbd.cu:45-441) as called by Orbor::init (orb.cpp:32-54). */
int orb_init(orb_ctx* ctx, const orb_params* params);

/* Number of pyramid levels the geometry rule gives for a w x h image
   (orb.cpp:148 for scale 2).  Host only. */
int orb_num_levels(const orb_ctx* ctx, int width, int height);

/* Detect, orient and describe one image (orb.cpp:57-101, orbd.cu:1564-1783).
   d_img: width x height uint8, row pitch `pitch` (the reference requires
   alignUp(width,128), orb.cpp:162).  Writes up to max_pts points to d_pts,
   32-byte descriptors to d_desc (may be NULL when compute_desc is 0) and the
   point count to *d_count (a device int).  Points are in raster order; if more
   survive NMS, the max_pts with the highest score are kept.  Asynchronous. */
int orb_detect_and_compute(orb_ctx* ctx, const uint8_t* d_img, int width, int height, int pitch,
                           ORB_POINT* d_pts, uint8_t* d_desc, int* d_count, int compute_desc);

/* The same for `nframes` images of any sizes in one launch per stage.
   Host arrays of length nframes hold the per-frame device pointers and sizes;
   d_counts is a device array of nframes ints; frame f writes up to max_pts
   points at d_pts[f] and descriptors at d_desc[f].  No counterpart in the
   reference (one image per call there). */
int orb_detect_and_compute_batch(orb_ctx* ctx, int nframes, const uint8_t* const* d_imgs,
                                 const int* widths, const int* heights, const int* pitches,
                                 ORB_POINT* const* d_pts, uint8_t* const* d_desc, int* d_counts,
                                 int compute_desc);

/* Reference match rule (orbd.cu:1471-1559, orb.cpp:104-113): for each of the
   n1 points of d_pts1 the best and second-best Hamming distance over d_desc2;
   match/distance = (best index, best distance) iff best < second && best < 64,
   else -1/-1.  Counts are host ints (n1 == 0 is a no-op).  Asynchronous. */
int orb_match(orb_ctx* ctx, ORB_POINT* d_pts1, const uint8_t* d_desc1, int n1,
              const uint8_t* d_desc2, int n2);

/* orb_match with the two counts read from device ints (as written by
   orb_detect_and_compute), so detection and matching need no host round trip;
   cap1/cap2 bound the counts. */
int orb_match_dcount(orb_ctx* ctx, ORB_POINT* d_pts1, const uint8_t* d_desc1, const int* d_n1,
                     int cap1, const uint8_t* d_desc2, const int* d_n2, int cap2);

/* npairs reference-rule matches in one launch: pair i matches the d_n1[i]
   descriptors at d_desc1[i] (at most cap1) against the d_n2[i] at d_desc2[i]
   (at most cap2) and writes match/distance into d_pts1[i], as
   orb_match_dcount does for one pair.  d_pts1, d_desc1, d_desc2 are host
   arrays of device pointers; d_n1 and d_n2 are device int arrays of npairs
   counts (the d_counts of a batched detection, for instance).  Not
   recordable into a graph.  No reference counterpart beyond Orbor::match
   (orb.cpp:104-113) per pair. */
int orb_match_batch(orb_ctx* ctx, int npairs, ORB_POINT* const* d_pts1, const uint8_t* const* d_desc1,
                    const int* d_n1, int cap1, const uint8_t* const* d_desc2, const int* d_n2, int cap2);

#define ORB_MATCH_POPCOUNT 0 /* VALU xor + popcount */
#define ORB_MATCH_MFMA 1     /* +/-1 INT8 contraction on the matrix cores */
#define ORB_MATCH_HAMMING2 2 /* VALU, OpenCV's NORM_HAMMING2: differing 2-bit pairs (WTA_K 3/4 codes) */

/* Exact brute-force nearest and second-nearest neighbours of nq query
   descriptors in a database of ndb (both 32-byte rows).  Ties go to the
   lowest index: (d1, idx1) is the minimum of (distance, index), (d2, idx2)
   the minimum over the other entries; with no entry idx = -1, d = 257.
   Any output pointer may be NULL.  Generalises gHammingMatch (orbd.cu:1471)
   to raw top-2 output.  method: ORB_MATCH_POPCOUNT and ORB_MATCH_MFMA give
   identical Hamming results; ORB_MATCH_HAMMING2 counts the 128 bit pairs
   (bits 2i, 2i+1) that differ, OpenCV's norm for WTA_K 3/4 descriptors. */
int orb_match_top2(orb_ctx* ctx, const uint8_t* d_query, int nq, const uint8_t* d_db, int ndb,
                   int method, int* d_idx1, int* d_d1, int* d_idx2, int* d_d2);

/* Largest nq orb_match_top2_noalloc accepts. */
#define ORB_NOALLOC_MAX_QUERIES 524288

/* orb_match_top2 without any allocation.  No counterpart in the reference:
   it is for callers that cannot allocate device memory (code under HIP graph
   capture, tight serving loops, harnesses that forbid hipMalloc).  Same
   arguments, results and methods as orb_match_top2, but the matcher
   workspace is a fixed device array of the library (about 10 MiB per GPU,
   held until the process ends) instead of the context's growing workspace:
   nq above ORB_NOALLOC_MAX_QUERIES is rejected with ORB_E_CAPACITY, and no
   hipMalloc, hipFree or memset is issued (what must start cleared is reset by
   a kernel).  Every call on one GPU shares that array, from any context:
   calls that overlap on one GPU (different streams or threads) race and give
   wrong results; serialise them.  Asynchronous on the context's stream. */
int orb_match_top2_noalloc(orb_ctx* ctx, const uint8_t* d_query, int nq, const uint8_t* d_db, int ndb,
                           int method, int* d_idx1, int* d_d1, int* d_idx2, int* d_d2);

/* A keypoint in the memory layout of OpenCV's cv::KeyPoint (28 bytes). */
typedef struct {
    float x, y;     /* pt: level coordinate x layerScale, layerScale = (float)pow(scale_factor, octave) */
    float size;     /* patch_size x layerScale */
    float angle;    /* degrees in [0, 360) */
    float response; /* the detection score */
    int octave;     /* pyramid level */
    int class_id;   /* -1 */
} orb_cv_keypoint;

/* OpenCV-layout export of a keypoint set written by orb_detect_and_compute
   with this context's current parameters: d_n (device int, or NULL for
   cap points) counts the points at d_pts, at most cap.  The retain_topn of
   orb_init, stored and unused by detection as in the reference (orb.cpp:41),
   is honoured here: when it is > 0 and below the count, the retain_topn
   highest scores are kept (ties in raster order).  Kept points stay in raster
   order; d_kp[0..*d_nout) receives them and, when d_desc and d_desc_out are
   both non-NULL, d_desc_out their descriptor rows.  d_kp and d_desc_out hold
   cap entries.  No reference counterpart: the reference hands back OrbPoints
   (orb_structures.h:8-35) that its demo draws field by field
   (main.cpp:27-38), beside a separate OpenCV ORB run (main.cpp:309-329). */
int orb_export_keypoints(orb_ctx* ctx, const ORB_POINT* d_pts, const uint8_t* d_desc, const int* d_n, int cap,
                         orb_cv_keypoint* d_kp, uint8_t* d_desc_out, int* d_nout);

/* Wait for all work on the context's stream (the reference synchronises after
   every launch, orbd.cu:1598-1791). */
int orb_sync(orb_ctx* ctx);

/* Last error message of the context (or of context creation when ctx is NULL). */
const char* orb_last_error(const orb_ctx* ctx);

/* Device scratch bytes the context currently holds (pyramids, maps, lists). */
size_t orb_workspace_bytes(const orb_ctx* ctx);

/* Seeded synthetic test images (host memory, no GPU).  Image A: `nshapes`
   filled rectangles/ellipses of intensity U[0,255] plus 3-octave value noise
   of amplitude 30.  orb_synth_warp: A warped by a similarity homography
   (rotation U[-15,15] deg, scale U[0.9,1.1], shift U[-20,20] px) plus N(0,3^2)
   noise; H (row-major 3x3, maps A coordinates to B) is written if non-NULL.
   Both write width x height bytes with row pitch `pitch`. */
int orb_synth_image(uint64_t seed, int width, int height, int pitch, int nshapes, uint8_t* out);
int orb_synth_warp(uint64_t seed, const uint8_t* src, int width, int height, int pitch,
                   uint8_t* out, double* H);

/* ------------------------------------------------------------------------
   Learned bucket matcher (no counterpart in the reference).  A database of
   256-bit descriptors is partitioned into buckets by k-modes in Hamming
   space; a BF16 MLP (256 -> hidden -> hidden -> nbuckets, GELU, FP32
   accumulate) trained on (query, exact nearest neighbour) pairs maps a query
   to buckets; a query searches its nprobe most likely buckets exactly, and
   falls back to the exact search of the whole database when the network's
   largest softmax probability is below `fallback`.  With centroid_scale
   g != 0 the logits also get g * <x, c_b> for every bucket b, with x and c_b
   the +/-1 expansions of the query and of the bucket's centroid (equal to
   g * (256 - 2 * Hamming distance)), so the untrained network (last layer
   zero) ranks buckets exactly as nearest-centroid probing does and training
   learns a correction to that ranking.  Results follow orb_match_top2's
   conventions (ties to the lowest database index). */
typedef struct {
    int nbuckets;          /* multiple of 64, <= 1024 */
    int hidden;            /* multiple of 64 */
    int nprobe;            /* buckets searched per query */
    float fallback;        /* exact search when max softmax < fallback */
    uint64_t seed;         /* weight initialisation, centroid choice, shuffling */
    float centroid_scale;  /* g above (rounded to bf16); 0: the plain MLP */
} orb_lindex_params;

typedef struct orb_lindex orb_lindex;

#define ORB_LINDEX_MLP 0        /* probe the network's top buckets (with fallback) */
#define ORB_LINDEX_CENTROIDS 1  /* probe the nearest centroids (no network, no fallback) */

int orb_lindex_create(orb_ctx* ctx, const orb_lindex_params* params, orb_lindex** out);
int orb_lindex_destroy(orb_lindex* lx);
/* Copy a database (ndb device descriptors) and partition it: `iters` k-modes
   steps from seeded initial centroids (or from loaded ones), then the final
   assignment and the bucket-sorted copy.  Synchronous. */
int orb_lindex_build(orb_lindex* lx, const uint8_t* d_db, int ndb, int iters);
/* Bucket offsets (nbuckets + 1 host ints), centroids (nbuckets x 32 host
   bytes) and the bucket of every database entry (ndb device ints); any NULL
   output is skipped. */
int orb_lindex_bucket_info(orb_lindex* lx, int* h_offsets, uint8_t* h_centroids, int* d_bucket_of);
/* Train on n queries whose exact nearest database entries are d_target
   (labels = their buckets): `epochs` shuffled passes of Adam steps with the
   given batch (multiple of 64); the step size decays from lr to lr/20 along a
   cosine over the whole run; h_epoch_loss (may be NULL)
   receives the mean loss of each epoch's last batch.  Synchronous. */
int orb_lindex_train(orb_lindex* lx, const uint8_t* d_q, const int* d_target, int n, int epochs, int batch,
                     float lr, float* h_epoch_loss);
/* Bucket logits of n queries (multiple of 64) into d_logits (n x nbuckets
   floats), rounded to FP16 (nearest even): the values the query path ranks
   its buckets by. */
int orb_lindex_logits(orb_lindex* lx, const uint8_t* d_q, int n, float* d_logits);
/* Gradients of the mean cross-entropy for n queries (multiple of 64) with
   labels d_labels, without updating: all parameters (W1 b1 W2 b2 W3 b3,
   row-major fp32) to h_grads, mean loss to *h_loss.  For checking. */
int orb_lindex_grads(orb_lindex* lx, const uint8_t* d_q, const int* d_labels, int n, float* h_grads, float* h_loss);
int orb_lindex_num_params(const orb_lindex* lx);
int orb_lindex_get_weights(orb_lindex* lx, float* h_w);
int orb_lindex_set_weights(orb_lindex* lx, const float* h_w);
/* The buckets each of nq queries (multiple of 64) probes, as orb_lindex_query
   would (d_probes: nq x nprobe ints, best first -- descending logit, or
   ascending centroid distance in ORB_LINDEX_CENTROIDS mode -- equal scores by
   ascending bucket id), and its max softmax probability (d_pmax: nq floats,
   1 in centroid mode).  Asynchronous on the context stream.  No counterpart in
   the reference (the learned matcher is an addition). */
int orb_lindex_probes(orb_lindex* lx, const uint8_t* d_q, int nq, int mode, int* d_probes, float* d_pmax);
/* Current number of probed buckets (set at creation, by orb_lindex_set_probe or by a load). */
int orb_lindex_nprobe(const orb_lindex* lx);
/* Change the number of probed buckets (1..nbuckets; at most 64 while a
   probability-mass rule is set) and the fallback threshold. */
int orb_lindex_set_probe(orb_lindex* lx, int nprobe, float fallback);
/* Probability-mass rule of the probes (mass in (0, 1]; 1, the default, probes
   all nprobe buckets): of the nprobe best buckets, ranked as
   orb_lindex_probes ranks them, a query probes the best-ranked prefix whose
   probability mass reaches `mass` -- a bucket is probed iff the weights of
   the buckets ranked before it sum to less than mass times the sum of all
   weights.  Weights: the softmax of the network's logits, or in
   ORB_LINDEX_CENTROIDS mode exp(-d / centroid_temperature) with d the
   Hamming distance to the centroid.  orb_lindex_probes then writes the
   probed buckets best first and -1 in the remaining slots.  Needs
   nprobe <= 64; a load whose nprobe is above 64 sets mass back to 1.  No
   counterpart in the reference (the learned matcher is an addition). */
int orb_lindex_set_mass(orb_lindex* lx, float mass, float centroid_temperature);
/* Nearest and second-nearest database entries of nq queries (multiple of 64);
   *h_nfallback (may be NULL) receives how many went to the exact search. */
int orb_lindex_query(orb_lindex* lx, const uint8_t* d_q, int nq, int mode, int* d_idx1, int* d_d1, int* d_idx2,
                     int* d_d2, int* h_nfallback);
/* Versioned file: "ORBLIDX1", version, shape, weights, centroids.  After a
   load, orb_lindex_build(lx, db, n, 0) assigns a database to the loaded
   centroids. */
int orb_lindex_save(orb_lindex* lx, const char* path);
int orb_lindex_load(orb_lindex* lx, const char* path);

/* Stage timing for measurement (no counterpart in the reference, whose
   GpuTimer wraps whole API calls, cuda_utils.h:81-108).  While enabled, each
   launch of a stage in stage_mask records a pair of HIP events on the
   context's stream; orb_profile_read synchronises, returns the summed
   milliseconds and launch counts per stage, and resets them. */
#define ORB_NUM_STAGES 8
#define ORB_STAGE_PYRAMID 0    /* raw + blurred levels */
#define ORB_STAGE_DETECT 1     /* FAST-9 + score into the key map */
#define ORB_STAGE_NMS 2        /* cross-level NMS + segment compaction */
#define ORB_STAGE_SELECT 3     /* per-frame scan / top-max_pts select */
#define ORB_STAGE_EMIT 4       /* keypoints in raster order */
#define ORB_STAGE_DESCRIBE 5   /* orientation + rBRIEF */
#define ORB_STAGE_MATCH 6      /* brute-force matcher main kernel */
#define ORB_STAGE_MATCH_MERGE 7
int orb_profile_enable(orb_ctx* ctx, unsigned stage_mask);
int orb_profile_read(orb_ctx* ctx, double* ms_sum, int* counts);

/* HIP graph capture (no counterpart in the reference, which synchronises
   after every launch, orbd.cu:1598-1791).  Between orb_capture_begin and
   orb_capture_end the context's calls are recorded into a graph instead of
   running; orb_graph_launch replays the recording on the context's stream.
   Each recorded call must have been made once before with the same frame
   sizes, parameters and buffer pointers (so no allocation, upload or tuning
   is needed), else it returns ORB_E_STATE.  Replays read and write the
   buffers the recorded calls named. */
typedef struct orb_graph orb_graph;
int orb_capture_begin(orb_ctx* ctx);
int orb_capture_end(orb_ctx* ctx, orb_graph** graph);
int orb_graph_launch(orb_graph* graph);
int orb_graph_destroy(orb_graph* graph);

/* Autotuner (no counterpart in the reference, whose launch shapes are fixed,
   orbd.cu:1564-1793).  Tile shapes of a stage are chosen per list of frame
   sizes on its first call: every variant is timed on the context's stream
   and the fastest kept; all variants give identical results.
   orb_tune_set forces a variant (-1 restores tuning); orb_tune_get reports
   the variant in use for the current frame set (-1 before its first call),
   the number of variants and, if ms is not null, the measured mean
   milliseconds per variant (0 where not measured). */
#define ORB_TUNE_PYRAMID 0     /* pyramid strip height */
int orb_tune_set(orb_ctx* ctx, int stage, int variant);
int orb_tune_get(orb_ctx* ctx, int stage, int* variant, int* nvariants, float* ms);

/* For tests, host only (no GPU): the 512 sampling points (x, y pairs) that
   orb_init uploads for a patch size (3..31) and wta_k 2, 3 or 4 (setPattern,
   orbd.cu:95-410). */
int orb_debug_pattern(int patch_size, int wta_k, int32_t* xy);

/* Device self-checks of the math helpers, for tests: n angles in d_in give
   sin, cos (own polynomials) and n (y, x) pairs give the centroid atan2. */
int orb_debug_sincos(orb_ctx* ctx, const float* d_in, int n, float* d_sin, float* d_cos);
int orb_debug_atan2(orb_ctx* ctx, const float* d_yx, int n, float* d_out);
/* For tests: copy frame `frame`'s flat pyramid of the last detect call
   (which 0: raw levels, 1: blurred levels; all levels with their row
   padding, *nbytes bytes) into d_out of cap bytes.  With d_out null only
   *nbytes is set. */
int orb_debug_pyramid(orb_ctx* ctx, int frame, int which, uint8_t* d_out, size_t cap, size_t* nbytes);

#ifdef __cplusplus
}
#endif
