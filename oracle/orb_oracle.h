/*
 * CPU oracle for the ORB detect / describe / match path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a scalar C restatement of the reference's
 * algorithm (Accustomer/CUDA-ORB, orbd.cu / orb.cpp), used by tests/, by
 * __graft_entry__.smoke() and by the cpu_baseline leg of bench.py as the
 * checker.  The product library never links, loads or calls it.
 *
 * Every function cites the reference file:line it follows.  Where the
 * reference is nondeterministic or undefined (atomic append order, the
 * in-place blur race, __sinf/__cosf, compiler FMA contraction, reads outside
 * the allocation) the oracle fixes one semantics and the GPU path reproduces
 * it bit for bit; those choices are listed at each function.
 */
#ifndef ORB_ORACLE_H
#define ORB_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_LEVELS 8

typedef struct {
    int x, y, octave;
    float score, angle;
    int match, distance;
} orc_point; /* byte-identical to orb_structures.h:8-17 (28 B) */

typedef struct {
    int nlevels;
    int ref_mode;            /* 1: scale 2, nearest decimation (reference); 0: bilinear resize */
    double scale;
    int w[ORC_MAX_LEVELS], h[ORC_MAX_LEVELS], p[ORC_MAX_LEVELS];
    long off[ORC_MAX_LEVELS];
    long total;              /* bytes of the flat pyramid (sum h*p) */
    int radius[ORC_MAX_LEVELS];
    double sc[ORC_MAX_LEVELS]; /* scale of level l relative to level 0 */
} orc_geom;

typedef struct {
    int noctaves, edge_threshold, wta_k, score_type, patch_size, fast_threshold, max_pts;
    double scale;
} orc_params;

/* tables (orbd.cu:57-63, 66-92, 95-410, 413-433, 436-441) */
void  orc_fast_lut(int threshold, uint8_t lut[512]);
void  orc_umax(int patch_size, int* umax /* patch/2+2 */);
void  orc_gauss(float k[4]);
float orc_harris_scale(void);
int   orc_pattern(int wta_k, int32_t* xy /* 1024 ints */);
int   orc_pattern_patch(int patch_size, int wta_k, int32_t* xy /* 1024 ints */);
void  orc_glibc_srand(unsigned seed);
int   orc_glibc_rand(void);

/* threads of the oracle's parallel loops (default 1); results do not depend on it */
void orc_set_threads(int n);
int  orc_get_threads(void);

/* math helpers (orbd.cu:986-998; own sin/cos replaces __sinf/__cosf at orbd.cu:1446-1447) */
float orc_fast_atan2(float y, float x);
void  orc_sincosf(float a, float* s, float* c);

/* geometry (orb.cpp:145-191, plus the scale-s extension) */
int  orc_geometry(int w, int h, int noctaves, double scale, orc_geom* g);
int  orc_level_to_l0(const orc_geom* g, int level, int v);
int  orc_l0_to_level(const orc_geom* g, int level, int v);

/* stages */
void orc_pyramid(const orc_geom* g, const uint8_t* img, uint8_t* raw, uint8_t* blur);
void orc_blur_level(const orc_geom* g, int level, const uint8_t* raw, uint8_t* blur);
void orc_detect_map(const orc_geom* g, const uint8_t* raw, int threshold, int score_type,
                    float* vmap, int* lmap);
int  orc_fast_corner(const uint8_t* raw, long total, long center, int pitch, const uint8_t* lut,
                     int threshold);
float orc_harris(const uint8_t* raw, long total, long center, int pitch);
float orc_fast_score(const uint8_t* raw, long total, long center, int pitch, int threshold);
int  orc_nms(const orc_geom* g, const float* vmap, const int* lmap, int border, int max_pts,
             orc_point* pts, int* n_survivors);
void orc_angles(const orc_geom* g, const uint8_t* raw, orc_point* pts, int n, int patch_size);
void orc_describe_patch(const orc_geom* g, const uint8_t* blur, const orc_point* pts, int n, int patch_size,
                        int wta_k, uint8_t* desc);
void orc_describe(const orc_geom* g, const uint8_t* blur, const orc_point* pts, int n, int wta_k,
                  uint8_t* desc);

/* whole detect+compute for one image, input pitch = alignUp(w,128) (orb.cpp:57-101) */
int orc_detect_and_compute(const orc_params* prm, const uint8_t* img, int w, int h,
                           orc_point* pts, uint8_t* desc, int* n_out, int* n_survivors);

/* matchers (orbd.cu:1471-1559) */
void orc_match(const uint8_t* d1, int n1, const uint8_t* d2, int n2, int* match, int* dist);
void orc_match_top2(const uint8_t* q, int nq, const uint8_t* db, int ndb, int nthreads,
                    int* idx1, int* d1, int* idx2, int* d2);
/* norm 0: Hamming; norm 1: OpenCV's NORM_HAMMING2 (differing bit pairs) */
void orc_match_top2_norm(const uint8_t* q, int nq, const uint8_t* db, int ndb, int nthreads, int norm,
                         int* idx1, int* d1, int* idx2, int* d2);

#ifdef __cplusplus
}
#endif
#endif
