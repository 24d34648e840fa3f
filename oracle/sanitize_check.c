/* Host sanitizer run of the oracle (test infrastructure): built with
   -fsanitize=address,undefined by `make -C oracle sanitize` and run by
   tests/test_sanitizers.py.  Exercises every stage at both pyramid scales,
   the three WTA_K modes, a patch size below 31, the cap, matching and the
   threaded loops; any out-of-bounds access, leak or undefined operation
   aborts with the sanitizer's report. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "orb_oracle.h"

/* a pitched image (row pitch p >= w, the padding zero), as the oracle takes it */
static void make_image(unsigned char* img, int w, int h, int p, unsigned seed)
{
    memset(img, 0, (size_t)p * h);
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            seed = seed * 1103515245u + 12345u;
            int blk = (((x / 12) ^ (y / 12)) & 1) ? 200 : 40;
            img[y * p + x] = (unsigned char)(blk + ((seed >> 16) & 15));
        }
}

int main(void)
{
    const int w = 300, h = 220, pitch = 384;   /* alignUp(w, 128) */
    unsigned char* a = malloc((size_t)pitch * h);
    unsigned char* b = malloc((size_t)pitch * h);
    make_image(a, w, h, pitch, 1);
    make_image(b, w, h, pitch, 2);
    int total = 0;
    for (int threads = 1; threads <= 4; threads += 3) {
        orc_set_threads(threads);
        for (int mode = 0; mode < 4; mode++) {
            orc_params p = {8, 31, 2, 0, 31, 20, 500, 1.2};
            if (mode == 1) { p.noctaves = 5; p.scale = 2.0; p.wta_k = 4; p.edge_threshold = 20; }
            if (mode == 2) { p.wta_k = 3; p.score_type = 1; p.max_pts = 40; }
            if (mode == 3) { p.patch_size = 13; p.wta_k = 4; }
            orc_point* pa = calloc(p.max_pts, sizeof(orc_point));
            orc_point* pb = calloc(p.max_pts, sizeof(orc_point));
            unsigned char* da = calloc((size_t)p.max_pts, 32);
            unsigned char* db = calloc((size_t)p.max_pts, 32);
            int na = 0, nb = 0, sa = 0, sb = 0;
            if (orc_detect_and_compute(&p, a, w, h, pa, da, &na, &sa) || orc_detect_and_compute(&p, b, w, h, pb, db, &nb, &sb)) {
                fprintf(stderr, "detect failed (mode %d)\n", mode);
                return 1;
            }
            int* m = calloc(na > 0 ? na : 1, sizeof(int));
            int* d = calloc(na > 0 ? na : 1, sizeof(int));
            orc_match(da, na, db, nb, m, d);
            int* o[4];
            for (int k = 0; k < 4; k++) o[k] = calloc(na > 0 ? na : 1, sizeof(int));
            orc_match_top2_norm(da, na, db, nb, threads, mode == 2, o[0], o[1], o[2], o[3]);
            total += na + nb;
            for (int k = 0; k < 4; k++) free(o[k]);
            free(m); free(d); free(pa); free(pb); free(da); free(db);
        }
    }
    int32_t xy[1024];
    for (int patch = 3; patch <= 31; patch++)
        for (int k = 2; k <= 4; k++)
            if (orc_pattern_patch(patch, k, xy)) return 1;
    free(a);
    free(b);
    printf("oracle sanitizer run: %d keypoints\n", total);
    return total > 0 ? 0 : 1;
}
