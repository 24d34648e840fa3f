/*
 * Host sanitizer driver for the C-ABI library.  `make -C cuda-orb_amd
 * sanitize` builds it, with address and undefined-behaviour checks on every
 * host translation unit of the library (the .cpp files and the host side of
 * the .hip files), into cuda-orb_amd/_build/asan/capi_check.
 *
 * Without a GPU it runs the host-only entry points (synthetic images,
 * sampling patterns) and the error paths.  With one it also runs, on the
 * context's stream: a three-frame detection batch and the same frames one by
 * one, the reference-rule match, both top-2 matchers and the allocation-free
 * matcher, keypoint export, stage profiling, the autotuner queries, a
 * recorded graph replayed twice, and the learned index (build, train,
 * logits, gradients, probes, queries in both modes, save and load).  Results
 * that must agree are compared; the first failed check exits with status 1.
 * Usage: capi_check [scratch-dir]   (the learned index is saved there)
 */
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "orb_capi.h"

static int g_checks = 0;

#define EXPECT(cond)                                                                     \
    do {                                                                                 \
        ++g_checks;                                                                      \
        if (!(cond)) {                                                                   \
            fprintf(stderr, "capi_check: %s:%d: check failed: %s\n", __FILE__, __LINE__, #cond); \
            exit(1);                                                                     \
        }                                                                                \
    } while (0)

#define HIPOK(call) EXPECT((call) == hipSuccess)

static void host_only_checks()
{
    /* synthetic images: deterministic, pitch respected, bad arguments refused */
    const int w = 333, h = 250, p = 384;
    std::vector<uint8_t> a(p * h, 7), b(p * h, 7), c(p * h), warp(p * h);
    EXPECT(orb_synth_image(11, w, h, p, 30, a.data()) == ORB_OK);
    EXPECT(orb_synth_image(11, w, h, p, 30, b.data()) == ORB_OK);
    EXPECT(a == b);
    for (int y = 0; y < h; y++)
        for (int x = w; x < p; x++) EXPECT(a[y * p + x] == 0);   /* padding is cleared */
    double H[9];
    EXPECT(orb_synth_warp(11, a.data(), w, h, p, warp.data(), H) == ORB_OK);
    EXPECT(H[8] == 1.0 || (H[8] > 0.999 && H[8] < 1.001));
    EXPECT(orb_synth_image(0, 0, 10, 16, 1, c.data()) != ORB_OK);
    EXPECT(orb_synth_image(0, 20, 10, 16, 1, c.data()) != ORB_OK);
    EXPECT(orb_synth_image(0, 8, 1, 16, 1, nullptr) != ORB_OK);

    /* sampling patterns, every patch size and WTA_K: the 31 table lies in
       [-15, 15]; the random points of smaller patches follow the reference's
       rand() % patch_size - -patch_size / 2 (orbd.cu:366-367), which lies in
       [patch/2, patch/2 + patch - 1] */
    std::vector<int32_t> xy(1024);
    for (int patch = 3; patch <= 31; patch++)
        for (int k = 2; k <= 4; k++) {
            EXPECT(orb_debug_pattern(patch, k, xy.data()) == ORB_OK);
            const int lo = patch == 31 ? -15 : patch / 2, hi = patch == 31 ? 15 : patch / 2 + patch - 1;
            const int used = k == 2 ? 1024 : 2 * 128 * k;   /* WTA_K 3/4: 128 tuples of k points */
            for (int i = 0; i < used; i++) EXPECT(xy[i] >= lo && xy[i] <= hi);
        }
    EXPECT(orb_debug_pattern(32, 2, xy.data()) == ORB_E_ARG);
    EXPECT(orb_debug_pattern(2, 2, xy.data()) == ORB_E_ARG);
    EXPECT(orb_debug_pattern(31, 5, xy.data()) == ORB_E_ARG);
    EXPECT(orb_debug_pattern(31, 2, nullptr) == ORB_E_ARG);

    /* null handles report instead of crashing */
    EXPECT(orb_init(nullptr, nullptr) == ORB_E_ARG);
    EXPECT(orb_sync(nullptr) == ORB_E_ARG);
    EXPECT(orb_match(nullptr, nullptr, nullptr, 0, nullptr, 0) == ORB_E_ARG);
    EXPECT(orb_match_top2_noalloc(nullptr, nullptr, 0, nullptr, 0, 0, nullptr, nullptr, nullptr, nullptr) ==
           ORB_E_ARG);
    EXPECT(orb_workspace_bytes(nullptr) == 0);
    EXPECT(orb_last_error(nullptr) != nullptr);
    EXPECT(orb_ctx_create(nullptr, 0) == ORB_E_ARG);
    EXPECT(orb_lindex_create(nullptr, nullptr, nullptr) != ORB_OK);
}

struct Dev {
    void* p = nullptr;
    explicit Dev(size_t n) { HIPOK(hipMalloc(&p, n ? n : 1)); }
    Dev(const Dev&) = delete;
    ~Dev() { (void)hipFree(p); }
    template <typename T> T* as() const { return static_cast<T*>(p); }
};

template <typename T> static std::vector<T> down(const Dev& d, size_t n)
{
    std::vector<T> h(n);
    if (n) HIPOK(hipMemcpy(h.data(), d.p, n * sizeof(T), hipMemcpyDeviceToHost));
    return h;
}

static bool same_points(const std::vector<orb::OrbPoint>& a, const std::vector<orb::OrbPoint>& b, int n)
{
    return memcmp(a.data(), b.data(), size_t(n) * sizeof(orb::OrbPoint)) == 0;
}

static void gpu_checks(orb_ctx* ctx, const std::string& dir)
{
    orb_params prm = {5, 31, 2, 0, 31, 20, 0, 3000, 2.0};
    EXPECT(orb_init(ctx, &prm) == ORB_OK);
    orb_params bad = prm;
    bad.patch_size = 33;
    EXPECT(orb_init(ctx, &bad) == ORB_E_ARG);
    bad = prm;
    bad.wta_k = 5;
    EXPECT(orb_init(ctx, &bad) == ORB_E_ARG);
    EXPECT(orb_num_levels(ctx, 640, 480) > 0);

    /* three frames: 640x480, its warp, and an odd 333x250 */
    const int cap = prm.max_pts, nf = 3;
    const int ws[nf] = {640, 640, 333}, hs[nf] = {480, 480, 250}, ps[nf] = {640, 640, 384};
    std::vector<std::vector<uint8_t>> host(nf);
    for (int f = 0; f < nf; f++) host[f].assign(size_t(ps[f]) * hs[f], 0);
    EXPECT(orb_synth_image(5, ws[0], hs[0], ps[0], 40, host[0].data()) == ORB_OK);
    EXPECT(orb_synth_warp(5, host[0].data(), ws[0], hs[0], ps[0], host[1].data(), nullptr) == ORB_OK);
    EXPECT(orb_synth_image(6, ws[2], hs[2], ps[2], 25, host[2].data()) == ORB_OK);
    std::vector<Dev*> img, pts, desc;
    for (int f = 0; f < nf; f++) {
        img.push_back(new Dev(host[f].size()));
        HIPOK(hipMemcpy(img[f]->p, host[f].data(), host[f].size(), hipMemcpyHostToDevice));
        pts.push_back(new Dev(cap * sizeof(orb::OrbPoint)));
        desc.push_back(new Dev(cap * 32));
    }
    Dev counts(nf * sizeof(int)), one_pts(cap * sizeof(orb::OrbPoint)), one_desc(cap * 32), one_n(sizeof(int));
    const uint8_t* ip[nf];
    orb::OrbPoint* pp[nf];
    uint8_t* dp[nf];
    for (int f = 0; f < nf; f++) {
        ip[f] = img[f]->as<uint8_t>();
        pp[f] = pts[f]->as<orb::OrbPoint>();
        dp[f] = desc[f]->as<uint8_t>();
    }
    EXPECT(orb_profile_enable(ctx, 0xffu) == ORB_OK);
    EXPECT(orb_detect_and_compute_batch(ctx, nf, ip, ws, hs, ps, pp, dp, counts.as<int>(), 1) == ORB_OK);
    EXPECT(orb_sync(ctx) == ORB_OK);
    std::vector<int> n = down<int>(counts, nf);
    for (int f = 0; f < nf; f++) EXPECT(n[f] > 0 && n[f] <= cap);
    double ms[ORB_NUM_STAGES];
    int cnt[ORB_NUM_STAGES];
    EXPECT(orb_profile_read(ctx, ms, cnt) == ORB_OK);
    EXPECT(cnt[ORB_STAGE_PYRAMID] > 0 && cnt[ORB_STAGE_DESCRIBE] > 0);
    EXPECT(orb_profile_enable(ctx, 0) == ORB_OK);
    int variant = -2, nvar = 0;
    EXPECT(orb_tune_get(ctx, ORB_TUNE_PYRAMID, &variant, &nvar, nullptr) == ORB_OK);
    EXPECT(nvar > 0 && variant >= -1 && variant < nvar);

    /* one frame at a time gives the batch's results */
    std::vector<std::vector<orb::OrbPoint>> bp(nf);
    std::vector<std::vector<uint8_t>> bd(nf);
    for (int f = 0; f < nf; f++) {
        bp[f] = down<orb::OrbPoint>(*pts[f], n[f]);
        bd[f] = down<uint8_t>(*desc[f], size_t(n[f]) * 32);
        EXPECT(orb_detect_and_compute(ctx, ip[f], ws[f], hs[f], ps[f], one_pts.as<orb::OrbPoint>(),
                                      one_desc.as<uint8_t>(), one_n.as<int>(), 1) == ORB_OK);
        EXPECT(orb_sync(ctx) == ORB_OK);
        EXPECT(down<int>(one_n, 1)[0] == n[f]);
        EXPECT(same_points(down<orb::OrbPoint>(one_pts, n[f]), bp[f], n[f]));
        EXPECT(down<uint8_t>(one_desc, size_t(n[f]) * 32) == bd[f]);
    }
    EXPECT(orb_detect_and_compute(ctx, ip[0], ws[0], hs[0], 100, pp[0], dp[0], counts.as<int>(), 1) == ORB_E_ARG);

    /* reference-rule match of the pair, host counts and device counts */
    EXPECT(orb_match(ctx, pp[0], dp[0], n[0], dp[1], n[1]) == ORB_OK);
    EXPECT(orb_sync(ctx) == ORB_OK);
    std::vector<orb::OrbPoint> m1 = down<orb::OrbPoint>(*pts[0], n[0]);
    int good = 0;
    for (int i = 0; i < n[0]; i++) {
        EXPECT((m1[i].match == -1) == (m1[i].distance == -1));
        EXPECT(m1[i].match < n[1] && m1[i].distance < 64);
        good += m1[i].match >= 0;
    }
    EXPECT(good > 0);
    EXPECT(orb_match_dcount(ctx, pp[0], dp[0], counts.as<int>(), cap, dp[1], counts.as<int>() + 1, cap) == ORB_OK);
    EXPECT(orb_sync(ctx) == ORB_OK);
    EXPECT(same_points(down<orb::OrbPoint>(*pts[0], n[0]), m1, n[0]));

    /* top-2 matchers: popcount == MFMA == allocation-free (twice, the second smaller) */
    const int nq = n[0], ndb = n[1];
    Dev o[3][4] = {{Dev(nq * 4), Dev(nq * 4), Dev(nq * 4), Dev(nq * 4)},
                   {Dev(nq * 4), Dev(nq * 4), Dev(nq * 4), Dev(nq * 4)},
                   {Dev(nq * 4), Dev(nq * 4), Dev(nq * 4), Dev(nq * 4)}};
    EXPECT(orb_match_top2(ctx, dp[0], nq, dp[1], ndb, ORB_MATCH_POPCOUNT, o[0][0].as<int>(), o[0][1].as<int>(),
                          o[0][2].as<int>(), o[0][3].as<int>()) == ORB_OK);
    EXPECT(orb_match_top2(ctx, dp[0], nq, dp[1], ndb, ORB_MATCH_MFMA, o[1][0].as<int>(), o[1][1].as<int>(),
                          o[1][2].as<int>(), o[1][3].as<int>()) == ORB_OK);
    for (int method : {ORB_MATCH_MFMA, ORB_MATCH_POPCOUNT}) {
        for (int rep = 0; rep < 2; rep++) {
            const int q = rep ? nq / 3 : nq;
            EXPECT(orb_match_top2_noalloc(ctx, dp[0], q, dp[1], ndb, method, o[2][0].as<int>(), o[2][1].as<int>(),
                                          o[2][2].as<int>(), o[2][3].as<int>()) == ORB_OK);
            EXPECT(orb_sync(ctx) == ORB_OK);
            for (int k = 0; k < 4; k++) {
                std::vector<int> want = down<int>(o[0][k], nq);
                EXPECT(down<int>(o[1][k], nq) == want);
                std::vector<int> got = down<int>(o[2][k], q);
                EXPECT(std::equal(got.begin(), got.end(), want.begin()));
            }
        }
    }
    EXPECT(orb_match_top2_noalloc(ctx, dp[0], ORB_NOALLOC_MAX_QUERIES + 64, dp[1], ndb, ORB_MATCH_MFMA,
                                  o[2][0].as<int>(), nullptr, nullptr, nullptr) == ORB_E_CAPACITY);
    EXPECT(orb_match_top2(ctx, dp[0], nq, dp[1], ndb, 7, o[0][0].as<int>(), nullptr, nullptr, nullptr) == ORB_E_ARG);

    /* OpenCV-layout export of frame 2 */
    Dev kp(cap * sizeof(orb_cv_keypoint)), dout(cap * 32), nout(sizeof(int));
    EXPECT(orb_export_keypoints(ctx, pp[2], dp[2], counts.as<int>() + 2, cap, kp.as<orb_cv_keypoint>(),
                                dout.as<uint8_t>(), nout.as<int>()) == ORB_OK);
    EXPECT(orb_sync(ctx) == ORB_OK);
    EXPECT(down<int>(nout, 1)[0] == n[2]);
    EXPECT(down<uint8_t>(dout, size_t(n[2]) * 32) == bd[2]);
    std::vector<orb_cv_keypoint> k = down<orb_cv_keypoint>(kp, n[2]);
    for (int i = 0; i < n[2]; i++) EXPECT(k[i].angle >= 0.f && k[i].angle < 360.f && k[i].class_id == -1);

    /* a recorded detection replayed twice equals the eager one */
    orb_graph* g = nullptr;
    EXPECT(orb_capture_begin(ctx) == ORB_OK);
    EXPECT(orb_detect_and_compute(ctx, ip[2], ws[2], hs[2], ps[2], one_pts.as<orb::OrbPoint>(),
                                  one_desc.as<uint8_t>(), one_n.as<int>(), 1) == ORB_OK);
    EXPECT(orb_capture_end(ctx, &g) == ORB_OK && g);
    for (int rep = 0; rep < 2; rep++) {
        HIPOK(hipMemset(one_pts.p, 0, cap * sizeof(orb::OrbPoint)));
        HIPOK(hipDeviceSynchronize());
        EXPECT(orb_graph_launch(g) == ORB_OK);
        EXPECT(orb_sync(ctx) == ORB_OK);
        EXPECT(down<int>(one_n, 1)[0] == n[2]);
        EXPECT(same_points(down<orb::OrbPoint>(one_pts, n[2]), bp[2], n[2]));
    }
    EXPECT(orb_graph_destroy(g) == ORB_OK);

    for (int f = 0; f < nf; f++) {
        delete img[f];
        delete pts[f];
        delete desc[f];
    }

    /* learned index on 16384 random descriptors, 2048 noisy queries */
    const int nd = 16384, nqs = 2048;
    std::vector<uint8_t> hdb(size_t(nd) * 32), hq(size_t(nqs) * 32);
    uint64_t s = 88172645463325252ull;
    auto next = [&] { s ^= s << 13; s ^= s >> 7; s ^= s << 17; return s; };
    for (auto& v : hdb) v = uint8_t(next());
    for (int i = 0; i < nqs; i++)
        for (int j = 0; j < 32; j++) hq[i * 32 + j] = hdb[(i * 7 % nd) * 32 + j] ^ uint8_t(1u << (next() % 8) & next());
    Dev ddb(hdb.size()), dq(hq.size()), lab(nqs * 4), li[4] = {Dev(nqs * 4), Dev(nqs * 4), Dev(nqs * 4), Dev(nqs * 4)};
    HIPOK(hipMemcpy(ddb.p, hdb.data(), hdb.size(), hipMemcpyHostToDevice));
    HIPOK(hipMemcpy(dq.p, hq.data(), hq.size(), hipMemcpyHostToDevice));
    EXPECT(orb_match_top2(ctx, dq.as<uint8_t>(), nqs, ddb.as<uint8_t>(), nd, ORB_MATCH_MFMA, lab.as<int>(), nullptr,
                          nullptr, nullptr) == ORB_OK);
    orb_lindex_params lp = {64, 64, 4, 0.2f, 1, 0.125f};
    orb_lindex* lx = nullptr;
    EXPECT(orb_lindex_create(ctx, &lp, &lx) == ORB_OK && lx);
    EXPECT(orb_lindex_build(lx, ddb.as<uint8_t>(), nd, 3) == ORB_OK);
    std::vector<int> off(lp.nbuckets + 1);
    std::vector<uint8_t> cen(size_t(lp.nbuckets) * 32);
    Dev bof(nd * 4);
    EXPECT(orb_lindex_bucket_info(lx, off.data(), cen.data(), bof.as<int>()) == ORB_OK);
    EXPECT(off[0] == 0 && off[lp.nbuckets] == nd);
    float loss[2] = {0, 0};
    EXPECT(orb_lindex_train(lx, dq.as<uint8_t>(), lab.as<int>(), nqs, 2, 256, 1e-3f, loss) == ORB_OK);
    EXPECT(loss[0] > 0.f && loss[1] > 0.f);
    const int np = orb_lindex_num_params(lx);
    EXPECT(np == 256 * 64 + 64 + 64 * 64 + 64 + 64 * 64 + 64);
    std::vector<float> wts(np), grads(np);
    EXPECT(orb_lindex_get_weights(lx, wts.data()) == ORB_OK);
    float gl = 0;
    EXPECT(orb_lindex_grads(lx, dq.as<uint8_t>(), lab.as<int>(), 256, grads.data(), &gl) == ORB_OK && gl > 0.f);
    Dev logits(size_t(nqs) * lp.nbuckets * 4), probes(size_t(nqs) * lp.nprobe * 4), pmax(nqs * 4);
    EXPECT(orb_lindex_logits(lx, dq.as<uint8_t>(), nqs, logits.as<float>()) == ORB_OK);
    EXPECT(orb_lindex_probes(lx, dq.as<uint8_t>(), nqs, ORB_LINDEX_MLP, probes.as<int>(), pmax.as<float>()) ==
           ORB_OK);
    int nfb = -1;
    std::vector<std::vector<int>> first(2);
    for (int mode : {ORB_LINDEX_MLP, ORB_LINDEX_CENTROIDS}) {
        EXPECT(orb_lindex_query(lx, dq.as<uint8_t>(), nqs, mode, li[0].as<int>(), li[1].as<int>(), li[2].as<int>(),
                                li[3].as<int>(), &nfb) == ORB_OK);
        EXPECT(nfb >= 0 && nfb <= nqs);
        first[mode] = down<int>(li[0], nqs);
        for (int v : first[mode]) EXPECT(v >= -1 && v < nd);
    }
    EXPECT(orb_lindex_set_probe(lx, lp.nbuckets + 1, 0.2f) == ORB_E_ARG);
    EXPECT(orb_lindex_set_probe(lx, 8, 0.1f) == ORB_OK && orb_lindex_nprobe(lx) == 8);

    /* save, load into a fresh index, assign the same database: same answers */
    const std::string path = dir + "/capi_check.lidx";
    EXPECT(orb_lindex_set_probe(lx, lp.nprobe, lp.fallback) == ORB_OK);
    EXPECT(orb_lindex_save(lx, path.c_str()) == ORB_OK);
    orb_lindex* lx2 = nullptr;
    EXPECT(orb_lindex_create(ctx, &lp, &lx2) == ORB_OK);
    EXPECT(orb_lindex_load(lx2, path.c_str()) == ORB_OK);
    EXPECT(orb_lindex_build(lx2, ddb.as<uint8_t>(), nd, 0) == ORB_OK);
    for (int mode : {ORB_LINDEX_MLP, ORB_LINDEX_CENTROIDS}) {
        EXPECT(orb_lindex_query(lx2, dq.as<uint8_t>(), nqs, mode, li[0].as<int>(), li[1].as<int>(), li[2].as<int>(),
                                li[3].as<int>(), &nfb) == ORB_OK);
        EXPECT(down<int>(li[0], nqs) == first[mode]);
    }
    EXPECT(orb_lindex_load(lx2, (dir + "/does-not-exist.lidx").c_str()) != ORB_OK);
    std::remove(path.c_str());
    EXPECT(orb_lindex_destroy(lx2) == ORB_OK);
    EXPECT(orb_lindex_destroy(lx) == ORB_OK);
}

int main(int argc, char** argv)
{
    host_only_checks();
    orb_ctx* ctx = nullptr;
    const int rc = orb_ctx_create(&ctx, 0);
    if (rc != ORB_OK) {
        EXPECT(ctx == nullptr && rc == ORB_E_HIP);
        printf("capi_check: no GPU (%s); %d host checks passed\n", orb_last_error(nullptr), g_checks);
        return 0;
    }
    gpu_checks(ctx, argc > 1 ? argv[1] : ".");
    EXPECT(orb_ctx_destroy(ctx) == ORB_OK);
    printf("capi_check: %d checks passed (host and GPU)\n", g_checks);
    return 0;
}
