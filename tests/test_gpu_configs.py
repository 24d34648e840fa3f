"""The benchmark workloads themselves, at full size, through the C ABI:

- dense matching, 65,536 queries x 262,144 random descriptors (queries are
  database rows with Binomial(256, 0.05) bit flips): the INT8-MFMA and
  popcount matchers equal each other element for element, the distances obey
  the flip bound, and 1,024 queries equal the oracle against the whole
  database;
- the 64-pair mixed-size batch (128 frames drawn from 640x480 ... 3840x2160,
  <= 5000 keypoints each) in one detect/describe call and one batched match:
  every frame's keypoints and descriptors, and every pair's matches, equal the
  oracle;
- the learned matcher on a 1,048,576-descriptor database extracted by the
  pipeline from 2000 synthetic 1280x720 frames (1024 buckets, hidden 512,
  centroid term 0.125, top-8 probe, fallback below max softmax 0.3): a 1,024
  query subset equals the exact bucketed oracle (top-2), the fallback
  queries the exact top-2 of the whole database, in both probe modes.

The workload builders are bench.py's, so the tests check what the benchmark
measures."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs an MI355X")]

THREADS = 16


@pytest.fixture(scope="module")
def torch():
    import torch as t
    return t


@pytest.fixture(scope="module")
def ctx():
    import orb_amd
    c = orb_amd.Context(0)
    yield c
    c.close()


def test_dense_match_full_size(torch, ctx, oracle):
    import orb_amd
    print("dense 65536 x 262144, popcount and MFMA", flush=True)
    rng = np.random.default_rng(2)
    ndb, nq = 262144, 65536
    db = rng.integers(0, 256, (ndb, 32), dtype=np.uint8)
    src = rng.integers(0, ndb, nq)
    flips = rng.random((nq, 256)) < 0.05
    q = db[src] ^ np.packbits(flips, axis=1, bitorder="little")
    dq, ddb = torch.from_numpy(q).cuda(), torch.from_numpy(db).cuda()
    pop = ctx.match_top2(dq, ddb, method=orb_amd.ORB_MATCH_POPCOUNT)
    mf = ctx.match_top2(dq, ddb, method=orb_amd.ORB_MATCH_MFMA)
    torch.cuda.synchronize()
    for name, a, b in zip(("idx1", "d1", "idx2", "d2"), pop, mf):
        assert torch.equal(a, b), name
    i1, d1, i2, d2 = (o.cpu().numpy() for o in mf)
    nflip = flips.sum(1)
    assert (d1 <= nflip).all() and (d1 <= d2).all() and (i1 != i2).all()
    assert ((i1 == src) | (d1 < nflip) | ((d1 == nflip) & (i1 < src))).all()
    sub = np.r_[0:512, nq - 512:nq]
    want = oracle.match_top2(q[sub], db, nthreads=THREADS)
    for g, w in zip((i1[sub], d1[sub], i2[sub], d2[sub]), want):
        assert np.array_equal(g, w)


def test_batch64_mixed_pairs_full_size(torch, ctx, oracle):
    import bench
    import orb_amd
    rng = np.random.default_rng(3)
    sizes = [(640, 480), (1280, 720), (1920, 1080), (3840, 2160)]
    pick = [sizes[i] for i in rng.integers(0, 4, 64)]
    assert (3840, 2160) in pick

    def make(i):
        w, h = pick[i]
        a = orb_amd.synth_image(300 + i, w, h, max(1, int(40 * w * h / 307200)))
        return a, orb_amd.synth_warp(300 + i, a, w, h)[0]

    with ThreadPoolExecutor(THREADS) as ex:
        imgs = list(ex.map(make, range(64)))
    host = [a for a, _ in imgs] + [b for _, b in imgs]
    ws = [w for w, _ in pick] * 2
    hs = [h for _, h in pick] * 2
    cap = 5000
    print(f"batch of 128 frames, {sum(w * h for w, h in zip(ws, hs)) / 1e6:.0f} Mpixel, cap {cap}", flush=True)
    ctx.init(noctaves=bench.LEVELS, edge_threshold=31, wta_k=2, score_type=0, patch_size=31, fast_threshold=20,
             max_pts=cap, scale_factor=bench.SCALE)
    frames = [torch.from_numpy(a).cuda() for a in host]
    pts = ctx.alloc_points(cap, frames=128)
    desc = ctx.alloc_desc(cap, frames=128)
    counts = torch.zeros(128, dtype=torch.int32, device="cuda")
    pl, dl = [pts[i] for i in range(128)], [desc[i] for i in range(128)]
    ctx.detect_and_compute_batch(frames, ws, hs, pl, dl, counts)
    ctx.match_batch(pl[:64], dl[:64], counts[:64], cap, dl[64:], counts[64:], cap)
    torch.cuda.synchronize()
    n = counts.cpu().numpy()
    got_pts = [orb_amd.points_to_numpy(pts[i], int(n[i])) for i in range(128)]
    got_desc = [desc[i, :int(n[i])].cpu().numpy() for i in range(128)]

    def ref(i):
        return oracle.detect_and_compute(host[i][:, :ws[i]].copy(), noctaves=bench.LEVELS, scale=bench.SCALE,
                                         max_pts=cap)

    print("oracle on all 128 frames", flush=True)
    with ThreadPoolExecutor(THREADS) as ex:
        want = list(ex.map(ref, range(128)))
    capped = 0
    for i in range(128):
        wp, wd, nsurv = want[i]
        capped += nsurv > cap
        assert int(n[i]) == len(wp) > 0, (i, ws[i], hs[i])
        for f in ("x", "y", "octave"):
            assert np.array_equal(got_pts[i][f], wp[f]), (i, f)
        for f in ("score", "angle"):
            assert np.array_equal(got_pts[i][f].view(np.uint32), wp[f].view(np.uint32)), (i, f)
        assert np.array_equal(got_desc[i], wd), i
    print(f"{capped} frames above the cap; keypoints per frame {n.min()}..{n.max()}", flush=True)

    print("oracle matches of the 64 pairs", flush=True)
    with ThreadPoolExecutor(THREADS) as ex:
        wm = list(ex.map(lambda i: oracle.match(want[i][1], want[64 + i][1]), range(64)))
    for i in range(64):
        assert np.array_equal(got_pts[i]["match"], wm[i][0]), i
        assert np.array_equal(got_pts[i]["distance"], wm[i][1]), i
    assert sum(int((m >= 0).sum()) for m, _ in wm) > 10000


def test_learned_matcher_full_size(torch, ctx, oracle):
    import bench
    import orb_amd
    from oracle import mlp
    w, h, nshapes, per = 1280, 720, 120, 525
    ndb, nq = 1 << 20, 65536
    ctx.init(noctaves=bench.LEVELS, edge_threshold=31, wta_k=2, score_type=0, patch_size=31, fast_threshold=20,
             max_pts=per, scale_factor=bench.SCALE)
    print("extracting the 1,048,576-descriptor database and 65,536 queries", flush=True)
    seeds = list(range(10000, 12000))
    db = bench.extract(ctx, torch, "cuda:0", seeds, w, h, nshapes, per, limit=ndb)
    more = 12000
    while db.shape[0] < ndb:
        extra = bench.extract(ctx, torch, "cuda:0", list(range(more, more + 64)), w, h, nshapes, per)
        db = torch.cat([db, extra])[:ndb]
        more += 64
    q = bench.extract(ctx, torch, "cuda:0", seeds, w, h, nshapes, per, warp_seed=2, limit=nq)
    assert db.shape[0] == ndb and q.shape[0] == nq
    ctx.init(noctaves=bench.LEVELS, max_pts=bench.CAP, scale_factor=bench.SCALE)
    print("building 1024 buckets, one training epoch", flush=True)
    lx = orb_amd.LearnedIndex(ctx, nbuckets=1024, hidden=512, nprobe=8, fallback=0.3, seed=4, centroid_scale=0.125)
    try:
        lx.build(db, iters=8)
        labels = ctx.match_top2(q, db, method=orb_amd.ORB_MATCH_MFMA)[0]
        lx.train(q, labels, epochs=1, batch=4096, lr=1e-3)
        off, cen, bof = lx.bucket_info()
        assert off[-1] == ndb
        db_h, q_h, bof = db.cpu().numpy(), q.cpu().numpy(), bof.cpu().numpy()
        sub = np.r_[0:512, nq - 512:nq]
        dsub = torch.from_numpy(q_h[sub]).cuda()

        print("network probes with fallback", flush=True)
        lx.set_probe(8, 0.3)
        got, nfb = lx.query(q, mode=orb_amd.ORB_LINDEX_MLP)
        assert 0 < nfb < nq
        logits = lx.logits(dsub).cpu().numpy().astype(np.float64)
        probes, _ = mlp.probe(logits, 8)
        pm = lx.probes(dsub)[1].cpu().numpy()
        fb = pm < np.float32(0.3)
        probed = mlp.bucket_search(q_h[sub], db_h, bof, probes)
        exact = oracle.match_top2(q_h[sub], db_h, nthreads=THREADS)
        want = [np.where(fb, e, p) for e, p in zip(exact, probed)]
        assert 0 < fb.sum() < len(sub)
        for name, g, wv in zip(("idx1", "d1", "idx2", "d2"), got, want):
            assert np.array_equal(g.cpu().numpy()[sub], wv), name

        print("centroid probes", flush=True)
        got, nfb = lx.query(q, mode=orb_amd.ORB_LINDEX_CENTROIDS)
        assert nfb == 0
        D = np.bitwise_count(q_h[sub].view(np.uint64)[:, None, :] ^ cen.view(np.uint64)[None, :, :]).sum(-1)
        cprobes = np.argsort(D, axis=1, kind="stable")[:, :8]
        want = mlp.bucket_search(q_h[sub], db_h, bof, cprobes)
        for name, g, wv in zip(("idx1", "d1", "idx2", "d2"), got, want):
            assert np.array_equal(g.cpu().numpy()[sub], wv), name
    finally:
        lx.close()
