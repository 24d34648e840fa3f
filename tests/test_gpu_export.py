"""GPU parity of the OpenCV-layout keypoint export (orb_export_keypoints) and of
the NORM_HAMMING2 matcher (orb_match_top2 with ORB_MATCH_HAMMING2), through the
C ABI, against the CPU oracle on the same inputs.  Every field is compared bit
for bit: the export's float fields are single IEEE operations on both sides."""
import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs an MI355X")]


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


def detect(torch, ctx, seed, w, h, scale, retain, wta_k=2, max_pts=1000):
    import orb_amd
    img = orb_amd.synth_image(seed, w, h, 40)
    ctx.init(noctaves=8 if scale != 2.0 else 5, edge_threshold=31, wta_k=wta_k, score_type=0, patch_size=31,
             fast_threshold=20, retain_topn=retain, max_pts=max_pts, scale_factor=scale)
    dimg = torch.from_numpy(img).cuda()
    pts = ctx.alloc_points(max_pts)
    desc = ctx.alloc_desc(max_pts)
    cnt = torch.zeros(1, dtype=torch.int32, device="cuda")
    ctx.detect_and_compute(dimg, w, h, pts, desc, cnt)
    return pts, desc, cnt


@pytest.mark.parametrize("scale,retain", [(1.2, -1), (1.2, 300), (1.2, 1), (2.0, 150), (2.0, -1), (1.2, 100000)])
def test_export_matches_oracle(torch, ctx, oracle, scale, retain):
    import orb_amd
    print(f"export: 640x480 seed 0, scale {scale}, retain_topn {retain}", flush=True)
    pts, desc, cnt = detect(torch, ctx, 0, 640, 480, scale, retain)
    kp, dout, nout = ctx.export_keypoints(pts, desc, cnt)
    torch.cuda.synchronize()
    n, m = int(cnt.item()), int(nout.item())
    assert n > 200
    P = orb_amd.points_to_numpy(pts, n)
    want, wdesc = oracle.export_cv(P, desc[:n].cpu().numpy(), retain, scale)
    assert m == len(want) == (min(n, retain) if retain > 0 else n)
    got = orb_amd.as_cv_keypoints(kp[:m])
    for f in want.dtype.names:
        assert got[f].tobytes() == want[f].tobytes(), f
    assert np.array_equal(dout[:m].cpu().numpy(), wdesc)


def test_export_counts_and_points_only(torch, ctx, oracle):
    """No count pointer exports all cap rows; a zero count exports nothing; without
    descriptors only points are written."""
    import orb_amd
    pts, desc, cnt = detect(torch, ctx, 1, 640, 480, 1.2, 100)
    torch.cuda.synchronize()
    n = int(cnt.item())
    head = pts[:n].clone()
    kp, dout, nout = ctx.export_keypoints(head, None, None)
    torch.cuda.synchronize()
    assert dout is None and int(nout.item()) == min(n, 100)
    want, _ = oracle.export_cv(orb_amd.points_to_numpy(head, n), None, 100, 1.2)
    assert orb_amd.as_cv_keypoints(kp[:len(want)]).tobytes() == want.tobytes()
    zero = torch.zeros(1, dtype=torch.int32, device="cuda")
    _, _, nout = ctx.export_keypoints(pts, desc, zero)
    torch.cuda.synchronize()
    assert int(nout.item()) == 0


@pytest.mark.parametrize("nq,ndb", [(1, 1), (70, 3), (300, 1000), (513, 4097)])
def test_hamming2_top2_matches_oracle(torch, ctx, oracle, nq, ndb):
    import orb_amd
    print(f"hamming2: nq {nq} ndb {ndb}", flush=True)
    rng = np.random.default_rng(nq * 7919 + ndb)
    db = rng.integers(0, 256, (ndb, 32), dtype=np.uint8)
    q = rng.integers(0, 256, (nq, 32), dtype=np.uint8)
    if ndb > 1:
        db[ndb - 1] = db[0]          # tie: the lower index wins
    q[0] = db[0]
    if nq > 2:
        q[1] = db[0] ^ np.uint8(0b11)   # one differing pair, two differing bits
    out = ctx.match_top2(torch.from_numpy(q).cuda(), torch.from_numpy(db).cuda(), method=orb_amd.ORB_MATCH_HAMMING2)
    torch.cuda.synchronize()
    want = oracle.match_top2(q, db, norm=1)
    for g, w_ in zip(out, want):
        assert np.array_equal(g.cpu().numpy(), w_)


def test_hamming2_on_wta4_descriptors(torch, ctx, oracle):
    """WTA_K = 4 descriptors (2-bit argmax codes) of a frame pair, matched with their norm."""
    import orb_amd
    pa, da, ca = detect(torch, ctx, 2, 640, 480, 1.2, -1, wta_k=4)
    torch.cuda.synchronize()
    na = int(ca.item())
    qa = da[:na].clone()
    pb, db_, cb = detect(torch, ctx, 3, 640, 480, 1.2, -1, wta_k=4)
    torch.cuda.synchronize()
    nb = int(cb.item())
    out = ctx.match_top2(qa, db_[:nb].contiguous(), method=orb_amd.ORB_MATCH_HAMMING2)
    torch.cuda.synchronize()
    want = oracle.match_top2(qa.cpu().numpy(), db_[:nb].cpu().numpy(), norm=1)
    for g, w_ in zip(out, want):
        assert np.array_equal(g.cpu().numpy(), w_)
