"""GPU parity: the HIP library (through its C ABI) against the CPU oracle on
the same seeded inputs.  Integer, index and byte outputs must be identical;
the float outputs (Harris score, angle) are defined by the same IEEE
operation sequence on both sides and must be bit-identical too."""
import math

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


def gpu_detect(torch, ctx, img_p, w, h, max_pts, compute_desc=True):
    import orb_amd
    dimg = torch.from_numpy(np.ascontiguousarray(img_p)).cuda()
    pts = ctx.alloc_points(max_pts)
    desc = ctx.alloc_desc(max_pts)
    cnt = torch.zeros(1, dtype=torch.int32, device="cuda")
    ctx.detect_and_compute(dimg, w, h, pts, desc, cnt, compute_desc)
    torch.cuda.synchronize()
    n = int(cnt.item())
    return orb_amd.points_to_numpy(pts, n), desc[:n].cpu().numpy()


def describe_mismatch(got, want, field):
    bad = np.nonzero(got[field] != want[field])[0]
    rows = [(int(i), int(got["x"][i]), int(got["y"][i]), int(got["octave"][i]), int(want["octave"][i]),
             hex(int(got["score"].view(np.uint32)[i])), hex(int(want["score"].view(np.uint32)[i])))
            for i in bad[:8]]
    return f"{field}: {len(bad)} of {len(got)} differ; (i, x, y, oct, oct_want, score, score_want): {rows}"


def assert_same_points(got, want):
    assert len(got) == len(want)
    for f in ("x", "y", "octave"):
        assert np.array_equal(got[f], want[f]), describe_mismatch(got, want, f)
    assert np.array_equal(got["score"].view(np.uint32), want["score"].view(np.uint32)), "score bits"
    assert np.array_equal(got["angle"].view(np.uint32), want["angle"].view(np.uint32)), "angle bits"


CASES = [
    # w, h, nshapes, seed, noctaves, scale, wta_k, score_type, max_pts
    (640, 480, 40, 0, 8, 1.2, 2, 0, 1000),
    (640, 480, 40, 0, 5, 2.0, 2, 0, 1000),
    (1920, 1080, 80, 1, 8, 1.2, 2, 0, 10000),
    (1920, 1080, 80, 1, 5, 2.0, 4, 0, 10000),
    (333, 257, 8, 7, 8, 1.2, 3, 0, 500),
    (640, 480, 40, 2, 8, 1.2, 4, 1, 1000),
    (640, 480, 40, 3, 8, 1.2, 2, 0, 150),      # cap well below the survivors: exact selection
    (1280, 720, 120, 3, 8, 1.25, 2, 0, 5000),
    (3840, 2160, 1080, 5, 8, 1.2, 2, 0, 2000),  # the largest batched frame size, cap below the survivors
    (3840, 2160, 1080, 5, 5, 2.0, 2, 0, 5000),
]


@pytest.mark.parametrize("w,h,nshapes,seed,nl,scale,wta,st,cap", CASES)
def test_detect_and_compute_matches_oracle(torch, ctx, oracle, w, h, nshapes, seed, nl, scale, wta, st, cap):
    import orb_amd
    print(f"case {w}x{h} seed={seed} levels={nl} scale={scale} wta={wta} score={st} cap={cap}", flush=True)
    img = orb_amd.synth_image(seed, w, h, nshapes)
    ctx.init(noctaves=nl, wta_k=wta, score_type=st, max_pts=cap, scale_factor=scale)
    got_p, got_d = gpu_detect(torch, ctx, img, w, h, cap)
    want_p, want_d, nsurv = oracle.detect_and_compute(img[:, :w].copy(), noctaves=nl, scale=scale, wta_k=wta,
                                                      score_type=st, max_pts=cap)
    assert len(want_p) > 0
    assert_same_points(got_p, want_p)
    assert np.array_equal(got_d, want_d)
    # the same call again (next epoch of the key map) gives the same bytes
    got_p2, got_d2 = gpu_detect(torch, ctx, img, w, h, cap)
    assert_same_points(got_p2, want_p)
    assert np.array_equal(got_d2, want_d)


@pytest.mark.parametrize("patch,wta,scale,nl", [(25, 2, 1.2, 8), (25, 4, 1.2, 8), (9, 3, 2.0, 5), (30, 2, 1.2, 8)])
def test_patch_sizes_below_31_match_oracle(torch, ctx, oracle, patch, wta, scale, nl):
    """Patch sizes other than 31: the umax disc of the patch for the angle and the reference's
    random sampling points (off centre, orbd.cu:359-369), read from the flat blurred pyramid with
    bytes outside it as 0; keypoints, angles and descriptors equal the oracle."""
    import orb_amd
    print(f"patch {patch} wta {wta} scale {scale}", flush=True)
    w, h = 640, 480
    img = orb_amd.synth_image(9, w, h, 40)
    ctx.init(noctaves=nl, wta_k=wta, patch_size=patch, max_pts=1000, scale_factor=scale)
    got_p, got_d = gpu_detect(torch, ctx, img, w, h, 1000)
    want_p, want_d, _ = oracle.detect_and_compute(img, noctaves=nl, scale=scale, wta_k=wta, patch_size=patch,
                                                  max_pts=1000)
    assert len(want_p) > 100
    assert_same_points(got_p, want_p)
    assert np.array_equal(got_d, want_d)
    with pytest.raises(orb_amd.OrbError):
        ctx.init(noctaves=nl, patch_size=33, max_pts=1000, scale_factor=scale)


@pytest.mark.parametrize("w,h", [(16, 16), (40, 40), (64, 48), (100, 100), (17, 300), (300, 17)])
def test_tiny_frames_match_oracle(torch, ctx, oracle, w, h):
    """Frames at the 16-pixel level floor and inside the 31-pixel detection border: the same
    (often empty) keypoint sets as the oracle, and no failure."""
    import orb_amd
    print(f"tiny {w}x{h}", flush=True)
    img = orb_amd.synth_image(6, w, h, 3)
    ctx.init(noctaves=8, max_pts=100, scale_factor=1.2)
    got_p, got_d = gpu_detect(torch, ctx, img, w, h, 100)
    want_p, want_d, _ = oracle.detect_and_compute(img[:, :w].copy(), noctaves=8, scale=1.2, max_pts=100)
    assert_same_points(got_p, want_p)
    assert np.array_equal(got_d, want_d)


def test_frame_too_small_for_reference_octaves_is_an_error(torch, ctx):
    """Scale 2 needs min(w, h) >= 80 for its first octave (orb.cpp:148); smaller frames are
    rejected with an error, and the context keeps working."""
    import orb_amd
    img = orb_amd.synth_image(6, 64, 48, 3)
    ctx.init(noctaves=5, max_pts=100, scale_factor=2.0)
    with pytest.raises(orb_amd.OrbError):
        gpu_detect(torch, ctx, img, 64, 48, 100)
    img = orb_amd.synth_image(6, 160, 120, 10)
    gpu_detect(torch, ctx, img, 160, 120, 100)


def test_input_padding_bytes_follow_the_flat_layout(torch, ctx, oracle):
    """Level-0 row padding is copied from the input (orbd.cu:1584), and reads that cross level
    boundaries see it; fill the padding with noise and compare."""
    import orb_amd
    w, h = 600, 400
    img = orb_amd.synth_image(5, w, h, 30)
    img[:, w:] = np.random.default_rng(1).integers(0, 256, (h, img.shape[1] - w), dtype=np.uint8)
    ctx.init(noctaves=5, max_pts=2000, scale_factor=2.0)
    got_p, got_d = gpu_detect(torch, ctx, img, w, h, 2000)
    # the oracle takes the same pitched bytes
    import ctypes
    prm = oracle.Params(5, 31, 2, 0, 31, 20, 2000, 2.0)
    pts = np.zeros(2000, oracle.POINT_DTYPE)
    desc = np.zeros((2000, 32), np.uint8)
    n = ctypes.c_int()
    ns = ctypes.c_int()
    oracle.lib().orc_detect_and_compute(ctypes.byref(prm), oracle.ptr(img), w, h, oracle.ptr(pts),
                                        oracle.ptr(desc), ctypes.byref(n), ctypes.byref(ns))
    assert_same_points(got_p, pts[: n.value])
    assert np.array_equal(got_d, desc[: n.value])


PYR_CASES = [
    # w, h, noctaves, scale: bilinear levels (scale 1.5 reaches level scales whose taps need one
    # load per column), scale-2 decimation, odd sizes and a pitch equal to the width
    (640, 480, 8, 1.2),
    (1920, 1080, 8, 1.2),
    (1920, 1080, 8, 1.5),
    (333, 257, 8, 1.2),
    (1280, 720, 8, 1.25),
    (600, 400, 5, 2.0),
    (1920, 1080, 5, 2.0),
    (1024, 200, 8, 1.2),
]


@pytest.mark.parametrize("w,h,nl,scale", PYR_CASES)
def test_pyramid_bytes_every_strip_variant(torch, ctx, oracle, w, h, nl, scale):
    """The raw and blurred flat pyramids, row padding included, are byte-identical to the oracle's
    for every strip height the autotuner may pick; the input's own padding bytes are noise."""
    import orb_amd
    print(f"pyramid {w}x{h} levels={nl} scale={scale}", flush=True)
    img = orb_amd.synth_image(11, w, h, 20)
    img[:, w:] = np.random.default_rng(3).integers(0, 256, (h, img.shape[1] - w), dtype=np.uint8)
    g = oracle.geometry(w, h, nl, scale)
    want_raw, want_blur = oracle.pyramid(g, img)
    ctx.init(noctaves=nl, max_pts=500, scale_factor=scale)
    try:
        for v in range(len(orb_amd.PYRAMID_STRIP_HEIGHTS)):
            ctx.tune_set(orb_amd.TUNE_PYRAMID, v)
            gpu_detect(torch, ctx, img, w, h, 500)
            raw = ctx.debug_pyramid(0, 0).cpu().numpy()
            blur = ctx.debug_pyramid(0, 1).cpu().numpy()
            for name, got, want in (("raw", raw, want_raw), ("blur", blur, want_blur)):
                bad = np.nonzero(got != want)[0]
                where = [(int(np.searchsorted(np.array(g.off[:g.nlevels]), i, "right") - 1), int(i)) for i in bad[:5]]
                assert len(bad) == 0, f"variant {v} {name}: {len(bad)} bytes differ, first (level, offset) {where}"
    finally:
        ctx.tune_set(orb_amd.TUNE_PYRAMID, -1)
    gpu_detect(torch, ctx, img, w, h, 500)
    chosen, ms = ctx.tune_get(orb_amd.TUNE_PYRAMID)
    assert 0 <= chosen < len(ms) and all(t > 0 for t in ms)


def test_constant_image_gives_no_points(torch, ctx):
    ctx.init(noctaves=8, max_pts=100, scale_factor=1.2)
    img = np.full((480, 640), 77, np.uint8)
    p, d = gpu_detect(torch, ctx, img, 640, 480, 100)
    assert len(p) == 0


def test_batch_matches_per_frame_oracle(torch, ctx, oracle):
    import orb_amd
    sizes = [(640, 480), (1280, 720), (333, 257), (1920, 1080), (640, 480)]
    cap = 3000
    ctx.init(noctaves=8, max_pts=cap, scale_factor=1.2)
    imgs, hosts = [], []
    for i, (w, h) in enumerate(sizes):
        a = orb_amd.synth_image(20 + i, w, h, max(8, w * h * 40 // 307200))
        hosts.append(a)
        imgs.append(torch.from_numpy(a).cuda())
    pts = ctx.alloc_points(cap, frames=len(sizes))
    desc = ctx.alloc_desc(cap, frames=len(sizes))
    cnt = torch.zeros(len(sizes), dtype=torch.int32, device="cuda")
    for rep in range(2):
        ctx.detect_and_compute_batch(imgs, [s[0] for s in sizes], [s[1] for s in sizes],
                                     [pts[i] for i in range(len(sizes))], [desc[i] for i in range(len(sizes))], cnt)
        torch.cuda.synchronize()
        for i, (w, h) in enumerate(sizes):
            print(f"batch frame {i} {w}x{h} rep {rep}", flush=True)
            n = int(cnt[i].item())
            want_p, want_d, _ = oracle.detect_and_compute(hosts[i][:, :w].copy(), noctaves=8, scale=1.2,
                                                          max_pts=cap)
            assert_same_points(orb_amd.points_to_numpy(pts[i], n), want_p)
            assert np.array_equal(desc[i][:n].cpu().numpy(), want_d)


def test_graph_replay_matches_eager_step(torch, ctx, oracle):
    """A recorded detect(2 frames) + match step replays to the eager results, several times, with
    eager work of other sizes and host synchronisations between the replays; recording a call that
    was not made before fails cleanly."""
    import orb_amd
    cap = 2000
    ctx.init(noctaves=8, max_pts=cap, scale_factor=1.2)
    a = orb_amd.synth_image(31, 640, 480, 40)
    b, _ = orb_amd.synth_warp(31, a, 640, 480)
    imgs = [torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()]
    pts = ctx.alloc_points(cap, frames=2)
    desc = ctx.alloc_desc(cap, frames=2)
    cnt = torch.zeros(2, dtype=torch.int32, device="cuda")

    def step():
        ctx.detect_and_compute_batch(imgs, [640, 640], [480, 480], [pts[0], pts[1]], [desc[0], desc[1]], cnt)
        ctx.match_dcount(pts[0], desc[0], cnt[0:1], cap, desc[1], cnt[1:2], cap)

    step()
    torch.cuda.synchronize()
    n0, n1 = (int(v) for v in cnt.cpu())
    want = [pts[0][:n0].clone(), pts[1][:n1].clone(), desc[0][:n0].clone(), desc[1][:n1].clone()]
    wp, wd, _ = oracle.detect_and_compute(a[:, :640].copy(), noctaves=8, scale=1.2, max_pts=cap)
    assert_same_points(orb_amd.points_to_numpy(pts[0], n0), wp)
    other = torch.from_numpy(orb_amd.synth_image(32, 333, 257, 8)).cuda()
    opts, odesc = ctx.alloc_points(cap), ctx.alloc_desc(cap)
    ocnt = torch.zeros(1, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()

    ctx.capture_begin()
    step()
    graph = ctx.capture_end()
    for rep in range(4):
        print(f"graph replay {rep}", flush=True)
        for t in (pts, desc, cnt):
            t.zero_()
        torch.cuda.synchronize()
        graph.launch()
        torch.cuda.synchronize()
        assert [int(v) for v in cnt.cpu()] == [n0, n1]
        got = [pts[0][:n0], pts[1][:n1], desc[0][:n0], desc[1][:n1]]
        for g, w_ in zip(got, want):
            assert torch.equal(g, w_)
        # eager work of another size on the same context between replays
        ctx.detect_and_compute(other, 333, 257, opts, odesc, ocnt)
        torch.cuda.synchronize()
    graph.close()
    # a call that would need an allocation cannot be recorded
    new_size = torch.from_numpy(orb_amd.synth_image(33, 320, 240, 8)).cuda()
    torch.cuda.synchronize()
    ctx.capture_begin()
    with pytest.raises(orb_amd.OrbError):
        ctx.detect_and_compute(new_size, 320, 240, opts, odesc, ocnt)
    try:
        ctx.capture_end().close()   # an empty recording; instantiating it may be refused
    except orb_amd.OrbError:
        pass
    # the context works normally afterwards
    ctx.detect_and_compute(new_size, 320, 240, opts, odesc, ocnt)
    torch.cuda.synchronize()


def test_match_reference_rule_and_device_counts(torch, ctx, oracle):
    import orb_amd
    w, h = 640, 480
    a = orb_amd.synth_image(0, w, h, 40)
    b, _ = orb_amd.synth_warp(0, a, w, h)
    ctx.init(noctaves=8, max_pts=1000, scale_factor=1.2)
    pa, da = gpu_detect(torch, ctx, a, w, h, 1000)
    pb, db = gpu_detect(torch, ctx, b, w, h, 1000)
    want_m, want_dist = oracle.match(da, db)
    pts1 = torch.from_numpy(pa.view(np.int32).reshape(-1, 7).copy()).cuda()
    ctx.match(pts1, torch.from_numpy(da).cuda(), len(pa), torch.from_numpy(db).cuda(), len(pb))
    got = orb_amd.points_to_numpy(pts1, len(pa))
    assert np.array_equal(got["match"], want_m) and np.array_equal(got["distance"], want_dist)
    assert (want_m >= 0).sum() > 100
    # device-count variant, counts larger buffers than used
    cap = 1500
    pts1b = torch.zeros((cap, 7), dtype=torch.int32, device="cuda")
    pts1b[: len(pa)] = pts1
    d1 = torch.zeros((cap, 32), dtype=torch.uint8, device="cuda"); d1[: len(pa)] = torch.from_numpy(da).cuda()
    d2 = torch.zeros((cap, 32), dtype=torch.uint8, device="cuda"); d2[: len(pb)] = torch.from_numpy(db).cuda()
    n1 = torch.tensor([len(pa)], dtype=torch.int32, device="cuda")
    n2 = torch.tensor([len(pb)], dtype=torch.int32, device="cuda")
    ctx.match_dcount(pts1b, d1, n1, cap, d2, n2, cap)
    got = orb_amd.points_to_numpy(pts1b, len(pa))
    assert np.array_equal(got["match"], want_m) and np.array_equal(got["distance"], want_dist)


def test_match_batch_matches_oracle_per_pair(torch, ctx, oracle):
    """Several pairs in one orb_match_batch launch, counts read on the device: each pair's
    match/distance equal the reference rule of the oracle; includes an empty database, an
    empty query set and a one-entry database; a second call with other counts reuses the
    merge state cleanly."""
    import orb_amd
    rng = np.random.default_rng(21)
    cap = 1200
    sizes = [(640, 700), (1100, 1200), (0, 300), (50, 0), (1, 1), (300, 1)]
    q = [rng.integers(0, 256, (cap, 32), dtype=np.uint8) for _ in sizes]
    db = [rng.integers(0, 256, (cap, 32), dtype=np.uint8) for _ in sizes]
    for i, (n1, n2) in enumerate(sizes):   # plant near-duplicates so some matches are accepted
        k = min(n1, n2)
        db[i][:k // 2] = q[i][:k // 2] ^ (rng.random((k // 2, 32)) < 0.02).astype(np.uint8)
    dq = [torch.from_numpy(x).cuda() for x in q]
    ddb = [torch.from_numpy(x).cuda() for x in db]
    pts = [torch.full((cap, 7), 7, dtype=torch.int32, device="cuda") for _ in sizes]
    for rnd in range(2):
        szs = sizes if rnd == 0 else [(n2 // 2, n1) for (n1, n2) in sizes]
        n1 = torch.tensor([a for a, _ in szs], dtype=torch.int32, device="cuda")
        n2 = torch.tensor([b for _, b in szs], dtype=torch.int32, device="cuda")
        ctx.match_batch(pts, dq, n1, cap, ddb, n2, cap)
        torch.cuda.synchronize()
        for i, (a, b) in enumerate(szs):
            got = orb_amd.points_to_numpy(pts[i], a)
            want_m, want_d = oracle.match(q[i][:a], db[i][:b])
            assert np.array_equal(got["match"], want_m), (rnd, i)
            assert np.array_equal(got["distance"], want_d), (rnd, i)


@pytest.mark.parametrize("method", [0, 1], ids=["popcount", "mfma"])
@pytest.mark.parametrize("nq,ndb", [(1, 1), (7, 300), (1000, 4097), (300, 70000), (5000, 0), (4100, 9000),
                                    (257, 63), (65, 129)])
def test_match_top2_matches_oracle(torch, ctx, oracle, nq, ndb, method):
    print(f"top2 method={method} nq={nq} ndb={ndb}", flush=True)
    rng = np.random.default_rng(nq * 7 + ndb)
    db = rng.integers(0, 256, (ndb, 32), dtype=np.uint8)
    if ndb > 10:
        db[ndb // 2] = db[3]        # duplicates: ties go to the lowest index
        db[ndb - 1] = db[3]
    q = rng.integers(0, 256, (nq, 32), dtype=np.uint8)
    if ndb > 10 and nq > 3:
        q[0] = db[3]
        q[1] = db[ndb - 1]
        q[2] = db[5] ^ 1
    out = ctx.match_top2(torch.from_numpy(q).cuda(), torch.from_numpy(db).cuda(), method=method)
    torch.cuda.synchronize()
    got = [o.cpu().numpy() for o in out]
    want = oracle.match_top2(q, db, nthreads=8)
    for g, w_ in zip(got, want):
        assert np.array_equal(g, w_)


@pytest.mark.parametrize("method", [0, 1], ids=["popcount", "mfma"])
def test_match_top2_near_duplicates_full_size_property(torch, ctx, oracle, method):
    """Queries that are copies of database rows with Binomial(256, 0.05) bit flips (the dense-matching
    workload at 1/4 of its database): the matcher's distance to the source row is the number of flips,
    so d1 <= flips and idx1 == source unless another row ties or beats it; cross-checked on a subset."""
    rng = np.random.default_rng(42)
    ndb, nq = 65536, 16384
    db = rng.integers(0, 256, (ndb, 32), dtype=np.uint8)
    src = rng.integers(0, ndb, nq)
    flips = rng.random((nq, 256)) < 0.05
    q = db[src] ^ np.packbits(flips, axis=1, bitorder="little")
    out = ctx.match_top2(torch.from_numpy(q).cuda(), torch.from_numpy(db).cuda(), method=method)
    torch.cuda.synchronize()
    i1, d1, i2, d2 = (o.cpu().numpy() for o in out)
    nflip = flips.sum(1)
    assert (d1 <= nflip).all() and (d1 <= d2).all()
    assert ((i1 == src) | (d1 < nflip) | ((d1 == nflip) & (i1 < src))).all()
    assert (i1 == src).mean() > 0.999
    sub = slice(0, 512)
    want = oracle.match_top2(q[sub], db, nthreads=8)
    for g, w_ in zip((i1[sub], d1[sub], i2[sub], d2[sub]), want):
        assert np.array_equal(g, w_)


def test_device_sincos_and_atan2_bit_exact(torch, ctx, oracle):
    a = np.concatenate([np.linspace(-math.pi, math.pi, 40001, dtype=np.float32),
                        np.array([0.0, -0.0, math.pi / 2, -math.pi / 2, math.pi / 4], np.float32)])
    s, c = ctx.debug_sincos(torch.from_numpy(a).cuda())
    torch.cuda.synchronize()
    want = np.array([oracle.sincosf(float(v)) for v in a], np.float32)
    assert np.array_equal(s.cpu().numpy().view(np.uint32), want[:, 0].view(np.uint32))
    assert np.array_equal(c.cpu().numpy().view(np.uint32), want[:, 1].view(np.uint32))
    rng = np.random.default_rng(3)
    yx = rng.integers(-3_000_000, 3_000_000, (20000, 2)).astype(np.float32)
    yx[:8] = [[0, 5], [5, 0], [0, -5], [-5, 0], [3, 3], [-3, -3], [7, -7], [1, 2]]
    got = ctx.debug_atan2(torch.from_numpy(yx).cuda()).cpu().numpy()
    want = np.array([oracle.fast_atan2(float(y), float(x)) for y, x in yx], np.float32)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
