"""Pins of the CPU oracle (oracle/orb_oracle.c) against what does not depend on
its own formulas: closed forms, the reference's data tables, glibc, libm,
scipy, textbook definitions and brute force on small inputs.  CPU only.

The reference ships no tests and no golden vectors (its data/ holds images
without expected outputs), so these checks are the oracle's pins.
"""
import ctypes
import hashlib
import math
import pathlib

import numpy as np
import pytest

REF = pathlib.Path("/root/reference")
PATTERN_SHA256 = "2164181aea6ff9ac426ca512d5130d15e1f6e3cd47b1cbdd568bbe1e55d49023"


# ----------------------------------------------------------------- tables

def test_fast_lut_closed_form(oracle):
    for t in (0, 7, 20, 254):
        lut = oracle.fast_lut(t)
        for i in range(-255, 256):
            want = 1 if i < -t else (2 if i > t else 0)
            assert lut[i + 255] == want


def test_umax_closed_form(oracle):
    u = oracle.umax(31)
    assert list(u[:17]) == [15, 15, 15, 15, 14, 14, 14, 13, 13, 12, 11, 10, 9, 8, 6, 3, 0]
    # 749-pixel disc (orbd.cu:1014-1030), symmetric under transposition
    assert sum(2 * u[abs(v)] + 1 for v in range(-15, 16)) == 749
    disc = {(x, y) for y in range(-15, 16) for x in range(-u[abs(y)], u[abs(y)] + 1)}
    assert disc == {(y, x) for (x, y) in disc}


def test_gauss_taps_closed_form(oracle):
    k = oracle.gauss()
    assert [hex(v) for v in k.view(np.uint32)] == ["0x3e5d4ae1", "0x3e434a3a", "0x3e06387f", "0x3d8fafb2"]
    ref = np.array([math.exp(-i * i / 8.0) for i in range(4)])
    ref /= ref[0] + 2 * ref[1:].sum()
    assert np.allclose(k, ref, rtol=3e-7, atol=0)


def test_harris_scale_closed_form(oracle):
    s = oracle.lib().orc_harris_scale()
    assert s == pytest.approx((1.0 / (4 * 7 * 255.0)) ** 4, rel=1e-6)


def test_glibc_rand_restatement_matches_libc(oracle):
    libc = ctypes.CDLL("libc.so.6")
    for seed in (0, 1, 0x12345678, 0x34985739):
        libc.srand(seed)
        want = [libc.rand() for _ in range(2000)]
        assert oracle.glibc_rand_seq(seed, 2000) == want


def test_pattern31_is_the_reference_table(oracle):
    pat = oracle.pattern(2)
    assert pat.shape == (512, 2)
    assert hashlib.sha256(bytes(int(v) & 255 for v in pat.ravel())).hexdigest() == PATTERN_SHA256
    assert pat.min() == -13 and pat.max() == 12
    if (REF / "orbd.cu").exists():
        import importlib.util
        spec = importlib.util.spec_from_file_location("mk", pathlib.Path(__file__).parent / "golden" / "make_pattern.py")
        mk = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mk)
        assert list(pat.ravel()) == mk.parse((REF / "orbd.cu").read_text())


def test_wta_patterns_are_distinct_tuples_of_the_table(oracle):
    base = {tuple(p) for p in oracle.pattern(2)}
    for k in (3, 4):
        pat = oracle.pattern(k)
        tuples = pat[: 128 * k].reshape(128, k, 2)
        for t in tuples:
            pts = [tuple(p) for p in t]
            assert len(set(pts)) == k and all(p in base for p in pts)


# ------------------------------------------------------------------- math

def test_fast_atan2_against_libm(oracle):
    rng = np.random.default_rng(0)
    worst = 0.0
    for _ in range(20000):
        y, x = (int(v) for v in rng.integers(-3_000_000, 3_000_000, 2))
        if x == 0 and y == 0:
            continue
        got = oracle.fast_atan2(float(y), float(x))
        worst = max(worst, abs(got - math.atan2(y, x)))
    # the reference's 7th-order polynomial (orbd.cu:992): max error about 2e-4 rad
    assert 1e-5 < worst < 2.5e-4
    for (y, x) in ((0, 1), (1, 0), (0, -1), (-1, 0), (1, 1), (-1, -1), (5, -5)):
        assert oracle.fast_atan2(float(y), float(x)) == pytest.approx(math.atan2(y, x), abs=2.5e-4)


def test_sincos_against_libm(oracle):
    worst = 0.0
    for a in np.linspace(-math.pi, math.pi, 20001, dtype=np.float32):
        s, c = oracle.sincosf(float(a))
        worst = max(worst, abs(s - math.sin(float(a))), abs(c - math.cos(float(a))))
    assert worst < 3e-7


# --------------------------------------------------------------- geometry

def test_geometry_reference_octaves(oracle):
    g = oracle.geometry(1920, 1080, 5, 2.0)
    assert g.nlevels == 4
    assert [g.p[i] for i in range(4)] == [1920, 1024, 512, 256]
    assert [g.w[i] for i in range(4)] == [1920, 960, 480, 240]
    g = oracle.geometry(640, 480, 5, 2.0)
    assert g.nlevels == 3 and [g.p[i] for i in range(3)] == [640, 384, 256]
    assert oracle.geometry(3840, 2160, 8, 2.0).nlevels == 5   # MAX_OCTAVE
    with pytest.raises(ValueError):
        oracle.geometry(79, 600, 5, 2.0)


def test_geometry_scale_1_2(oracle):
    g = oracle.geometry(1920, 1080, 8, 1.2)
    sizes = [(g.w[i], g.h[i]) for i in range(g.nlevels)]
    assert sizes == [(1920, 1080), (1600, 900), (1333, 750), (1111, 625), (926, 521), (772, 434),
                     (643, 362), (536, 301)]
    assert sum(w * h for w, h in sizes) == 6419321
    assert [g.radius[i] for i in range(8)] == [1, 1, 1, 2, 2, 2, 3, 4]
    off = 0
    for i in range(8):
        assert g.off[i] == off and g.p[i] % 128 == 0 and g.p[i] >= g.w[i]
        off += g.h[i] * g.p[i]
    assert g.total == off


def test_level_coordinate_round_trip(oracle):
    import ctypes as C
    for scale in (1.2, 1.5, 2.0):
        g = oracle.geometry(1280, 720, 8, scale)
        for l in range(g.nlevels):
            for v in range(g.w[l]):
                x0 = oracle.lib().orc_level_to_l0(C.byref(g), l, v)
                assert oracle.lib().orc_l0_to_level(C.byref(g), l, x0) == v


# ---------------------------------------------------------------- pyramid

def _rand_img(h, w, seed):
    return np.random.default_rng(seed).integers(0, 256, (h, w), dtype=np.uint8)


def test_pyramid_reference_decimation(oracle):
    img = _rand_img(480, 640, 1)
    g = oracle.geometry(640, 480, 5, 2.0)
    raw, _ = oracle.pyramid(g, oracle.pitched(img))
    for l in range(g.nlevels):
        lev = raw[g.off[l]: g.off[l] + g.h[l] * g.p[l]].reshape(g.h[l], g.p[l])
        assert np.array_equal(lev[:, : g.w[l]], img[:: 1 << l, :: 1 << l][: g.h[l], : g.w[l]])
        assert not lev[:, g.w[l]:].any()


def test_pyramid_bilinear_against_float_resize(oracle):
    img = _rand_img(240, 320, 2).astype(np.float64)
    g = oracle.geometry(320, 240, 8, 1.2)
    raw, _ = oracle.pyramid(g, oracle.pitched(img.astype(np.uint8)))
    for l in range(1, g.nlevels):
        w, h = g.w[l], g.h[l]
        lev = raw[g.off[l]: g.off[l] + h * g.p[l]].reshape(h, g.p[l])[:, :w].astype(np.float64)
        sx = np.clip((np.arange(w) + 0.5) * 320 / w - 0.5, 0, 319)
        sy = np.clip((np.arange(h) + 0.5) * 240 / h - 0.5, 0, 239)
        x0 = np.floor(sx).astype(int); y0 = np.floor(sy).astype(int)
        x1 = np.minimum(x0 + 1, 319); y1 = np.minimum(y0 + 1, 239)
        fx = sx - x0; fy = sy - y0
        top = img[y0][:, x0] * (1 - fx) + img[y0][:, x1] * fx
        bot = img[y1][:, x0] * (1 - fx) + img[y1][:, x1] * fx
        ref = top * (1 - fy)[:, None] + bot * fy[:, None]
        assert np.abs(lev - ref).max() <= 2.0   # 8-bit fixed-point weights


def test_blur_against_scipy_mirror(oracle):
    from scipy import ndimage
    img = _rand_img(97, 131, 3)
    g = oracle.geometry(131, 97, 1, 1.2)
    raw, blur = oracle.pyramid(g, oracle.pitched(img))
    k = oracle.gauss().astype(np.float64)
    k1d = np.array([k[3], k[2], k[1], k[0], k[1], k[2], k[3]])
    ref = ndimage.correlate1d(img.astype(np.float64), k1d, axis=1, mode="mirror")
    ref = ndimage.correlate1d(ref, k1d, axis=0, mode="mirror")
    got = blur[: 97 * g.p[0]].reshape(97, g.p[0])[:, :131]
    diff = np.abs(got.astype(np.float64) - ref)
    assert diff.max() <= 1.0 and (diff > 0.5 + 1e-3).sum() == 0


def test_blur_impulse_and_constant(oracle):
    img = np.zeros((64, 64), np.uint8)
    img[32, 32] = 255
    g = oracle.geometry(64, 64, 1, 1.2)
    _, blur = oracle.pyramid(g, oracle.pitched(img))
    b = blur[: 64 * 128].reshape(64, 128)[:, :64].astype(int)
    k = oracle.gauss().astype(np.float64)
    kk = np.array([k[3], k[2], k[1], k[0], k[1], k[2], k[3]])
    ref = np.rint(255 * np.outer(kk, kk))
    assert np.abs(b[29:36, 29:36] - ref).max() <= 1
    assert b.sum() - b[29:36, 29:36].sum() == 0
    img = np.full((40, 50), 173, np.uint8)
    g = oracle.geometry(50, 40, 1, 1.2)
    _, blur = oracle.pyramid(g, oracle.pitched(img))
    assert (blur[: 40 * 128].reshape(40, 128)[:, :50] == 173).all()


# -------------------------------------------------------------- detection

CIRCLE = [(0, 3), (1, 3), (2, 2), (3, 1), (3, 0), (3, -1), (2, -2), (1, -3), (0, -3), (-1, -3),
          (-2, -2), (-3, -1), (-3, 0), (-3, 1), (-2, 2), (-1, 3)]


def _run9(flags):
    f = flags + flags
    return any(all(f[i:i + 9]) for i in range(16))


def test_fast_segment_test_is_textbook_fast9(oracle):
    """FAST-9: >= 9 contiguous circle pixels all darker than c-t or all brighter than c+t."""
    rng = np.random.default_rng(4)
    lut = oracle.fast_lut(20)
    hits = 0
    for trial in range(4000):
        patch = np.zeros((7, 128), np.uint8)
        c = int(rng.integers(30, 226))
        patch[3, 3] = c
        mode = trial % 3
        start = int(rng.integers(0, 16))
        run = int(rng.integers(6, 13))
        vals = []
        for k, (dx, dy) in enumerate(CIRCLE):
            if mode == 0:
                v = int(rng.integers(0, 256))
            else:
                inside = (k - start) % 16 < run
                v = (c - 25 if mode == 1 else c + 25) if inside else int(rng.integers(c - 20, c + 21))
                v = max(0, min(255, v))
            patch[3 + dy, 3 + dx] = v
            vals.append(v)
        flat = patch.ravel()
        got = oracle.lib().orc_fast_corner(oracle.ptr(flat), flat.size, 3 * 128 + 3, 128, oracle.ptr(lut), 20)
        want = _run9([v < c - 20 for v in vals]) or _run9([v > c + 20 for v in vals])
        assert bool(got) == want, (c, vals)
        hits += want
    assert hits > 500


def test_fast_score_is_the_largest_threshold_keeping_a_9_arc(oracle):
    """FAST_SCORE (OpenCV cornerScore<16>, orbd.cu:538-648) in closed form: with d_k = centre -
    ring_k, the score is max(t, max over 9-arcs of min d, max over 9-arcs of -max d) - 1, i.e. one
    less than the smallest margin by which some contiguous 9-arc is all brighter or all darker."""
    rng = np.random.default_rng(14)
    t = 20
    for trial in range(6):
        img = rng.integers(0, 256, (40, 128)).astype(np.uint8)
        if trial % 2:   # smooth regions with a few strong corners, so large scores occur
            img = (rng.integers(0, 2, (40, 128)) * 200 + rng.integers(0, 30, (40, 128))).astype(np.uint8)
        flat = img.ravel()
        for y in range(3, 37):
            for x in range(3, 125):
                c = int(img[y, x])
                d = np.array([c - int(img[y + dy, x + dx]) for dx, dy in CIRCLE])
                arcs = np.stack([np.roll(d, -s)[:9] for s in range(16)])
                want = max(t, int(arcs.min(1).max()), int((-arcs.max(1)).max())) - 1
                got = oracle.lib().orc_fast_score(oracle.ptr(flat), flat.size, y * 128 + x, 128, t)
                assert got == want, (trial, x, y)


def _wta_codes(B, cx, cy, pat, k):
    """2-bit argmax codes at angle 0 straight from the pattern table (values distinct)."""
    codes, distinct = [], []
    for b in range(32):
        idx = ([[12 * b + 3 * q + i for i in range(3)] for q in range(3)] + [[12 * b + 8, 12 * b + 9, 12 * b + 11]]
               if k == 3 else [[16 * b + 4 * q + i for i in range(4)] for q in range(4)])
        for tup in idx:
            v = [int(B[cy + pat[j][1], cx + pat[j][0]]) for j in tup]
            codes.append(int(np.argmax(v)))
            distinct.append(len(set(v)) == len(v))
    return np.array(codes), np.array(distinct)


@pytest.mark.parametrize("k", [3, 4])
def test_descriptor_wta_codes_are_argmax_of_pattern_tuples(oracle, k):
    """WTA_K 3/4 (orbd.cu:1364-1434): each 2-bit code is the index of the brightest sample of its
    tuple; for WTA_K 3 the fourth tuple of every byte is samples 8, 9, 11 (the reference's
    indices).  Checked wherever the tuple's values are distinct, so no tie rule is assumed."""
    rng = np.random.default_rng(20 + k)
    img = rng.integers(0, 256, (128, 128)).astype(np.uint8)
    g = oracle.geometry(128, 128, 1, 1.2)
    _, blur = oracle.pyramid(g, oracle.pitched(img))
    B = blur[: 128 * 128].reshape(128, 128)
    pat = oracle.pattern(k)
    for cx, cy in ((60, 70), (40, 41), (85, 90)):
        pt = np.zeros(1, oracle.POINT_DTYPE)
        pt["x"], pt["y"] = cx, cy
        desc = oracle.describe(g, blur, pt, k)[0]
        got = np.stack([(desc >> s) & 3 for s in (0, 2, 4, 6)], -1).ravel()
        want, distinct = _wta_codes(B, cx, cy, pat, k)
        assert distinct.mean() > 0.5
        assert np.array_equal(got[distinct], want[distinct])


def test_harris_against_float64_with_wrapped_ints(oracle):
    """Harris response vs an independent numpy evaluation (int64 Sobel and box sums wrapped to
    int32, float64 tail); the float32 result must agree to a few ulp, and the sample must include
    windows whose products wrap."""
    rng = np.random.default_rng(5)
    s = (1.0 / (4 * 7 * 255.0)) ** 4
    wrapped = 0
    for trial in range(300):
        p = 64
        img = rng.integers(0, 256, (16, p), dtype=np.uint8)
        if trial % 2:
            img[:, :8] = 0
            img[:, 8:] = 255
            img[:8, :] //= 3
        a = img.astype(np.int64)
        cy, cx = 8, 8
        dxx = dyy = dxy = 0
        for yy in range(cy - 3, cy + 4):
            for xx in range(cx - 3, cx + 4):
                dx = 2 * (a[yy, xx + 1] - a[yy, xx - 1]) + (a[yy - 1, xx + 1] - a[yy - 1, xx - 1]) + (
                    a[yy + 1, xx + 1] - a[yy + 1, xx - 1])
                dy = 2 * (a[yy + 1, xx] - a[yy - 1, xx]) + (a[yy + 1, xx - 1] - a[yy - 1, xx - 1]) + (
                    a[yy + 1, xx + 1] - a[yy - 1, xx + 1])
                dxx += dx * dx; dyy += dy * dy; dxy += dx * dy
        A64 = dxx * dyy - dxy * dxy
        B64 = (dxx + dyy) ** 2
        A = np.int64(A64).astype(np.uint64).astype(np.uint32).astype(np.int32)
        B = np.int64(B64).astype(np.uint64).astype(np.uint32).astype(np.int32)
        wrapped += (int(A) != A64) or (int(B) != B64)
        want = (float(np.float32(A)) - float(np.float32(0.04)) * float(np.float32(B))) * float(np.float32(s))
        flat = img.ravel()
        got = oracle.lib().orc_harris(oracle.ptr(flat), flat.size, cy * p + cx, p)
        assert got == pytest.approx(want, rel=4e-7, abs=1e-30)
    assert wrapped > 50


def test_nms_brute_force_small_images(oracle):
    rng = np.random.default_rng(6)
    for trial in range(6):
        h, w = 96, 128
        img = (rng.integers(0, 2, (h // 8, w // 8)) * 200 + 20).astype(np.uint8)
        img = np.kron(img, np.ones((8, 8), np.uint8)) + rng.integers(0, 15, (h, w)).astype(np.uint8)
        g = oracle.geometry(w, h, 1, 1.2)
        raw, _ = oracle.pyramid(g, oracle.pitched(img))
        vmap, lmap = oracle.detect_map(g, raw, 20, 0)
        p = g.p[0]
        V = vmap.reshape(h, p)
        # every positive map entry is exactly the FAST-9 corners with positive Harris
        lut = oracle.fast_lut(20)
        for y in range(3, h - 3):
            for x in range(3, w - 3):
                c = y * p + x
                corner = oracle.lib().orc_fast_corner(oracle.ptr(raw), g.total, c, p, oracle.ptr(lut), 20)
                s = oracle.lib().orc_harris(oracle.ptr(raw), g.total, c, p) if corner else 0.0
                assert V[y, x] == (np.float32(s) if s > 0 else 0.0)
        pts, nsurv = oracle.nms(g, vmap, lmap, border=10, max_pts=100000)
        want = []
        for y in range(10, h - 10):
            for x in range(10, w - 10):
                v = V[y, x]
                if v <= 0:
                    continue
                sup = False
                for dy in (-1, 0, 1):
                    for dx in (-1, 0, 1):
                        if (dx or dy) and V[y + dy, x + dx] > 0 and (
                                V[y + dy, x + dx] > v or (V[y + dy, x + dx] == v and dx <= 0 and dy <= 0)):
                            sup = True
                if not sup:
                    want.append((x, y))
        assert [(int(a), int(b)) for a, b in zip(pts["x"], pts["y"])] == want
        assert nsurv == len(want)


def _nms_brute(V, L, radii, border, w, h):
    """Survivors of the reference's square NMS (orbd.cu:945-966) by direct enumeration: a point
    (y, x) with score v > 0 on level l is dropped iff some (dx, dy) != 0 with |dx|, |dy| <= radii[l]
    holds v' > 0 with v' > v, or v' == v and dx <= 0 and dy <= 0.  Rows outside the map read 0."""
    hh, p = V.shape
    out = []
    ys, xs = np.nonzero(V > 0)
    for y, x in sorted(zip(ys.tolist(), xs.tolist())):
        if not (border <= y < h - border and border <= x < w - border):
            continue
        v, r = V[y, x], radii[L[y, x]]
        sup = False
        for dy in range(-r, r + 1):
            yy = y + dy
            if yy < 0 or yy >= hh:
                continue
            for dx in range(-r, r + 1):
                if dx == 0 and dy == 0:
                    continue
                nv = V[yy, x + dx]
                if nv > 0 and (nv > v or (nv == v and dx <= 0 and dy <= 0)):
                    sup = True
                    break
            if sup:
                break
        if not sup:
            out.append((x, y))
    return out


def test_nms_brute_force_scale12_eight_levels(oracle):
    """Detection maps of 128x128 images over 8 levels at scale 1.2, so corners of every level
    (radius rn(1.2^l): 1,1,1,2,2,2,3,4) compete in one level-0 map."""
    rng = np.random.default_rng(16)
    radii = [max(1, int(math.floor(1.2 ** l + 0.5))) for l in range(8)]
    assert radii == [1, 1, 1, 2, 2, 2, 3, 4]
    levels_seen = set()
    for trial in range(4):
        h = w = 128
        blk = [4, 6, 8, 12][trial]
        img = (rng.integers(0, 2, (h // blk + 1, w // blk + 1)) * 180 + 30).astype(np.uint8)
        img = np.kron(img, np.ones((blk, blk), np.uint8))[:h, :w] + rng.integers(0, 12, (h, w)).astype(np.uint8)
        g = oracle.geometry(w, h, 8, 1.2)
        assert g.nlevels == 8 and list(g.radius[:8]) == radii
        raw, _ = oracle.pyramid(g, oracle.pitched(img))
        vmap, lmap = oracle.detect_map(g, raw, 20, 0)
        V, L = vmap.reshape(h, g.p[0]), lmap.reshape(h, g.p[0])
        levels_seen |= set(L[V > 0].tolist())
        pts, nsurv = oracle.nms(g, vmap, lmap, border=8, max_pts=100000)
        want = _nms_brute(V, L, radii, 8, w, h)
        assert [(int(a), int(b)) for a, b in zip(pts["x"], pts["y"])] == want
        assert nsurv == len(want)
        assert list(pts["octave"]) == [int(L[y, x]) for x, y in want]
    assert {0, 3, 5, 6, 7} <= levels_seen, levels_seen


def test_nms_brute_force_scale2_five_octaves_with_ties(oracle):
    """The reference's octave radii 2^l up to 16 (orbd.cu:941), on a synthetic level-0 map of a
    1280x1280 frame (5 octaves) with few distinct scores, so ties at every offset occur."""
    rng = np.random.default_rng(17)
    g = oracle.geometry(1280, 1280, 5, 2.0)
    radii = [1 << l for l in range(5)]
    assert g.nlevels == 5 and list(g.radius[:5]) == radii
    h, w, p = g.h[0], g.w[0], g.p[0]
    for trial in range(3):
        V = np.zeros((h, p), np.float32)
        L = np.zeros((h, p), np.int32)
        n = 6000
        ys, xs = rng.integers(0, h, n), rng.integers(0, w, n)
        V[ys, xs] = rng.integers(1, 5, n).astype(np.float32) * np.float32(0.5)
        L[ys, xs] = rng.integers(0, 5, n)
        pts, nsurv = oracle.nms(g, V.ravel(), L.ravel(), border=31, max_pts=100000)
        want = _nms_brute(V, L, radii, 31, w, h)
        assert [(int(a), int(b)) for a, b in zip(pts["x"], pts["y"])] == want
        assert nsurv == len(want) > 100


def test_nms_cap_keeps_best_scores_in_raster_order(oracle):
    img = oracle_synth(640, 480, 40, 11)
    pts_all, _, ns = oracle.detect_and_compute(img, max_pts=100000)
    assert len(pts_all) == ns > 300
    cap = 300
    pts, _, ns2 = oracle.detect_and_compute(img, max_pts=cap)
    assert ns2 == ns and len(pts) == cap
    order = np.lexsort((np.arange(len(pts_all)), -pts_all["score"]))[:cap]
    want = pts_all[np.sort(order)]
    assert np.array_equal(pts, want)


def test_constant_image_has_no_keypoints(oracle):
    pts, desc, ns = oracle.detect_and_compute(np.full((480, 640), 90, np.uint8))
    assert len(pts) == 0 and ns == 0


# --------------------------------------------------- angle and descriptors

def test_centroid_angle_of_ramps(oracle):
    g = oracle.geometry(256, 256, 1, 1.2)
    for ang in np.linspace(-3.0, 3.0, 13):
        yy, xx = np.mgrid[0:256, 0:256]
        img = np.clip(128 + 3.0 * ((xx - 128) * math.cos(ang) + (yy - 128) * math.sin(ang)), 0, 255)
        raw, _ = oracle.pyramid(g, oracle.pitched(img.astype(np.uint8)))
        pt = np.zeros(1, oracle.POINT_DTYPE)
        pt["x"], pt["y"] = 128, 128
        got = oracle.angles(g, raw, pt)["angle"][0]
        d = (got - ang + math.pi) % (2 * math.pi) - math.pi
        assert abs(d) < 0.02, (ang, got)


def test_angle_rotates_with_the_patch(oracle):
    rng = np.random.default_rng(7)
    g = oracle.geometry(128, 128, 1, 1.2)
    for _ in range(20):
        img = rng.integers(0, 256, (128, 128)).astype(np.uint8)
        pt = np.zeros(1, oracle.POINT_DTYPE)
        pt["x"], pt["y"] = 64, 64
        a0 = oracle.angles(g, oracle.pyramid(g, oracle.pitched(img))[0], pt)["angle"][0]
        rot = np.ascontiguousarray(np.rot90(img, k=-1))   # 90 deg clockwise on screen (y down)
        # rot90(k=-1): new[y, x] = old[127 - x, y]; centre (64,64) maps to (63, 64)
        pt2 = pt.copy()
        pt2["x"], pt2["y"] = 63, 64
        a1 = oracle.angles(g, oracle.pyramid(g, oracle.pitched(rot))[0], pt2)["angle"][0]
        d = (a1 - a0 - math.pi / 2 + math.pi) % (2 * math.pi) - math.pi
        assert abs(d) < 2e-4


def test_descriptor_at_angle_zero_is_the_plain_pattern(oracle):
    rng = np.random.default_rng(8)
    img = rng.integers(0, 256, (128, 128)).astype(np.uint8)
    g = oracle.geometry(128, 128, 1, 1.2)
    _, blur = oracle.pyramid(g, oracle.pitched(img))
    B = blur[: 128 * 128].reshape(128, 128)
    pt = np.zeros(1, oracle.POINT_DTYPE)
    pt["x"], pt["y"] = 60, 70
    desc = oracle.describe(g, blur, pt, 2)[0]
    pat = oracle.pattern(2)
    bits = [int(B[70 + pat[2 * k][1], 60 + pat[2 * k][0]] < B[70 + pat[2 * k + 1][1], 60 + pat[2 * k + 1][0]])
            for k in range(256)]
    assert np.array_equal(np.unpackbits(desc, bitorder="little"), np.array(bits, np.uint8))


def test_descriptor_rotation_by_quarter_turn(oracle):
    """At angle +pi/2 (the own sin/cos gives exactly 1 and 0 up to 1 ulp) each point (x, y)
    samples (-y, x)."""
    rng = np.random.default_rng(9)
    img = rng.integers(0, 256, (128, 128)).astype(np.uint8)
    g = oracle.geometry(128, 128, 1, 1.2)
    _, blur = oracle.pyramid(g, oracle.pitched(img))
    B = blur[: 128 * 128].reshape(128, 128)
    pt = np.zeros(1, oracle.POINT_DTYPE)
    pt["x"], pt["y"], pt["angle"] = 60, 70, np.float32(math.pi / 2)
    desc = oracle.describe(g, blur, pt, 2)[0]
    pat = oracle.pattern(2)
    val = lambda q: B[70 + q[0], 60 - q[1]]
    bits = [int(val(pat[2 * k]) < val(pat[2 * k + 1])) for k in range(256)]
    assert np.array_equal(np.unpackbits(desc, bitorder="little"), np.array(bits, np.uint8))


def test_descriptor_wta_codes_in_range(oracle):
    img = oracle_synth(320, 240, 20, 12)
    for k in (3, 4):
        pts, desc, _ = oracle.detect_and_compute(img, wta_k=k)
        codes = np.stack([(desc >> s) & 3 for s in (0, 2, 4, 6)], -1)
        assert len(pts) > 20
        if k == 3:
            assert codes.max() <= 2
        assert len(np.unique(codes)) == (3 if k == 3 else 4)


# --------------------------------------------------------------- matching

def test_match_top2_brute_force_with_ties(oracle):
    rng = np.random.default_rng(10)
    db = rng.integers(0, 256, (300, 32), dtype=np.uint8)
    db[150] = db[20]          # duplicate entries: ties must go to the lowest index
    db[299] = db[20]
    q = np.concatenate([db[[20, 5, 150]], rng.integers(0, 256, (50, 32), dtype=np.uint8)])
    i1, d1, i2, d2 = oracle.match_top2(q, db)
    bits_q = np.unpackbits(q, axis=1)
    bits_db = np.unpackbits(db, axis=1)
    D = (bits_q[:, None, :] != bits_db[None, :, :]).sum(-1)
    for r in range(len(q)):
        order = np.lexsort((np.arange(300), D[r]))
        assert (i1[r], d1[r], i2[r], d2[r]) == (order[0], D[r, order[0]], order[1], D[r, order[1]])
    assert (i1[0], d1[0], i2[0], d2[0]) == (20, 0, 150, 0)


def test_match_reference_rule(oracle):
    rng = np.random.default_rng(11)
    d2 = rng.integers(0, 256, (200, 32), dtype=np.uint8)
    d1 = d2[rng.permutation(200)[:80]].copy()
    d1[:10] ^= 1                    # distance 10 from its source, others far away
    d1[10] = d2[3]; d2[4] = d2[3]   # a tied best is never accepted
    m, d = oracle.match(d1, d2)
    i1, b1, i2, b2 = oracle.match_top2(d1, d2)
    for r in range(80):
        ok = b1[r] < b2[r] and b1[r] < 64
        assert (m[r], d[r]) == ((i1[r], b1[r]) if ok else (-1, -1))
    assert m[10] == -1
    assert (m[:10] >= 0).all()
    m, d = oracle.match(d1, d2[:0])
    assert (m == -1).all() and (d == -1).all()


def test_hamming_metric_properties(oracle):
    rng = np.random.default_rng(12)
    x = rng.integers(0, 256, (40, 32), dtype=np.uint8)
    D = np.zeros((40, 40), int)
    for i in range(40):
        _, d1, _, _ = oracle.match_top2(x[i:i + 1], x[i:i + 1])
        assert d1[0] == 0
        for j in range(40):
            D[i, j] = oracle.match_top2(x[i:i + 1], x[j:j + 1])[1][0]
    assert (D == D.T).all()
    assert all(D[i, k] <= D[i, j] + D[j, k] for i in range(10) for j in range(10) for k in range(10))


def test_hamming2_closed_form_and_bounds(oracle):
    """NORM_HAMMING2 counts the bit pairs (2i, 2i+1) that differ: hand cases, a brute-force
    count over the unpacked bits, and ham2 <= ham <= 2 ham2."""
    z = np.zeros((1, 32), np.uint8)
    for byte, bits, want in ((0, 0b11, 1), (0, 0b110, 2), (5, 0b11000000, 1), (31, 0xff, 4)):
        b = z.copy()
        b[0, byte] = bits
        assert oracle.match_top2(z, b, norm=1)[1][0] == want
        assert oracle.match_top2(z, b, norm=0)[1][0] == bin(bits).count("1")
    rng = np.random.default_rng(13)
    q = rng.integers(0, 256, (30, 32), dtype=np.uint8)
    db = rng.integers(0, 256, (200, 32), dtype=np.uint8)
    db[77] = db[12]                       # tie: the lower index wins
    q[0] = db[12]
    bq = np.unpackbits(q, axis=1, bitorder="little").reshape(30, 128, 2)
    bd = np.unpackbits(db, axis=1, bitorder="little").reshape(200, 128, 2)
    D2 = (bq[:, None] != bd[None]).any(-1).sum(-1)
    D1 = (bq[:, None] != bd[None]).sum((-1, -2))
    assert ((D2 <= D1) & (D1 <= 2 * D2)).all()
    i1, d1, i2, d2 = oracle.match_top2(q, db, norm=1)
    for r in range(30):
        order = np.lexsort((np.arange(200), D2[r]))
        assert (i1[r], d1[r], i2[r], d2[r]) == (order[0], D2[r, order[0]], order[1], D2[r, order[1]])
    assert (i1[0], d1[0], i2[0], d2[0]) == (12, 0, 77, 0)


def export_points(oracle, scale, rng, n):
    """Points at known level coordinates, mapped to level 0 by the detector's own rule."""
    import ctypes
    g = oracle.geometry(1920, 1080, 8 if scale != 2.0 else 5, scale)
    pts = np.zeros(n, oracle.POINT_DTYPE)
    lv = rng.integers(0, g.nlevels, n)
    lxy = np.zeros((n, 2), int)
    for i in range(n):
        l = int(lv[i])
        lx, ly = int(rng.integers(3, g.w[l] - 3)), int(rng.integers(3, g.h[l] - 3))
        lxy[i] = (lx, ly)
        pts["x"][i] = oracle.lib().orc_level_to_l0(ctypes.byref(g), l, lx)
        pts["y"][i] = oracle.lib().orc_level_to_l0(ctypes.byref(g), l, ly)
        pts["octave"][i] = l
    pts["score"] = rng.integers(1, 40, n).astype(np.float32) * np.float32(1e-3)   # many ties
    pts["angle"] = rng.uniform(-math.pi, math.pi, n).astype(np.float32)
    return pts, lxy


@pytest.mark.parametrize("scale", [1.2, 2.0])
def test_export_cv_coordinates_angles_retain(oracle, scale):
    """The export recovers each point's level coordinate (round trip through the detector's
    level -> level-0 rule), scales it by float32(scale**l), keeps OpenCV's degree range, and
    with retain_topn keeps exactly the top scores with ties to the earlier point, in order,
    descriptors following."""
    rng = np.random.default_rng(14)
    pts, lxy = export_points(oracle, scale, rng, 400)
    desc = rng.integers(0, 256, (400, 32), dtype=np.uint8)
    kp, d = oracle.export_cv(pts, desc, -1, scale)
    assert len(kp) == 400 and np.array_equal(d, desc)
    lsc = np.array([np.float32(scale ** l) for l in pts["octave"]], np.float32)
    assert np.array_equal(kp["x"], lxy[:, 0].astype(np.float32) * lsc)
    assert np.array_equal(kp["y"], lxy[:, 1].astype(np.float32) * lsc)
    if scale == 2.0:
        assert np.array_equal(kp["x"], pts["x"].astype(np.float32))
    assert np.array_equal(kp["size"], np.float32(31) * lsc)
    assert ((kp["angle"] >= 0) & (kp["angle"] < 360)).all()
    want = np.degrees(pts["angle"].astype(np.float64)) % 360.0
    err = np.abs(kp["angle"] - want)
    assert (np.minimum(err, 360 - err) < 1e-3).all()
    assert (kp["octave"] == pts["octave"]).all() and (kp["class_id"] == -1).all()
    for k in (1, 57, 399, 400, 1000):
        kk, dk = oracle.export_cv(pts, desc, k, scale)
        best = sorted(range(400), key=lambda i: (-float(pts["score"][i]), i))[:min(k, 400)]
        keep = sorted(best)
        assert np.array_equal(kk["x"], kp["x"][keep]) and np.array_equal(kk["y"], kp["y"][keep])
        assert np.array_equal(kk["response"], pts["score"][keep])
        assert np.array_equal(dk, desc[keep])


# -------------------------------------------------------------- end to end

def oracle_synth(w, h, nshapes, seed):
    from orb_amd import synth_image
    return synth_image(seed, w, h, nshapes)[:, :w].copy()


def test_pair_matches_agree_with_ground_truth_homography(oracle):
    from orb_amd import synth_image, synth_warp
    a = synth_image(0, 640, 480, 40)
    b, H = synth_warp(0, a, 640, 480)
    pa, da, _ = oracle.detect_and_compute(a[:, :640].copy(), max_pts=1000)
    pb, db, _ = oracle.detect_and_compute(b[:, :640].copy(), max_pts=1000)
    m, d = oracle.match(da, db)
    ok = m >= 0
    assert ok.sum() > 100
    src = np.stack([pa["x"][ok], pa["y"][ok], np.ones(ok.sum())], 0)
    dst = H @ src
    dst = dst[:2] / dst[2]
    err = np.hypot(dst[0] - pb["x"][m[ok]], dst[1] - pb["y"][m[ok]])
    assert (err < 8.0).mean() > 0.5   # a wrong stage gives ~0 geometric inliers


def test_threaded_oracle_is_identical_to_single_thread(oracle):
    """The oracle's OpenMP loops (pyramid rows, detection rows within a level, keypoints, match
    queries) write disjoint outputs: 8 threads give the same keypoints, descriptors and matches as
    1, at both pyramid scales."""
    a = oracle_synth(640, 480, 40, 12)
    b = oracle_synth(640, 480, 40, 13)
    try:
        runs = {}
        for nt in (1, 8):
            oracle.set_threads(nt)
            r = []
            for scale, nl in ((1.2, 8), (2.0, 5)):
                pa, da, na = oracle.detect_and_compute(a, noctaves=nl, scale=scale, max_pts=700)
                pb, db, _ = oracle.detect_and_compute(b, noctaves=nl, scale=scale, max_pts=700)
                r.append((pa.tobytes(), da.tobytes(), na, oracle.match(da, db)))
            runs[nt] = r
        for (x, y) in zip(runs[1], runs[8]):
            assert x[0] == y[0] and x[1] == y[1] and x[2] == y[2]
            assert np.array_equal(x[3][0], y[3][0]) and np.array_equal(x[3][1], y[3][1])
    finally:
        oracle.set_threads(1)


def test_small_patch_pattern_is_the_references_libc_draw(oracle):
    """Patch sizes other than 31 use 512 random points, srand(0x34985739) and each coordinate
    rand() % p - -p / 2 in C (orbd.cu:359-369), i.e. rand() % p + p // 2: checked against libc's own
    generator, and WTA_K 3/4 tuples are made of those points."""
    libc = ctypes.CDLL("libc.so.6")
    for p in (25, 8, 30):
        libc.srand(0x34985739)
        want = [libc.rand() % p + p // 2 for _ in range(1024)]
        got = oracle.pattern(2, p)
        assert got.ravel().tolist() == want
        assert got.min() >= p // 2 and got.max() <= p - 1 + p // 2
        base = {tuple(v) for v in got}
        for k in (3, 4):
            tup = oracle.pattern(k, p)
            assert {tuple(v) for v in tup[: 128 * k]} <= base
    with pytest.raises(ValueError):
        oracle.pattern(2, 33)
