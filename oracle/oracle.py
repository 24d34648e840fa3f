"""ctypes view of the CPU oracle (oracle/orb_oracle.c) -- TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, as the checker.  The product package never does.
"""
import ctypes
import pathlib
import subprocess

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "liborb_oracle.so"

POINT_DTYPE = np.dtype([("x", "<i4"), ("y", "<i4"), ("octave", "<i4"), ("score", "<f4"),
                        ("angle", "<f4"), ("match", "<i4"), ("distance", "<i4")])
assert POINT_DTYPE.itemsize == 28

MAX_LEVELS = 8


class Geom(ctypes.Structure):
    _fields_ = [("nlevels", ctypes.c_int), ("ref_mode", ctypes.c_int), ("scale", ctypes.c_double),
                ("w", ctypes.c_int * MAX_LEVELS), ("h", ctypes.c_int * MAX_LEVELS),
                ("p", ctypes.c_int * MAX_LEVELS), ("off", ctypes.c_long * MAX_LEVELS),
                ("total", ctypes.c_long), ("radius", ctypes.c_int * MAX_LEVELS),
                ("sc", ctypes.c_double * MAX_LEVELS)]


class Params(ctypes.Structure):
    _fields_ = [("noctaves", ctypes.c_int), ("edge_threshold", ctypes.c_int),
                ("wta_k", ctypes.c_int), ("score_type", ctypes.c_int),
                ("patch_size", ctypes.c_int), ("fast_threshold", ctypes.c_int),
                ("max_pts", ctypes.c_int), ("scale", ctypes.c_double)]


def build():
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = ctypes.CDLL(str(LIB_PATH))
        P = ctypes.c_void_p
        I = ctypes.c_int
        L.orc_fast_atan2.restype = ctypes.c_float
        L.orc_fast_atan2.argtypes = [ctypes.c_float, ctypes.c_float]
        L.orc_sincosf.argtypes = [ctypes.c_float, P, P]
        L.orc_harris_scale.restype = ctypes.c_float
        L.orc_harris.restype = ctypes.c_float
        L.orc_harris.argtypes = [P, ctypes.c_long, ctypes.c_long, I]
        L.orc_fast_score.restype = ctypes.c_float
        L.orc_fast_score.argtypes = [P, ctypes.c_long, ctypes.c_long, I, I]
        L.orc_fast_corner.argtypes = [P, ctypes.c_long, ctypes.c_long, I, P, I]
        L.orc_geometry.argtypes = [I, I, I, ctypes.c_double, ctypes.POINTER(Geom)]
        L.orc_level_to_l0.argtypes = [ctypes.POINTER(Geom), I, I]
        L.orc_l0_to_level.argtypes = [ctypes.POINTER(Geom), I, I]
        L.orc_pyramid.argtypes = [ctypes.POINTER(Geom), P, P, P]
        L.orc_blur_level.argtypes = [ctypes.POINTER(Geom), I, P, P]
        L.orc_detect_map.argtypes = [ctypes.POINTER(Geom), P, I, I, P, P]
        L.orc_nms.argtypes = [ctypes.POINTER(Geom), P, P, I, I, P, P]
        L.orc_angles.argtypes = [ctypes.POINTER(Geom), P, P, I, I]
        L.orc_describe.argtypes = [ctypes.POINTER(Geom), P, P, I, I, P]
        L.orc_detect_and_compute.argtypes = [ctypes.POINTER(Params), P, I, I, P, P, P, P]
        L.orc_match.argtypes = [P, I, P, I, P, P]
        L.orc_match_top2.argtypes = [P, I, P, I, I, P, P, P, P]
        L.orc_match_top2_norm.argtypes = [P, I, P, I, I, I, P, P, P, P]
        L.orc_pattern.argtypes = [I, P]
        L.orc_pattern_patch.argtypes = [I, I, P]
        L.orc_glibc_srand.argtypes = [ctypes.c_uint]
        L.orc_umax.argtypes = [I, P]
        L.orc_gauss.argtypes = [P]
        L.orc_fast_lut.argtypes = [I, P]
        L.orc_set_threads.argtypes = [I]
        _lib = L
    return _lib


def ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def set_threads(n):
    """Threads of the oracle's parallel loops (1 = scalar); outputs are identical for any n."""
    lib().orc_set_threads(int(n))


def align_up(a, b=128):
    return a if a % b == 0 else a - a % b + b


def geometry(w, h, noctaves=8, scale=1.2):
    g = Geom()
    if lib().orc_geometry(w, h, noctaves, scale, ctypes.byref(g)) != 0:
        raise ValueError("invalid geometry")
    return g


def pitched(img):
    """Copy an (h, w) uint8 image into an (h, alignUp(w,128)) zero-padded buffer."""
    h, w = img.shape
    out = np.zeros((h, align_up(w)), np.uint8)
    out[:, :w] = img
    return out


def pyramid(g, img_pitched):
    raw = np.zeros(g.total, np.uint8)
    blur = np.zeros(g.total, np.uint8)
    lib().orc_pyramid(ctypes.byref(g), ptr(np.ascontiguousarray(img_pitched)), ptr(raw), ptr(blur))
    return raw, blur


def detect_map(g, raw, threshold=20, score_type=0):
    n = g.h[0] * g.p[0]
    vmap = np.zeros(n, np.float32)
    lmap = np.zeros(n, np.int32)
    lib().orc_detect_map(ctypes.byref(g), ptr(raw), threshold, score_type, ptr(vmap), ptr(lmap))
    return vmap, lmap


def nms(g, vmap, lmap, border=31, max_pts=1000):
    pts = np.zeros(max(max_pts, 1), POINT_DTYPE)
    nsurv = ctypes.c_int(0)
    n = lib().orc_nms(ctypes.byref(g), ptr(vmap), ptr(lmap), border, max_pts, ptr(pts),
                      ctypes.byref(nsurv))
    return pts[:n].copy(), nsurv.value


def angles(g, raw, pts, patch_size=31):
    pts = np.ascontiguousarray(pts.copy())
    lib().orc_angles(ctypes.byref(g), ptr(raw), ptr(pts), len(pts), patch_size)
    return pts


def describe(g, blur, pts, wta_k=2):
    pts = np.ascontiguousarray(pts)
    desc = np.zeros((len(pts), 32), np.uint8)
    lib().orc_describe(ctypes.byref(g), ptr(blur), ptr(pts), len(pts), wta_k, ptr(desc))
    return desc


def detect_and_compute(img, noctaves=8, scale=1.2, edge_threshold=31, wta_k=2, score_type=0,
                       patch_size=31, fast_threshold=20, max_pts=1000):
    """img: (h, w) uint8 (unpitched) or (h, alignUp(w,128)) pitched with explicit w."""
    h, w = img.shape
    buf = pitched(img)
    prm = Params(noctaves, edge_threshold, wta_k, score_type, patch_size, fast_threshold,
                 max_pts, scale)
    pts = np.zeros(max(max_pts, 1), POINT_DTYPE)
    desc = np.zeros((max(max_pts, 1), 32), np.uint8)
    n = ctypes.c_int(0)
    ns = ctypes.c_int(0)
    rc = lib().orc_detect_and_compute(ctypes.byref(prm), ptr(buf), w, h, ptr(pts), ptr(desc),
                                      ctypes.byref(n), ctypes.byref(ns))
    if rc != 0:
        raise ValueError("oracle rejected the parameters")
    return pts[:n.value].copy(), desc[:n.value].copy(), ns.value


def match(desc1, desc2):
    d1 = np.ascontiguousarray(desc1, np.uint8)
    d2 = np.ascontiguousarray(desc2, np.uint8)
    m = np.zeros(len(d1), np.int32)
    d = np.zeros(len(d1), np.int32)
    lib().orc_match(ptr(d1), len(d1), ptr(d2), len(d2), ptr(m), ptr(d))
    return m, d


def match_top2(q, db, nthreads=1, norm=0):
    """Exact top-2 (idx1, d1, idx2, d2); norm 0 Hamming, 1 NORM_HAMMING2 (orbd.cu:1471-1559 generalised)."""
    q = np.ascontiguousarray(q, np.uint8)
    db = np.ascontiguousarray(db, np.uint8)
    out = [np.zeros(len(q), np.int32) for _ in range(4)]
    lib().orc_match_top2_norm(ptr(q), len(q), ptr(db), len(db), nthreads, norm, *[ptr(o) for o in out])
    return tuple(out)  # idx1, d1, idx2, d2


CV_KEYPOINT_DTYPE = np.dtype([("x", "<f4"), ("y", "<f4"), ("size", "<f4"), ("angle", "<f4"),
                              ("response", "<f4"), ("octave", "<i4"), ("class_id", "<i4")])


def export_cv(pts, desc, retain_topn, scale, patch_size=31):
    """OpenCV-layout export of points in raster order (the library's orb_export_keypoints).

    Kept: all points, or with 0 < retain_topn < n the retain_topn best by score (ties to the
    earlier point, the rule of the max_pts cut), in their original order.  Per point of level l:
    the level coordinate is recovered from the level-0 one (x >> l at scale 2, else
    floor(x / s^l + 0.5) with s^l the running product detection uses), pt = coordinate *
    float32(pow(scale, l)) in float32, size = patch_size * that, angle = radians * 180/pi in
    float32 moved into [0, 360), response = score, octave = l, class_id = -1."""
    n = len(pts)
    keep = np.ones(n, bool)
    if 0 < retain_topn < n:
        order = np.lexsort((np.arange(n), -pts["score"].astype(np.float64)))
        keep[:] = False
        keep[order[:retain_topn]] = True
    idx = np.nonzero(keep)[0]
    out = np.zeros(len(idx), CV_KEYPOINT_DTYPE)
    ref = scale == 2.0
    for o, i in enumerate(idx):
        l = int(pts["octave"][i])
        x, y = int(pts["x"][i]), int(pts["y"][i])
        if ref:
            lx, ly = x >> l, y >> l
        else:
            sc = 1.0
            for _ in range(l):
                sc *= scale
            lx, ly = int(np.floor(x / sc + 0.5)), int(np.floor(y / sc + 0.5))
        lsc = np.float32(scale ** l)
        out["x"][o] = np.float32(lx) * lsc
        out["y"][o] = np.float32(ly) * lsc
        out["size"][o] = np.float32(patch_size) * lsc
        deg = np.float32(pts["angle"][i]) * np.float32(57.2957795130823208768)
        if deg < 0:
            deg = deg + np.float32(360.0)
        if deg >= 360:
            deg = deg - np.float32(360.0)
        out["angle"][o] = deg
        out["response"][o] = pts["score"][i]
        out["octave"][o] = l
        out["class_id"][o] = -1
    d = None if desc is None else np.ascontiguousarray(desc[idx])
    return out, d


def fast_atan2(y, x):
    return lib().orc_fast_atan2(y, x)


def sincosf(a):
    s = ctypes.c_float()
    c = ctypes.c_float()
    lib().orc_sincosf(a, ctypes.byref(s), ctypes.byref(c))
    return s.value, c.value


def pattern(wta_k=2, patch_size=31):
    out = np.zeros(1024, np.int32)
    if lib().orc_pattern_patch(patch_size, wta_k, ptr(out)) != 0:
        raise ValueError("wta_k / patch_size")
    return out.reshape(512, 2)


def glibc_rand_seq(seed, n):
    lib().orc_glibc_srand(seed)
    return [lib().orc_glibc_rand() for _ in range(n)]


def umax(patch_size=31):
    out = np.zeros(patch_size // 2 + 2, np.int32)
    lib().orc_umax(patch_size, ptr(out))
    return out


def gauss():
    out = np.zeros(4, np.float32)
    lib().orc_gauss(ptr(out))
    return out


def fast_lut(t):
    out = np.zeros(512, np.uint8)
    lib().orc_fast_lut(t, ptr(out))
    return out
