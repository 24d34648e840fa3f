"""Python binding of liborb_amd.so (the MI355X ORB library's C ABI, include/orb_capi.h).

The compute path is the HIP library; this module only passes device pointers
(torch tensors on ``cuda``) and sizes through ctypes, on torch's current
stream.  Loading fails loudly if the library has not been built.

The class :class:`Orbor` mirrors the reference's ``orb::Orbor`` (orb.h:11-76):
``init``, ``detect_and_compute`` (detectAndCompute), ``match``.
"""
import ctypes
import os
import pathlib

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = HERE / "liborb_amd.so"

ORB_OK = 0
ORB_MATCH_POPCOUNT = 0
ORB_MATCH_MFMA = 1
ORB_MATCH_HAMMING2 = 2
NOALLOC_MAX_QUERIES = 524288   # ORB_NOALLOC_MAX_QUERIES of orb_capi.h
HARRIS_SCORE = 0
FAST_SCORE = 1

# OrbPoint, orb_structures.h:8-17 (28 bytes)
POINT_DTYPE = np.dtype([("x", "<i4"), ("y", "<i4"), ("octave", "<i4"), ("score", "<f4"),
                        ("angle", "<f4"), ("match", "<i4"), ("distance", "<i4")])
assert POINT_DTYPE.itemsize == 28

# orb_cv_keypoint: the memory layout of OpenCV's cv::KeyPoint (28 bytes)
CV_KEYPOINT_DTYPE = np.dtype([("x", "<f4"), ("y", "<f4"), ("size", "<f4"), ("angle", "<f4"),
                              ("response", "<f4"), ("octave", "<i4"), ("class_id", "<i4")])
assert CV_KEYPOINT_DTYPE.itemsize == 28

EXPORTS = [
    "orb_ctx_create", "orb_ctx_destroy", "orb_set_stream", "orb_init", "orb_num_levels",
    "orb_detect_and_compute", "orb_detect_and_compute_batch", "orb_match", "orb_match_dcount",
    "orb_match_top2", "orb_match_top2_noalloc", "orb_match_batch", "orb_export_keypoints", "orb_sync", "orb_last_error", "orb_workspace_bytes", "orb_synth_image",
    "orb_synth_warp", "orb_debug_pattern", "orb_debug_sincos", "orb_debug_atan2", "orb_profile_enable", "orb_profile_read",
    "orb_tune_set", "orb_tune_get", "orb_debug_pyramid", "orb_capture_begin", "orb_capture_end",
    "orb_graph_launch", "orb_graph_destroy",
    "orb_lindex_create", "orb_lindex_destroy", "orb_lindex_build", "orb_lindex_bucket_info", "orb_lindex_train",
    "orb_lindex_logits", "orb_lindex_grads", "orb_lindex_num_params", "orb_lindex_get_weights",
    "orb_lindex_set_weights", "orb_lindex_query", "orb_lindex_probes", "orb_lindex_nprobe", "orb_lindex_save", "orb_lindex_load", "orb_lindex_set_probe",
    "orb_lindex_set_mass",
]

ORB_LINDEX_MLP = 0
ORB_LINDEX_CENTROIDS = 1

TUNE_PYRAMID = 0
PYRAMID_STRIP_HEIGHTS = (18, 26, 34, 42, 58)
NUM_STAGES = 8
STAGES = ["pyramid", "detect", "nms", "select", "emit", "describe", "match", "match_merge"]


class OrbError(RuntimeError):
    pass


class LindexParams(ctypes.Structure):
    _fields_ = [("nbuckets", ctypes.c_int), ("hidden", ctypes.c_int), ("nprobe", ctypes.c_int),
                ("fallback", ctypes.c_float), ("seed", ctypes.c_uint64), ("centroid_scale", ctypes.c_float)]


class Params(ctypes.Structure):
    _fields_ = [("noctaves", ctypes.c_int), ("edge_threshold", ctypes.c_int),
                ("wta_k", ctypes.c_int), ("score_type", ctypes.c_int),
                ("patch_size", ctypes.c_int), ("fast_threshold", ctypes.c_int),
                ("retain_topn", ctypes.c_int), ("max_pts", ctypes.c_int),
                ("scale_factor", ctypes.c_double)]


_lib = None


def lib():
    """Load liborb_amd.so; raises if it is missing (no fallback path exists)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise OrbError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(str(LIB_PATH))
        P, I, V = ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p
        L.orb_ctx_create.argtypes = [ctypes.POINTER(P), I]
        L.orb_ctx_destroy.argtypes = [P]
        L.orb_set_stream.argtypes = [P, V]
        L.orb_init.argtypes = [P, ctypes.POINTER(Params)]
        L.orb_num_levels.argtypes = [P, I, I]
        L.orb_detect_and_compute.argtypes = [P, V, I, I, I, V, V, V, I]
        L.orb_detect_and_compute_batch.argtypes = [P, I, V, V, V, V, V, V, V, I]
        L.orb_match.argtypes = [P, V, V, I, V, I]
        L.orb_match_dcount.argtypes = [P, V, V, V, I, V, V, I]
        L.orb_match_top2.argtypes = [P, V, I, V, I, I, V, V, V, V]
        L.orb_match_top2_noalloc.argtypes = [P, V, I, V, I, I, V, V, V, V]
        L.orb_export_keypoints.argtypes = [P, V, V, V, I, V, V, V]
        L.orb_match_batch.argtypes = [P, I, V, V, V, I, V, V, I]
        L.orb_sync.argtypes = [P]
        L.orb_last_error.argtypes = [P]
        L.orb_last_error.restype = ctypes.c_char_p
        L.orb_workspace_bytes.argtypes = [P]
        L.orb_workspace_bytes.restype = ctypes.c_size_t
        L.orb_synth_image.argtypes = [ctypes.c_uint64, I, I, I, I, V]
        L.orb_synth_warp.argtypes = [ctypes.c_uint64, V, I, I, I, V, V]
        L.orb_debug_pattern.argtypes = [I, I, V]
        L.orb_debug_sincos.argtypes = [P, V, I, V, V]
        L.orb_debug_atan2.argtypes = [P, V, I, V]
        L.orb_profile_enable.argtypes = [P, ctypes.c_uint]
        L.orb_profile_read.argtypes = [P, V, V]
        L.orb_tune_set.argtypes = [P, I, I]
        L.orb_tune_get.argtypes = [P, I, V, V, V]
        L.orb_debug_pyramid.argtypes = [P, I, I, V, ctypes.c_size_t, V]
        L.orb_capture_begin.argtypes = [P]
        L.orb_capture_end.argtypes = [P, ctypes.POINTER(P)]
        L.orb_graph_launch.argtypes = [P]
        L.orb_graph_destroy.argtypes = [P]
        L.orb_lindex_create.argtypes = [P, ctypes.POINTER(LindexParams), ctypes.POINTER(P)]
        L.orb_lindex_destroy.argtypes = [P]
        L.orb_lindex_build.argtypes = [P, V, I, I]
        L.orb_lindex_bucket_info.argtypes = [P, V, V, V]
        L.orb_lindex_train.argtypes = [P, V, V, I, I, I, ctypes.c_float, V]
        L.orb_lindex_logits.argtypes = [P, V, I, V]
        L.orb_lindex_grads.argtypes = [P, V, V, I, V, V]
        L.orb_lindex_num_params.argtypes = [P]
        L.orb_lindex_get_weights.argtypes = [P, V]
        L.orb_lindex_set_weights.argtypes = [P, V]
        L.orb_lindex_query.argtypes = [P, V, I, I, V, V, V, V, V]
        L.orb_lindex_probes.argtypes = [P, V, I, I, V, V]
        L.orb_lindex_nprobe.argtypes = [P]
        L.orb_lindex_save.argtypes = [P, ctypes.c_char_p]
        L.orb_lindex_load.argtypes = [P, ctypes.c_char_p]
        L.orb_lindex_set_probe.argtypes = [P, I, ctypes.c_float]
        L.orb_lindex_set_mass.argtypes = [P, ctypes.c_float, ctypes.c_float]
        _lib = L
    return _lib


def align_up(a, b=128):
    return a if a % b == 0 else a - a % b + b


def _np_ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def synth_image(seed, width, height, nshapes, pitch=None):
    """Seeded synthetic image A as an (height, pitch) uint8 array (host only)."""
    pitch = pitch or align_up(width)
    out = np.zeros((height, pitch), np.uint8)
    if lib().orb_synth_image(seed, width, height, pitch, nshapes, _np_ptr(out)) != ORB_OK:
        raise OrbError("orb_synth_image: bad arguments")
    return out


def synth_warp(seed, src, width, height):
    """Warp of a pitched image plus noise; returns (image, H 3x3 mapping A to B)."""
    src = np.ascontiguousarray(src)
    pitch = src.shape[1]
    out = np.zeros_like(src)
    H = np.zeros(9, np.float64)
    if lib().orb_synth_warp(seed, _np_ptr(src), width, height, pitch, _np_ptr(out), _np_ptr(H)) != ORB_OK:
        raise OrbError("orb_synth_warp: bad arguments")
    return out, H.reshape(3, 3)


def _dptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


class Context:
    """An orb_ctx on one device, bound to torch's current stream at each call."""

    def __init__(self, device=0):
        import torch
        self._torch = torch
        self.device = device
        h = ctypes.c_void_p()
        if lib().orb_ctx_create(ctypes.byref(h), device) != ORB_OK:
            raise OrbError(lib().orb_last_error(None).decode())
        self.h = h
        self.params = None

    def close(self):
        if self.h:
            lib().orb_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != ORB_OK:
            raise OrbError(f"{what}: {lib().orb_last_error(self.h).decode()} (code {rc})")

    def _bind_stream(self):
        s = self._torch.cuda.current_stream(self.device).cuda_stream
        self._check(lib().orb_set_stream(self.h, ctypes.c_void_p(s)), "orb_set_stream")

    def init(self, noctaves=5, edge_threshold=31, wta_k=2, score_type=HARRIS_SCORE, patch_size=31,
             fast_threshold=20, retain_topn=-1, max_pts=1000, scale_factor=2.0):
        self._bind_stream()
        p = Params(noctaves, edge_threshold, wta_k, score_type, patch_size, fast_threshold,
                   retain_topn, max_pts, scale_factor)
        self._check(lib().orb_init(self.h, ctypes.byref(p)), "orb_init")
        self.params = p
        return self

    def num_levels(self, width, height):
        return lib().orb_num_levels(self.h, width, height)

    def alloc_points(self, n, frames=None):
        shape = (n, 7) if frames is None else (frames, n, 7)
        return self._torch.empty(shape, dtype=self._torch.int32, device=f"cuda:{self.device}")

    def alloc_desc(self, n, frames=None):
        shape = (n, 32) if frames is None else (frames, n, 32)
        return self._torch.empty(shape, dtype=self._torch.uint8, device=f"cuda:{self.device}")

    def detect_and_compute(self, img, width, height, pts, desc, count, compute_desc=True):
        """img: device uint8 (height, pitch) tensor; pts (max_pts,7) int32; desc (max_pts,32) uint8;
        count: device int32 tensor of >= 1 element.  Asynchronous."""
        self._bind_stream()
        pitch = img.shape[1] if img.dim() == 2 else width
        self._check(lib().orb_detect_and_compute(
            self.h, _dptr(img), width, height, pitch, _dptr(pts), _dptr(desc) if compute_desc else None,
            _dptr(count), 1 if compute_desc else 0), "orb_detect_and_compute")

    def detect_and_compute_batch(self, imgs, widths, heights, pts_list, desc_list, counts,
                                 compute_desc=True):
        self._bind_stream()
        n = len(imgs)
        arr = lambda vals, t: (t * n)(*vals)
        ptrs = arr([im.data_ptr() for im in imgs], ctypes.c_void_p)
        ws = arr(widths, ctypes.c_int)
        hs = arr(heights, ctypes.c_int)
        ps = arr([im.shape[1] for im in imgs], ctypes.c_int)
        pp = arr([p.data_ptr() for p in pts_list], ctypes.c_void_p)
        dp = arr([d.data_ptr() for d in desc_list], ctypes.c_void_p) if compute_desc else None
        self._check(lib().orb_detect_and_compute_batch(
            self.h, n, ptrs, ws, hs, ps, pp, dp, _dptr(counts), 1 if compute_desc else 0),
            "orb_detect_and_compute_batch")

    def match(self, pts1, desc1, n1, desc2, n2):
        self._bind_stream()
        self._check(lib().orb_match(self.h, _dptr(pts1), _dptr(desc1), n1, _dptr(desc2), n2), "orb_match")

    def match_dcount(self, pts1, desc1, d_n1, cap1, desc2, d_n2, cap2):
        self._bind_stream()
        self._check(lib().orb_match_dcount(self.h, _dptr(pts1), _dptr(desc1), _dptr(d_n1), cap1,
                                           _dptr(desc2), _dptr(d_n2), cap2), "orb_match_dcount")

    def match_batch(self, pts1_list, desc1_list, d_n1, cap1, desc2_list, d_n2, cap2):
        """orb_match_batch: pair i matches desc1_list[i] (count d_n1[i]) against desc2_list[i]
        (count d_n2[i]) into pts1_list[i], all pairs in one launch.  d_n1, d_n2: device int32
        tensors of len(pts1_list) counts.  Asynchronous."""
        self._bind_stream()
        n = len(pts1_list)
        arr = lambda vals: (ctypes.c_void_p * n)(*vals)
        self._check(lib().orb_match_batch(self.h, n, arr([p.data_ptr() for p in pts1_list]),
                                          arr([d.data_ptr() for d in desc1_list]), _dptr(d_n1), cap1,
                                          arr([d.data_ptr() for d in desc2_list]), _dptr(d_n2), cap2),
                    "orb_match_batch")

    def match_top2(self, q, db, method=ORB_MATCH_POPCOUNT, out=None):
        """Exact top-2 of each query row in db; returns (idx1, d1, idx2, d2) int32 device tensors."""
        self._bind_stream()
        torch = self._torch
        nq = q.shape[0]
        if out is None:
            out = [torch.empty(nq, dtype=torch.int32, device=q.device) for _ in range(4)]
        self._check(lib().orb_match_top2(self.h, _dptr(q), nq, _dptr(db), db.shape[0], method,
                                         *[_dptr(o) for o in out]), "orb_match_top2")
        return out

    def match_top2_noalloc(self, q, db, out, method=ORB_MATCH_POPCOUNT):
        """orb_match_top2_noalloc into the four int32 device tensors of out (idx1, d1, idx2, d2):
        the library's fixed scratch array instead of an allocated workspace."""
        self._bind_stream()
        self._check(lib().orb_match_top2_noalloc(self.h, _dptr(q), q.shape[0], _dptr(db), db.shape[0], method,
                                                 *[_dptr(o) for o in out]), "orb_match_top2_noalloc")
        return out

    def export_keypoints(self, pts, desc=None, count=None, kp=None, desc_out=None, nout=None):
        """OpenCV-layout export (orb_export_keypoints) of the points in pts ((cap, 7) int32, device),
        count a device int32 tensor (None: all cap rows), honouring the retain_topn given to init.
        Returns (kp (cap, 7) int32 tensor holding orb_cv_keypoint rows, desc_out or None, nout
        device int32 tensor of 1 element).  Asynchronous."""
        self._bind_stream()
        torch = self._torch
        cap = pts.shape[0]
        if kp is None:
            kp = torch.empty((cap, 7), dtype=torch.int32, device=pts.device)
        if desc is not None and desc_out is None:
            desc_out = torch.empty((cap, 32), dtype=torch.uint8, device=pts.device)
        if nout is None:
            nout = torch.zeros(1, dtype=torch.int32, device=pts.device)
        self._check(lib().orb_export_keypoints(self.h, _dptr(pts), _dptr(desc), _dptr(count), cap, _dptr(kp),
                                               _dptr(desc_out) if desc is not None else None, _dptr(nout)),
                    "orb_export_keypoints")
        return kp, desc_out, nout

    def sync(self):
        self._check(lib().orb_sync(self.h), "orb_sync")

    def profile_enable(self, stage_mask):
        """Record HIP events around every launch of the stages in stage_mask (bit i = STAGES[i])."""
        self._check(lib().orb_profile_enable(self.h, stage_mask), "orb_profile_enable")

    def profile_read(self):
        """Synchronise and return {stage: (summed ms, launches)} since the last read."""
        ms = np.zeros(NUM_STAGES, np.float64)
        cnt = np.zeros(NUM_STAGES, np.int32)
        self._check(lib().orb_profile_read(self.h, _np_ptr(ms), _np_ptr(cnt)), "orb_profile_read")
        return {STAGES[i]: (float(ms[i]), int(cnt[i])) for i in range(NUM_STAGES)}

    def capture_begin(self):
        """Record the following calls into a HIP graph (orb_capture_begin)."""
        self._bind_stream()
        self._check(lib().orb_capture_begin(self.h), "orb_capture_begin")

    def capture_end(self):
        """End the recording; returns a Graph whose launch() replays it on the context's stream."""
        g = ctypes.c_void_p()
        self._check(lib().orb_capture_end(self.h, ctypes.byref(g)), "orb_capture_end")
        return Graph(self, g)

    def tune_set(self, stage, variant):
        """Force tile-shape variant `variant` of a tuned stage (-1: autotune)."""
        self._check(lib().orb_tune_set(self.h, stage, variant), "orb_tune_set")

    def tune_get(self, stage=0):
        """(variant in use or -1, measured mean ms per variant) of a tuned stage."""
        v = ctypes.c_int(-1)
        n = ctypes.c_int(0)
        ms = np.zeros(16, np.float32)
        self._check(lib().orb_tune_get(self.h, stage, ctypes.byref(v), ctypes.byref(n), _np_ptr(ms)),
                    "orb_tune_get")
        return v.value, [float(x) for x in ms[: n.value]]

    def workspace_bytes(self):
        return lib().orb_workspace_bytes(self.h)

    def debug_sincos(self, angles):
        self._bind_stream()
        s = self._torch.empty_like(angles)
        c = self._torch.empty_like(angles)
        self._check(lib().orb_debug_sincos(self.h, _dptr(angles), angles.numel(), _dptr(s), _dptr(c)),
                    "orb_debug_sincos")
        return s, c

    def debug_pyramid(self, frame=0, which=0):
        """uint8 tensor: the frame's flat raw (0) or blurred (1) pyramid of the last detect call."""
        self._bind_stream()
        n = ctypes.c_size_t(0)
        self._check(lib().orb_debug_pyramid(self.h, frame, which, None, 0, ctypes.byref(n)), "orb_debug_pyramid")
        out = self._torch.empty(n.value, dtype=self._torch.uint8, device="cuda")
        self._check(lib().orb_debug_pyramid(self.h, frame, which, _dptr(out), n.value, ctypes.byref(n)),
                    "orb_debug_pyramid")
        return out

    def debug_atan2(self, yx):
        self._bind_stream()
        out = self._torch.empty(yx.shape[0], dtype=self._torch.float32, device=yx.device)
        self._check(lib().orb_debug_atan2(self.h, _dptr(yx), yx.shape[0], _dptr(out)), "orb_debug_atan2")
        return out


class Graph:
    """A recorded sequence of context calls (orb_graph)."""

    def __init__(self, ctx, h):
        self.ctx = ctx
        self.h = h

    def launch(self):
        self.ctx._bind_stream()
        self.ctx._check(lib().orb_graph_launch(self.h), "orb_graph_launch")

    def close(self):
        if self.h:
            lib().orb_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class LearnedIndex:
    """Learned bucket matcher (orb_lindex_*): k-modes buckets over a descriptor database, a BF16 MLP
    that routes queries to buckets, bucketed exact search with exact fallback."""

    def __init__(self, ctx, nbuckets=1024, hidden=512, nprobe=8, fallback=0.3, seed=4, centroid_scale=0.0):
        self.ctx = ctx
        torch = ctx._torch
        self._torch = torch
        self.prm = LindexParams(nbuckets, hidden, nprobe, fallback, seed, centroid_scale)
        h = ctypes.c_void_p()
        ctx._bind_stream()
        ctx._check(lib().orb_lindex_create(ctx.h, ctypes.byref(self.prm), ctypes.byref(h)), "orb_lindex_create")
        self.h = h
        self.nbuckets, self.hidden = nbuckets, hidden

    def close(self):
        if self.h:
            lib().orb_lindex_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _c(self, rc, what):
        self.ctx._check(rc, what)

    def build(self, db, iters=8):
        self.ctx._bind_stream()
        self._c(lib().orb_lindex_build(self.h, _dptr(db), db.shape[0], iters), "orb_lindex_build")
        self.ndb = db.shape[0]

    def bucket_info(self):
        torch = self._torch
        off = np.zeros(self.nbuckets + 1, np.int32)
        cen = np.zeros((self.nbuckets, 32), np.uint8)
        bof = torch.empty(self.ndb, dtype=torch.int32, device=f"cuda:{self.ctx.device}")
        self._c(lib().orb_lindex_bucket_info(self.h, _np_ptr(off), _np_ptr(cen), _dptr(bof)), "orb_lindex_bucket_info")
        return off, cen, bof

    def train(self, q, target, epochs=3, batch=4096, lr=1e-3):
        self.ctx._bind_stream()
        losses = np.zeros(epochs, np.float32)
        self._c(lib().orb_lindex_train(self.h, _dptr(q), _dptr(target), q.shape[0], epochs, batch, lr,
                                       _np_ptr(losses)), "orb_lindex_train")
        return losses

    def logits(self, q):
        self.ctx._bind_stream()
        out = self._torch.empty((q.shape[0], self.nbuckets), dtype=self._torch.float32, device=q.device)
        self._c(lib().orb_lindex_logits(self.h, _dptr(q), q.shape[0], _dptr(out)), "orb_lindex_logits")
        return out

    def grads(self, q, labels):
        self.ctx._bind_stream()
        g = np.zeros(self.num_params(), np.float32)
        loss = ctypes.c_float()
        self._c(lib().orb_lindex_grads(self.h, _dptr(q), _dptr(labels), q.shape[0], _np_ptr(g), ctypes.byref(loss)),
                "orb_lindex_grads")
        return g, loss.value

    def num_params(self):
        return lib().orb_lindex_num_params(self.h)

    def get_weights(self):
        w = np.zeros(self.num_params(), np.float32)
        self._c(lib().orb_lindex_get_weights(self.h, _np_ptr(w)), "orb_lindex_get_weights")
        return w

    def set_weights(self, w):
        w = np.ascontiguousarray(w, np.float32)
        self._c(lib().orb_lindex_set_weights(self.h, _np_ptr(w)), "orb_lindex_set_weights")

    def split(self, flat):
        """Flat parameter vector -> dict W1 (256,H), b1, W2 (H,H), b2, W3 (H,C), b3."""
        H, C = self.hidden, self.nbuckets
        shapes = [("W1", (256, H)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (H, C)), ("b3", (C,))]
        out, o = {}, 0
        for name, shp in shapes:
            n = int(np.prod(shp))
            out[name] = flat[o:o + n].reshape(shp)
            o += n
        return out

    def query(self, q, mode=ORB_LINDEX_MLP, out=None):
        self.ctx._bind_stream()
        torch = self._torch
        if out is None:
            out = [torch.empty(q.shape[0], dtype=torch.int32, device=q.device) for _ in range(4)]
        nfb = ctypes.c_int(0)
        self._c(lib().orb_lindex_query(self.h, _dptr(q), q.shape[0], mode, *[_dptr(o) for o in out],
                                       ctypes.byref(nfb)), "orb_lindex_query")
        return out, nfb.value

    def probes(self, q, mode=ORB_LINDEX_MLP):
        """(probes, pmax): the nprobe buckets each query probes, best first (ties to the lower
        bucket id), and its max softmax probability (1 in centroid mode)."""
        self.ctx._bind_stream()
        torch = self._torch
        pr = torch.empty((q.shape[0], lib().orb_lindex_nprobe(self.h)), dtype=torch.int32, device=q.device)
        pm = torch.empty(q.shape[0], dtype=torch.float32, device=q.device)
        self._c(lib().orb_lindex_probes(self.h, _dptr(q), q.shape[0], mode, _dptr(pr), _dptr(pm)),
                "orb_lindex_probes")
        return pr, pm

    def set_probe(self, nprobe, fallback):
        self._c(lib().orb_lindex_set_probe(self.h, nprobe, fallback), "orb_lindex_set_probe")

    def set_mass(self, mass, centroid_temperature=4.0):
        """Probe only the best-ranked prefix of the nprobe buckets whose probability mass reaches
        `mass` (1 = every one); centroid-mode weights exp(-distance / centroid_temperature)."""
        self._c(lib().orb_lindex_set_mass(self.h, mass, centroid_temperature), "orb_lindex_set_mass")

    def save(self, path):
        self._c(lib().orb_lindex_save(self.h, str(path).encode()), "orb_lindex_save")

    def load(self, path):
        self._c(lib().orb_lindex_load(self.h, str(path).encode()), "orb_lindex_load")


def points_to_numpy(pts_tensor, n):
    """First n OrbPoints of an (N,7) int32 device tensor as a structured array."""
    a = pts_tensor[:n].contiguous().cpu().numpy()
    return a.view(POINT_DTYPE).reshape(-1)


def as_cv_keypoints(kp):
    """(n, 7) int32 tensor/array of orb_cv_keypoint rows -> numpy structured array (CV_KEYPOINT_DTYPE)."""
    a = kp.cpu().numpy() if hasattr(kp, "cpu") else np.asarray(kp)
    return np.ascontiguousarray(a, dtype=np.int32).view(CV_KEYPOINT_DTYPE).reshape(-1)


class Orbor:
    """Mirror of the reference's orb::Orbor (orb.h:11-76) for Python callers."""

    def __init__(self, device=0):
        self.ctx = Context(device)
        self.max_pts = 1000
        self._pts = self._desc = self._count = None

    def init(self, noctaves=5, edge_threshold=31, wta_k=2, score_type=HARRIS_SCORE, patch_size=31,
             fast_threshold=20, retain_topn=-1, max_pts=1000, scale_factor=2.0):
        self.ctx.init(noctaves, edge_threshold, wta_k, score_type, patch_size, fast_threshold,
                      retain_topn, max_pts, scale_factor)
        self.max_pts = max_pts
        return self

    def detect_and_compute(self, img, width, height, compute_desc=True):
        """img: device (height, pitch) uint8 tensor.  Returns (points ndarray, desc device tensor)."""
        torch = self.ctx._torch
        pts = self.ctx.alloc_points(self.max_pts)
        desc = self.ctx.alloc_desc(self.max_pts)
        count = torch.zeros(1, dtype=torch.int32, device=img.device)
        self.ctx.detect_and_compute(img, width, height, pts, desc, count, compute_desc)
        n = int(count.item())
        return points_to_numpy(pts, n), desc[:n], pts

    def match(self, pts1_dev, n1, desc1, desc2, n2):
        self.ctx.match(pts1_dev, desc1, n1, desc2, n2)
        return points_to_numpy(pts1_dev, n1)
