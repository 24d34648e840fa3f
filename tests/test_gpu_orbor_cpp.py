"""The drop-in C++ API (include/orb.h, orb::Orbor) on the GPU: the demo program
cuda-orb_amd/orb_amd/orb_demo runs the reference's protocol (main.cpp:141-271)
and its final points and descriptors must equal the oracle's.  Inputs: the
reference's own sample images data/left.pgm and data/right.pgm (kept as
fixtures in tests/golden/) with the reference's demo parameters, and a seeded
synthetic pair with the bilinear pyramid."""
import pathlib
import subprocess

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs an MI355X")]

ROOT = pathlib.Path(__file__).resolve().parents[1]
DEMO = ROOT / "cuda-orb_amd" / "orb_amd" / "orb_demo"
GOLDEN = ROOT / "tests" / "golden"


def read_pgm(path):
    data = path.read_bytes()
    parts = data.split(maxsplit=4)
    w, h = int(parts[1]), int(parts[2])
    return np.frombuffer(parts[4][: w * h], np.uint8).reshape(h, w)


def describe_mismatch(got, want, field):
    bad = np.nonzero(got[field] != want[field])[0]
    rows = [(int(i), int(got["x"][i]), int(got["y"][i]), int(got["octave"][i]), int(want["octave"][i]),
             hex(int(got["score"].view(np.uint32)[i])), hex(int(want["score"].view(np.uint32)[i])))
            for i in bad[:8]]
    return f"{field}: {len(bad)} of {len(got)} differ; (i, x, y, oct, oct_want, score, score_want): {rows}"


def run_demo(tmp_path, img1, img2, extra):
    cmd = [str(DEMO), "0", img1, img2, str(tmp_path)] + extra
    print("running", " ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    print(res.stdout, res.stderr, flush=True)
    assert res.returncode == 0
    pts = [np.fromfile(tmp_path / f"points{i}.bin", dtype=np.int32).reshape(-1, 7) for i in (1, 2)]
    desc = [np.fromfile(tmp_path / f"desc{i}.bin", dtype=np.uint8).reshape(-1, 32) for i in (1, 2)]
    return pts, desc, res.stdout


def check(oracle, pts, desc, imgs, **kw):
    outs = []
    for p, d, img in zip(pts, desc, imgs):
        want_p, want_d, _ = oracle.detect_and_compute(img, **kw)
        got = p.copy().view(oracle.POINT_DTYPE).reshape(-1)
        assert len(got) == len(want_p) > 0
        for f in ("x", "y", "octave"):
            assert np.array_equal(got[f], want_p[f]), describe_mismatch(got, want_p, f)
        assert np.array_equal(got["score"].view(np.uint32), want_p["score"].view(np.uint32))
        assert np.array_equal(got["angle"].view(np.uint32), want_p["angle"].view(np.uint32))
        assert np.array_equal(d, want_d)
        outs.append((got, want_d))
    m, dist = oracle.match(outs[0][1], outs[1][1])
    assert np.array_equal(outs[0][0]["match"], m) and np.array_equal(outs[0][0]["distance"], dist)
    return m


def read_ppm(path):
    data = path.read_bytes()
    parts = data.split(maxsplit=4)
    assert parts[0] == b"P6" and parts[3] == b"255"
    w, h = int(parts[1]), int(parts[2])
    return np.frombuffer(parts[4][: w * h * 3], np.uint8).reshape(h, w, 3)


def test_reference_images_reference_parameters(tmp_path, oracle):
    left, right = GOLDEN / "left.pgm", GOLDEN / "right.pgm"
    pts, desc, out = run_demo(tmp_path, str(left), str(right), ["--repeats", "5", "--draw", "1"])
    assert "Time of detection and computation" in out
    img1, img2 = read_pgm(left), read_pgm(right)
    m = check(oracle, pts, desc, [img1, img2], noctaves=5, scale=2.0, wta_k=4, max_pts=10000)
    assert (m >= 0).sum() > 200
    # the drawings (main.cpp:27-99): the images in grey, a coloured ring around each keypoint,
    # a coloured line from each accepted match to its partner in the right half
    kp = read_ppm(tmp_path / "keypoints1.ppm")
    assert kp.shape == img1.shape + (3,)
    grey = (kp[..., 0] == kp[..., 1]) & (kp[..., 1] == kp[..., 2])
    assert (kp[grey][:, 0] == img1[grey]).all()
    p1 = pts[0]
    ring = [(p1[i, 1], p1[i, 0] + 1) for i in range(len(p1)) if 2 <= p1[i, 0] < img1.shape[1] - 2]
    assert np.mean([not grey[y, x] for y, x in ring]) > 0.5
    mt = read_ppm(tmp_path / "matches.ppm")
    assert mt.shape == (max(img1.shape[0], img2.shape[0]), img1.shape[1] + img2.shape[1], 3)
    left_half = mt[:, : img1.shape[1]]
    assert (~((left_half[..., 0] == left_half[..., 1]) & (left_half[..., 1] == left_half[..., 2]))).sum() > 1000


def test_synthetic_pair_bilinear_pyramid(tmp_path, oracle):
    import orb_amd
    pts, desc, _ = run_demo(tmp_path, "synth:640:480:0", "warp", ["--scale", "1.2", "--levels", "8", "--wta", "2",
                                                                   "--max-pts", "1000", "--repeats", "3"])
    a = orb_amd.synth_image(0, 640, 480, 40, pitch=640)
    b, _ = orb_amd.synth_warp(0, a, 640, 480)
    check(oracle, pts, desc, [a, b], noctaves=8, scale=1.2, wta_k=2, max_pts=1000)
