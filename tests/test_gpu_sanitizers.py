"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer with the
GPU present (device code unchanged; the sanitizers instrument host code
only).  Both executables are built by `make -C cuda-orb_amd sanitize`:

- tests/sanitize/capi_check.cpp drives every C-ABI entry point family (batch
  and single detection, reference-rule and top-2 matching, the
  allocation-free matcher, export, profiling, tuning, graph replay, the
  learned index with save/load) and checks the results that must agree;
- the demo (orb::Orbor, the reference's protocol) runs on the reference's
  sample images with drawing on, so the C++ API and the picture writer run
  instrumented too."""
import os
import pathlib
import subprocess

import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs an MI355X")]

ROOT = pathlib.Path(__file__).resolve().parents[1]
ASAN = ROOT / "cuda-orb_amd" / "_build" / "asan"
GOLDEN = ROOT / "tests" / "golden"
# leaks are not checked here: the HIP runtime keeps allocations until exit
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0", UBSAN_OPTIONS="print_stacktrace=1")


def _run(cmd):
    print("running", " ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True, env=ENV, timeout=240)
    print(res.stdout[-4000:], res.stderr[-4000:], flush=True)
    return res


def test_capi_check_clean_under_asan_ubsan(tmp_path):
    res = _run([str(ASAN / "capi_check"), str(tmp_path)])
    assert res.returncode == 0
    assert "checks passed (host and GPU)" in res.stdout and "ERROR" not in res.stderr


def test_demo_clean_under_asan_ubsan(tmp_path):
    res = _run([str(ASAN / "orb_demo"), "0", str(GOLDEN / "left.pgm"), str(GOLDEN / "right.pgm"), str(tmp_path),
                "--repeats", "3", "--draw", "1"])
    assert res.returncode == 0 and "ERROR" not in res.stderr
    for name in ("keypoints1.ppm", "keypoints2.ppm", "matches.ppm", "points1.bin", "desc2.bin"):
        assert (tmp_path / name).stat().st_size > 0, name
