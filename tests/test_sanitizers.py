"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer, on the CPU:

- the C oracle (oracle/orb_oracle.c) driven through detection, description,
  matching and top-2 by oracle/sanitize_check.c (make -C oracle sanitize),
  with leak checking;
- the C-ABI library's host code (every .cpp and the host side of every .hip,
  make -C cuda-orb_amd sanitize) driven by tests/sanitize/capi_check.cpp
  through its host-only entry points and error paths, with leak checking.
The same driver's GPU half runs under the sanitizers in
test_gpu_sanitizers.py."""
import os
import pathlib
import subprocess

import pytest

from conftest import gpu_available

ROOT = pathlib.Path(__file__).resolve().parents[1]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")


def _run(cmd, **kw):
    res = subprocess.run(cmd, capture_output=True, text=True, env=ENV, **kw)
    print(res.stdout, res.stderr, flush=True)
    return res


def test_oracle_clean_under_asan_ubsan(tmp_path):
    assert _run(["make", "-s", "-C", str(ROOT / "oracle"), "sanitize"], timeout=300).returncode == 0
    res = _run([str(ROOT / "oracle" / "_build" / "sanitize_check")], timeout=300)
    assert res.returncode == 0
    assert "keypoints" in res.stdout and "ERROR" not in res.stderr


@pytest.mark.skipif(gpu_available(), reason="the GPU half runs in test_gpu_sanitizers.py")
def test_capi_host_paths_clean_under_asan_ubsan(tmp_path):
    assert _run(["make", "-s", "-j8", "-C", str(ROOT / "cuda-orb_amd"), "sanitize"], timeout=900).returncode == 0
    res = _run([str(ROOT / "cuda-orb_amd" / "_build" / "asan" / "capi_check"), str(tmp_path)], timeout=300)
    assert res.returncode == 0
    assert "host checks passed" in res.stdout and "ERROR" not in res.stderr
