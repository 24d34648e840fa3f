"""CPU checks of the C-ABI library: it loads, exports every symbol include/*.h
declares, its host-only synthetic generator is deterministic, and its error
paths report instead of crashing when no GPU is present."""
import ctypes
import hashlib
import json
import pathlib
import re

import numpy as np
import pytest

from conftest import gpu_available

ROOT = pathlib.Path(__file__).resolve().parents[1]
GOLDEN = ROOT / "tests" / "golden"


def declared_functions():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = h.read_text()
        for m in re.finditer(r"^\s*(?:int|size_t|const char\*)\s+(orb_\w+)\s*\(", text, re.M):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    import orb_amd
    lib = orb_amd.lib()
    names = declared_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) <= set(orb_amd.EXPORTS) | set(names)


def test_orbor_cpp_symbols_exported():
    import subprocess
    out = subprocess.run(["nm", "-DC", str(ROOT / "cuda-orb_amd/orb_amd/liborb_amd.so")],
                         capture_output=True, text=True, check=True).stdout
    for sym in ("orb::Orbor::init", "orb::Orbor::detectAndCompute", "orb::Orbor::match",
                "orb::Orbor::initOrbData", "orb::Orbor::freeOrbData"):
        assert sym in out, sym


def test_synthetic_images_match_committed_hashes():
    import orb_amd
    fixtures = json.loads((GOLDEN / "synth_sha256.json").read_text())
    for case in fixtures:
        a = orb_amd.synth_image(case["seed"], case["w"], case["h"], case["nshapes"])
        b, H = orb_amd.synth_warp(case["seed"], a, case["w"], case["h"])
        assert hashlib.sha256(a.tobytes()).hexdigest() == case["a"]
        assert hashlib.sha256(b.tobytes()).hexdigest() == case["b"]
        assert np.allclose(H.ravel(), case["H"], rtol=0, atol=1e-12)


def test_synth_rejects_bad_arguments():
    import orb_amd
    lib = orb_amd.lib()
    buf = np.zeros(16, np.uint8)
    assert lib.orb_synth_image(0, 0, 10, 16, 1, orb_amd._np_ptr(buf)) != 0
    assert lib.orb_synth_image(0, 20, 10, 16, 1, orb_amd._np_ptr(buf)) != 0   # pitch < width
    assert lib.orb_synth_image(0, 8, 1, 16, 1, None) != 0


def test_null_context_errors_are_reported():
    import orb_amd
    lib = orb_amd.lib()
    assert lib.orb_init(None, None) == -1
    assert lib.orb_sync(None) == -1
    assert lib.orb_match(None, None, None, 0, None, 0) == -1
    assert lib.orb_workspace_bytes(None) == 0
    assert isinstance(lib.orb_last_error(None), bytes)


@pytest.mark.skipif(gpu_available(), reason="checks the no-GPU error path")
def test_context_creation_without_gpu_fails_cleanly():
    import orb_amd
    lib = orb_amd.lib()
    h = ctypes.c_void_p()
    rc = lib.orb_ctx_create(ctypes.byref(h), 0)
    assert rc != 0 and not h.value
    assert lib.orb_last_error(None)


def _pattern(lib, wta_k, patch_size=31):
    import orb_amd
    xy = np.zeros(1024, np.int32)
    assert lib.orb_debug_pattern(patch_size, wta_k, orb_amd._np_ptr(xy)) == 0
    return xy.reshape(512, 2)


def test_product_patterns_equal_the_oracle(oracle):
    """The sampling points orb_init uploads (WTA_K 2: the reference table; 3/4: tuples drawn by the
    library's own TYPE_3 generator) equal the oracle's, whose generator is pinned against libc."""
    import orb_amd
    lib = orb_amd.lib()
    for patch in (31, 25, 8):
        for wta_k in (2, 3, 4):
            assert np.array_equal(_pattern(lib, wta_k, patch), oracle.pattern(wta_k, patch)), (patch, wta_k)
    buf = orb_amd._np_ptr(np.zeros(1024, np.int32))
    assert lib.orb_debug_pattern(31, 5, buf) == -1
    assert lib.orb_debug_pattern(32, 2, buf) == -1 and lib.orb_debug_pattern(2, 2, buf) == -1


def test_pattern_leaves_the_callers_rand_stream_alone():
    import orb_amd
    lib = orb_amd.lib()
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(4242)
    want = [libc.rand() for _ in range(6)]
    libc.srand(4242)
    got = [libc.rand(), libc.rand()]
    _pattern(lib, 3)
    got.append(libc.rand())
    _pattern(lib, 4)
    got += [libc.rand() for _ in range(3)]
    assert got == want


def test_concurrent_pattern_builds_agree(oracle):
    """Eight threads build WTA_K 3 and 4 patterns at once (ctypes releases the GIL during the
    call): each result equals the single-threaded one, so no generator state is shared."""
    import threading

    import orb_amd
    lib = orb_amd.lib()
    want = {k: oracle.pattern(k) for k in (3, 4)}
    bad = []

    def work(k):
        for _ in range(50):
            if not np.array_equal(_pattern(lib, k), want[k]):
                bad.append(k)

    threads = [threading.Thread(target=work, args=(3 + i % 2,)) for i in range(8)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not bad
