"""orb_match_top2_noalloc, the allocation-free variant of orb_match_top2 (for
callers that cannot allocate device memory): through the C ABI its outputs
equal orb_match_top2's bit for bit, for every method, on two calls in a row
(the second smaller, so stale scratch contents from the first call must not
leak), it never grows the context's workspace, and it rejects query counts
above its fixed limit."""
import ctypes

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs an MI355X")]

METHODS = {"popcount": 0, "mfma": 1, "hamming2": 2}


def _inputs(rng, nq, ndb):
    db = rng.integers(0, 256, (ndb, 32), dtype=np.uint8)
    db[ndb // 2] = db[3]               # ties go to the lowest index
    src = rng.integers(0, ndb, nq)
    flips = rng.random((nq, 256)) < 0.05
    q = db[src] ^ np.packbits(flips, axis=1, bitorder="little")
    q[:nq // 8] = rng.integers(0, 256, (nq // 8, 32), dtype=np.uint8)   # some far queries
    return q, db


@pytest.mark.parametrize("method", list(METHODS.values()), ids=list(METHODS))
def test_noalloc_equals_top2_on_two_calls(method):
    import torch

    import orb_amd
    print(f"noalloc method={method}", flush=True)
    ctx = orb_amd.Context(0)
    other = orb_amd.Context(0)       # the allocating reference path, in its own context
    try:
        rng = np.random.default_rng(1234 + method)
        # the second call is smaller in both dimensions
        for nq, ndb in ((20000, 70000), (3000, 5000)):
            q, db = _inputs(rng, nq, ndb)
            dq, ddb = torch.from_numpy(q).cuda(), torch.from_numpy(db).cuda()
            out = [torch.full((nq,), -7, dtype=torch.int32, device="cuda") for _ in range(4)]
            ctx.match_top2_noalloc(dq, ddb, out, method=method)
            want = other.match_top2(dq, ddb, method=method)
            torch.cuda.synchronize()
            for name, g, w in zip(("idx1", "d1", "idx2", "d2"), out, want):
                assert torch.equal(g, w), (nq, ndb, name)
            assert ctx.workspace_bytes() == 0, "the allocation-free path grew the context workspace"
    finally:
        ctx.close()
        other.close()


def test_noalloc_rejects_queries_above_limit():
    import torch

    import orb_amd
    lib = orb_amd.lib()
    ctx = orb_amd.Context(0)
    try:
        limit = orb_amd.NOALLOC_MAX_QUERIES
        db = torch.zeros((64, 32), dtype=torch.uint8, device="cuda")
        rc = lib.orb_match_top2_noalloc(ctx.h, ctypes.c_void_p(db.data_ptr()), limit + 1,
                                        ctypes.c_void_p(db.data_ptr()), 64, 0, None, None, None, None)
        assert rc == -5   # ORB_E_CAPACITY, before anything is launched
        assert b"allocation-free limit" in lib.orb_last_error(ctx.h)
        assert ctx.workspace_bytes() == 0
    finally:
        ctx.close()
