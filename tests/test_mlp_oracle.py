"""Pins of the learned matcher's CPU oracle (oracle/mlp.py), checked against
computations that do not reuse its formulas:

- the hand-derived backward pass against torch autograd in float64 and
  against central finite differences, with the bf16 rounding switched off
  (the float64 network the kernels approximate), with and without the
  centroid term;
- the forward pass against torch's own erf GELU and cross-entropy;
- the bf16 rounding helper against torch's float32 -> bfloat16 conversion;
- the bucketed exact top-2 against the C oracle's exact matcher (all buckets
  probed), and against hand-built cases with ties and one-entry buckets.
"""
import numpy as np
import pytest
import torch

from oracle import mlp


def _params(rng, hidden, nb, scale=0.2):
    return {"W1": rng.normal(0, scale, (256, hidden)), "b1": rng.normal(0, 0.1, hidden),
            "W2": rng.normal(0, scale, (hidden, hidden)), "b2": rng.normal(0, 0.1, hidden),
            "W3": rng.normal(0, scale, (hidden, nb)), "b3": rng.normal(0, 0.1, nb)}


def _torch_loss(params, desc, labels, cg=None):
    """The same network written with torch's building blocks (erf GELU, cross_entropy)."""
    T = {k: torch.tensor(v, dtype=torch.float64, requires_grad=True) for k, v in params.items()}
    bits = torch.tensor(np.unpackbits(desc, axis=1, bitorder="little"), dtype=torch.float64)
    X = bits * 2 - 1
    A1 = torch.nn.functional.gelu(X @ T["W1"] + T["b1"], approximate="none")
    A2 = torch.nn.functional.gelu(A1 @ T["W2"] + T["b2"], approximate="none")
    Z3 = A2 @ T["W3"] + T["b3"]
    if cg is not None:
        Z3 = Z3 + X @ torch.tensor(cg, dtype=torch.float64).T
    loss = torch.nn.functional.cross_entropy(Z3, torch.tensor(labels, dtype=torch.long))
    loss.backward()
    return float(loss.detach()), {k: t.grad.numpy() for k, t in T.items()}, Z3.detach().numpy()


@pytest.mark.parametrize("with_centroids", [False, True], ids=["plain", "centroid-term"])
def test_backward_matches_torch_autograd(with_centroids):
    rng = np.random.default_rng(3)
    hidden, nb, n = 48, 40, 33
    params = _params(rng, hidden, nb)
    desc = rng.integers(0, 256, (n, 32), dtype=np.uint8)
    labels = rng.integers(0, nb, n)
    cg = mlp.centroid_term(rng.integers(0, 256, (nb, 32), dtype=np.uint8), 0.25) if with_centroids else None
    loss, g, f = mlp.loss_and_grads(params, desc, labels, cg, rounding=False)
    want_loss, want_g, want_z3 = _torch_loss(params, desc, labels, cg)
    assert loss == pytest.approx(want_loss, rel=1e-12)
    np.testing.assert_allclose(f["Z3"], want_z3, rtol=1e-12, atol=1e-12)
    for k in ("W1", "b1", "W2", "b2", "W3", "b3"):
        np.testing.assert_allclose(g[k], want_g[k], rtol=1e-9, atol=1e-13, err_msg=k)


@pytest.mark.parametrize("with_centroids", [False, True], ids=["plain", "centroid-term"])
def test_backward_matches_finite_differences(with_centroids):
    """Central differences of the oracle's own float64 loss on a tiny network: every parameter of
    every tensor, step 1e-6, so an error in any one gradient entry (a transposed product, a missing
    bias term, a wrong GELU derivative) shows."""
    rng = np.random.default_rng(4)
    hidden, nb, n = 5, 4, 7
    params = _params(rng, hidden, nb, scale=0.5)
    desc = rng.integers(0, 256, (n, 32), dtype=np.uint8)
    labels = rng.integers(0, nb, n)
    cg = mlp.centroid_term(rng.integers(0, 256, (nb, 32), dtype=np.uint8), 0.25) if with_centroids else None
    _, g, _ = mlp.loss_and_grads(params, desc, labels, cg, rounding=False)
    h = 1e-6
    for k in ("W1", "b1", "W2", "b2", "W3", "b3"):
        flat = params[k].reshape(-1)
        num = np.zeros_like(flat)
        for i in range(flat.size):
            old = flat[i]
            flat[i] = old + h
            lp = mlp.loss_and_grads(params, desc, labels, cg, rounding=False)[0]
            flat[i] = old - h
            lm = mlp.loss_and_grads(params, desc, labels, cg, rounding=False)[0]
            flat[i] = old
            num[i] = (lp - lm) / (2 * h)
        np.testing.assert_allclose(g[k].reshape(-1), num, rtol=1e-5, atol=1e-9, err_msg=k)


def test_gelu_and_its_derivative_match_torch():
    z = np.concatenate([np.linspace(-8, 8, 4001), [0.0, -1e-8, 1e-8, 30.0, -30.0]])
    t = torch.tensor(z, dtype=torch.float64, requires_grad=True)
    y = torch.nn.functional.gelu(t, approximate="none")
    y.sum().backward()
    np.testing.assert_allclose(mlp.gelu(z), y.detach().numpy(), rtol=1e-14, atol=1e-15)
    np.testing.assert_allclose(mlp.gelu_grad(z), t.grad.numpy(), rtol=1e-12, atol=1e-15)


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(5)
    x = rng.normal(0, 3, 200000).astype(np.float32)
    # exact halfway cases (ties to even both ways), denormals, large values, signed zeros
    u = rng.integers(0, 1 << 15, 2000, dtype=np.uint32) << 16 | 0x8000
    x = np.concatenate([x, u.view(np.float32), np.float32([0.0, -0.0, 1e-40, -1e-40, 3.0e38, -3.0e38, 1.0])])
    x = x[np.isfinite(x)]
    want = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    got = mlp.bf16(x)
    assert np.array_equal(got, want)


def test_bits_expansion_is_lsb_first():
    desc = np.zeros((1, 32), np.uint8)
    desc[0, 0] = 0b00000101   # bits 0 and 2
    desc[0, 31] = 0b10000000  # bit 255
    x = mlp.bits_pm1(desc)[0]
    assert set(np.nonzero(x > 0)[0]) == {0, 2, 255}
    assert ((x == 1) | (x == -1)).all()


def test_bucketed_top2_equals_exact_matcher_when_every_bucket_is_probed(oracle):
    rng = np.random.default_rng(6)
    db = rng.integers(0, 256, (3000, 32), dtype=np.uint8)
    db[1700] = db[12]                 # a tied pair
    q = db[rng.integers(0, 3000, 200)] ^ np.packbits(rng.random((200, 256)) < 0.1, axis=1, bitorder="little")
    q[0] = db[12]
    bucket_of = rng.integers(0, 16, 3000)
    probes = np.tile(np.arange(16), (200, 1))
    got = mlp.bucket_search(q, db, bucket_of, probes)
    want = oracle.match_top2(q, db, nthreads=4)
    for a, b in zip(got, want):
        assert np.array_equal(a, b)
    assert got[0][0] == 12 and got[2][0] == 1700 and got[1][0] == got[3][0] == 0


def test_bucketed_top2_restricts_to_probed_buckets():
    db = np.zeros((6, 32), np.uint8)
    db[:, 0] = [0b0, 0b1, 0b11, 0b111, 0b1111, 0b0]   # distances 0,1,2,3,4,0 from a zero query
    bucket_of = np.array([0, 1, 1, 2, 3, 3])
    q = np.zeros((4, 32), np.uint8)
    probes = np.array([[1, 1], [2, 2], [3, 3], [0, 0]])
    i1, d1, i2, d2 = mlp.bucket_search(q, db, bucket_of, probes)
    assert list(i1) == [1, 3, 5, 0] and list(d1) == [1, 3, 0, 0]
    assert list(i2) == [2, -1, 4, -1] and list(d2) == [2, 257, 4, 257]


def test_probe_mass_rule_against_brute_force():
    """The mass rule of the oracle (probe_mass) against a direct restatement over every prefix:
    for each row, the kept buckets are the shortest best-first prefix of the top nprobe whose
    weight reaches mass x the row total (all nprobe if none does), ties ranked by bucket id;
    a hand case with ties fixes the order."""
    rng = np.random.default_rng(12)
    for trial in range(200):
        C = int(rng.integers(4, 40))
        k = int(rng.integers(1, C + 1))
        scores = rng.integers(-3, 4, C).astype(np.float64) if trial % 2 else rng.normal(0, 2, C)
        w = np.exp(scores - scores.max())
        mass = float(rng.choice([0.3, 0.5, 0.9, 0.99, 1.0]))
        got = mlp.probe_mass(scores[None, :], w[None, :], k, mass)[0]
        ranked = sorted(range(C), key=lambda b: (-scores[b], b))[:k]
        want = []
        for b in ranked:
            want.append(b)
            if sum(w[c] for c in want) >= mass * w.sum():
                break
        assert list(got[:len(want)]) == want and (got[len(want):] == -1).all(), (trial, got, want)
    # ties: equal scores ranked by bucket id; mass 0.5 of four equal weights keeps two
    got = mlp.probe_mass(np.zeros((1, 4)), np.ones((1, 4)), 4, 0.5)[0]
    assert list(got) == [0, 1, -1, -1]
