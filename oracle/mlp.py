"""CPU oracle of the learned bucket matcher -- TEST INFRASTRUCTURE.

Pinned by tests/test_mlp_oracle.py: the backward pass against torch autograd
and central finite differences of the float64 network (rounding off), the
bf16 helper against torch's bfloat16 conversion, gelu against torch's erf
GELU, the loss against torch's cross-entropy, and the bucketed top-2 against
brute force and the exact C matcher.

Only tests/ and bench.py's cpu_baseline leg use this module, as the checker.
It restates, in float64 numpy, the network and training step the HIP kernels
compute (cuda-orb_amd/csrc/orb_mlp.hip), with the same bf16 rounding points,
so GPU gradients can be compared to it within a stated tolerance, and the
exact bucketed search used to check the probe results.

Network (no counterpart in the reference; it is the learned part the
workload adds): X = +/-1 expansion of the 256 descriptor bits,
Z1 = X W1 + b1, A1 = bf16(gelu(Z1)), Z2 = A1 W2 + b2, A2 = bf16(gelu(Z2)),
Z3 = A2 W3 + b3 (bucket logits), plus with a centroid scale g the term
X Cg^T (Cg = the bucket centroids' bits as +/-g), loss = mean cross-entropy.
Backward: dZ3 = bf16((softmax(Z3) - onehot) / n), dW3 = A2^T dZ3,
db3 = colsum(dZ3), dZ2 = bf16((dZ3 W3^T) * gelu'(Z2)), ... down to dW1.
gelu(x) = x * Phi(x) with Phi the standard normal CDF (erf form).
"""
import math

import numpy as np
from scipy.special import erf

SQRT1_2 = 1.0 / math.sqrt(2.0)
INV_SQRT_2PI = 1.0 / math.sqrt(2.0 * math.pi)


def bf16(x):
    """Round float values to bfloat16 (round to nearest even), returned as float64."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def bits_pm1(desc):
    """(n, 32) uint8 descriptors -> (n, 256) float64 of +/-1 (bit k = byte k//8, bit k%8)."""
    bits = np.unpackbits(np.asarray(desc, np.uint8), axis=1, bitorder="little")
    return bits.astype(np.float64) * 2.0 - 1.0


def gelu(z):
    return z * 0.5 * (1.0 + erf(z * SQRT1_2))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z * SQRT1_2)) + z * INV_SQRT_2PI * np.exp(-0.5 * z * z)


def centroid_term(centroids, scale):
    """(C, 32) uint8 centroids -> (C, 256) float64 of +/-bf16(scale)."""
    return bits_pm1(centroids) * bf16(np.array([scale]))[0]


def _identity(x):
    return np.asarray(x, np.float64)


def forward(params, desc, cg=None, rounding=True):
    """params: dict of float64 arrays W1 (256,H), b1, W2 (H,H), b2, W3 (H,C), b3 (bf16-valued weights);
    cg: optional centroid term matrix from centroid_term().  rounding=False drops the bf16 rounding
    of the activations (the exact float64 network, for checking the backward pass)."""
    rnd = bf16 if rounding else _identity
    X = bits_pm1(desc)
    Z1 = X @ params["W1"] + params["b1"]
    A1 = rnd(gelu(Z1))
    Z2 = A1 @ params["W2"] + params["b2"]
    A2 = rnd(gelu(Z2))
    Z3 = A2 @ params["W3"]
    if cg is not None:
        Z3 = Z3 + X @ cg.T
    Z3 = Z3 + params["b3"]
    return {"X": X, "Z1": Z1, "A1": A1, "Z2": Z2, "A2": A2, "Z3": Z3}


def loss_and_grads(params, desc, labels, cg=None, rounding=True):
    """Mean softmax cross-entropy of the bucket labels and its gradients with respect to every
    parameter, rounding the backward products where the kernels do (rounding=False: none)."""
    rnd = bf16 if rounding else _identity
    f = forward(params, desc, cg, rounding)
    n = len(labels)
    Z3 = f["Z3"]
    m = Z3.max(1, keepdims=True)
    e = np.exp(Z3 - m)
    P = e / e.sum(1, keepdims=True)
    loss = float(-np.log(P[np.arange(n), labels]).mean())
    G = P.copy()
    G[np.arange(n), labels] -= 1.0
    dZ3 = rnd(G / n)
    g = {}
    g["W3"] = f["A2"].T @ dZ3
    g["b3"] = dZ3.sum(0)
    dZ2 = rnd((dZ3 @ params["W3"].T) * gelu_grad(f["Z2"]))
    g["W2"] = f["A1"].T @ dZ2
    g["b2"] = dZ2.sum(0)
    dZ1 = rnd((dZ2 @ params["W2"].T) * gelu_grad(f["Z1"]))
    g["W1"] = f["X"].T @ dZ1
    g["b1"] = dZ1.sum(0)
    return loss, g, f


def probe(logits, nprobe=8):
    """Top-nprobe buckets by logit (ties to the lower bucket id) and the max softmax probability."""
    order = np.lexsort((np.broadcast_to(np.arange(logits.shape[1]), logits.shape), -logits), axis=1)
    m = logits.max(1, keepdims=True)
    e = np.exp(logits - m)
    pmax = 1.0 / e.sum(1)
    return order[:, :nprobe], pmax


def probe_mass(scores, weights, nprobe, mass):
    """The probability-mass rule over the top-nprobe buckets: buckets ranked by score (higher
    first, ties to the lower bucket id); the bucket of rank r is kept iff the weights of ranks
    0 .. r-1 sum to less than mass times the row's total weight (float64 here).  Returns
    (n, nprobe) bucket ids, the kept prefix best first and -1 after it."""
    scores = np.asarray(scores, np.float64)
    weights = np.asarray(weights, np.float64)
    order = np.lexsort((np.broadcast_to(np.arange(scores.shape[1]), scores.shape), -scores), axis=1)[:, :nprobe]
    w = np.take_along_axis(weights, order, axis=1)
    before = np.cumsum(w, axis=1) - w
    keep = before < mass * weights.sum(1, keepdims=True)
    return np.where(keep, order, -1)


def bucket_search(q, db, bucket_of, buckets):
    """Exact top-2 of each query restricted to the database entries whose bucket is in its probe
    list (entries below 0 are empty slots), ties to the lowest database index.  Returns (idx1, d1,
    idx2, d2); a missing neighbour is -1 / 257 (orb_match_top2's convention)."""
    q64 = np.ascontiguousarray(q, np.uint8).view(np.uint64)
    db64 = np.ascontiguousarray(db, np.uint8).view(np.uint64)
    bucket_of = np.asarray(bucket_of)
    nb = int(max(bucket_of.max() + 1 if len(bucket_of) else 0, np.max(buckets) + 1 if np.size(buckets) else 0))
    members = np.argsort(bucket_of, kind="stable")          # entries of each bucket in index order
    starts = np.concatenate([[0], np.cumsum(np.bincount(bucket_of, minlength=nb))])
    out = [np.full(len(q), v, np.int64) for v in (-1, 257, -1, 257)]
    for i in range(len(q)):
        probed = np.unique(buckets[i])
        probed = probed[probed >= 0]
        cand = np.sort(np.concatenate([members[starts[b]:starts[b + 1]] for b in probed]))
        if len(cand) == 0:
            continue
        d = np.bitwise_count(db64[cand] ^ q64[i]).sum(1).astype(np.int64)
        order = np.lexsort((cand, d))
        out[0][i], out[1][i] = cand[order[0]], d[order[0]]
        if len(cand) > 1:
            out[2][i], out[3][i] = cand[order[1]], d[order[1]]
    return tuple(out)
