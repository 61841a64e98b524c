"""NumPy FP64 restatement of the learned warm start -- TEST INFRASTRUCTURE ONLY.

The reference has no learned initialiser: it exposes ``init_q`` (src/driver.jl:56,184-185) and starts
from random Q otherwise.  This module restates, in float64 NumPy, the pieces that the product computes
with hand-written kernels (csrc/hk_mlp.hip), so that tests can check them:

* ``gauss_matrix`` / ``features``: the seeded Gaussian random projection of allele-2 dosages;
* ``forward`` / ``loss_grad``: the 3-layer MLP with BF16-rounded weight matrices, softmax output and the
  permutation-matched cross-entropy (exhaustive search over label permutations);
* ``adam``: one Adam update of the FP32 parameter block.

Only ``tests/`` imports this module.  Pinning: ``loss_grad`` is checked against central finite
differences and the permutation search against an explicit loop over relabelled targets
(tests/test_oracle_warmstart.py); the projection's Gaussian matrix against its first two moments.
"""
from __future__ import annotations

import itertools

import numpy as np

D_FEAT = 256
_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _mix64(z: np.ndarray) -> np.ndarray:
    z = z + np.uint64(0x9E3779B97F4A7C15)
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def gauss_matrix(S: int, seed: int, D: int = D_FEAT) -> np.ndarray:
    """R (S x D): Box-Muller on the two 24-bit halves of a counter hash of (seed, s * D + d)."""
    with np.errstate(over="ignore"):
        t = np.arange(S * D, dtype=np.uint64)
        h = _mix64(_mix64(np.array([seed], dtype=np.uint64))[0] ^ t)
    u1 = ((h >> np.uint64(40)).astype(np.float64) + 0.5) * 2.0 ** -24
    u2 = ((h & np.uint64(0xFFFFFF)).astype(np.float64) + 0.5) * 2.0 ** -24
    return (np.sqrt(-2.0 * np.log(u1)) * np.cos(6.2831853 * u2)).reshape(S, D)


def dosages(codes: np.ndarray) -> np.ndarray:
    """PLINK 2-bit codes -> allele-2 dosage (0, 1, 2), NaN where missing (code 1)."""
    d = np.where(codes == 3, 2.0, np.where(codes == 2, 1.0, 0.0))
    return np.where(codes == 1, np.nan, d)


def features(codes: np.ndarray, n_snp: int, seed: int, D: int = D_FEAT, rows=None) -> np.ndarray:
    """F (I x D) = S^-1/2 sum_{s observed} (g_is - mu_s) R_s over the first n_snp SNPs.

    codes: (n_snp_total x I) PLINK codes, SNP-major.  mu_s is the observed mean dosage of SNP s over all
    individuals; rows (optional) selects the individuals whose features are returned."""
    C = codes[:n_snp]
    cnt = np.count_nonzero(C != 1, axis=1)
    tot = np.count_nonzero(C == 2, axis=1) + 2 * np.count_nonzero(C == 3, axis=1)
    mu = np.where(cnt > 0, tot / np.maximum(cnt, 1), 0.0)
    G = dosages(C if rows is None else C[:, rows]).T  # I x S
    Z = np.where(np.isnan(G), 0.0, G - mu)
    return Z @ gauss_matrix(n_snp, seed, D) / np.sqrt(n_snp)


def bf16(x: np.ndarray) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest, ties to even) and return them as float64."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + np.uint64(0x7FFF) + ((u >> np.uint64(16)) & np.uint64(1))) & np.uint64(0xFFFF0000)
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(w: np.ndarray, offsets, D: int, H: int, K: int):
    sizes = [(D, H), (H,), (H, H), (H,), (H, K), (K,)]
    return [np.asarray(w[o:o + int(np.prod(s))], dtype=np.float64).reshape(s) for o, s in zip(offsets, sizes)]


def forward(w, offsets, D, H, K, X, round_bf16=True):
    W1, b1, W2, b2, W3, b3 = unpack(w, offsets, D, H, K)
    if round_bf16:
        W1, W2, W3 = bf16(W1), bf16(W2), bf16(W3)
    X = np.asarray(X, dtype=np.float64)
    H1 = np.maximum(X @ W1 + b1, 0.0)
    H2 = np.maximum(H1 @ W2 + b2, 0.0)
    Z = H2 @ W3 + b3
    Z = Z - Z.max(axis=1, keepdims=True)
    P = np.exp(Z)
    P /= P.sum(axis=1, keepdims=True)
    return P, (X, H1, H2, W1, W2, W3)


def best_perm(P, T):
    """perm minimising -sum_i sum_k T[i, perm[k]] log P[i, k]: K <= 8 enumerates the permutations (first
    minimum in lexicographic order); larger K solves the assignment problem on the K x K cost matrix with
    scipy.optimize.linear_sum_assignment, keeping the identity when it is as good."""
    K = P.shape[1]
    L = np.log(np.maximum(P, 1e-30))
    if K > 8:
        from scipy.optimize import linear_sum_assignment

        C = -(L.T @ np.asarray(T, dtype=np.float64))  # C[k, k2] = -sum_i T[i, k2] log P[i, k]
        _, cols = linear_sum_assignment(C)
        v, v_id = float(C[np.arange(K), cols].sum()), float(np.trace(C))
        return (cols, v) if v < v_id else (np.arange(K), v_id)
    best, best_v = None, np.inf
    for perm in itertools.permutations(range(K)):
        v = -float(np.sum(T[:, list(perm)] * L))
        if v < best_v:
            best, best_v = perm, v
    return np.array(best), best_v


def loss_grad(w, offsets, D, H, K, X, T, match=True, round_bf16=True):
    """(loss, perm, grad) of the mean (matched) cross-entropy; grad has the parameter-block layout of w."""
    n = X.shape[0]
    P, (X, H1, H2, W1, W2, W3) = forward(w, offsets, D, H, K, X, round_bf16)
    T = np.asarray(T, dtype=np.float64)
    if match:
        perm, v = best_perm(P, T)
    else:
        perm = np.arange(K)
        v = -float(np.sum(T * np.log(np.maximum(P, 1e-30))))
    dZ = (P - T[:, perm]) / n
    gW3 = H2.T @ dZ
    gb3 = dZ.sum(axis=0)
    dH2 = (dZ @ W3.T) * (H2 > 0)
    gW2 = H1.T @ dH2
    gb2 = dH2.sum(axis=0)
    dH1 = (dH2 @ W2.T) * (H1 > 0)
    gW1 = X.T @ dH1
    gb1 = dH1.sum(axis=0)
    g = np.zeros(np.asarray(w).size)
    for o, a in zip(offsets, [gW1, gb1, gW2, gb2, gW3, gb3]):
        g[o:o + a.size] = a.ravel()
    return v / n, perm, g


def adam(w, g, m, v, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (bias-corrected); returns (w, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    c1, c2 = 1 - b1 ** step, 1 - b2 ** step
    return w - lr * (m / c1) / (np.sqrt(v / c2) + eps), m, v
