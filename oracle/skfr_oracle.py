"""CPU restatement of the SKFR marker pre-selection -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg import this module, to
check or to time the CPU path; the product library (``hk_skfr`` in libhaplo_hip.so) never routes through it.

Reference: src/driver.jl:84-131 (``_filter_SKFR``) calls SparseKmeansFeatureRanking.jl's
``sparsekmeans1`` (skfr_mode = :global), ``sparsekmeans2`` (:local) and ``sparsekmeans_repeat``
(tries > 1) on ``ImputedSnpMatrix{Float64}(g, K; blocksize=2)``.  That package is a registered Julia
dependency (Project.toml: SparseKmeansFeatureRanking, compat 0.1) that is not vendored under the
reference, so this file restates the published method (Zhang, Lange & Xu, "Simple and scalable sparse
k-means clustering via feature ranking", NeurIPS 2020) over SNP-pair blocks, in the same floating-point
operation order as the device engine, so labels and markers agree exactly:

* ``standardise``: per-SNP mean / standard deviation over the observed entries (integer sums),
  z = (x - m) / s, 0 for a constant SNP;
* ``kmeanspp``: k-means++ starts from a splitmix64 stream, squared distances over all SNPs with missing
  z taken as 0, summed in chunks of ``CHUNK`` blocks (SNPs ascending, then chunks ascending);
* ``skfr``: center update (mean of the observed z per cluster), block scores
  sum_c n_c (theta_c,2b^2 + theta_c,2b+1^2) (global) or per cluster (local), stable top-s, theta = 0
  off the kept blocks, reassignment by sum_kept theta^2 - sum_kept (2 z) theta with a missing z = the
  center of the current cluster (both sums per partial of ``assign_chunk`` kept blocks, partials of
  sum theta^2 - sum (2 z) theta added in order), stop when no label moves; tries keep the smallest within-cluster sum of squares;
  markers are the kept blocks' SNP pairs in decreasing order of the all-cluster score of the last center
  update (ties to the smaller block), as ``aims`` is documented (docs/src/index.md:287).

Parity against SparseKmeansFeatureRanking.jl itself is unpinned (its source and its RNG are not here).
The restatement is pinned by tests/test_oracle_skfr.py: with every block kept and complete data it is
Lloyd's k-means (scikit-learn ``KMeans(algorithm="lloyd", tol=0)`` from the same starts gives the same
labels and inertia); its output is a fixed point of the method, checked by brute-force loops over
individuals; on planted structure it selects exactly the informative blocks; splitmix64 matches the
published test vector.
"""
from __future__ import annotations

import numpy as np

CHUNK = 1024          # blocks per partial sum of the all-SNP distances (device: SK_CHUNK)
ACH_MIN = 256         # kept blocks per partial of the reassignment sums, at least (device: SK_ACH_MIN)
MAX_CLUSTERS = 16    # the device's per-lane cluster arrays (the published method has no limit)
_MASK = (1 << 64) - 1


class SplitMix64:
    """splitmix64 (Steele, Lea & Flood 2014); ``uniform`` = top 53 bits / 2^53."""

    def __init__(self, seed: int):
        self.s = int(seed) & _MASK

    def next(self) -> int:
        self.s = (self.s + 0x9E3779B97F4A7C15) & _MASK
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK
        return z ^ (z >> 31)

    def uniform(self) -> float:
        return float(self.next() >> 11) * (1.0 / 9007199254740992.0)


def dosages(codes: np.ndarray) -> np.ndarray:
    """PLINK 2-bit codes (n_snps x I, as ``plink.unpack_codes``) -> allele-2 dosages (I x n_snps) int8,
    -1 = missing (0b00 -> 0, 0b10 -> 1, 0b11 -> 2, 0b01 missing)."""
    lut = np.array([0, -1, 1, 2], dtype=np.int8)
    return np.ascontiguousarray(lut[np.asarray(codes, dtype=np.uint8)].T)


def standardise(D: np.ndarray):
    """(mean, sd, Z): Z = (D - mean) / sd over observed entries, 0 where sd = 0, NaN where missing."""
    obs = D >= 0
    Di = np.where(obs, D, 0).astype(np.int64)
    n = obs.sum(axis=0, dtype=np.int64)
    s = Di.sum(axis=0)
    ss = (Di * Di).sum(axis=0)
    m = np.zeros(D.shape[1])
    v = np.zeros(D.shape[1])
    nz = n > 0
    m[nz] = s[nz].astype(np.float64) / n[nz].astype(np.float64)
    v[nz] = ss[nz].astype(np.float64) / n[nz].astype(np.float64) - m[nz] * m[nz]
    sd = np.where(v > 0.0, np.sqrt(np.where(v > 0.0, v, 0.0)), 0.0)
    Df = D.astype(np.float64)
    Z = np.divide(Df - m, sd, out=np.zeros_like(Df), where=sd > 0.0)
    Z[~obs] = np.nan
    return m, sd, Z


def _chunked_dist(Z0: np.ndarray, rows: np.ndarray, J: int) -> np.ndarray:
    """Squared distances (I x nc) of the rows of Z0 (missing already 0) to the center rows (nc x 2J), in
    the device order: SNPs ascending within chunks of CHUNK blocks, chunk partials added in order."""
    I = Z0.shape[0]
    nc = rows.shape[0]
    tot = np.zeros((I, nc))
    for j0 in range(0, J, CHUNK):
        acc = np.zeros((I, nc))
        for col in range(2 * j0, 2 * min(J, j0 + CHUNK)):
            e = Z0[:, col][:, None] - rows[:, col][None, :]
            acc = acc + e * e
        tot = tot + acc
    return tot


def kmeanspp(Z: np.ndarray, k: int, rng: SplitMix64) -> list:
    """Starting individuals: the first uniform, each further one with probability proportional to the
    squared distance (all SNPs, missing -> 0) to the nearest start so far."""
    I, S = Z.shape
    J = S // 2
    Z0 = np.where(np.isnan(Z), 0.0, Z)
    starts = []
    d2 = None
    for c in range(k):
        if c == 0:
            r = int(rng.uniform() * float(I))
            r = min(r, I - 1)
        else:
            cum = np.cumsum(d2)          # sequential: cum[-1] is the left-to-right total
            u = rng.uniform() * cum[-1]
            hit = np.nonzero(cum > u)[0]
            r = int(hit[0]) if hit.size else I - 1
        starts.append(r)
        if c + 1 < k:
            d = _chunked_dist(Z0, Z0[r][None, :], J)[:, 0]
            d2 = d if c == 0 else np.where(d < d2, d, d2)
    return starts


def centers(D: np.ndarray, labels: np.ndarray, k: int, m: np.ndarray, sd: np.ndarray) -> np.ndarray:
    """theta (k x 2J): mean standardised dosage over the observed entries of each cluster."""
    obs = D >= 0
    Di = np.where(obs, D, 0).astype(np.int64)
    th = np.zeros((k, D.shape[1]))
    for c in range(k):
        mem = labels == c
        n = obs[mem].sum(axis=0, dtype=np.int64)
        s = Di[mem].sum(axis=0)
        ok = (n > 0) & (sd > 0.0)
        th[c, ok] = (s[ok].astype(np.float64) / n[ok].astype(np.float64) - m[ok]) / sd[ok]
    return th


def scores(theta: np.ndarray, csize: np.ndarray, local: bool) -> np.ndarray:
    """Block scores: (1 x J) global sum over clusters, or (k x J) per cluster."""
    k = theta.shape[0]
    per = []
    for c in range(k):
        sc = 0.0 + theta[c, 0::2] * theta[c, 0::2]
        sc = sc + theta[c, 1::2] * theta[c, 1::2]
        per.append(float(csize[c]) * sc)
    if local:
        return np.stack(per)
    g = np.zeros_like(per[0])
    for v in per:
        g = g + v
    return g[None, :]


def top_blocks(score_row: np.ndarray, s: int) -> np.ndarray:
    """The s best blocks of one score row; a tie goes to the smaller block."""
    return np.argsort(-score_row, kind="stable")[:s]


def assign_chunk(maxnb: int) -> int:
    """Kept blocks per partial of the reassignment sums: at least ACH_MIN, at most 64 partials."""
    return max(ACH_MIN, -(-maxnb // 64))


def assign(Z: np.ndarray, theta: np.ndarray, labels: np.ndarray, blocks: np.ndarray, ach: int):
    """(new labels, per-individual WSS) over the kept blocks (ascending), summed in the device order:
    blocks in list order within partials of `ach` blocks, partials added in order."""
    I = Z.shape[0]
    k = theta.shape[0]
    cost = np.zeros((I, k))
    for t0 in range(0, len(blocks), ach):
        # chunk constant sum theta^2 and cross term sum (2 z) theta, each in list order
        a_c = np.zeros(k)
        acc = np.zeros((I, k))
        for b in blocks[t0:t0 + ach]:
            for a in (0, 1):
                col = 2 * int(b) + a
                th = theta[:, col]
                a_c = a_c + th * th
                z = Z[:, col]
                z2 = 2.0 * np.where(np.isnan(z), theta[labels, col], z)
                acc = acc + z2[:, None] * th
        cost = cost + (a_c[None, :] - acc)
    best = np.argmin(cost, axis=1)
    w = np.zeros(I)
    for t0 in range(0, len(blocks), ach):
        acc = np.zeros(I)
        for b in blocks[t0:t0 + ach]:
            for a in (0, 1):
                col = 2 * int(b) + a
                z = Z[:, col]
                th = theta[best, col]
                e = np.where(np.isnan(z), th, z) - th
                acc = acc + e * e
        w = w + acc
    return best.astype(np.int32), w


def skfr(D: np.ndarray, k: int, sparsity: int, *, mode: str = "global", max_iter: int = 50, tries: int = 1,
         seed: int = 1, init=None, init_labels=None, return_history: bool = False) -> dict:
    """SKFR on the dosage matrix D (I x 2J, int8, -1 missing).  Returns clusters (I,), aims (0-based SNP
    indices, most important block first), wss, iters (and the per-try results when return_history).
    init_labels: the first try starts from these cluster labels (no starting centers, no first assignment)."""
    I, S = D.shape
    if S % 2:
        raise ValueError("D must hold whole SNP blocks (an even number of columns)")
    J = S // 2
    if not (1 <= k <= MAX_CLUSTERS and k <= I and 1 <= sparsity <= J and max_iter >= 1 and tries >= 1):
        raise ValueError("invalid SKFR arguments")
    if mode not in ("global", "local"):
        raise ValueError("mode must be 'global' or 'local'")
    local = mode == "local"
    ach = assign_chunk(min(J, k * sparsity) if local else sparsity)
    m, sd, Z = standardise(D)
    Z0 = np.where(np.isnan(Z), 0.0, Z)
    rng = SplitMix64(seed)
    best = None
    history = []
    for tr in range(tries):
        if tr == 0 and init_labels is not None:
            starts = None
            labels = np.asarray(init_labels, dtype=np.int32).copy()
        else:
            starts = list(init) if (tr == 0 and init is not None) else kmeanspp(Z, k, rng)
            labels = np.argmin(_chunked_dist(Z0, Z0[starts], J), axis=1).astype(np.int32)
        it = 0
        for it in range(1, max_iter + 1):
            csize = np.bincount(labels, minlength=k)
            theta = centers(D, labels, k, m, sd)
            sc = scores(theta, csize, local)
            gscore = scores(theta, csize, False)[0]
            keep = np.zeros((k, J), dtype=bool)
            for c in range(sc.shape[0]):
                keep[c, top_blocks(sc[c], sparsity)] = True
            if not local:
                keep[:] = keep[0]
            for c in range(k):
                theta[c, 0::2][~keep[c]] = 0.0
                theta[c, 1::2][~keep[c]] = 0.0
            blocks = np.nonzero(keep.any(axis=0))[0]
            new, w = assign(Z, theta, labels, blocks, ach)
            moved = int((new != labels).sum())
            labels = new
            wsum = float(np.cumsum(w)[-1])
            if moved == 0:
                break
        key = np.where(keep.any(axis=0), gscore, -1.0)
        ranked = np.argsort(-key, kind="stable")[: blocks.size]
        res = dict(clusters=labels.copy(), blocks=ranked, wss=wsum, iters=it, starts=starts)
        history.append(res)
        if tr == 0 or wsum < best["wss"]:
            best = res
    aims = np.stack([2 * best["blocks"], 2 * best["blocks"] + 1], axis=1).ravel().astype(np.int32)
    out = dict(clusters=best["clusters"], aims=aims, wss=best["wss"], iters=best["iters"])
    if return_history:
        out["history"] = history
    return out


def skfr_path(D: np.ndarray, k: int, sparsities, *, mode: str = "global", max_iter: int = 50, tries: int = 1,
              seed: int = 1) -> list:
    """sparsekmeans_path restated: SKFR for each of a decreasing list of sparsities, the first with `tries`
    seeded starts, each next one started from the previous result's cluster labels."""
    out, prev = [], None
    for t, sp in enumerate(sparsities):
        r = skfr(D, k, int(sp), mode=mode, max_iter=max_iter, tries=tries if t == 0 else 1, seed=seed,
                 init_labels=prev)
        out.append(r)
        prev = r["clusters"]
    return out
