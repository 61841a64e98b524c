"""SKFR ancestry-informative-marker pre-selection on the MI355X (``hk_skfr`` of libhaplo_hip.so).

Host mirror of the calls src/driver.jl:84-131 makes into SparseKmeansFeatureRanking.jl:
``ImputedSnpMatrix{Float64}(g, K; rng, blocksize=2)``, ``sparsekmeans1`` (skfr_mode = :global),
``sparsekmeans2`` (:local) and ``sparsekmeans_repeat`` (tries > 1).  The clustering runs on the device
over the packed genotype streams; these wrappers only move labels and marker indices.

Conventions of this library: labels are 0..k-1 and markers are 0-based SNP indices (the reference's are
1-based Julia indices); ``rng`` is an integer seed of the splitmix64 stream that draws the k-means++ starts.
The tuples keep the reference's arity with ``None`` for the cluster centers, which its call sites discard.
"""
from __future__ import annotations

import numpy as np

from . import plink
from .engine import HipAdmixData


class ImputedSnpMatrix:
    """The genotype blocks of a PLINK fileset on the device, prepared for SKFR with k clusters.

    blocksize=2 (the only block size the reference uses, driver.jl:99): SNPs 2b and 2b+1 form block b.  A
    fileset with an odd SNP count gets an all-missing SNP appended to close the last block; it is never
    reported as a marker."""

    def __init__(self, g, k: int, rng: int = 1, blocksize: int = 2):
        if blocksize != 2:
            raise NotImplementedError("only blocksize=2 (SNP pairs) is supported")
        fs = g if isinstance(g, plink.PlinkFileset) else plink.PlinkFileset(str(g))
        self.n_individuals, self.n_snps = fs.shape
        bps = fs.bytes_per_snp
        payload = np.asarray(fs.data[: self.n_snps * bps])
        nsnp = self.n_snps
        if nsnp % 2:
            payload = np.concatenate([payload, np.full(bps, 0x55, dtype=np.uint8)])
            nsnp += 1
        self.k = int(k)
        self.seed = int(rng)
        self.J = nsnp // 2
        self.eng = HipAdmixData(payload, self.n_individuals, self.J, self.k, bytes_per_snp=bps, n_snp=nsnp)

    def close(self):
        self.eng.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def run(self, sparsity: int, mode: str, max_iter: int, tries: int, init=None) -> dict:
        r = self.eng.skfr(self.k, sparsity, mode=mode, max_iter=max_iter, tries=tries, seed=self.seed, init=init)
        r["aims"] = r["aims"][r["aims"] < self.n_snps]
        return r


def sparsekmeans1(X: ImputedSnpMatrix, sparsity: int, max_iter: int = 1000, squares=None):
    """Global feature ranking: (clusters, None, aims, wss, iters)."""
    r = X.run(sparsity, "global", max_iter, 1)
    return r["clusters"], None, r["aims"], r["wss"], r["iters"]


def sparsekmeans2(X: ImputedSnpMatrix, sparsity: int, max_iter: int = 1000, squares=None):
    """Feature ranking per cluster (markers = union of the clusters' blocks): (clusters, None, aims, wss,
    iters)."""
    r = X.run(sparsity, "local", max_iter, 1)
    return r["clusters"], None, r["aims"], r["wss"], r["iters"]


def sparsekmeans_repeat(X: ImputedSnpMatrix, sparsity: int, iter: int = 20, max_inner_iter: int = 50,
                        ftn=sparsekmeans1):
    """Best of `iter` seeded runs by within-cluster sum of squares: (clusters, None, aims, wss, iters, iter)."""
    mode = "local" if ftn is sparsekmeans2 else "global"
    r = X.run(sparsity, mode, max_inner_iter, iter)
    return r["clusters"], None, r["aims"], r["wss"], r["iters"], iter


def sparsekmeans_path(X: ImputedSnpMatrix, sparsities, iter: int = 20, max_inner_iter: int = 50, ftn=sparsekmeans1):
    """SKFR along a decreasing list of sparsities (the call at src/driver.jl:122): the first sparsity is the
    best of `iter` seeded runs, each next one starts from the previous one's labels.  Returns (clusters,
    aims): one label vector and one marker list per sparsity."""
    mode = "local" if ftn is sparsekmeans2 else "global"
    rs = X.eng.skfr_path(X.k, sparsities, mode=mode, max_iter=max_inner_iter, tries=iter, seed=X.seed)
    return [r["clusters"] for r in rs], [r["aims"][r["aims"] < X.n_snps] for r in rs]
