"""SKFR marker pre-selection on the GPU (hk_skfr through the C ABI) against the CPU restatement."""
import os

import numpy as np
import pytest

import skfr_oracle as so
from haplo_amd import HipAdmixData, HipError, plink, run_admixture
from haplo_amd.sim import simulate

pytestmark = pytest.mark.gpu

_CODE = np.array([1, 0, 2, 3], dtype=np.uint8)  # dosage -1, 0, 1, 2 -> PLINK code


def planted(I, J, k, informative, miss=0.0, seed=0, delta=0.35):
    rng = np.random.default_rng(seed)
    pop = np.arange(I) % k
    rng.shuffle(pop)
    base = rng.uniform(0.05, 0.95, size=2 * J)
    f = np.tile(base, (k, 1))
    for b in informative:
        for a in (0, 1):
            f[:, 2 * b + a] = np.clip(base[2 * b + a] + rng.uniform(-delta, delta, size=k), 0.02, 0.98)
    D = rng.binomial(2, f[pop]).astype(np.int8)
    if miss > 0:
        D[rng.random(D.shape) < miss] = -1
    return D, pop


def engine_for(D, K=None):
    I, S = D.shape
    payload = plink.pack_codes(_CODE[D.T.astype(np.int64) + 1])
    return HipAdmixData(payload, I, S // 2, K or 2)


CASES = [
    # I, J, k, sparsity, mode, tries, miss, seed, init
    (301, 150, 3, 10, "global", 1, 0.03, 5, None),
    (256, 1500, 4, 40, "global", 2, 0.02, 8, None),
    (97, 37, 1, 5, "global", 1, 0.0, 2, None),
    (200, 60, 5, 60, "global", 1, 0.01, 4, [0, 40, 80, 120, 160]),
    (300, 120, 3, 8, "local", 3, 0.02, 6, None),
    (123, 45, 8, 1, "local", 1, 0.1, 3, None),
    (500, 2100, 6, 300, "local", 1, 0.05, 9, None),
    (400, 3000, 4, 700, "global", 1, 0.02, 13, None),   # three reassignment partials
    (2500, 300, 3, 30, "global", 2, 0.03, 14, None),     # three center-count partials
    (9, 5, 3, 2, "global", 1, 0.1, 15, None),            # fewer individuals / blocks than a genotype word
    (16, 8, 2, 8, "local", 2, 0.3, 16, [3, 11]),         # heavy missingness, given starts, then a k-means++ try
    (7, 3, 7, 3, "global", 1, 0.0, 17, None),            # k = I
    (600, 200, 12, 40, "global", 1, 0.02, 21, None),     # k above 8
    (400, 300, 16, 60, "local", 1, 0.02, 22, None),      # the largest k
]


@pytest.mark.parametrize("case", CASES, ids=[f"I{c[0]}_J{c[1]}_k{c[2]}_s{c[3]}_{c[4]}_t{c[5]}" for c in CASES])
def test_skfr_matches_oracle(case):
    I, J, k, s, mode, tries, miss, seed, init = case
    D, _ = planted(I, J, k, informative=range(0, J, max(1, J // max(s, 1))), miss=miss, seed=seed)
    print(f"skfr case I={I} J={J} k={k} s={s} {mode} tries={tries}", flush=True)
    ref = so.skfr(D, k, s, mode=mode, tries=tries, seed=seed, init=init, max_iter=50)
    with engine_for(D, K=max(k, 1)) as eng:
        got = eng.skfr(k, s, mode=mode, max_iter=50, tries=tries, seed=seed, init=init)
    assert np.array_equal(got["clusters"], ref["clusters"])
    assert np.array_equal(got["aims"], ref["aims"])
    assert got["iters"] == ref["iters"]
    assert got["wss"] == pytest.approx(ref["wss"], rel=1e-12, abs=0.0)


def test_skfr_fixed_point_at_scale():
    """I=4000, J=20000: the device result is a fixed point of the method (checked with numpy)."""
    I, J, k, s = 4000, 20000, 4, 500
    D, pop = planted(I, J, k, informative=range(0, J, 37), miss=0.02, seed=12, delta=0.3)
    with engine_for(D, K=k) as eng:
        r = eng.skfr(k, s, mode="global", max_iter=100, tries=1, seed=3)
    assert r["iters"] < 100
    m, sd, Z = so.standardise(D)
    lab = r["clusters"].astype(np.int64)
    obs = D >= 0
    oh = np.eye(k)[lab]                                    # I x k
    n = oh.T @ obs.astype(np.float64)                      # k x 2J observed counts
    sm = oh.T @ np.where(obs, D, 0).astype(np.float64)     # k x 2J dosage sums
    with np.errstate(invalid="ignore", divide="ignore"):
        th = np.where((n > 0) & (sd > 0), (sm / np.maximum(n, 1) - m) / np.where(sd > 0, sd, 1), 0.0)
    size = oh.sum(axis=0)
    score = (size[:, None] * (th[:, 0::2] ** 2 + th[:, 1::2] ** 2)).sum(axis=0)
    order = r["aims"][0::2] // 2
    assert np.array_equal(r["aims"], np.stack([2 * order, 2 * order + 1], 1).ravel())
    kept = np.unique(order)
    assert kept.size == s
    assert np.all(np.diff(score[order]) <= 1e-9 * score[order[0]])   # most important block first
    rest = np.setdiff1d(np.arange(J), kept)
    assert score[kept].min() >= score[rest].max() - 1e-9 * score[kept].min()
    cols = np.stack([2 * kept, 2 * kept + 1], 1).ravel()
    Zk = Z[:, cols]
    Zk = np.where(np.isnan(Zk), th[lab][:, cols], Zk)
    cost = (Zk ** 2).sum(1)[:, None] - 2 * Zk @ th[:, cols].T + (th[:, cols] ** 2).sum(1)[None, :]
    assert np.all(cost[np.arange(I), lab] <= cost.min(axis=1) + 1e-8 * np.abs(cost).max())
    # the planted structure is found: the kept blocks are informative ones
    assert np.all(kept % 37 == 0)


def test_skfr_rejects_bad_arguments():
    D, _ = planted(40, 20, 2, informative=[0, 5], seed=1)
    with engine_for(D) as eng:
        for kw in (dict(k=17, sparsity=2), dict(k=2, sparsity=21), dict(k=2, sparsity=0), dict(k=0, sparsity=2)):
            with pytest.raises(HipError):
                eng.skfr(kw["k"], kw["sparsity"])
        with pytest.raises(HipError):
            eng.skfr(2, 2, init=[0, 40])
        with pytest.raises(HipError):
            eng.skfr(2, 2, max_iter=0)


@pytest.mark.parametrize("odd", [False, True])
def test_run_admixture_with_sparsity(tmp_path, odd):
    """run_admixture(sparsity=...) (driver.jl:58-79): SKFR on the whole fileset, the filtered fileset
    holds exactly the selected SNPs, and the fit runs on it."""
    I, J, K, s = 400, 150, 3, 30
    sim = simulate(I, J, K, fst=0.1, miss=0.01, seed=4)
    nsnp = 2 * J - (1 if odd else 0)
    bps = (I + 3) // 4
    path = str(tmp_path / "data.bed")
    plink.write_bed(path, sim.bed[: nsnp * bps], I, nsnp)
    print(f"run_admixture sparsity={s} odd={odd}", flush=True)
    # the fit reads the first 2 * Jfit SNPs of the filtered fileset (one short of 2s when the block closed
    # by the padding SNP is selected)
    Jfit = s - 1 if odd else s
    d, clusters, aims = run_admixture(path, I, Jfit, K, sparsity=s, skfr_tries=2, rng=7, admix_n_iter=30)
    des = str(tmp_path / f"data_{K}_{2 * s}aims.bed")
    assert os.path.exists(des)
    assert aims.size <= 2 * s and np.all(aims < nsnp)
    codes = plink.PlinkFileset(path).codes()
    D = so.dosages(codes)
    if odd:
        D = np.concatenate([D, np.full((I, 1), -1, dtype=np.int8)], axis=1)
    ref = so.skfr(D, K, s, tries=2, seed=7, max_iter=50)
    assert np.array_equal(clusters, ref["clusters"])
    assert np.array_equal(aims, ref["aims"][ref["aims"] < nsnp])
    out = plink.PlinkFileset(des)
    assert out.n_snps == aims.size
    assert np.array_equal(out.codes(), codes[np.sort(aims)])
    assert np.isfinite(d.ll_new) and d.q.shape == (K, I)


@pytest.mark.parametrize("mode,tries", [("global", 2), ("local", 1)])
def test_skfr_path_matches_oracle(mode, tries):
    """hk_skfr_path (sparsekmeans_path): decreasing sparsities, each started on the device from the previous
    labels; labels, markers, iterations and WSS equal the CPU restatement's path at every sparsity."""
    I, J, k = 400, 600, 4
    D, _ = planted(I, J, k, informative=range(0, J, 5), miss=0.02, seed=23)
    sp = [120, 60, 25, 8]
    print(f"skfr path {mode} tries={tries}", flush=True)
    ref = so.skfr_path(D, k, sp, mode=mode, tries=tries, seed=3, max_iter=50)
    with engine_for(D, K=k) as eng:
        got = eng.skfr_path(k, sp, mode=mode, max_iter=50, tries=tries, seed=3)
        with pytest.raises(HipError):
            eng.skfr_path(k, [10, 20])
    for g, r in zip(got, ref):
        assert np.array_equal(g["clusters"], r["clusters"])
        assert np.array_equal(g["aims"], r["aims"])
        assert g["iters"] == r["iters"]
        assert g["wss"] == pytest.approx(r["wss"], rel=1e-12, abs=0.0)


def test_filter_skfr_path_writes_one_fileset_per_sparsity(tmp_path):
    """_filter_SKFR with a decreasing vector of sparsities (driver.jl:108-131): one filtered fileset per
    sparsity, named {prefix}_{K}_{s}aims, with the path's markers."""
    from haplo_amd.admix import _filter_SKFR

    I, J, K = 300, 200, 3
    sim = simulate(I, J, K, fst=0.1, miss=0.01, seed=6)
    path = str(tmp_path / "d.bed")
    plink.write_bed(path, sim.bed, I, 2 * J)
    outs, clusters, aims = _filter_SKFR(path, K, [40, 20], rng=9, tries=2)
    D = so.dosages(plink.PlinkFileset(path).codes())
    ref = so.skfr_path(D, K, [40, 20], tries=2, seed=9, max_iter=50)
    for o, cl, am, r, s in zip(outs, clusters, aims, ref, (40, 20)):
        assert o.endswith(f"d_{K}_{s}aims.bed") and os.path.exists(o)
        assert np.array_equal(cl, r["clusters"]) and np.array_equal(am, r["aims"])
        assert plink.PlinkFileset(o).n_snps == am.size
