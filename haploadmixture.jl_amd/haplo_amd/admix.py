"""Reference-shaped entry points: ``run_admixture``, ``_admixture_base``, ``AdmixData2``.

Mirrors src/driver.jl:38-201 (keyword names, defaults and control flow) with the likelihood engine on
the MI355X.  Differences, each forced by the build:
  * ``rng`` is an integer seed of the build's counter RNG (the reference takes a Julia AbstractRNG);
  * ``use_gpu=False`` is rejected: the CPU path of the reference is not part of this library;
  * SKFR labels and markers are 0-based (``skfr`` module);
  * ``T2=np.float32`` selects the MIXED engine (FP32 per-pair products, FP64 accumulation), the
    analogue of the reference's ``T2=Float32`` GPU arrays (driver.jl:149, cuda/structs.jl:19-30).
Addition: ``warm_start=True`` fills the ``init_q`` hook (driver.jl:56,184-185) with the learned
initialiser (``warmstart``): Q0 from the shipped network for this K, then EM on P with Q0 held instead
of ``init_em!``.  Without a network for K, or when ``init_q`` is given, the usual start runs.
``warm_start="skfr"`` fills it from sparse k-means labels on the panel instead (``warmstart.skfr_q0``; any K).
"""
from __future__ import annotations

import math

import numpy as np

from . import plink, skfr, warmstart
from .engine import HipAdmixData
from .sim import init_params


class AdmixData2:
    """Host view of the fit state (src/structs.jl:14-139): q (K x I), p (K x 4J), x = [q p], lls."""

    def __init__(self, I, J, K, Q=3, rng=1):
        self.I, self.J, self.K, self.Q = int(I), int(J), int(K), int(Q)
        self.q, self.p = init_params(self.I, self.J, self.K, seed=int(rng))
        self.doublemissing = None
        self.ll_prev = math.nan
        self.ll_new = math.nan
        self.ll_tmp = math.nan
        self.trace = np.zeros(0)
        self.iters = 0
        self.converged = False

    @property
    def x(self):
        return np.concatenate([self.q, self.p], axis=1)


def _admixture_base(filename, I, J, K, *, n_iter=1000, rtol=1e-7, rng=1, em_iters=5, T=np.float64,
                    T2=np.float64, Q=3, use_gpu=True, verbose=False, fix_q=False, fix_p=False,
                    init_q=None, init_p=None, engine_opts=None, warm_start=False, guarantee_increase=False,
                    mode="ZAL", doublemissing="file"):
    """src/driver.jl:133-201.  guarantee_increase and mode (:ZAL / :LBQN) are admixture_qn!'s keywords
    (algorithms_outer.jl:39-41).  doublemissing: "file" (the reference: AdmixData2 counts double-missing
    pairs over every SNP pair of the file, structs.jl:169, and the EM normaliser 2(J - dm) uses them even
    when the fit reads only the first 2J SNPs) or "fit" (over the fit's J blocks only)."""
    if doublemissing not in ("file", "fit"):
        raise ValueError("doublemissing must be 'file' or 'fit'")
    if mode not in ("ZAL", "LBQN"):
        raise ValueError("mode must be 'ZAL' or 'LBQN'")
    if not use_gpu:
        raise NotImplementedError("this library is the GPU engine; use_gpu=False has no implementation here")
    if T is not np.float64:
        raise NotImplementedError("parameters and QP run in FP64 (T=Float64)")
    precision = "mixed" if T2 is np.float32 else "fp64"
    print("Loading genotype data...")
    payload, n_ind, n_snps, bps = plink.read_bed(filename)
    if I > n_ind or 2 * J > n_snps:
        raise ValueError(f"file has {n_ind} individuals and {n_snps} SNPs; asked for I={I}, 2J={2 * J}")
    print(f"Loaded {I} samples and {2 * J} SNPs")
    d = AdmixData2(I, J, K, Q, rng=rng)
    eng = HipAdmixData(np.asarray(payload[: 2 * J * bps]), I, J, K, bytes_per_snp=bps, n_snp=2 * J,
                       precision=precision)
    try:
        if doublemissing == "file" and n_snps > 2 * J:
            eng.count_double_missing_file(np.asarray(payload), n_snps, bytes_per_snp=bps)
        d.doublemissing = eng.double_missing()
        eng.set_params(d.q, d.p)
        base = dict(Qn=Q, fix_q=int(fix_q), fix_p=int(fix_p), verbose=int(verbose),
                    guarantee_increase=int(guarantee_increase), qn_mode=int(mode == "LBQN"))
        if engine_opts:
            base.update(engine_opts)
        ll_start = math.nan
        if warm_start not in (False, True, "network", "skfr"):
            raise ValueError("warm_start must be False, True / 'network' or 'skfr'")
        use_net = warm_start in (True, "network") and init_q is None
        mlp = warmstart.load_default(K) if use_net else None
        if use_net and mlp is None:
            print(f"No warm-start network for K={K}; starting from random Q")
        q_warm = None
        if mlp is not None:
            q_warm = warmstart.predict_q(eng, mlp)
            mlp.close()
        elif warm_start == "skfr" and init_q is None:
            q_warm = warmstart.skfr_q0(eng, seed=rng)
        if q_warm is not None:
            # learned or SKFR-label Q0, then 2 * em_iters EM updates of P with Q0 held (each costs one P pass
            # and one log-likelihood pass, a third of a full EM update), in place of init_em!
            eng.set_params(q_warm, d.p)
            r = eng.fit(eng.default_opts(**{**base, "em_iters": max(em_iters, 1) * 2, "em_fix_q": 1, "n_iter": 0,
                                            "fix_q": 0, "fix_p": 0}))
            ll_start = r["ll"]
        # init_em! unless both init_q and init_p are given (driver.jl:181-183); it runs without fix_q/fix_p.
        elif init_q is None or init_p is None:
            r = eng.fit(eng.default_opts(**{**base, "em_iters": em_iters, "n_iter": 0, "fix_q": 0, "fix_p": 0}))
            ll_start = r["ll"]
        if init_q is not None or init_p is not None:
            q, p = eng.get_params()
            if init_q is not None:
                q = np.asfortranarray(init_q, dtype=np.float64)
            if init_p is not None:
                p = np.asfortranarray(init_p, dtype=np.float64)
            eng.set_params(q, p)
        # admixture_qn! (driver.jl:195-197): starts from d.ll_new as left by init_em! (NaN -> evaluated).
        r = eng.fit(eng.default_opts(**{**base, "em_iters": 0, "n_iter": n_iter, "rtol": rtol}), ll_start=ll_start)
        d.q, d.p = eng.get_params()
        d.ll_new = r["ll"]
        d.trace = r["trace"]
        d.iters = r["iters"]
        d.converged = bool(r["converged"])
        d.ll_prev = d.trace[-2] if d.trace.size > 1 else r["ll_init"]
        d.fit_info = r
    finally:
        eng.close()
    return d


def _filter_SKFR(filename, K, sparsity, *, rng=1, prefix=None, tries=10, max_inner_iter=50, ftn=None):
    """src/driver.jl:84-131: SKFR on every individual and SNP of the fileset, then des.bed/.bim/.fam with
    the selected SNPs, des = f"{prefix}_{K}_{2 * sparsity}aims".  Returns (des + ".bed", clusters, aims);
    for a decreasing list of sparsities, lists of each (des = f"{prefix}_{K}_{s}aims", as the reference
    names them)."""
    if not str(filename).endswith(".bed"):
        raise ValueError("filename should end with .bed")
    prefix = str(filename)[:-4] if prefix is None else prefix
    ftn = skfr.sparsekmeans1 if ftn is None else ftn
    many = not np.isscalar(sparsity)
    sparsities = [int(s) for s in (sparsity if many else [sparsity])]
    if many and sparsities != sorted(sparsities, reverse=True):
        raise ValueError("sparsities should be decreasing")
    fs = plink.PlinkFileset(str(filename))
    outs, clusters, aims = [], [], []
    with skfr.ImputedSnpMatrix(fs, K, rng=rng, blocksize=2) as ism:
        if many:  # driver.jl:108-131: sparsekmeans_path, each sparsity started from the previous labels
            cls, ams = skfr.sparsekmeans_path(ism, sparsities, iter=tries, max_inner_iter=max_inner_iter)
            runs = list(zip(sparsities, cls, ams))
        elif tries == 1:
            cl, _, am, _, _ = ftn(ism, sparsities[0], max_iter=max_inner_iter, squares=False)
            runs = [(sparsities[0], cl, am)]
        else:
            cl, _, am, _, _, _ = skfr.sparsekmeans_repeat(ism, sparsities[0], iter=tries,
                                                          max_inner_iter=max_inner_iter, ftn=ftn)
            runs = [(sparsities[0], cl, am)]
        for s, cl, am in runs:
            des = f"{prefix}_{K}_{s if many else 2 * s}aims"
            print(des)
            outs.append(plink.filter(str(filename), np.ones(fs.n_individuals, dtype=bool), np.sort(am), des))
            clusters.append(cl)
            aims.append(am)
    if many:
        return outs, clusters, aims
    return outs[0], clusters[0], aims[0]


def run_admixture(filename, I, J, K, *, rng=1, sparsity=None, prefix=None, skfr_tries=1, skfr_max_inner_iter=50,
                  skfr_mode="global", admix_n_iter=1000, admix_rtol=1e-7, admix_n_em_iters=5, T=np.float64,
                  T2=np.float64, Q=3, use_gpu=True, verbose=False, progress_bar=False, fix_q=False, fix_p=False,
                  init_q=None, init_p=None, warm_start=False, guarantee_increase=False, doublemissing="file"):
    """src/driver.jl:38-79.  Returns (d, clusters, aims); clusters/aims are None without SKFR.
    doublemissing: see _admixture_base ("file" = the reference's whole-file counts)."""
    if not str(filename).endswith(".bed"):
        raise ValueError("filename should end with .bed")
    print(f"Using {filename} as input.")
    if sparsity is not None:
        if skfr_mode == "global":
            ftn = skfr.sparsekmeans1
        elif skfr_mode == "local":
            ftn = skfr.sparsekmeans2
        else:
            raise ValueError("skfr_mode can only be 'global' or 'local'")
        admix_input, clusters, aims = _filter_SKFR(filename, K, sparsity, rng=rng, prefix=prefix, tries=skfr_tries,
                                                   max_inner_iter=skfr_max_inner_iter, ftn=ftn)
    else:
        admix_input, clusters, aims = filename, None, None
    d = _admixture_base(admix_input, I, J, K, n_iter=admix_n_iter, rtol=admix_rtol, rng=rng,
                        em_iters=admix_n_em_iters, T=T, T2=T2, Q=Q, use_gpu=use_gpu, verbose=verbose,
                        fix_q=fix_q, fix_p=fix_p, init_q=init_q, init_p=init_p, warm_start=warm_start,
                        guarantee_increase=guarantee_increase, doublemissing=doublemissing)
    return d, clusters, aims
