"""Model selection over K: a grid of fits (K values x random restarts) on one dataset, scored by AIC.

The reference fits one K per ``run_admixture`` call (src/driver.jl:38) and only mentions AIC for choosing
K (README.md:9).  Here all fits share one genotype upload (``HipAdmixData.shared``), the chunk shapes of
each K's passes are chosen by the autotuner (``hk_autotune``), and the fits run concurrently
(``hk_fit_batch``): every fit's QN iteration is captured into ONE graph per round, replayed until each
fit's device-side stopping rule has switched its kernels off (with batch_rounds=False, each fit's
iteration is enqueued on its own stream as soon as its previous one completed).

AIC = -2 ll + 2 (I (K - 1) + 3 K J): I (K-1) free admixture proportions and 3 free haplotype frequencies
per (population, block).
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from .engine import HipAdmixData, fit_batch
from .sim import init_params


def n_params(I: int, J: int, K: int) -> int:
    return I * (K - 1) + 3 * K * J


def aic(ll: float, I: int, J: int, K: int) -> float:
    return -2.0 * ll + 2.0 * n_params(I, J, K)


@dataclass
class SelectResult:
    table: list = field(default_factory=list)  # one dict per fit
    best_by_k: dict = field(default_factory=dict)
    best_k: int = 0
    wall_s: float = 0.0
    tuning: dict = field(default_factory=dict)


def _seed(K: int, r: int, base: int) -> int:
    return base * 1000003 + 131 * K + r


def select_k(bed, I, J, Ks=(2, 3, 4, 5, 6, 7, 8), restarts=4, seed=1, precision="mixed", rtol=1e-7,
             n_iter=1000, em_iters=5, Q=3, autotune=True, batched=True, use_graph=True,
             batch_rounds=True) -> SelectResult:
    """batched: hk_fit_batch over all fits (batch_rounds: one captured graph per round holding every fit's
    iteration, else each fit's iteration enqueued on its own as soon as its previous one completed);
    otherwise the fits run one after another."""
    Ks = list(Ks)
    base = HipAdmixData(bed, I, J, Ks[0], precision=precision)
    engines, specs = [], []
    try:
        for K in Ks:
            for r in range(restarts):
                e = base if (K == Ks[0] and r == 0) else HipAdmixData.shared(base, K, precision)
                q0, p0 = init_params(I, J, K, seed=_seed(K, r, seed))
                e.set_params(q0, p0)
                engines.append(e)
                specs.append((K, r))
        out = SelectResult()
        if autotune:
            for K in Ks:
                first = next(e for e, (k, _) in zip(engines, specs) if k == K)
                qc, pc = first.autotune(reps=2)
                qs, ps = first.get_pass_shape()
                out.tuning[K] = (qc, pc, qs, ps)
                for e, (k, _) in zip(engines, specs):
                    if k == K and e is not first:
                        e.set_tuning(qc, pc)
                        e.set_pass_shape(qs, ps)
        opts = [HipAdmixData.default_opts(em_iters=em_iters, n_iter=n_iter, rtol=rtol, Qn=Q,
                                          use_graph=int(use_graph), batch_rounds=int(batch_rounds)) for _ in engines]
        t0 = time.perf_counter()
        if batched:
            res, fin = fit_batch(engines, opts)
        else:
            res, fin = [], []
            for e, o in zip(engines, opts):
                res.append(e.fit(o, trace=False))
                fin.append(time.perf_counter() - t0)
            fin = np.array(fin)
        out.wall_s = time.perf_counter() - t0
        for (K, r), rr, tf in zip(specs, res, fin):
            out.table.append(dict(K=K, restart=r, ll=rr["ll"], iters=rr["iters"], converged=bool(rr["converged"]),
                                  seconds=rr["seconds"], finished_at_s=float(tf), aic=aic(rr["ll"], I, J, K)))
        for K in Ks:
            rows = [t for t in out.table if t["K"] == K]
            out.best_by_k[K] = min(rows, key=lambda t: t["aic"])
        out.best_k = min(out.best_by_k.values(), key=lambda t: t["aic"])["K"]
        return out
    finally:
        for e in engines:
            if e is not base:
                e.close()
        base.close()
