"""GPU tests at the workloads of the five benchmark configurations.

C1  plain method, FP64: I=500, J=2,000, K=3, Fst 0.1, 1% missing, rtol 1e-7 -- the whole fit against the
    CPU restatement: same iteration count, log-likelihood trace <= 1e-10 per iteration, converged
    log-likelihood <= 1e-9, Q and P <= 1e-6.
C4  batched model selection: I=3,000, J=20,000, data with K=5, K = 2..8 x 4 restarts (28 fits of uneven
    K, one genotype upload): every fit of the batch equals the same fit run alone, bit for bit; AIC picks
    K=5; and an FP64 fit per K tracks the CPU restatement for one EM and two QN iterations (<= 1e-10).
C5  learned warm start on held-out I=10,000, J=100,000, K=5 data: features of a subset of individuals
    against the NumPy restatement (SNP means over all individuals; <= 2e-5 of the largest feature), and
    fewer iterations and less wall time (network time included) to 1e-6 of the plain method's converged
    log-likelihood.
(C2 and C3 are covered in test_gpu_parity.py.)
"""
import functools

import numpy as np
import pytest

import warmstart_oracle as ow
from gen import rel, unpack_codes
from haplo_amd import HipAdmixData, fit_batch
from haplo_amd import warmstart as ws
from haplo_amd.select import _seed, aic
from haplo_amd.sim import init_params, simulate

pytestmark = pytest.mark.gpu


def test_c1_fp64_fit_matches_oracle(oracle):
    I, J, K = 500, 2000, 3
    print(f"C1 fit I={I} J={J} K={K}", flush=True)
    s = simulate(I, J, K, fst=0.1, miss=0.01, seed=1)
    q0, p0 = init_params(I, J, K, seed=101)
    oq, op, oll, oit, otr = oracle.fit(s.bed, I, J, q0, p0, rtol=1e-7, n_iter=1000)
    with HipAdmixData(s.bed, I, J, K, precision="fp64") as e:
        e.set_params(q0, p0)
        r = e.fit(rtol=1e-7, n_iter=1000)
        gq, gp = e.get_params()
    assert r["converged"] and r["qp_failures"] == 0 and r["qp_iter_limit"] == 0
    assert r["iters"] == oit, (r["iters"], oit)
    assert rel(r["trace"], otr) <= 1e-10
    assert abs(r["ll"] - oll) <= 1e-9 * abs(oll)
    assert rel(gq, oq, floor=1.0) <= 1e-6 and rel(gp, op, floor=1.0) <= 1e-6


C4_I, C4_J, C4_KS, C4_RESTARTS = 3000, 20000, tuple(range(2, 9)), 4


@functools.lru_cache(maxsize=1)
def _c4_data():
    return simulate(C4_I, C4_J, 5, fst=0.05, miss=0.01, seed=4)


@pytest.mark.slow
def test_c4_batch_equals_single_fits_and_aic_picks_true_k():
    I, J = C4_I, C4_J
    print(f"C4 batch I={I} J={J} K={C4_KS} x {C4_RESTARTS}", flush=True)
    s = _c4_data()
    base = HipAdmixData(s.bed, I, J, C4_KS[0], precision="mixed")
    engines, specs, inits = [], [], []
    try:
        for K in C4_KS:
            for r in range(C4_RESTARTS):
                e = base if not engines else HipAdmixData.shared(base, K, "mixed")
                q0, p0 = init_params(I, J, K, seed=_seed(K, r, 1))
                e.set_params(q0, p0)
                engines.append(e)
                specs.append((K, r))
                inits.append((q0, p0))
        for K in C4_KS:  # per-K chunk shapes from the autotuner, as select_k does
            group = [e for e, (k, _) in zip(engines, specs) if k == K]
            qc, pc = group[0].autotune(reps=2)
            qs, ps = group[0].get_pass_shape()
            for e in group[1:]:
                e.set_tuning(qc, pc)
                e.set_pass_shape(qs, ps)
            group[0].set_params(*inits[specs.index((K, 0))])
        opts = [HipAdmixData.default_opts(rtol=1e-7, n_iter=1000) for _ in engines]
        res, fin = fit_batch(engines, opts)
        print("C4 batch done; single fits", flush=True)
        for e, (q0, p0), o, rb, sp in zip(engines, inits, opts, res, specs):
            e.set_params(q0, p0)
            ra = e.fit(o, trace=False)
            assert ra["iters"] == rb["iters"] and ra["ll"] == rb["ll"], sp
            assert rb["converged"] and rb["qp_failures"] == 0 and rb["qp_iter_limit"] == 0, sp
    finally:
        for e in engines[1:]:
            e.close()
        base.close()
    best = {K: min(aic(r["ll"], I, J, k) for r, (k, _) in zip(res, specs) if k == K) for K in C4_KS}
    assert min(best, key=best.get) == 5, best
    assert np.all(fin > 0)


@pytest.mark.slow
@pytest.mark.parametrize("K", C4_KS)
def test_c4_fp64_fit_per_k_tracks_oracle(K, oracle):
    I, J = C4_I, C4_J
    print(f"C4 oracle fit K={K}", flush=True)
    s = _c4_data()
    q0, p0 = init_params(I, J, K, seed=_seed(K, 0, 1))
    _, _, _, oit, otr = oracle.fit(s.bed, I, J, q0, p0, em_iters=1, n_iter=2, rtol=0.0)
    with HipAdmixData(s.bed, I, J, K, precision="fp64") as e:
        e.set_params(q0, p0)
        r = e.fit(em_iters=1, n_iter=2, rtol=0.0)
    assert oit == 2 and len(r["trace"]) == 2
    assert rel(r["trace"], otr) <= 1e-10


@pytest.mark.slow
def test_c5_warm_start_features_and_gain():
    mlp = ws.load_default(5)
    if mlp is None:
        pytest.skip("no trained K=5 warm-start network")
    I, J, K = 10000, 100000, 5
    print(f"C5 warm start I={I} J={J} K={K}", flush=True)
    sd = ws.simulate_world(I, J, K, 0.05, miss=0.01, data_seed=4242, truth=False)
    try:
        with HipAdmixData(sd.bed, I, J, K, precision="mixed") as e:
            nb = min(J, ws.PANEL_BLOCKS)
            F = ws.features(e)
            rows = np.arange(3, I, 41)
            ref = ow.features(unpack_codes(sd.bed, I, 2 * nb), 2 * nb, ws.FEATURE_SEED, rows=rows)
            assert np.max(np.abs(F[rows] - ref)) <= 2e-5 * np.max(np.abs(ref))
            ws.predict_q(e, mlp)  # set-up: the network's projection scratch for this panel size
            q0, p0 = init_params(I, J, K, seed=11)
            e.set_params(q0, p0)
            plain = e.fit(rtol=1e-7, n_iter=1000)
            warm = ws.fit_warm(e, mlp, rtol=1e-7, n_iter=1000)
    finally:
        mlp.close()
    ll = plain["ll"]
    ip, tp = ws.time_to_rel(plain["trace"], plain["t_trace"], ll)
    iw, tw = ws.time_to_rel(warm["trace"], warm["t_trace"], ll)
    print(f"C5 plain {ip} it {tp:.3f}s; warm {iw} it {tw}s + predict {warm['predict_seconds']:.3f}s", flush=True)
    assert iw is not None and iw < ip
    assert tw + warm["predict_seconds"] < tp
