"""GPU: graph-captured iterations, shared genotype contexts, concurrent batches, autotuner, AIC selection."""
import numpy as np
import pytest

from gen import case, rel
from haplo_amd import HipAdmixData, fit_batch, select_k
from haplo_amd.sim import init_params, simulate

pytestmark = pytest.mark.gpu


def _fit(e, q, p, **kw):
    e.set_params(q, p)
    return e.fit(**kw)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_graph_replay_matches_eager(prec):
    """A QN iteration replayed from a captured hipGraph gives bitwise the eager iterates (several replays,
    with host synchronisation between them)."""
    print(f"graph vs eager {prec}", flush=True)
    s, q, p = case(300, 400, 4, fst=0.1, miss=0.02)
    with HipAdmixData(s.bed, 300, 400, 4, precision=prec) as e:
        a = _fit(e, q, p, n_iter=12, rtol=0.0, use_graph=0)
        qa, pa = e.get_params()
        b = _fit(e, q, p, n_iter=12, rtol=0.0, use_graph=1)
        qb, pb = e.get_params()
        c = _fit(e, q, p, n_iter=12, rtol=0.0, use_graph=1)  # replays a cached graph
    assert np.array_equal(a["trace"], b["trace"]) and np.array_equal(b["trace"], c["trace"])
    assert np.array_equal(qa, qb) and np.array_equal(pa, pb)


def test_shared_context_matches_own_context():
    s, q, p = case(257, 300, 3, miss=0.03)
    with HipAdmixData(s.bed, 257, 300, 5) as base:
        sh = HipAdmixData.shared(base, 3, "fp64")
        r1 = _fit(sh, q, p, n_iter=8, rtol=0.0)
        sh.close()
    with HipAdmixData(s.bed, 257, 300, 3) as own:
        r2 = _fit(own, q, p, n_iter=8, rtol=0.0)
    assert np.array_equal(r1["trace"], r2["trace"])


@pytest.mark.parametrize("rounds", [1, 0], ids=["round_graph", "event_driven"])
def test_batch_matches_sequential_fits(rounds):
    """hk_fit_batch over uneven K gives the same fits as one-at-a-time hk_fit, in both batch forms: one
    captured graph per round with device-side stopping rules (fits of 3..29 iterations, so most rounds
    replay converged fits' gated kernels), and each fit's iteration enqueued as its previous one completed."""
    I, J = 400, 500
    s = simulate(I, J, 4, fst=0.1, miss=0.02, seed=9)
    Ks = [2, 3, 5, 3]
    inits = [init_params(I, J, K, seed=50 + t) for t, K in enumerate(Ks)]
    with HipAdmixData(s.bed, I, J, 2) as base:
        engs = [base] + [HipAdmixData.shared(base, K, "mixed" if t % 2 else "fp64") for t, K in enumerate(Ks[1:], 1)]
        for e, (q0, p0) in zip(engs, inits):
            e.set_params(q0, p0)
        opts = [HipAdmixData.default_opts(rtol=1e-8, n_iter=200, batch_rounds=rounds) for _ in engs]
        res, fin = fit_batch(engs, opts)
        seq = [_fit(e, q0, p0, rtol=1e-8, n_iter=200) for e, (q0, p0) in zip(engs, inits)]
        for e in engs[1:]:
            e.close()
    for a, b in zip(res, seq):
        assert a["iters"] == b["iters"] and a["ll"] == b["ll"]
    assert np.all(fin > 0)


@pytest.mark.parametrize("K", [4, 6, 7])
def test_pass_shapes_give_identical_curvature(K):
    """Both variants of the FP32 curvature passes (register prefetch of the next staging step or not) do the same
    arithmetic in the same order: bitwise the same XtX / Xtz; sides without a second variant are rejected."""
    s, q, p = case(700, 600, K, miss=0.02)
    with HipAdmixData(s.bed, 700, 600, K, precision="mixed") as e:
        e.set_params(q, p)
        a = (e.update_q(), e.update_p())
        e.set_pass_shape(1, 1 if K <= 6 else 0)
        b = (e.update_q(), e.update_p())
        if K == 7:
            with pytest.raises(Exception):
                e.set_pass_shape(0, 1)
    for x, y in zip(a, b):
        assert np.array_equal(x[0], y[0]) and np.array_equal(x[1], y[1])


def test_autotune_keeps_results():
    s, q, p = case(600, 900, 5, miss=0.02)
    with HipAdmixData(s.bed, 600, 900, 5, precision="fp64") as e:
        e.set_params(q, p)
        ll0 = e.loglikelihood()
        X0 = e.update_p()
        qc, pc = e.autotune(reps=1)
        assert qc >= 1 and pc >= 1 and e.get_tuning() == (qc, pc)
        e.set_params(q, p)
        assert abs(e.loglikelihood() - ll0) <= 1e-12 * abs(ll0)
        X1 = e.update_p()
        assert rel(X1[0], X0[0]) <= 1e-12 and rel(X1[1], X0[1]) <= 1e-12


def test_select_k_prefers_true_k():
    """AIC over K = 2..5 x 2 restarts on data simulated with K = 3."""
    I, J = 600, 1500
    s = simulate(I, J, 3, fst=0.15, miss=0.01, seed=21)
    out = select_k(s.bed, I, J, Ks=(2, 3, 4, 5), restarts=2, precision="mixed", rtol=1e-6)
    assert len(out.table) == 8 and all(t["converged"] for t in out.table)
    lls = {K: out.best_by_k[K]["ll"] for K in (2, 3, 4, 5)}
    assert lls[3] > lls[2] and lls[4] >= lls[3] - 1e-6 * abs(lls[3])
    assert out.best_k == 3, [(K, out.best_by_k[K]["aic"]) for K in (2, 3, 4, 5)]


def test_round_graph_batch_respects_n_iter_and_modes():
    """Round-graph batch with uneven stopping: fits capped by n_iter (3 and 7 iterations), one with
    rtol 0 run to its cap, fix_q and LBQN options, an inactive fit (n_iter = 0); each equals its single fit."""
    I, J = 300, 400
    s = simulate(I, J, 3, fst=0.1, miss=0.02, seed=5)
    specs = [(2, dict(n_iter=3, rtol=0.0)), (3, dict(n_iter=7, rtol=1e-8)), (4, dict(n_iter=40, rtol=1e-9)),
             (3, dict(n_iter=5, rtol=0.0, fix_q=1)), (3, dict(n_iter=6, rtol=1e-10, qn_mode=1)),
             (2, dict(n_iter=0))]
    with HipAdmixData(s.bed, I, J, 2, precision="fp64") as base:
        engs = [base] + [HipAdmixData.shared(base, K, "fp64") for K, _ in specs[1:]]
        inits = [init_params(I, J, K, seed=70 + t) for t, (K, _) in enumerate(specs)]
        opts = [HipAdmixData.default_opts(batch_rounds=1, **kw) for _, kw in specs]
        for e, (q0, p0) in zip(engs, inits):
            e.set_params(q0, p0)
        res, _ = fit_batch(engs, opts)
        seq = [_fit(e, q0, p0, **kw) for e, (q0, p0), (_, kw) in zip(engs, inits, specs)]
        for e in engs[1:]:
            e.close()
    for a, b, (K, kw) in zip(res, seq, specs):
        assert a["iters"] == b["iters"] and a["ll"] == b["ll"] and a["converged"] == b["converged"], (K, kw)
    assert res[0]["iters"] == 3 and res[1]["iters"] <= 7 and res[5]["iters"] == 0
