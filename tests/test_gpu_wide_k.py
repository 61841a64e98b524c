"""GPU parity for 9 <= K <= 16: passes over windows of packed curvature columns, the P-side EM pass on the
matrix cores and the dense solver for every P row QP, against the CPU restatement (which has no K limit
below 32).  Tolerances as for K <= 8: FP64 <= 1e-12 per element, MIXED <= 2e-5, row steps <= 1e-9 / 1e-8,
FP64 fit traces <= 1e-10 per iteration."""
import numpy as np
import pytest

from gen import case, rel
from haplo_amd import HipAdmixData, fit_batch
from haplo_amd._native import HipError
from haplo_amd.sim import init_params, simulate

pytestmark = pytest.mark.gpu

WIDE = [(130, 70, 9), (260, 100, 10), (101, 45, 12), (77, 40, 16)]


def _eng(s, q, p, prec="fp64"):
    e = HipAdmixData(s.bed, s.I, s.J, s.K, precision=prec)
    e.set_params(q, p)
    return e


@pytest.mark.parametrize("I,J,K", WIDE)
def test_wide_fp64_passes_match_oracle(I, J, K, oracle):
    print(f"wide fp64 passes I={I} J={J} K={K}", flush=True)
    s, q, p = case(I, J, K)
    with _eng(s, q, p) as e:
        ll, rows = e.loglikelihood(rows=True)
        assert abs(ll - oracle.loglik(s.bed, I, J, q, p)) <= 1e-12 * abs(ll)
        assert rel(rows, oracle.loglik_rows(s.bed, I, J, q, p)) <= 1e-12
        XtX, Xtz = e.update_q()
        oX, oz = oracle.update_q(s.bed, I, J, q, p)
        assert rel(XtX, oX) <= 1e-12 and rel(Xtz, oz) <= 1e-12
        XtX, Xtz = e.update_p()
        oX, oz = oracle.update_p(s.bed, I, J, q, p)
        assert rel(XtX, oX) <= 1e-12 and rel(Xtz, oz) <= 1e-12
        qn, pn = e.em()
        oq, op = oracle.em(s.bed, I, J, q, p)
        assert rel(qn, oq) <= 1e-12 and rel(pn, op) <= 1e-12


@pytest.mark.parametrize("I,J,K", WIDE)
def test_wide_mixed_passes_match_oracle(I, J, K, oracle):
    print(f"wide mixed passes I={I} J={J} K={K}", flush=True)
    s, q, p = case(I, J, K)
    with _eng(s, q, p, "mixed") as e:
        ll = e.loglikelihood()
        assert abs(ll - oracle.loglik(s.bed, I, J, q, p)) <= 2e-6 * abs(ll)
        XtX, Xtz = e.update_q()
        oX, oz = oracle.update_q(s.bed, I, J, q, p)
        assert rel(XtX, oX) <= 2e-5 and rel(Xtz, oz) <= 2e-5
        XtX, Xtz = e.update_p()
        oX, oz = oracle.update_p(s.bed, I, J, q, p)
        assert rel(XtX, oX) <= 2e-5 and rel(Xtz, oz) <= 2e-5
        qn, pn = e.em()
        oq, op = oracle.em(s.bed, I, J, q, p)
        assert rel(qn, oq) <= 2e-5 and rel(pn, op) <= 2e-5


def test_wide_chunked_passes_match_oracle(oracle):
    """Several i- and j-chunks (partial sums reduced over chunks, FP32 partials flushed within a chunk) at
    K = 11 against the oracle."""
    I, J, K = 700, 300, 11
    print(f"wide chunked I={I} J={J} K={K}", flush=True)
    s, q, p = case(I, J, K, miss=0.02)
    with _eng(s, q, p, "mixed") as e:
        e.set_tuning(5, 3)
        XtX, Xtz = e.update_p()
        oX, oz = oracle.update_p(s.bed, I, J, q, p)
        assert rel(XtX, oX) <= 2e-5 and rel(Xtz, oz) <= 2e-5
        XtX, Xtz = e.update_q()
        oX, oz = oracle.update_q(s.bed, I, J, q, p)
        assert rel(XtX, oX) <= 2e-5 and rel(Xtz, oz) <= 2e-5


def _well_conditioned_blocks(XtX, K, J, ratio=1e-6):
    """Blocks whose four haplotype curvatures A_l all have min / max eigenvalue above ratio.  With few
    individuals per population some A_l are singular (a haplotype compatible with fewer individuals than K),
    and the P-side QP of such a block is flat along directions that move mass between haplotypes, where two
    correct solvers differ by far more than round-off (parity unpinned for those blocks)."""
    A = np.moveaxis(XtX.reshape(K, K, J, 4), (0, 1), (2, 3))  # J x 4 x K x K
    ev = np.linalg.eigvalsh(A)
    return np.all(ev[..., 0] > ratio * ev[..., -1], axis=1)


@pytest.mark.parametrize("I,J,K", [(400, 30, 9), (1000, 24, 12), (1500, 16, 16)])
def test_wide_row_steps_match_oracle(I, J, K, oracle):
    """update_q!/update_p! QP + step + projection: K-variable Q QPs, and 4K-variable P QPs on the four-lane
    block-structured solver (blocks with a singular haplotype curvature on the dense one-wave solver in the
    same launch), penalty and equality forms.  P results are compared on the blocks with well-conditioned
    haplotype curvature (at least 90 % of them here) and checked to be simplex points everywhere."""
    print(f"wide row steps I={I} J={J} K={K}", flush=True)
    s, q, p = case(I, J, K)
    q = oracle.project_q(q)
    p = oracle.project_p(p)
    with _eng(s, q, p) as e:
        XtX, Xtz = oracle.update_q(s.bed, I, J, q, p)
        oq, fails = oracle.q_step(XtX, Xtz, q)
        assert fails == 0
        assert rel(e.q_step(), oq, floor=1.0) <= 1e-9
        XtX, Xtz = oracle.update_p(s.bed, I, J, oq, p)
        good = _well_conditioned_blocks(XtX, K, J)
        assert good.mean() >= 0.9
        cols = np.repeat(good, 4)
        for pen in (True, False):
            op, fails = oracle.p_step(XtX, Xtz, p, penalty=pen)
            assert fails == 0
            gp = e.p_step(q_pass=oq, penalty=pen)
            assert rel(gp[:, cols], op[:, cols], floor=1.0) <= 1e-8, pen
            blk = gp.reshape(K, J, 4)
            assert np.all(blk >= 1e-5 - 1e-15) and np.max(np.abs(blk.sum(axis=2) - 1.0)) <= 1e-12


@pytest.mark.parametrize("K", [10, 14])
def test_wide_fit_first_iterations_track_oracle(K, oracle):
    """init_em! + QN iterations of the FP64 device loop, iteration by iteration against the CPU loop."""
    I, J = 300, 150
    print(f"wide fit trace K={K}", flush=True)
    s, q, p = case(I, J, K, fst=0.1, miss=0.02)
    _, _, _, _, otr = oracle.fit(s.bed, I, J, q, p, em_iters=2, n_iter=4, rtol=0.0)
    with _eng(s, q, p) as e:
        r = e.fit(em_iters=2, n_iter=4, rtol=0.0)
    assert len(r["trace"]) == len(otr) == 4
    assert rel(r["trace"], otr) <= 1e-10


def test_wide_mixed_fit_converges_to_oracle(oracle):
    """A whole MIXED fit at K = 9 reaches the FP64 CPU loop's converged log-likelihood to 1e-6."""
    I, J, K = 240, 160, 9
    print(f"wide mixed fit K={K}", flush=True)
    s, q, p = case(I, J, K, fst=0.1, miss=0.01)
    _, _, oll, _, _ = oracle.fit(s.bed, I, J, q, p, rtol=1e-8, n_iter=400)
    with _eng(s, q, p, "mixed") as e:
        r = e.fit(rtol=1e-8, n_iter=400)
    assert r["qp_failures"] == 0
    assert abs(r["ll"] - oll) <= 1e-6 * abs(oll), (r["ll"], oll)
    tr = r["trace"]
    assert np.all(np.diff(tr) >= -1e-7 * abs(tr[-1]))


def test_wide_graph_batch_and_single_fits_agree():
    """Captured-graph iterations equal eager ones bitwise, and a batch over K = 9, 16 (one of each precision)
    equals the same fits run alone."""
    I, J = 200, 120
    s = simulate(I, J, 4, fst=0.1, miss=0.02, seed=9)
    with HipAdmixData(s.bed, I, J, 12, precision="mixed") as e:
        q0, p0 = init_params(I, J, 12, seed=5)
        e.set_params(q0, p0)
        a = e.fit(n_iter=6, rtol=0.0, use_graph=0)
        e.set_params(q0, p0)
        b = e.fit(n_iter=6, rtol=0.0, use_graph=1)
        assert np.array_equal(a["trace"], b["trace"])
    Ks, precs = [9, 16], ["fp64", "mixed"]
    inits = [init_params(I, J, K, seed=40 + K) for K in Ks]
    with HipAdmixData(s.bed, I, J, Ks[0], precision=precs[0]) as base:
        engs = [base, HipAdmixData.shared(base, Ks[1], precs[1])]
        for en, (q0, p0) in zip(engs, inits):
            en.set_params(q0, p0)
        opts = [HipAdmixData.default_opts(rtol=1e-8, n_iter=60) for _ in engs]
        res, _ = fit_batch(engs, opts)
        seq = []
        for en, (q0, p0), o in zip(engs, inits, opts):
            en.set_params(q0, p0)
            seq.append(en.fit(o, trace=False))
        engs[1].close()
    for x, y in zip(res, seq):
        assert x["iters"] == y["iters"] and x["ll"] == y["ll"]


def test_wide_fp32_accumulation_mode(oracle):
    """HK_PREC_FP32 (FP32 products and FP32 sums over the whole i or j range) at K = 10: within FP32-accumulation
    error of the FP64 oracle."""
    I, J, K = 2000, 48, 10
    print(f"wide fp32 accumulation I={I} J={J} K={K}", flush=True)
    s, q, p = case(I, J, K, miss=0.02)
    with _eng(s, q, p, "fp32acc") as e:
        ll = e.loglikelihood()
        assert abs(ll - oracle.loglik(s.bed, I, J, q, p)) <= 2e-6 * abs(ll)
        XtX, Xtz = e.update_p()
        oX, oz = oracle.update_p(s.bed, I, J, q, p)
        assert rel(XtX, oX) <= 1e-3 and rel(Xtz, oz) <= 1e-3
        XtX, Xtz = e.update_q()
        oX, oz = oracle.update_q(s.bed, I, J, q, p)
        assert rel(XtX, oX) <= 1e-3 and rel(Xtz, oz) <= 1e-3


def test_skfr_cluster_limit_is_checked():
    """SKFR keeps per-lane cluster arrays of 16: k = 17 is rejected with a message, the context stays usable."""
    s = simulate(64, 40, 3, seed=2)
    with HipAdmixData(s.bed, 64, 40, 12) as e:
        with pytest.raises(HipError):
            e.skfr(17, 10)
        r = e.skfr(3, 10, max_iter=5, seed=1)
        assert len(r["clusters"]) == 64


def test_k_above_limit_is_rejected():
    s = simulate(20, 10, 2, seed=1)
    with pytest.raises(HipError):
        HipAdmixData(s.bed, 20, 10, 17)
