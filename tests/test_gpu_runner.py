"""GPU: the runner level of the drop-in boundary -- the reference's own control flow over the device arrays.

RunnerLevel replays init_em! and admixture_qn! (src/algorithms_outer.jl:1-172) with em! / update_q! /
update_p! (src/algorithms_inner.jl:29-355) exactly as they run with d_cu !== nothing: every genotype pass
goes through the named device arrays of the C ABI (hk_buf_upload / hk_buf_download / hk_buf_copy,
hk_em_buf, hk_update_q_buf, hk_update_p_buf, hk_loglik_buf), and the host part that stays in Julia --
EM normalisation, row QPs, steps, projections, ZAL -- comes from the CPU restatement, as the Julia host code
would compute it.  Its log-likelihood trace must match the CPU restatement's whole fit to 1e-10.

Also here: the double-missing counts over the whole file (count_double_missing!(dm, g) on the full SnpArray,
src/structs.jl:169) and the EM normaliser 2(J - dm) they feed (src/algorithms_inner.jl:66).
"""
import numpy as np
import pytest

from gen import case, rel
from haplo_amd import HipAdmixData
from haplo_amd.sim import init_params, simulate

pytestmark = pytest.mark.gpu


def _flat(q, p):
    return np.concatenate([q.ravel(order="F"), p.ravel(order="F")])


class RunnerLevel:
    def __init__(self, e: HipAdmixData, oracle, dm, Qn=3):
        self.e, self.o, self.dm, self.Qn = e, oracle, dm, Qn
        self.K, self.I, self.J = e.K, e.I, e.J

    # em!(d, g; d_cu) (algorithms_inner.jl:29-84)
    def em(self, q, p):
        e = self.e
        e.buf_upload("q", q)                                   # copyto_sync!([d_cu.q, d_cu.p], ...) :36
        e.buf_upload("p", p)
        e.em_buf(fix_q=False)                                  # em!(d_cu, g_cu) :38
        qn, pn = e.buf_download("q_next"), e.buf_download("p_next")  # :42
        qn, pn = self.o.em_normalize(qn, pn, self.dm, self.J)  # :66-76
        e.buf_upload("q_next", qn)                             # :77-81
        e.buf_upload("p_next", pn)
        return qn, pn

    # init_em! (algorithms_outer.jl:1-22)
    def init_em(self, q, p, iters):
        ll = np.nan
        for _ in range(iters):
            q_next, p_next = self.em(q, p)
            p, q = p_next.copy(order="F"), q_next.copy(order="F")   # :6-7
            self.e.buf_copy("p", "p_next")                          # :9-10
            self.e.buf_copy("q", "q_next")
            ll = self.e.loglikelihood_buf("q", "p")                 # :11
            p, q = self.o.project_p(p), self.o.project_q(q)          # :15-16
        return q, p, ll

    # update_q! / update_p! with d_cu (algorithms_inner.jl:121-130, 247-256; QP and step on the host)
    def update_q(self, q, p):
        e = self.e
        e.buf_upload("q", q)
        e.buf_upload("p", p)
        e.update_q_buf()
        XtX, Xtz = e.buf_download("XtX_q"), e.buf_download("Xtz_q")
        assert np.all(np.isfinite(XtX)) and np.all(np.isfinite(Xtz))  # :132-133
        qn, fails = self.o.q_step(XtX, Xtz, q, penalty=False)
        assert fails == 0
        return qn

    def update_p(self, q, p):
        e = self.e
        e.buf_upload("q", q)
        e.buf_upload("p", p)
        e.update_p_buf()
        XtX, Xtz = e.buf_download("XtX_p"), e.buf_download("Xtz_p")
        pn, fails = self.o.p_step(XtX, Xtz, p, penalty=True)
        assert fails == 0
        return pn

    def loglik(self, q, p):  # copyto_sync!([d_cu.p, d_cu.q], ...) + loglikelihood(d_cu, g_cu, d_cu.q, d_cu.p)
        self.e.buf_upload("p", p)
        self.e.buf_upload("q", q)
        return self.e.loglikelihood_buf("q", "p")

    # admixture_qn! (algorithms_outer.jl:38-172), :ZAL mode
    def qn(self, q, p, ll_new, iters):
        K, I = self.K, self.I
        N = K * (I + 4 * self.J)
        U = np.zeros((N, self.Qn), order="F")
        V = np.zeros((N, self.Qn), order="F")
        trace = []
        for i in range(1, iters + 1):
            q_next = self.update_q(q, p)                 # :72
            p_next = self.update_p(q_next, p)            # :78
            q_next2 = self.update_q(q_next, p_next)      # :95
            p_next2 = self.update_p(q_next2, p_next)     # :101
            ll_basic = self.loglik(q_next2, p_next2)     # :119-126
            x, x1, x2 = _flat(q, p), _flat(q_next, p_next), _flat(q_next2, p_next2)
            xt, rc = self.o.zal(U, V, x, x1, x2, i, self.Qn)  # :129-133
            assert rc == 0
            q_tmp = xt[: K * I].reshape((K, I), order="F")
            p_tmp = xt[K * I:].reshape((K, 4 * self.J), order="F")
            p_tmp, q_tmp = self.o.project_p(p_tmp), self.o.project_q(q_tmp)  # :143-144
            ll_qn = self.loglik(q_tmp, p_tmp)             # :146-153
            if ll_basic < ll_qn:                          # :154-160
                q, p, ll_new = q_tmp, p_tmp, ll_qn
            else:
                q, p, ll_new = q_next2, p_next2, ll_basic
            trace.append(ll_new)
        return q, p, np.array(trace)


@pytest.mark.parametrize("I,J,K", [(150, 120, 4), (301, 130, 6)])
def test_runner_level_replay_matches_oracle_fit(I, J, K, oracle):
    print(f"runner level I={I} J={J} K={K}", flush=True)
    s, q0, p0 = case(I, J, K, fst=0.1, miss=0.02)
    oq, op, oll, oit, otr = oracle.fit(s.bed, I, J, q0, p0, em_iters=2, n_iter=5, rtol=0.0)
    with HipAdmixData(s.bed, I, J, K, precision="fp64") as e:
        rl = RunnerLevel(e, oracle, e.double_missing())
        q, p, ll = rl.init_em(np.asfortranarray(q0), np.asfortranarray(p0), 2)
        q, p, tr = rl.qn(q, p, ll, 5)
    assert oit == 5
    assert rel(tr, otr) <= 1e-10
    assert rel(q, oq, floor=1.0) <= 1e-9 and rel(p, op, floor=1.0) <= 1e-9


def test_named_arrays_roundtrip_copy_and_passes(oracle):
    """Upload / download / device copy of each array; the passes on the named arrays equal the host-pointer
    runners (hk_update_q / hk_update_p / hk_em / hk_loglik) bit for bit."""
    I, J, K = 97, 53, 3
    s, q, p = case(I, J, K)
    with HipAdmixData(s.bed, I, J, K, precision="fp64") as e:
        for name in HipAdmixData.BUFFERS:
            a = np.random.default_rng(len(name)).random(e.buf_shape(name))
            e.buf_upload(name, a)
            assert np.array_equal(e.buf_download(name), a), name
        e.buf_upload("q", q)
        e.buf_upload("p_next", p)
        e.buf_copy("p", "p_next")
        assert np.array_equal(e.buf_download("p"), p)
        with pytest.raises(Exception):
            e.buf_copy("p", "q")
        e.update_q_buf()
        e.update_p_buf()
        ll_buf = e.loglikelihood_buf("q", "p")
        XtXq, Xtzq = e.buf_download("XtX_q"), e.buf_download("Xtz_q")
        XtXp, Xtzp = e.buf_download("XtX_p"), e.buf_download("Xtz_p")
        e.em_buf()
        qn, pn = e.buf_download("q_next"), e.buf_download("p_next")
        e.set_params(q, p)
        assert ll_buf == e.loglikelihood()
        A, B = e.update_q()
        assert np.array_equal(A, XtXq) and np.array_equal(B, Xtzq)
        A, B = e.update_p()
        assert np.array_equal(A, XtXp) and np.array_equal(B, Xtzp)
        A, B = e.em()
        assert np.array_equal(A, qn) and np.array_equal(B, pn)
    oX, _ = oracle.update_q(s.bed, I, J, q, p)
    assert rel(XtXq, oX) <= 1e-12


def test_double_missing_over_whole_file_and_em_normaliser(oracle):
    """A file with more SNP pairs than the fit uses: the context's counts cover the fit's J blocks; after
    hk_count_double_missing_file they cover every complete pair of the file (an odd last SNP is left out),
    as AdmixData2's count_double_missing!(dm, g) does; one EM update then normalises Q by 2(J - dm_file),
    which the CPU restatement reproduces with the same counts and not with the first-J ones."""
    I, Jf, J, K = 203, 900, 300, 3
    s = simulate(I, Jf, K, fst=0.1, miss=0.3, seed=12)
    bps = (I + 3) // 4
    q0, p0 = init_params(I, J, K, seed=4)
    dm_fit = oracle.count_double_missing(s.bed, I, J)
    dm_file = oracle.count_double_missing(s.bed, I, Jf)
    dm_odd = oracle.count_double_missing(s.bed, I, Jf - 1)
    assert np.any(dm_file != dm_fit)
    with HipAdmixData(s.bed, I, J, K, bytes_per_snp=bps, n_snp=2 * Jf, precision="fp64") as e:
        assert np.array_equal(e.double_missing(), dm_fit)
        assert np.array_equal(e.count_double_missing_file(s.bed, 2 * Jf - 1), dm_odd)
        assert np.array_equal(e.count_double_missing_file(s.bed, 2 * Jf), dm_file)
        e.set_params(q0, p0)
        e.fit(em_iters=1, n_iter=0)
        gq, _ = e.get_params()
    qn, pn = oracle.em(s.bed, I, J, q0, p0)
    q_file, _ = oracle.em_normalize(qn, pn, dm_file, J)
    q_fit, _ = oracle.em_normalize(qn, pn, dm_fit, J)
    assert rel(gq, oracle.project_q(q_file), floor=1.0) <= 1e-12
    assert np.max(np.abs(gq - oracle.project_q(q_fit))) > 1e-6
