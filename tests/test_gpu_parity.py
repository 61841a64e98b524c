"""GPU parity: the HIP engine (through the C ABI) against the CPU oracle on identical inputs.

Tolerances: FP64 engine <= 1e-12 relative on log-likelihood, gradients and curvature (the reference's
own precision, src/loops.jl in Float64); genotype decode and double-missing counts bit-exact; the MIXED
engine (FP32 per-pair products, FP64 accumulation) <= 2e-5 relative.
"""
import math

import numpy as np
import pytest

from gen import case, rel, unpack_codes
from haplo_amd import HipAdmixData

pytestmark = pytest.mark.gpu

SMALL = [(13, 7, 2), (13, 7, 3), (64, 40, 1), (50, 33, 4), (301, 130, 5), (257, 65, 6), (129, 71, 7), (77, 90, 8)]


def _eng(s, q, p, prec="fp64"):
    e = HipAdmixData(s.bed, s.I, s.J, s.K, precision=prec)
    e.set_params(q, p)
    return e


@pytest.mark.parametrize("I,J,K", SMALL)
def test_decode_and_double_missing_bitexact(I, J, K, oracle):
    print(f"decode case I={I} J={J} K={K}", flush=True)
    s, q, p = case(I, J, K, miss=0.2)
    with _eng(s, q, p) as e:
        assert np.array_equal(e.decode(), unpack_codes(s.bed, I, 2 * J))
        assert np.array_equal(e.double_missing(), oracle.count_double_missing(s.bed, I, J))


@pytest.mark.parametrize("I,J,K", SMALL)
def test_fp64_passes_match_oracle(I, J, K, oracle):
    print(f"fp64 passes I={I} J={J} K={K}", flush=True)
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
        qn2, pn2 = e.em(fix_q=True)
        assert qn2 is None and rel(pn2, op) <= 1e-12


@pytest.mark.parametrize("I,J,K", [(301, 130, 5), (257, 65, 6), (77, 90, 8), (190, 70, 4), (130, 45, 7), (64, 33, 2)])
def test_mixed_passes_match_oracle(I, J, K, oracle):
    print(f"mixed passes I={I} J={J} K={K}", flush=True)
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


@pytest.mark.parametrize("K", list(range(1, 17)))
def test_mixed_loglik_rows_every_k(K, oracle):
    """The MIXED log-likelihood pass (X = P^T Q from three BF16 terms of each FP32 parameter on the matrix
    cores, ceil(6K / 16) instructions per set) per individual against the oracle at every compiled K: FP32
    per-pair products with an FP64 product and sum keep each row within 1e-6 relative.  J = 37 leaves a
    partial last genotype word, I = 131 padding individuals in the last workgroup."""
    I, J = 131, 37
    s, q, p = case(I, J, K)
    with _eng(s, q, p, "mixed") as e:
        ll, rows = e.loglikelihood(rows=True)
    orows = oracle.loglik_rows(s.bed, I, J, q, p)
    assert rel(rows, orows) <= 1e-6
    assert abs(ll - orows.sum()) <= 1e-6 * abs(ll)


@pytest.mark.parametrize("I,J,K", [(1, 1, 2), (1, 4, 3), (2, 1, 1), (5, 2, 4), (3, 9, 6), (1, 17, 8)])
def test_degenerate_shapes_match_oracle(I, J, K, oracle):
    """One individual, one SNP block, and half of all genotypes missing: every lane but a few of each pass
    is padding (individuals past I, blocks past J), so the padded words must contribute exactly nothing.
    Both engines against the oracle on every pass output."""
    print(f"degenerate shape I={I} J={J} K={K}", flush=True)
    s, q, p = case(I, J, K, miss=0.5)
    olr = oracle.loglik_rows(s.bed, I, J, q, p)
    oXq, ozq = oracle.update_q(s.bed, I, J, q, p)
    oXp, ozp = oracle.update_p(s.bed, I, J, q, p)
    for prec, tol, ltol in (("fp64", 1e-12, 1e-12), ("mixed", 2e-5, 1e-6)):
        with _eng(s, q, p, prec) as e:
            ll, rows = e.loglikelihood(rows=True)
            assert rel(rows, olr) <= ltol, prec
            assert abs(ll - olr.sum()) <= ltol * abs(olr.sum()), prec
            XtX, Xtz = e.update_q()
            assert rel(XtX, oXq) <= tol and rel(Xtz, ozq) <= tol, prec
            XtX, Xtz = e.update_p()
            assert rel(XtX, oXp) <= tol and rel(Xtz, ozp) <= tol, prec
            assert np.array_equal(e.double_missing(), oracle.count_double_missing(s.bed, I, J))


def _random_shape(seed):
    r = np.random.default_rng(1000 + seed)
    return (int(r.integers(1, 701)), int(r.integers(1, 301)), int(r.integers(1, 17)),
            float(r.choice([0.0, 0.05, 0.3])), float(r.choice([0.02, 0.1, 0.3])))


@pytest.mark.parametrize("seed", range(16))
def test_random_shapes_match_oracle(seed, oracle):
    """Seeded random shapes (I up to 700, J up to 300, K 1..16, missingness 0 / 5 / 30 %, Fst 0.02-0.3):
    log-likelihood rows, both curvature passes and EM of both engines against the oracle, element by element."""
    I, J, K, miss, fst = _random_shape(seed)
    print(f"random shape seed={seed} I={I} J={J} K={K} miss={miss} fst={fst}", flush=True)
    s, q, p = case(I, J, K, seed=50 + seed, miss=miss, fst=fst, init_seed=70 + seed)
    olr = oracle.loglik_rows(s.bed, I, J, q, p)
    oXq, ozq = oracle.update_q(s.bed, I, J, q, p)
    oXp, ozp = oracle.update_p(s.bed, I, J, q, p)
    oq, op = oracle.em(s.bed, I, J, q, p)
    for prec, tol, ltol in (("fp64", 1e-12, 1e-12), ("mixed", 2e-5, 1e-6)):
        with _eng(s, q, p, prec) as e:
            ll, rows = e.loglikelihood(rows=True)
            assert rel(rows, olr) <= ltol and abs(ll - olr.sum()) <= ltol * abs(olr.sum()), prec
            XtX, Xtz = e.update_q()
            assert rel(XtX, oXq) <= tol and rel(Xtz, ozq) <= tol, prec
            XtX, Xtz = e.update_p()
            assert rel(XtX, oXp) <= tol and rel(Xtz, ozp) <= tol, prec
            if prec == "fp64":
                qn, pn = e.em()
                assert rel(qn, oq) <= 1e-12 and rel(pn, op) <= 1e-12


def test_fp32_accumulation_mode(oracle):
    """The reference's T2 = Float32 semantics (HK_PREC_FP32): FP32 products and FP32 gradient / curvature
    sums, each over the whole i or j range.  Its passes stay within FP32-accumulation error of the FP64
    oracle, and on the P-side curvature (a sum over 20,000 individuals) the MIXED engine -- FP64
    accumulation of the same FP32 products -- is clearly closer, which is the point of the comparison."""
    I, J, K = 20000, 64, 4
    print(f"fp32 accumulation I={I} J={J} K={K}", flush=True)
    s, q, p = case(I, J, K, miss=0.02)
    oX, oz = oracle.update_p(s.bed, I, J, q, p)
    oXq, ozq = oracle.update_q(s.bed, I, J, q, p)
    oll = oracle.loglik(s.bed, I, J, q, p)
    errs = {}
    for prec in ("fp32acc", "mixed"):
        with _eng(s, q, p, prec) as e:
            assert abs(e.loglikelihood() - oll) <= 2e-6 * abs(oll)
            XtX, Xtz = e.update_p()
            XtXq, Xtzq = e.update_q()
            errs[prec] = rel(XtX, oX)
            assert max(rel(Xtz, oz), rel(XtXq, oXq), rel(Xtzq, ozq)) <= (1e-3 if prec == "fp32acc" else 2e-5)
    assert errs["fp32acc"] <= 1e-3 and errs["mixed"] <= 2e-5, errs
    assert errs["mixed"] * 4 <= errs["fp32acc"], errs


@pytest.mark.parametrize("I,J,K", [(13, 7, 2), (50, 33, 4), (301, 130, 5), (257, 65, 6), (77, 90, 8)])
def test_row_steps_match_oracle(I, J, K, oracle):
    """update_q!/update_p! QP + step + projection on the device vs the CPU restatement."""
    print(f"row steps I={I} J={J} K={K}", flush=True)
    s, q, p = case(I, J, K)
    q = oracle.project_q(q)
    p = oracle.project_p(p)
    with _eng(s, q, p) as e:
        XtX, Xtz = oracle.update_q(s.bed, I, J, q, p)
        oq, fails = oracle.q_step(XtX, Xtz, q)
        assert fails == 0
        assert rel(e.q_step(), oq, floor=1.0) <= 1e-9
        XtX, Xtz = oracle.update_p(s.bed, I, J, oq, p)
        op, fails = oracle.p_step(XtX, Xtz, p, penalty=True)
        assert fails == 0
        assert rel(e.p_step(q_pass=oq, penalty=True), op, floor=1.0) <= 1e-8
        op2, _ = oracle.p_step(XtX, Xtz, p, penalty=False)
        assert rel(e.p_step(q_pass=oq, penalty=False), op2, floor=1.0) <= 1e-8


@pytest.mark.parametrize("K", [3, 6])
def test_p_step_four_lane_path_matches_oracle(K, oracle):
    """P-side row QPs at a size where every block's curvature is well conditioned, so the four-lanes-
    per-block solver (not its dense fallback) takes them; penalty and equality modes, parameters
    from a few fit iterations so that many bounds are active."""
    I, J = 1500, 400
    print(f"p step four-lane path I={I} J={J} K={K}", flush=True)
    s, q, p = case(I, J, K, miss=0.02)
    q, p, _, _, _ = oracle.fit(s.bed, I, J, q, p, em_iters=3, n_iter=2, rtol=0.0)
    q, p = np.asfortranarray(q), np.asfortranarray(p)
    with _eng(s, q, p) as e:
        XtX, Xtz = oracle.update_p(s.bed, I, J, q, p)
        for pen in (True, False):
            op, fails = oracle.p_step(XtX, Xtz, p, penalty=pen)
            assert fails == 0
            assert rel(e.p_step(q_pass=q, penalty=pen), op, floor=1.0) <= 1e-8, pen


def _well_conditioned_blocks(XtX, K, J, ratio=1e-6):
    """Blocks whose four haplotype curvatures A_l all have min / max eigenvalue above ratio.  With few
    individuals, rare haplotypes make some A_l singular or nearly so; the P-side QP of such a block can be
    flat or ill-conditioned along directions that move mass between haplotypes, where two correct solvers
    differ by far more than round-off (parity unpinned for those blocks)."""
    A = np.moveaxis(XtX.reshape(K, K, J, 4), (0, 1), (2, 3))  # J x 4 x K x K
    ev = np.linalg.eigvalsh(A)
    return np.all(ev[..., 0] > ratio * ev[..., -1], axis=1)


@pytest.mark.parametrize("K", [3, 6])
def test_p_step_work_queue_beyond_resident_slots(K, oracle):
    """More block QPs than four-lane slots resident on the device (J = 100,000 > 65,536 slots at K = 3,
    32,768 at K = 6), so most slots solve several quads in turn.  The result of every block is
    bitwise the one a context holding only the first 20,000 blocks (one quad per slot) gives,
    and two calls agree bitwise; against the oracle, the blocks with well-conditioned curvature agree to
    1e-8 and every block is a valid simplex point.  The dense fallback list (singular haplotype
    curvature, 64 individuals) is exercised in the same calls."""
    I, J, Js = 64, 100000, 20000
    print(f"p step work queue I={I} J={J} K={K}", flush=True)
    s, q, p = case(I, J, K, miss=0.02, fst=0.1)
    q = oracle.project_q(q)
    p = oracle.project_p(p)
    with _eng(s, q, p) as e:
        a = e.p_step(q_pass=q, penalty=True)
        b = e.p_step(q_pass=q, penalty=True)
        tuning = e.get_tuning()
    assert np.array_equal(a, b)
    bps = (I + 3) // 4
    with HipAdmixData(s.bed[: 2 * Js * bps], I, Js, K, precision="fp64") as e:
        e.set_tuning(*tuning)  # same i-chunks: the same curvature sums, bit for bit
        e.set_params(q, p[:, : 4 * Js])
        assert np.array_equal(e.p_step(q_pass=q, penalty=True), a[:, : 4 * Js])
    XtX, Xtz = oracle.update_p(s.bed, I, J, q, p)
    op, fails = oracle.p_step(XtX, Xtz, p, penalty=True)
    assert fails == 0
    good = _well_conditioned_blocks(XtX, K, J)
    assert good.mean() > 0.5
    cols = np.repeat(good, 4)
    assert rel(a[:, cols], op[:, cols], floor=1.0) <= 1e-8
    blk = a.reshape(K, J, 4)
    assert np.all(blk >= 1e-5 - 1e-15) and np.max(np.abs(blk.sum(axis=2) - 1.0)) <= 1e-12


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_fit_matches_oracle_fit(prec, oracle):
    """init_em! + admixture_qn! on the device vs the CPU restatement of the same loop."""
    I, J, K = 200, 300, 3
    print(f"fit parity {prec}", flush=True)
    s, q, p = case(I, J, K, fst=0.1, miss=0.01)
    oq, op, oll, oit, otr = oracle.fit(s.bed, I, J, q, p, rtol=1e-9, n_iter=300)
    with _eng(s, q, p, prec) as e:
        r = e.fit(rtol=1e-9, n_iter=300)
        assert r["qp_failures"] == 0 and r["qp_iter_limit"] == 0
        tol = 1e-9 if prec == "fp64" else 1e-6
        assert abs(r["ll"] - oll) <= tol * abs(oll), (r["ll"], oll, r["iters"], oit)
        # the log-likelihood never decreases across accepted iterations beyond round-off (MIXED: the
        # FP32 per-pair likelihoods put ~1e-8 relative noise on each evaluation)
        tr = r["trace"]
        noise = 1e-9 if prec == "fp64" else 1e-7
        assert np.all(np.diff(tr) >= -noise * abs(tr[-1]))


def test_fit_first_iterations_track_oracle(oracle):
    """Iteration-by-iteration agreement of the FP64 device loop with the CPU loop."""
    I, J, K = 150, 120, 4
    s, q, p = case(I, J, K, fst=0.1, miss=0.02)
    _, _, _, _, otr = oracle.fit(s.bed, I, J, q, p, n_iter=6, rtol=0.0)
    with _eng(s, q, p) as e:
        r = e.fit(n_iter=6, rtol=0.0)
    assert len(r["trace"]) == 6
    assert rel(r["trace"], otr) <= 1e-10


@pytest.mark.parametrize("seed", range(6))
def test_random_fits_track_oracle(seed, oracle):
    """Seeded random shapes (I up to 300, J up to 150, K 2..12): the FP64 device loop's log-likelihood trace over
    the EM warm-up and 8 QN iterations against the CPU loop, iteration by iteration."""
    r = np.random.default_rng(2000 + seed)
    I, J, K = int(r.integers(20, 301)), int(r.integers(10, 151)), int(r.integers(2, 13))
    fst, miss = float(r.choice([0.05, 0.1, 0.2])), float(r.choice([0.0, 0.02, 0.1]))
    print(f"random fit seed={seed} I={I} J={J} K={K} fst={fst} miss={miss}", flush=True)
    s, q, p = case(I, J, K, seed=90 + seed, fst=fst, miss=miss, init_seed=110 + seed)
    _, _, _, _, otr = oracle.fit(s.bed, I, J, q, p, n_iter=8, rtol=0.0)
    with _eng(s, q, p) as e:
        res = e.fit(n_iter=8, rtol=0.0)
    assert len(res["trace"]) == len(otr) == 8
    assert rel(res["trace"], otr) <= 1e-10


@pytest.mark.parametrize("opts", [dict(fix_q=True), dict(fix_p=True), dict(q_penalty=True),
                                  dict(p_penalty=False), dict(guarantee_increase=True), dict(qn_mode=1)],
                         ids=["fix_q", "fix_p", "q_penalty", "p_equality", "guarantee_increase", "lbqn"])
def test_fit_modes_track_oracle(opts, oracle):
    """admixture_qn! modes (fix_q / fix_p: algorithms_outer.jl:71-116; penalty variants: :58-64 and
    algorithms_inner.jl:158-168,295-343; guarantee_increase: algorithms_inner.jl:170-199,314-343),
    iteration by iteration against the CPU loop with the same options (FP64 engine)."""
    print(f"fit mode {opts}", flush=True)
    I, J, K = 150, 120, 4
    s, q, p = case(I, J, K, fst=0.1, miss=0.02)
    _, _, _, _, otr = oracle.fit(s.bed, I, J, q, p, n_iter=6, rtol=0.0, **opts)
    o = {k: int(v) for k, v in opts.items()}
    with _eng(s, q, p) as e:
        r = e.fit(n_iter=6, rtol=0.0, **o)
    assert r["qp_failures"] == 0
    assert len(r["trace"]) == 6
    assert rel(r["trace"], otr) <= 1e-10


def test_guarantee_increase_past_convergence(oracle):
    """Step halving where it acts: 30 iterations of a tiny problem run past convergence, so that steps
    are at round-off scale and halvings / exhausted searches occur (the oracle's composition test counts
    them).  Halving decisions compare log-likelihoods to the last bit, so device and CPU may take
    different factors there; the log-likelihood trace agrees to 1e-10 and parameters to 1e-6."""
    I, J, K = 30, 20, 2
    s, q, p = case(I, J, K, miss=0.05, fst=0.1)
    oq, op, oll, oit, otr = oracle.fit(s.bed, I, J, q, p, em_iters=0, n_iter=30, rtol=0.0, guarantee_increase=True)
    with _eng(s, q, p) as e:
        r = e.fit(em_iters=0, n_iter=30, rtol=0.0, guarantee_increase=1, verbose=1)
        gq, gp = e.get_params()
    assert len(r["trace"]) == 30
    assert rel(r["trace"], otr) <= 1e-10
    assert np.max(np.abs(gq - oq)) <= 1e-6 and np.max(np.abs(gp - op)) <= 1e-6
    tr = np.concatenate([[r["ll_init"]], r["trace"]])
    assert np.all(np.diff(tr) >= -1e-12 * abs(tr[-1]))


def test_deterministic_repeat():
    """Atomics-free reductions: two runs are bitwise identical."""
    s, q, p = case(1000, 800, 5)
    with _eng(s, q, p) as e:
        a = e.update_q()
        b = e.update_q()
        c = e.update_p()
        d = e.update_p()
        l1, l2 = e.loglikelihood(), e.loglikelihood()
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert np.array_equal(c[0], d[0]) and np.array_equal(c[1], d[1])
    assert l1 == l2


def test_error_paths_leave_context_usable(oracle):
    """Rejected calls (NaN / Inf parameters, algorithms_inner.jl:132-133; bad fit options; bad pass
    kind) return their status without touching the device state: the context's parameters and
    results afterwards are those of the last accepted call."""
    from haplo_amd import HipError
    print("error paths", flush=True)
    s, q, p = case(50, 33, 4)
    with _eng(s, q, p) as e:
        ll0 = e.loglikelihood()
        for bad_q, bad_p in ((np.nan, None), (None, np.inf)):
            qb, pb = q.copy(), p.copy()
            if bad_q is not None:
                qb[2, 17] = bad_q
            if bad_p is not None:
                pb[1, 40] = bad_p
            with pytest.raises(HipError) as ex:
                e.set_params(qb, pb)
            assert ex.value.code == -4 and "non-finite" in str(ex.value)
        q1, p1 = e.get_params()
        assert np.array_equal(q1, q) and np.array_equal(p1, p)
        with pytest.raises(HipError) as ex:
            e.fit(Qn=0)
        assert ex.value.code == -1
        with pytest.raises(HipError) as ex:
            e.time_pass(7)
        assert ex.value.code == -1
        with pytest.raises(ValueError):
            e.set_params(q[:, :-1], p)
        assert e.loglikelihood() == ll0
    assert abs(ll0 - oracle.loglik(s.bed, s.I, s.J, q, p)) <= 1e-12 * abs(ll0)


@pytest.mark.slow
def test_c2_scale_fp64_parity(oracle):
    """Single-evaluation parity at the FP64 GPU workload size: I=5000, J=50,000, K=4 (1e-12)."""
    I, J, K = 5000, 50000, 4
    print("C2-scale parity", flush=True)
    s, q, p = case(I, J, K, fst=0.05, miss=0.01, seed=2)
    with _eng(s, q, p) as e:
        ll = e.loglikelihood()
        XtXq, Xtzq = e.update_q()
        XtXp, Xtzp = e.update_p()
        assert np.array_equal(e.double_missing(), oracle.count_double_missing(s.bed, I, J))
    assert abs(ll - oracle.loglik(s.bed, I, J, q, p)) <= 1e-12 * abs(ll)
    oX, oz = oracle.update_q(s.bed, I, J, q, p)
    assert rel(XtXq, oX) <= 1e-12 and rel(Xtzq, oz) <= 1e-12
    oX, oz = oracle.update_p(s.bed, I, J, q, p)
    assert rel(XtXp, oX) <= 1e-12 and rel(Xtzp, oz) <= 1e-12


def _block_double_missing(bed, I, J):
    """Per block j, the number of individuals whose two genotypes are both missing (BED code 01), counted
    straight from the packed bytes (no decode): a slot is missing when its low bit is set and its high bit
    is clear."""
    bps = (I + 3) // 4
    tab = np.array([bin(v).count("1") for v in range(256)], dtype=np.uint8)
    b = np.asarray(bed[: 2 * J * bps]).reshape(2 * J, bps)
    out = np.zeros(J, dtype=np.int64)
    step = 20000
    for j0 in range(0, J, step):
        blk = b[2 * j0: 2 * min(J, j0 + step)]
        m = blk & ~(blk >> 1) & np.uint8(0x55)
        out[j0: j0 + blk.shape[0] // 2] = tab[m[0::2] & m[1::2]].sum(axis=1, dtype=np.int64)
    return out


@pytest.mark.slow
def test_full_size_invariants():
    """Size-independent identities at the mixed-precision workload shape (I=20k, J=200k, K=6), where the
    oracle would take hours.  The posterior counts c_h of a pair sum to 2 unless both SNPs are missing, so
      Q side, per individual i:  sum_k q_ki Xtz_q[k,i] = sum_k q_ki (XtX_q q)_ki = 2 (J - dm_i)
      P side, per block j:       sum_h sum_k p_kh Xtz_p[k,4j+h] = sum_h sum_k p_kh (XtX_p p)_kh = 2 (I - dmb_j)
    (dm_i, dmb_j: double-missing pairs of individual i / block j, dmb counted from the BED bytes here), and
    the unnormalised EM sums satisfy the same: sum_k q_next[k,i] = 2 (J - dm_i),
    sum_{k,h} p_next[k,4j+h] = 2 (I - dmb_j).  MIXED engine: 1e-5 of 2J (2I)."""
    I, J, K = 20000, 200000, 6
    print("full-size invariants", flush=True)
    s, q, p = case(I, J, K, fst=0.05, miss=0.02, seed=5)
    dmb = _block_double_missing(s.bed, I, J)
    with _eng(s, q, p, "mixed") as e:
        dm = e.double_missing()
        XtX, Xtz = e.update_q()
        lhs = np.einsum("ki,ki->i", q, Xtz)
        assert np.max(np.abs(lhs - 2 * (J - dm)) / (2 * J)) <= 1e-5
        quad = np.einsum("ki,kli,li->i", q, XtX, q)
        assert np.max(np.abs(quad - 2 * (J - dm)) / (2 * J)) <= 1e-5
        del XtX, Xtz
        XtXp, Xtzp = e.update_p()
        lin_p = np.einsum("kc,kc->c", p, Xtzp).reshape(J, 4).sum(axis=1)
        assert np.max(np.abs(lin_p - 2 * (I - dmb)) / (2 * I)) <= 1e-5
        quad_p = np.einsum("kc,klc,lc->c", p, XtXp, p).reshape(J, 4).sum(axis=1)
        assert np.max(np.abs(quad_p - 2 * (I - dmb)) / (2 * I)) <= 1e-5
        del XtXp, Xtzp
        qn, pn = e.em()
        assert np.max(np.abs(qn.sum(axis=0) - 2 * (J - dm)) / (2 * J)) <= 1e-5
        assert np.max(np.abs(pn.reshape(K, J, 4).sum(axis=(0, 2)) - 2 * (I - dmb)) / (2 * I)) <= 1e-5
        ll = e.loglikelihood()
        assert math.isfinite(ll) and ll < 0


@pytest.mark.slow
def test_full_size_mixed_matches_fp64_engine():
    """At the mixed-precision workload shape (I=20k, J=200k, K=6), where the oracle would take hours, the MIXED
    passes against the FP64 engine (itself within 1e-12 of the oracle up to I=5k, J=50k) on the same
    parameters, element by element: log-likelihood rows and total, update_q and update_p curvature and
    gradients within 2e-5 relative (FP32 per-pair products, FP64 accumulation over 200,000 blocks or 20,000
    individuals)."""
    I, J, K = 20000, 200000, 6
    print("full-size mixed vs fp64", flush=True)
    s, q, p = case(I, J, K, fst=0.05, miss=0.02, seed=5)
    out = {}
    for prec in ("fp64", "mixed"):
        with _eng(s, q, p, prec) as e:
            ll, rows = e.loglikelihood(rows=True)
            uq = e.update_q()
            up = e.update_p()
        out[prec] = (ll, rows, uq, up)
    (l64, r64, q64, p64), (l32, r32, q32, p32) = out["fp64"], out["mixed"]
    assert abs(l32 - l64) <= 1e-8 * abs(l64)
    assert rel(r32, r64) <= 2e-6
    assert rel(q32[0], q64[0]) <= 2e-5 and rel(q32[1], q64[1]) <= 2e-5
    assert rel(p32[0], p64[0]) <= 2e-5 and rel(p32[1], p64[1]) <= 2e-5


@pytest.mark.slow
def test_genotype_streams_beyond_2_32_words():
    """I = 70,000 x J = 500,000 blocks: each packed stream holds J/8 x Ip = 4.38e9 words, past the 2^32
    work-items a single dispatch can name, so the repack kernels must stride over a capped grid.  The BED is a
    pattern with closed-form results (every genotype homozygous allele 1, the last 8 blocks missing for everyone),
    so every word of both streams is checked through the passes: double-missing counts (Q-side stream) = 8;
    with q = 1/K and p = 1/4 every observed pair has likelihood 1/16 and X_0 = 1/4, so the log-likelihood is
    I (J - 8) log(1/16) and the P-side gradient (block-major stream) is Xtz_p[k, 4j] = 2 / X_0 / K * I for the
    observed blocks, 0 for every other slot (FP64 engine, exact)."""
    I, J, K = 70000, 500000, 2
    bps = (I + 3) // 4
    print(f"streams beyond 2^32 words: I={I} J={J}", flush=True)
    bed = np.zeros(2 * J * bps, dtype=np.uint8)  # BED code 00: homozygous allele 1
    bed[2 * (J - 8) * bps:] = 0x55               # code 01 (missing) for the last 16 SNPs
    q = np.full((K, I), 1.0 / K)
    p = np.full((K, 4 * J), 0.25)
    with HipAdmixData(bed, I, J, K, precision="fp64") as e:
        del bed
        e.set_params(q, p)
        dm = e.double_missing()
        assert np.all(dm == 8)
        ll = e.loglikelihood()
        ref = I * (J - 8) * math.log(1.0 / 16.0)
        assert abs(ll - ref) <= 1e-12 * abs(ref), (ll, ref)
        _, Xtz = e.update_p()
    want = np.zeros((K, J, 4))
    want[:, : J - 8, 0] = 2.0 / 0.25 / K * I
    assert np.array_equal(Xtz.reshape(K, J, 4), want)
