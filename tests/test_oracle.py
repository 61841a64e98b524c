"""Pins of the CPU oracle (oracle/hk_oracle.c) -- runs without a GPU.

Each check is chosen so that a plausible slip in the literal restatement of src/loops.jl (a dropped
term, a wrong index or sign, a transposed operand) fails: brute-force enumeration of ordered haplotype
pairs, exhaustive QP / projection enumeration, finite differences, and exact identities.
"""
import glob
import itertools
import os

import numpy as np
import pytest

from gen import case, rel, unpack_codes

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = sorted(glob.glob(os.path.join(HERE, "golden", "tiny_*.npz")))


def test_decode_table_exhaustive(brute):
    """Every byte value x every 2-bit slot against the PLINK 1 specification (loops.jl:55-83)."""
    bed = np.arange(256, dtype=np.uint8)  # one SNP column of 256 bytes = 1024 individuals
    G = brute.decode_bed(bed, 1024, 1)[:, 0]
    for byte in range(256):
        for slot in range(4):
            code = (byte >> (2 * slot)) & 3
            want = {0: 0, 1: -1, 2: 1, 3: 2}[code]
            assert G[4 * byte + slot] == want
    assert np.array_equal(unpack_codes(bed, 1024, 1)[0], [(b >> (2 * s)) & 3 for b in range(256) for s in range(4)])


@pytest.mark.parametrize("path", GOLD, ids=[os.path.basename(p) for p in GOLD])
def test_oracle_matches_golden_bruteforce(path, oracle):
    z = np.load(path)
    I, J = int(z["I"]), int(z["J"])
    bed, q, p = z["bed"], np.asfortranarray(z["q"]), np.asfortranarray(z["p"])
    assert abs(oracle.loglik(bed, I, J, q, p) - z["ll"]) <= 1e-12 * abs(z["ll"])
    assert rel(oracle.loglik_rows(bed, I, J, q, p), z["ll_rows"]) <= 1e-12
    XtX, Xtz = oracle.update_q(bed, I, J, q, p)
    assert rel(XtX, z["XtXq"]) <= 1e-12 and rel(Xtz, z["Xtzq"]) <= 1e-12
    XtX, Xtz = oracle.update_p(bed, I, J, q, p)
    assert rel(XtX, z["XtXp"]) <= 1e-12 and rel(Xtz, z["Xtzp"]) <= 1e-12
    qn, pn = oracle.em(bed, I, J, q, p)
    assert rel(qn, z["emq"]) <= 1e-12 and rel(pn, z["emp"]) <= 1e-12
    assert np.array_equal(oracle.count_double_missing(bed, I, J), z["dm"])


def test_pair_loglik_all_genotype_classes(brute):
    """The reference's branchless @coefs_likelihood equals its branchy closed form (loops.jl:178-208)
    and the ordered-pair enumeration on all 16 genotype classes."""
    rng = np.random.default_rng(0)
    for _ in range(20):
        qp = rng.dirichlet(np.ones(4))
        qp00, qp01, qp10, qp11 = qp
        closed = {
            (0, 0): 2 * np.log(qp00), (0, 2): 2 * np.log(qp01), (2, 0): 2 * np.log(qp10), (2, 2): 2 * np.log(qp11),
            (-1, 0): 2 * np.log(qp00 + qp10), (-1, 2): 2 * np.log(qp01 + qp11), (0, -1): 2 * np.log(qp00 + qp01),
            (2, -1): 2 * np.log(qp10 + qp11), (-1, 1): np.log(qp00 + qp10) + np.log(qp01 + qp11),
            (1, -1): np.log(qp00 + qp01) + np.log(qp10 + qp11), (0, 1): np.log(qp00) + np.log(qp01),
            (2, 1): np.log(qp10) + np.log(qp11), (1, 0): np.log(qp00) + np.log(qp10),
            (1, 2): np.log(qp01) + np.log(qp11), (1, 1): np.log(qp00 * qp11 + qp01 * qp10), (-1, -1): 0.0,
        }
        for (g1, g2), v in closed.items():
            assert abs(brute.pair_loglik(g1, g2, qp) - v) <= 1e-13 * max(1.0, abs(v))


def _pair_table():
    """The kernels' per-nibble table (hk_pair.h fill_pair_lut), restated for the checks below:
    nibble -> (h s1_00, h s1_01, h s1_11, e), (s2_00, s2_01, s2_11, 2 h s1_01)."""
    def snp_s(c):
        hi, lo = c >> 1, c & 1
        return np.array([float(hi ^ 1), float(hi ^ lo), float(lo)])
    t1, t2 = np.zeros((16, 4)), np.zeros((16, 4))
    for nib in range(16):
        c1, c2 = nib & 3, nib >> 2
        h = 0.5 if (c1 == 2 or c2 == 2) else 1.0
        t1[nib, :3] = h * snp_s(c1)
        t2[nib, :3] = snp_s(c2)
        t2[nib, 3] = 2 * h * snp_s(c1)[1]
        if nib == 5:  # both SNPs missing
            t1[nib] = [0.0, 0.0, 0.0, 1.0]
            t2[nib, 3] = 0.0
    return t1, t2


def _pair_form(X, nib, t1, t2):
    """PairForm of hk_pair.h: Y = S1' X S2 and L' = <X, Y> + e."""
    a, b = t1[nib], t2[nib]
    z00, z01 = X[0] * b[0] + X[1] * b[1], X[0] * b[1] + X[1] * b[2]
    z10, z11 = X[2] * b[0] + X[3] * b[1], X[2] * b[1] + X[3] * b[2]
    Y = np.array([a[0] * z00 + a[1] * z10, a[0] * z01 + a[1] * z11, a[1] * z00 + a[2] * z10, a[1] * z01 + a[2] * z11])
    return Y, float(X @ Y) + a[3]


def _pair_lik(X, nib, t1, t2):
    """pair_lik of hk_pair.h: the quadratic form G without Y."""
    a, b = t1[nib], t2[nib]
    z00, z01 = X[0] * b[0] + X[1] * b[1], X[0] * b[1] + X[1] * b[2]
    z10, z11 = X[2] * b[0] + X[3] * b[1], X[2] * b[1] + X[3] * b[2]
    g00, g01, g11 = X[0] * z00 + X[1] * z01, X[0] * z10 + X[1] * z11, X[2] * z10 + X[3] * z11
    return a[0] * g00 + b[3] * g01 + a[2] * g11 + a[3]


def _lik_table():
    """The likelihood passes' factored table (hk_pair.h fill_lik_lut), restated: nibble -> alpha, beta, (dh, e)
    with L' = (alpha . X)(beta . X) + dh (X_00 X_11 + X_01 X_10) + e."""
    def u(c, x):  # rank-one factor of a homozygous or missing code
        return 1.0 if c == 1 else float((c == 0 and x == 0) or (c == 3 and x == 1))
    al, be, de = np.zeros((16, 4)), np.zeros((16, 4)), np.zeros((16, 2))
    for nib in range(16):
        c1, c2 = nib & 3, nib >> 2
        for tab, out in ((0, al), (1, be)):
            for f in range(4):
                a, b = f >> 1, f & 1
                if nib == 5 or (c1 == 2 and c2 == 2):
                    v = 0.0
                elif c1 == 2:
                    v = float(a == tab) * u(c2, b)
                elif c2 == 2:
                    v = u(c1, a) * float(b == tab)
                else:
                    v = u(c1, a) * u(c2, b)
                out[nib, f] = v
        de[nib] = (1.0 if nib == 10 else 0.0, 1.0 if nib == 5 else 0.0)
    return al, be, de


def _lik_factored(X, nib, tab):
    """lik_factored of hk_pair.h."""
    al, be, de = tab
    return (al[nib] @ X) * (be[nib] @ X) + de[nib, 0] * (X[0] * X[3] + X[1] * X[2]) + de[nib, 1]


def test_pair_table_forms_match_bruteforce(brute):
    """Every nibble (both BED codes) and random mixtures: the table-driven per-pair forms of the kernels
    equal the ordered-haplotype-pair enumeration -- likelihood L' (with the reference's 1/2 per het SNP,
    1 for a double-missing pair) in all three of its forms (with Y, without, and the factored table), gradient weights 2 Y_h / L' = c_h / qp_h and
    curvature weights Y_h / (L' qp_h) = c_h / (2 qp_h^2) (0 for a double-missing pair) -- so no genotype
    class can drop a term."""
    geno = {0: 0, 1: -1, 2: 1, 3: 2}  # BED code -> allele-2 count (-1 missing), loops.jl:1-2
    t1, t2 = _pair_table()
    lik = _lik_table()
    rng = np.random.default_rng(11)
    for _ in range(25):
        X = rng.dirichlet(np.ones(4))
        for c1, c2 in itertools.product(range(4), repeat=2):
            nib = c1 | (c2 << 2)
            Y, L = _pair_form(X, nib, t1, t2)
            want = brute.pair_loglik(geno[c1], geno[c2], X)
            assert abs(np.log(L) - want) <= 1e-13 * max(1.0, abs(want)), (c1, c2)
            assert abs(np.log(_pair_lik(X, nib, t1, t2)) - want) <= 1e-13 * max(1.0, abs(want)), (c1, c2)
            assert abs(np.log(_lik_factored(X, nib, lik)) - want) <= 1e-13 * max(1.0, abs(want)), (c1, c2)
            c = brute.pair_counts(geno[c1], geno[c2], X)
            assert np.allclose(2 * Y / L, c / X, rtol=1e-13, atol=1e-15), (c1, c2)
            assert np.allclose(Y / (L * X), c / (2 * X * X), rtol=1e-13, atol=1e-15), (c1, c2)


def test_counts_sum_to_two(brute):
    rng = np.random.default_rng(1)
    for g1, g2 in itertools.product([-1, 0, 1, 2], repeat=2):
        c = brute.pair_counts(g1, g2, rng.dirichlet(np.ones(4)))
        assert abs(c.sum() - (0.0 if (g1 < 0 and g2 < 0) else 2.0)) < 1e-14


def test_gradient_is_finite_difference_of_loglik(oracle):
    s, q, p = case(11, 9, 3, miss=0.1)
    _, Xtzq = oracle.update_q(s.bed, 11, 9, q, p)
    _, Xtzp = oracle.update_p(s.bed, 11, 9, q, p)
    h = 1e-6
    for (k, i) in [(0, 0), (2, 5), (1, 10)]:
        qa, qb = q.copy(order="F"), q.copy(order="F")
        qa[k, i] += h
        qb[k, i] -= h
        fd = (oracle.loglik(s.bed, 11, 9, qa, p) - oracle.loglik(s.bed, 11, 9, qb, p)) / (2 * h)
        assert abs(fd - Xtzq[k, i]) <= 1e-6 * max(1.0, abs(fd))
    for (k, c) in [(0, 0), (2, 13), (1, 35)]:
        pa, pb = p.copy(order="F"), p.copy(order="F")
        pa[k, c] += h
        pb[k, c] -= h
        fd = (oracle.loglik(s.bed, 11, 9, q, pa) - oracle.loglik(s.bed, 11, 9, q, pb)) / (2 * h)
        assert abs(fd - Xtzp[k, c]) <= 1e-6 * max(1.0, abs(fd))


def test_pass_identities(oracle):
    """sum_k q_ki Xtz_q[k,i] = q' XtX_q q = 2 (J - dm_i); curvature symmetric PSD; EM = param .* grad."""
    I, J, K = 40, 30, 4
    s, q, p = case(I, J, K, miss=0.1)
    dm = oracle.count_double_missing(s.bed, I, J)
    XtX, Xtz = oracle.update_q(s.bed, I, J, q, p)
    assert np.allclose(np.einsum("ki,ki->i", q, Xtz), 2 * (J - dm), rtol=1e-13)
    assert np.allclose(np.einsum("ki,kli,li->i", q, XtX, q), 2 * (J - dm), rtol=1e-13)
    assert np.allclose(XtX, XtX.transpose(1, 0, 2), rtol=1e-14, atol=0)
    assert min(np.linalg.eigvalsh(XtX[:, :, i]).min() for i in range(I)) > -1e-9
    qn, pn = oracle.em(s.bed, I, J, q, p)
    assert rel(qn, q * Xtz) <= 1e-13
    XtXp, Xtzp = oracle.update_p(s.bed, I, J, q, p)
    assert rel(pn, p * Xtzp) <= 1e-13


def test_qp_matches_bruteforce_fixture(oracle):
    z = np.load(os.path.join(HERE, "golden", "qp_bruteforce.npz"))
    for t in range(len(z["n"])):
        n = int(z["n"][t])
        grp = z["grp"][t][:n]
        x, it = oracle.qp_solve(z["H"][t][:n, :n], z["g"][t][:n], z["lo"][t][:n], z["hi"][t][:n],
                                grp if (grp >= 0).any() else None)
        assert it >= 0
        assert np.max(np.abs(x - z["x"][t][:n])) <= 1e-9


def test_qp_kkt_random(oracle):
    """KKT of the box + sum-zero QP at the returned point, including P-style penalty matrices."""
    rng = np.random.default_rng(3)
    for t in range(60):
        K = int(rng.integers(1, 5))
        n = 4 * K if t % 2 else K + 2
        A = rng.standard_normal((n, 2 * n))
        H = A @ A.T / n + (1e3 * np.kron(np.ones((4, 4)), np.eye(K)) if t % 2 else 0)
        g = rng.standard_normal(n) * 5
        x0 = rng.uniform(1e-5, 1, n)
        lo, hi = -x0, 1 - x0
        grp = None if t % 2 else np.zeros(n, dtype=np.int32)
        x, it = oracle.qp_solve(H, g, lo, hi, grp)
        assert it >= 0
        assert np.all(x >= lo - 1e-12) and np.all(x <= hi + 1e-12)
        r = H @ x + g
        lam = 0.0
        free = (x > lo + 1e-10) & (x < hi - 1e-10)
        if grp is not None:
            assert abs(x.sum()) <= 1e-10
            lam = -r[free].mean() if free.any() else 0.0
        kk = r + lam
        scale = 1e-7 * (np.abs(g).max() + np.abs(H).max())
        assert np.all(np.abs(kk[free]) <= scale)
        assert np.all(kk[(x <= lo + 1e-10) & ~free] >= -scale)
        assert np.all(kk[(x >= hi - 1e-10) & ~free] <= scale)


@pytest.mark.parametrize("K", [12, 16])
def test_wide_k_gradient_identities_and_p_qp_kkt(oracle, K):
    """K above 8 (the GPU's windowed passes and dense P row QPs are checked against these): the gradients are
    finite differences of the log-likelihood, sum_k q_ki Xtz_q[k,i] = q' XtX_q q = 2 (J - dm_i), and the
    4K-variable P row QP (penalty and K group equalities) satisfies its KKT conditions."""
    I, J = 23, 9
    s, q, p = case(I, J, K, miss=0.1)
    XtXq, Xtzq = oracle.update_q(s.bed, I, J, q, p)
    XtXp, Xtzp = oracle.update_p(s.bed, I, J, q, p)
    h = 1e-6
    for (k, i) in [(0, 0), (K - 1, 7), (K // 2, 22)]:
        qa, qb = q.copy(order="F"), q.copy(order="F")
        qa[k, i] += h
        qb[k, i] -= h
        fd = (oracle.loglik(s.bed, I, J, qa, p) - oracle.loglik(s.bed, I, J, qb, p)) / (2 * h)
        assert abs(fd - Xtzq[k, i]) <= 1e-6 * max(1.0, abs(fd))
    for (k, c) in [(0, 0), (K - 1, 13), (K // 2, 35)]:
        pa, pb = p.copy(order="F"), p.copy(order="F")
        pa[k, c] += h
        pb[k, c] -= h
        fd = (oracle.loglik(s.bed, I, J, q, pa) - oracle.loglik(s.bed, I, J, q, pb)) / (2 * h)
        assert abs(fd - Xtzp[k, c]) <= 1e-6 * max(1.0, abs(fd))
    dm = oracle.count_double_missing(s.bed, I, J)
    assert np.allclose(np.einsum("ki,ki->i", q, Xtzq), 2 * (J - dm), rtol=1e-12)
    assert np.allclose(np.einsum("ki,kli,li->i", q, XtXq, q), 2 * (J - dm), rtol=1e-12)
    rng = np.random.default_rng(K)
    n = 4 * K
    for t in range(6):
        A = rng.standard_normal((n, 2 * n))
        pen = t % 2 == 0
        H = A @ A.T / n + (1e3 * np.kron(np.ones((4, 4)), np.eye(K)) if pen else 0)
        g = rng.standard_normal(n) * 5
        x0 = rng.uniform(1e-5, 1, n)
        lo, hi = -x0, 1 - x0
        grp = None if pen else (np.arange(n) % K).astype(np.int32)
        x, it = oracle.qp_solve(H, g, lo, hi, grp)
        assert it >= 0
        assert np.all(x >= lo - 1e-12) and np.all(x <= hi + 1e-12)
        r = H @ x + g
        free = (x > lo + 1e-10) & (x < hi - 1e-10)
        lam = np.zeros(n)
        if grp is not None:
            for c in range(K):
                m = grp == c
                assert abs(x[m].sum()) <= 1e-10
                f = m & free
                lam[m] = -r[f].mean() if f.any() else 0.0
        kk = r + lam
        scale = 1e-7 * (np.abs(g).max() + np.abs(H).max())
        assert np.all(np.abs(kk[free]) <= scale)
        assert np.all(kk[(x <= lo + 1e-10) & ~free] >= -scale)
        assert np.all(kk[(x >= hi - 1e-10) & ~free] <= scale)


def test_projection_matches_bruteforce_fixture(oracle):
    z = np.load(os.path.join(HERE, "golden", "projection_bruteforce.npz"))
    for t in range(len(z["n"])):
        n = int(z["n"][t])
        x = oracle.project_simplex(z["v"][t][:n], 1e-5)
        assert np.max(np.abs(x - z["x"][t][:n])) <= 1e-12
        assert abs(x.sum() - 1) <= 1e-14 and x.min() >= 1e-5 - 1e-18


def test_projection_idempotent_and_docs_values(oracle):
    """Projected columns stay put; a one-hot column becomes (1e-5, ..., 1 - (K-1) 1e-5) as printed in
    docs/src/index.md:300-303 (0.99997 for K = 4)."""
    x = oracle.project_simplex(np.array([0.0, 0.0, 1.0, 0.0]), 1e-5)
    assert np.allclose(x, [1e-5, 1e-5, 0.99997, 1e-5], rtol=0, atol=1e-15)
    rng = np.random.default_rng(5)
    for _ in range(20):
        v = oracle.project_simplex(rng.standard_normal(6), 1e-5)
        assert np.array_equal(oracle.project_simplex(v, 1e-5), v) or rel(oracle.project_simplex(v, 1e-5), v, floor=1.0) < 1e-15


def test_zal_closed_form(oracle):
    """ZAL extrapolation against x1 - V (U'U - U'V)^{-1} U'(x0 - x1), with the column cycling of
    update_UV! (iterations 1..5, Q = 3)."""
    rng = np.random.default_rng(9)
    N, Qn = 50, 3
    U = np.zeros((N, Qn), order="F")
    V = np.zeros((N, Qn), order="F")
    Uref = np.zeros((N, Qn))
    Vref = np.zeros((N, Qn))
    for it in range(1, 6):
        x0, x1, x2 = rng.standard_normal((3, N))
        out, rc = oracle.zal(U, V, x0, x1, x2, it, Qn)
        col = (it - 1) % Qn
        Uref[:, col] = x1 - x0
        Vref[:, col] = x2 - x1
        nc = min(it, Qn)
        Up, Vp = Uref[:, :nc], Vref[:, :nc]
        want = x1 - Vp @ np.linalg.solve(Up.T @ Up - Up.T @ Vp, Up.T @ (x0 - x1))
        assert rc == 0
        assert rel(out, want, floor=1.0) <= 1e-10


@pytest.mark.parametrize("mode", [0, 1], ids=["ZAL", "LBQN"])
def test_quasi_newton_exact_on_affine_maps(oracle, mode):
    """Both quasi-Newton modes are multisecant Broyden updates from H0 = -I of the inverse Jacobian of
    G(x) = F(x) - x.  For an affine contraction F(x) = A x + b the secants satisfy V = A U exactly, so once
    the U columns span the space the update lands on the fixed point (I - A)^-1 b -- for ZAL through
    (U'(I - A)U)^-1, for LBQN through ((A - I)U)'((A - I)U) -- whatever the iteration point.  A dropped
    term, a sign or a transposed product breaks this."""
    rng = np.random.default_rng(5)
    n = 3
    A = rng.normal(size=(n, n))
    A *= 0.6 / np.max(np.abs(np.linalg.eigvals(A)))
    b = rng.normal(size=n)
    xstar = np.linalg.solve(np.eye(n) - A, b)
    U = np.zeros((n, 3), order="F")
    V = np.zeros((n, 3), order="F")
    for it in (1, 2, 3):
        x0 = rng.normal(size=n)
        x1 = A @ x0 + b
        x2 = A @ x1 + b
        out, rc = oracle.qn_step(U, V, x0, x1, x2, it, Qn=3, mode=mode)
        assert rc == 0
    assert np.max(np.abs(out - xstar)) <= 1e-10 * max(1.0, np.max(np.abs(xstar)))
    # with fewer secants than dimensions the step is not exact (the test above is not vacuous)
    U1 = np.zeros((n, 3), order="F")
    V1 = np.zeros((n, 3), order="F")
    x0 = rng.normal(size=n)
    out1, _ = oracle.qn_step(U1, V1, x0, A @ x0 + b, A @ (A @ x0 + b) + b, 1, Qn=3, mode=mode)
    assert np.max(np.abs(out1 - xstar)) > 1e-6


def test_lbqn_fit_reaches_the_zal_optimum(oracle):
    """admixture_qn!(mode=:LBQN) on the CPU restatement: the log-likelihood never decreases across
    iterations (the accept rule keeps the better of the plain and the extrapolated point) and the fit
    converges to the same optimum as the default ZAL mode."""
    I, J, K = 120, 90, 3
    s, q, p = case(I, J, K, fst=0.1, miss=0.02)
    _, _, ll_zal, _, _ = oracle.fit(s.bed, I, J, q, p, rtol=1e-10, n_iter=500)
    _, _, ll_lb, it_lb, tr = oracle.fit(s.bed, I, J, q, p, rtol=1e-10, n_iter=500, qn_mode=1)
    assert np.all(np.diff(tr) >= -1e-9 * abs(tr[-1]))
    assert abs(ll_lb - ll_zal) <= 1e-7 * abs(ll_zal), (ll_lb, ll_zal, it_lb)


def test_oracle_fit_monotone_and_converges(oracle):
    """init_em! + admixture_qn! on a small simulated set: accepted ll never decreases, converges."""
    I, J, K = 60, 80, 3
    s, q, p = case(I, J, K, miss=0.02, fst=0.15)
    q1, p1, ll, iters, tr = oracle.fit(s.bed, I, J, q, p, rtol=1e-8, n_iter=500)
    assert iters < 500
    assert np.all(np.diff(tr) >= -1e-10 * abs(ll))
    assert np.allclose(q1.sum(0), 1) and q1.min() >= 1e-5 - 1e-15
    assert np.allclose(p1.reshape(K, 4, J, order="F").sum(1), 1)
    assert abs(oracle.loglik(s.bed, I, J, q1, p1) - ll) <= 1e-10 * abs(ll)


def test_equality_qp_semidefinite_curvature(oracle, brute):
    """A P-side block with an exactly singular haplotype curvature: with the K group-sum equalities the
    problem is still strictly convex on the feasible set.  Adding sigma (11' x I_K), which vanishes
    there, must return the unique optimum (checked against KKT enumeration of the unaugmented QP)."""
    rng = np.random.default_rng(11)
    K = 2
    n = 4 * K
    for t in range(10):
        H = np.zeros((n, n))
        for l in range(4):
            B = rng.standard_normal((K, 1 if l == t % 4 else 3))
            H[l * K:(l + 1) * K, l * K:(l + 1) * K] = B @ B.T  # block l rank 1: singular
        g = rng.standard_normal(n)
        x0 = rng.dirichlet(np.ones(4), size=K).T.reshape(-1)
        lo, hi = -x0, 1 - x0
        grp = np.array([i % K for i in range(n)], dtype=np.int32)
        sigma = np.diag(H).max()
        Ha = H + sigma * np.kron(np.ones((4, 4)), np.eye(K))
        x, it = oracle.qp_solve(Ha, g, lo, hi, grp)
        xb = brute.qp_bruteforce(H, g, lo, hi, grp)
        assert it >= 0
        assert np.max(np.abs(x - xb)) <= 1e-8, (t, x, xb)


def _gi_fit_py(oracle, bed, I, J, q, p, n_iter, Qn=3, eps=1e-5):
    """admixture_qn! with guarantee_increase (algorithms_outer.jl:65-171; algorithms_inner.jl:170-199,
    314-343), composed in Python from the oracle's single-step pieces; returns (q, p, trace)."""
    K = q.shape[0]
    N = K * (I + 4 * J)
    U = np.zeros((N, Qn), order="F")
    V = np.zeros((N, Qn), order="F")

    def halve(base, delta, out, llf, ll_tmp, is_q):
        f = 1.0
        while True:
            if f < 1.0:
                halvings[is_q] += 1
                out = oracle.project_q(base + f * delta) if is_q else oracle.project_p(base + f * delta)
            ll = llf(out)
            if (ll >= ll_tmp) if is_q else (ll > ll_tmp):
                return out, ll
            f /= 2
            if f < 1e-3:
                return out, ll_tmp

    halvings = {True: 0, False: 0}
    ll = lambda a, b: oracle.loglik(bed, I, J, a, b)  # noqa: E731
    ll_new = ll(q, p)
    trace = []
    for it in range(1, n_iter + 1):
        ll_tmp = ll_new
        qn, dq, _ = oracle.q_step_delta(*oracle.update_q(bed, I, J, q, p), q)
        qn, ll_tmp = halve(q, dq, qn, lambda z: ll(z, p), ll_tmp, True)
        pn, dp, _ = oracle.p_step_delta(*oracle.update_p(bed, I, J, qn, p), p)
        pn, ll_tmp = halve(p, dp, pn, lambda z: ll(qn, z), ll_tmp, False)
        qn2, dq, _ = oracle.q_step_delta(*oracle.update_q(bed, I, J, qn, pn), qn)
        qn2, ll_tmp = halve(qn, dq, qn2, lambda z: ll(z, pn), ll_tmp, True)
        pn2, dp, _ = oracle.p_step_delta(*oracle.update_p(bed, I, J, qn2, pn), pn)
        pn2, ll_tmp = halve(pn, dp, pn2, lambda z: ll(qn2, z), ll_tmp, False)
        ll_basic = ll(qn2, pn2)
        flat = lambda a, b: np.concatenate([a, b], axis=1).ravel(order="F")  # noqa: E731
        xt, _ = oracle.zal(U, V, flat(q, p), flat(qn, pn), flat(qn2, pn2), it, Qn)
        xt = xt.reshape(K, I + 4 * J, order="F")
        qt, pt = oracle.project_q(xt[:, :I]), oracle.project_p(xt[:, I:])
        ll_qn = ll(qt, pt)
        if ll_basic < ll_qn:
            q, p, ll_new = qt, pt, ll_qn
        else:
            q, p, ll_new = qn2, pn2, ll_basic
        trace.append(ll_new)
    return q, p, np.array(trace), halvings


@pytest.mark.parametrize("I,J,K,n", [(40, 30, 3, 4), (30, 20, 2, 30)])
def test_guarantee_increase_loop_matches_composition(oracle, I, J, K, n):
    """hko_fit(guarantee_increase) against the same loop composed from the single-step pieces: pins which
    iterate, which step and which comparison (>= for Q, > for P) every halving uses.  The 30-iteration
    case runs past convergence, where steps are tiny and halvings (and exhausted searches) happen."""
    s, q, p = case(I, J, K, miss=0.05, fst=0.1)
    qc, pc, llc, iters, trc = oracle.fit(s.bed, I, J, q, p, em_iters=0, n_iter=n, rtol=0.0, guarantee_increase=True)
    qr, pr, trr, halvings = _gi_fit_py(oracle, s.bed, I, J, q, p, n)
    assert iters == n
    assert rel(trc, trr) <= 1e-12
    assert rel(qc, qr, floor=1.0) <= 1e-10 and rel(pc, pr, floor=1.0) <= 1e-10
    if n > 10:
        assert halvings[True] > 0 and halvings[False] > 0


def test_guarantee_increase_never_decreases(oracle):
    """With step halving every accepted iterate is at least as likely as the last (from a random start,
    no EM warm-up, where unguarded SQP steps can overshoot)."""
    for seed in (3, 4, 5):
        I, J, K = 50, 40, 4
        s, q, p = case(I, J, K, miss=0.02, fst=0.2, seed=seed)
        _, _, ll, iters, tr = oracle.fit(s.bed, I, J, q, p, em_iters=0, n_iter=30, rtol=1e-10,
                                         guarantee_increase=True)
        ll0 = oracle.loglik(s.bed, I, J, q, p)
        full = np.concatenate([[ll0], tr])
        assert np.all(np.diff(full) >= -1e-12 * abs(ll)), seed
