"""Brute-force pins for the oracle -- TEST INFRASTRUCTURE ONLY (numpy, tiny inputs).

These re-derive the per-pair quantities of the reference's ``src/loops.jl`` from first principles,
not from its formulas, so that a wrong sign, index or dropped term in the literal restatement
(``oracle/hk_oracle.c``) fails a test:

* genotype decode from the PLINK 1 BED specification (2-bit codes, SNP-major, LSB-first);
* the likelihood of a block as a sum over ORDERED haplotype pairs (h1, h2) compatible with the two
  genotypes, divided by 2 when either SNP is heterozygous (the reference's ``@coefs_likelihood``,
  ``loops.jl:85-104``, keeps one of the two mirror orderings of a phase-ambiguous pair);
* posterior expected haplotype counts c_h (``@coefs``, ``loops.jl:114-155``) as the posterior mean
  of [h1 = h] + [h2 = h] over those pairs;
* double-missing blocks contribute nothing (weight 0, log-likelihood 0).
"""
from __future__ import annotations

import itertools

import numpy as np

# PLINK 1 BED: 00 hom allele-1, 01 missing, 10 het, 11 hom allele-2  ->  allele-2 count or None.
PLINK_CODE = {0b00: 0, 0b01: None, 0b10: 1, 0b11: 2}
HAPS = [(0, 0), (0, 1), (1, 0), (1, 1)]  # h = 2a + b, a = allele at SNP 1, b = allele at SNP 2


def decode_bed(bed: np.ndarray, I: int, nsnp: int):
    """Genotype matrix (I x nsnp) of allele-2 counts, -1 for missing, straight from the spec."""
    bps = (I + 3) // 4
    G = np.empty((I, nsnp), dtype=np.int64)
    for s in range(nsnp):
        col = bed[s * bps:(s + 1) * bps]
        for i in range(I):
            code = (int(col[i // 4]) >> (2 * (i % 4))) & 3
            g = PLINK_CODE[code]
            G[i, s] = -1 if g is None else g
    return G


def compatible_pairs(g1: int, g2: int):
    out = []
    for h1, h2 in itertools.product(range(4), range(4)):
        a1, b1 = HAPS[h1]
        a2, b2 = HAPS[h2]
        if (g1 < 0 or a1 + a2 == g1) and (g2 < 0 or b1 + b2 == g2):
            out.append((h1, h2))
    return out


def pair_loglik(g1: int, g2: int, qp: np.ndarray) -> float:
    if g1 < 0 and g2 < 0:
        return 0.0
    L = sum(qp[h1] * qp[h2] for h1, h2 in compatible_pairs(g1, g2))
    if g1 == 1 or g2 == 1:
        L /= 2.0
    return float(np.log(L))


def pair_counts(g1: int, g2: int, qp: np.ndarray) -> np.ndarray:
    c = np.zeros(4)
    if g1 < 0 and g2 < 0:
        return c
    pairs = compatible_pairs(g1, g2)
    L = sum(qp[h1] * qp[h2] for h1, h2 in pairs)
    for h1, h2 in pairs:
        w = qp[h1] * qp[h2] / L
        c[h1] += w
        c[h2] += w
    return c


def loglik(G: np.ndarray, q: np.ndarray, p: np.ndarray) -> float:
    I = G.shape[0]
    J = G.shape[1] // 2
    K = q.shape[0]
    P = p.reshape(K, 4, J, order="F").transpose(0, 2, 1)  # P[k, j, h] = p[k, 4j + h]
    tot = 0.0
    for i in range(I):
        for j in range(J):
            qp = q[:, i] @ P[:, j, :]
            tot += pair_loglik(G[i, 2 * j], G[i, 2 * j + 1], qp)
    return tot


def passes(G: np.ndarray, q: np.ndarray, p: np.ndarray):
    """Gradient/curvature sums of update_q_loop!/update_p_loop! and EM sums, from brute-force c_h."""
    I = G.shape[0]
    J = G.shape[1] // 2
    K = q.shape[0]
    P = p.reshape(K, 4, J, order="F").transpose(0, 2, 1)
    XtXq = np.zeros((K, K, I))
    Xtzq = np.zeros((K, I))
    XtXp = np.zeros((K, K, 4 * J))
    Xtzp = np.zeros((K, 4 * J))
    emq = np.zeros((K, I))
    emp = np.zeros((K, 4 * J))
    for i in range(I):
        for j in range(J):
            qp = q[:, i] @ P[:, j, :]
            c = pair_counts(G[i, 2 * j], G[i, 2 * j + 1], qp)
            for h in range(4):
                ph = P[:, j, h]
                w = c[h] / qp[h]
                v = c[h] / qp[h] ** 2
                Xtzq[:, i] += w * ph
                XtXq[:, :, i] += v * np.outer(ph, ph)
                Xtzp[:, 4 * j + h] += w * q[:, i]
                XtXp[:, :, 4 * j + h] += v * np.outer(q[:, i], q[:, i])
                emq[:, i] += w * q[:, i] * ph
                emp[:, 4 * j + h] += w * q[:, i] * ph
    return dict(XtXq=XtXq, Xtzq=Xtzq, XtXp=XtXp, Xtzp=Xtzp, emq=emq, emp=emp)


def project_simplex_bruteforce(v: np.ndarray, eps: float) -> np.ndarray:
    """Projection onto {x >= eps, sum x = 1} by enumerating which coordinates sit at the floor."""
    n = v.size
    best, bestd = None, np.inf
    for mask in itertools.product([0, 1], repeat=n):
        free = np.array(mask, dtype=bool)
        if not free.any():
            continue
        x = np.full(n, eps)
        rem = 1.0 - eps * (~free).sum()
        tau = (v[free].sum() - rem) / free.sum()
        x[free] = v[free] - tau
        if (x[free] < eps - 1e-15).any():
            continue
        d = np.sum((x - v) ** 2)
        if d < bestd:
            best, bestd = x, d
    return best


def qp_bruteforce(H, g, lo, hi, grp=None):
    """Strictly convex box(+group-sum) QP by enumerating every (free / at-lower / at-upper)
    assignment, solving the equality-constrained KKT system for each and keeping the feasible
    point of least objective.  Exponential: n <= 6."""
    n = len(g)
    groups = sorted(set(int(t) for t in grp if t >= 0)) if grp is not None else []
    best, bestf = None, np.inf
    for st in itertools.product([0, -1, 1], repeat=n):
        st = np.array(st)
        free = st == 0
        x = np.zeros(n)
        x[st == -1] = lo[st == -1]
        x[st == 1] = hi[st == 1]
        nf = int(free.sum())
        E = []
        rhs = []
        for c in groups:
            members = np.array([grp[t] == c for t in range(n)])
            e = members[free].astype(float)
            if e.sum() == 0:
                if abs(x[members].sum()) > 1e-12:
                    E = None
                    break
                continue
            E.append(e)
            rhs.append(-x[members & ~free].sum())
        if E is None:
            continue
        m = len(E)
        Hff = H[np.ix_(free, free)]
        r = g[free] + H[np.ix_(free, ~free)] @ x[~free]
        A = np.zeros((nf + m, nf + m))
        A[:nf, :nf] = Hff
        b = np.zeros(nf + m)
        b[:nf] = -r
        for t in range(m):
            A[:nf, nf + t] = E[t]
            A[nf + t, :nf] = E[t]
            b[nf + t] = rhs[t]
        try:
            sol = np.linalg.solve(A, b) if nf + m > 0 else np.zeros(0)
        except np.linalg.LinAlgError:
            continue
        x[free] = sol[:nf]
        if (x < lo - 1e-12).any() or (x > hi + 1e-12).any():
            continue
        f = 0.5 * x @ H @ x + g @ x
        if f < bestf - 1e-15:
            best, bestf = x.copy(), f
    return best
