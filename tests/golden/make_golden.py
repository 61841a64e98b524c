"""Writes tests/golden/*.npz: tiny-case vectors computed by first-principles brute force.

The reference's only golden values (test/runtests.jl:8,10) need Julia, StableRNG streams and SnpArrays'
EUR_subset.bed, none of which exist here, so the committed vectors come from ``oracle/brute.py``:
enumeration of ordered haplotype pairs for the likelihood and posterior counts, exhaustive active-set
enumeration for the QPs, exhaustive floor-set enumeration for the simplex projection.  Inputs are
drawn by the build's seeded generator.  Run: ``python tests/golden/make_golden.py``.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "haploadmixture.jl_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import brute  # noqa: E402
from haplo_amd.sim import init_params, simulate  # noqa: E402

CASES = [
    # name, I, J, K, missing rate, fst, seed
    ("tiny_I13_J7_K2", 13, 7, 2, 0.15, 0.1, 101),
    ("tiny_I9_J5_K3_miss", 9, 5, 3, 0.45, 0.2, 102),
    ("tiny_I4_J3_K1", 4, 3, 1, 0.10, 0.1, 103),
    ("tiny_I21_J11_K4", 21, 11, 4, 0.05, 0.05, 104),
]


def passes_case(name, I, J, K, miss, fst, seed):
    s = simulate(I, J, K, fst=fst, miss=miss, seed=seed)
    q, p = init_params(I, J, K, seed=seed + 1000)
    G = brute.decode_bed(s.bed, I, 2 * J)
    P = p.reshape(K, 4, J, order="F").transpose(0, 2, 1)
    rows = np.array([sum(brute.pair_loglik(G[i, 2 * j], G[i, 2 * j + 1], q[:, i] @ P[:, j, :]) for j in range(J))
                     for i in range(I)])
    b = brute.passes(G, q, p)
    dm = np.array([sum(1 for j in range(J) if G[i, 2 * j] < 0 and G[i, 2 * j + 1] < 0) for i in range(I)])
    np.savez(os.path.join(HERE, name + ".npz"), I=I, J=J, K=K, bed=s.bed, q=q, p=p, G=G, ll=rows.sum(),
             ll_rows=rows, dm=dm, **b)


def qp_cases(seed=7, n_cases=40):
    rng = np.random.default_rng(seed)
    recs = []
    for t in range(n_cases):
        n = int(rng.integers(2, 6))
        A = rng.standard_normal((n, n))
        H = A @ A.T + 0.05 * np.eye(n)
        g = rng.standard_normal(n) * 2.0
        x0 = rng.dirichlet(np.ones(n))
        lo, hi = -x0, 1.0 - x0
        grp = np.zeros(n, dtype=np.int32) if t % 2 == 0 else -np.ones(n, dtype=np.int32)
        x = brute.qp_bruteforce(H, g, lo, hi, grp if t % 2 == 0 else None)
        recs.append((n, H, g, lo, hi, grp, x))
    N = max(r[0] for r in recs)
    pad = lambda v: np.pad(v, (0, N - v.size))  # noqa: E731
    np.savez(os.path.join(HERE, "qp_bruteforce.npz"),
             n=np.array([r[0] for r in recs]),
             H=np.stack([np.pad(r[1], ((0, N - r[0]), (0, N - r[0]))) for r in recs]),
             g=np.stack([pad(r[2]) for r in recs]), lo=np.stack([pad(r[3]) for r in recs]),
             hi=np.stack([pad(r[4]) for r in recs]),
             grp=np.stack([np.pad(r[5], (0, N - r[0]), constant_values=-1) for r in recs]),
             x=np.stack([pad(r[6]) for r in recs]))


def projection_cases(seed=8, n_cases=60):
    rng = np.random.default_rng(seed)
    vs, xs, ns = [], [], []
    for t in range(n_cases):
        n = int(rng.integers(2, 7))
        v = rng.standard_normal(n) * 0.6 + 1.0 / n
        if t % 5 == 0:
            v[0] = 1e-7  # a coordinate that must land on the floor
        vs.append(np.pad(v, (0, 6 - n)))
        xs.append(np.pad(brute.project_simplex_bruteforce(v, 1e-5), (0, 6 - n)))
        ns.append(n)
    np.savez(os.path.join(HERE, "projection_bruteforce.npz"), n=np.array(ns), v=np.stack(vs), x=np.stack(xs))


if __name__ == "__main__":
    for c in CASES:
        passes_case(*c)
    qp_cases()
    projection_cases()
    print("wrote", sorted(f for f in os.listdir(HERE) if f.endswith(".npz")))
