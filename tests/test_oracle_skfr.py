"""Pins of the SKFR restatement (oracle/skfr_oracle.py) on the CPU.

SparseKmeansFeatureRanking.jl is not vendored with the reference, so the oracle is checked against what the
method fixes: the splitmix64 published vector, Lloyd's k-means when every block is kept (scikit-learn),
the fixed-point conditions of the returned solution (brute-force loops over individuals), and recovery of
planted informative blocks.
"""
import numpy as np
import pytest

import skfr_oracle as so


def planted(I, J, k, informative, miss=0.0, seed=0, delta=0.35):
    """Dosages (I x 2J, -1 missing) of k equal-sized populations whose allele frequencies differ only on the
    blocks listed in `informative`; (D, population labels)."""
    rng = np.random.default_rng(seed)
    pop = np.arange(I) % k
    rng.shuffle(pop)
    base = rng.uniform(0.2, 0.8, size=2 * J)
    f = np.tile(base, (k, 1))
    for b in informative:
        for a in (0, 1):
            f[:, 2 * b + a] = np.clip(base[2 * b + a] + rng.uniform(-delta, delta, size=k), 0.02, 0.98)
    D = rng.binomial(2, f[pop]).astype(np.int8)
    if miss > 0:
        D[rng.random(D.shape) < miss] = -1
    return D, pop


def agreement(a, b):
    """Fraction of individuals whose label matches under the best relabelling of a."""
    from itertools import permutations

    a = np.asarray(a)
    b = np.asarray(b)
    k = max(a.max(), b.max()) + 1
    return max(np.mean(np.asarray(p)[a] == b) for p in permutations(range(k)))


def test_splitmix64_published_vector():
    r = so.SplitMix64(1234567)
    assert [r.next() for _ in range(5)] == [6457827717110365317, 3203168211198807973, 9817491932198370423,
                                            4593380528125082431, 16408922859458223821]


def test_all_blocks_kept_is_lloyd_kmeans():
    """sparsity = J and complete data: SKFR is Lloyd's k-means on the standardised dosages."""
    sk = pytest.importorskip("sklearn.cluster")
    D, _ = planted(240, 150, 3, informative=range(40), seed=1)
    init = [3, 77, 150]
    r = so.skfr(D, 3, 150, init=init, max_iter=300)
    _, _, Z = so.standardise(D)
    km = sk.KMeans(n_clusters=3, init=Z[init], n_init=1, algorithm="lloyd", tol=0.0, max_iter=300).fit(Z)
    assert np.array_equal(r["clusters"], km.labels_)
    assert r["wss"] == pytest.approx(km.inertia_, rel=1e-9)
    assert np.array_equal(np.sort(r["aims"]), np.arange(300))


def _bruteforce_fixed_point(D, res, k, s, local):
    """Check with per-individual Python loops that (labels, kept blocks) is a fixed point of the method:
    centers from the labels rank the kept blocks at the top, and every individual sits in a cluster of
    minimal cost (missing entries imputed with its own cluster's center)."""
    I, S = D.shape
    J = S // 2
    labels = [int(c) for c in res["clusters"]]
    mean, sd = [], []
    for j in range(S):
        vals = [int(D[i, j]) for i in range(I) if D[i, j] >= 0]
        m = sum(vals) / len(vals) if vals else 0.0
        v = sum(x * x for x in vals) / len(vals) - m * m if vals else 0.0
        mean.append(m)
        sd.append(v ** 0.5 if v > 0 else 0.0)
    theta = [[0.0] * S for _ in range(k)]
    size = [labels.count(c) for c in range(k)]
    for c in range(k):
        for j in range(S):
            vals = [int(D[i, j]) for i in range(I) if labels[i] == c and D[i, j] >= 0]
            if vals and sd[j] > 0:
                theta[c][j] = (sum(vals) / len(vals) - mean[j]) / sd[j]
    aims = [int(a) for a in res["aims"]]
    assert aims[1::2] == [a + 1 for a in aims[0::2]] and all(a % 2 == 0 for a in aims[0::2])
    order = [a // 2 for a in aims[0::2]]
    kept = sorted(order)
    score = [[size[c] * (theta[c][2 * b] ** 2 + theta[c][2 * b + 1] ** 2) for b in range(J)] for c in range(k)]
    total = [sum(score[c][b] for c in range(k)) for b in range(J)]
    # markers come most important first
    assert all(total[x] >= total[y] - 1e-12 * max(1.0, total[x]) for x, y in zip(order, order[1:]))
    rows = score if local else [[sum(score[c][b] for c in range(k)) for b in range(J)]]
    keep = [[False] * J for _ in range(k)]
    for ci, row in enumerate(rows):
        top = sorted(range(J), key=lambda b: (-row[b], b))[:s]
        lo = min(row[b] for b in top)
        hi = max([row[b] for b in range(J) if b not in top] or [-1.0])
        assert lo >= hi - 1e-12 * max(1.0, abs(lo))
        for c in (range(k) if not local else [ci]):
            for b in top:
                keep[c][b] = True
    assert kept == sorted(set(b for c in range(k) for b in range(J) if keep[c][b]))
    for c in range(k):
        for b in range(J):
            if not keep[c][b]:
                theta[c][2 * b] = theta[c][2 * b + 1] = 0.0
    for i in range(I):
        cost = []
        for c in range(k):
            t = 0.0
            for b in kept:
                for j in (2 * b, 2 * b + 1):
                    z = (D[i, j] - mean[j]) / sd[j] if (D[i, j] >= 0 and sd[j] > 0) else (
                        0.0 if D[i, j] >= 0 else theta[labels[i]][j])
                    t += (z - theta[c][j]) ** 2
            cost.append(t)
        assert cost[labels[i]] <= min(cost) + 1e-9 * max(1.0, min(cost))


@pytest.mark.parametrize("local", [False, True])
def test_solution_is_a_fixed_point(local):
    k, s = 3, 5
    D, _ = planted(45, 24, k, informative=[1, 4, 9, 16, 20], miss=0.05, seed=7)
    r = so.skfr(D, k, s, mode="local" if local else "global", seed=11, max_iter=100)
    assert r["iters"] < 100
    _bruteforce_fixed_point(D, r, k, s, local)


@pytest.mark.parametrize("mode", ["global", "local"])
def test_planted_blocks_are_selected(mode):
    """Populations differ on 20 of 120 blocks: those are the markers, and the populations come back (up to
    the few individuals 40 SNPs cannot place)."""
    inf = list(range(0, 120, 6))
    D, pop = planted(300, 120, 3, informative=inf, miss=0.02, seed=5, delta=0.4)
    r = so.skfr(D, 3, len(inf), mode=mode, seed=3, tries=3)
    assert agreement(r["clusters"], pop) >= 0.98
    blocks = sorted(set(int(a) // 2 for a in r["aims"]))
    if mode == "global":
        assert blocks == inf
    else:
        assert set(inf) <= set(blocks)


def test_tries_keep_the_best_run():
    D, _ = planted(150, 60, 4, informative=range(12), miss=0.03, seed=9, delta=0.2)
    one = so.skfr(D, 4, 10, seed=21, tries=1)
    many = so.skfr(D, 4, 10, seed=21, tries=4, return_history=True)
    h = many["history"]
    assert np.array_equal(h[0]["clusters"], one["clusters"]) and h[0]["wss"] == one["wss"]
    assert many["wss"] == min(x["wss"] for x in h)
    assert many["wss"] <= one["wss"]


def test_chunked_distances_match_direct_sums():
    """The chunked, ordered distance sum equals a plain per-individual distance up to rounding."""
    rng = np.random.default_rng(2)
    J = 2 * so.CHUNK + 37
    Z0 = rng.standard_normal((9, 2 * J))
    rows = rng.standard_normal((3, 2 * J))
    d = so._chunked_dist(Z0, rows, J)
    ref = ((Z0[:, None, :] - rows[None, :, :]) ** 2).sum(axis=2)
    assert np.allclose(d, ref, rtol=1e-12)


def test_invalid_arguments():
    D, _ = planted(20, 10, 2, informative=[0], seed=1)
    for kw in (dict(k=0, sparsity=2), dict(k=17, sparsity=2), dict(k=2, sparsity=11), dict(k=2, sparsity=0)):
        with pytest.raises(ValueError):
            so.skfr(D, kw["k"], kw["sparsity"])
    with pytest.raises(ValueError):
        so.skfr(D, 2, 2, mode="other")


def test_path_warm_starts_from_previous_labels():
    """skfr_path: the first sparsity is skfr with its tries; each next one is skfr started from the previous
    result's labels; a start from a fixed point at the same sparsity returns it after one iteration."""
    I, J, k = 300, 120, 3
    D, _ = planted(I, J, k, informative=range(0, 40), miss=0.02, seed=21)
    path = so.skfr_path(D, k, [40, 20, 10], tries=2, seed=5)
    first = so.skfr(D, k, 40, tries=2, seed=5)
    assert np.array_equal(path[0]["clusters"], first["clusters"]) and np.array_equal(path[0]["aims"], first["aims"])
    for prev, cur, s in zip(path, path[1:], (20, 10)):
        ref = so.skfr(D, k, s, init_labels=prev["clusters"])
        assert np.array_equal(cur["clusters"], ref["clusters"]) and np.array_equal(cur["aims"], ref["aims"])
        assert cur["aims"].size == 2 * s
    again = so.skfr(D, k, 10, init_labels=path[2]["clusters"])
    assert again["iters"] == 1 and np.array_equal(again["clusters"], path[2]["clusters"])
