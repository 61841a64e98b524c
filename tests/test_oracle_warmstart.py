"""Checks that pin the NumPy warm-start restatement (oracle/warmstart_oracle.py) before it judges the kernels."""
import itertools

import numpy as np
import pytest

import warmstart_oracle as ow


def _net(D, H, K, seed):
    rng = np.random.default_rng(seed)
    sizes = [D * H, H, H * H, H, H * K, K]
    offs, o = [], 0
    for s in sizes:
        offs.append(o)
        o += (s + 63) // 64 * 64
    w = np.zeros(o)
    for o_, s in zip(offs, sizes):
        w[o_:o_ + s] = rng.normal(0, 0.7, s)
    return w, offs


@pytest.mark.parametrize("match", [False, True])
def test_gradient_matches_finite_differences(match):
    D, H, K, n = 6, 5, 3, 7
    w, offs = _net(D, H, K, 1)
    rng = np.random.default_rng(2)
    X = rng.normal(size=(n, D))
    T = rng.dirichlet(np.ones(K), size=n)
    loss, perm, g = ow.loss_grad(w, offs, D, H, K, X, T, match=match, round_bf16=False)
    idx = [i for o, s in zip(offs, [D * H, H, H * H, H, H * K, K]) for i in range(o, o + s)]
    h = 1e-6
    for i in idx:
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        # the matched permutation is held fixed: the loss is piecewise smooth in w
        lp = ow.loss_grad(wp, offs, D, H, K, X, T[:, perm] if match else T, match=False, round_bf16=False)[0]
        lm = ow.loss_grad(wm, offs, D, H, K, X, T[:, perm] if match else T, match=False, round_bf16=False)[0]
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])
    # padding between blocks carries no gradient
    mask = np.ones(w.size, bool)
    mask[idx] = False
    assert np.all(g[mask] == 0)


def test_permutation_search_is_exhaustive_minimum():
    rng = np.random.default_rng(5)
    K, n = 4, 40
    P = rng.dirichlet(np.ones(K), size=n)
    sigma = np.array([2, 0, 3, 1])
    T = P[:, np.argsort(sigma)] * 0.9 + 0.1 / K
    perm, v = ow.best_perm(P, T)
    vals = {p: -np.sum(T[:, list(p)] * np.log(P)) for p in itertools.permutations(range(K))}
    assert v == pytest.approx(min(vals.values()), rel=1e-15)
    # T's columns are P's relabelled by sigma^-1: the search undoes it
    assert np.array_equal(perm, sigma)


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7 + 2 ** -8, 1.0 + 2 ** -9, -3.0 - 2 ** -7 - 2 ** -9], np.float32)
    got = ow.bf16(x)
    # ties go to the even mantissa; below half an ulp rounds down; above rounds up
    assert list(got) == [1.0, 1.0, 1.0 + 2 ** -6, 1.0, -3.0 - 2 ** -6]


def test_gauss_matrix_moments_and_seed():
    R = ow.gauss_matrix(400, 77)
    assert abs(R.mean()) < 0.01 and abs(R.var() - 1.0) < 0.01
    assert abs(np.mean(R ** 4) - 3.0) < 0.05  # Gaussian kurtosis
    assert not np.allclose(R, ow.gauss_matrix(400, 78))


def test_features_centre_and_ignore_missing():
    rng = np.random.default_rng(4)
    S, I = 30, 9
    codes = rng.choice([0, 2, 3], size=(S, I)).astype(np.uint8)
    codes[:, 4] = codes[:, 2]          # duplicated individual
    codes[:, 6] = 1                    # everything missing
    F = ow.features(codes, S, 77)
    assert np.allclose(F[4], F[2])
    assert np.all(F[6] == 0)
    # centring: with complete data the features sum to zero over individuals
    full = np.delete(np.arange(I), 6)
    F2 = ow.features(codes[:, full], S, 77)
    assert np.allclose(F2.sum(axis=0), 0, atol=1e-12)
    assert F.shape == (I, ow.D_FEAT)
