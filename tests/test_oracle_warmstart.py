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


@pytest.mark.parametrize("K", [9, 10])
def test_assignment_for_large_K_is_exhaustive_minimum(K):
    """Above K = 8 best_perm solves the assignment problem instead of enumerating: its permutation and value
    equal the minimum over all K! permutations (vectorised brute force), and relabelled targets are undone."""
    rng = np.random.default_rng(K)
    n = 60
    P = rng.dirichlet(np.ones(K), size=n)
    T = rng.dirichlet(np.ones(K) * 0.3, size=n)
    perm, v = ow.best_perm(P, T)
    C = -(np.log(P).T @ T)  # C[k, k2]: cost of giving P's column k the target column k2
    allp = np.array(list(itertools.permutations(range(K))), dtype=np.int8)
    vals = C[np.arange(K), allp].sum(axis=1)
    i = int(np.argmin(vals))
    assert np.array_equal(perm, allp[i]) and np.sum(vals <= vals[i] + 1e-9 * abs(vals[i])) == 1
    assert v == pytest.approx(vals[i], rel=1e-13)
    assert v == pytest.approx(-np.sum(T[:, perm] * np.log(P)), rel=1e-13)
    sigma = rng.permutation(K)
    perm2, _ = ow.best_perm(P, P[:, np.argsort(sigma)] * 0.9 + 0.1 / K)
    assert np.array_equal(perm2, sigma)
    # the identity is kept when it is optimal
    perm3, _ = ow.best_perm(P, P * 0.9 + 0.1 / K)
    assert np.array_equal(perm3, np.arange(K))


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


def _torch():
    return pytest.importorskip("torch")


@pytest.mark.parametrize("match", [False, True])
def test_forward_and_gradient_match_torch_autograd(match):
    """forward / loss_grad against torch's own softmax cross-entropy and autograd in float64, with the same
    BF16-rounded weight matrices (the gradient w.r.t. the FP32 masters is the one w.r.t. the rounded
    weights: the rounding passes the gradient straight through)."""
    torch = _torch()
    D, H, K, n = 9, 7, 4, 23
    w, offs = _net(D, H, K, 11)
    rng = np.random.default_rng(12)
    X = rng.normal(size=(n, D))
    T = rng.dirichlet(np.ones(K) * 0.7, size=n)
    P, _ = ow.forward(w, offs, D, H, K, X)
    loss, perm, g = ow.loss_grad(w, offs, D, H, K, X, T, match=match)
    W1, b1, W2, b2, W3, b3 = ow.unpack(w, offs, D, H, K)
    leaves = [torch.tensor(a, dtype=torch.float64, requires_grad=True)
              for a in (ow.bf16(W1), b1, ow.bf16(W2), b2, ow.bf16(W3), b3)]
    tW1, tb1, tW2, tb2, tW3, tb3 = leaves
    Xt = torch.tensor(X, dtype=torch.float64)
    Z = torch.relu(torch.relu(Xt @ tW1 + tb1) @ tW2 + tb2) @ tW3 + tb3
    Pt = torch.softmax(Z, dim=1)
    assert np.max(np.abs(Pt.detach().numpy() - P)) <= 1e-13
    Tt = torch.tensor(T[:, perm], dtype=torch.float64)
    lt = -(Tt * torch.log_softmax(Z, dim=1)).sum() / n
    lt.backward()
    assert abs(lt.item() - loss) <= 1e-12 * abs(loss)
    for leaf, o in zip(leaves, offs):
        gt = leaf.grad.numpy().ravel()
        assert np.max(np.abs(g[o:o + gt.size] - gt)) <= 1e-12 * max(1.0, np.max(np.abs(gt)))
    if match:  # the matched permutation is the one with the smallest torch loss among all relabellings
        Zd = torch.log_softmax(Z.detach(), dim=1)
        best = min(itertools.permutations(range(K)), key=lambda p: float(-(torch.tensor(T[:, list(p)]) * Zd).sum()))
        assert tuple(perm) == best


def test_adam_matches_torch_optimizer():
    """Three bias-corrected Adam steps against torch.optim.Adam (same lr, betas, eps) in float64."""
    torch = _torch()
    rng = np.random.default_rng(13)
    w = rng.normal(size=50)
    grads = [rng.normal(size=50) * s for s in (1.0, 0.1, 3.0)]
    t = torch.tensor(w, dtype=torch.float64, requires_grad=True)
    opt = torch.optim.Adam([t], lr=1e-3, betas=(0.9, 0.999), eps=1e-8)
    m = np.zeros_like(w)
    v = np.zeros_like(w)
    for step, g in enumerate(grads, 1):
        opt.zero_grad()
        t.grad = torch.tensor(g, dtype=torch.float64)
        opt.step()
        w, m, v = ow.adam(w, g, m, v, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8)
        assert np.max(np.abs(t.detach().numpy() - w)) <= 1e-15 * max(1.0, np.max(np.abs(w)))


@pytest.mark.parametrize("K", range(2, 17))
def test_shipped_network_has_its_training_record(K):
    """A network ships for every K from 2 to 16, each with the record of the run that produced it
    (haplo_amd.warmstart.train_default): 200 plain-method fits, the loss per epoch, the device time of each
    stage."""
    import json
    import os

    from haplo_amd import warmstart as ws

    path = ws.default_weights_path(K)
    rec = json.load(open(os.path.splitext(path)[0] + ".json"))
    z = np.load(path)
    assert rec["weights"] == os.path.basename(path) and int(z["K"]) == rec["K"] == K
    assert rec["network"].startswith(f"256-512-512-{K},")
    assert rec["datasets"]["n"] == 200 and rec["rows"] == 200 * rec["datasets"]["I"]
    hist = rec["loss_per_epoch"]
    assert len(hist) == rec["training"]["epochs"] and hist[-1] < hist[0]
    assert all(v > 0 for v in rec["seconds"].values())
