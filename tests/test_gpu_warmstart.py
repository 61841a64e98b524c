"""GPU parity of the learned warm start (csrc/hk_mlp.hip) against the NumPy restatement
(oracle/warmstart_oracle.py), and the warm-started fit against the plain one.

Tolerances: features <= 2e-5 of their largest magnitude (FP32 sums over <= 2J terms); MLP outputs
<= 2e-5 absolute on probabilities; gradients <= 2e-5 of the largest gradient entry on samples away from
the ReLU kinks (FP32 MFMA accumulation against FP64, same BF16-rounded weights); Adam <= 1e-6 relative.
"""
import numpy as np
import pytest

import warmstart_oracle as ow
from gen import case, unpack_codes
from haplo_amd import HipAdmixData
from haplo_amd import warmstart as ws
from haplo_amd._native import HipError

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("I,J,nb", [(97, 53, 40), (97, 53, 53), (300, 20, 1), (513, 64, 64), (300, 2600, 2500)])
def test_features_match_numpy(I, J, nb):
    print(f"features I={I} J={J} nb={nb}", flush=True)
    s, q, p = case(I, J, 3, miss=0.1)
    with HipAdmixData(s.bed, I, J, 3, precision="mixed") as e:
        F = ws.features(e, n_blocks=nb, seed=77)
    ref = ow.features(unpack_codes(s.bed, I, 2 * J), 2 * nb, 77)
    assert F.shape == (I, 256)
    assert np.max(np.abs(F - ref)) <= 2e-5 * np.max(np.abs(ref))


def test_device_prediction_matches_host_path():
    """hk_mlp_predict (features on the device, fed to the network in place) equals features() + forward(),
    on two calls in a row with different sizes (the scratch is reused)."""
    m, _ = _random_mlp(256, 64, 3, seed=5)
    for I, J, nb in ((400, 300, 250), (150, 80, 80)):
        print(f"predict I={I} J={J} nb={nb}", flush=True)
        s, q, p = case(I, J, 3, miss=0.05)
        with HipAdmixData(s.bed, I, J, 3, precision="mixed") as e:
            ref = m.forward(ws.features(e, n_blocks=nb))
            got = ws.predict_probs(e, m, n_blocks=nb)
        assert got.shape == (I, 3)
        assert np.max(np.abs(got - ref)) <= 1e-6
    m.close()


def test_device_prediction_rejects_bad_inputs():
    s, q, p = case(40, 10, 2)
    m, _ = _random_mlp(37, 16, 2, seed=3)  # input width is not the feature width
    m256, _ = _random_mlp(256, 16, 2, seed=3)
    with HipAdmixData(s.bed, 40, 10, 2) as e:
        with pytest.raises(HipError, match="input width"):
            ws.predict_probs(e, m, n_blocks=10)
        with pytest.raises(HipError):
            ws.predict_probs(e, m256, n_blocks=11)
    m.close()
    m256.close()


def test_features_reject_bad_block_count():
    s, q, p = case(40, 10, 2)
    with HipAdmixData(s.bed, 40, 10, 2) as e:
        for nb in (0, 11):
            with pytest.raises(HipError):
                ws.features(e, n_blocks=nb)


def _random_mlp(D, H, K, seed):
    m = ws.MLP(K, din=D, width=H, seed=seed)
    rng = np.random.default_rng(seed)
    w = np.zeros(m.nparam, dtype=np.float32)
    for o, sz, sc in zip(m.offsets, [D * H, H, H * H, H, H * K, K],
                         [1 / np.sqrt(D), 0.1, 1 / np.sqrt(H), 0.1, 2 / np.sqrt(H), 0.3]):
        w[o:o + sz] = rng.normal(0, sc, sz)
    m.set_params(w)
    return m, w


DIMS = [(256, 512, 5, 300), (256, 48, 3, 37), (37, 40, 8, 19), (256, 512, 2, 500), (256, 64, 12, 300),
        (40, 48, 16, 70)]


@pytest.mark.parametrize("D,H,K,n", DIMS)
def test_mlp_forward_matches_numpy(D, H, K, n):
    print(f"mlp forward D={D} H={H} K={K} n={n}", flush=True)
    m, w = _random_mlp(D, H, K, 3)
    X = np.random.default_rng(4).normal(size=(n, D)).astype(np.float32)
    P = m.forward(X)
    ref, _ = ow.forward(w, m.offsets, D, H, K, X)
    assert np.max(np.abs(P - ref)) <= 2e-5
    assert np.allclose(P.sum(axis=1), 1.0, atol=1e-5)
    m.close()


def _away_from_kinks(w, offsets, D, H, K, X, n, margin=1e-4):
    """The first n rows of X whose hidden pre-activations all sit >= margin from the ReLU kink, so that
    FP32 and FP64 agree on every ReLU derivative (a unit at |z| ~ 1e-7 flips between them)."""
    W1, b1, W2, b2, _, _ = ow.unpack(w, offsets, D, H, K)
    z1 = X.astype(np.float64) @ ow.bf16(W1) + b1
    z2 = np.maximum(z1, 0) @ ow.bf16(W2) + b2
    ok = (np.abs(z1).min(axis=1) >= margin) & (np.abs(z2).min(axis=1) >= margin)
    assert ok.sum() >= n
    return X[ok][:n]


@pytest.mark.parametrize("D,H,K,n", DIMS)
@pytest.mark.parametrize("match", [0, 1])
def test_mlp_gradient_matches_numpy(D, H, K, n, match):
    print(f"mlp grad D={D} H={H} K={K} n={n} match={match}", flush=True)
    m, w = _random_mlp(D, H, K, 5)
    rng = np.random.default_rng(6)
    N = n + 23
    X = _away_from_kinks(w, m.offsets, D, H, K, rng.normal(size=(2 * N, D)).astype(np.float32), N)
    T = rng.dirichlet(np.ones(K) * 0.5, size=N).astype(np.float32)
    m.load_data(X, T)
    row0 = 23
    loss, perm = m.grad(row0, n, match=bool(match))
    _, g = m.params(grad=True)
    lref, pref, gref = ow.loss_grad(w, m.offsets, D, H, K, X[row0:row0 + n], T[row0:row0 + n], match=bool(match))
    assert np.array_equal(perm, pref)
    assert abs(loss - lref) <= 1e-5 * abs(lref)
    assert np.max(np.abs(g - gref)) <= 2e-5 * np.max(np.abs(gref))
    m.close()


@pytest.mark.parametrize("K", [5, 13])
def test_mlp_matching_recovers_relabelled_targets(K):
    """Targets whose columns are the network's own output relabelled: the matching (the enumeration at K = 5,
    the assignment solver at K = 13) returns the relabelling."""
    D, H, n = 256, 64, 200
    m, w = _random_mlp(D, H, K, 7)
    X = np.random.default_rng(8).normal(size=(n, D)).astype(np.float32)
    P = m.forward(X)
    sigma = np.array([3, 0, 4, 1, 2]) if K == 5 else np.random.default_rng(9).permutation(K)
    T = (0.8 * P[:, np.argsort(sigma)] + 0.2 / K).astype(np.float32)
    m.load_data(X, T)
    _, perm = m.grad(0, n, match=True)
    assert np.array_equal(perm, sigma)
    m.close()


def test_mlp_adam_matches_numpy():
    D, H, K, n = 256, 96, 4, 64
    m, w = _random_mlp(D, H, K, 9)
    rng = np.random.default_rng(10)
    X = rng.normal(size=(n, D)).astype(np.float32)
    T = rng.dirichlet(np.ones(K), size=n).astype(np.float32)
    m.load_data(X, T)
    wr = w.astype(np.float64)
    mr = np.zeros_like(wr)
    vr = np.zeros_like(wr)
    for step in (1, 2, 3):
        m.grad(0, n)
        _, g = m.params(grad=True)
        m.adam(lr=1e-3)
        wr, mr, vr = ow.adam(wr, g.astype(np.float64), mr, vr, step, lr=1e-3)
        wg = m.params()
        assert np.max(np.abs(wg - wr)) <= 1e-6 * np.max(np.abs(wr)), step
        wr = wg.astype(np.float64)
    m.close()


def test_mlp_rejects_bad_arguments():
    m = ws.MLP(3, din=16, width=16)
    m.load_data(np.zeros((10, 16), np.float32), np.full((10, 3), 1 / 3, np.float32))
    with pytest.raises(HipError):
        m.grad(5, 6)
    with pytest.raises(ValueError):
        m.set_params(np.zeros(3, np.float32))
    m.close()
    for K in (1, 17):  # 2 <= K <= 16
        with pytest.raises(HipError):
            ws.MLP(K, din=16, width=16)


def test_em_with_q_held_matches_oracle(oracle):
    I, J, K = 61, 37, 3
    s, q, p = case(I, J, K)
    with HipAdmixData(s.bed, I, J, K) as e:
        e.set_params(q, p)
        r = e.fit(em_iters=3, em_fix_q=1, n_iter=0)
        qg, pg = e.get_params()
    # em_p! (algorithms_inner.jl:40) three times: P-only EM update, normalisation, ll, projection
    dm = oracle.count_double_missing(s.bed, I, J)
    qr, pr = q.copy(order="F"), p.copy(order="F")
    for _ in range(3):
        _, pn = oracle.em(s.bed, I, J, qr, pr, fix_q=True)
        _, pr = oracle.em_normalize(qr.copy(order="F"), pn, dm, J)
        ll = oracle.loglik(s.bed, I, J, qr, pr)
        pr = oracle.project_p(pr)
        qr = oracle.project_q(qr)
    assert np.max(np.abs(qg - q)) <= 1e-15
    assert np.max(np.abs(qg - qr)) <= 1e-15
    assert np.max(np.abs(pg - pr)) <= 1e-12
    assert r["ll_init"] == pytest.approx(ll, rel=1e-12)


@pytest.mark.parametrize("K", range(2, 17))
def test_warm_start_reaches_plain_optimum_in_fewer_iterations(K):
    """The shipped network of each K (held-out data of its population structure): the warm-started fit ends
    at least as high as the plain one and reaches 1e-6 of the better optimum in fewer iterations."""
    import os

    path = ws.default_weights_path(K)
    if not os.path.exists(path):
        pytest.skip("no trained warm-start weights")
    mlp = ws.MLP.load(path)
    I, J = 3000, 30000
    sd = ws.simulate_world(I, J, K, 0.08, world_seed=ws.WORLD_SEED, data_seed=987654, truth=False)
    with HipAdmixData(sd.bed, I, J, K, precision="mixed") as e:
        q0, p0 = ws.init_params(I, J, K, seed=11)
        e.set_params(q0, p0)
        plain = e.fit(rtol=1e-7, n_iter=500)
        warm = ws.fit_warm(e, mlp, rtol=1e-7, n_iter=500)
    ll_ref = max(plain["ll"], warm["ll"])
    it_p, _ = ws.time_to_rel(plain["trace"], plain["t_trace"], ll_ref, 1e-6)
    it_w, _ = ws.time_to_rel(warm["trace"], warm["t_trace"], ll_ref, 1e-6)
    print(f"plain {plain['iters']} it ll={plain['ll']:.6f} (1e-6 at {it_p}); "
          f"warm {warm['iters']} it ll={warm['ll']:.6f} (1e-6 at {it_w})", flush=True)
    assert warm["ll"] >= plain["ll"] - 1e-6 * abs(plain["ll"])
    assert it_w is not None and (it_p is None or it_w < it_p)
    mlp.close()


def test_driver_warm_start_and_fallback(tmp_path, capsys, monkeypatch):
    """run_admixture(warm_start=True): the shipped K=5 network fills init_q; with the weights directory
    pointed at an empty folder there is no network for K=9 and the random start runs.  Both fits reach the
    same kind of optimum as the plain driver."""
    import os

    import haplo_amd
    from haplo_amd import plink

    if not os.path.exists(ws.default_weights_path(5)):
        pytest.skip("no trained warm-start weights")
    I, J = 600, 2000
    sd = ws.simulate_world(I, J, 5, 0.1, world_seed=ws.WORLD_SEED, data_seed=31337, truth=False)
    f = str(tmp_path / "w.bed")
    plink.write_bed(f, sd.bed, I, 2 * J)
    plain, _, _ = haplo_amd.run_admixture(f, I, J, 5, T2=np.float32)
    warm, _, _ = haplo_amd.run_admixture(f, I, J, 5, T2=np.float32, warm_start=True)
    assert warm.ll_new >= plain.ll_new - 1e-5 * abs(plain.ll_new)
    assert np.allclose(warm.q.sum(axis=0), 1.0)
    capsys.readouterr()
    monkeypatch.setattr(ws, "DATA_DIR", str(tmp_path / "no_networks"))
    fb, _, _ = haplo_amd.run_admixture(f, I, J, 9, T2=np.float32, warm_start=True)
    assert "No warm-start network for K=9" in capsys.readouterr().out
    assert np.isfinite(fb.ll_new)


@pytest.mark.parametrize("I,J,K,fst,alpha", [(3000, 30000, 6, 0.1, 0.1), (2000, 20000, 10, 0.05, 0.5)])
def test_skfr_start_reaches_plain_optimum(I, J, K, fst, alpha):
    """The SKFR-label start (no network; any K): Q0 weights 0.8 on each individual's sparse k-means cluster,
    EM on P with Q0 held, then QN.  It ends at least as high as the plain random start (on panels with sharp
    clusters the random start stops in a lower optimum) and Q0 is on the simplex."""
    from haplo_amd.sim import simulate

    s = simulate(I, J, K, fst=fst, miss=0.01, alpha_q=alpha, seed=21)
    with HipAdmixData(s.bed, I, J, K, precision="mixed") as e:
        q0, p0 = ws.init_params(I, J, K, seed=11)
        e.set_params(q0, p0)
        plain = e.fit(rtol=1e-7, n_iter=500)
        Q0 = ws.skfr_q0(e)
        assert Q0.shape == (K, I) and np.allclose(Q0.sum(axis=0), 1.0) and Q0.min() > 0
        assert np.all(np.isclose(Q0.max(axis=0), ws.SKFR_Q0_WEIGHT))
        sk = ws.fit_skfr_start(e, rtol=1e-7, n_iter=500)
    print(f"K={K}: plain {plain['iters']} it ll={plain['ll']:.6f}; skfr start {sk['iters']} it ll={sk['ll']:.6f} "
          f"(+{sk['init_seconds']:.3f}s)", flush=True)
    assert sk["qp_failures"] == 0 and sk["qp_iter_limit"] == 0
    assert sk["ll"] >= plain["ll"] - 1e-6 * abs(plain["ll"])


def test_driver_skfr_start(tmp_path):
    """run_admixture(warm_start="skfr") at K=9, where no network ships: the SKFR-label start fills init_q and
    the fit ends near the plain driver's optimum."""
    import haplo_amd
    from haplo_amd import plink

    I, J, K = 600, 2000, 9
    sd = ws.simulate_world(I, J, K, 0.1, world_seed=ws.WORLD_SEED, data_seed=4711, truth=False)
    f = str(tmp_path / "s.bed")
    plink.write_bed(f, sd.bed, I, 2 * J)
    plain, _, _ = haplo_amd.run_admixture(f, I, J, K, T2=np.float32)
    sk, _, _ = haplo_amd.run_admixture(f, I, J, K, T2=np.float32, warm_start="skfr")
    assert np.allclose(sk.q.sum(axis=0), 1.0) and np.isfinite(sk.ll_new)
    assert sk.ll_new >= plain.ll_new - 1e-4 * abs(plain.ll_new)
    with pytest.raises(ValueError):
        haplo_amd.run_admixture(f, I, J, K, T2=np.float32, warm_start="sometimes")


@pytest.mark.slow
def test_training_from_scratch_learns_a_useful_start():
    """A 256-512-512-3 network trained from random weights by the library's own kernels (hk_mlp_grad /
    hk_mlp_adam) on the converged Q of plain-method fits of 12 datasets (I=800, J=4,000, K=3, one
    population structure, Fst ~ U[0.05, 0.15]): the epoch loss falls, and on a held-out dataset of the same
    structure the trained network's Q0 is closer to the truth than the untrained network's, and its warm
    start reaches 1e-6 of the plain fit's converged log-likelihood in fewer iterations than the random start."""
    import itertools

    K, I, J = 3, 800, 4000
    print("training set (12 plain fits)", flush=True)
    ts = ws.make_training_set(n_datasets=12, I=I, J=J, K=K, fst_range=(0.05, 0.15), seed=500, group=6)
    untrained = ws.MLP(K, seed=4)
    print("training", flush=True)
    mlp, hist = ws.train(ts, K, epochs=25, batch=200, lr=1e-3, seed=4)
    print(f"loss per epoch {np.round(hist, 4).tolist()}", flush=True)
    assert hist[-1] < 0.9 * hist[0] and min(hist[-5:]) < min(hist[:5])
    sd = ws.simulate_world(I, J, K, 0.1, miss=0.01, data_seed=9999, truth=True)

    def rmse(Q0):
        return min(float(np.sqrt(np.mean((Q0[list(pm)] - sd.q_true) ** 2))) for pm in itertools.permutations(range(K)))

    with HipAdmixData(sd.bed, I, J, K, precision="mixed") as e:
        r_trained, r_untrained = rmse(ws.predict_q(e, mlp)), rmse(ws.predict_q(e, untrained))
        q0, p0 = ws.init_params(I, J, K, seed=11)
        e.set_params(q0, p0)
        plain = e.fit(rtol=1e-7, n_iter=500)
        warm = ws.fit_warm(e, mlp, rtol=1e-7, n_iter=500)
    ll = plain["ll"]
    ip, _ = ws.time_to_rel(plain["trace"], plain["t_trace"], ll)
    iw, _ = ws.time_to_rel(warm["trace"], warm["t_trace"], ll)
    print(f"Q0 rmse trained {r_trained:.4f} untrained {r_untrained:.4f}; iterations to 1e-6: plain {ip} warm {iw}",
          flush=True)
    assert r_trained < 0.5 * r_untrained
    assert iw is not None and iw < ip
    mlp.close()
    untrained.close()
