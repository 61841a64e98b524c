"""Learned warm start for Q (the reference's ``init_q`` hook, src/driver.jl:56,184-185).

A 3-layer MLP (256 -> 512 -> 512 -> K, ReLU, softmax; BF16 weights, FP32 accumulation) maps a seeded
Gaussian random projection of an individual's SNP dosages (first ``PANEL_BLOCKS`` blocks of the panel) to
its ancestry proportions.  It is trained with hand-written kernels (libhaplo_hip.so: ``hk_features``,
``hk_mlp_*``) on (projection, converged Q) pairs that the plain method produces on simulated datasets of
one population structure (``sim.simulate_world``), with a permutation-matched cross-entropy (fit labels
are arbitrary).  The warm-started fit then runs EM on P with Q held at the prediction (``em_fix_q``)
before the usual QN/SQP loop; random initialisation stays the fallback.

A second, data-driven start needs no network and works for any K the library fits (``skfr_q0``,
``fit_skfr_start``): Q0 from the cluster labels of sparse k-means with feature ranking on the panel's own
genotypes (``hk_skfr``, the SKFR step of ``run_admixture(sparsity=...)``), weight ``SKFR_Q0_WEIGHT`` on an
individual's cluster and the rest spread evenly over the others, then the same EM on P with Q0 held.
"""
from __future__ import annotations

import ctypes
import os
import time
from dataclasses import dataclass

import numpy as np

from . import _native
from ._native import HipError, ptr
from .engine import HipAdmixData, fit_batch
from .sim import init_params, simulate_world

D_FEAT = 256
WIDTH = 512
PANEL_BLOCKS = 20000
FEATURE_SEED = 77
WORLD_SEED = 2024
DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")


def features(engine: HipAdmixData, n_blocks: int | None = None, seed: int = FEATURE_SEED) -> np.ndarray:
    nb = min(engine.J, PANEL_BLOCKS) if n_blocks is None else int(n_blocks)
    F = np.zeros((engine.I, D_FEAT), dtype=np.float32)
    engine._check(engine._L.hk_features(engine._h, nb, seed, ptr(F)))
    return F


class MLP:
    def __init__(self, K: int, din: int = D_FEAT, width: int = WIDTH, seed: int = 1):
        self._L = _native.hip()
        self.K, self.din, self.width = int(K), int(din), int(width)
        h = ctypes.c_void_p()
        rc = self._L.hk_mlp_create(ctypes.byref(h), self.din, self.width, self.K, seed)
        if rc != 0:
            raise HipError(rc, "hk_mlp_create failed")
        self._h = h
        n = ctypes.c_int64()
        self.offsets = np.zeros(6, dtype=np.int64)
        self._check(self._L.hk_mlp_num_params(self._h, ctypes.byref(n), ptr(self.offsets)))
        self.nparam = n.value
        self.step = 0

    def _check(self, rc):
        if rc != 0:
            raise HipError(rc, self._L.hk_mlp_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            self._L.hk_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # parameters -------------------------------------------------------------------------------------
    def params(self, grad=False):
        w = np.zeros(self.nparam, dtype=np.float32)
        g = np.zeros(self.nparam, dtype=np.float32) if grad else None
        self._check(self._L.hk_mlp_get_params(self._h, ptr(w), ptr(g)))
        return (w, g) if grad else w

    def set_params(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        if w.size != self.nparam:
            raise ValueError("parameter block size mismatch")
        self._check(self._L.hk_mlp_set_params(self._h, ptr(w)))

    def unpack(self, flat):
        """Views (W1, b1, W2, b2, W3, b3) of a flat parameter block."""
        D, H, K = self.din, self.width, self.K
        sizes = [(D, H), (H,), (H, H), (H,), (H, K), (K,)]
        out = []
        for o, sh in zip(self.offsets, sizes):
            out.append(flat[o:o + int(np.prod(sh))].reshape(sh))
        return out

    def save(self, path):
        np.savez(path, w=self.params(), K=self.K, din=self.din, width=self.width)

    @classmethod
    def load(cls, path):
        z = np.load(path)
        m = cls(int(z["K"]), int(z["din"]), int(z["width"]))
        m.set_params(z["w"])
        return m

    # compute ------------------------------------------------------------------------------------------
    def forward(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        P = np.zeros((X.shape[0], self.K), dtype=np.float32)
        self._check(self._L.hk_mlp_forward(self._h, ptr(X), X.shape[0], ptr(P)))
        return P

    def load_data(self, X, T):
        self._X = np.ascontiguousarray(X, dtype=np.float32)
        self._T = None if T is None else np.ascontiguousarray(T, dtype=np.float32)
        self._check(self._L.hk_mlp_load_data(self._h, ptr(self._X), ptr(self._T), self._X.shape[0]))

    def grad(self, row0, n, match=True):
        loss = ctypes.c_float()
        perm = np.zeros(self.K, dtype=np.int32)
        self._check(self._L.hk_mlp_grad(self._h, int(row0), int(n), int(match), ctypes.byref(loss), ptr(perm)))
        return loss.value, perm

    def adam(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.step += 1
        self._check(self._L.hk_mlp_adam(self._h, lr, b1, b2, eps, self.step))


# ------------------------------------------------------------------------------------------------------
# training data, training, warm-started fits
# ------------------------------------------------------------------------------------------------------

@dataclass
class TrainingSet:
    X: np.ndarray        # (N, 256) features
    T: np.ndarray        # (N, K) converged Q (rows on the simplex)
    groups: list         # (row0, n) per dataset
    fit_seconds: float
    fsts: list


def make_training_set(n_datasets=200, I=2000, J=20000, K=5, fst_range=(0.02, 0.15), miss=0.01, seed=1000,
                      world_seed=WORLD_SEED, precision="mixed", rtol=1e-6, group=8, log=None) -> TrainingSet:
    rng = np.random.default_rng(seed)
    fsts = list(rng.uniform(*fst_range, size=n_datasets))
    Xs, Ts, groups = [], [], []
    row = 0
    t_fit = 0.0
    for g0 in range(0, n_datasets, group):
        engs = []
        for d in range(g0, min(n_datasets, g0 + group)):
            s = simulate_world(I, J, K, fsts[d], miss=miss, world_seed=world_seed, data_seed=seed + d, truth=False)
            e = HipAdmixData(s.bed, I, J, K, precision=precision)
            q0, p0 = init_params(I, J, K, seed=seed + 7919 * d)
            e.set_params(q0, p0)
            engs.append(e)
        t0 = time.perf_counter()
        fit_batch(engs, [HipAdmixData.default_opts(rtol=rtol, n_iter=500) for _ in engs])
        t_fit += time.perf_counter() - t0
        for e in engs:
            q, _ = e.get_params()
            Xs.append(features(e))
            Ts.append(np.ascontiguousarray(q.T, dtype=np.float32))
            groups.append((row, I))
            row += I
            e.close()
        if log:
            log(f"training data: {min(n_datasets, g0 + group)}/{n_datasets} datasets")
    return TrainingSet(np.concatenate(Xs), np.concatenate(Ts), groups, t_fit, fsts)


def train(ts: TrainingSet, K: int, epochs=10, batch=500, lr=1e-3, seed=3, log=None) -> tuple:
    mlp = MLP(K, seed=seed)
    mlp.load_data(ts.X, ts.T)
    rng = np.random.default_rng(seed)
    hist = []
    steps = [(r0 + b, min(batch, n - b)) for (r0, n) in ts.groups for b in range(0, n, batch)]
    for ep in range(epochs):
        order = rng.permutation(len(steps))
        tot = 0.0
        for t in order:
            r0, n = steps[t]
            loss, _ = mlp.grad(r0, n, match=True)
            mlp.adam(lr=lr)
            tot += loss * n
        hist.append(tot / ts.X.shape[0])
        if log:
            log(f"epoch {ep + 1}/{epochs}: loss {hist[-1]:.5f}")
    return mlp, hist


def predict_probs(engine: HipAdmixData, mlp: MLP, n_blocks: int | None = None, seed: int = FEATURE_SEED) -> np.ndarray:
    """softmax(MLP(features)) (I x K) with the features computed on the device and fed to the network in
    place (hk_mlp_predict); equal to ``mlp.forward(features(engine))``."""
    nb = min(engine.J, PANEL_BLOCKS) if n_blocks is None else int(n_blocks)
    P = np.zeros((engine.I, mlp.K), dtype=np.float32)
    engine._check(engine._L.hk_mlp_predict(mlp._h, engine._h, nb, seed, ptr(P)))
    return P


def predict_q(engine: HipAdmixData, mlp: MLP, floor=1e-5) -> np.ndarray:
    """Q0 (K x I) for the engine's individuals: softmax(MLP(features)), floored and renormalised."""
    P = predict_probs(engine, mlp).astype(np.float64)
    P = np.maximum(P, floor)
    P /= P.sum(axis=1, keepdims=True)
    return np.asfortranarray(P.T)


def fit_warm(engine: HipAdmixData, mlp: MLP, p_seed=11, em_iters=10, **fit_kw):
    """Warm-started fit: Q0 from the network, P random, ``em_iters`` EM updates of P with Q0 held (each
    costs one P pass and one log-likelihood pass), then QN/SQP on both.  ``predict_seconds`` in the result
    is the feature projection + network time, which comes on top of ``seconds``."""
    t0 = time.perf_counter()
    q0 = predict_q(engine, mlp)
    _, p0 = init_params(engine.I, engine.J, engine.K, seed=p_seed)
    engine.set_params(q0, p0)
    t_pred = time.perf_counter() - t0
    r = engine.fit(em_iters=em_iters, em_fix_q=1, **fit_kw)
    r["predict_seconds"] = t_pred
    return r


def default_weights_path(K=5, world_seed=WORLD_SEED):
    return os.path.join(DATA_DIR, f"warmstart_world{world_seed}_K{K}.npz")


def load_default(K: int):
    """The shipped network for K clusters, or None when there is none (callers then start at random)."""
    path = default_weights_path(K)
    return MLP.load(path) if os.path.exists(path) else None


def fit_auto(engine: HipAdmixData, mlp: MLP | None = None, em_iters=10, **fit_kw):
    """Warm-started fit when a network for engine.K is at hand, else the plain fit from the engine's
    current (q, p) -- the random start stays the fallback.  result["warm_start"] says which ran."""
    if mlp is None:
        mlp = load_default(engine.K)
    if mlp is None or mlp.K != engine.K or mlp.din != D_FEAT:
        r = engine.fit(**fit_kw)
        r["warm_start"] = False
        return r
    r = fit_warm(engine, mlp, em_iters=em_iters, **fit_kw)
    r["warm_start"] = True
    return r


SKFR_Q0_WEIGHT = 0.8
SKFR_MAX_BLOCKS = 5000


def skfr_q0(engine: HipAdmixData, sparsity: int | None = None, weight: float = SKFR_Q0_WEIGHT, seed: int = 1,
            max_iter: int = 20) -> np.ndarray:
    """Q0 (K x I) from sparse k-means labels on the engine's genotypes (hk_skfr, K clusters, the top
    ``sparsity`` blocks -- at most SKFR_MAX_BLOCKS by default): ``weight`` on an individual's cluster,
    (1 - weight) / (K - 1) on each other one."""
    K, I = engine.K, engine.I
    if K == 1:
        return np.ones((1, I), order="F")
    s = min(engine.J, SKFR_MAX_BLOCKS) if sparsity is None else int(sparsity)
    lab = engine.skfr(K, s, max_iter=max_iter, tries=1, seed=seed)["clusters"]
    q0 = np.full((K, I), (1.0 - weight) / (K - 1))
    q0[lab, np.arange(I)] = weight
    return np.asfortranarray(q0)


def fit_skfr_start(engine: HipAdmixData, p_seed=11, em_iters=10, **fit_kw):
    """Fit from the SKFR-label start: Q0 = skfr_q0, P random, ``em_iters`` EM updates of P with Q0 held, then
    QN/SQP on both.  ``init_seconds`` in the result is the sparse k-means time, on top of ``seconds``."""
    t0 = time.perf_counter()
    q0 = skfr_q0(engine)
    _, p0 = init_params(engine.I, engine.J, engine.K, seed=p_seed)
    engine.set_params(q0, p0)
    t_init = time.perf_counter() - t0
    r = engine.fit(em_iters=em_iters, em_fix_q=1, **fit_kw)
    r["init_seconds"] = t_init
    return r


def time_to_rel(trace, t_trace, ll_ref, rel=1e-6):
    """(iterations, seconds) at which a fit's log-likelihood first reached ll_ref - rel * |ll_ref|
    (a fit that ends above ll_ref reaches it too)."""
    hit = np.nonzero(np.asarray(trace) >= ll_ref - rel * abs(ll_ref))[0]
    if hit.size == 0:
        return None, None
    return int(hit[0]) + 1, float(t_trace[hit[0]])


def _say(msg):
    print(msg, flush=True)


def train_default(out_path=None, K=5, n_datasets=200, epochs=10, log=_say):
    """Regenerates the committed weights for K clusters: 200 datasets of the world-seed population structure
    (I=2000, J=20,000, Fst ~ U[0.02, 0.15]) fitted by the plain method (batched, MIXED, rtol 1e-6), then
    `epochs` passes of Adam over their (features, converged Q) rows.  Writes the weights and, next to them, a
    JSON record of the run (settings, the loss per epoch and the device time of each stage)."""
    import json

    t0 = time.perf_counter()
    ts = make_training_set(n_datasets=n_datasets, K=K, log=log)
    t_data = time.perf_counter() - t0
    t0 = time.perf_counter()
    mlp, hist = train(ts, K, epochs=epochs, log=log)
    t_train = time.perf_counter() - t0
    out_path = out_path or default_weights_path(K)
    os.makedirs(os.path.dirname(os.path.abspath(out_path)), exist_ok=True)
    mlp.save(out_path)
    record = {
        "weights": os.path.basename(out_path), "world_seed": WORLD_SEED, "K": K,
        "datasets": {"n": n_datasets, "I": 2000, "J": 20000, "fst": [min(ts.fsts), max(ts.fsts)],
                     "fst_draw": "U[0.02, 0.15], numpy default_rng(1000)", "missing": 0.01,
                     "data_seeds": [1000, 1000 + n_datasets - 1]},
        "targets": "converged Q of the plain method (hk_fit_batch, MIXED, rtol 1e-6, at most 500 iterations)",
        "features": f"seeded Gaussian projection of the first {PANEL_BLOCKS} SNP pairs to {D_FEAT} (seed {FEATURE_SEED})",
        "network": f"{D_FEAT}-{WIDTH}-{WIDTH}-{K}, ReLU, softmax; BF16 weights, FP32 accumulation",
        "training": {"epochs": epochs, "batch": 500, "optimizer": "Adam lr 1e-3 betas (0.9, 0.999) eps 1e-8",
                     "loss": "cross-entropy against the best label permutation per dataset batch", "seed": 3},
        "loss_per_epoch": [round(float(v), 6) for v in hist],
        "rows": int(ts.X.shape[0]),
        "seconds": {"plain_fits": round(ts.fit_seconds, 2), "training_data_total": round(t_data, 2),
                    "training": round(t_train, 2)},
    }
    with open(os.path.splitext(out_path)[0] + ".json", "w") as f:
        json.dump(record, f, indent=1)
    mlp.close()
    return out_path, hist, ts


if __name__ == "__main__":
    # python -m haplo_amd.warmstart [--out DIR] [K ...]: train the network of each K (default 5)
    import argparse

    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=None, help="directory for the weights and records (default: the package data)")
    ap.add_argument("K", nargs="*", type=int, default=[5])
    a = ap.parse_args()
    for k in a.K:
        dst = os.path.join(a.out, os.path.basename(default_weights_path(k))) if a.out else None
        path, hist, ts = train_default(out_path=dst, K=k)
        print(f"wrote {path}; final loss {hist[-1]:.5f}; fits {ts.fit_seconds:.1f}s", flush=True)
