#!/usr/bin/env python3
"""bench.py -- time-to-converge of the HaploADMIXTURE likelihood engine on one MI355X.

A "step" is one complete fit on one synthetic dataset resident in HBM: init_em! (5 EM passes) +
admixture_qn! (QN-accelerated block SQP, six genotype passes per iteration) until the reference's
stopping rule (relative log-likelihood change < rtol, default 1e-7), starting from the same seeded
parameters every step.  Default workload: I=20,000 individuals, J=200,000 SNP blocks (400k SNPs,
2-bit packed, 2.0 GB), K=6, Fst=0.05, 2% missing, MIXED precision (FP32 per-pair products, FP64
accumulation / log-likelihood / QP / QN).

Printed JSON (one line): value = seconds per fit (whole job, one GPU).  Extra fields: the plain FP64
GPU fit of the same data (time, log-likelihood, relative gap), the time at which the MIXED fit's
log-likelihood first came within 1e-6 relative of the FP64 converged value, iterations/s, the dominant
kernel's roofline (HIP events on the engine stream) and the CPU baseline (the FP64 C restatement of the
reference's loops, timed on a J-subset on the host cores and scaled).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "haploadmixture.jl_amd"))

import numpy as np  # noqa: E402

from haplo_amd import HipAdmixData  # noqa: E402
from haplo_amd.sim import init_params, simulate  # noqa: E402

WORKLOADS = {
    # name: I, J, K, Fst, missing, data seed, precision
    "mixed_I20k_J200k_K6": dict(I=20000, J=200000, K=6, fst=0.05, miss=0.02, seed=3, prec="mixed"),
    "fp64_I5k_J50k_K4": dict(I=5000, J=50000, K=4, fst=0.05, miss=0.01, seed=2, prec="fp64"),
    "cpu_I500_J2k_K3": dict(I=500, J=2000, K=3, fst=0.10, miss=0.01, seed=1, prec="fp64"),
    # K above 8: windowed curvature passes; four-lane P row QPs spilling to scratch
    "mixed_I20k_J50k_K12": dict(I=20000, J=50000, K=12, fst=0.05, miss=0.02, seed=5, prec="mixed"),
    # batched model selection: K = 2..8 x 4 restarts on data simulated with K = 5
    "select_I3k_J20k_K2-8x4": dict(I=3000, J=20000, K=5, fst=0.05, miss=0.01, seed=4, prec="mixed"),
    # learned warm start: held-out datasets of the training population structure (world seed 2024)
    "warmstart_I10k_J100k_K5": dict(I=10000, J=100000, K=5, fst=0.05, miss=0.01, seed=4242, prec="mixed"),
    # SKFR marker pre-selection (run_admixture(sparsity=...)) on the default fit dataset: k = 6, 10,000 of
    # 200,000 blocks kept (:global), at most 50 reassignments, one try
    "skfr_I20k_J200k_K6_s10k": dict(I=20000, J=200000, K=6, fst=0.05, miss=0.02, seed=3, prec="mixed",
                                    sparsity=10000, sample_I=2000, sample_J=20000),
}
PEAK_TFLOPS = {"mixed": 157.3, "fp64": 78.6}  # MI355X FP32 / FP64 vector peaks (dense)


def log(msg):
    sys.stderr.write(f"[bench {time.strftime('%H:%M:%S')}] {msg}\n")
    sys.stderr.flush()


def flops_per_pair(K: int, kind: str) -> int:
    """Algorithmic flops per (individual, block) pair (FMA = 2, division / log = 1)."""
    if kind == "update":
        return 16 * K + 4 * K * (K + 1) + 24
    if kind == "loglik":
        return 8 * K + 8
    return 24 * K + 16  # EM


def cpu_baseline(bed, I, J, K, q0, p0, gpu_iters, em_iters, j_sample):
    """Time the FP64 C restatement (oracle) on the first j_sample blocks, all individuals: one EM
    iteration and one QN iteration; scale by J / j_sample and the GPU run's iteration counts."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    oracle.lib()
    bps = (I + 3) // 4
    sub = np.ascontiguousarray(bed[: 2 * j_sample * bps])
    qs = np.asfortranarray(q0)
    ps = np.asfortranarray(p0[:, : 4 * j_sample])
    t0 = time.perf_counter()
    oracle.fit(sub, I, j_sample, qs, ps, em_iters=1, n_iter=0)
    t_em = time.perf_counter() - t0
    t0 = time.perf_counter()
    oracle.fit(sub, I, j_sample, qs, ps, em_iters=0, n_iter=1, rtol=0.0)
    t_it = time.perf_counter() - t0
    scale = J / j_sample
    est = scale * (t_em * em_iters + t_it * gpu_iters)
    return dict(value=est, unit="s (time-to-converge, extrapolated)", cores=oracle.threads(), kind="port",
                cpu_model=cpu_model(),
                sample=(f"I={I}, first {j_sample} of {J} blocks, K={K}: 1 EM iteration {t_em:.2f}s + 1 QN "
                        f"iteration {t_it:.2f}s of the FP64 C + OpenMP restatement of loops.jl, scaled "
                        f"x{scale:g} blocks x ({em_iters} EM + {gpu_iters} QN iterations of the GPU run)"),
                t_em_sample_s=t_em, t_iter_sample_s=t_it)


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def select_bench(a, w, prec):
    """28 fits (K = 2..8 x 4 restarts) of one dataset: concurrent batch (shared genotype, per-K autotuned
    chunk shapes, one captured graph per fit per iteration) against the same 28 fits run one at a time."""
    from haplo_amd.select import select_k

    I, J = w["I"], w["J"]
    log(f"generating {a.workload}: I={I} J={J} true K={w['K']}")
    data = simulate(I, J, w["K"], fst=w["fst"], miss=w["miss"], seed=w["seed"], truth=False)
    Ks = list(range(2, 9))
    log("warm-up (kernel loading, graph capture)")
    select_k(data.bed, I, J, Ks=Ks, restarts=1, precision=prec, rtol=a.rtol, n_iter=3, autotune=False)
    log("concurrent batch: one captured graph per round")
    b = select_k(data.bed, I, J, Ks=Ks, restarts=4, precision=prec, rtol=a.rtol, n_iter=a.max_iter, batched=True,
                 use_graph=not a.eager, batch_rounds=not a.eager)
    log(f"batch wall {b.wall_s:.2f}s best K={b.best_k}")
    ev = select_k(data.bed, I, J, Ks=Ks, restarts=4, precision=prec, rtol=a.rtol, n_iter=a.max_iter, batched=True,
                  use_graph=not a.eager, batch_rounds=False)
    log(f"event-driven batch wall {ev.wall_s:.2f}s best K={ev.best_k}")
    sq = select_k(data.bed, I, J, Ks=Ks, restarts=4, precision=prec, rtol=a.rtol, n_iter=a.max_iter, batched=False,
                  use_graph=not a.eager)
    log(f"sequential wall {sq.wall_s:.2f}s best K={sq.best_k}")
    lat_b = np.array([t["finished_at_s"] for t in b.table])
    lat_s = np.array([t["seconds"] for t in sq.table])
    out = {
        "metric": "model selection wall time (s), K=2..8 x 4 restarts, I=3k J=20k",
        "value": round(b.wall_s, 3), "unit": "s", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": round(1000 * b.wall_s, 1), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32-products/f64-accum" if prec == "mixed" else "f64",
        "data": "synthetic (seeded generator, true K=5)",
        "config": {"workload": a.workload, "I": I, "J": J, "Ks": Ks, "restarts": 4, "fits": len(b.table),
                   "precision": prec, "rtol": a.rtol,
                   "batch": "eager per-fit enqueue" if a.eager else "one captured graph per round, device stops"},
        "event_driven_batch_wall_s": round(ev.wall_s, 3),
        "sequential_wall_s": round(sq.wall_s, 3),
        "speedup_batch_over_sequential": round(sq.wall_s / b.wall_s, 3),
        "per_fit_latency_batch_s": {"p50": round(float(np.percentile(lat_b, 50)), 3),
                                    "p95": round(float(np.percentile(lat_b, 95)), 3)},
        "per_fit_latency_sequential_s": {"p50": round(float(np.percentile(lat_s, 50)), 3),
                                         "p95": round(float(np.percentile(lat_s, 95)), 3)},
        "aic_by_k": {str(K): round(b.best_by_k[K]["aic"], 1) for K in Ks},
        "ll_by_k": {str(K): b.best_by_k[K]["ll"] for K in Ks},
        "best_k": b.best_k,
        "autotuned_by_k": {str(K): dict(zip(("q_chunks", "p_chunks", "q_shape", "p_shape"), v))
                           for K, v in b.tuning.items()},
        "iters_by_fit": [t["iters"] for t in b.table],
    }
    print(json.dumps(out), flush=True)


def warmstart_bench(a, w, prec):
    """Plain fit (random start, init_em!, QN) against the warm-started fit (network Q0, EM on P with Q0
    held, QN) on held-out data: iterations and wall time until the log-likelihood reaches
    ll_plain - 1e-6 |ll_plain|, ll_plain = the plain method's converged value.  The main line is the
    Fst=0.05 dataset; Fst=0.02 and 0.12 datasets are reported beside it, and each row also gives the
    SKFR-label start (Q0 from sparse k-means labels, no network) for comparison."""
    import itertools

    from haplo_amd import warmstart as ws

    I, J, K = w["I"], w["J"], w["K"]
    mlp = ws.MLP.load(a.weights) if a.weights else ws.load_default(K)
    if mlp is None:
        raise SystemExit(f"no warm-start network for K={K} in {ws.DATA_DIR}")
    log("warm-up (kernel loading)")
    small = ws.simulate_world(500, 2000, K, 0.05, data_seed=1, truth=False)
    with HipAdmixData(small.bed, 500, 2000, K, precision=prec) as e:
        e.set_params(*init_params(500, 2000, K, seed=1))
        e.fit(rtol=a.rtol, n_iter=5)
        ws.fit_warm(e, mlp, rtol=a.rtol, n_iter=5)
    rows = []
    for fst, seed in [(w["fst"], w["seed"]), (0.02, w["seed"] + 2), (0.12, w["seed"] + 1)]:
        log(f"generating I={I} J={J} K={K} Fst={fst} (world seed {ws.WORLD_SEED}, data seed {seed})")
        sd = ws.simulate_world(I, J, K, fst, miss=w["miss"], data_seed=seed, truth=True)
        with HipAdmixData(sd.bed, I, J, K, precision=prec) as e:
            q0, p0 = init_params(I, J, K, seed=11)
            # set-up, untimed like the context's: the network's projection scratch for this panel size
            # (~250 MB at I=10k) and the seeded projection matrix are made on its first prediction
            ws.predict_q(e, mlp)
            plain_runs, warm_runs, skfr_runs = [], [], []
            for _ in range(a.steps):
                e.set_params(q0, p0)
                plain_runs.append(e.fit(rtol=a.rtol, n_iter=a.max_iter))
                warm_runs.append(ws.fit_warm(e, mlp, rtol=a.rtol, n_iter=a.max_iter))
                skfr_runs.append(ws.fit_skfr_start(e, rtol=a.rtol, n_iter=a.max_iter))
            plain, warm, sk = plain_runs[-1], warm_runs[-1], skfr_runs[-1]
            Q0 = ws.predict_q(e, mlp)
        ll = plain["ll"]
        ip, tp = ws.time_to_rel(plain["trace"], plain["t_trace"], ll)
        iw, tw = ws.time_to_rel(warm["trace"], warm["t_trace"], ll)
        tp_med = float(np.median([ws.time_to_rel(r["trace"], r["t_trace"], r["ll"])[1] for r in plain_runs]))
        tw_med = float(np.median([ws.time_to_rel(r["trace"], r["t_trace"], ll)[1] + r["predict_seconds"]
                                  for r in warm_runs]))
        isk, _ = ws.time_to_rel(sk["trace"], sk["t_trace"], ll)
        tsk = [ws.time_to_rel(r["trace"], r["t_trace"], ll)[1] for r in skfr_runs]
        tsk_med = (None if any(t is None for t in tsk) else
                   float(np.median([t + r["init_seconds"] for t, r in zip(tsk, skfr_runs)])))
        rmse = min(float(np.sqrt(np.mean((Q0[list(pm)] - sd.q_true) ** 2))) for pm in itertools.permutations(range(K)))
        row = {"fst": fst, "data_seed": seed, "plain_iters_to_target": ip, "plain_s_to_target": round(tp_med, 4),
               "plain_iters_total": plain["iters"], "plain_s_total": round(plain["seconds"], 4),
               "warm_iters_to_target": iw, "warm_s_to_target_incl_predict": round(tw_med, 4),
               "warm_iters_total": warm["iters"], "warm_s_total": round(warm["seconds"], 4),
               "predict_s": round(warm["predict_seconds"], 4), "em_p_iters": 10,
               "skfr_start_iters_to_target": isk,
               "skfr_start_s_to_target_incl_init": None if tsk_med is None else round(tsk_med, 4),
               "skfr_start_ll_minus_plain": round(sk["ll"] - ll, 2),
               "ll_plain": ll, "ll_warm_minus_plain": round(warm["ll"] - ll, 2), "q0_rmse_vs_truth": round(rmse, 4)}
        log(f"Fst={fst}: plain {ip} it / {tp_med:.3f}s; warm {iw} it / {tw_med:.3f}s; "
            f"ll warm-plain {warm['ll'] - ll:.1f}; skfr start {isk} it / "
            f"{'-' if tsk_med is None else format(tsk_med, '.3f')}s")
        rows.append(row)
    mlp.close()
    m = rows[0]
    out = {
        "metric": "warm-start time to 1e-6 rel of the plain converged loglik (s), I=10k J=100k K=5",
        "value": m["warm_s_to_target_incl_predict"], "unit": "s", "n_gpus": 1, "steps": a.steps, "warmup": 1,
        "ms_per_step": round(1000 * m["warm_s_to_target_incl_predict"], 2), "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f32-products/f64-accum; MLP bf16 weights/f32 accum",
        "data": "synthetic held-out datasets of the training population structure (seeded world generator)",
        "config": {"workload": a.workload, "I": I, "J": J, "K": K, "fst": w["fst"], "missing": w["miss"],
                   "precision": prec, "rtol": a.rtol, "network": "256-512-512-5, trained on 200 x (I=2000, "
                   "J=20000, K=5, Fst~U[0.02,0.15])"},
        "plain_s_to_target": m["plain_s_to_target"],
        "speedup_time": round(m["plain_s_to_target"] / m["warm_s_to_target_incl_predict"], 3),
        "iteration_reduction": (None if m["warm_iters_to_target"] is None else
                                round(m["plain_iters_to_target"] / m["warm_iters_to_target"], 3)),
        "datasets": rows,
    }
    print(json.dumps(out), flush=True)


def skfr_bench(a, w):
    """SKFR on the full dataset (wall time per call, genotypes resident in HBM), the device time of its
    kernel classes (HIP events on the engine stream), and the CPU restatement against the GPU on a bounded
    sample (first sample_I individuals x first sample_J blocks, sparsity scaled with J)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import skfr_oracle

    from haplo_amd import plink

    I, J, K, s = w["I"], w["J"], w["K"], w["sparsity"]
    log(f"generating {a.workload}: I={I} J={J} Fst={w['fst']} miss={w['miss']}")
    data = simulate(I, J, K, fst=w["fst"], miss=w["miss"], seed=w["seed"], truth=False)
    eng = HipAdmixData(data.bed, I, J, K, precision="mixed")
    log("warm-up (kernel loading)")
    eng.skfr(K, s, max_iter=2, seed=1)
    eng.sync()
    runs = []
    for st in range(a.steps):
        t0 = time.perf_counter()
        r = eng.skfr(K, s, max_iter=50, seed=1, profile=True)
        runs.append((time.perf_counter() - t0, r))
        log(f"step {st}: {runs[-1][0]:.3f}s, {r['iters']} iterations, wss {r['wss']:.6g}")
    wall = float(np.median([t for t, _ in runs]))
    r = runs[-1][1]
    pr = r["profile"]
    eng.close()
    # the center update reads the block-major stream once: 0.5 B per (individual, block) pair, plus its
    # center (k x 2J) and score (J) writes
    cen_ms = pr["centers_ms"] / max(pr["centers_launches"], 1)
    cen_bytes = 0.5 * I * J + 8.0 * (2 * K * J + J)
    dist_ms = pr["dist_ms"] / max(pr["dist_launches"], 1)
    # measured HBM bytes per center update (rocprofv3 FETCH_SIZE / WRITE_SIZE, scaled) from profiles/
    cen_name = "center update (k_sk_member + k_sk_counts + k_sk_centers)"
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        t = json.load(open(tpath)).get(a.workload, {}).get(cen_name)
        traffic = t["hbm_bytes_per_launch"] if t else None
    roof = {"kernel": cen_name, "bound": "hbm", "achieved": round(cen_bytes / (cen_ms * 1e-3) / 1e9, 1),
            "peak": 8000.0, "unit": "GB/s", "frac": None, "traffic": traffic,
            "avg_launch_ms": round(cen_ms, 4), "algorithmic_bytes_per_launch": cen_bytes}
    roof["frac"] = round(roof["achieved"] / roof["peak"], 4)
    # bounded CPU sample: the NumPy restatement and the GPU on the same sub-problem
    Is, Js = w["sample_I"], w["sample_J"]
    ss = max(1, s * Js // J)
    bps = (I + 3) // 4
    codes = plink.unpack_codes(data.bed, I, np.arange(2 * Js))[:, :Is]
    D = skfr_oracle.dosages(codes)
    log(f"CPU sample: I={Is} J={Js} sparsity={ss}")
    t0 = time.perf_counter()
    ref = skfr_oracle.skfr(D, K, ss, max_iter=50, seed=1)
    t_cpu = time.perf_counter() - t0
    with HipAdmixData(np.ascontiguousarray(data.bed[: 2 * Js * bps]), Is, Js, K, bytes_per_snp=bps,
                      n_snp=2 * Js) as es:
        es.skfr(K, ss, max_iter=2, seed=1)
        t0 = time.perf_counter()
        gs = es.skfr(K, ss, max_iter=50, seed=1)
        t_gpu_s = time.perf_counter() - t0
    same = bool(np.array_equal(gs["clusters"], ref["clusters"]) and np.array_equal(gs["aims"], ref["aims"]))
    log(f"sample: CPU {t_cpu:.2f}s, GPU {t_gpu_s:.4f}s, identical labels/markers: {same}")
    out = {
        "metric": f"SKFR marker selection time (s), I=20k J=200k k={K} sparsity={s}",
        "value": round(wall, 4), "unit": "s", "n_gpus": 1, "steps": a.steps, "warmup": 1,
        "ms_per_step": round(1000 * wall, 2), "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (seeded generator, K=6 populations)",
        "config": {"workload": a.workload, "I": I, "J": J, "k": K, "sparsity": s, "mode": "global", "tries": 1,
                   "max_iter": 50},
        "iters": r["iters"], "wss": r["wss"],
        "device_ms": {k2: (round(v, 3) if isinstance(v, float) else v) for k2, v in pr.items()},
        "dist_avg_launch_ms": round(dist_ms, 4),
        "roofline": roof,
        "cpu_baseline": {"value": round(t_cpu, 3), "unit": "s", "cores": 1, "kind": "port",
                         "sample": f"first {Is} individuals x first {Js} blocks, sparsity {ss}, k={K}: the NumPy "
                                   f"restatement (single-threaded Python loop over SNP columns)",
                         "gpu_same_sample_s": round(t_gpu_s, 4), "identical_result": same},
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--workload", default="mixed_I20k_J200k_K6", choices=sorted(WORKLOADS))
    ap.add_argument("--prec", choices=["mixed", "fp64"], default=None)
    ap.add_argument("--rtol", type=float, default=1e-7)
    ap.add_argument("--max-iter", type=int, default=1000)
    ap.add_argument("--no-fp64-ref", action="store_true", help="skip the plain FP64 GPU fit")
    ap.add_argument("--no-fp32acc-ref", action="store_true",
                    help="skip the fit with the reference's T2 = Float32 semantics (FP32 accumulation)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-blocks", type=int, default=10000)
    ap.add_argument("--profile-steps", type=int, default=3,
                    help="fits run after the timed ones with HIP events around every pass (per-kernel times)")
    ap.add_argument("--eager", action="store_true",
                    help="enqueue every iteration's kernels directly instead of replaying captured graphs (for "
                         "kernel traces under rocprofv3)")
    ap.add_argument("--weights", default=None, help="warm start: network weights (.npz) instead of the shipped ones")
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="after the timed fits, run a plain copy kernel of this size (counter calibration)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on exactly one GPU; --gpus {a.gpus} is not supported (the engine is single-GPU; "
              "a split over SNP blocks with one all-reduce per update_q pass is its multi-GPU design)", file=sys.stderr)
        sys.exit(2)
    w = dict(WORKLOADS[a.workload])
    prec = a.prec or w["prec"]
    if a.workload.startswith("select"):
        return select_bench(a, w, prec)
    if a.workload.startswith("warmstart"):
        return warmstart_bench(a, w, prec)
    if a.workload.startswith("skfr"):
        return skfr_bench(a, w)
    I, J, K = w["I"], w["J"], w["K"]
    log(f"generating {a.workload}: I={I} J={J} K={K} Fst={w['fst']} miss={w['miss']}")
    t0 = time.perf_counter()
    data = simulate(I, J, K, fst=w["fst"], miss=w["miss"], seed=w["seed"], truth=False)
    q0, p0 = init_params(I, J, K, seed=w["seed"] + 100)
    log(f"generated in {time.perf_counter() - t0:.1f}s; uploading")
    HipAdmixData(data.bed, I, J, K, precision=prec).close()  # HIP runtime and code objects loaded, untimed
    t0 = time.perf_counter()
    eng = HipAdmixData(data.bed, I, J, K, precision=prec)  # PCIe upload + repack + double-missing counts
    eng.sync()
    setup_s = time.perf_counter() - t0
    log(f"engine ready in {setup_s:.3f}s (q_chunks, p_chunks) = {eng.get_tuning()}")

    def one_fit(profile=0):
        # verbose=0: no per-EM-iteration progress lines, so the EM warm-up evaluates only the log-likelihood
        # the fit uses (the last one); every QN iteration's log-likelihoods are computed as always
        eng.set_params(q0, p0)
        return eng.fit(rtol=a.rtol, n_iter=a.max_iter, profile=profile, verbose=0, use_graph=int(not a.eager))

    for wi in range(a.warmup):
        r = one_fit()
        log(f"warmup {wi}: {r['iters']} iters, {r['seconds']:.2f}s")
    eng.sync()
    t0 = time.perf_counter()
    timed = []
    for si in range(a.steps):
        r = one_fit(profile=0)  # the production path: each QN iteration replayed from its captured hipGraph
        timed.append(r)
        log(f"step {si}: ll={r['ll']:.10g} iters={r['iters']} converged={r['converged']} {r['seconds']:.2f}s")
    eng.sync()
    wall = time.perf_counter() - t0
    ms_per_step = 1000.0 * wall / a.steps
    r = timed[-1]
    # Per-kernel device times: HIP events on the engine stream around every pass, in fits enqueued eagerly
    # (the same kernels in the same order; events cannot bracket single nodes of a replayed graph).
    runs = []
    for si in range(max(a.profile_steps, 1)):
        rp = one_fit(profile=1)
        runs.append(rp)
        log(f"profiled fit {si}: {rp['iters']} iters {rp['seconds']:.3f}s")
    prof_s = float(np.mean([x["seconds"] for x in runs]))

    # dominant pass kernel (HIP events on the engine stream inside the profiled fits)
    fam = {
        "update_p pass (k_pside curvature)": (sum(x["t_ppass_ms"] for x in runs), sum(x["n_ppass"] for x in runs), "update"),
        "update_q pass (k_qside curvature)": (sum(x["t_qpass_ms"] for x in runs), sum(x["n_qpass"] for x in runs), "update"),
        # MIXED: k_ll_mm, both log-likelihoods of a QN iteration in one launch; FP64: k_qside, one per launch
        ("log-likelihood pass (k_ll_mm)" if prec == "mixed" else "log-likelihood pass (k_qside)"):
            (sum(x["t_ll_ms"] for x in runs), sum(x["n_ll"] for x in runs), "loglik"),
    }
    name, (tot_ms, nl, kind) = max(fam.items(), key=lambda kv: kv[1][0])
    avg_ms = tot_ms / max(nl, 1)
    # a MIXED log-likelihood launch evaluates two parameter sets over one genotype sweep
    fl = flops_per_pair(K, kind) * I * J * (2 if kind == "loglik" and prec == "mixed" else 1)
    achieved = fl / (avg_ms * 1e-3) / 1e12
    peak = PEAK_TFLOPS[prec]
    nprof = len(runs)
    breakdown = {k: round(v[0] / nprof, 2) for k, v in fam.items()}
    breakdown["QP q (ms)"] = round(sum(x["t_qstep_ms"] for x in runs) / nprof, 2)
    breakdown["QP p (ms)"] = round(sum(x["t_pstep_ms"] for x in runs) / nprof, 2)
    breakdown["EM (ms)"] = round(sum(x["t_em_ms"] for x in runs) / nprof, 2)
    breakdown["ZAL+proj (ms)"] = round(sum(x["t_qn_ms"] for x in runs) / nprof, 2)

    # HBM bytes per launch of the dominant kernel: not measurable inside this process (PMC counters need
    # rocprofv3); taken from the committed profile, whose own average launch time is compared with this
    # run's so that a stale figure shows
    traffic, traffic_src = None, None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath) and prec == "mixed":
        t = json.load(open(tpath)).get(a.workload, {}).get(name)
        if t:
            traffic = t["hbm_bytes_per_launch"]
            prof_avg = t.get("avg_launch_ms")
            traffic_src = {"file": "profiles/traffic.json", "profile": t.get("source"), "kernel": t.get("kernel"),
                           "profile_avg_launch_ms": prof_avg,
                           "stale": (None if prof_avg is None else abs(prof_avg - avg_ms) > 0.1 * avg_ms)}
    headline = a.workload == "mixed_I20k_J200k_K6" and prec == "mixed"
    out = {
        "metric": ("time-to-converge (s) at 1e-6 rel loglik, I=20k J=200k K=6; HBM GB/s" if headline else
                   f"time-to-converge (s) at rtol {a.rtol:g}, I={I} J={J} K={K}, {prec}"),
        "value": round(ms_per_step / 1000.0, 4),
        "context_setup_s": round(setup_s, 4),
        "unit": "s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_per_step, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-products/f64-accum" if prec == "mixed" else "f64",
        "data": "synthetic (seeded haplotype-admixture generator, Balding-Nichols Fst)",
        "config": {"workload": a.workload, "I": I, "J": J, "K": K, "fst": w["fst"], "missing": w["miss"],
                   "precision": prec, "rtol": a.rtol, "em_iters": 5, "Q": 3, "parallelism": "single GPU"},
        "iters": r["iters"],
        "converged": bool(r["converged"]),
        "ll": r["ll"],
        "iters_per_s": round(r["iters"] / r["seconds"], 3),
        "profiled_fit_s": round(prof_s, 4),
        "kernel_ms_per_fit": breakdown,
        "genotype_GBps": round(0.5 * I * J * (6 * r["iters"] + 15) / (ms_per_step * 1e-3) / 1e9, 1),
        "roofline": {"bound": "mfma", "achieved": round(achieved, 2), "peak": peak, "unit": "TFLOP/s",
                     "frac": round(achieved / peak, 4), "traffic": traffic, "traffic_source": traffic_src,
                     "kernel": name, "avg_launch_ms": round(avg_ms, 3), "launches": nl,
                     "flops_per_launch": fl,
                     "timing": f"HIP events on the engine stream, {nprof} profiled fits after the timed ones",
                     "note": ("compute-bound (instruction issue): v_mfma_f32_16x16x4f32 plus FP32 VALU, which "
                              "gfx950 overlaps little; MI355X's FP32 peak is 157.3 TFLOP/s for matrix and vector "
                              "alike"
                              if prec == "mixed" else
                              "compute-bound: v_mfma_f64_16x16x4f64 plus FP64 VALU; FP64 peak 78.6 TFLOP/s")},
    }
    if a.calib_copy_mb:
        out["copy_calibration"] = {"bytes": a.calib_copy_mb << 20, "reps": 3,
                                   "GBps": round(eng.copy_bench(a.calib_copy_mb << 20, 3), 1)}
    if not a.no_fp64_ref and prec == "mixed":
        qm, pm = eng.get_params()
        eng.close()
        log("plain FP64 GPU fit (reference precision) on the same data")
        e64 = HipAdmixData(data.bed, I, J, K, precision="fp64")
        e64.set_params(q0, p0)
        r64 = e64.fit(rtol=a.rtol, n_iter=a.max_iter, profile=0, verbose=0)
        q64, p64 = e64.get_params()
        e64.close()
        gap = abs(r["ll"] - r64["ll"]) / abs(r64["ll"])
        tr, tt = r["trace"], r["t_trace"]
        hit = np.nonzero(tr >= r64["ll"] - 1e-6 * abs(r64["ll"]))[0]
        out["fp64_gpu"] = {"seconds": round(r64["seconds"], 3), "ll": r64["ll"], "iters": r64["iters"],
                           "rel_ll_gap_mixed_vs_fp64": gap,
                           "max_abs_dq_mixed_vs_fp64": float(np.max(np.abs(qm - q64))),
                           "max_abs_dp_mixed_vs_fp64": float(np.max(np.abs(pm - p64))),
                           "speedup_mixed_over_fp64": round(r64["seconds"] / r["seconds"], 3)}
        out["time_to_1e-6_of_fp64_ll_s"] = round(float(tt[hit[0]]), 3) if hit.size else None
        log(f"fp64: ll={r64['ll']:.10g} iters={r64['iters']} {r64['seconds']:.2f}s; gap {gap:.2e}; max|dQ| "
            f"{out['fp64_gpu']['max_abs_dq_mixed_vs_fp64']:.2e} max|dP| {out['fp64_gpu']['max_abs_dp_mixed_vs_fp64']:.2e}")
        if not a.no_fp32acc_ref:
            log("FP32-accumulation fit (the reference's T2 = Float32 semantics) on the same data")
            e32 = HipAdmixData(data.bed, I, J, K, precision="fp32acc")
            e32.set_params(q0, p0)
            r32 = e32.fit(rtol=a.rtol, n_iter=a.max_iter, profile=0, verbose=0)
            e32.close()
            out["fp32_accumulation_gpu"] = {
                "seconds": round(r32["seconds"], 3), "ll": r32["ll"], "iters": r32["iters"],
                "converged": bool(r32["converged"]),
                "rel_ll_gap_vs_fp64": abs(r32["ll"] - r64["ll"]) / abs(r64["ll"]),
                "note": "FP32 products and FP32 gradient/curvature sums (reference T2 = Float32); "
                        "the MIXED engine adds FP64 accumulation"}
            log(f"fp32 accumulation: ll={r32['ll']:.10g} iters={r32['iters']} gap "
                f"{out['fp32_accumulation_gpu']['rel_ll_gap_vs_fp64']:.2e}")
    if not a.no_cpu_baseline:
        js = min(a.cpu_sample_blocks, J)
        log(f"CPU baseline: oracle on {js} blocks")
        out["cpu_baseline"] = cpu_baseline(data.bed, I, J, K, q0, p0, r["iters"], 5, js)
        out["speedup_vs_cpu"] = round(out["cpu_baseline"]["value"] / out["value"], 1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
