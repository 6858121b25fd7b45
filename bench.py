#!/usr/bin/env python3
"""Benchmark of the reweight-and-fill + LLH hot path on one MI355X.

Workload (default): 5M events in 4 samples (2M/1.5M/1M/0.5M) of 30x50 bins,
300 parameters, 12 dials per event from the seven-type mix (uniform, general,
compact splines, graphs, norms, tabulated, kriged), Barlow-Beeston OA2021
likelihood against the Asimov histogram -- seeded synthetic data from
gundam_amd/csrc/synth.cpp (workload 2).

A step is one LLH evaluation as a fitter issues it: upload the parameter
point and its lookup tables, mirror/clamp, evaluate every dial, form the
event weights, fill the histograms, evaluate the LLH and fetch it to the host.
The device tables are resident in HBM before the timed region starts.

Prints one JSON line (rank 0): the plain FP64 path as "value", plus the
batched path (K points per launch), the fill kernel's roofline position and
the CPU oracle baseline on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published
MFMA_BF16_PEAK_TFLOPS = 2500.0  # MI355X dense BF16 MFMA, published
FP64_PEAK_TFLOPS = 78.6  # MI355X FP64, published
# FP64 operations of one 64-point launch on the 100k-event set's points
# (SQ_INSTS_VALU_ADD_F64 and _MUL_F64 x 64 lanes, _FMA_F64 x 128, read per
# dispatch by rocprofv3 on the same points; the count does not depend on the
# launch shape)
C3_FP64_FLOP_PER_LAUNCH = 1.179e9


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def read_traffic(path):
    """HBM bytes per fill launch measured by rocprofv3 --pmc (FETCH_SIZE +
    WRITE_SIZE, scaled by the copy-kernel calibration), from a committed
    profile summary; None when absent."""
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(workload, n_sample, n_full, seed, threads, evals=20, full_wl=None):
    """The oracle timed on the host cores.  With full_wl (the bench's own
    workload, binned), the whole event set; otherwise a regenerated subset of
    n_sample events, scaled to the full count."""
    from gundam_amd import Workload
    from oracle import pyoracle

    if full_wl is not None:
        wl = full_wl
        t = wl.tables
        n_sample = t.n_events
    else:
        wl = Workload(workload, n_events=n_sample, seed=seed)
        t = wl.tables
        t.bin[:] = pyoracle.bin_events(wl.vars, wl.edges)
    t0 = time.perf_counter()
    orc = pyoracle.Oracle(t, n_threads=threads)
    build_s = time.perf_counter() - t0
    pts = wl.random_points(evals + 3, 0.5, seed=seed + 1)
    tabs = wl.table_values(pts) if t.table_len else None
    _, data, _ = orc.propagate(wl.nominal(), want_weights=False)
    times = []
    for i, th in enumerate(pts):
        t0 = time.perf_counter()
        if tabs is not None:
            orc.set_table_values(tabs[i])
        _, s, e = orc.propagate(th, want_weights=False)
        pyoracle.llh("barlow", data, s, e, th, wl.prior_mean, wl.prior_sigma)
        dt = time.perf_counter() - t0
        if i >= 3:
            times.append(dt)
    med = float(np.median(times))
    binned = int((t.bin >= 0).sum())
    # the 100k-event model whole (the reference's CPU-runnable case)
    w1 = Workload(1, seed=seed)
    t1 = w1.tables
    t1.bin[:] = pyoracle.bin_events(w1.vars, w1.edges)
    o1 = pyoracle.Oracle(t1, n_threads=threads)
    _, d1, _ = o1.propagate(w1.nominal(), want_weights=False)
    times1 = []
    for i, th in enumerate(w1.random_points(evals + 3, 0.5, seed=seed + 2)):
        t0 = time.perf_counter()
        _, s1, e1 = o1.propagate(th, want_weights=False)
        pyoracle.llh("poisson", d1, s1, e1, th, w1.prior_mean, w1.prior_sigma)
        if i >= 3:
            times1.append(time.perf_counter() - t0)
    o1.close()
    orc.close()
    full = n_sample == n_full
    return {
        "workload1_evals_per_s": 1.0 / float(np.median(times1)),
        "value": (1.0 / med) * (n_sample / n_full),
        "unit": "LLH evals/s" if full else "LLH evals/s (scaled to the full event count)",
        "cores": threads,
        "cpu_model": cpu_model(),
        "kind": "port",
        "oracle_build_s": build_s,
        "sample": (f"workload {workload}, all {n_sample} events ({binned} binned), median "
                   f"of {len(times)} evaluations after 3 warm-ups = {med*1e3:.2f} ms"
                   if full else
                   f"workload {workload} restricted to {n_sample} events "
                   f"({binned} binned), median of {len(times)} evaluations "
                   f"after 3 warm-ups = {med*1e3:.2f} ms, scaled linearly by "
                   f"{n_sample}/{n_full} events"),
    }


def plain_workload1(args, log):
    """The 100k-event model (20 parameters, 400 bins, 4 monotonic splines +
    1 norm per event, Poisson LLH against the Asimov histogram) one point per
    evaluation, as a fitter calls it: about 31 MB of algorithmic bytes, which
    the 256 MB Infinity Cache holds, so launch latency dominates.  Reported:
    evaluations/s eager and with the hipGraph replay, and the fill kernel's
    time and algorithmic bandwidth."""
    from gundam_amd import CacheManager, Workload

    wl = Workload(1, seed=args.seed)
    t = wl.tables
    wl.bin_events_on_device(0)
    cm = CacheManager(0)
    cm.build(t)
    cm.propagate_parameters(wl.nominal())
    cm.set_likelihood("poisson", cm.hist(0)[0].copy(), wl.prior_mean, wl.prior_sigma)
    pts = wl.random_points(200, 0.5, seed=args.seed + 31)
    out = {"workload": "100k events, 20 parameters, 400 bins, 4 monotonic splines + 1 "
                       "norm per event, Poisson LLH (Asimov)",
           "algorithmic_bytes": int(t.algorithmic_bytes(8))}
    for graph in (False, True):
        cm.enable_graph(graph)
        for i in range(10):
            cm.set_parameters(pts[i])
            cm.fill()
            cm.llh()
        t0 = time.perf_counter()
        for i in range(10, 200):
            cm.set_parameters(pts[i])
            cm.fill()
            cm.llh()
        el = (time.perf_counter() - t0) / 190
        out["evals_per_s_graph" if graph else "evals_per_s"] = 1.0 / el
    cm.enable_graph(False)
    cm.enable_timing(True)
    for i in range(20):
        cm.set_parameters(pts[i])
        cm.fill()
        cm.llh()
    tm = cm.timing()
    fill_ms = tm["fill_ms"] / max(1, tm["count"])
    out["fill_kernel_ms"] = fill_ms
    out["fill_achieved_GBs"] = (t.algorithmic_bytes(8) - 16 * t.n_bins) / (fill_ms * 1e-3) / 1e9
    out["launch"] = cm.launch_config()
    log(f"c1 (100k events, 1 point): {out['evals_per_s']:.0f} evals/s eager, "
        f"{out['evals_per_s_graph']:.0f} with graph replay; fill {fill_ms*1e3:.1f} us "
        f"({out['fill_achieved_GBs']:.0f} GB/s algorithmic)")
    cm.close()
    return out


def batched_workload3(args, log):
    """The batched measurement on the 100k-event set re-split into samples of
    50k/30k/15k/5k (20x20 bins each): K = 64 points per launch -- 32 chain
    points (prior + N(0,1)) and a 32-point central-difference stencil (+-h,
    h = 1e-3, on parameters 0-15 around one centre) -- captured in a hipGraph,
    launch shape autotuned on K = 64 launches."""
    from gundam_amd import CacheManager, Workload
    from gundam_amd.autotune import autotune

    wl = Workload(3, seed=args.seed)
    t = wl.tables
    wl.bin_events_on_device(0)
    K = 64
    rng = np.random.default_rng(args.seed + 3)
    chains = wl.prior_mean + wl.prior_sigma * rng.standard_normal((32, wl.n_parameters))
    centre = wl.prior_mean + 0.3 * wl.prior_sigma * rng.standard_normal(wl.n_parameters)
    stencil = np.repeat(centre[None, :], 32, axis=0)
    for j in range(16):
        stencil[2 * j, j] += 1e-3
        stencil[2 * j + 1, j] -= 1e-3
    pts = np.concatenate([chains, stencil])
    if args.no_autotune:
        tuned = {"block_threads": args.block or 384, "tile_words": args.tile_words or 2048,
                 "blocks_per_cu": args.blocks_per_cu or 2, "ring_buffers": args.ring or 3}
    else:
        tuned = autotune(t, pts, None, n_events=t.n_events, reps=3)
    cm = CacheManager(0, block_threads=tuned["block_threads"],
                      tile_words=tuned["tile_words"],
                      blocks_per_cu=tuned["blocks_per_cu"],
                      ring_buffers=tuned["ring_buffers"])
    cm.build(t, max_points=K)
    cm.propagate_parameters(wl.nominal())
    cm.set_likelihood("barlow", cm.hist(0)[0].copy(), wl.prior_mean, wl.prior_sigma)
    cm.enable_graph(True)
    cm.set_parameters(pts)
    cm.fill().get()
    reps = 20
    lat = []
    for _ in range(reps):
        t0 = time.perf_counter()
        cm.set_parameters(pts)
        cm.fill()
        cm.llh()
        lat.append(time.perf_counter() - t0)
    med = float(np.median(lat))
    # the K-point kernel's own time (HIP events around it, the same launches
    # without the graph) and its FP64 rate
    cm.enable_graph(False)
    cm.enable_timing(True)
    for _ in range(reps):
        cm.set_parameters(pts)
        cm.fill()
        cm.llh()
    tm = cm.timing()
    cm.enable_timing(False)
    kern_ms = tm["fill_ms"] / max(1, tm["count"])
    tflops = C3_FP64_FLOP_PER_LAUNCH / (kern_ms * 1e-3) / 1e12
    log(f"batched C3 (100k events, K=64, graph): {K/med:.0f} evals/s, "
        f"{med*1e3:.3f} ms/batch, launch {cm.launch_config()}; K-point kernel "
        f"{kern_ms*1e3:.1f} us, {tflops:.2f} FP64 TFLOP/s "
        f"({tflops / FP64_PEAK_TFLOPS:.1%} of peak)")
    return {"workload": "100k events in 4 samples (50k/30k/15k/5k) x 400 bins, 20 "
                        "parameters, 4 monotonic splines + 1 norm per event, BB OA2021",
            "points_per_launch": K, "evals_per_s": K / med,
            "latency_ms_per_batch": med * 1e3, "hipgraph": True,
            "launch": cm.launch_config(),
            "kernel_ms": kern_ms,
            "fp64_flop_per_launch": C3_FP64_FLOP_PER_LAUNCH,
            "fp64_flop_source": "SQ_INSTS_VALU_{ADD,MUL}_F64 x 64 + SQ_INSTS_VALU_FMA_F64 x "
                                "128 per launch on these 64 points "
                                "(profiles/fill_points_kernel_workload3_k64.md)",
            "fp64_tflops": tflops, "fp64_peak_tflops": FP64_PEAK_TFLOPS,
            "fp64_frac": tflops / FP64_PEAK_TFLOPS}


def surrogate_da(args, cm, wl, data, err, log, device_tables=False):
    """Delayed-acceptance MCMC with the learned surrogate on the 5M-event
    model, against the plain exact GPU chain.

    Training pairs come from the exact path during burn-in: 32 batched exact
    chains started from Laplace draws (every evaluated point is kept), topped
    up with draws from the Laplace approximation.  The network is
    n_par -> 512 -> 512 -> n_bins (GELU, BF16 MFMA).  Each method then runs
    for the same wall time from the burn-in end states with the adapted
    proposal covariance.  "Exact-equivalent accepted samples/s" is the
    squared-jump rate (Mahalanobis in the Laplace scales) divided by the
    plain chain's squared jump per accepted move: the number of plain-chain
    accepted moves that cover the same distance.  Raw accepted moves/s are
    reported beside it."""
    from gundam_amd.mcmc import (ExactTarget, SurrogateTarget, adapt_cov, ess_batch_means,
                                 laplace_scales, run_batched_exact, run_delayed, run_plain,
                                 var_ratio)
    from gundam_amd.surrogate import Surrogate

    tv = wl.table_values if wl.tables.table_len and not device_tables else None
    mode = wl.nominal()
    d = mode.size
    ex = ExactTarget(cm, tv)
    sc = laplace_scales(ex, mode)
    rng = np.random.default_rng(args.seed + 41)
    C_ = 32
    t0 = time.perf_counter()
    ex.record = True
    burn = run_batched_exact(ex, mode + sc * rng.standard_normal((C_, d)),
                             (2.38 ** 2 / d) * np.diag(sc ** 2), args.c5_burn, rng)
    extra = args.c5_pairs - sum(a.shape[0] for a in ex.rec_theta)
    for j0 in range(0, max(0, extra), 16384):  # same draws as one (extra, d) block
        m = min(16384, extra - j0)
        ex(mode + sc * rng.standard_normal((m, d)))
        if extra > 16384:
            log(f"c5: {j0 + m}/{extra} Laplace-draw training points")
    ex.record = False
    X, Y = ex.recorded()
    gen_s = time.perf_counter() - t0
    cov = adapt_cov(burn.samples[burn.samples.shape[0] // 2:], sc)
    log(f"c5: {X.shape[0]} training pairs in {gen_s:.1f}s (burn-in acceptance "
        f"{burn.accepted / burn.proposals:.3f})")
    sur = Surrogate(d, wl.tables.n_bins, hidden=512, seed=args.seed + 2)
    t0 = time.perf_counter()
    losses = sur.fit(X, Y, epochs=args.c5_epochs, batch=256, lr=1e-3, log=log, log_every=50)
    train_s = time.perf_counter() - t0
    del X, Y
    ex.rec_theta, ex.rec_sums = [], []
    surr = SurrogateTarget(sur, data, var_ratio(data, err), wl.prior_mean, wl.prior_sigma)
    test = mode + sc * rng.standard_normal((512, d))
    dchi2 = 2.0 * (surr(test) - ex(test))
    cen = np.abs(dchi2 - np.median(dchi2))
    log(f"c5: trained {args.c5_epochs} epochs in {train_s:.1f}s, loss {losses[0]:.3f} -> "
        f"{losses[-1]:.4f}; surrogate - exact chi2: offset {np.median(dchi2):.2f}, "
        f"|centred| p50 {np.median(cen):.2f} p90 {np.quantile(cen, 0.9):.2f}")
    start = burn.samples[-1]
    T = args.c5_seconds
    big = 10 ** 7
    res = {
        "plain": run_plain(ex, start[0].copy(), cov, big, rng, max_seconds=T),
        "batched_exact": run_batched_exact(ex, start.copy(), cov, big, rng, max_seconds=T),
        "da": run_delayed(ex, surr, start.copy(), cov, big, rng, n_sub=1, max_seconds=T),
        "da_sub": run_delayed(ex, surr, start.copy(), cov, big, rng, n_sub=args.c5_sub,
                              max_seconds=T),
    }
    plain = res["plain"]
    jump_per_acc = plain.sq_jumps(sc) / max(1, plain.accepted)
    out = {}
    for k, r in res.items():
        sj = r.sq_jumps(sc) / r.seconds
        ess = ess_batch_means(r.samples)
        kept = r.samples.shape[0] - r.samples.shape[0] // 10
        out[k] = {"ess_per_s": {"min": float(ess.min() / r.seconds),
                                "median": float(np.median(ess) / r.seconds)},
                  "ess_batches_per_chain": max(2, int(np.sqrt(kept))),
                  "accepted_per_s": r.accepted_per_s,
                  "exact_equivalent_accepted_per_s": sj / jump_per_acc,
                  "exact_evals_per_s": r.exact_evals / r.seconds,
                  "exact_evals_per_accepted": r.exact_evals / max(1, r.accepted),
                  "stage2_acceptance": r.accepted / max(1, r.stage1_passed),
                  "iterations": int(r.samples.shape[0]), "chains": int(r.samples.shape[1]),
                  "seconds": r.seconds}
        log(f"c5 {k}: ESS/s min {out[k]['ess_per_s']['min']:.2f} median "
            f"{out[k]['ess_per_s']['median']:.2f}, {r.accepted_per_s:.1f} accepted/s, "
            f"{sj / jump_per_acc:.1f} exact-equivalent/s, "
            f"{r.exact_evals / max(1, r.accepted):.2f} exact evals per accepted move")
    out["da"]["n_sub"] = 1
    out["da_sub"]["n_sub"] = args.c5_sub
    base = out["plain"]["exact_equivalent_accepted_per_s"]
    # posterior means of the delayed-acceptance chains against the batched
    # exact chains (second halves), in Laplace sigmas
    def mean2(r):
        return r.samples[r.samples.shape[0] // 2:].reshape(-1, d).mean(0)
    ref = mean2(res["batched_exact"])
    agree = {k: {"median": float(np.median(np.abs(mean2(res[k]) - ref) / sc)),
                 "max": float(np.max(np.abs(mean2(res[k]) - ref) / sc))}
             for k in ("da", "da_sub")}
    b = res["batched_exact"].samples
    h = b.shape[1] // 2
    split = np.abs(b[b.shape[0] // 2:, :h].reshape(-1, d).mean(0) -
                   b[b.shape[0] // 2:, h:].reshape(-1, d).mean(0)) / sc
    return {
        "model": f"{d} -> 512 -> 512 -> {wl.tables.n_bins} (GELU, BF16 MFMA, FP32 accumulate)",
        "training_pairs": int(args.c5_pairs), "pair_generation_s": gen_s,
        "epochs": args.c5_epochs, "train_s": train_s,
        "loss_first_last": [float(losses[0]), float(losses[-1])],
        "surrogate_minus_exact_chi2": {"offset": float(np.median(dchi2)),
                                       "p50_abs_centred": float(np.median(cen)),
                                       "p90_abs_centred": float(np.quantile(cen, 0.9))},
        "methods": out,
        "gain_ess_per_s": {k: {q: out[k]["ess_per_s"][q] / out["plain"]["ess_per_s"][q]
                               for q in ("min", "median")}
                           for k in ("batched_exact", "da", "da_sub")},
        "ess_method": ("batch means (Flegal & Jones 2010), per parameter over all chains, "
                       "first 10% of each run dropped, floor(sqrt(n)) batches per chain; "
                       "runs much shorter than the integrated autocorrelation time "
                       "overstate ESS"),
        "gain_exact_equivalent": {k: out[k]["exact_equivalent_accepted_per_s"] / base
                                  for k in ("batched_exact", "da", "da_sub")},
        "gain_accepted_per_s": {k: out[k]["accepted_per_s"] / out["plain"]["accepted_per_s"]
                                for k in ("batched_exact", "da", "da_sub")},
        "posterior_mean_diff_sigma": agree,
        "exact_split_half_diff_sigma": {"median": float(np.median(split)),
                                        "max": float(np.max(split))},
        "seconds_per_method": T,
    }


def precision_c4(args, log):
    """FP32 dial mode against the FP64 path on the 20M-event workload (6
    samples of 60x50 bins, 500 parameters, 15 dials/event): max per-bin
    relative error of the bin sums and |dLLH| (Barlow-Beeston against the
    FP64 Asimov histogram) at points drawn from the priors (shape parameters
    N(0,1), norms 1 + 0.1 N(0,1)), K points per launch; plus the plain fill
    kernel of each mode.
    The FP64 path is the reference here: it matches the CPU oracle to 1e-10
    on the bin sums (tests/test_gpu_parity.py)."""
    from gundam_amd import CacheManager, Workload

    t0 = time.time()
    wl = Workload(4, seed=args.seed)
    t = wl.tables
    wl.bin_events_on_device(0)
    log(f"c4: generated {t.n_events} events, {t.n_dials} dials in {time.time()-t0:.1f}s")
    K = 64
    cms = {}
    shapes = {}
    from gundam_amd.autotune import autotune
    tp = wl.prior_points(1, seed=args.seed + 5)
    for prec in ("fp64", "fp32"):
        # FP32 records are half the size: the main workload's FP64 shape
        # does not carry over, so that mode is always tuned
        if args.c4_autotune or not args.tile_words or prec == "fp32":
            tuned = autotune(t, tp, wl.table_values(tp), precision=prec,
                             n_events=args.tune_events, reps=3)
        else:
            tuned = {"tile_words": args.tile_words, "block_threads": args.block,
                     "blocks_per_cu": args.blocks_per_cu, "ring_buffers": args.ring}
        shapes[prec] = {k: tuned[k] for k in ("tile_words", "block_threads",
                                              "blocks_per_cu", "ring_buffers")}
        t0 = time.time()
        cms[prec] = CacheManager(0, precision=prec, **shapes[prec])
        cms[prec].build(t, max_points=K)
        log(f"c4: {prec} shape {shapes[prec]} built in {time.time()-t0:.1f}s, "
            f"{cms[prec].counts()}, resident {cms[prec].resident_memory()/2**30:.2f} GiB")
    nominal = wl.nominal()
    cms["fp64"].propagate_parameters(nominal, wl.table_values(nominal))
    data = cms["fp64"].hist(0)[0].copy()
    for cm in cms.values():
        cm.set_likelihood("barlow", data, wl.prior_mean, wl.prior_sigma)
    fill = {}
    for prec, cm in cms.items():
        p1 = wl.prior_points(6, seed=args.seed + 7)
        tb = wl.table_values(p1)
        cm.enable_timing(True)
        for i in range(6):
            cm.set_parameters(p1[i])
            cm.set_table_values(tb[i])
            cm.fill()
            cm.llh()
        tm = cm.timing()
        cm.enable_timing(False)
        ms = tm["fill_ms"] / max(1, tm["count"])
        gbs = t.algorithmic_bytes(8 if prec == "fp64" else 4) / (ms * 1e-3) / 1e9
        fill[prec] = {"fill_kernel_ms": ms, "achieved_GBs": gbs, "frac": gbs / HBM_PEAK_GBS}
        log(f"c4: {prec} plain fill {ms:.3f} ms ({gbs:.0f} GB/s)")
    n = args.c4_points
    pts = wl.prior_points(n, seed=args.seed + 4)
    rel_max = np.zeros(n)
    rel2_max = np.zeros(n)
    dllh = np.zeros(n)
    llh64 = np.zeros(n)
    zero_mismatch = 0
    t0 = time.time()
    for i in range(0, n, K):
        th = pts[i:i + K]
        tb = wl.table_values(th)
        res = {}
        for prec, cm in cms.items():
            cm.set_parameters(th)
            cm.set_table_values(tb)
            cm.fill()
            res[prec] = ([cm.hist(k) for k in range(th.shape[0])], cm.llh()[0][:th.shape[0]].copy())
        for k in range(th.shape[0]):
            s64, e64 = res["fp64"][0][k]
            s32, e32 = res["fp32"][0][k]
            nz = s64 != 0
            zero_mismatch += int(np.count_nonzero(s32[~nz]))
            rel_max[i + k] = np.max(np.abs(s32[nz] - s64[nz]) / np.abs(s64[nz]))
            rel2_max[i + k] = np.max(np.abs(e32[nz] - e64[nz]) / np.abs(e64[nz]))
        l64, l32 = res["fp64"][1], res["fp32"][1]
        dllh[i:i + th.shape[0]] = np.abs(l32 - l64)
        llh64[i:i + th.shape[0]] = l64
    log(f"c4: {n} points in {time.time()-t0:.1f}s; max per-bin rel error "
        f"{rel_max.max():.3e}, max |dLLH| {dllh.max():.3e}")
    for cm in cms.values():
        cm.close()
    return {
        "workload": (f"{t.n_events} events in 6 samples (8M/5M/3M/2M/1M/1M) x 3000 bins, "
                     f"{wl.n_parameters} parameters, {t.n_dials} dials"),
        "points": n, "points_per_launch": K,
        "max_rel_bin_sum_error": float(rel_max.max()),
        "median_of_per_point_max_rel_error": float(np.median(rel_max)),
        "max_rel_sqrt_sumw2_error": float(rel2_max.max()),
        "max_abs_dllh": float(dllh.max()),
        "median_abs_dllh": float(np.median(dllh)),
        "max_rel_dllh": float(np.max(dllh / np.abs(llh64))),
        "llh_range": [float(llh64.min()), float(llh64.max())],
        "nonzero_where_fp64_zero": zero_mismatch,
        "plain_fill": fill,
        "launch": shapes,
    }


def mlp_kernels(args, log, n_in=300, hidden=512, n_out=6000, rows=16384, batch=256,
                epochs=3):
    """The surrogate network's training kernels on their own: n_in -> 512 ->
    512 -> n_out (the 5M-event model's 300 parameters and 6000 bins), Adam
    minibatches of 256 over `rows` synthetic pairs resident on the device.
    FLOPs per row and step: 2 (3 (n_in h + h^2 + h n_out) - n_in h) (forward,
    weight gradients, and input gradients of the two upper layers), against
    the dense BF16 MFMA peak."""
    from gundam_amd.surrogate import MLP
    rng = np.random.default_rng(args.seed + 97)
    x = rng.standard_normal((rows, n_in)).astype(np.float32)
    tgt = rng.standard_normal((rows, n_out)).astype(np.float32)
    m = MLP(n_in, hidden, n_out, seed=args.seed + 5)
    m.set_data(x, tgt)
    m.train(1, batch, 1e-3, 0)  # warm-up epoch
    t0 = time.perf_counter()
    m.train(epochs, batch, 1e-3, 1)
    el = time.perf_counter() - t0
    m.close()
    steps = epochs * (rows // batch)
    flop = 2.0 * (3 * (n_in * hidden + hidden * hidden + hidden * n_out) - n_in * hidden)
    tflops = flop * batch * steps / el / 1e12
    log(f"surrogate MLP training: {el / steps * 1e6:.1f} us per step of {batch} rows, "
        f"{tflops:.1f} TFLOP/s ({tflops / MFMA_BF16_PEAK_TFLOPS:.2%} of the BF16 MFMA peak)")
    # the same kernels on 4096-row minibatches: what they reach once a launch
    # carries enough work (the training itself uses 256-row minibatches)
    big_rows, big_batch, big_epochs = 16384, 4096, 5
    m = MLP(n_in, hidden, n_out, seed=args.seed + 5)
    m.set_data(x[:big_rows], tgt[:big_rows])
    m.train(1, big_batch, 1e-3, 0)
    t0 = time.perf_counter()
    m.train(big_epochs, big_batch, 1e-3, 1)
    el_big = time.perf_counter() - t0
    m.close()
    steps_big = big_epochs * (big_rows // big_batch)
    tflops_big = flop * big_batch * steps_big / el_big / 1e12
    log(f"surrogate MLP kernels at {big_batch}-row minibatches: "
        f"{el_big / steps_big * 1e6:.1f} us per step, {tflops_big:.1f} TFLOP/s "
        f"({tflops_big / MFMA_BF16_PEAK_TFLOPS:.2%})")
    return {"network": f"{n_in}-{hidden}-{hidden}-{n_out}", "batch": batch,
            "steps": steps, "us_per_step": el / steps * 1e6, "tflops": tflops,
            "peak_tflops": MFMA_BF16_PEAK_TFLOPS, "frac": tflops / MFMA_BF16_PEAK_TFLOPS,
            "large_batch": {"batch": big_batch, "steps": steps_big,
                            "us_per_step": el_big / steps_big * 1e6, "tflops": tflops_big,
                            "frac": tflops_big / MFMA_BF16_PEAK_TFLOPS},
            "bound": ("256 rows: launch and load latency (13 launches of 3-19 us per "
                      "5.3-GFLOP step); 4096 rows: the three large GEMMs (363-520 TFLOP/s, "
                      "179 us) and the memory-bound split-K reductions (with the Adam "
                      "steps), column sums and row gather (63 us)")}


def posterior_agreement(args, log):
    """Posterior means of the delayed-acceptance chains (surrogate sub-chains
    of args.c5_sub steps) against batched exact chains on the 100k-event
    model (20 parameters; --post-workload 2: the 5M-event model, 300
    parameters, device table refresh), each run for args.post_seconds from the
    same burn-in states, in segments of at most a minute with a progress line
    after each: per-parameter difference of the means in units of the exact
    chains' posterior sigma, with its Monte Carlo standard error from batch
    means (20 batches per chain set)."""
    from gundam_amd import CacheManager, Workload
    from gundam_amd.mcmc import (ExactTarget, SurrogateTarget, adapt_cov, ess_batch_means,
                                 laplace_scales, run_batched_exact, run_delayed, var_ratio)
    from gundam_amd.surrogate import Surrogate

    wk = 2 if args.post_workload == 2 else 1
    wl = Workload(wk, seed=args.seed + 1)
    t = wl.tables
    wl.bin_events_on_device(0)
    if wk == 2:  # the launch shape the autotuner picks for this event set
        cm = CacheManager(0, block_threads=384, tile_words=3840, blocks_per_cu=2,
                          ring_buffers=2)
    else:
        cm = CacheManager(0)
    cm.build(t, max_points=64)
    if t.table_len:
        from gundam_amd import _native as N
        cm.add_table_update(N.TABLE_UPDATE_COS, N.TABLE_UPDATE_COS_KERNEL, wl.table_jobs())
    log(f"posterior: workload {wk} built, {t.n_events} events, {t.n_parameters} parameters")
    mode = wl.nominal()
    cm.propagate_parameters(mode)
    data, w2 = (a.copy() for a in cm.hist(0))
    cm.set_likelihood("barlow", data, wl.prior_mean, wl.prior_sigma)
    ex = ExactTarget(cm)
    sc = laplace_scales(ex, mode)
    d = mode.size
    rng = np.random.default_rng(args.seed + 77)
    C_ = 32
    ex.record = True
    burn = run_batched_exact(ex, mode + sc * rng.standard_normal((C_, d)),
                             (2.38 ** 2 / d) * np.diag(sc ** 2), args.post_burn, rng)
    ex(mode + sc * rng.standard_normal((max(0, 16000 - ex.points), d)))
    ex.record = False
    X, Y = ex.recorded()
    cov = adapt_cov(burn.samples[burn.samples.shape[0] // 2:], sc)
    sur = Surrogate(d, t.n_bins, hidden=512, seed=args.seed + 3)
    sur.fit(X, Y, epochs=300 if wk == 2 else 200, batch=256, lr=1e-3)
    surr = SurrogateTarget(sur, data, var_ratio(data, w2), wl.prior_mean, wl.prior_sigma)
    log(f"posterior: burn-in, {X.shape[0]} training pairs and the surrogate done")
    start = burn.samples[-1]
    T = args.post_seconds
    big = 10 ** 8

    class Joined:  # the segments of one chain set, joined
        def __init__(self, samples):
            self.samples = samples

    def segmented(name, run):
        parts, th, done = [], start.copy(), 0.0
        while done < T:
            r = run(th, min(60.0, T - done))
            parts.append(r.samples)
            th = r.samples[-1].copy()
            done += r.seconds
            log(f"posterior: {name} chains {done:.0f}/{T:.0f} s, "
                f"{sum(p.shape[0] for p in parts)} iterations")
        return Joined(np.concatenate(parts))

    res = {"exact": segmented("exact", lambda th, sec: run_batched_exact(
               ex, th, cov, big, rng, max_seconds=sec)),
           "delayed": segmented("delayed", lambda th, sec: run_delayed(
               ex, surr, th, cov, big, rng, n_sub=args.c5_sub, max_seconds=sec))}

    def batch_stats(smp, nb=20):
        smp = smp[smp.shape[0] // 10:]  # drop 10% as further burn-in
        n = smp.shape[0] // nb * nb
        b = smp[:n].reshape(nb, -1, smp.shape[1], d).mean(axis=(1, 2))
        return smp.reshape(-1, d).mean(0), b.std(0, ddof=1) / np.sqrt(nb)

    m_ex, se_ex = batch_stats(res["exact"].samples)
    m_da, se_da = batch_stats(res["delayed"].samples)
    sigma = res["exact"].samples[res["exact"].samples.shape[0] // 10:].reshape(-1, d).std(0)
    diff = (m_da - m_ex) / sigma
    se = np.sqrt(se_ex ** 2 + se_da ** 2) / sigma
    out = {
        "workload": ("100k events, 20 parameters, 400 bins, Barlow-Beeston OA2021 (Asimov)"
                     if wk == 1 else
                     "5M events, 300 parameters, 6000 bins, Barlow-Beeston OA2021 (Asimov)"),
        "chains": C_, "seconds_per_method": T, "n_sub": args.c5_sub,
        "iterations": {k: int(r.samples.shape[0]) for k, r in res.items()},
        "max_abs_mean_diff_sigma": float(np.max(np.abs(diff))),
        "median_abs_mean_diff_sigma": float(np.median(np.abs(diff))),
        "max_mc_standard_error_sigma": float(np.max(se)),
        # the exact chains' own batch-means standard error of the mean
        "exact_mc_standard_error_sigma": {"median": float(np.median(se_ex / sigma)),
                                          "max": float(np.max(se_ex / sigma))},
        "ess_batch_means": {k: {"min": float(e.min()), "median": float(np.median(e))}
                            for k, e in ((k, ess_batch_means(r.samples))
                                         for k, r in res.items())},
        "params_within_0p05_sigma": int(np.sum(np.abs(diff) < 0.05)),
        "max_abs_z": float(np.max(np.abs(diff / se))),
        "burn_in_iterations": int(args.post_burn),
        # median over parameters of the second half's spread over the first
        # half's (after the 10% drop): > 1 while a chain set still widens
        "width_growth": {k: float(np.median(
            r.samples[r.samples.shape[0] // 10:][(r.samples.shape[0] * 9 // 10) // 2:]
            .reshape(-1, d).std(0) /
            r.samples[r.samples.shape[0] // 10:][:(r.samples.shape[0] * 9 // 10) // 2]
            .reshape(-1, d).std(0))) for k, r in res.items()},
        "std_ratio_delayed_over_exact": [
            float(x) for x in np.round(res["delayed"].samples[res["delayed"].samples.shape[0] // 10:]
                                       .reshape(-1, d).std(0) / sigma, 4)],
    }
    log(f"posterior: max |dmean| {out['max_abs_mean_diff_sigma']:.3f} sigma "
        f"(MC s.e. <= {out['max_mc_standard_error_sigma']:.3f}), "
        f"{out['params_within_0p05_sigma']}/{d} within 0.05 sigma, max |z| {out['max_abs_z']:.2f}")
    cm.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", type=int, default=2)
    ap.add_argument("--events", type=int, default=0, help="0: full size")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"])
    ap.add_argument("--batch", type=int, default=64,
                    help="points per launch for the batched figure (0: skip)")
    ap.add_argument("--cpu-sample", type=int, default=0,
                    help="events of the CPU baseline (0: the whole workload)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--tile-words", type=int, default=0)
    ap.add_argument("--blocks-per-cu", type=int, default=0)
    ap.add_argument("--ring", type=int, default=0, help="LDS tile buffers (2 or 3)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-c3", action="store_true",
                    help="skip the 100k-event measurements (K=64 batched, single point)")
    ap.add_argument("--tune-events", type=int, default=500000)
    ap.add_argument("--no-c5", action="store_true",
                    help="skip the surrogate delayed-acceptance MCMC leg")
    ap.add_argument("--c5-pairs", type=int, default=200000)
    ap.add_argument("--c5-burn", type=int, default=120)
    ap.add_argument("--c5-epochs", type=int, default=60)
    ap.add_argument("--c5-sub", type=int, default=16)
    ap.add_argument("--c5-seconds", type=float, default=8.0)
    ap.add_argument("--host-tables", action="store_true",
                    help="compute lookup tables with the host callback and upload them "
                         "every step (default: device table refresh)")
    ap.add_argument("--no-c4", action="store_true",
                    help="skip the FP32-vs-FP64 precision study on the 20M-event workload")
    ap.add_argument("--c4-points", type=int, default=128)
    ap.add_argument("--c4-autotune", action="store_true",
                    help="autotune the precision study's launch shape per precision "
                         "(default: the main workload's shape)")
    ap.add_argument("--only-c4", action="store_true",
                    help="run only the precision study and print its JSON")
    ap.add_argument("--only-posterior", action="store_true",
                    help="run only the posterior-agreement study (100k-event model)")
    ap.add_argument("--post-seconds", type=float, default=60.0)
    ap.add_argument("--post-burn", type=int, default=400,
                    help="batched exact burn-in iterations of the posterior study")
    ap.add_argument("--post-workload", type=int, default=1, choices=[1, 2],
                    help="posterior study on the 100k-event (1) or 5M-event (2) model")
    ap.add_argument("--only-c3", action="store_true",
                    help="run only the 64-point batched leg on the 100k-event set (with "
                         "--no-autotune: the shape from --tile-words/--block/"
                         "--blocks-per-cu/--ring, default 2048/384/2/3)")
    ap.add_argument("--only-mlp", action="store_true",
                    help="run only the surrogate network's training-kernel timing")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; "
                          "--gpus must be 1 (multi-GPU sharding is a design "
                          "note, not built)"}))
        sys.exit(2)

    from gundam_amd import CacheManager, Workload

    if args.only_c4:
        print(json.dumps({"precision_c4": precision_c4(args, log)}), flush=True)
        return
    if args.only_posterior:
        print(json.dumps({"posterior_agreement": posterior_agreement(args, log)}), flush=True)
        return
    if args.only_mlp:
        print(json.dumps({"surrogate_mlp": mlp_kernels(args, log)}), flush=True)
        return
    if args.only_c3:
        print(json.dumps({"batched_c3": batched_workload3(args, log)}), flush=True)
        return

    t0 = time.time()
    wl = Workload(args.workload, n_events=args.events, seed=args.seed)
    t = wl.tables
    log(f"generated workload {args.workload}: {t.n_events} events, "
        f"{t.n_dials} dials, {t.n_bins} bins in {time.time()-t0:.1f}s")
    t0 = time.time()
    wl.bin_events_on_device(0)
    log(f"binned on device in {time.time()-t0:.1f}s, "
        f"{int((t.bin < 0).sum())} unbinned")

    word = 8 if args.precision == "fp64" else 4
    alg_bytes = t.algorithmic_bytes(word)
    fill_bytes = alg_bytes - 16 * t.n_bins  # the finalize kernel's share

    K = max(1, args.batch)
    nominal = wl.nominal()
    tuned = None
    if not args.no_autotune and not (args.block or args.tile_words or args.blocks_per_cu):
        from gundam_amd.autotune import autotune
        tp = wl.random_points(1, 0.5, seed=args.seed + 5)
        tuned = autotune(t, tp, wl.table_values(tp) if t.table_len else None,
                         precision=args.precision, n_events=args.tune_events,
                         reps=3, log=log)
        args.block, args.tile_words, args.blocks_per_cu, args.ring = (
            tuned["block_threads"], tuned["tile_words"], tuned["blocks_per_cu"],
            tuned["ring_buffers"])
        log(f"autotuned in {tuned['tune_seconds']:.1f}s over {tuned['candidates']} "
            f"shapes: {args.tile_words} words/tile, {args.block}+64 threads, "
            f"{args.blocks_per_cu} blocks/CU, {args.ring} buffers")
    t0 = time.time()
    cm = CacheManager(0, precision=args.precision, block_threads=args.block,
                      tile_words=args.tile_words,
                      blocks_per_cu=args.blocks_per_cu, ring_buffers=args.ring)
    cm.build(t, max_points=K)
    counts = cm.counts()
    log(f"built in {time.time()-t0:.1f}s: {counts}, resident "
        f"{cm.resident_memory()/2**30:.2f} GiB, launch {cm.launch_config()}")

    # lookup tables: refreshed on the device at every fill from the mirrored
    # parameters (gcm_add_table_update), or with --host-tables computed by the
    # host callback and uploaded with the parameters, as the reference does
    device_tables = bool(t.table_len) and not args.host_tables
    if device_tables:
        from gundam_amd import _native as N
        cm.add_table_update(N.TABLE_UPDATE_COS, N.TABLE_UPDATE_COS_KERNEL, wl.table_jobs())
    cm.propagate_parameters(nominal, None if device_tables else wl.table_values(nominal))
    data, data_err = (a.copy() for a in cm.hist(0))  # Asimov data at nominal
    cm.set_likelihood("barlow", data, wl.prior_mean, wl.prior_sigma)

    n_pts = args.warmup + args.steps
    pts = wl.random_points(n_pts, 0.5, seed=args.seed + 17)
    tabs = wl.table_values(pts) if t.table_len and not device_tables else None

    def one(i):
        cm.set_parameters(pts[i])
        if tabs is not None:
            cm.set_table_values(tabs[i])
        cm.fill()
        return cm.llh()[0][0]

    for i in range(args.warmup):
        one(i)
    cm.enable_timing(True)
    t0 = time.perf_counter()
    llhs = [one(args.warmup + i) for i in range(args.steps)]
    elapsed = time.perf_counter() - t0
    tm = cm.timing()
    cm.enable_timing(False)
    if not np.all(np.isfinite(llhs)):
        raise RuntimeError("non-finite LLH in the timed region")
    ms_per_step = elapsed / args.steps * 1e3
    evals_per_s = args.steps / elapsed
    fill_ms = tm["fill_ms"] / max(1, tm["count"])
    fin_ms = tm["finalize_ms"] / max(1, tm["count"])
    achieved = fill_bytes / (fill_ms * 1e-3) / 1e9
    log(f"plain: {ms_per_step:.3f} ms/eval, fill kernel {fill_ms:.3f} ms "
        f"({achieved:.0f} GB/s), finalize {fin_ms:.3f} ms")

    batched = None
    if args.batch > 1:
        bp = wl.random_points(K, 0.5, seed=args.seed + 99)
        btab = wl.table_values(bp) if t.table_len and not device_tables else None
        cm.set_parameters(bp)
        if btab is not None:
            cm.set_table_values(btab)
        cm.fill()
        cm.llh()
        reps = 3
        cm.enable_timing(True)
        t0 = time.perf_counter()
        for _ in range(reps):
            cm.set_parameters(bp)
            if btab is not None:
                cm.set_table_values(btab)
            cm.fill()
            cm.llh()
        el = (time.perf_counter() - t0) / reps
        tmb = cm.timing()
        cm.enable_timing(False)
        batched = {"points_per_launch": K, "evals_per_s": K / el,
                   "latency_ms_per_batch": el * 1e3,
                   "fill_kernel_ms": tmb["fill_ms"] / max(1, tmb["count"])}
        log(f"batched K={K}: {K/el:.1f} evals/s, {el*1e3:.2f} ms/batch")

    batched_c3 = None if args.no_c3 else batched_workload3(args, log)
    plain_c1 = None if args.no_c3 else plain_workload1(args, log)

    c4 = None if args.no_c4 else precision_c4(args, log)

    c5 = None
    if not args.no_c5 and args.batch >= 32:
        c5 = surrogate_da(args, cm, wl, data, data_err, log, device_tables)
    mlp = None if args.no_c5 else mlp_kernels(args, log)

    cpu = None
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        n_full = t.n_events
        sample = min(args.cpu_sample, n_full) if args.cpu_sample > 0 else n_full
        log(f"cpu baseline: oracle on {sample} events, {threads} threads")
        cpu = cpu_baseline(args.workload, sample, n_full, args.seed, threads,
                           full_wl=wl if sample == n_full else None)

    traffic = read_traffic(os.path.join(
        ROOT, "profiles", f"fill_kernel_workload{args.workload}_{args.precision}_traffic.txt"))
    out = {
        "metric": "LLH evals/s, 5M events x 300 dials: plain vs batched vs surrogate; % HBM peak",
        "value": evals_per_s,
        "unit": "LLH evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32 dials, f64 sums",
        "data": "synthetic (seeded generator, gundam_amd/csrc/synth.cpp)",
        "config": {
            "workload": ("plain reweight+fill+BB LLH: 5M events in 4 samples "
                         "(2M/1.5M/1M/0.5M) x 1500 bins, 300 parameters, 12 "
                         "dials/event (uniform/general/compact splines, graphs, "
                         "norms, tabulated, kriged)") if args.workload == 2 and not args.events
                        else f"workload {args.workload}, {t.n_events} events",
            "events": int(t.n_events),
            "binned_events": int(counts["events"]),
            "dials": int(t.n_dials),
            "bins": int(t.n_bins),
            "parameters": int(t.n_parameters),
            "precision": args.precision,
            "launch": cm.launch_config(),
            "autotune": tuned,
            "lookup_tables": ("device refresh (gcm_add_table_update)" if device_tables
                              else "host callback + upload per step"),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "kernel": "fill_kernel",
            "algorithmic_bytes_per_launch": fill_bytes,
            "avg_kernel_ms": fill_ms,
        },
        "batched": batched,
        "batched_c3": batched_c3,
        "plain_c1": plain_c1,
        "surrogate_da": c5,
        "surrogate_mlp": mlp,
        "precision_c4": c4,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
