#!/usr/bin/env python3
"""Benchmark of the MI355X period scan: LM pair-fits per second.

Workload (default): the plain FP64 full scan of one synthetic asteroid with
10 dense lightcurves of 150 points + 1,200 sparse points (2,700 points,
22 relative curves), lmax = mmax = 6 (49 coefficients), nrows = 6
(288 facets), 5,000 trial frequencies over 2-40 h x 10 initial poles =
50,000 LM problems with 54 free parameters, 50 LM iterations each.

A step fits one batch of 2,500 frequencies (all 10 poles each: 25,000
pair-fits, one persistent launch) through the product's scan entry point
(ps_start); the steps walk through the scan and wrap around, so the default
10 steps fit it 5 times.  The inputs are resident on the device before the
timed region; the convexity-weight search runs once, untimed, to fix conwR.

The JSON line also carries the fit kernel's roofline (algorithmic FP64
flops of the fitted pairs / kernel time measured with HIP events on the
library's stream) and a CPU baseline: the repository's CPU FP64 oracle on a
seeded sample of pairs of the same workload, on the host's cores.

    python bench.py [--gpus 1] [--steps 10] [--warmup 1] [--freqs-per-step 2500] [--cpu-pairs 512] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "LM pair-fits/sec on 1 MI355X (FP64, warm vs plain) + % FP64 peak, same poles"
FP64_PEAK_TFLOPS = 78.6  # MI355X dense FP64 (vector and matrix)
# HBM bytes per pair-fit of the fit kernel, from FETCH_SIZE / WRITE_SIZE
# counter passes scaled by a plain copy of known size.
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "ps_fit_kernel_traffic.json")
WARM_DATA = os.path.join(ROOT, "ps_claude_opt_amd", "data")  # warm.weights_for(lmax): warm_lmax<lmax>.npz


def measured_traffic(preset: str, pairs_per_launch: float):
    try:
        with open(TRAFFIC_FILE) as f:
            e = json.load(f).get(preset)
    except (OSError, ValueError):
        return None
    if not e:
        return None
    return round((e["fetch_bytes_per_pair"] + e["write_bytes_per_pair"]) * pairs_per_launch)


def _affinity() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def _cores() -> int:
    """The host cores this process has: its CPU affinity, or the share the
    host allots it (OMP_NUM_THREADS) when that is smaller."""
    n = _affinity()
    share = os.environ.get("OMP_NUM_THREADS", "")
    if share.isdigit() and int(share) > 0:
        n = min(n, int(share))
    return max(1, n)


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--freqs-per-step", type=int, default=2500,
                    help="trial frequencies per step (one launch); 0 = the scan divided into --steps steps")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--preset", default="c2")
    ap.add_argument("--cpu-pairs", type=int, default=512)
    ap.add_argument("--asteroids", type=int, default=64, help="batched workload (--preset c3): asteroids")
    ap.add_argument("--max-freqs", type=int, default=0, help="scan only the first N frequencies (0 = all)")
    ap.add_argument("--conw-r", type=float, default=0.0, help="use this conwR instead of searching for it")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="skip the launch-configuration autotuner")
    ap.add_argument("--no-warm", action="store_true", help="skip the warm-start comparison")
    ap.add_argument("--no-screen", action="store_true", help="skip the learned frequency screen's scan")
    ap.add_argument("--warm-iters", type=int, default=20, help="LM iterations per pair of the warm-started scan")
    ap.add_argument("--warm-refit", type=int, default=10,
                    help="frequencies of the warm-started scan always refitted with full cold fits")
    ap.add_argument("--warm-band", type=float, default=0.15,
                    help="also refit every frequency whose truncated rms is within this fraction of the best")
    ap.add_argument("--work-dir", default=os.path.join(ROOT, "gpurun_out", "bench"))
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run 5 plain device copies of MIB MiB (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.copy_probe > 0:
        return copy_probe(args.copy_probe)
    if args.preset == "c3":
        if args.gpus != 1:
            print("bench.py: this build measures one MI355X.", file=sys.stderr)
            return 2
        return bench_batch(args)
    if args.gpus != 1:
        print("bench.py: this build measures one MI355X; pairs shard across GPUs by frequency "
              "(see DESIGN.md), but multi-GPU runs are not part of it.", file=sys.stderr)
        return 2

    import numpy as np

    import ps_claude_opt_amd as ps
    from ps_claude_opt_amd import metrics, synth

    os.makedirs(args.work_dir, exist_ok=True)
    path = os.path.join(args.work_dir, f"{args.preset}.in")
    seed = synth.config_seed({"c1": 1, "c2": 2, "c5": 5}.get(args.preset, 0))
    truth = synth.write_asteroid(path, args.preset, seed)
    print(f"# input {path}: {truth['n_points']} points, {truth['n_curves']} curves, "
          f"{truth['n_freq']} frequencies", file=sys.stderr, flush=True)

    scan = ps.Scan(0)
    scan.prepare_file(path)
    info = scan.info
    n_max = info["n_max"]
    if args.max_freqs > 0:
        n_max = min(n_max, args.max_freqs)
    t0 = time.perf_counter()
    conw_r = args.conw_r if args.conw_r > 0 else scan.conw_search()
    t_conw = time.perf_counter() - t0
    print(f"# conw search: conwR {conw_r:.6g} in {t_conw:.2f} s", file=sys.stderr, flush=True)

    steps = max(1, args.steps)
    # one step = one launch of `batch` frequencies (10 poles each); the steps
    # walk through the scan and wrap around.  2,500 frequencies (25,000
    # pair-fits) per launch: the launch tail, where the last workgroups
    # finish their pairs while the others have no work left, costs about 4 %
    # at 5,000 pair-fits per launch and 0.2 % here.
    batch = max(1, min(n_max, args.freqs_per_step if args.freqs_per_step > 0 else -(-n_max // steps)))
    heavy = scan.info["ndata"] * scan.info["nfac"] > 4_000_000  # lmax 8: 20,800 points x 512 facets
    tune = autotune(scan, np.zeros(1, np.int32), np.array([n_max]), conw_r, args.no_tune, k=3072 if heavy else 12288)
    # warmup on the first batch (untimed)
    for _ in range(args.warmup):
        scan.start(conw_r, 1, batch)
    step_ms, kern_ms = [], []
    pairs = vis = iters = evals = 0
    lines = []
    t_total = 0.0
    for k in range(steps):
        n0 = 1 + (k * batch) % n_max
        cnt = min(batch, n_max - n0 + 1)
        t0 = time.perf_counter()
        lines += scan.start(conw_r, n0, cnt)
        dt = time.perf_counter() - t0
        t_total += dt
        st = scan.last_stats()
        step_ms.append(dt * 1e3)
        kern_ms.append(st["kernel_ms"])
        pairs += st["pairs"]
        vis += st["vis_pairs"]
        iters += st["iterations"]
        evals += st["evaluations"]
        print(f"# step {k}: freqs {n0}..{n0 + cnt - 1}, {st['pairs']} pairs, {dt:.3f} s "
              f"(kernel {st['kernel_ms'] / 1e3:.3f} s)", file=sys.stderr, flush=True)
    value = pairs / t_total
    N, F, C, M = info["ndata"] + 3, info["nfac"], info["ncoef"], info["mfit"]
    total_flops = metrics.flops_of_fits(N, F, C, M, vis, evals, iters)
    kern_s = sum(kern_ms) / 1e3
    achieved = total_flops / kern_s / 1e12
    best = min(lines, key=lambda L: L.rms)
    result = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "pair-fits/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(float(np.mean(step_ms)), 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded generator, ps_claude_opt_amd/synth.py)",
        "config": {
            "workload": f"plain FP64 scan {args.preset}: {info['ndata']} points in {info['n_curves']} relative "
                        f"curves, lmax={synth.PRESETS[args.preset].lmax}, {F} facets, {n_max} frequencies x 10 poles, "
                        f"{M} free parameters, {info['n_iter_max']} LM iterations",
            "pairs_timed": pairs,
            "frequencies_per_step": batch,
            "variant": "plain (cold start)",
            "conw_r": conw_r,
            "conw_search_s": round(t_conw, 3),
            "launch": tune,
        },
        "roofline": {
            "bound": "mfma",
            "achieved": round(achieved, 3),
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / FP64_PEAK_TFLOPS, 4),
            "traffic": measured_traffic(args.preset, pairs / steps),
            "traffic_unit": "bytes per launch (counter passes, profiles/ps_fit_kernel_traffic.json)",
            "kernel": "ps_fit_kernel",
            "kernel_ms_per_launch": round(float(np.mean(kern_ms)), 3),
            "gflop_per_pair_fit": round(total_flops / pairs / 1e9, 4),
            "visible_facets_per_point": round(vis / max(evals, 1) / N, 2),
        },
        "best_line": {"period_h": best.period_h, "rms": best.rms, "lambda": best.lambda_deg,
                      "beta": best.beta_deg, "truth_period_h": truth["period_h"],
                      "truth_lambda": truth["lambda_deg"], "truth_beta": truth["beta_deg"]},
    }
    weights = os.path.join(WARM_DATA, f"warm_lmax{synth.PRESETS[args.preset].lmax}.npz")
    if not os.path.exists(weights):
        weights = os.path.join(WARM_DATA, "warm_lmax6.npz")
    if not args.no_warm and n_max == info["n_max"] and os.path.exists(weights):
        result["warm"] = bench_warm(scan, conw_r, n_max, best, args.warm_iters, args.warm_refit, args.warm_band,
                                    weights, screen_ok=args.preset == "c2" and not args.no_screen)
    scan.close()
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(path, conw_r, n_max, args.cpu_pairs)
    print(json.dumps(result), flush=True)
    return 0


def bench_warm(scan, conw_r: float, n_max: int, plain_best, n_iter: int, refit: int, band: float,
               weights: str, screen_ok: bool = False) -> dict:
    """The learned warm start on the same scan: every pair runs n_iter LM
    iterations from the network's prediction, then the `refit` best
    frequencies and every frequency within `band` of the best truncated rms
    are refitted cold (the plain scan's own 50-iteration fits), which decide
    the reported winner.  The same truncated scan from the
    reference's start is timed beside it (what the network adds), and both
    winners are compared with the plain scan's (the same printed lambda and
    beta, period within 1e-5 relative, rms within 5e-4, chi2 within 0.5).
    How often each finds the plain winner on held-out asteroids is measured
    by ps_claude_opt_amd/warm_eval.py.  With screen_ok, the learned frequency
    screen's scan is timed as well: only its kept frequencies get truncated
    fits (config-2-style asteroids, the ones it was trained on)."""
    import numpy as np

    from ps_claude_opt_amd import warm

    ws = warm.WarmStart.load(weights)
    out = {"n_iter": n_iter, "refit_min": refit, "refit_band": band, "weights": os.path.relpath(weights, ROOT)}
    modes = [("warm", ws), ("truncated_cold", None)]
    if ws.y_mean.size != scan.info["ncoef"] - 1 + 4:
        # the shipped network predicts the lmax-6 coefficients only: the
        # truncated scan from the reference's start still runs beside the
        # plain one
        out["warm"] = {"skipped": f"the network predicts {ws.y_mean.size} parameters; this problem has "
                                  f"{scan.info['ncoef'] - 1 + 4}"}
        modes = modes[1:]
    net = ws.network(scan) if len(modes) == 2 else None
    for name, w in modes:
        t0 = time.perf_counter()
        res, refitted = warm.warm_scan(scan, w, net if w else None, conw_r, 1, n_max, n_iter, refit=refit, band=band)
        dt = time.perf_counter() - t0
        b = warm.best_line(res, 1, n_max, among=refitted)
        b["chi2"] = b["rms"] * b["rms"] * scan.info["ndata"]  # the output line's chi2 (start_CUDA.cu:1222-1226)
        agree = (round(b["lambda_deg"]) == round(plain_best.lambda_deg) and round(b["beta_deg"]) == round(plain_best.beta_deg)
                 and abs(b["period_h"] - plain_best.period_h) <= 1e-5 * plain_best.period_h
                 and abs(b["rms"] - plain_best.rms) <= 5e-4 and abs(b["chi2"] - plain_best.chi2) <= 0.5)
        out[name] = {"value": round(n_max * 10 / dt, 3), "unit": "pair-fits/s", "seconds": round(dt, 3),
                     "refitted_frequencies": int(len(refitted)),
                     "best_line": b, "agrees_with_plain": bool(agree)}
        print(f"# {name}: {n_max * 10} pairs in {dt:.2f} s, best {b['period_h']:.6f} h rms {b['rms']:.6f} "
              f"agrees {agree}", file=sys.stderr, flush=True)
    if net is not None:
        net.close()
    if screen_ok:
        # the learned frequency screen: fold features and the screen network
        # for every frequency, truncated cold fits of the kept ones, cold
        # refits as above (ps_claude_opt_amd/screen.py)
        from ps_claude_opt_amd import screen

        sc = screen.Screen.load()
        t0 = time.perf_counter()
        snet = sc.network(scan)
        res, refitted, kept = screen.screened_scan(scan, sc, snet, conw_r, 1, n_max, n_iter, refit=refit, band=band)
        snet.close()
        dt = time.perf_counter() - t0
        b = warm.best_line(res, 1, n_max, among=refitted)
        b["chi2"] = b["rms"] * b["rms"] * scan.info["ndata"]
        agree = (round(b["lambda_deg"]) == round(plain_best.lambda_deg) and round(b["beta_deg"]) == round(plain_best.beta_deg)
                 and abs(b["period_h"] - plain_best.period_h) <= 1e-5 * plain_best.period_h
                 and abs(b["rms"] - plain_best.rms) <= 5e-4 and abs(b["chi2"] - plain_best.chi2) <= 0.5)
        out["screened"] = {"value": round(n_max * 10 / dt, 3), "unit": "pair-fits/s", "seconds": round(dt, 3),
                           "keep": sc.keep, "kept_frequencies": int(kept.size),
                           "refitted_frequencies": int(len(refitted)), "weights": "ps_claude_opt_amd/data/screen.npz",
                           "best_line": b, "agrees_with_plain": bool(agree),
                           "vs_truncated_cold": round(out["truncated_cold"]["seconds"] / dt, 3)}
        print(f"# screened: {n_max * 10} pairs in {dt:.2f} s ({kept.size} of {n_max} frequencies fitted), best "
              f"{b['period_h']:.6f} h rms {b['rms']:.6f} agrees {agree}", file=sys.stderr, flush=True)
    return out


def autotune(scan, probs, n_max, conw, skip: bool, k: int = 12288) -> dict:
    """Picks the launch configuration (wavefronts per pair, workgroups per
    CU, facet-pass claim) on a seeded sample of k of the workload's pairs
    (12,288, or 3,072 for the lmax-8 workload, whose pairs cost 20 times
    more), before the timed region.  The configurations give bit-identical
    results."""
    import numpy as np

    if skip:
        return dict(scan.get_launch(), tuned=False)
    rng = np.random.default_rng(777)
    # a whole number of rounds of pairs for every launch shape on 256 CUs (2,
    # 3, 4, 6, 8 or 12 resident workgroups per CU all divide 12,288; 2, 3, 4
    # or 12 divide 3,072), so that no shape gains or loses from the size of
    # its last round
    pi = rng.integers(0, len(probs), k)
    p = np.asarray(probs)[pi]
    nl = np.array([rng.integers(1, int(np.asarray(n_max)[j]) + 1) for j in pi])
    ml = rng.integers(1, 11, k)
    cw = np.broadcast_to(np.asarray(conw, dtype=np.float64), (len(probs),))[pi] if np.ndim(conw) else conw
    best = scan.autotune(p, nl, ml, cw)
    print(f"# autotune: {best}", file=sys.stderr, flush=True)
    return dict(best, tuned=True, sample_pairs=k)


def bench_batch(args) -> int:
    """The batched ragged workload: 64 asteroids (300-8000 points, 1000-4000
    frequencies each) in one persistent launch per step.  A step fits the
    next --freqs-per-step frequencies of every asteroid (all 10 poles); the
    timed region covers --max-freqs frequencies per asteroid (default 100,
    a bounded slice of the 1.6 M-pair scan)."""
    import numpy as np

    import ps_claude_opt_amd as ps
    from ps_claude_opt_amd import metrics, synth

    d = os.path.join(args.work_dir, "c3")
    truths = synth.write_batch(d, args.asteroids)
    paths = [t["path"] for t in truths]
    scan = ps.Scan(0)
    scan.prepare_files(paths)
    bi = scan.batch_info
    n = len(bi)
    t0 = time.perf_counter()
    conw = np.full(n, args.conw_r) if args.conw_r > 0 else scan.conw_search_batch()
    t_conw = time.perf_counter() - t0
    print(f"# {n} asteroids, {sum(b['ndata'] for b in bi)} points; conw search {t_conw:.2f} s", file=sys.stderr,
          flush=True)
    per_ast = args.max_freqs if args.max_freqs > 0 else 100
    tune = autotune(scan, np.arange(n), np.array([b["n_max"] for b in bi]), conw, args.no_tune)
    steps = max(1, args.steps)
    fps = max(1, -(-per_ast // steps))

    def items(f0, cnt):
        prob, nl, ml = [], [], []
        for i, b in enumerate(bi):
            for f in range(f0, min(f0 + cnt, per_ast + 1, b["n_max"] + 1)):
                for m in range(1, 11):
                    prob.append(i)
                    nl.append(f)
                    ml.append(m)
        return np.array(prob), np.array(nl), np.array(ml)

    for _ in range(args.warmup):
        p_, n_, m_ = items(1, fps)
        scan.fit_items(p_, n_, m_, conw[p_])
    vis = np.zeros(n)
    ev = np.zeros(n)
    it = np.zeros(n)
    pairs = 0
    t_total = 0.0
    kern = []
    step_ms = []
    for k in range(steps):
        p_, n_, m_ = items(1 + k * fps, fps)
        if len(p_) == 0:
            break
        t0 = time.perf_counter()
        R = scan.fit_items(p_, n_, m_, conw[p_])
        dt = time.perf_counter() - t0
        t_total += dt
        step_ms.append(dt * 1e3)
        st = scan.last_stats()
        kern.append(st["kernel_ms"])
        pairs += len(p_)
        np.add.at(vis, p_, R["n_vis"])
        np.add.at(it, p_, R["niter"])
        np.add.at(ev, p_, np.where(R["niter"] > 0, R["niter"] + 1, 0))
        print(f"# step {k}: {len(p_)} pairs, {dt:.3f} s (kernel {st['kernel_ms'] / 1e3:.3f} s)", file=sys.stderr,
              flush=True)
    info0 = scan.info
    F, C, M = info0["nfac"], info0["ncoef"], info0["mfit"]
    flops = sum(metrics.flops_of_fits(bi[i]["ndata"] + 3, F, C, M, vis[i], ev[i], it[i]) for i in range(n))
    kern_s = sum(kern) / 1e3
    achieved = flops / kern_s / 1e12
    scan.close()
    result = {
        "metric": METRIC,
        "value": round(pairs / t_total, 3),
        "unit": "pair-fits/s",
        "n_gpus": 1,
        "steps": len(step_ms),
        "warmup": args.warmup,
        "ms_per_step": round(float(np.mean(step_ms)), 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded generator, ps_claude_opt_amd/synth.py)",
        "config": {
            "workload": f"batched ragged scan: {n} asteroids, {min(b['ndata'] for b in bi)}-"
                        f"{max(b['ndata'] for b in bi)} points, lmax=6, {F} facets, first {per_ast} of "
                        f"{min(b['n_max'] for b in bi)}-{max(b['n_max'] for b in bi)} frequencies x 10 poles each, "
                        f"one persistent launch per step",
            "pairs_timed": pairs,
            "conw_search_s": round(t_conw, 3),
            "launch": tune,
        },
        "roofline": {
            "bound": "mfma",
            "achieved": round(achieved, 3),
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / FP64_PEAK_TFLOPS, 4),
            "traffic": None,
            "kernel": "ps_fit_kernel",
            "kernel_ms_per_launch": round(float(np.mean(kern)), 3),
        },
    }
    print(json.dumps(result), flush=True)
    return 0


def copy_probe(mib: int) -> int:
    """Five device-to-device copies of a known size.  Under a counter
    profiler their FETCH_SIZE / WRITE_SIZE per dispatch give the factor
    between the counters and the bytes a kernel really moves."""
    import torch

    n = mib * (1 << 20) // 8
    x = torch.arange(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    ok = bool(torch.equal(x, y))
    print(json.dumps({"copy_probe_bytes_read": n * 8, "copy_probe_bytes_written": n * 8, "ok": ok}), flush=True)
    return 0 if ok else 1


def cpu_baseline(path: str, conw_r: float, n_max: int, n_pairs: int) -> dict:
    """The CPU FP64 oracle on a seeded sample of pairs of the same workload."""
    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    O = pyoracle.Problem(path)
    rng = np.random.default_rng(12345)
    nl = rng.integers(1, n_max + 1, n_pairs)
    ml = rng.integers(1, 11, n_pairs)
    cores = _cores()
    print(f"# cpu baseline: {n_pairs} pairs on {cores} threads", file=sys.stderr, flush=True)
    _, wall = O.fit_pairs(nl, ml, conw_r, conw_r, threads=cores)
    n1 = max(1, min(8, n_pairs))
    _, wall1 = O.fit_pairs(nl[:n1], ml[:n1], conw_r, conw_r, threads=1)
    return {
        "value": round(n_pairs / wall, 4),
        "unit": "pair-fits/s",
        "cores": cores,
        "nproc": os.cpu_count(),
        "affinity": _affinity(),
        "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
        "kind": "port",
        "sample": f"{n_pairs} seeded random (frequency, pole) pairs of the same scan, {wall:.1f} s wall "
                  f"on {cores} threads (the host cores this process has, {_cpu_model()}); the full scan would take "
                  f"{n_max * 10 / (n_pairs / wall) / 3600:.2f} h",
        "single_core_value": round(n1 / wall1, 4),
        "single_core_sample": f"the first {n1} of those pairs on one thread, {wall1:.1f} s",
    }


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


if __name__ == "__main__":
    sys.exit(main())
