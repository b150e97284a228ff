#!/usr/bin/env python3
"""Benchmark of the MI355X period scan: LM pair-fits per second.

Workload (default): the plain FP64 full scan of one synthetic asteroid with
10 dense lightcurves of 150 points + 1,200 sparse points (2,700 points,
22 relative curves), lmax = mmax = 6 (49 coefficients), nrows = 6
(288 facets), 5,000 trial frequencies over 2-40 h x 10 initial poles =
50,000 LM problems with 54 free parameters, 50 LM iterations each.

A step fits one batch of frequencies (all 10 poles each) through the
product's scan entry point (ps_start); the default 10 steps cover the whole
5,000-frequency scan.  The inputs are resident on the device before the
timed region; the convexity-weight search runs once, untimed, to fix conwR.

The JSON line also carries the fit kernel's roofline (algorithmic FP64
flops of the fitted pairs / kernel time measured with HIP events on the
library's stream) and a CPU baseline: the repository's CPU FP64 oracle on a
seeded sample of pairs of the same workload, on the host's cores.

    python bench.py [--gpus 1] [--steps 10] [--warmup 1] [--cpu-pairs 32] [--no-cpu]
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


def measured_traffic(preset: str, pairs_per_launch: float):
    try:
        with open(TRAFFIC_FILE) as f:
            e = json.load(f).get(preset)
    except (OSError, ValueError):
        return None
    if not e:
        return None
    return round((e["fetch_bytes_per_pair"] + e["write_bytes_per_pair"]) * pairs_per_launch)


def _cores() -> int:
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--preset", default="c2")
    ap.add_argument("--cpu-pairs", type=int, default=128)
    ap.add_argument("--max-freqs", type=int, default=0, help="scan only the first N frequencies (0 = all)")
    ap.add_argument("--conw-r", type=float, default=0.0, help="use this conwR instead of searching for it")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--work-dir", default=os.path.join(ROOT, "gpurun_out", "bench"))
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run 5 plain device copies of MIB MiB (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.copy_probe > 0:
        return copy_probe(args.copy_probe)
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
    batch = max(1, -(-n_max // steps))  # ceil: the default 10 steps cover the whole scan
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
    scan.close()
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(path, conw_r, n_max, args.cpu_pairs)
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
    _, wall = O.fit_pairs(nl, ml, conw_r, conw_r, threads=cores)
    return {
        "value": round(n_pairs / wall, 4),
        "unit": "pair-fits/s",
        "cores": cores,
        "kind": "port",
        "sample": f"{n_pairs} seeded random (frequency, pole) pairs of the same scan, {wall:.1f} s wall "
                  f"on {cores} threads of the host CPU; the full scan would take "
                  f"{n_max * 10 / (n_pairs / wall) / 3600:.2f} h",
    }


if __name__ == "__main__":
    sys.exit(main())
