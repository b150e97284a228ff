"""Trains a warm-start network on the GPU and reports what it gains on
held-out asteroids.

    python -m ps_claude_opt_amd.warm_train --asteroids 20000 --pairs 40 \
        --out ps_claude_opt_amd/data/warm_lmax6.npz
    python -m ps_claude_opt_amd.warm_train --preset c5s --asteroids 2000 --pairs 40 \
        --out ps_claude_opt_amd/data/warm_lmax8.npz

Training data: cold LM fits (the input file's own 50-iteration stop mode) of
seeded random (frequency, pole) pairs of synthetic asteroids of the preset
(c2: config-2-style, lmax 6; c5s: config 5's geometry with fewer sparse
points, lmax 8; seeds 20261203 + 4000 + i: disjoint from the benchmark's
asteroids), each
fitted with its initial convexity weight conw / escl^2.  Features and targets
as in warm.py; the network is trained by the library's kernels (BF16 / FP32
accumulation, Adam).  Validation: on held-out asteroids, the rms reached by
a few LM iterations from the network's start versus from the reference's
start, both against the full cold fit.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from multiprocessing import Pool

import numpy as np


def _write_one(args):
    path, index, preset = args
    from ps_claude_opt_amd import synth

    truth = synth.write_asteroid(path, preset, synth.config_seed(4, index))
    return path, truth["period_h"]


def _one_thread():
    # each worker renders with one BLAS thread (the workers are the parallelism)
    try:
        from threadpoolctl import threadpool_limits

        threadpool_limits(1)
    except ImportError:
        pass


def generate(directory: str, first: int, count: int, workers: int, preset: str = "c2"):
    os.makedirs(directory, exist_ok=True)
    jobs = [(os.path.join(directory, f"a{first + i:05d}.in"), first + i, preset) for i in range(count)]
    out = []
    t0 = time.time()
    with Pool(workers, initializer=_one_thread) as pool:
        for r in pool.imap(_write_one, jobs, chunksize=16):
            out.append(r)
            if len(out) % 1000 == 0:
                print(f"# {len(out)} of {count} asteroids written, {time.time() - t0:.0f} s", flush=True)
    return out


def true_freq_index(scan, periods_h) -> np.ndarray:
    """Each problem's scan frequency nearest to its true rotation frequency
    (the scan's periods are 1 / (f_start - (n - 1) f_step) hours)."""
    from ps_claude_opt_amd import screen

    return np.array([screen.true_index(b, p) for b, p in zip(scan.batch_info, periods_h)])


def initial_conw(scan) -> np.ndarray:
    return np.array([b["conw"] / b["escl"] / b["escl"] for b in scan.batch_info])


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--asteroids", type=int, default=800)
    ap.add_argument("--preset", default="c2", help="synthetic asteroid preset (c2: lmax 6; c5s: lmax 8)")
    ap.add_argument("--validation", type=int, default=16)
    ap.add_argument("--pairs", type=int, default=40, help="training pairs per asteroid")
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--batch", type=int, default=512)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--near-frac", type=float, default=0.0,
                    help="fraction of each asteroid's pairs drawn near its true frequency")
    ap.add_argument("--near-width", type=int, default=2, help="frequency steps either side of the true one")
    ap.add_argument("--workers", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--work-dir", default="gpurun_out/warm")
    ap.add_argument("--out", default="gpurun_out/warm/warm_lmax6.npz")
    args = ap.parse_args(argv)

    import ps_claude_opt_amd as ps
    from ps_claude_opt_amd import warm

    t0 = time.time()
    train = generate(os.path.join(args.work_dir, "train"), 0, args.asteroids, args.workers, args.preset)
    val = generate(os.path.join(args.work_dir, "val"), 100000, args.validation, args.workers, args.preset)
    train_paths, val_paths = [p for p, _ in train], [p for p, _ in val]
    print(f"# generated {len(train_paths)} + {len(val_paths)} asteroids in {time.time() - t0:.1f} s", flush=True)
    report = dict(preset=args.preset, asteroids=args.asteroids, pairs_per_asteroid=args.pairs, epochs=args.epochs, batch=args.batch,
                  lr=args.lr, near_frac=args.near_frac, near_width=args.near_width)
    with ps.Scan(0) as s:
        t0 = time.time()
        s.prepare_files(train_paths)
        near = true_freq_index(s, [t for _, t in train])
        feats, tg, res, prob = warm.make_training_set(s, args.pairs, initial_conw(s), seed=11, near_n=near,
                                                      near_frac=args.near_frac, near_width=args.near_width)
        report.update(training_pairs=int(len(feats)), fit_s=round(time.time() - t0, 2))
        print(f"# {len(feats)} training pairs in {time.time() - t0:.1f} s", flush=True)
        t0 = time.time()
        ws, losses = warm.train_warm_start(s, feats, tg, prob, epochs=args.epochs, batch=args.batch, lr=args.lr,
                                           meta=dict(features=int(feats.shape[1]), outputs=int(tg.shape[1]),
                                                     asteroids=args.asteroids, seed_base="config_seed(4, i)"))
        report.update(train_s=round(time.time() - t0, 2), loss_first=float(losses[0]), loss_last=float(losses[-1]),
                      epochs_run=len(losses), val_loss=ws.meta["val_loss"],
                      shrink_mean=float(ws.shrink.mean()))
        print(f"# trained in {time.time() - t0:.1f} s: loss {losses[0]:.4f} -> {losses[-1]:.4f}", flush=True)
        os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
        ws.save(args.out)
        # validation on held-out asteroids: random pairs, and pairs within
        # near_width frequency steps of the true frequency
        s.prepare_files(val_paths)
        rng = np.random.default_rng(99)
        k = 200
        pr = np.repeat(np.arange(len(val_paths)), k)
        ml = rng.integers(1, 11, pr.size)
        cw = initial_conw(s)[pr]
        vnear = true_freq_index(s, [t for _, t in val])
        sets = {"random": np.concatenate([rng.integers(1, b["n_max"] + 1, k) for b in s.batch_info]),
                "near": np.concatenate([np.clip(vnear[i] + rng.integers(-args.near_width, args.near_width + 1, k), 1,
                                                b["n_max"]) for i, b in enumerate(s.batch_info)])}
        net = ws.network(s)
        val = {}
        for name, nl in sets.items():
            full, _ = s.fit_items_from(pr, nl, ml, cw, cw, None)
            a0 = ws.initial_params(s, net, pr, nl, ml)
            val[name] = {}
            for it in (5, 10, 15, 20, 30):
                c, _ = s.fit_items_from(pr, nl, ml, cw, cw, None, n_iter=it)
                w, _ = s.fit_items_from(pr, nl, ml, cw, cw, a0, n_iter=it)
                ok = np.isfinite(full["dev"]) & (full["dev"] < 1e30)
                gc = np.abs(c["dev"] - full["dev"])[ok]
                gw = np.abs(w["dev"] - full["dev"])[ok]
                val[name][str(it)] = dict(cold_median_gap=float(np.median(gc)), warm_median_gap=float(np.median(gw)),
                                          cold_within_5e4=float(np.mean(gc <= 5e-4)),
                                          warm_within_5e4=float(np.mean(gw <= 5e-4)))
                print(f"# {name} pairs, {it} iterations: |dev - dev_full| median cold {np.median(gc):.2e} "
                      f"warm {np.median(gw):.2e}; within 5e-4: cold {np.mean(gc <= 5e-4):.2f} "
                      f"warm {np.mean(gw <= 5e-4):.2f}", flush=True)
        net.close()
        report["validation"] = val
    with open(os.path.splitext(args.out)[0] + "_report.json", "w") as f:
        json.dump(report, f, indent=1)
    print(json.dumps(report), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
