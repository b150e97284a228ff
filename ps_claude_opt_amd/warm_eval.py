"""Held-out check of the warm-started scan: does it find the plain scan's
winner, and what does the network add over the same truncated scan from
the reference's start?

    python -m ps_claude_opt_amd.warm_eval --asteroids 16 --iters 15,20 \
        --out ps_claude_opt_amd/data/warm_eval_report.json

For each seeded config-2-style asteroid (seeds 20261203 + 2000 + 300 + i,
disjoint from the benchmark's and the training asteroids), the plain scan
(every pair to the input file's own stop condition) gives the winning
frequency.  Then every pair runs n_iter LM iterations, once from the
network's start (warm) and once from the reference's start (cold), and the
report records where the plain winner ranks among the frequencies ordered by
their best truncated rms.

A warm scan that refits its best-ranked frequencies with full cold fits,
and takes its winner among those refits (plain fits, bit for bit),
reproduces the plain winner exactly when the winner is in the refit set:
for a fixed count when its rank is below `refit`; the report also gives
warm.refit_set's adaptive set (at least 10, and every frequency within 10 %
of the best truncated rms).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np


def winner_ranks(scan, ws, net, conw_r: float, n_max: int, iters, dump=None):
    """(plain winner frequency index, {mode+n_iter: rank of the winner});
    dump (a dict) receives the per-frequency best rms of the plain scan
    ("plain") and of every truncated screen ("warm20", "cold15", ...)."""
    from . import warm

    lines = scan.start(conw_r, 1, n_max)
    rms = np.array([L.rms for L in lines])
    f_best = int(np.nanargmin(rms))
    if dump is not None:
        dump["plain"] = rms
    n = np.repeat(np.arange(1, n_max + 1), 10)
    m = np.tile(np.arange(1, 11), n_max)
    p = np.zeros(n.size, np.int32)
    a0w = ws.initial_params(scan, net, p, n, m)
    ranks = {}
    for mode, a0 in (("warm", a0w), ("cold", None)):
        for it in iters:
            res, _ = scan.fit_items_from(p, n, m, conw_r, conw_r, a0, n_iter=it)
            best = warm.best_lines(res, n_max)
            r = res["dev"].reshape(n_max, 10)[np.arange(n_max), best]
            if dump is not None:
                dump[f"{mode}{it}"] = r
            order = np.argsort(np.where(np.isnan(r), np.inf, r), kind="stable")
            ranks[f"{mode}{it}"] = int(np.where(order == f_best)[0][0])
            # the adaptive refit set of warm.warm_scan: found, and its size
            fr = warm.refit_set(r)
            ranks[f"{mode}{it}_adaptive"] = [bool(f_best in set(fr.tolist())), int(fr.size)]
    return f_best, ranks


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--asteroids", type=int, default=16)
    ap.add_argument("--iters", default="15,20")
    ap.add_argument("--refit", default="10,50", help="refit counts to summarise")
    ap.add_argument("--work-dir", default="gpurun_out/warm_eval")
    ap.add_argument("--out", default="gpurun_out/warm_eval/warm_eval_report.json")
    ap.add_argument("--weights", default=None)
    ap.add_argument("--first", type=int, default=0, help="index of the first held-out asteroid")
    ap.add_argument("--dump", default=None, help="npz of per-frequency best rms of every scan")
    args = ap.parse_args(argv)
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import ps_claude_opt_amd as ps
    from ps_claude_opt_amd import synth, warm

    iters = [int(x) for x in args.iters.split(",")]
    refits = [int(x) for x in args.refit.split(",")]
    ws = warm.WarmStart.load(args.weights or warm.DEFAULT_WEIGHTS)
    os.makedirs(args.work_dir, exist_ok=True)
    rows = []
    dumps = {}
    for i in range(args.first, args.first + args.asteroids):
        path = os.path.join(args.work_dir, f"asteroid_{i}.in")
        synth.write_asteroid(path, "c2", synth.config_seed(2, 300 + i))
        t0 = time.time()
        d = {} if args.dump else None
        with ps.Scan(0) as s:
            s.prepare_file(path)
            cw = s.conw_search()
            net = ws.network(s)
            f_best, ranks = winner_ranks(s, ws, net, cw, s.info["n_max"], iters, d)
            net.close()
        if d is not None:
            dumps.update({f"a{i}_{k}": v for k, v in d.items()})
        rows.append(dict(asteroid=i, winner_freq=f_best + 1, ranks=ranks, seconds=round(time.time() - t0, 1)))
        print(json.dumps(rows[-1]), flush=True)
    if args.dump:
        os.makedirs(os.path.dirname(os.path.abspath(args.dump)), exist_ok=True)
        np.savez_compressed(args.dump, **dumps)
    summary = {}
    for mode in ("warm", "cold"):
        for it in iters:
            for rf in refits:
                ok = sum(r["ranks"][f"{mode}{it}"] < rf for r in rows)
                summary[f"{mode} n_iter={it} refit={rf}"] = f"{ok}/{len(rows)}"
            ad = [r["ranks"][f"{mode}{it}_adaptive"] for r in rows]
            summary[f"{mode} n_iter={it} adaptive refit set"] = (
                f"{sum(a[0] for a in ad)}/{len(rows)}, mean {np.mean([a[1] for a in ad]):.0f} "
                f"max {max(a[1] for a in ad)} frequencies refitted")
    report = dict(asteroids=args.asteroids, iters=iters, refits=refits, rows=rows, found_winner=summary)
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    with open(args.out, "w") as f:
        json.dump(report, f, indent=1)
    print(json.dumps(summary, indent=1), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
