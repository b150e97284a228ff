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


def winner_ranks(scan, ws, net, conw_r: float, n_max: int, iters, dump=None, lines=None, modes=("warm", "cold")):
    """(plain winner frequency index, {mode+n_iter: rank of the winner});
    dump (a dict) receives the per-frequency best rms of the plain scan
    ("plain") and of every truncated screen ("warm20", "cold15", ...).
    lines: the plain scan's lines if already run."""
    from . import warm

    if lines is None:
        lines = scan.start(conw_r, 1, n_max)
    rms = np.array([L.rms for L in lines])
    f_best = int(np.nanargmin(rms))
    if dump is not None:
        dump["plain"] = rms
    n = np.repeat(np.arange(1, n_max + 1), 10)
    m = np.tile(np.arange(1, 11), n_max)
    p = np.zeros(n.size, np.int32)
    a0w = ws.initial_params(scan, net, p, n, m) if "warm" in modes else None
    ranks = {}
    for mode, a0 in [x for x in (("warm", a0w), ("cold", None)) if x[0] in modes]:
        for it in iters:
            res, _ = scan.fit_items_from(p, n, m, conw_r, conw_r, a0, n_iter=it)
            best = warm.best_lines(res, n_max)
            r = res["dev"].reshape(n_max, 10)[np.arange(n_max), best]
            if dump is not None:
                dump[f"{mode}{it}"] = r
            order = np.argsort(np.where(np.isnan(r), np.inf, r), kind="stable")
            ranks[f"{mode}{it}"] = int(np.where(order == f_best)[0][0])
            # the adaptive refit set of warm.warm_scan: found, and its size
            fr = warm.refit_frequencies(scan.info, r, 1) - 1
            ranks[f"{mode}{it}_adaptive"] = [bool(f_best in set(fr.tolist())), int(fr.size)]
    return f_best, ranks


def screen_check(scan, sc, net, conw_r: float, n_max: int, lines, n_iter: int = 20, dump=None) -> dict:
    """The learned screen on one asteroid whose plain scan gave `lines`:
    the plain winner's rank under the screen's score, whether the kept set
    holds it, and whether the screened scan (screen.screened_scan) reports
    the plain winner's line."""
    from . import screen, warm

    rms = np.array([L.rms for L in lines])
    f_best = int(np.nanargmin(rms))
    scores = sc.scores(scan, net, 1, n_max)
    if dump is not None:
        dump["screen"] = scores
    t0 = time.perf_counter()
    res, refitted, kept = screen.screened_scan(scan, sc, net, conw_r, 1, n_max, n_iter=n_iter)
    dt = time.perf_counter() - t0
    b = warm.best_line(res, 1, n_max, among=refitted)
    L = lines[f_best]
    agrees = (b["freq_index"] == f_best + 1 and b["rms"] == L.rms and b["lambda_deg"] == L.lambda_deg
              and b["beta_deg"] == L.beta_deg)
    return dict(winner_rank=screen.rank_of(scores, f_best), kept=int(kept.size),
                in_kept=bool(f_best + 1 in set(kept.tolist())), refitted=int(refitted.size),
                screened_winner=b["freq_index"], agrees=bool(agrees), screened_s=round(dt, 3))


def screen_batch_run(paths, sc, n_iter: int = 20) -> list:
    """One prepared batch of asteroids: the plain batched scan and the
    batched screened scan, both timed; per asteroid whether the screened
    scan reports the plain winner's line."""
    import ps_claude_opt_amd as ps
    from . import screen, warm

    with ps.Scan(0) as s:
        s.prepare_files(paths)
        cw = s.conw_search_batch()
        t0 = time.perf_counter()
        plain = s.start_batch(cw)
        t_plain = time.perf_counter() - t0
        t0 = time.perf_counter()
        net = sc.network(s)
        got = screen.screened_scan_batch(s, sc, net, cw, n_iter=n_iter)
        net.close()
        t_screen = time.perf_counter() - t0
    pairs = sum(10 * len(L) for L in plain)
    rows = []
    for i, (res, refitted, kept) in enumerate(got):
        rms = np.array([L.rms for L in plain[i]])
        f_best = int(np.nanargmin(rms))
        b = warm.best_line(res, 1, len(plain[i]), among=refitted)
        L = plain[i][f_best]
        rows.append(dict(winner_freq=f_best + 1, kept=bool(f_best + 1 in set(kept.tolist())), refitted=int(refitted.size),
                         agrees=bool(b["freq_index"] == f_best + 1 and b["rms"] == L.rms
                                     and b["lambda_deg"] == L.lambda_deg and b["beta_deg"] == L.beta_deg)))
    return [dict(asteroids=len(paths), pairs=pairs, plain_s=round(t_plain, 3), screened_s=round(t_screen, 3),
                 plain_pair_fits_per_s=round(pairs / t_plain, 1), screened_pair_fits_per_s=round(pairs / t_screen, 1),
                 rows=rows)]


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
    ap.add_argument("--modes", default="warm,cold,screen",
                    help="warm / cold: truncated screens from the network's and the reference's start; "
                         "screen: the learned frequency screen")
    ap.add_argument("--screen", default=None, help="screen weights (default screen.DEFAULT_WEIGHTS)")
    ap.add_argument("--extra-screens", default="", help="comma-separated screen weights whose scores are dumped too")
    ap.add_argument("--batch-size", type=int, default=16, help="asteroids per prepared batch (--modes screen_batch)")
    args = ap.parse_args(argv)
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import ps_claude_opt_amd as ps
    from ps_claude_opt_amd import screen, synth, warm

    modes = set(args.modes.split(","))
    sc = screen.Screen.load(args.screen or screen.DEFAULT_WEIGHTS) if modes & {"screen", "screen_batch"} else None
    iters = [int(x) for x in args.iters.split(",")]
    refits = [int(x) for x in args.refit.split(",")]
    ws = warm.WarmStart.load(args.weights or warm.DEFAULT_WEIGHTS)
    os.makedirs(args.work_dir, exist_ok=True)
    rows = []
    dumps = {}
    if modes == {"screen_batch"}:
        batches = []
        for b0 in range(args.first, args.first + args.asteroids, args.batch_size):
            paths = []
            for i in range(b0, min(args.first + args.asteroids, b0 + args.batch_size)):
                path = os.path.join(args.work_dir, f"asteroid_{i}.in")
                synth.write_asteroid(path, "c2", synth.config_seed(2, 300 + i))
                paths.append(path)
            batches += screen_batch_run(paths, sc)
            print(json.dumps({k: v for k, v in batches[-1].items() if k != "rows"}), flush=True)
        allr = [r for b in batches for r in b["rows"]]
        summary = dict(screen_batch=dict(
            asteroids=len(allr), batch_size=args.batch_size, keep=sc.keep,
            screened_scan_reports_plain_winner=f"{sum(r['agrees'] for r in allr)}/{len(allr)}",
            plain_pair_fits_per_s=round(sum(b["pairs"] for b in batches) / sum(b["plain_s"] for b in batches), 1),
            screened_pair_fits_per_s=round(sum(b["pairs"] for b in batches) / sum(b["screened_s"] for b in batches), 1)))
        report = dict(asteroids=args.asteroids, first=args.first, modes=["screen_batch"], batches=batches,
                      found_winner=summary)
        os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
        with open(args.out, "w") as f:
            json.dump(report, f, indent=1)
        print(json.dumps(summary, indent=1), flush=True)
        return 0
    for i in range(args.first, args.first + args.asteroids):
        path = os.path.join(args.work_dir, f"asteroid_{i}.in")
        synth.write_asteroid(path, "c2", synth.config_seed(2, 300 + i))
        t0 = time.time()
        d = {} if args.dump else None
        with ps.Scan(0) as s:
            s.prepare_file(path)
            cw = s.conw_search()
            row = dict(asteroid=i)
            lines = s.start(cw, 1, s.info["n_max"])
            if modes & {"warm", "cold"}:
                net = ws.network(s) if "warm" in modes else None
                f_best, ranks = winner_ranks(s, ws, net, cw, s.info["n_max"], iters, d, lines, modes)
                if net is not None:
                    net.close()
                row.update(winner_freq=f_best + 1, ranks=ranks)
            if sc is not None:
                snet = sc.network(s)
                row["screen"] = screen_check(s, sc, snet, cw, s.info["n_max"], lines, dump=d)
                for j, ep in enumerate(x for x in args.extra_screens.split(",") if x):
                    esc = screen.Screen.load(ep)
                    enet = esc.network(s)
                    if d is not None:
                        d[f"screen_x{j}"] = esc.scores(s, enet, 1, s.info["n_max"])
                    enet.close()
                snet.close()
                row.setdefault("winner_freq", int(np.nanargmin([L.rms for L in lines])) + 1)
        if d is not None:
            d["plain"] = np.array([L.rms for L in lines])
            dumps.update({f"a{i}_{k}": v for k, v in d.items()})
        row["seconds"] = round(time.time() - t0, 1)
        rows.append(row)
        print(json.dumps(rows[-1]), flush=True)
    if args.dump:
        os.makedirs(os.path.dirname(os.path.abspath(args.dump)), exist_ok=True)
        np.savez_compressed(args.dump, **dumps)
    summary = {}
    for mode in ("warm", "cold"):
        if mode not in modes:
            continue
        for it in iters:
            for rf in refits:
                ok = sum(r["ranks"][f"{mode}{it}"] < rf for r in rows)
                summary[f"{mode} n_iter={it} refit={rf}"] = f"{ok}/{len(rows)}"
            ad = [r["ranks"][f"{mode}{it}_adaptive"] for r in rows]
            summary[f"{mode} n_iter={it} adaptive refit set"] = (
                f"{sum(a[0] for a in ad)}/{len(rows)}, mean {np.mean([a[1] for a in ad]):.0f} "
                f"max {max(a[1] for a in ad)} frequencies refitted")
    if sc is not None:
        sr = [r["screen"] for r in rows]
        summary["screen"] = dict(
            keep=sc.keep, kept=sr[0]["kept"] if sr else 0,
            plain_winner_kept=f"{sum(x['in_kept'] for x in sr)}/{len(sr)}",
            screened_scan_reports_plain_winner=f"{sum(x['agrees'] for x in sr)}/{len(sr)}",
            winner_rank_max=max(x["winner_rank"] for x in sr), winner_rank_median=float(np.median([x["winner_rank"] for x in sr])),
            refitted_mean=float(np.mean([x["refitted"] for x in sr])), refitted_max=max(x["refitted"] for x in sr),
            screened_s_mean=float(np.mean([x["screened_s"] for x in sr])))
    report = dict(asteroids=args.asteroids, first=args.first, modes=sorted(modes), iters=iters, refits=refits,
                  rows=rows, found_winner=summary)
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    with open(args.out, "w") as f:
        json.dump(report, f, indent=1)
    print(json.dumps(summary, indent=1), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
