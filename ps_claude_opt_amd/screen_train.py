"""Trains the learned frequency screen on the GPU and reports how it ranks
the true frequency of held-out asteroids.

    python -m ps_claude_opt_amd.screen_train --asteroids 2048 \
        --out ps_claude_opt_amd/data/screen.npz

Training data: seeded config-2-style asteroids (seeds config_seed(4, 50000 +
i): disjoint from the benchmark's, the warm-start network's training set and
the held-out evaluation's).  Per asteroid, `--random` trial frequencies
drawn uniformly and every frequency within `--near` steps of the true one
and of its aliases (2, 1/2, 1.5, 3, 2.5 and 2/3 times the period: the plain
scan's winner is often one of them); all 10 poles of each are fitted cold,
as the plain scan fits them (`--iters` > 0: that many LM iterations
instead), with the asteroid's initial convexity weight.  A frequency's
target is the log of its best rms over the poles
(start_CUDA.cu:1200-1210's choice of the pole) minus the median of that log
over the asteroid's random frequencies, capped at 0 (the spread among the
poorer half of the frequencies does not matter for the ranking of the best
ones); its input the fold features
(ps_screen_features).  The network (GELU MLP, screen.N_LAYERS x N_HIDDEN) is
trained by the library's kernels (BF16 operands, FP32 accumulation, Adam),
with early stopping on held-out asteroids.

The inputs are screen.network_inputs of the asteroid's whole grid (the fold
features and each power relative to the grid's median).

Report: on the validation asteroids, the rank among all trial frequencies of
the sampled near-alias frequency with the lowest rms (where the plain
scan's winner lies), under the network's score and under single fold
features alone (what the network adds over the feature it leans on most).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np


ALIASES = (1.0, 2.0, 0.5, 1.5, 3.0, 2.5, 2.0 / 3.0)  # period multiples whose neighbourhood is sampled


def build_set(scan, truths, n_random: int, near: int, iters: int, seed: int, keep_full):
    """(features, targets, asteroid ids, is_near, {asteroid: full-grid
    features} for the asteroids keep_full(i) selects) of the scan's batch."""
    from ps_claude_opt_amd import screen, warm
    from ps_claude_opt_amd.warm_train import initial_conw

    rng = np.random.default_rng(seed)
    cw_all = initial_conw(scan)
    prob, nl, near_flag, feats_full, rows_f = [], [], [], {}, []
    for i, (b, t) in enumerate(zip(scan.batch_info, truths)):
        fr = rng.choice(np.arange(1, b["n_max"] + 1), size=min(n_random, b["n_max"]), replace=False)
        # windows around the true period and its aliases: a convex model
        # fits a rotational lightcurve about as well at twice, half and 1.5
        # times the period, and the plain scan's winner is often one of them
        nr = []
        for mult in ALIASES:
            nt = screen.true_index(b, t * mult)
            lo, hi = max(1, nt - near), min(b["n_max"], nt + near)
            if abs(24.0 / (t * mult) - 24.0 * (b["f_start"] - (nt - 1) * b["f_step"])) < 24.0 * b["f_step"]:
                nr += [n for n in range(lo, hi + 1) if n not in nr]
        nr = np.array(nr, np.int64)
        freqs = np.concatenate([fr, nr])
        F = screen.network_inputs(scan.screen_features(1, b["n_max"], prob=i))
        if keep_full(i):
            feats_full[i] = F
        rows_f.append(F[freqs - 1])
        prob += [i] * freqs.size
        nl += list(freqs)
        near_flag += [False] * fr.size + [True] * nr.size
    prob, nl, near_flag = np.array(prob), np.array(nl), np.array(near_flag)
    p10 = np.repeat(prob, 10)
    n10 = np.repeat(nl, 10)
    m10 = np.tile(np.arange(1, 11), prob.size)
    cw = cw_all[p10]
    res, _ = scan.fit_items_from(p10, n10, m10, cw, cw, None, n_iter=iters)
    best = warm.best_lines(res, prob.size)
    rms = res["dev"].reshape(prob.size, 10)[np.arange(prob.size), best]
    lr = np.log(np.where(np.isfinite(rms) & (rms > 0) & (rms < 1e30), rms, np.nan))
    tg = np.zeros(prob.size, np.float32)
    for i in np.unique(prob):
        sel = prob == i
        med = np.nanmedian(lr[sel & ~near_flag])
        tg[sel] = lr[sel] - med
    # only how far a frequency falls below the typical one matters for the
    # ranking of the best ones: the spread among poor frequencies is cut off
    tg = np.minimum(tg, 0.0)
    ok = np.isfinite(tg)
    return np.concatenate(rows_f)[ok], tg[ok], prob[ok], near_flag[ok], feats_full, nl[ok]


def near_rank(scores: np.ndarray, n_best: int) -> int:
    """Rank (0 = best) among all trial frequencies of frequency n_best
    (1-based): the sampled frequency near the true period or an alias whose
    best truncated rms is lowest, the stand-in for the plain scan's
    winner."""
    from ps_claude_opt_amd import screen

    return screen.rank_of(scores, n_best - 1)


def is_validation(a: int) -> bool:
    """Asteroid a of the training run is held out (one in eight)."""
    return a % 8 == 7


def train(scan, x, y, groups, n_hidden, n_layers, epochs, batch, lr, seed=0, patience=20):
    from ps_claude_opt_amd import screen, warm

    val = np.array([is_validation(int(a)) for a in groups])
    tr = ~val
    xm, xs = x[tr].mean(0), x[tr].std(0)
    xs = np.where(xs > 1e-6, xs, 1.0).astype(np.float32)
    ym, ys = float(y[tr].mean()), float(y[tr].std() or 1.0)
    xt, yt = (x[tr] - xm) / xs, ((y[tr] - ym) / ys)[:, None]
    xv, yv = (x[val] - xm) / xs, ((y[val] - ym) / ys)[:, None]
    net = warm.Mlp(scan, x.shape[1], 1, n_hidden, n_layers, seed=seed)
    best, best_loss, since, losses = net.get_params(), np.inf, 0, []
    for e in range(epochs):
        losses.append(float(net.train(xt, yt, 1, batch, lr, seed + 1 + e)[0]))
        vl = float(np.mean((net.forward(xv) - yv) ** 2))
        if vl < best_loss:
            best, best_loss, since = net.get_params(), vl, 0
        else:
            since += 1
            if since >= patience:
                break
    net.close()
    sc = screen.Screen(best, xm.astype(np.float32), xs, ym, ys, n_hidden=n_hidden, n_layers=n_layers,
                       meta=dict(epochs_run=len(losses), val_loss=best_loss, train_rows=int(tr.sum()),
                                 val_rows=int(val.sum()), batch=batch, lr=lr))
    return sc, np.array(losses)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--asteroids", type=int, default=2048)
    ap.add_argument("--chunk", type=int, default=256, help="asteroids per prepared batch")
    ap.add_argument("--random", type=int, default=64, help="random trial frequencies per asteroid")
    ap.add_argument("--near", type=int, default=3, help="frequencies either side of the true one")
    ap.add_argument("--iters", type=int, default=0,
                    help="LM iterations of the target fits (0: the input file's own stop condition, the plain fit)")
    ap.add_argument("--hidden", type=int, default=128)
    ap.add_argument("--layers", type=int, default=2)
    ap.add_argument("--epochs", type=int, default=300)
    ap.add_argument("--batch", type=int, default=512)
    ap.add_argument("--lr", type=float, default=3e-4)
    ap.add_argument("--keep", type=float, default=None, help="keep fraction stored with the network (default screen.KEEP)")
    ap.add_argument("--workers", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--work-dir", default="gpurun_out/screen")
    ap.add_argument("--out", default="gpurun_out/screen/screen.npz")
    args = ap.parse_args(argv)
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import ps_claude_opt_amd as ps
    from ps_claude_opt_amd import screen
    from ps_claude_opt_amd.warm_train import generate

    t0 = time.time()
    gen = generate(os.path.join(args.work_dir, "train"), 50000, args.asteroids, args.workers, "c2")
    print(f"# generated {len(gen)} asteroids in {time.time() - t0:.1f} s", flush=True)
    X, Y, G, NEAR, full, best_n = [], [], [], [], {}, {}
    with ps.Scan(0) as s:
        t0 = time.time()
        for c0 in range(0, len(gen), args.chunk):
            part = gen[c0:c0 + args.chunk]
            s.prepare_files([p for p, _ in part])
            x, y, g, nr, ff, fq = build_set(s, [t for _, t in part], args.random, args.near, args.iters, seed=c0,
                                        keep_full=lambda i, c0=c0: is_validation(c0 + i))
            X.append(x)
            Y.append(y)
            G.append(g + c0)
            NEAR.append(nr)
            full.update({c0 + i: f for i, f in ff.items()})
            for i in ff:
                sel = (g == i) & nr
                best_n[c0 + i] = int(fq[sel][np.argmin(y[sel])])
            print(f"# asteroids {c0}..{c0 + len(part) - 1}: {len(y)} rows, {time.time() - t0:.1f} s", flush=True)
        X, Y, G, NEAR = np.concatenate(X), np.concatenate(Y), np.concatenate(G), np.concatenate(NEAR)
        t1 = time.time()
        sc, losses = train(s, X, Y, G, args.hidden, args.layers, args.epochs, args.batch, args.lr)
        print(f"# trained in {time.time() - t1:.1f} s: {len(losses)} epochs, loss {losses[0]:.4f} -> "
              f"{losses[-1]:.4f}, held-out {sc.meta['val_loss']:.4f}", flush=True)
        # ranks on the held-out training asteroids: network and single features
        net = sc.network(s)
        ranks_net, ranks_feat = [], {}
        for a in sorted(full):
            F = full[a]
            ranks_net.append(near_rank(sc.predict(net, F), best_n[a]))
            for j in range(F.shape[1]):
                if np.ptp(F[:, j]) > 0:
                    ranks_feat.setdefault(j, []).append(near_rank(-F[:, j], best_n[a]))
        net.close()
    rn = np.array(ranks_net)
    best_feat = min(ranks_feat, key=lambda j: np.percentile(ranks_feat[j], 95))
    rf = np.array(ranks_feat[best_feat])
    n_grid = int(np.median([f.shape[0] for f in full.values()]))
    report = dict(asteroids=args.asteroids, random=args.random, near=args.near, iters=args.iters,
                  hidden=args.hidden, layers=args.layers, rows=int(len(Y)), epochs_run=len(losses),
                  loss_first=float(losses[0]), loss_last=float(losses[-1]), val_loss=sc.meta["val_loss"],
                  seconds=round(time.time() - t0, 1), validation_asteroids=len(rn), grid=n_grid,
                  network_rank=dict(median=float(np.median(rn)), p95=float(np.percentile(rn, 95)),
                                    max=int(rn.max()), within_5pct=float(np.mean(rn < 0.05 * n_grid)),
                                    within_10pct=float(np.mean(rn < 0.10 * n_grid))),
                  best_single_feature=dict(index=int(best_feat), median=float(np.median(rf)),
                                           p95=float(np.percentile(rf, 95)), max=int(rf.max()),
                                           within_5pct=float(np.mean(rf < 0.05 * n_grid)),
                                           within_10pct=float(np.mean(rf < 0.10 * n_grid))))
    sc.keep = args.keep if args.keep is not None else screen.KEEP
    sc.meta.update(trained_on=f"{args.asteroids} c2-style asteroids, config_seed(4, 50000 + i)",
                   iters=args.iters or "plain fits", aliases=list(ALIASES), network_rank=report["network_rank"])
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    sc.save(args.out)
    with open(os.path.splitext(args.out)[0] + "_report.json", "w") as f:
        json.dump(report, f, indent=1)
    print(json.dumps(report), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
