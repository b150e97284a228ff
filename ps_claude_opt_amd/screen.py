"""Learned frequency screen of the warm-started period scan.

The plain scan fits all 10 poles at every trial frequency.  The screened
scan first scores every frequency with a small MLP and fits only the
best-scored fraction: there, every pole runs a truncated LM fit, and the
frequencies whose truncated rms lies near the best one are refitted cold
to the input file's own stop condition (warm.refit_set).  The reported
winner is the best of those refits, so it is a plain fit, bit for bit.

The network's input is pole-independent: the lightcurves folded at the
trial frequency, as harmonic powers (ps_screen_features in
include/ps_mi355x.h).  Its target is the log of the best rms of the 10
poles at that frequency, relative to the asteroid's median over random
frequencies: it is trained, by the library's own MLP kernels, on the cold
fits that the plain scan itself produces on seeded synthetic asteroids, at
random frequencies and around the true period and its aliases
(screen_train.py).  The plain scan has no counterpart of this
step (start_CUDA.cu:792-1245 fits every frequency); the screen only decides
which frequencies are fitted, never what a fit returns.

The shipped network was trained on config-2-style scans: 5,000 trial
frequencies over 2-40 h, whose step shifts the phase by about 2.5 turns over
the 3-year span of the data.  On a narrow grid of fine steps centred on the
true frequency (config 1: 200 steps of half a turn) every frequency folds
coherently and the score does not separate them; the screen is not meant
for such scans.
"""
from __future__ import annotations

import dataclasses
import json
import os
from typing import Optional

import numpy as np

from . import warm

N_HIDDEN = 128
N_LAYERS = 2
KEEP = 0.50      # fraction of the scan's frequencies the screened scan fits
KEEP_MIN = 50    # and at least this many
DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "screen.npz")


@dataclasses.dataclass
class Screen:
    """A trained screen network with its input standardisation."""
    params: np.ndarray
    x_mean: np.ndarray
    x_std: np.ndarray
    y_mean: float
    y_std: float
    keep: float = KEEP
    keep_min: int = KEEP_MIN
    n_hidden: int = N_HIDDEN
    n_layers: int = N_LAYERS
    meta: Optional[dict] = None

    @property
    def n_in(self) -> int:
        return int(self.x_mean.size)

    def save(self, path: str):
        np.savez(path, params=self.params, x_mean=self.x_mean, x_std=self.x_std,
                 y=np.array([self.y_mean, self.y_std]), keep=np.array([self.keep, self.keep_min]),
                 shape=np.array([self.n_hidden, self.n_layers]), meta=np.array(json.dumps(self.meta or {})))

    @classmethod
    def load(cls, path: str = DEFAULT_WEIGHTS) -> "Screen":
        d = np.load(path, allow_pickle=False)
        return cls(params=d["params"], x_mean=d["x_mean"], x_std=d["x_std"], y_mean=float(d["y"][0]),
                   y_std=float(d["y"][1]), keep=float(d["keep"][0]), keep_min=int(d["keep"][1]),
                   n_hidden=int(d["shape"][0]), n_layers=int(d["shape"][1]), meta=json.loads(str(d["meta"])))

    def network(self, scan) -> warm.Mlp:
        net = warm.Mlp(scan, self.n_in, 1, self.n_hidden, self.n_layers)
        net.set_params(self.params)
        return net

    def predict(self, net: warm.Mlp, feats: np.ndarray) -> np.ndarray:
        """Predicted relative log truncated rms per feature row (lower is
        better), FP32 forward pass."""
        y = net.forward((feats - self.x_mean) / self.x_std)[:, 0]
        return y * self.y_std + self.y_mean

    def scores(self, scan, net: warm.Mlp, n_start: int, n_count: int, prob: int = 0) -> np.ndarray:
        """Scores of frequencies n_start .. n_start + n_count - 1 (the
        relative inputs are taken over exactly this range)."""
        return self.predict(net, network_inputs(scan.screen_features(n_start, n_count, prob)))

    def n_keep(self, n_count: int) -> int:
        return int(min(n_count, max(self.keep_min, int(np.ceil(self.keep * n_count)))))


N_POWER = 24  # leading fold features that are harmonic powers (ps_screen_features row layout)


def network_inputs(feats: np.ndarray) -> np.ndarray:
    """The network's input rows for the frequencies of one scan: the fold
    features, then each power feature relative to the scan's own grid,
    (x - median) / (1.4826 MAD + 1e-3).  A rotational lightcurve shows as
    powers far above the asteroid's typical frequency, whatever the
    amplitude of its curves."""
    f = np.asarray(feats, np.float32)
    pw = f[:, :N_POWER]
    med = np.median(pw, axis=0)
    mad = np.median(np.abs(pw - med), axis=0)
    rel = (pw - med) / (1.4826 * mad + 1e-3)
    return np.concatenate([f, rel.astype(np.float32)], axis=1)


def keep_set(scores: np.ndarray, n_keep: int) -> np.ndarray:
    """The n_keep best-scored frequencies (0-based, ascending); NaN scores
    rank last, ties by frequency index."""
    s = np.where(np.isnan(scores), np.inf, scores)
    order = np.argsort(s, kind="stable")
    return np.sort(order[:n_keep])


def screened_scan(scan, sc: Screen, net: warm.Mlp, conw_r: float, n_start: int, n_count: int, n_iter: int,
                  refit: int = warm.REFIT_MIN, band: float = warm.REFIT_BAND, ws=None, ws_net=None,
                  alias_seeds: int = warm.ALIAS_SEEDS):
    """Screened scan of problem 0.  Keeps sc.n_keep(n_count) frequencies by
    the network's score, runs every pole of those for n_iter LM iterations
    (from the reference's start, or from the warm-start network ws if
    given), then refits cold, with the problem's own stop condition,
    warm.refit_frequencies of the kept ones: refit_set's frequencies and
    the alias windows of the `alias_seeds` best truncated ones (which may
    lie outside the kept set: a plain scan's winner is often an alias of
    the frequency the screen ranks first).  Returns (results [n_count*10]
    frequency-major, NaN dev where a frequency was not fitted; refitted
    frequency indices; kept frequency indices), indices n_start-based.  The
    winner is warm.best_line(res, ..., among=refitted)."""
    from . import RESULT_DTYPE

    scores = sc.scores(scan, net, n_start, n_count)
    kept = keep_set(scores, sc.n_keep(n_count))
    res = np.zeros(n_count * 10, dtype=RESULT_DTYPE)
    res["dev"] = np.nan
    n = np.repeat(n_start + kept, 10)
    m = np.tile(np.arange(1, 11), kept.size)
    p = np.zeros(n.size, np.int32)
    a0 = ws.initial_params(scan, ws_net, p, n, m) if ws is not None else None
    part, _ = scan.fit_items_from(p, n, m, conw_r, conw_r, a0, n_iter=n_iter)
    rows = (np.repeat(kept, 10) * 10 + np.tile(np.arange(10), kept.size))
    res[rows] = part
    best = warm.best_lines(part, kept.size)
    rms = np.full(n_count, np.nan)
    rms[kept] = part["dev"].reshape(kept.size, 10)[np.arange(kept.size), best]
    fr = warm.refit_frequencies(scan.info, rms, n_start, refit, band, alias_seeds, among=n_start + kept) - n_start
    if fr.size:
        n2 = np.repeat(n_start + fr, 10)
        m2 = np.tile(np.arange(1, 11), fr.size)
        cold, _ = scan.fit_items_from(np.zeros(n2.size, np.int32), n2, m2, conw_r, conw_r, None)
        for k, f in enumerate(fr):
            res[f * 10:(f + 1) * 10] = cold[k * 10:(k + 1) * 10]
    return res, n_start + fr, n_start + kept


def screened_scan_batch(scan, sc: Screen, net: warm.Mlp, conw_r, n_iter: int, refit: int = warm.REFIT_MIN,
                        band: float = warm.REFIT_BAND, alias_seeds: int = warm.ALIAS_SEEDS):
    """screened_scan over every problem of a prepared batch (the whole grid
    of each), with the truncated fits of all problems in one persistent
    launch and their cold refits in a second one.  conw_r: one value or one
    per problem.  Returns, per problem, (results [n_max*10], refitted
    1-based frequencies, kept 1-based frequencies)."""
    from . import RESULT_DTYPE

    infos = scan.batch_info
    cw = np.broadcast_to(np.asarray(conw_r, np.float64), (len(infos),))
    kept = []
    for i, b in enumerate(infos):
        scores = sc.scores(scan, net, 1, b["n_max"], prob=i)
        kept.append(keep_set(scores, sc.n_keep(b["n_max"])))
    p = np.concatenate([np.full(k.size * 10, i, np.int32) for i, k in enumerate(kept)])
    n = np.concatenate([np.repeat(k + 1, 10) for k in kept])
    m = np.concatenate([np.tile(np.arange(1, 11), k.size) for k in kept])
    part, _ = scan.fit_items_from(p, n, m, cw[p], cw[p], None, n_iter=n_iter)
    out, refits, off = [], [], 0
    for i, (b, k) in enumerate(zip(infos, kept)):
        res = np.zeros(b["n_max"] * 10, dtype=RESULT_DTYPE)
        res["dev"] = np.nan
        pk = part[off:off + k.size * 10]
        off += k.size * 10
        res[(np.repeat(k, 10) * 10 + np.tile(np.arange(10), k.size))] = pk
        best = warm.best_lines(pk, k.size)
        rms = np.full(b["n_max"], np.nan)
        rms[k] = pk["dev"].reshape(k.size, 10)[np.arange(k.size), best]
        refits.append(warm.refit_frequencies(b, rms, 1, refit, band, alias_seeds, among=k + 1))
        out.append(res)
    p2 = np.concatenate([np.full(f.size * 10, i, np.int32) for i, f in enumerate(refits)])
    n2 = np.concatenate([np.repeat(f, 10) for f in refits])
    m2 = np.concatenate([np.tile(np.arange(1, 11), f.size) for f in refits])
    if p2.size:
        cold, _ = scan.fit_items_from(p2, n2, m2, cw[p2], cw[p2], None)
        off = 0
        for i, f in enumerate(refits):
            for j, fr in enumerate(f):
                out[i][(fr - 1) * 10:fr * 10] = cold[off + j * 10:off + (j + 1) * 10]
            off += f.size * 10
    return [(out[i], refits[i], kept[i] + 1) for i in range(len(infos))]


def true_index(info: dict, period_h: float) -> int:
    """1-based index of the trial frequency nearest to 1 / period_h on the
    problem's grid f_start - (n - 1) f_step (1/h)."""
    n = int(round((info["f_start"] - 1.0 / period_h) / info["f_step"])) + 1
    return min(max(n, 1), info["n_max"])


def rank_of(scores: np.ndarray, index0: int) -> int:
    """Rank (0 = best) of frequency index0 (0-based) under the scores."""
    s = np.where(np.isnan(scores), np.inf, scores)
    order = np.argsort(s, kind="stable")
    return int(np.where(order == index0)[0][0])
