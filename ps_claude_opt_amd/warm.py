"""Learned warm start of the period scan.

A small MLP (3 hidden layers of 256 GELU units, 52 outputs) predicts, for
each (trial frequency, initial pole) pair, how far that pair's LM fit moves
the shape coefficients (parameters 2..C) and phase-function / Lambert
parameters (A, D, k, log cl) from the reference's initialisation (the
ellipsoid of ellfit.c and the input file's phase function, Start.cu:
4346-4418), from the lightcurves folded at the trial period, the pole and
the viewing geometry (features: ps_warm_features in include/ps_mi355x.h).
The fit then starts at initialisation + prediction; the LM iterations
themselves stay FP64 and unchanged.  Only free parameters move.

The network is trained on the GPU by the library's own kernels (BF16 operands,
FP32 accumulation, Adam) on converged fits that the plain scan produces on
seeded synthetic asteroids, and evaluated in FP32.  The plain cold-start scan
stays the judge: ``warm_scan`` refits the best frequencies of a warm scan
cold, so every reported winner is a cold fit.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import json
import os
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import _abi

N_HIDDEN = 256
N_LAYERS = 3
# Weights trained on seeded lmax-6 asteroids (288 facets, 54 free parameters).
DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "warm_lmax6.npz")


def weights_for(lmax: int) -> str:
    """The shipped weights for shapes of degree lmax (the file may be absent)."""
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", f"warm_lmax{lmax}.npz")


class Mlp:
    """The network kernels through the C ABI (ps_mlp_*)."""

    def __init__(self, scan, n_in: int, n_out: int, n_hidden: int = N_HIDDEN, n_layers: int = N_LAYERS,
                 seed: int = 0):
        self._lib = scan._lib
        self._scan = scan
        h = C.c_void_p()
        scan._check(self._lib.ps_mlp_create(scan._h, n_in, n_hidden, n_layers, n_out, seed, C.byref(h)))
        self._h = h
        self.n_in, self.n_out, self.n_hidden, self.n_layers = n_in, n_out, n_hidden, n_layers

    def close(self):
        if self._h:
            self._lib.ps_mlp_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def n_params(self) -> int:
        return int(self._lib.ps_mlp_n_params(self._h))

    def get_params(self) -> np.ndarray:
        p = np.zeros(self.n_params, np.float32)
        self._scan._check(self._lib.ps_mlp_get_params(self._h, _abi.fptr(p)))
        return p

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.n_params
        self._scan._check(self._lib.ps_mlp_set_params(self._h, _abi.fptr(p)))

    def forward(self, x: np.ndarray, bf16: bool = False) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros((x.shape[0], self.n_out), np.float32)
        self._scan._check(self._lib.ps_mlp_forward(self._h, x.shape[0], _abi.fptr(x), _abi.fptr(y), int(bf16)))
        return y

    def loss_grad(self, x: np.ndarray, t: np.ndarray):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        g = np.zeros(self.n_params, np.float32)
        loss = C.c_float(0.0)
        self._scan._check(self._lib.ps_mlp_loss_grad(self._h, x.shape[0], _abi.fptr(x), _abi.fptr(t), C.byref(loss),
                                                     _abi.fptr(g)))
        return float(loss.value), g

    def train(self, x: np.ndarray, t: np.ndarray, epochs: int, batch: int = 512, lr: float = 1e-3,
              seed: int = 0) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        losses = np.zeros(max(epochs, 1), np.float32)
        self._scan._check(self._lib.ps_mlp_train(self._h, x.shape[0], _abi.fptr(x), _abi.fptr(t), epochs, batch,
                                                 lr, seed, _abi.fptr(losses)))
        return losses[:epochs]


def targets_from_params(a: np.ndarray, ncoef: int) -> np.ndarray:
    """The predicted parameters of parameter vectors: 2..C, then A, D, k,
    log cl (C-1 + 4 values; 52 at lmax = 6)."""
    a = np.atleast_2d(a)
    return np.concatenate([a[:, 1:ncoef], a[:, ncoef + 3:ncoef + 7]], axis=1)


@dataclasses.dataclass
class WarmStart:
    """A trained warm-start network with its standardisation."""
    params: np.ndarray
    x_mean: np.ndarray
    x_std: np.ndarray
    y_mean: np.ndarray
    y_std: np.ndarray
    y_lo: Optional[np.ndarray] = None  # predictions are clipped to the training targets' range
    y_hi: Optional[np.ndarray] = None
    shrink: Optional[np.ndarray] = None  # per output: prediction = mean + shrink * (network - mean)
    n_hidden: int = N_HIDDEN
    n_layers: int = N_LAYERS
    meta: Optional[dict] = None

    @property
    def n_in(self) -> int:
        return int(self.x_mean.size)

    @property
    def n_out(self) -> int:
        return int(self.y_mean.size)

    def save(self, path: str):
        np.savez(path, params=self.params, x_mean=self.x_mean, x_std=self.x_std, y_mean=self.y_mean,
                 y_std=self.y_std, y_lo=self.y_lo, y_hi=self.y_hi, shrink=self.shrink,
                 shape=np.array([self.n_hidden, self.n_layers]),
                 meta=np.array(json.dumps(self.meta or {})))

    @classmethod
    def load(cls, path: str) -> "WarmStart":
        d = np.load(path, allow_pickle=False)
        return cls(params=d["params"], x_mean=d["x_mean"], x_std=d["x_std"], y_mean=d["y_mean"], y_std=d["y_std"],
                   y_lo=d["y_lo"], y_hi=d["y_hi"], shrink=d["shrink"], n_hidden=int(d["shape"][0]),
                   n_layers=int(d["shape"][1]),
                   meta=json.loads(str(d["meta"])))

    def network(self, scan) -> Mlp:
        net = Mlp(scan, self.n_in, self.n_out, self.n_hidden, self.n_layers)
        net.set_params(self.params)
        return net

    def predict(self, net: Mlp, feats: np.ndarray) -> np.ndarray:
        """Predicted parameter moves (raw units) of feature rows, FP32."""
        y = net.forward((feats - self.x_mean) / self.x_std) * self.y_std + self.y_mean
        if self.shrink is not None:
            y = self.y_mean + self.shrink * (y - self.y_mean)
        if self.y_lo is not None:
            y = np.clip(y, self.y_lo, self.y_hi)
        return y

    def initial_params(self, scan, net: Mlp, prob, n_list, m_list) -> np.ndarray:
        """Warm initial parameter vectors [items][ma]: the reference's
        initialisation with the free shape and phase-function parameters
        moved by the network's FP32 prediction."""
        feats = scan.warm_features(prob, n_list, m_list)
        y = self.predict(net, feats)
        a0 = scan.initial_params(prob, n_list, m_list)
        Cn = scan.info["ncoef"]
        ia = scan.free_params()
        for j in range(Cn - 1):
            if ia[1 + j]:
                a0[:, 1 + j] += y[:, j]
        for j in range(4):
            if ia[Cn + 3 + j]:
                a0[:, Cn + 3 + j] += y[:, Cn - 1 + j]
        return a0


def fit_standardisation(x: np.ndarray, y: np.ndarray):
    xm, xs = x.mean(0), x.std(0)
    ym, ys = y.mean(0), y.std(0)
    xs = np.where(xs > 1e-6, xs, 1.0).astype(np.float32)
    ys = np.where(ys > 1e-9, ys, 1.0).astype(np.float32)
    return xm.astype(np.float32), xs, ym.astype(np.float32), ys


def make_training_set(scan, n_pairs_per_problem: int, conw_r: Sequence[float], seed: int = 0,
                      near_n: Optional[Sequence[int]] = None, near_frac: float = 0.0, near_width: int = 2):
    """Cold fits of seeded random (frequency, pole) pairs of every problem of
    the scan's batch; returns (features, targets = final - initial
    parameters, results, problem index) of the fits that converged to finite
    parameters.  With near_n (each problem's true frequency index), a
    fraction near_frac of a problem's pairs is drawn within near_width
    frequency steps of it: the pairs whose converged shape is the
    asteroid's, the ones a truncated scan must rank first."""
    rng = np.random.default_rng(seed)
    prob, nl, ml = [], [], []
    for i, b in enumerate(scan.batch_info):
        n = rng.integers(1, b["n_max"] + 1, n_pairs_per_problem)
        if near_n is not None and near_frac > 0:
            k = int(round(near_frac * n_pairs_per_problem))
            n[:k] = np.clip(near_n[i] + rng.integers(-near_width, near_width + 1, k), 1, b["n_max"])
        prob += [i] * n_pairs_per_problem
        nl += list(n)
        ml += list(rng.integers(1, 11, n_pairs_per_problem))
    prob, nl, ml = np.array(prob), np.array(nl), np.array(ml)
    cw = np.asarray(conw_r, np.float64)[prob]
    res, a_final = scan.fit_items_from(prob, nl, ml, cw, cw, None, want_final=True)
    feats = scan.warm_features(prob, nl, ml)
    a_init = scan.initial_params(prob, nl, ml)
    Cn = scan.info["ncoef"]
    # runaway fits overflow the FP32 cast to inf; the finiteness test below drops them
    with np.errstate(over="ignore"):
        tg = (targets_from_params(a_final, Cn) - targets_from_params(a_init, Cn)).astype(np.float32)
    ok = np.isfinite(tg).all(1) & np.isfinite(res["dev"]) & (res["dev"] < 1e30)
    # fits that ran away (phase-function parameters in the 1e20s and beyond
    # with a finite rms) are no examples of where a fit goes: keep rows whose
    # every move lies within 10 interquartile ranges of the quartiles
    q1, q3 = np.percentile(tg[ok], 25, axis=0), np.percentile(tg[ok], 75, axis=0)
    lo, hi = q1 - 10 * (q3 - q1) - 1e-6, q3 + 10 * (q3 - q1) + 1e-6
    ok &= ((tg >= lo) & (tg <= hi)).all(1)
    return feats[ok], tg[ok], res[ok], prob[ok]


def train_warm_start(scan, feats: np.ndarray, targets: np.ndarray, groups: Optional[np.ndarray] = None,
                     epochs: int = 60, batch: int = 512, lr: float = 1e-3, seed: int = 0, val_frac: float = 0.2,
                     patience: int = 8, meta: Optional[dict] = None):
    """Trains the network on the GPU (BF16, Adam).  The rows of a held-out
    fraction of the groups (asteroids) pick the epoch (early stopping on the
    FP32 validation loss) and a per-output shrinkage factor towards the mean
    move, fitted in closed form on the same rows.  Returns (WarmStart,
    per-epoch training losses)."""
    rng = np.random.default_rng(seed)
    if groups is None:
        groups = np.arange(len(feats))
    ug = np.unique(groups)
    val_g = rng.choice(ug, max(1, int(round(val_frac * len(ug)))), replace=False)
    val = np.isin(groups, val_g)
    tr = ~val
    xm, xs, ym, ys = fit_standardisation(feats[tr], targets[tr])
    xt, yt = (feats[tr] - xm) / xs, (targets[tr] - ym) / ys
    xv, yv = (feats[val] - xm) / xs, (targets[val] - ym) / ys
    net = Mlp(scan, feats.shape[1], targets.shape[1], seed=seed)
    best, best_loss, since = net.get_params(), np.inf, 0
    losses = []
    for e in range(epochs):
        losses.append(float(net.train(xt, yt, 1, batch, lr, seed + 1 + e)[0]))
        vl = float(np.mean((net.forward(xv) - yv) ** 2))
        if vl < best_loss:
            best, best_loss, since = net.get_params(), vl, 0
        else:
            since += 1
            if since >= patience:
                break
    net.set_params(best)
    pv = net.forward(xv)  # standardised predictions on the validation rows
    num = (pv * yv).sum(0)
    den = (pv * pv).sum(0)
    shrink = np.clip(np.where(den > 0, num / np.maximum(den, 1e-30), 0.0), 0.0, 1.0).astype(np.float32)
    lo, hi = np.percentile(targets[tr], 0.5, axis=0), np.percentile(targets[tr], 99.5, axis=0)
    ws = WarmStart(best, xm, xs, ym, ys, lo.astype(np.float32), hi.astype(np.float32), shrink,
                   meta=dict(meta or {}, epochs_run=len(losses), batch=batch, lr=lr, val_loss=best_loss,
                             train_rows=int(tr.sum()), val_rows=int(val.sum())))
    net.close()
    return ws, np.array(losses)


def best_lines(res: np.ndarray, n_freq: int):
    """Index of the best pole per frequency (start_CUDA.cu:1200-1210) for
    results ordered frequency-major, 10 poles each."""
    dev = res["dev"].reshape(n_freq, 10)
    best = np.zeros(n_freq, np.int64)
    for f in range(n_freq):
        b = 0
        for p in range(1, 10):
            if not np.isnan(dev[f, p]) and not (dev[f, p] >= dev[f, b]):
                b = p
        best[f] = b
    return best


REFIT_MIN = 10      # frequencies always refitted cold
REFIT_BAND = 0.15   # and every frequency whose truncated rms is within 15 % of the best one


def refit_set(rms: np.ndarray, refit: int = REFIT_MIN, band: float = REFIT_BAND) -> np.ndarray:
    """Frequencies (0-based, sorted) a warm scan refits cold: the `refit`
    lowest truncated rms, and every frequency whose truncated rms is at most
    (1 + band) times the lowest.  The band makes the set adaptive: where
    the truncated screen separates its best frequencies clearly, few are
    refitted; where many lie close to the best, all of those are."""
    r = np.where(np.isnan(rms), np.inf, rms)
    order = np.argsort(r, kind="stable")
    k = min(refit, r.size)
    if r.size and np.isfinite(r[order[0]]):
        k = max(k, int(np.searchsorted(r[order], r[order[0]] * (1.0 + band), side="right")))
    return np.sort(order[:k])


# Period aliases of a lightcurve fit: a convex model fits a rotational
# lightcurve about as well at these multiples of the period, and which of
# them a plain scan ranks first is decided by small rms differences.
ALIAS_MULTIPLES = (2.0, 0.5, 1.5, 2.0 / 3.0, 3.0, 1.0 / 3.0, 2.5, 0.4)
ALIAS_SEEDS = 3   # the best truncated frequencies whose aliases are refitted too
ALIAS_WIDTH = 3   # frequency steps either side of each alias


def alias_windows(info: dict, n_list, width: int = ALIAS_WIDTH, mults=ALIAS_MULTIPLES) -> np.ndarray:
    """The 1-based trial frequencies within `width` steps of each alias
    (period times mults) of the frequencies n_list, on the problem's grid
    f_start - (n - 1) f_step (1/h); sorted, unique, inside 1 .. n_max."""
    out = set()
    for n in np.atleast_1d(n_list):
        period_h = 1.0 / (info["f_start"] - (int(n) - 1) * info["f_step"])
        for m in mults:
            k = int(round((info["f_start"] - 1.0 / (m * period_h)) / info["f_step"])) + 1
            for j in range(k - width, k + width + 1):
                if 1 <= j <= info["n_max"]:
                    out.add(j)
    return np.array(sorted(out), np.int64)


def refit_frequencies(info: dict, rms: np.ndarray, n_start: int, refit: int = REFIT_MIN, band: float = REFIT_BAND,
                      alias_seeds: int = ALIAS_SEEDS, alias_width: int = ALIAS_WIDTH, among=None) -> np.ndarray:
    """Frequencies (n_start-based, sorted) to refit cold after a truncated
    screen whose best truncated rms per frequency is `rms` (n_start ..;
    NaN where a frequency was not fitted): refit_set's frequencies, plus
    the alias windows of the `alias_seeds` best ones, within the scanned
    range."""
    r = np.asarray(rms, np.float64)
    if among is not None:
        sub = np.asarray(among, np.int64) - n_start
        base = sub[refit_set(r[sub], refit, band)]
    else:
        base = refit_set(r, refit, band)
    rr = np.where(np.isnan(r), np.inf, r)
    order = np.argsort(rr, kind="stable")
    seeds = [n_start + int(j) for j in order[:alias_seeds] if np.isfinite(rr[j])]
    al = alias_windows(info, seeds, alias_width) if seeds and alias_seeds > 0 else np.zeros(0, np.int64)
    al = al[(al >= n_start) & (al < n_start + r.size)] - n_start
    return n_start + np.union1d(base, al).astype(np.int64)


def warm_scan(scan, ws: Optional[WarmStart], net: Optional[Mlp], conw_r: float, n_start: int, n_count: int,
              n_iter: int, refit: int = REFIT_MIN, band: float = REFIT_BAND, alias_seeds: int = ALIAS_SEEDS):
    """Warm-started scan of problem 0: every pair runs n_iter LM iterations
    from the network's prediction (ws None: from the reference's start);
    then the frequencies of refit_frequencies (the `refit` lowest truncated
    rms, all within `band` of the lowest, and the alias windows of the
    `alias_seeds` lowest) are fitted again cold, with the problem's own stop
    condition (all 10 poles), and those cold results replace the truncated
    ones.  Returns (results [n_count*10]
    frequency-major, refitted frequency indices).  The scan's winner is the
    best of the refitted frequencies (best_line(..., among=refitted)):
    those are plain fits, and a warm-started truncated fit may end below a
    converged cold one."""
    n = np.repeat(np.arange(n_start, n_start + n_count), 10)
    m = np.tile(np.arange(1, 11), n_count)
    p = np.zeros(n.size, np.int32)
    a0 = ws.initial_params(scan, net, p, n, m) if ws is not None else None
    res, _ = scan.fit_items_from(p, n, m, conw_r, conw_r, a0, n_iter=n_iter)
    best = best_lines(res, n_count)
    rms = res["dev"].reshape(n_count, 10)[np.arange(n_count), best]
    fr = refit_frequencies(scan.info, rms, n_start, refit, band, alias_seeds) - n_start
    if fr.size:
        n2 = np.repeat(n_start + fr, 10)
        m2 = np.tile(np.arange(1, 11), fr.size)
        cold, _ = scan.fit_items_from(np.zeros(n2.size, np.int32), n2, m2, conw_r, conw_r, None)
        for k, f in enumerate(fr):
            res[f * 10:(f + 1) * 10] = cold[k * 10:(k + 1) * 10]
    return res, n_start + fr


def best_line(res: np.ndarray, n_start: int, n_count: int, among=None) -> dict:
    """The overall best (lowest rms) line of a scan's results; `among`
    (frequency indices, n_start-based) restricts the choice to those
    frequencies, e.g. the cold refits of a warm scan."""
    best = best_lines(res, n_count)
    rows = np.arange(n_count) * 10 + best
    rms = res["dev"][rows]
    rms = np.where(np.isnan(rms), np.inf, rms)
    if among is not None:
        keep = np.zeros(n_count, bool)
        keep[np.asarray(among, np.int64) - n_start] = True
        rms = np.where(keep, rms, np.inf)
    f = int(np.argmin(rms))
    r = res[rows[f]]
    return dict(freq_index=n_start + f, pole=int(best[f]) + 1, period_h=24.0 * float(r["period"]),
                rms=float(r["dev"]), lambda_deg=float(r["lambda_deg"]), beta_deg=float(r["beta_deg"]),
                chisq=float(r["chisq"]))
