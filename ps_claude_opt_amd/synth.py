"""Seeded synthetic asteroids written as ``period_search_in`` files.

The generator draws a convex shape (Gaussian log-curvature series over the
unnormalised Laplace basis of the reference's ``sphfunc.c``), a pole, a
rotation period and viewing geometries, and renders relative lightcurves
with Lommel-Seeliger + Lambert scattering and the exponential-linear phase
function of the reference's model (Start.cu:2095-2195, 2333-2403).  The
rendering is an independent numpy forward model on a fine Fibonacci
sphere, not the fitting code, so a scan has to rediscover the truth.

Input-file grammar: period_search_BOINC.cpp:343-533.  Lightcurves are
written as relative (flag 0), the only kind the reference's GPU path
computes, except the last ``n_abs`` curves, which are written as calibrated
(flag 1, absolute): their brightness is the model flux of the truth, whose
size-scale coefficient is 0.

The truth shape's Gaussian image is closed (a convex body exists) and its
rotation is referenced to the epoch the parser derives, so the truth is a
parameter vector of the fitted model itself (``truth_params``).
"""
from __future__ import annotations

import dataclasses
import json
import math
from typing import Dict, List, Optional

import numpy as np

SEED_BASE = 20261203


@dataclasses.dataclass
class SynthConfig:
    n_dense: int = 4
    pts_dense: int = 120
    dense_hours: float = 8.0
    n_sparse: int = 200
    sparse_years: float = 3.0
    sparse_group: int = 100          # points per relative sparse curve
    lmax: int = 6
    nrows: int = 6
    n_freq: int = 200                # trial frequencies in the scan
    p_min_h: Optional[float] = None  # fixed scan range (hours); None = centred on truth
    p_max_h: Optional[float] = None
    stop_condition: float = 50.0
    n_iter_min: int = 0
    alamda_incr: float = 10.0
    alamda_start: float = 0.001
    conw: float = 0.1
    noise: float = 0.01
    phase_min_deg: float = 5.0
    phase_max_deg: float = 25.0
    period_range_h: tuple = (4.0, 12.0)
    cl_true: float = 0.1
    par_true: tuple = (0.5, 0.1, -0.5)     # A, D, k of the phase function
    par_init: tuple = (0.5, 0.1, -0.5)
    par_free: tuple = (1, 1, 1)
    cl_init: float = 0.1
    cl_free: int = 0
    n_abs: int = 0                   # last n_abs curves written as calibrated (absolute)


# the BASELINE.json workloads (configs 1, 2, 5; 3 and 4 reuse the generator)
PRESETS: Dict[str, SynthConfig] = {
    "c1": SynthConfig(),
    "c2": SynthConfig(n_dense=10, pts_dense=150, n_sparse=1200, sparse_group=100,
                      n_freq=5000, p_min_h=2.0, p_max_h=40.0),
    "c5": SynthConfig(n_dense=4, pts_dense=200, n_sparse=20000, sparse_years=10.0,
                      sparse_group=250, lmax=8, nrows=8, n_freq=20000,
                      p_min_h=2.0, p_max_h=100.0),
    # config 5's observing geometry (4 dense curves, sparse points over 10
    # years, lmax 8) with 2,000 sparse points: the lmax-8 warm-start
    # training asteroids
    "c5s": SynthConfig(n_dense=4, pts_dense=200, n_sparse=2000, sparse_years=10.0,
                       sparse_group=250, lmax=8, nrows=8, n_freq=20000,
                       p_min_h=2.0, p_max_h=100.0),
    # small cases for tests
    "tiny": SynthConfig(n_dense=2, pts_dense=40, n_sparse=40, sparse_group=40,
                        lmax=3, nrows=3, n_freq=6),
    "small": SynthConfig(n_dense=3, pts_dense=60, n_sparse=60, sparse_group=30,
                         lmax=4, nrows=4, n_freq=12),
}


def _fib_sphere(n: int):
    i = np.arange(n) + 0.5
    z = 1.0 - 2.0 * i / n
    r = np.sqrt(1.0 - z * z)
    ph = i * math.pi * (3.0 - math.sqrt(5.0))
    return np.stack([r * np.cos(ph), r * np.sin(ph), z], axis=1)


def _assoc_legendre(lmax: int, x: np.ndarray):
    """P_l^m(x) without the Condon-Shortley phase (as sphfunc.c)."""
    s = np.sqrt(np.maximum(0.0, 1.0 - x * x))
    P = np.zeros((lmax + 1, lmax + 1, x.size))
    P[0, 0] = 1.0
    for m in range(1, lmax + 1):
        P[m, m] = P[m - 1, m - 1] * (2 * m - 1) * s
    for m in range(0, lmax):
        P[m + 1, m] = x * (2 * m + 1) * P[m, m]
    for m in range(0, lmax + 1):
        for l in range(m + 2, lmax + 1):
            P[l, m] = ((2 * l - 1) * x * P[l - 1, m] - (l + m - 1) * P[l - 2, m]) / (l - m)
    return P


def _laplace_basis(lmax: int, nrm: np.ndarray):
    """Columns in the reference ordering: m-major, cos then sin (sphfunc.c:65-86)."""
    th = np.arccos(np.clip(nrm[:, 2], -1, 1))
    ph = np.arctan2(nrm[:, 1], nrm[:, 0])
    P = _assoc_legendre(lmax, np.cos(th))
    cols, lvals = [], []
    for m in range(lmax + 1):
        for l in range(m, lmax + 1):
            cols.append(P[l, m] * np.cos(m * ph)); lvals.append(l)
            if m:
                cols.append(P[l, m] * np.sin(m * ph)); lvals.append(l)
    return np.stack(cols, axis=1), np.array(lvals)


def _rot_body(lam, beta, omega, phi0, t, vec):
    """Ecliptic -> body frame: Rz(phi) P(pi/2 - beta) Rz(lam), phi = omega t + phi0."""
    bp = math.pi / 2 - beta
    cl, sl, cb, sb = math.cos(lam), math.sin(lam), math.cos(bp), math.sin(bp)
    ux = cl * vec[:, 0] + sl * vec[:, 1]
    uy = -sl * vec[:, 0] + cl * vec[:, 1]
    uz = vec[:, 2]
    vx = cb * ux - sb * uz
    vy = uy
    vz = sb * ux + cb * uz
    phi = omega * t + phi0
    cf, sf = np.cos(phi), np.sin(phi)
    return np.stack([cf * vx + sf * vy, -sf * vx + cf * vy, vz], axis=1)


def _unit(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def _random_geometry(rng, n, phase_min, phase_max):
    """n (sun, earth) unit direction pairs near the ecliptic with a solar phase
    angle in [phase_min, phase_max] degrees."""
    lon = rng.uniform(0, 2 * math.pi, n)
    lat = np.radians(rng.uniform(-10, 10, n))
    sun = np.stack([np.cos(lat) * np.cos(lon), np.cos(lat) * np.sin(lon), np.sin(lat)], axis=1)
    # perpendicular direction, mostly in the ecliptic
    zax = np.array([0.0, 0.0, 1.0])
    perp = _unit(np.cross(sun, zax))
    tilt = np.radians(rng.uniform(-20, 20, n))
    perp = _unit(np.cos(tilt)[:, None] * perp + np.sin(tilt)[:, None] * np.cross(sun, perp))
    a = np.radians(rng.uniform(phase_min, phase_max, n))
    earth = np.cos(a)[:, None] * sun + np.sin(a)[:, None] * perp
    return sun, _unit(earth)


class Truth(dict):
    pass


def l1_columns(lmax: int):
    """Indices of the three l = 1 columns of the Laplace basis in the
    reference ordering (sphfunc.c:65-86): P_1^0 = z, P_1^1 cos(phi) = x,
    P_1^1 sin(phi) = y (no Condon-Shortley phase)."""
    return 1, lmax + 1, lmax + 2


CLOSURE_TOL = 1e-12


def close_gaussian_image(Y: np.ndarray, nrm: np.ndarray, coef: np.ndarray, lmax: int) -> np.ndarray:
    """Sets the l = 1 coefficients so that the facet areas exp(Y @ coef)
    satisfy sum_f A_f n_f = 0, the closure condition of a convex body's
    Gaussian image (Minkowski): then a convex polyhedron with these normals
    and areas exists.

    The l = 1 terms of the series are exp(g . n): the closure is the zero
    gradient of the strictly convex Phi(g) = sum_f exp(h_f + g . n_f), solved
    by Newton's method."""
    if lmax < 1:
        raise ValueError("closing the Gaussian image needs lmax >= 1")
    iz, ix, iy = l1_columns(lmax)
    c = coef.copy()
    c[[iz, ix, iy]] = 0.0
    h = Y @ c
    h -= h.max()
    # the basis's own l = 1 columns (the normal's components up to rounding),
    # so the solved coefficients close the areas exp(Y @ c) exactly as used
    Y1 = Y[:, [ix, iy, iz]]
    g = np.zeros(3)
    w = np.exp(h)
    for _ in range(100):
        grad = nrm.T @ w
        if np.linalg.norm(grad) <= 1e-15 * w.sum():
            break
        hess = (nrm * w[:, None]).T @ Y1
        step = np.linalg.solve(hess, grad)
        # damped Newton: halve the step until |sum A n| decreases
        gn = np.linalg.norm(grad)
        t = 1.0
        while True:
            w_new = np.exp(h + Y1 @ (g - t * step))
            if np.linalg.norm(nrm.T @ w_new) < gn or t < 1e-12:
                break
            t *= 0.5
        g = g - t * step
        w = w_new
    c[ix], c[iy], c[iz] = g
    area = np.exp(Y @ c)
    gap = np.linalg.norm(nrm.T @ area) / area.sum()
    if not gap < CLOSURE_TOL:
        raise AssertionError(f"Gaussian image not closed: |sum A n| / sum A = {gap:.3e}")
    return c


def make_asteroid(cfg: SynthConfig, seed: int):
    """Returns (text of period_search_in, truth dict)."""
    rng = np.random.default_rng(seed)
    lmax = cfg.lmax
    # shape: log-curvature coefficients, sigma 0.3/(l+1)
    nrm = _fib_sphere(4000)
    Y, lv = _laplace_basis(lmax, nrm)
    # sigma 0.3/(l+1) per unit-rms basis function (the unnormalised columns
    # span several decades in magnitude)
    rms = np.sqrt(np.mean(Y * Y, axis=0))
    shape_c = rng.normal(0.0, 1.0, Y.shape[1]) * (0.3 / (lv + 1.0)) / rms
    shape_c[0] = 0.0
    shape_c = close_gaussian_image(Y, nrm, shape_c, lmax)
    area = (4 * math.pi / nrm.shape[0]) * np.exp(Y @ shape_c)
    period_h = rng.uniform(*cfg.period_range_h)
    lam = rng.uniform(0, 2 * math.pi)
    beta = math.asin(rng.uniform(-1, 1))
    omega = 2 * math.pi / (period_h / 24.0)       # rad / day
    phi0 = 0.0
    A, D, k = cfg.par_true
    jd_start = 2455000.5 + rng.uniform(0, 100)
    span = cfg.sparse_years * 365.25
    curves: List[dict] = []
    # dense curves: fixed geometry over dense_hours
    for c in range(cfg.n_dense):
        t0 = jd_start + rng.uniform(0, span - 1)
        t = t0 + np.sort(rng.uniform(0, cfg.dense_hours / 24.0, cfg.pts_dense))
        sun, earth = _random_geometry(rng, 1, cfg.phase_min_deg, cfg.phase_max_deg)
        curves.append(dict(t=t, sun=np.repeat(sun, t.size, 0), earth=np.repeat(earth, t.size, 0)))
    # sparse points grouped into relative curves
    if cfg.n_sparse:
        ts = np.sort(jd_start + rng.uniform(0, span, cfg.n_sparse))
        sun, earth = _random_geometry(rng, cfg.n_sparse, cfg.phase_min_deg, cfg.phase_max_deg)
        for g0 in range(0, cfg.n_sparse, cfg.sparse_group):
            sl_ = slice(g0, min(cfg.n_sparse, g0 + cfg.sparse_group))
            curves.append(dict(t=ts[sl_], sun=sun[sl_], earth=earth[sl_]))
    # times as the file stores them (8 decimals); the rotation phase is
    # referenced to the epoch the parser derives from them, (int) of the
    # earliest time (period_search_BOINC.cpp:537-544), so the truth is a
    # parameter vector of the fitted model itself
    for cv in curves:
        cv["t"] = np.round(cv["t"], 8)
    jd0 = float(math.floor(min(float(cv["t"].min()) for cv in curves)))
    # render
    for cv in curves:
        t = cv["t"]
        g = _rot_body(lam, beta, omega, phi0, t - jd0, cv["earth"])
        g0 = _rot_body(lam, beta, omega, phi0, t - jd0, cv["sun"])
        mu = g @ nrm.T
        mu0 = g0 @ nrm.T
        vis = (mu > 0) & (mu0 > 0)
        mu = np.where(vis, mu, 0.0)
        mu0 = np.where(vis, mu0, 0.0)
        den = np.where(vis, mu + mu0, 1.0)
        s = np.where(vis, mu * mu0 * (cfg.cl_true + 1.0 / den), 0.0)
        br = s @ area
        alph = np.arccos(np.clip(np.sum(cv["earth"] * cv["sun"], axis=1), -1, 1))
        scale = 1.0 + A * np.exp(-alph / D) + k * alph
        flux = br * scale
        flux = flux * (1.0 + cfg.noise * rng.standard_normal(flux.size))
        cv["flux"] = flux
        # distances (AU): Sun 2-3 AU, Earth 1-2 AU
        cv["dsun"] = rng.uniform(2.0, 3.0, t.size)
        cv["dearth"] = rng.uniform(1.0, 2.0, t.size)
    jd_all = np.concatenate([cv["t"] for cv in curves])
    jd_min, jd_max = float(jd_all.min()), float(jd_all.max())
    f_true = 24.0 / period_h * (1 / 24.0)  # 1/hours
    f_step = 0.5 / (jd_max - jd_min) / 24.0
    if cfg.p_min_h is not None:
        fs, fe = 1.0 / cfg.p_min_h, 1.0 / cfg.p_max_h
        coef = (fs - fe) / ((cfg.n_freq - 0.5) * f_step)
        p_start, p_end = cfg.p_min_h, cfg.p_max_h
    else:
        coef = 1.0
        f_true = 1.0 / period_h
        fs = f_true + (cfg.n_freq / 2.0 - 0.3) * f_step
        fe = f_true - (cfg.n_freq / 2.0 - 0.7) * f_step
        p_start, p_end = 1.0 / fs, 1.0 / fe
    header = _header(cfg, p_start, coef, p_end)
    # verify the scan length the parser will derive, nudging the coefficient
    for _ in range(50):
        n_max = _n_max_from_text(header, jd_min, jd_max)
        if n_max == cfg.n_freq:
            break
        coef *= (n_max / cfg.n_freq) if n_max > 0 else 1.0
        if n_max > cfg.n_freq:
            coef *= 1.0 + 1e-7
        else:
            coef *= 1.0 - 1e-7
        header = _header(cfg, p_start, coef, p_end)
    lines = [header, f"{len(curves)}"]
    for ci, cv in enumerate(curves):
        calibrated = ci >= len(curves) - cfg.n_abs
        lines.append(f"{cv['t'].size} {1 if calibrated else 0}")
        for i in range(cv["t"].size):
            e0 = cv["sun"][i] * cv["dsun"][i]
            e = cv["earth"][i] * cv["dearth"][i]
            lines.append(f"{cv['t'][i]:.8f} {cv['flux'][i]:.10e} "
                         f"{e0[0]:.12f} {e0[1]:.12f} {e0[2]:.12f} "
                         f"{e[0]:.12f} {e[1]:.12f} {e[2]:.12f}")
    text = "\n".join(lines) + "\n"
    truth = dict(seed=seed, period_h=period_h, lambda_deg=math.degrees(lam) % 360.0,
                 beta_deg=math.degrees(beta), n_points=int(jd_all.size),
                 n_curves=len(curves), n_freq=cfg.n_freq, lmax=lmax, nrows=cfg.nrows,
                 jd0=jd0, shape=[float(x) for x in shape_c], par=list(cfg.par_true), cl=cfg.cl_true,
                 noise=cfg.noise)
    return text, truth


def truth_params(truth: dict) -> np.ndarray:
    """The truth as a parameter vector of the fitted model (Appendix B
    order: shape coefficients, beta' = 90 deg - beta, lambda, omega [rad/day],
    phase-function A, D, k, log of the Lambert weight, Lommel-Seeliger
    weight 1).  The rendering uses the parser's epoch and phi0 = 0, so this
    vector reproduces the noise-free data up to the facet discretisation."""
    shape = np.asarray(truth["shape"], float)
    a = np.concatenate([shape, np.zeros(8)])
    C = shape.size
    a[C] = math.radians(90.0 - truth["beta_deg"])
    a[C + 1] = math.radians(truth["lambda_deg"])
    a[C + 2] = 2.0 * math.pi * 24.0 / truth["period_h"]
    a[C + 3:C + 6] = truth["par"]
    a[C + 6] = math.log(truth["cl"])
    a[C + 7] = 1.0
    return a


def _header(cfg: SynthConfig, p_start, coef, p_end):
    A, D, k = cfg.par_init
    fa, fd, fk = cfg.par_free
    return "\n".join([
        f"{p_start:.12f} {coef:.12f} {p_end:.12f} 1  period start/step/stop (1 = free)",
        "0  epoch of zero time t0",
        "0  initial fixed rotation angle fi0",
        f"{cfg.conw}  the weight factor for conv. reg.",
        f"{cfg.lmax} {cfg.lmax}  degree and order of the Laplace series",
        f"{cfg.nrows}  nr. of triangulation rows per octant",
        f"{A} {fa}  initial guesses for phase funct. params.",
        f"{D} {fd}",
        f"{k} {fk}",
        f"{cfg.cl_init} {cfg.cl_free}  initial Lambert coeff. (L-S=1)",
        f"{cfg.stop_condition}  stop condition",
        f"{cfg.n_iter_min}  minimum number of iterations",
        f"{cfg.alamda_incr}  Alamda multiplicative factor",
        f"{cfg.alamda_start}  initial Alamda",
    ])


def _n_max_from_text(header: str, jd_min: float, jd_max: float) -> int:
    first = header.split("\n", 1)[0].split()
    ps, coef, pe = float(first[0]), float(first[1]), float(first[2])
    fs, fe = 1.0 / ps, 1.0 / pe
    fstep = 0.5 / (jd_max - jd_min) / 24 * coef
    return int((fs - fe) / fstep) + 1


def config_seed(config: int, index: int = 0) -> int:
    return SEED_BASE + 1000 * config + index


def batch_configs(n: int = 64, seed: Optional[int] = None) -> List[SynthConfig]:
    """The batched ragged workload: n asteroids of the same generator with
    point counts log-uniform in 300-8000 (dense : sparse = 1 : 2, dense
    curves of up to 150 points, sparse curves of 100) and 1000-4000 trial
    frequencies each, centred on the true period."""
    rng = np.random.default_rng(config_seed(3) if seed is None else seed)
    out = []
    for _ in range(n):
        npts = int(round(math.exp(rng.uniform(math.log(300.0), math.log(8000.0)))))
        dense = npts // 3
        n_dense = max(1, -(-dense // 150))
        pts_dense = max(10, dense // n_dense)
        n_sparse = max(10, npts - n_dense * pts_dense)
        n_freq = int(rng.integers(1000, 4001))
        out.append(dataclasses.replace(PRESETS["c1"], n_dense=n_dense, pts_dense=pts_dense, n_sparse=n_sparse,
                                       sparse_group=100, n_freq=n_freq))
    return out


def write_batch(directory: str, n: int = 64, seed: Optional[int] = None,
                overrides: Optional[dict] = None) -> List[dict]:
    """Writes asteroid_000.in ... for the batched workload; returns the truth
    dicts (with "path").  Asteroid i uses seed config_seed(3, i)."""
    import os

    os.makedirs(directory, exist_ok=True)
    truths = []
    for i, cfg in enumerate(batch_configs(n, seed)):
        if overrides:
            cfg = dataclasses.replace(cfg, **overrides)
        text, truth = make_asteroid(cfg, config_seed(3, i))
        path = os.path.join(directory, f"asteroid_{i:03d}.in")
        with open(path, "w") as f:
            f.write(text)
        truth.update(path=path, preset="c3", index=i)
        truths.append(truth)
    return truths


def write_asteroid(path: str, preset: str = "c1", seed: Optional[int] = None,
                   overrides: Optional[dict] = None) -> dict:
    cfg = dataclasses.replace(PRESETS[preset], **(overrides or {}))
    if seed is None:
        seed = config_seed({"c1": 1, "c2": 2, "c5": 5}.get(preset, 0))
    text, truth = make_asteroid(cfg, seed)
    with open(path, "w") as f:
        f.write(text)
    truth["preset"] = preset
    return truth


if __name__ == "__main__":
    import argparse

    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("out")
    ap.add_argument("--preset", default="c1", choices=sorted(PRESETS))
    ap.add_argument("--seed", type=int, default=None)
    ap.add_argument("--set", action="append", default=[], help="key=value override")
    a = ap.parse_args()
    ov = {}
    for kv in a.set:
        k, v = kv.split("=", 1)
        ov[k] = json.loads(v)
    print(json.dumps(write_asteroid(a.out, a.preset, a.seed, ov)))
