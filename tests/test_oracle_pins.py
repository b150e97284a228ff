"""Independent pins of the oracle functions that a bit-exact GPU comparison
cannot pin: the synthetic truth, Iter2's rms, FinishPole's dark-facet
percentage, the convexity pseudo-curve rows and the convexity-weight
decision.

  * the generator's shapes close their Gaussian image (a convex body exists)
    and the model evaluated at the true parameters reproduces noise-free
    data down to a facet-discretisation floor that falls as nrows^-2;
  * a fit started at the truth ends at that floor without noise and at the
    noise level with 1 % noise, at the true pole and period;
  * Iter2 (Start.cu:5129-5248): dev^2 * ndata + |sum A n|^2 conw_r^2 equals
    the fit's chi^2, with chi^2 and the areas recomputed in numpy from the
    tables (not from the oracle's own rows);
  * FinishPole (Start.cu:5056-5125): dark = 100 |sum A n| / sum A from numpy
    areas; zero for a pure l = 0 shape (a closed polyhedron);
  * the convexity rows (Start.cu:2539-2612) against numpy
    sum_f A_f n_j,f Y_c(f) and against finite differences of sum_f A_f n_j,f;
  * the conw decision (start_CUDA.cu:747-789) on crafted dark values.
"""
import math

import numpy as np
import pytest

import pyoracle as po
from ps_claude_opt_amd import synth
from test_oracle_model import _normal_equations_numpy


def _np_area(P, a):
    """curv (Start.cu:421-528) in numpy: Darea * exp(sum_k a_k Y_k)."""
    C, F = P.info["ncoef"], P.info["nfac"]
    Y = P.table("d_sphere").reshape(C, F)
    return P.table("d_area") * np.exp(np.asarray(a[:C]) @ Y)


def _np_normals(P):
    return P.table("normal").reshape(3, P.info["nfac"])


def _np_convexity_rows(P, a):
    """The three pseudo-curve rows: y_j = sum_f A_f n_jf and
    J_j[c] = sum_f A_f n_jf Y_c(f) for shape columns c, 0 beyond."""
    C, F, ma = P.info["ncoef"], P.info["nfac"], P.info["ma"]
    A = _np_area(P, a)
    n = _np_normals(P)
    Y = P.table("d_sphere").reshape(C, F)
    rows = np.zeros((3, ma + 1))
    rows[:, 0] = n @ A
    rows[:, 1:C + 1] = (n * A) @ Y.T
    return rows


def _np_chisq(P, a, conw_r2):
    """chi^2 of the renormalised model (relative curves: c = sig L / sum y)
    plus the convexity rows, from the oracle's model brightness column and
    numpy convexity rows."""
    N = P.info["ndata"]
    _, _, _, _, J = P.mrqcof(a, conw_r2, jrows=True)
    y = J[:N, 0]
    lp = P.table("l_points").astype(int)
    sig, br = P.table("sig"), P.table("brightness")
    w = P.table("sigr2") * P.table("weight")
    chi = 0.0
    p0 = 0
    for L in lp:
        c = sig[p0:p0 + L] * L / y[p0:p0 + L].sum()
        r = br[p0:p0 + L] - c * y[p0:p0 + L]
        chi += float(r @ (w[p0:p0 + L] * r))
        p0 += L
    v = _np_convexity_rows(P, a)[:, 0]
    return chi + conw_r2 * float(v @ v)


@pytest.fixture(scope="module")
def c1_clean(tmp_path_factory):
    path = str(tmp_path_factory.mktemp("truth") / "c1_clean.in")
    t = synth.write_asteroid(path, "c1", None, {"noise": 0.0})
    return po.Problem(path), t


@pytest.fixture(scope="module")
def small(tmp_path_factory):
    path = str(tmp_path_factory.mktemp("pins") / "small.in")
    t = synth.write_asteroid(path, "small")
    return po.Problem(path), t


# ---------------------------------------------------------------- generator

@pytest.mark.parametrize("lmax", [1, 3, 6, 8])
def test_generator_closes_gaussian_image(lmax):
    nrm = synth._fib_sphere(4000)
    Y, lv = synth._laplace_basis(lmax, nrm)
    rng = np.random.default_rng(lmax)
    # the generator's draw: sigma 0.3 / (l + 1) per unit-rms basis column,
    # three times larger so the image starts far from closed
    c = rng.normal(0, 0.9, Y.shape[1]) / (lv + 1.0) / np.sqrt(np.mean(Y * Y, axis=0))
    c[0] = 0.0
    before = np.linalg.norm(nrm.T @ np.exp(Y @ c)) / np.exp(Y @ c).sum()
    cc = synth.close_gaussian_image(Y, nrm, c, lmax)
    A = np.exp(Y @ cc)
    assert np.linalg.norm(nrm.T @ A) / A.sum() < 1e-12 < before
    # only the l = 1 coefficients move
    moved = np.flatnonzero(cc != c)
    assert set(moved) <= set(synth.l1_columns(lmax))
    # the l = 1 columns are the Cartesian components of the normal
    iz, ix, iy = synth.l1_columns(lmax)
    assert np.allclose(Y[:, [ix, iy, iz]], nrm, atol=1e-14)


def test_truth_dark_falls_with_resolution(tmp_path):
    """Sum A n of the true shape on the fitted triangulation tends to 0 as
    nrows grows (the truth is closed on the fine sphere)."""
    darks = []
    for nrows in (4, 8, 12):
        path = str(tmp_path / f"c1_{nrows}.in")
        t = synth.write_asteroid(path, "c1", None, {"noise": 0.0, "nrows": nrows})
        P = po.Problem(path)
        a = synth.truth_params(t)
        A = _np_area(P, a)
        darks.append(100 * np.linalg.norm(_np_normals(P) @ A) / A.sum())
    assert darks[0] > darks[1] > darks[2]
    assert darks[2] < 0.05


def test_model_at_truth_reaches_discretisation_floor(tmp_path):
    """dev of the model at the true parameters on noise-free data: the
    quadrature error of nrows rows per octant, O(nrows^-2)."""
    devs = {}
    for nrows in (4, 8, 12):
        path = str(tmp_path / f"c1_{nrows}.in")
        t = synth.write_asteroid(path, "c1", None, {"noise": 0.0, "nrows": nrows})
        P = po.Problem(path)
        a = synth.truth_params(t)
        cw = P.info["conw"] / P.info["escl"] ** 2
        _, _, chi, _, _ = P.mrqcof(a, cw * cw)
        devs[nrows] = P.finish(a, chi, cw)[0]
    assert devs[12] < 7e-4
    # second order: tripling nrows cuts the floor by well over 5x
    assert devs[4] / devs[12] > 10.0
    assert devs[4] / devs[8] > 3.0


def test_fit_from_truth_noise_free(c1_clean):
    P, t = c1_clean
    a = synth.truth_params(t)
    cw = P.info["conw"] / P.info["escl"] ** 2
    (r,), af = P.fit_params(a, cw, cw)
    assert r.niter == 50
    # discretisation floor of 288 facets (nrows 6)
    assert r.dev < 2e-3
    assert abs(r.lambda_deg - t["lambda_deg"]) < 0.5
    assert abs(r.beta_deg - t["beta_deg"]) < 0.5
    assert abs(24 * r.period / t["period_h"] - 1) < 1e-6


def test_fit_from_truth_with_noise(tmp_path):
    path = str(tmp_path / "c1_noisy.in")
    t = synth.write_asteroid(path, "c1")
    assert t["noise"] == 0.01
    P = po.Problem(path)
    a = synth.truth_params(t)
    cw = P.info["conw"] / P.info["escl"] ** 2
    (r,), _ = P.fit_params(a, cw, cw)
    assert abs(r.dev - 0.01) < 0.1 * 0.01
    assert abs(r.lambda_deg - t["lambda_deg"]) < 2.0
    assert abs(r.beta_deg - t["beta_deg"]) < 2.0
    assert abs(24 * r.period / t["period_h"] - 1) < 1e-5


def test_fit_from_truth_absolute_curves(tmp_path):
    """With two calibrated sparse curves the size scale is fitted too: it
    stays at the truth's 0 up to the facet discretisation."""
    path = str(tmp_path / "c1_abs.in")
    t = synth.write_asteroid(path, "c1", None, {"noise": 0.0, "n_abs": 2})
    P = po.Problem(path)
    assert P.info["onlyrel"] == 0 and P.table("ia")[0] == 1
    a = synth.truth_params(t)
    cw = P.info["conw"] / P.info["escl"] ** 2
    (r,), af = P.fit_params(a, cw, cw)
    assert r.dev < 2e-3
    assert abs(af[0][0]) < 0.05
    assert abs(r.lambda_deg - t["lambda_deg"]) < 0.5
    assert abs(r.beta_deg - t["beta_deg"]) < 0.5


def test_scan_near_truth_finds_true_pole(tmp_path):
    """The best line of the frequencies next to the true one, 10 poles
    each from the reference's initialisation, recovers the pole (or its
    ecliptic mirror lambda + 180, which near-ecliptic geometry cannot tell
    apart) and the period."""
    path = str(tmp_path / "c1_scan.in")
    t = synth.write_asteroid(path, "c1")
    P = po.Problem(path)
    f_true = 1.0 / t["period_h"]
    n_true = int(round((P.info["f_start"] - f_true) / P.info["f_step"])) + 1
    ns = [n for n in range(n_true - 1, n_true + 2) for _ in range(10)]
    ms = [m for _ in range(3) for m in range(1, 11)]
    cw = P.info["conw"] / P.info["escl"] ** 2
    res, _ = P.fit_pairs(ns, ms, cw, cw, threads=0)
    best = min(res, key=lambda r: r.dev)
    assert best.dev < 0.0105
    assert abs(24 * best.period / t["period_h"] - 1) < 1e-5
    dl = min(abs((best.lambda_deg - t["lambda_deg"] + 180) % 360 - 180),
             abs((best.lambda_deg - t["lambda_deg"]) % 360 - 180))
    assert dl < 3.0
    assert abs(best.beta_deg - t["beta_deg"]) < 3.0


# ---------------------------------------------------------------- curv / rows

def test_curv_matches_numpy_general_shape(small):
    P, _ = small
    a = P.initial_params(2, 4)
    a[:P.info["ncoef"]] += np.random.default_rng(5).normal(0, 0.1, P.info["ncoef"])
    _, _, _, area, _ = P.mrqcof(a, 0.0)
    ref = _np_area(P, a)
    assert np.abs(area / ref - 1).max() < 1e-13


@pytest.mark.parametrize("seed", [0, 1])
def test_convexity_rows_match_numpy(small, seed):
    P, _ = small
    N, C = P.info["ndata"], P.info["ncoef"]
    a = P.initial_params(3, 7)
    a[:C] += np.random.default_rng(seed).normal(0, 0.1, C)
    J = P.mrqcof(a, 0.5, jrows=True)[4]
    got = J[N:N + 3]
    ref = _np_convexity_rows(P, a)
    scale = np.abs(ref[:, 1:C + 1]).max()
    assert np.abs(got[:, 0] - ref[:, 0]).max() < 1e-13 * _np_area(P, a).sum()
    assert np.abs(got[:, 1:C + 1] - ref[:, 1:C + 1]).max() < 1e-13 * scale
    assert np.all(got[:, C + 1:] == 0.0)


def test_convexity_rows_are_derivatives_of_sum_an(small):
    P, _ = small
    N, C = P.info["ndata"], P.info["ncoef"]
    a = P.initial_params(3, 7)
    a[:C] += np.random.default_rng(3).normal(0, 0.1, C)
    J = P.mrqcof(a, 0.5, jrows=True)[4][N:N + 3]
    for col in range(2, C + 1):
        h = 1e-6
        ap, am = a.copy(), a.copy()
        ap[col - 1] += h
        am[col - 1] -= h
        yp = P.mrqcof(ap, 0.5, jrows=True)[4][N:N + 3, 0]
        ym = P.mrqcof(am, 0.5, jrows=True)[4][N:N + 3, 0]
        fd = (yp - ym) / (2 * h)
        assert np.abs(fd - J[:, col]).max() < 1e-7 * max(1.0, np.abs(J[:, col]).max()), col


# ---------------------------------------------------------------- Iter2 / FinishPole

@pytest.mark.parametrize("conw_dev", ["scan", "precalc"])
def test_iter2_rms_identity(small, conw_dev):
    """The fit's final dev, chi^2 and parameters satisfy
    dev^2 * ndata + |sum A n|^2 conw_r_dev^2 = chi^2 with chi^2 and A from
    numpy, dev over the data points only (the reference's ndata - 3)."""
    P, _ = small
    N = P.info["ndata"]
    cw = 0.3
    cw_dev = cw if conw_dev == "scan" else 0.0
    ns, ms = [2, 5, 9], [1, 6, 10]
    res, af = P.fit_pairs_from(ns, ms, cw, cw_dev, want_final=True)
    for r, a in zip(res, af):
        chi = _np_chisq(P, a, cw * cw)
        assert abs(r.chisq - chi) < 1e-11 * chi
        A = _np_area(P, a)
        v = _np_normals(P) @ A
        lhs = r.dev ** 2 * N + float(v @ v) * cw_dev ** 2
        assert abs(lhs - chi) < 1e-11 * chi
        # a plausible slip, ndata + 3 in the divisor, is far outside
        assert abs(r.dev ** 2 * (N + 3) + float(v @ v) * cw_dev ** 2 - chi) > 1e-4 * chi
    # dark = 100 |sum A n| / sum A.  |sum A n| is Iter2's, of the accepted
    # parameters; sum A is of the areas the last curv call left behind
    # (Start.cu:5056-5125 reads Areag, which mrqcof's curv overwrites at the
    # trial parameters, Start.cu:546): after a final rejected step those are
    # the rejected trial's.  The trajectory tells which case a pair is in.
    for n, m, r, a in zip(ns, ms, res, af):
        _, traj = P.fit_pair(n, m, cw, cw_dev, traj_cap=60)
        lam = traj[:r.niter, 0]
        prev = lam[-2] if r.niter > 1 else P.info["alamda_start"]
        accepted_last = lam[-1] < prev
        A = _np_area(P, a)
        v = _np_normals(P) @ A
        ratio = r.dark / (100 * np.linalg.norm(v) / A.sum())
        if accepted_last:
            assert abs(ratio - 1) < 1e-9
        else:
            # the rejected trial: a + solve(alpha (1 + lambda) on the
            # diagonal, beta) at the accepted a, lambda before its last x10
            alpha, beta, _ = _normal_equations_numpy(P, a, cw * cw)
            lam_used = lam[-1] / P.info["alamda_incr"]
            Ad = alpha.copy()
            Ad[np.diag_indices_from(Ad)] *= 1 + lam_used
            atry = a.copy()
            atry[P.table("ia").astype(bool)] += np.linalg.solve(Ad, beta)
            ref = 100 * np.linalg.norm(v) / _np_area(P, atry).sum()
            assert abs(r.dark / ref - 1) < 1e-7


def test_finish_dark_zero_for_l0_shape(small):
    P, _ = small
    a = P.initial_params(1, 1)
    C = P.info["ncoef"]
    a[:C] = 0.0
    a[0] = 0.7
    dev, chck, dark = P.finish(a, 1.0, 0.0)
    A = _np_area(P, a)
    assert chck < 1e-13 * A.sum()
    assert dark < 1e-11
    # l = 1 shift: the areas tilt towards g, the image opens by a known amount
    iz, ix, iy = synth.l1_columns(P.info["lmax"])
    a[ix], a[iy], a[iz] = 0.2, -0.1, 0.05
    dev, chck, dark = P.finish(a, 10.0, 0.5)
    A = _np_area(P, a)
    v = _np_normals(P) @ A
    assert abs(chck - np.linalg.norm(v)) < 1e-13 * A.sum()
    assert abs(dark - 100 * np.linalg.norm(v) / A.sum()) < 1e-12 * dark
    # the open image points along g (to first order in the small tilt)
    g = np.array([0.2, -0.1, 0.05])
    assert float(v @ g) / (np.linalg.norm(v) * np.linalg.norm(g)) > 0.99
    # dev = sqrt((chi^2 - |v|^2 conw^2) / ndata)
    assert abs(dev - math.sqrt((10.0 - float(v @ v) * 0.25) / P.info["ndata"])) < 1e-14


# ---------------------------------------------------------------- conw decision

def _results(darks, devs=None):
    out = []
    for i, d in enumerate(darks):
        r = po.PairResult()
        r.reported = 1
        r.dark = d
        r.dev = 1.0 if devs is None else devs[i]
        out.append(r)
    return out


def test_conw_decision_on_crafted_darks():
    # 10 periods x 10 poles, all poles equal: the mean of the first poles
    darks = [0.99] * 100
    assert po.conw_decide(_results(darks), 0.25) == (0.25, 1)
    # mean exactly 1 doubles (>= 1 is "still not good")
    assert po.conw_decide(_results([1.0] * 100), 0.25) == (0.5, 0)
    # the mean is taken over the best pole of each period: pole 3 wins
    darks, devs = [], []
    for n in range(10):
        for m in range(10):
            devs.append(0.5 if m == 3 else 1.0)
            darks.append(0.2 if m == 3 else 50.0)
    assert po.conw_decide(_results(darks, devs), 0.25) == (0.25, 1)
    # nine periods at 0 and one at 15: mean 1.5 -> double
    assert po.conw_decide(_results([0.0] * 90 + [15.0] * 10), 0.1) == (0.2, 0)
    # 0.9 mean (nine periods at 1.0, one at 0.0) accepts
    assert po.conw_decide(_results([1.0] * 90 + [0.0] * 10), 0.1) == (0.1, 1)
    # a period with no reported pole adds nothing but still counts in the mean
    rs = _results([5.0] * 100)
    for r in rs[:90]:
        r.reported = 0
    assert po.conw_decide(rs, 0.1) == (0.1, 1)  # 5.0 / 10 = 0.5
