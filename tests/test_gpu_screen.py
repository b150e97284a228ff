"""The learned frequency screen: its fold features against an independent
numpy computation, and the screened scan's refits against plain fits."""
import os

import numpy as np
import pytest

import ps_claude_opt_amd as ps
from ps_claude_opt_amd import _abi, screen, synth, warm

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _numpy_features(path: str, f_start: float, f_step: float, n_first: int, n_count: int) -> np.ndarray:
    """ps_screen_features restated with one complex exponential per point
    and frequency (no rotation recurrence)."""
    import sys

    sys.path.insert(0, os.path.join(os.path.dirname(HERE), "oracle"))
    import pyoracle

    P = pyoracle.Problem(path)
    N = P.info["ndata"]
    tim = P.table("tim")[:N]
    br = P.table("brightness")[:N]
    lp = P.table("l_points").astype(int)
    curves, off = [], 0
    for n in lp:
        if n >= 8:
            y = br[off:off + n] / br[off:off + n].mean() - 1.0
            if np.sum(y * y) > 0:
                curves.append((tim[off:off + n], y, tim[off + n - 1] - tim[off] <= 2.0))
        off += n
    out = np.zeros((n_count, _abi.load().ps_screen_n_features()), np.float64)
    for r in range(n_count):
        nu = 24.0 * (f_start - (n_first + r - 1) * f_step)
        period_h = 24.0 / nu
        acc = {g: dict(mean=np.zeros(4), mx=np.zeros(4), s2=np.zeros(4), var=0.0, n=0, pts=0) for g in (0, 1)}
        overlap = 0
        for t, y, dense in curves:
            g = 0 if dense else 1
            ph = 2 * np.pi * np.outer(np.arange(1, 5), nu * t)
            S = (y[None, :] * np.exp(1j * ph)).sum(1)
            v = float(np.sum(y * y))
            P_k = np.abs(S) ** 2 / v
            a = acc[g]
            a["mean"] += P_k
            a["mx"] = np.maximum(a["mx"], P_k / t.size)
            a["s2"] += np.abs(S) ** 2
            a["var"] += v
            a["n"] += 1
            a["pts"] += t.size
            if dense and 24.0 * (t[-1] - t[0]) > period_h:
                overlap += 1
        for g in (0, 1):
            a = acc[g]
            if a["n"] == 0:
                continue
            out[r, 12 * g:12 * g + 4] = np.log(a["mean"] / a["n"] + 1e-6)
            out[r, 12 * g + 4:12 * g + 8] = np.log(a["mx"] + 1e-6)
            out[r, 12 * g + 8:12 * g + 12] = np.log(a["s2"] / a["var"] + 1e-6)
        nd, ns = acc[0]["n"], acc[1]["n"]
        out[r, 24] = np.log(period_h / 10.0)
        out[r, 25] = overlap / nd if nd else 0.0
        out[r, 26] = np.log(acc[0]["pts"] / nd) if nd else 0.0
        out[r, 27] = np.log(acc[1]["pts"] / ns) if ns else 0.0
        out[r, 28] = np.log(1.0 + nd)
        out[r, 29] = np.log(1.0 + ns)
    return out


def test_screen_features_match_numpy(tmp_path):
    path = str(tmp_path / "c1.in")
    synth.write_asteroid(path, "c1", 11)
    with ps.Scan(0) as s:
        s.prepare_file(path)
        n_max = s.info["n_max"]
        got = s.screen_features(1, n_max)
        part = s.screen_features(37, 20)
        f_start, f_step = s.info["f_start"], s.info["f_step"]
    ref = _numpy_features(path, f_start, f_step, 1, n_max)
    assert got.shape == (n_max, 32) and got.dtype == np.float32
    # FP32 output of FP64 sums; the rotation recurrence agrees with direct
    # exponentials to ~1e-12
    assert np.abs(got - ref).max() < 2e-5 * max(1.0, np.abs(ref).max())
    assert np.array_equal(part, got[36:56]) or np.abs(part - got[36:56]).max() < 1e-6


def test_screen_features_errors(tmp_path):
    path = str(tmp_path / "c1.in")
    synth.write_asteroid(path, "c1", 11)
    with ps.Scan(0) as s:
        s.prepare_file(path)
        n_max = s.info["n_max"]
        for args in ((1, n_max + 1, 0), (0, 5, 0), (1, 5, 1)):
            with pytest.raises(ps.PsError):
                s.screen_features(*args[:2], prob=args[2])
        assert s.screen_features(1, 0).shape == (0, 32)


@pytest.mark.skipif(not os.path.exists(screen.DEFAULT_WEIGHTS), reason="no trained screen shipped")
def test_screened_scan_refits_are_plain_fits(tmp_path):
    """On the config-2 asteroid (5,000 frequencies over 2-40 h, the kind of
    scan the screen was trained on): the screened scan's kept set holds the
    plain winner, its refitted lines are the plain scan's bit for bit, and
    its winner is the plain winner."""
    path = str(tmp_path / "c2.in")
    synth.write_asteroid(path, "c2", synth.config_seed(2))
    sc = screen.Screen.load()
    with ps.Scan(0) as s:
        s.prepare_file(path)
        n_max = s.info["n_max"]
        cw = s.conw_search()
        plain = s.start(cw, 1, n_max)
        net = sc.network(s)
        res, refitted, kept = screen.screened_scan(s, sc, net, cw, 1, n_max, n_iter=20)
        net.close()
    rms = np.array([L.rms for L in plain])
    f_best = int(np.nanargmin(rms)) + 1
    assert f_best in set(kept.tolist())
    assert kept.size == sc.n_keep(n_max)
    best = warm.best_line(res, 1, n_max, among=refitted)
    assert best["freq_index"] == f_best
    L = plain[f_best - 1]
    assert best["rms"] == L.rms and best["lambda_deg"] == L.lambda_deg and best["beta_deg"] == L.beta_deg
    # every refitted frequency's best pole is the plain scan's line
    bl = warm.best_lines(res, n_max)
    for f in refitted:
        r = res[(f - 1) * 10 + bl[f - 1]]
        assert r["dev"] == plain[f - 1].rms
    # frequencies neither kept nor refitted (alias refits may lie outside
    # the kept set) were not fitted
    idle = np.setdiff1d(np.setdiff1d(np.arange(1, n_max + 1), kept), refitted)
    assert idle.size > 0 and np.all(np.isnan(res["dev"].reshape(n_max, 10)[idle - 1]))


@pytest.mark.skipif(not os.path.exists(screen.DEFAULT_WEIGHTS), reason="no trained screen shipped")
def test_screened_batch_matches_plain_and_single_scans(tmp_path):
    """A batch of four config-2-style asteroids (the held-out evaluation's
    first four): the batched screened scan reports each asteroid's plain
    winner line, and asteroid 0's results equal the single-asteroid
    screened scan's bit for bit."""
    paths = []
    for i in range(4):
        p = str(tmp_path / f"a{i}.in")
        synth.write_asteroid(p, "c2", synth.config_seed(2, 300 + i))
        paths.append(p)
    sc = screen.Screen.load()
    with ps.Scan(0) as s:
        s.prepare_files(paths)
        cw = s.conw_search_batch()
        plain = s.start_batch(cw)
        net = sc.network(s)
        got = screen.screened_scan_batch(s, sc, net, cw, n_iter=20)
        net.close()
    for i, (res, refitted, kept) in enumerate(got):
        n_max = len(plain[i])
        rms = np.array([L.rms for L in plain[i]])
        f_best = int(np.nanargmin(rms)) + 1
        b = warm.best_line(res, 1, n_max, among=refitted)
        L = plain[i][f_best - 1]
        assert (b["freq_index"], b["rms"], b["lambda_deg"], b["beta_deg"]) == (f_best, L.rms, L.lambda_deg, L.beta_deg), i
        assert kept.size == sc.n_keep(n_max)
    with ps.Scan(0) as s:
        s.prepare_file(paths[0])
        net = sc.network(s)
        res0, fr0, kept0 = screen.screened_scan(s, sc, net, float(cw[0]), 1, s.info["n_max"], n_iter=20)
        net.close()
    res_b, fr_b, kept_b = got[0]
    assert np.array_equal(kept0, kept_b) and np.array_equal(fr0, fr_b)
    for f in ("dev", "chisq", "lambda_deg", "beta_deg", "period", "niter"):
        assert np.array_equal(res0[f], res_b[f], equal_nan=True), f
