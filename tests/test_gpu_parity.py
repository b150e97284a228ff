"""MI355X kernels against the CPU FP64 oracle, through the C ABI.

The kernels and the oracle fix the same operation order, so every check here
is bit-for-bit: device exp/exp2/sin/cos, facet areas, raw Jacobian rows, the
augmented normal equations (alpha, beta, chi^2), the damped Gauss-Jordan,
whole LM fits (dev, dark, period, lambda, beta), the convexity-weight search
and complete period_search_out files, for relative and for mixed relative /
calibrated (absolute) lightcurves.  Full-size configurations are checked on
seeded samples of pairs.
"""
import os

import numpy as np
import pytest

import pyoracle as po

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def scan():
    import ps_claude_opt_amd as ps

    s = ps.Scan(0)
    yield s
    s.close()


def _prep(scan, path):
    scan.prepare_file(path)
    return po.Problem(path)


def _perturb(a, C, seed):
    a = a.copy()
    rng = np.random.default_rng(seed)
    a[1:C] += rng.normal(0, 0.05, C - 1)
    a[C:C + 3] += rng.normal(0, 0.01, 3)
    return a


def test_device_math_bit_identical(scan):
    rng = np.random.default_rng(7)
    x = np.concatenate([np.linspace(-40, 40, 40001), rng.uniform(-700, 700, 20000), rng.uniform(-1e5, 1e5, 2000)])
    e, e2, s, c = scan.math_probe(x)
    oe = np.array([po.exp(v) for v in x])
    oe2 = np.array([po.exp2(v) for v in x])
    osc = np.array([po.sincos(v) for v in x])
    assert np.array_equal(e, oe, equal_nan=True)
    assert np.array_equal(e2, oe2, equal_nan=True)
    assert np.array_equal(s, osc[:, 0], equal_nan=True)
    assert np.array_equal(c, osc[:, 1], equal_nan=True)


@pytest.mark.parametrize("name", ["tiny", "small", "c1", "c5"])
def test_eval_bit_identical(scan, inputs, name):
    path, _ = inputs(name)
    O = _prep(scan, path)
    C = O.info["ncoef"]
    for n, m, seed in [(1, 1, None), (2, 5, 1), (O.info["n_max"], 10, 2)]:
        a = O.initial_params(n, m)
        if seed is not None:
            a = _perturb(a, C, seed)
        cw2 = 0.37 ** 2
        A0, B0, c0, ar0, _ = O.mrqcof(a, cw2)
        A1, B1, c1, ar1, _ = scan.eval(a, cw2)
        assert np.array_equal(ar0, ar1, equal_nan=True), (name, n, m)
        assert np.array_equal(A0, A1, equal_nan=True), (name, n, m, np.abs(A0 - A1).max())
        assert np.array_equal(B0, B1, equal_nan=True), (name, n, m)
        assert c0 == c1 or (np.isnan(c0) and np.isnan(c1))


@pytest.mark.parametrize("name", ["tiny", "c1"])
def test_jacobian_rows_bit_identical(scan, inputs, name):
    path, _ = inputs(name)
    O = _prep(scan, path)
    N = O.info["ndata"]
    a = _perturb(O.initial_params(3, 7), O.info["ncoef"], 3)
    J0 = O.mrqcof(a, 0.0, jrows=True)[4][:N]
    J1 = scan.eval(a, 0.0, jrows=True)[4][:N]
    assert np.array_equal(J0, J1, equal_nan=True)


def _gaussj_cases(scan, inputs):
    rng = np.random.default_rng(11)
    cases = []
    for n in (8, 33, 54, 86):
        M = rng.normal(size=(n, n))
        cases.append((M @ M.T + n * np.eye(n), rng.normal(size=n), 0.01))
    # ties in |entries| exercise the last-occurrence pivot rule
    T = np.round(rng.normal(size=(20, 20)) * 2) / 2
    cases.append((T @ T.T + np.eye(20), rng.normal(size=20), 0.0))
    cases.append((np.eye(12), np.arange(12.0), 0.0))
    # real damped normal matrix
    path, _ = inputs("c1")
    O = _prep(scan, path)
    A, B, _, _, _ = O.mrqcof(O.initial_params(4, 2), 0.3 ** 2)
    cases.append((A, B, 0.001))
    sing = np.eye(6)
    sing[3, 3] = 0.0
    cases.append((sing, np.ones(6), 0.0))
    return cases


def test_gaussj_bit_identical(scan, inputs):
    for A_, b_, lam in _gaussj_cases(scan, inputs):
        x0, e0 = po.gaussj(A_, b_, lam)
        x1, e1 = scan.gaussj(A_, b_, lam)
        assert e0 == e1
        assert np.array_equal(x0, x1, equal_nan=True), (len(b_), np.abs(x0 - x1).max())


def test_gaussj_every_wave_count(scan, inputs):
    """The solver of each kernel variant (1, 2, 4 or 8 waves per workgroup,
    of which 1, 2, 4 and 4 own rows) agrees with the oracle bit for bit on
    the same systems."""
    before = scan.get_launch()
    try:
        for w in (1, 2, 8, 4):
            scan.set_launch(w, 0)
            for A_, b_, lam in _gaussj_cases(scan, inputs):
                x0, e0 = po.gaussj(A_, b_, lam)
                x1, e1 = scan.gaussj(A_, b_, lam)
                assert e0 == e1, (w, len(b_))
                assert np.array_equal(x0, x1, equal_nan=True), (w, len(b_), np.abs(x0 - x1).max())
    finally:
        scan.set_launch(before["waves_per_pair"], before["wg_per_cu"])


def _same(a, b):
    return all(x == y or (np.isnan(x) and np.isnan(y)) for x, y in zip(a, b))


def _compare_pairs(O, R, nl, ml, cw, cw_dev):
    res, _ = O.fit_pairs(nl, ml, cw, cw_dev, threads=min(16, os.cpu_count() or 1))
    for i, r in enumerate(res):
        g = R[i]
        assert _same((r.dev, r.dark, r.period, r.lambda_deg, r.beta_deg, r.chisq, r.niter),
                     (g["dev"], g["dark"], g["period"], g["lambda_deg"], g["beta_deg"], g["chisq"], g["niter"])), (
            nl[i], ml[i])
        assert r.n_vis == g["n_vis"], (nl[i], ml[i], r.n_vis, g["n_vis"])


def test_fit_pairs_bit_identical_tiny(scan, inputs):
    path, _ = inputs("tiny")
    O = _prep(scan, path)
    nl = [n for n in range(1, O.info["n_max"] + 1) for _ in range(10)]
    ml = [m for _ in range(O.info["n_max"]) for m in range(1, 11)]
    for cw, cwd in [(0.5, 0.5), (0.5, 0.0)]:
        R = scan.fit_pairs(nl, ml, cw, cwd)
        _compare_pairs(O, R, nl, ml, cw, cwd)


@pytest.mark.parametrize("preset,count", [("c1", 40), ("c2", 24)])
def test_fit_pairs_bit_identical_sample(scan, inputs, preset, count):
    path, _ = inputs(preset)
    O = _prep(scan, path)
    rng = np.random.default_rng(5)
    nl = list(rng.integers(1, O.info["n_max"] + 1, count))
    ml = list(rng.integers(1, 11, count))
    cw = 0.8
    R = scan.fit_pairs(nl, ml, cw, cw)
    _compare_pairs(O, R, nl, ml, cw, cw)


def test_fit_pairs_bit_identical_lmax8(scan, inputs):
    """The widest kernel variant (lmax = 8: 81 coefficients, 86 free
    parameters, 96-column rows, 512 facets) on a config-5-shaped asteroid
    with fewer sparse points, so the oracle finishes in seconds."""
    path, _ = inputs("c5_small", "c5", None, n_sparse=2000)
    O = _prep(scan, path)
    assert O.info["mfit"] > 64 and O.info["nfac"] == 512
    nl, ml = [1, 700, 9000, 19999], [1, 5, 8, 10]
    cw = 0.8
    R = scan.fit_pairs(nl, ml, cw, cw)
    _compare_pairs(O, R, nl, ml, cw, cw)


def test_threshold_stop_mode_bit_identical(scan, inputs):
    path, _ = inputs("tiny_thr", "tiny", None, stop_condition=0.001, n_iter_min=2)
    O = _prep(scan, path)
    nl = [1, 2, 3, 4, 5, 6] * 2
    ml = [1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 1, 2]
    R = scan.fit_pairs(nl, ml, 0.5, 0.5)
    _compare_pairs(O, R, nl, ml, 0.5, 0.5)
    assert (R["niter"] < 60).any()


@pytest.mark.parametrize("name", ["tiny", "small"])
def test_whole_application_identical(scan, inputs, tmp_path, name):
    import ps_claude_opt_amd as ps

    path, _ = inputs(name)
    O = po.Problem(path)
    out_o = tmp_path / f"{name}_oracle.out"
    cwo = O.run(str(out_o), threads=min(16, os.cpu_count() or 1))
    scan.prepare_file(path)
    cwg = scan.conw_search()
    assert cwg == cwo
    out_g = tmp_path / f"{name}_gpu.out"
    ps.write_output(str(out_g), scan.start(cwg))
    assert out_g.read_text() == out_o.read_text()


def test_golden_fixture_reproduced(scan, golden, tmp_path):
    import ps_claude_opt_amd as ps

    scan.prepare_file(golden("tiny.in"))
    cw = scan.conw_search()
    out = tmp_path / "tiny.out"
    ps.write_output(str(out), scan.start(cw))
    assert out.read_text() == open(golden("tiny.out")).read()


def test_conw_search_and_first_lines_c1(scan, inputs):
    path, _ = inputs("c1")
    O = _prep(scan, path)
    cw_o = O.conw_search(threads=min(16, os.cpu_count() or 1))
    cw_g = scan.conw_search()
    assert cw_g == cw_o
    lines = scan.start(cw_g, 1, 12)
    assert lines[0].dark_or_conw == (cw_g * scan.info["escl"]) * scan.info["escl"]
    nl = [n for n in range(1, 13) for _ in range(10)]
    ml = [m for _ in range(12) for m in range(1, 11)]
    res, _ = O.fit_pairs(nl, ml, cw_o, cw_o, threads=min(16, os.cpu_count() or 1))
    for i, L in enumerate(lines):
        grp = res[10 * i:10 * i + 10]
        devs = [r.dev for r in grp]
        best = 0
        for p in range(1, 10):
            if not np.isnan(devs[p]) and not (devs[p] >= devs[best]):
                best = p
        r = grp[best]
        assert (L.period_h, L.rms, L.lambda_deg, L.beta_deg) == (24 * r.period, r.dev, r.lambda_deg, r.beta_deg)


def test_error_paths(scan, tmp_path):
    import ps_claude_opt_amd as ps

    s = ps.Scan(0)
    try:
        with pytest.raises(ps.PsError) as e:
            s.start(0.5, 1, 3)
        assert e.value.code == -5
    finally:
        s.close()
    # a free-parameter mask with a gap (phase-function A fixed, D and k
    # free) is outside what the kernels support
    src = open(os.path.join(os.path.dirname(__file__), "golden", "tiny.in")).read().splitlines()
    assert src[6].split()[1] == "1"
    src[6] = src[6].split()[0] + " 0  phase A fixed"
    p = tmp_path / "gap.in"
    p.write_text("\n".join(src) + "\n")
    with pytest.raises(ps.PsError) as e:
        scan.prepare_file(str(p))
    assert e.value.code == -4


# ---- calibrated (absolute) lightcurves: no renormalisation, the size scale
# (parameter 1) free (Start.cu:3363-3692; period_search_BOINC.cpp:343-533)

@pytest.mark.parametrize("name,preset,n_abs", [("small_abs", "small", 1), ("c1_abs", "c1", 2)])
def test_absolute_eval_bit_identical(scan, inputs, name, preset, n_abs):
    path, _ = inputs(name, preset, None, n_abs=n_abs)
    O = _prep(scan, path)
    assert O.info["onlyrel"] == 0 and O.table("ia")[0] == 1
    C = O.info["ncoef"]
    for n, m, seed in [(1, 1, None), (3, 6, 4)]:
        a = O.initial_params(n, m)
        if seed is not None:
            a = _perturb(a, C, seed)
        A0, B0, c0, ar0, J0 = O.mrqcof(a, 0.3 ** 2, jrows=True)
        A1, B1, c1, ar1, J1 = scan.eval(a, 0.3 ** 2, jrows=True)
        assert np.array_equal(ar0, ar1)
        assert np.array_equal(J0, J1)
        assert np.array_equal(A0, A1), (name, n, m, np.abs(A0 - A1).max())
        assert np.array_equal(B0, B1)
        assert c0 == c1


def test_absolute_fit_pairs_bit_identical(scan, inputs):
    path, _ = inputs("c1_abs", "c1", None, n_abs=2)
    O = _prep(scan, path)
    rng = np.random.default_rng(9)
    nl = list(rng.integers(1, O.info["n_max"] + 1, 20))
    ml = list(rng.integers(1, 11, 20))
    R = scan.fit_pairs(nl, ml, 0.6, 0.6)
    _compare_pairs(O, R, nl, ml, 0.6, 0.6)


def test_absolute_whole_application_identical(scan, inputs, tmp_path):
    import ps_claude_opt_amd as ps

    path, _ = inputs("tiny_abs", "tiny", None, n_abs=1)
    O = po.Problem(path)
    out_o = tmp_path / "tiny_abs_oracle.out"
    cwo = O.run(str(out_o), threads=min(16, os.cpu_count() or 1))
    scan.prepare_file(path)
    cwg = scan.conw_search()
    assert cwg == cwo
    out_g = tmp_path / "tiny_abs_gpu.out"
    ps.write_output(str(out_g), scan.start(cwg))
    assert out_g.read_text() == out_o.read_text()
