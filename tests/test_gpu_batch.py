"""Batched ragged scans: several asteroids (different point counts, scan
ranges and conwR) fitted in one persistent launch.  Every pair's result must
be bit-identical to the same pair fitted alone and to the CPU oracle, the
batched convexity-weight search and scan lines must equal the per-asteroid
ones, and asteroids that do not share a shape basis are rejected."""
import os

import numpy as np
import pytest

import pyoracle as po

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)


@pytest.fixture(scope="module")
def batch_files(tmp_path_factory):
    from ps_claude_opt_amd import synth

    d = tmp_path_factory.mktemp("batch")
    # same shape basis (lmax 3, nrows 3), ragged point counts and scan lengths
    specs = [dict(n_dense=1, pts_dense=30, n_sparse=25, n_freq=5),
             dict(n_dense=3, pts_dense=50, n_sparse=130, n_freq=8),
             dict(n_dense=2, pts_dense=17, n_sparse=41, n_freq=6)]
    paths = []
    for i, ov in enumerate(specs):
        p = str(d / f"a{i}.in")
        synth.write_asteroid(p, "tiny", synth.config_seed(3, 900 + i), ov)
        paths.append(p)
    return paths


def _same(a, b):
    return a == b or (np.isnan(a) and np.isnan(b))


FIELDS = ("dev", "dark", "period", "lambda_deg", "beta_deg", "chisq", "niter", "n_vis")


def test_batch_items_equal_single_and_oracle(batch_files):
    import ps_claude_opt_amd as ps

    rng = np.random.default_rng(3)
    with ps.Scan(0) as b:
        b.prepare_files(batch_files)
        assert len(b.batch_info) == 3
        prob = rng.integers(0, 3, 60)
        nl = np.array([rng.integers(1, b.batch_info[p]["n_max"] + 1) for p in prob])
        ml = rng.integers(1, 11, 60)
        cw = np.array([0.3, 0.7, 1.1])[prob]
        R = b.fit_items(prob, nl, ml, cw, cw * 0.5)
    for p, path in enumerate(batch_files):
        sel = np.nonzero(prob == p)[0]
        with ps.Scan(0) as s:
            s.prepare_file(path)
            S = s.fit_pairs(nl[sel], ml[sel], float(cw[sel[0]]), float(cw[sel[0]] * 0.5))
        O = po.Problem(path)
        ores, _ = O.fit_pairs(nl[sel], ml[sel], float(cw[sel[0]]), float(cw[sel[0]] * 0.5), threads=THREADS)
        for k, i in enumerate(sel):
            for f in FIELDS:
                assert _same(R[i][f], S[k][f]), (p, i, f)
                assert _same(R[i][f], float(getattr(ores[k], f))), (p, i, f)


def test_batch_conw_search_and_scan_equal_single(batch_files):
    import ps_claude_opt_amd as ps

    with ps.Scan(0) as b:
        b.prepare_files(batch_files)
        cw = b.conw_search_batch()
        lines = b.start_batch(cw)
    for p, path in enumerate(batch_files):
        with ps.Scan(0) as s:
            s.prepare_file(path)
            cw1 = s.conw_search()
            assert cw1 == cw[p]
            single = s.start(cw1)
        assert [ps.format_line(L) for L in lines[p]] == [ps.format_line(L) for L in single]


def test_batch_rejects_different_basis(batch_files, tmp_path):
    import ps_claude_opt_amd as ps
    from ps_claude_opt_amd import synth

    other = str(tmp_path / "other.in")
    synth.write_asteroid(other, "small", synth.config_seed(3, 950))
    with ps.Scan(0) as b:
        with pytest.raises(ps.PsError) as e:
            b.prepare_files([batch_files[0], other])
        assert e.value.code == -4


def test_launch_configurations_bit_identical(inputs):
    """1, 2, 4 or 8 wavefronts per pair, 1 or all resident workgroups per
    CU, 16 to 256 points per facet-pass claim: the same results bit for bit
    (only which wavefront computes what changes)."""
    import ps_claude_opt_amd as ps

    path, _ = inputs("c1")
    rng = np.random.default_rng(21)
    nl = rng.integers(1, 201, 48)
    ml = rng.integers(1, 11, 48)
    with ps.Scan(0) as s:
        s.prepare_file(path)
        opts = s.launch_options()
        assert set(opts) == {1, 2, 4, 8}, opts
        ref = None
        for w, g, t in [(4, 0, 16), (4, 1, 16), (8, 0, 16), (8, 1, 16), (2, 0, 16), (1, 0, 16), (4, 0, 64),
                        (4, 0, 128), (2, 0, 256)]:
            s.set_launch(w, g)
            s.set_tile_points(t)
            got = s.get_launch()
            assert got["waves_per_pair"] == w and got["tile_points"] == t
            R = s.fit_pairs(nl, ml, 0.6)
            if ref is None:
                ref = R
            else:
                for f in FIELDS:
                    assert all(_same(a, b) for a, b in zip(ref[f], R[f])), (w, g, t, f)
        best = s.autotune(np.zeros(48, np.int32), nl, ml, 0.6)
        assert best["waves_per_pair"] in opts and 1 <= best["wg_per_cu"] <= opts[best["waves_per_pair"]]
        assert best["tile_points"] in (16, 64, 128, 256) and best["kernel_ms"] > 0
        got = s.get_launch()
        assert got == {k: best[k] for k in ("waves_per_pair", "wg_per_cu", "tile_points")}


@pytest.fixture(scope="module")
def config3_files(tmp_path_factory):
    """The batched ragged workload: 64 seeded lmax-6 asteroids with 300 to
    8,000 points (dense : sparse = 1 : 2) and 1,000-4,000 frequencies each."""
    from ps_claude_opt_amd import synth

    d = str(tmp_path_factory.mktemp("c3"))
    truths = synth.write_batch(d, 64)
    return [t["path"] for t in truths], truths


def test_config3_workload_sample(config3_files):
    """One persistent launch over all 64 asteroids of the ragged workload:
    a seeded sample of 128 items (two per asteroid) is bit-identical to the
    CPU oracle and, for eight asteroids, to the same pairs fitted alone; the
    autotuner then runs on the batch and its choice gives the same results."""
    import ps_claude_opt_amd as ps

    paths, truths = config3_files
    npts = [t["n_points"] for t in truths]
    assert min(npts) >= 300 and max(npts) <= 8200 and max(npts) > 4000
    rng = np.random.default_rng(33)
    with ps.Scan(0) as b:
        b.prepare_files(paths)
        info = b.batch_info
        assert len(info) == 64 and all(1000 <= x["n_max"] <= 4000 for x in info)
        prob = np.repeat(np.arange(64), 2)
        nl = np.array([rng.integers(1, info[p]["n_max"] + 1) for p in prob])
        ml = rng.integers(1, 11, prob.size)
        cw = 0.5 + 0.01 * prob
        R = b.fit_items(prob, nl, ml, cw, cw)
        best = b.autotune(prob, nl, ml, cw)
        assert best["kernel_ms"] > 0
        R2 = b.fit_items(prob, nl, ml, cw, cw)
    for f in FIELDS:
        assert all(_same(x, y) for x, y in zip(R[f], R2[f])), f
    from concurrent.futures import ThreadPoolExecutor

    def oracle(p):
        sel = np.nonzero(prob == p)[0]
        O = po.Problem(paths[p])
        ores, _ = O.fit_pairs(nl[sel], ml[sel], float(cw[sel[0]]), float(cw[sel[0]]), threads=2)
        return p, sel, ores

    # the oracle's ctypes calls release the GIL: asteroids run side by side
    with ThreadPoolExecutor(max_workers=max(1, THREADS // 2)) as ex:
        for p, sel, ores in ex.map(oracle, range(64)):
            for k, i in enumerate(sel):
                for f in FIELDS:
                    assert _same(R[i][f], float(getattr(ores[k], f))), (p, i, f)
    for p in range(0, 64, 8):
        sel = np.nonzero(prob == p)[0]
        with ps.Scan(0) as s:
            s.prepare_file(paths[p])
            S = s.fit_pairs(nl[sel], ml[sel], float(cw[sel[0]]), float(cw[sel[0]]))
        for k, i in enumerate(sel):
            for f in FIELDS:
                assert _same(R[i][f], S[k][f]), (p, i, f)
