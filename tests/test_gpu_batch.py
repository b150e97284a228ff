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
    """4 or 8 wavefronts per pair, 1 or 2 workgroups per CU: the same
    results bit for bit (only which wavefront computes what changes)."""
    import ps_claude_opt_amd as ps

    path, _ = inputs("c1")
    rng = np.random.default_rng(21)
    nl = rng.integers(1, 201, 48)
    ml = rng.integers(1, 11, 48)
    with ps.Scan(0) as s:
        s.prepare_file(path)
        ref = None
        for w, g in [(4, 0), (4, 1), (8, 0), (8, 1)]:
            s.set_launch(w, g)
            got = s.get_launch()
            assert got["waves_per_pair"] == w
            R = s.fit_pairs(nl, ml, 0.6)
            if ref is None:
                ref = R
            else:
                for f in FIELDS:
                    assert all(_same(a, b) for a, b in zip(ref[f], R[f])), (w, g, f)
        best = s.autotune(np.zeros(48, np.int32), nl, ml, 0.6)
        assert best["waves_per_pair"] in (4, 8) and best["wg_per_cu"] >= 1 and best["kernel_ms"] > 0
        assert s.get_launch() == {"waves_per_pair": best["waves_per_pair"], "wg_per_cu": best["wg_per_cu"]}
