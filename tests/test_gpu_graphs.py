"""HIP-graph mode: every fit launch (work-counter reset + persistent LM
kernel) captured into a graph and replayed while its arguments stay the
same.  Scan lines, raw pair results, the convexity-weight search, batched
scans and fixed-iteration warm fits must be bit-identical with graphs on and
off; the graph is replayed several times with the library's own eager
kernels and host synchronisations between the replays, and every replay is
checked; a changed launch (other item count) is re-captured, not replayed.
"""
import dataclasses

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("dev", "dark", "period", "lambda_deg", "beta_deg", "chisq", "niter", "n_vis")


def _same_results(a, b):
    for f in FIELDS:
        x, y = np.asarray(a[f]), np.asarray(b[f])
        assert np.array_equal(x, y, equal_nan=np.issubdtype(x.dtype, np.floating)), f


def _same_lines(a, b):
    assert len(a) == len(b)
    for x, y in zip(a, b):
        assert np.array_equal(np.array(dataclasses.astuple(x)), np.array(dataclasses.astuple(y)), equal_nan=True)


def _eager_work(s, rng):
    """Eager kernels of the library on the same stream, each synchronised."""
    info = s.info
    a = rng.normal(0.0, 0.1, info["ma"])
    a[:3] += 1.0
    alpha, beta, chisq, _, _ = s.eval(a, 0.1)
    assert np.isfinite(chisq)
    s.gaussj(alpha + np.eye(len(beta)), beta, 0.01)


def test_graph_scan_replays_equal_eager(inputs):
    import ps_claude_opt_amd as ps

    rng = np.random.default_rng(11)
    with ps.Scan(0) as s:
        s.prepare_file(inputs("tiny")[0])
        n_max = s.info["n_max"]
        cw_eager = s.conw_search()
        lines_eager = s.start(cw_eager)
        n = rng.integers(1, n_max + 1, 23)
        m = rng.integers(1, 11, 23)
        pairs_eager = s.fit_pairs(n, m, cw_eager, 0.5 * cw_eager)

        s.set_graphs(True)
        assert s.conw_search() == cw_eager
        lines = s.start(cw_eager)
        _same_lines(lines, lines_eager)
        st1 = s.graph_stats()
        assert st1["captures"] >= 1 and st1["replays"] >= st1["captures"]
        for _ in range(3):
            _eager_work(s, rng)
            _same_lines(s.start(cw_eager), lines_eager)
        st2 = s.graph_stats()
        # the same scan step again: replayed, never re-captured
        assert st2["captures"] == st1["captures"]
        assert st2["replays"] >= st1["replays"] + 3

        # another item count: a new capture, still the eager results
        _same_results(s.fit_pairs(n, m, cw_eager, 0.5 * cw_eager), pairs_eager)
        assert s.graph_stats()["captures"] == st2["captures"] + 1
        _eager_work(s, rng)
        _same_results(s.fit_pairs(n, m, cw_eager, 0.5 * cw_eager), pairs_eager)

        # and back to eager launches
        s.set_graphs(False)
        _same_lines(s.start(cw_eager), lines_eager)


def test_graph_batch_and_warm_fits_equal_eager(inputs):
    import ps_claude_opt_amd as ps
    from ps_claude_opt_amd import synth

    paths = [inputs("tiny")[0],
             inputs("tiny_b", preset="tiny", seed=synth.config_seed(3, 950),
                    n_dense=2, pts_dense=23, n_sparse=37, n_freq=7)[0]]
    rng = np.random.default_rng(5)
    out = {}
    for graphs in (False, True):
        with ps.Scan(0) as s:
            s.prepare_files(paths)
            s.set_graphs(graphs)
            cw = s.conw_search_batch()
            lines = [s.start_batch(cw) for _ in range(2 if graphs else 1)]
            prob = np.repeat([0, 1], 12)
            nl = np.concatenate([np.arange(1, 13) % s.batch_info[p]["n_max"] + 1 for p in (0, 1)])
            ml = np.tile(np.arange(1, 13) % 10 + 1, 2)
            a0 = s.initial_params(prob, nl, ml)
            a0 = a0 * (1.0 + 0.01 * np.random.default_rng(9).standard_normal(a0.shape))
            warm = []
            for _ in range(2 if graphs else 1):
                res, af = s.fit_items_from(prob, nl, ml, cw[prob], cw[prob], a0, want_final=True, n_iter=7)
                warm.append((res, af))
                if graphs:
                    _eager_work_batch(s, rng)
            out[graphs] = (cw, lines, warm)
            if graphs:
                assert s.graph_stats()["replays"] > s.graph_stats()["captures"]
    cw0, lines0, warm0 = out[False]
    cw1, lines1, warm1 = out[True]
    assert np.array_equal(cw0, cw1)
    for L in lines1:
        for a, b in zip(L, lines0[0]):
            _same_lines(a, b)
    for res, af in warm1:
        _same_results(res, warm0[0][0])
        assert np.array_equal(af, warm0[0][1], equal_nan=True)


def _eager_work_batch(s, rng):
    p = rng.integers(0, 2, 5)
    s.warm_features(p, np.ones(5, np.int32), np.arange(1, 6))
