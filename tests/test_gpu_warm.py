"""The warm-start network kernels and the warm-started fits.

Network: the GPU forward pass (FP32 and BF16 operands), the BF16 gradients
of the mean-squared error and one Adam step are checked against the FP64
restatement in oracle/mlp_ref.py (itself pinned by finite differences), with
tolerances for the operand precision; training must reduce the loss.
Fits: a warm fit started from the reference's own initial vector is
bit-identical to the cold fit, the returned final parameters are the ones
the results report, the features equal an independent numpy computation,
and the training pipeline generalises to held-out asteroids.
"""
import numpy as np
import pytest

import mlp_ref
import pyoracle as po

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def scan(inputs):
    import ps_claude_opt_amd as ps

    s = ps.Scan(0)
    s.prepare_file(inputs("tiny")[0])
    yield s
    s.close()


def _net(scan, n_in=40, n_hidden=32, n_layers=3, n_out=12, seed=3):
    from ps_claude_opt_amd import warm

    return warm.Mlp(scan, n_in, n_out, n_hidden, n_layers, seed=seed)


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


def test_forward_fp32_and_bf16_match_fp64(scan):
    net = _net(scan)
    ref = mlp_ref.Mlp(40, 32, 3, 12)
    p = net.get_params().astype(np.float64)
    x = np.random.default_rng(0).normal(size=(70, 40)).astype(np.float32)
    y_ref, _, _ = ref.forward(p, x.astype(np.float64))
    assert _rel(net.forward(x, bf16=False), y_ref) < 1e-5
    e16 = _rel(net.forward(x, bf16=True), y_ref)
    assert 1e-5 < e16 < 2e-2  # BF16 operands: visibly rounded, still close
    net.close()


def test_bf16_gradients_match_fp64(scan):
    net = _net(scan)
    ref = mlp_ref.Mlp(40, 32, 3, 12)
    p = net.get_params().astype(np.float64)
    rng = np.random.default_rng(1)
    x = rng.normal(size=(96, 40)).astype(np.float32)
    t = rng.normal(size=(96, 12)).astype(np.float32)
    loss, g = net.loss_grad(x, t)
    l_ref, g_ref = ref.loss_grad(p, x.astype(np.float64), t.astype(np.float64))
    assert abs(loss - l_ref) < 2e-2 * l_ref
    # per layer: weights and biases
    o = 0
    for k, n in zip(ref.dims[:-1], ref.dims[1:]):
        for size in (k * n, n):
            assert _rel(g[o:o + size], g_ref[o:o + size]) < 5e-2, (k, n, size)
            o += size
    net.close()


def test_shipped_network_shape_forward_and_gradients(scan):
    """The shipped lmax-6 network's shape, 528 -> 256 x 3 -> 52, with its
    trained parameters: forward in FP32 and BF16 and the BF16 gradients
    against the FP64 restatement, on standardised features of real
    (frequency, pole) items."""
    from ps_claude_opt_amd import warm

    ws = warm.WarmStart.load(warm.DEFAULT_WEIGHTS)
    assert (ws.n_in, ws.n_hidden, ws.n_layers, ws.n_out) == (528, 256, 3, 52)
    net = ws.network(scan)
    ref = mlp_ref.Mlp(528, 256, 3, 52)
    p = net.get_params().astype(np.float64)
    assert np.array_equal(net.get_params(), ws.params)
    rng = np.random.default_rng(4)
    x = rng.normal(size=(96, 528)).astype(np.float32)
    y_ref, _, _ = ref.forward(p, x.astype(np.float64))
    assert _rel(net.forward(x, bf16=False), y_ref) < 1e-5
    assert _rel(net.forward(x, bf16=True), y_ref) < 3e-2
    t = rng.normal(size=(96, 52)).astype(np.float32)
    loss, g = net.loss_grad(x, t)
    l_ref, g_ref = ref.loss_grad(p, x.astype(np.float64), t.astype(np.float64))
    assert abs(loss - l_ref) < 2e-2 * l_ref
    o = 0
    for k, n in zip(ref.dims[:-1], ref.dims[1:]):
        for size in (k * n, n):
            assert _rel(g[o:o + size], g_ref[o:o + size]) < 5e-2, (k, n, size)
            o += size
    net.close()


def test_first_adam_step_and_training_reduce_loss(scan):
    net = _net(scan, seed=5)
    ref = mlp_ref.Mlp(40, 32, 3, 12)
    p0 = net.get_params().astype(np.float64)
    rng = np.random.default_rng(2)
    x = rng.normal(size=(64, 40)).astype(np.float32)
    t = rng.normal(size=(64, 12)).astype(np.float32)
    _, g_ref = ref.loss_grad(p0, x.astype(np.float64), t.astype(np.float64))
    net.train(x, t, epochs=1, batch=64, lr=1e-3, seed=0)  # one full-batch step
    step = net.get_params().astype(np.float64) - p0
    # a first Adam step moves a parameter by -lr * sign(gradient)
    big = np.abs(g_ref) > 0.05 * np.abs(g_ref).max()
    assert np.allclose(step[big], -1e-3 * np.sign(g_ref[big]), atol=2e-6)
    # a smooth target of the inputs is learnt
    teacher = mlp_ref.Mlp(40, 16, 1, 12)
    tp = rng.normal(0, 0.3, teacher.n_params())
    xs = rng.normal(size=(2048, 40)).astype(np.float32)
    ys = teacher.forward(tp, xs.astype(np.float64))[0].astype(np.float32)
    net2 = _net(scan, seed=6)
    losses = net2.train(xs, ys, epochs=40, batch=128, lr=2e-3, seed=1)
    assert losses[-1] < 0.1 * losses[0]
    net.close()
    net2.close()


def test_fit_from_reference_initialisation_is_bit_identical(scan):
    n = np.array([1, 2, 3, 4, 5, 6] * 2)
    m = np.arange(1, 13) % 10 + 1
    p = np.zeros(12, np.int32)
    cold = scan.fit_pairs(n, m, 0.5, 0.5)
    a0 = scan.initial_params(p, n, m)
    warm, _ = scan.fit_items_from(p, n, m, 0.5, 0.5, a0)
    for f in ("dev", "dark", "period", "lambda_deg", "beta_deg", "chisq", "niter", "n_vis"):
        assert np.array_equal(cold[f], warm[f], equal_nan=True), f


def test_warm_fits_bit_identical_to_oracle(scan, inputs):
    """Fits from perturbed starting parameters, in the problem's own stop
    mode and with a fixed iteration count, equal the oracle's bit for bit
    (final parameters included)."""
    path, _ = inputs("tiny")
    O = po.Problem(path)
    n = np.array([1, 2, 3, 4, 5, 6, 2, 5])
    m = np.array([1, 3, 5, 7, 9, 10, 2, 8])
    p = np.zeros(8, np.int32)
    a0 = scan.initial_params(p, n, m)
    C = scan.info["ncoef"]
    rng = np.random.default_rng(4)
    a0[:, 1:C] += rng.normal(0, 0.05, (8, C - 1))
    a0[:, C + 3:C + 6] *= 1 + rng.normal(0, 0.05, (8, 3))
    for n_iter in (0, 7):
        g, ga = scan.fit_items_from(p, n, m, 0.5, 0.5, a0, want_final=True, n_iter=n_iter)
        o, oa = O.fit_pairs_from(n, m, 0.5, 0.5, a0, threads=8, want_final=True, n_iter=n_iter)
        for i in range(8):
            for f in ("dev", "dark", "period", "lambda_deg", "beta_deg", "chisq", "niter", "n_vis"):
                assert g[i][f] == float(getattr(o[i], f)) or (np.isnan(g[i][f]) and np.isnan(getattr(o[i], f))), (
                    n_iter, i, f)
        assert np.array_equal(ga, oa, equal_nan=True)
        if n_iter:
            assert (g["niter"] == n_iter).all()


def test_final_parameters_match_reported_pole_and_period(scan):
    n = np.array([1, 3, 5, 2, 4, 6])
    m = np.array([1, 4, 7, 9, 2, 10])
    p = np.zeros(6, np.int32)
    r, a = scan.fit_items_from(p, n, m, 0.5, 0.5, None, want_final=True)
    C = scan.info["ncoef"]
    assert np.array_equal(r["period"], (2 * np.pi) / a[:, C + 2])
    lt = np.degrees(a[:, C + 1])
    assert np.allclose(r["lambda_deg"], lt + np.where(lt < 0, 360, 0), rtol=0, atol=1e-12)
    assert np.allclose(r["beta_deg"], 90 - np.degrees(a[:, C]), rtol=0, atol=1e-12)


def test_features_match_numpy(scan, inputs):
    path, _ = inputs("tiny")
    O = po.Problem(path)
    N = O.info["ndata"]
    tim, br = O.table("tim"), O.table("brightness")
    ee, ee0 = O.table("ee").reshape(3, N), O.table("ee0").reshape(3, N)
    lp = O.table("l_points").astype(int)
    n = np.array([1, 4, 6])
    m = np.array([2, 5, 10])
    f = scan.warm_features(np.zeros(3, np.int32), n, m)
    assert f.shape == (3, 528)
    off = np.concatenate([[0], np.cumsum(lp)])
    curves = [c for c in range(len(lp)) if lp[c] >= 16 and tim[off[c + 1] - 1] - tim[off[c]] <= 2.0][:8]
    sel = np.concatenate([np.arange(off[c], off[c + 1]) for c in curves])
    for k in range(3):
        P = 1.0 / (O.info["f_start"] - (n[k] - 1) * O.info["f_step"])
        exp = np.zeros(528)
        for ci, c in enumerate(curves):
            j = np.arange(off[c], off[c + 1])
            ph = tim[j] * 24.0 / P
            ph -= np.floor(ph)
            b = np.clip((ph * 64).astype(int), 0, 63)
            rel = br[j] / br[j].mean() - 1.0
            for q in range(64):
                if (b == q).any():
                    exp[64 * ci + q] = rel[b == q].mean()
        lam = [0, 90, 180, 270, 60, 180, 300, 60, 180, 300][m[k] - 1]
        bet = [0, 0, 0, 0, 60, 60, 60, -60, -60, -60][m[k] - 1]
        lr, br_ = np.radians(lam), np.radians(bet)
        exp[512:515] = [np.cos(br_) * np.cos(lr), np.cos(br_) * np.sin(lr), np.sin(br_)]
        exp[515:518] = ee[:, sel].mean(1)
        exp[518:521] = ee0[:, sel].mean(1)
        exp[521] = np.arccos(np.clip((ee[:, sel] * ee0[:, sel]).sum(0), -1, 1)).mean()
        exp[522] = np.log(P / 10.0)
        assert np.allclose(f[k], exp, rtol=1e-5, atol=1e-6), np.abs(f[k] - exp).max()


def test_training_pipeline_generalises(tmp_path):
    """Train on cold fits of 24 small seeded asteroids (lmax 3): training
    lowers the loss, and on held-out asteroids fits started from the
    (shrunk, clipped) prediction stay finite and end about as close to the
    full cold fit as cold-started fits of the same length.  At this size the
    features carry little about a held-out asteroid; the gain at the
    benchmark's scale is measured by warm_train and bench.py."""
    import ps_claude_opt_amd as ps
    from ps_claude_opt_amd import synth, warm

    ov = dict(n_dense=3, pts_dense=40, n_sparse=40, sparse_group=40, lmax=3, nrows=3, n_freq=40)
    paths = []
    for i in range(28):
        p = str(tmp_path / f"w{i}.in")
        synth.write_asteroid(p, "tiny", synth.config_seed(4, 500 + i), ov)
        paths.append(p)
    with ps.Scan(0) as s:
        s.prepare_files(paths)
        cw = np.full(28, 0.3)
        feats, tg, _, prob = warm.make_training_set(s, 80, cw, seed=1)
        train = prob < 24
        ws, losses = warm.train_warm_start(s, feats[train], tg[train], prob[train], epochs=80, batch=128, lr=2e-3)
        assert losses[-1] < losses[0]
        assert np.isfinite(ws.meta["val_loss"]) and np.isfinite(ws.params).all()
        rng = np.random.default_rng(9)
        pr = np.repeat(np.arange(24, 28), 40)
        nl = rng.integers(1, 41, 160)
        ml = rng.integers(1, 11, 160)
        net = ws.network(s)
        full, _ = s.fit_items_from(pr, nl, ml, 0.3, 0.3, None)
        few_cold, _ = s.fit_items_from(pr, nl, ml, 0.3, 0.3, None, n_iter=6)
        a0 = ws.initial_params(s, net, pr, nl, ml)
        few_warm, _ = s.fit_items_from(pr, nl, ml, 0.3, 0.3, a0, n_iter=6)
        net.close()
    ok = np.isfinite(full["dev"])
    assert np.isfinite(few_warm["dev"][ok]).mean() >= np.isfinite(few_cold["dev"][ok]).mean() - 0.02
    gap_cold = np.nanmedian(np.abs(few_cold["dev"] - full["dev"])[ok])
    gap_warm = np.nanmedian(np.abs(few_warm["dev"] - full["dev"])[ok])
    assert gap_warm < 1.25 * gap_cold, (gap_warm, gap_cold)


def test_warm_scan_refits_are_plain_fits(inputs):
    """A warm scan's refitted frequencies carry the plain fits bit for bit,
    and the warm winner (the best refit) is the plain winner whenever the
    plain winner's frequency was refitted."""
    import ps_claude_opt_amd as ps
    from ps_claude_opt_amd import warm

    path, _ = inputs("c1")
    with ps.Scan(0) as s:
        s.prepare_file(path)
        n_max = min(s.info["n_max"], 80)
        cw = 0.5
        n = np.repeat(np.arange(1, n_max + 1), 10)
        m = np.tile(np.arange(1, 11), n_max)
        p = np.zeros(n.size, np.int32)
        plain, _ = s.fit_items_from(p, n, m, cw, cw, None)
        ws = warm.WarmStart.load(warm.DEFAULT_WEIGHTS)
        net = ws.network(s)
        res, fr = warm.warm_scan(s, ws, net, cw, 1, n_max, n_iter=4, refit=12, band=0.0, alias_seeds=0)
        net.close()
    assert fr.size == 12
    rows = ((fr - 1)[:, None] * 10 + np.arange(10)).ravel()
    for f in ("dev", "dark", "period", "lambda_deg", "beta_deg", "chisq", "niter"):
        assert np.array_equal(res[f][rows], plain[f][rows], equal_nan=True), f
    p_best = warm.best_line(plain, 1, n_max)
    w_best = warm.best_line(res, 1, n_max, among=fr)
    if p_best["freq_index"] in set(fr.tolist()):
        assert w_best == p_best
    else:
        assert w_best["rms"] >= p_best["rms"]


def test_warm_scan_config2_first_500_frequencies(inputs):
    """The warm scan on a config-2 workload (the c2 preset's asteroid, its
    first 500 of 5,000 frequencies, 10 poles each) with the shipped lmax-6
    network: the refitted frequencies carry the plain fits bit for bit, and
    the scan reports the plain winner."""
    import ps_claude_opt_amd as ps
    from ps_claude_opt_amd import warm

    path, _ = inputs("c2")
    with ps.Scan(0) as s:
        s.prepare_file(path)
        assert s.info["ndata"] == 2700 and s.info["n_max"] == 5000
        cw = s.conw_search()
        n_max = 500
        n = np.repeat(np.arange(1, n_max + 1), 10)
        m = np.tile(np.arange(1, 11), n_max)
        plain, _ = s.fit_items_from(np.zeros(n.size, np.int32), n, m, cw, cw, None)
        ws = warm.WarmStart.load(warm.DEFAULT_WEIGHTS)
        net = ws.network(s)
        res, fr = warm.warm_scan(s, ws, net, cw, 1, n_max, n_iter=20)
        net.close()
    rows = ((fr - 1)[:, None] * 10 + np.arange(10)).ravel()
    for f in ("dev", "dark", "period", "lambda_deg", "beta_deg", "chisq", "niter"):
        assert np.array_equal(res[f][rows], plain[f][rows], equal_nan=True), f
    assert warm.best_line(res, 1, n_max, among=fr) == warm.best_line(plain, 1, n_max)


def test_lmax8_network_warm_fits_bit_identical_to_oracle(inputs):
    """The shipped lmax-8 network (config 5's 84 predicted parameters) runs
    on a config-5-shaped asteroid, and fits from its prediction equal the
    oracle's bit for bit, final parameters included."""
    import ps_claude_opt_amd as ps
    from ps_claude_opt_amd import warm

    path, _ = inputs("c5_small", "c5", None, n_sparse=2000)
    O = po.Problem(path)
    n = np.array([1, 700, 9000, 19999])
    m = np.array([1, 5, 8, 10])
    p = np.zeros(4, np.int32)
    ws = warm.WarmStart.load(warm.weights_for(8))
    with ps.Scan(0) as s:
        s.prepare_file(path)
        assert s.info["ncoef"] == 81 and ws.n_out == 84
        net = ws.network(s)
        a0 = ws.initial_params(s, net, p, n, m)
        net.close()
        cold = s.initial_params(p, n, m)
        g, ga = s.fit_items_from(p, n, m, 0.8, 0.8, a0, want_final=True, n_iter=5)
    assert np.isfinite(a0).all() and not np.array_equal(a0, cold)
    o, oa = O.fit_pairs_from(n, m, 0.8, 0.8, a0, threads=8, want_final=True, n_iter=5)
    for i in range(4):
        for f in ("dev", "dark", "period", "lambda_deg", "beta_deg", "chisq", "niter", "n_vis"):
            assert g[i][f] == float(getattr(o[i], f)) or (np.isnan(g[i][f]) and np.isnan(getattr(o[i], f))), (i, f)
    assert np.array_equal(ga, oa, equal_nan=True)
