"""The shipped warm-start weights: one network per shape degree, each with
the feature width of ps_warm_features and the predicted-parameter count of
its lmax (the free shape coefficients and the four phase-function and
Lambert parameters)."""
import numpy as np
import pytest

from ps_claude_opt_amd import synth, warm


@pytest.mark.parametrize("lmax", [6, 8])
def test_shipped_weights_shapes(lmax):
    ws = warm.WarmStart.load(warm.weights_for(lmax))
    assert ws.n_in == 528
    assert ws.n_out == (lmax + 1) ** 2 - 1 + 4
    for a in (ws.params, ws.x_mean, ws.x_std, ws.y_mean, ws.y_std, ws.y_lo, ws.y_hi, ws.shrink):
        assert np.isfinite(a).all()
    assert (ws.x_std > 0).all() and (ws.y_std > 0).all()
    assert ((ws.shrink >= 0) & (ws.shrink <= 1)).all()
    assert (ws.y_lo <= ws.y_hi).all()


def test_default_weights_are_the_lmax6_ones():
    assert warm.DEFAULT_WEIGHTS == warm.weights_for(6)
    assert synth.PRESETS["c2"].lmax == 6 and synth.PRESETS["c5s"].lmax == 8 == synth.PRESETS["c5"].lmax


def test_true_frequency_index_inverts_the_scan_grid():
    """warm_train's near-truth sampling: the scan's n-th period is
    1 / (f_start - (n - 1) f_step) days, and the true period maps back to
    its nearest n."""
    from types import SimpleNamespace

    from ps_claude_opt_amd.warm_train import true_freq_index

    bi = [dict(f_start=12.0, f_step=0.002), dict(f_start=6.5, f_step=0.0005)]
    n_true = [1234, 77]
    periods_h = [24.0 / (b["f_start"] - (n - 1) * b["f_step"]) for b, n in zip(bi, n_true)]
    got = true_freq_index(SimpleNamespace(batch_info=bi), periods_h)
    assert list(got) == n_true
    # a period between two grid points maps to the nearer one
    b = bi[0]
    f_mid = b["f_start"] - (1234 - 1 + 0.3) * b["f_step"]
    assert true_freq_index(SimpleNamespace(batch_info=[b]), [24.0 / f_mid])[0] == 1234
