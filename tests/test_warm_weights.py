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
