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


def test_true_frequency_index_on_a_parsed_scan_grid(tmp_path):
    """warm_train's and the screen's true-frequency index, on the grid the
    library itself parses from a period_search_in (host-only ps_input_read;
    the kernels' period of frequency n is 1 / (f_start - (n - 1) f_step)
    hours).  The c1 generator centres its grid on the true frequency, n_freq
    / 2 - 0.3 steps below f_start, so the index is n_freq / 2 + 1."""
    import ctypes as C
    from types import SimpleNamespace

    from ps_claude_opt_amd import _abi, screen
    from ps_claude_opt_amd.warm_train import true_freq_index

    lib = _abi.load()
    cases = [("c1", 1), ("c1", 7), ("small", 3)]
    for preset, seed in cases:
        path = str(tmp_path / f"{preset}_{seed}.in")
        t = synth.write_asteroid(path, preset, seed)
        h = C.c_void_p()
        err = C.create_string_buffer(256)
        assert lib.ps_input_read(path.encode(), C.byref(h), err, 256) == 0, err.value
        view = _abi.PsProblem()
        conw = C.c_double(0.0)
        n_max = C.c_int(0)
        assert lib.ps_input_view(h, C.byref(view), C.byref(conw), C.byref(n_max)) == 0
        b = dict(f_start=view.f_start, f_step=view.f_step, n_max=n_max.value)
        lib.ps_input_free(h)
        n = int(true_freq_index(SimpleNamespace(batch_info=[b]), [t["period_h"]])[0])
        assert n == screen.true_index(b, t["period_h"])
        # nearest grid frequency, in 1/h
        f_n = b["f_start"] - (n - 1) * b["f_step"]
        assert abs(f_n - 1.0 / t["period_h"]) <= 0.5 * b["f_step"] * (1 + 1e-9)
        if preset == "c1":
            assert n_max.value == synth.PRESETS["c1"].n_freq
            assert n == synth.PRESETS["c1"].n_freq // 2 + 1
