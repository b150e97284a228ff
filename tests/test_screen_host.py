"""Host logic of the screened and warm scans: alias windows, refit and keep
sets, the screen's input rows and its weight file (no GPU)."""
import numpy as np

from ps_claude_opt_amd import screen, warm

# the config-2 grid: 5,000 frequencies from 1/2 h to 1/40 h
INFO = dict(f_start=0.5, f_step=(0.5 - 1 / 40.0) / 4999.5, n_max=5000)


def _period(n):
    return 1.0 / (INFO["f_start"] - (n - 1) * INFO["f_step"])


def test_alias_windows_hold_the_nearest_alias_frequencies():
    for n in (1200, 3679, 4565):
        got = set(warm.alias_windows(INFO, [n], width=0).tolist())
        for m in warm.ALIAS_MULTIPLES:
            f = 1.0 / (m * _period(n))  # the alias's frequency, 1/h
            if not (INFO["f_start"] - (INFO["n_max"] - 1) * INFO["f_step"] <= f <= INFO["f_start"]):
                continue
            # brute force: the grid frequency nearest to the alias
            grid = INFO["f_start"] - np.arange(INFO["n_max"]) * INFO["f_step"]
            k = int(np.argmin(np.abs(grid - f))) + 1
            assert k in got, (n, m, k)
        # width w adds w frequencies either side (inside the grid)
        wide = warm.alias_windows(INFO, [n], width=2)
        assert set(got) <= set(wide.tolist()) and wide.min() >= 1 and wide.max() <= INFO["n_max"]
    # 2P of P/2 is P: the round trip lands within a step
    n = 3000
    half = warm.alias_windows(INFO, [n], width=0, mults=(0.5,))
    back = warm.alias_windows(INFO, half, width=0, mults=(2.0,))
    assert np.min(np.abs(back - n)) <= 1


def test_refit_frequencies_extend_refit_set_with_alias_windows():
    rng = np.random.default_rng(1)
    rms = rng.uniform(0.02, 0.05, 5000)
    rms[3678] = 0.010  # frequency 3679, the best
    rms[[10, 20]] = np.nan
    base = warm.refit_set(rms)
    got = warm.refit_frequencies(INFO, rms, 1)
    assert set((base + 1).tolist()) <= set(got.tolist())
    al = warm.alias_windows(INFO, [3679])
    assert set(al.tolist()) <= set(got.tolist())
    assert got.min() >= 1 and got.max() <= 5000 and np.all(np.diff(got) > 0)
    assert warm.refit_frequencies(INFO, rms, 1, alias_seeds=0).tolist() == (base + 1).tolist()
    # a sub-range: indices stay inside it
    sub = warm.refit_frequencies(INFO, rms[3000:4000], 3001)
    assert sub.min() >= 3001 and sub.max() <= 4000 and 3679 in set(sub.tolist())
    # restricted to the fitted frequencies (NaN elsewhere)
    part = np.full(5000, np.nan)
    kept = np.arange(3600, 3700)
    part[kept] = rms[kept]
    got2 = warm.refit_frequencies(INFO, part, 1, among=kept + 1)
    assert 3679 in set(got2.tolist()) and set(al.tolist()) <= set(got2.tolist())


def test_keep_set_and_network_inputs():
    s = np.array([0.3, np.nan, -1.0, 0.2, -1.0, 5.0])
    assert screen.keep_set(s, 3).tolist() == [2, 3, 4]
    assert screen.keep_set(s, 6).tolist() == list(range(6))
    rng = np.random.default_rng(2)
    f = rng.normal(size=(400, 32)).astype(np.float32)
    x = screen.network_inputs(f)
    assert x.shape == (400, 32 + screen.N_POWER) and x.dtype == np.float32
    assert np.array_equal(x[:, :32], f)
    assert np.allclose(np.median(x[:, 32:], axis=0), 0.0, atol=1e-5)
    # a frequency far above the grid's typical power stands out
    f[7, 13] += 50.0
    assert screen.network_inputs(f)[7, 32 + 13] > 20.0


def test_screen_weights_round_trip(tmp_path):
    sc = screen.Screen(np.arange(10, dtype=np.float32), np.zeros(56, np.float32), np.ones(56, np.float32), -0.5, 2.0,
                       keep=0.3, keep_min=40, meta={"a": 1})
    p = str(tmp_path / "s.npz")
    sc.save(p)
    r = screen.Screen.load(p)
    assert np.array_equal(r.params, sc.params) and r.keep == 0.3 and r.keep_min == 40 and r.meta == {"a": 1}
    assert r.n_keep(5000) == 1500 and r.n_keep(100) == 40 and r.n_keep(10) == 10


def test_shipped_screen_loads():
    sc = screen.Screen.load()
    assert sc.n_in == 32 + screen.N_POWER and 0 < sc.keep <= 1
    assert sc.params.dtype == np.float32 and np.isfinite(sc.params).all()
