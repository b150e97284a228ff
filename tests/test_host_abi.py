"""Host side of the product: the C ABI library, the input parser and the
geometry tables (CPU only; no device calls).

  * libps_mi355x.so loads and exports every function include/ps_mi355x.h
    declares;
  * ps_input_read builds tables bit-identical to the CPU oracle's
    (period_search_BOINC.cpp:343-866, trifac/areanorm/sphfunc/ellfit);
  * the best-pole rule (start_CUDA.cu:1200-1210) on crafted NaN/tie sets;
  * the output line format (start_CUDA.cu:1222-1226);
  * malformed inputs return PS_E_INPUT instead of exiting.
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

import pyoracle as po
from ps_claude_opt_amd import Line, _abi, format_line

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    lib = _abi.load()
    header = open(os.path.join(ROOT, "include", "ps_mi355x.h")).read()
    names = set(re.findall(r"\b(ps_[a-z0-9_]+)\s*\(", header))
    assert len(names) >= 19
    for n in sorted(names):
        assert hasattr(lib, n), n
    assert names == set(_abi.SIGNATURES), names ^ set(_abi.SIGNATURES)
    # the command line's BOINC shim, in the same library
    boinc = set(re.findall(r"\b(ps_boinc_[a-z0-9_]+)\s*\(", open(os.path.join(ROOT, "include", "ps_boinc.h")).read()))
    assert len(boinc) == 7
    for n in sorted(boinc):
        assert hasattr(lib, n), n


def _read_product(path):
    lib = _abi.load()
    h = C.c_void_p()
    err = C.create_string_buffer(256)
    rc = lib.ps_input_read(path.encode(), C.byref(h), err, 256)
    return rc, h, err.value.decode()


def _view(h):
    lib = _abi.load()
    v = _abi.PsProblem()
    conw = C.c_double()
    nmax = C.c_int()
    assert lib.ps_input_view(h, C.byref(v), C.byref(conw), C.byref(nmax)) == 0
    return v, conw.value, nmax.value


def _arr(ptr, n):
    return np.ctypeslib.as_array(ptr, shape=(n,)).copy()


@pytest.mark.parametrize("name,over", [("tiny", {}), ("small", {}), ("c1", {}), ("tiny", {"lmax": 6, "nrows": 8}),
                                       ("small", {"n_abs": 1})])
def test_product_tables_equal_oracle(inputs, name, over):
    path, _ = inputs(name + "_t" if over else name, name, None, **over)
    rc, h, err = _read_product(path)
    assert rc == 0, err
    try:
        v, conw, nmax = _view(h)
        O = po.Problem(path)
        N, F, Cc = v.ndata, v.nfac, v.ncoef
        assert (N, F, Cc) == (O.info["ndata"], O.info["nfac"], O.info["ncoef"])
        assert nmax == O.info["n_max"]
        assert conw == O.info["conw"]
        assert v.escl == O.info["escl"]
        assert (v.f_start, v.f_end, v.f_step) == (O.info["f_start"], O.info["f_end"], O.info["f_step"])
        assert v.phi0 == O.info["phi0"]
        same = lambda a, b: np.array_equal(a, b)
        assert same(_arr(v.tim, N), O.table("tim"))
        assert same(_arr(v.ee, 3 * N), O.table("ee"))
        assert same(_arr(v.ee0, 3 * N), O.table("ee0"))
        assert same(_arr(v.brightness, N), O.table("brightness"))
        assert same(_arr(v.sig, N), O.table("sig"))
        assert same(_arr(v.sigr2, N), O.table("sigr2"))
        assert same(_arr(v.weight, N), O.table("weight"))
        assert same(_arr(v.normal, 3 * F), O.table("normal"))
        assert same(_arr(v.d_area, F), O.table("d_area"))
        assert same(_arr(v.d_sphere, Cc * F), O.table("d_sphere"))
        assert same(_arr(v.cg_first, Cc), O.table("cg_first"))
        assert same(_arr(v.pleg, (v.mmax + 1) * (v.lmax + 1) * F), O.table("pleg"))
        ma = Cc + 8
        assert same(np.ctypeslib.as_array(v.ia, shape=(ma,)), O.table("ia").astype(np.int32))
    finally:
        _abi.load().ps_input_free(h)


def test_malformed_inputs_return_errors(tmp_path):
    lib = _abi.load()
    p = tmp_path / "bad.in"
    p.write_text("not a header\n")
    rc, h, err = _read_product(str(p))
    assert rc == _abi.PS_E_INPUT and err
    rc, h, err = _read_product(str(tmp_path / "missing.in"))
    assert rc == _abi.PS_E_INPUT
    # truncated data section
    good = open(os.path.join(ROOT, "tests", "golden", "tiny.in")).read().splitlines()
    p.write_text("\n".join(good[:40]) + "\n")
    rc, h, err = _read_product(str(p))
    assert rc == _abi.PS_E_INPUT


def _results(devs, reported=None):
    arr = (_abi.PsPairResult * 10)()
    for i, d in enumerate(devs):
        arr[i].dev = d
        arr[i].reported = 1 if reported is None else reported[i]
    return arr


@pytest.mark.parametrize("devs,expect", [
    ([5, 4, 3, 2, 1, 6, 7, 8, 9, 10], 4),
    ([1, 1, 1, 1, 1, 1, 1, 1, 1, 1], 0),        # ties keep the earliest pole
    ([2, 1, 1, 3, 4, 5, 6, 7, 8, 1], 1),
    ([np.nan, 3, 2, 9, 9, 9, 9, 9, 9, 9], 2),  # a NaN incumbent is replaced
    ([3, np.nan, 2, 9, 9, 9, 9, 9, 9, 9], 2),  # a NaN never wins
    ([np.nan] * 10, 0),
])
def test_best_pole_rule(devs, expect):
    lib = _abi.load()
    assert lib.ps_select_best_pole(_results(devs)) == expect


def test_best_pole_skips_unreported():
    lib = _abi.load()
    rep = [0, 0, 1, 1, 1, 1, 1, 1, 1, 1]
    assert lib.ps_select_best_pole(_results([0.1, 0.1, 5, 4, 9, 9, 9, 9, 9, 9], rep)) == 3
    assert lib.ps_select_best_pole(_results([1] * 10, [0] * 10)) == -1


def test_output_line_format():
    L = Line(12.345678912345, 0.0123456789, 1.5, 2.25, 359.5, -44.5)
    s = format_line(L)
    assert s == "12.34567891  0.012346  1.500000  2.2  360  -45\n"
    assert format_line(Line(3.0, 0.1, 0.2, 0.0, 0.49, -0.49)) == "3.00000000  0.100000  0.200000  0.0    0   -0\n"
