"""Host-side pieces of the drop-in command line, on the CPU.

  * the output comparator (ps_claude_opt_amd/compare_out.py) applies the
    reference's acceptance rule (hip/compare_out.py:17-52): CR stripped,
    equal line counts, dark-facet / lambda / beta equal as strings, worst
    period / rms / chi^2 differences reported; --tolerance adds the
    north-star envelope (1e-5 relative, 5e-4, 0.5);
  * the BOINC slot-directory shim (include/ps_boinc.h): init_data.xml,
    soft links, checkpoint timing, fraction done, boinc_finish_called
    (period_search_BOINC.cpp:102-118, 255-281, 598-624).
"""
import ctypes as C
import os

import pytest

from ps_claude_opt_amd import _abi, compare_out

LINES = ["4.25753696  0.015333  0.028211  0.1  306   63",
         "4.25781153  0.015786  0.029903  0.4  295   54",
         "4.25823631  0.017221  0.035586  1.4   62   57"]


def _write(path, lines, crlf=False):
    with open(path, "wb") as f:
        f.write(("\r\n" if crlf else "\n").join(lines).encode() + (b"\r\n" if crlf else b"\n"))
    return str(path)


def test_compare_identical_and_crlf(tmp_path):
    a = _write(tmp_path / "a", LINES)
    b = _write(tmp_path / "b", LINES, crlf=True)
    c = compare_out.compare_files(a, b)
    assert c.identical == 3 and c.pole_mismatches == 0 and c.passes(tolerance=True)
    assert compare_out.main([a, b, "--tolerance"]) == 0


def test_compare_numeric_differences_reported(tmp_path):
    t = list(LINES)
    t[1] = "4.25781160  0.015886  0.031903  0.4  295   54"
    c = compare_out.compare_files(_write(tmp_path / "t", t), _write(tmp_path / "r", LINES))
    assert c.identical == 2 and c.pole_mismatches == 0
    assert c.worst_period == pytest.approx(7e-8, abs=1e-12)
    assert c.worst_rms == pytest.approx(1e-4, abs=1e-12)
    assert c.worst_chi2 == pytest.approx(0.002, abs=1e-12)
    assert c.passes() and c.passes(tolerance=True)
    # outside the envelope: rms off by 1e-3
    t[1] = "4.25781153  0.016786  0.029903  0.4  295   54"
    c = compare_out.compare_files(_write(tmp_path / "t", t), _write(tmp_path / "r", LINES))
    assert c.passes() and not c.passes(tolerance=True)
    # period off by 2e-5 relative
    t[1] = f"{4.25781153 * (1 + 2e-5):.8f}  0.015786  0.029903  0.4  295   54"
    c = compare_out.compare_files(_write(tmp_path / "t", t), _write(tmp_path / "r", LINES))
    assert c.worst_period_rel > 1e-5 and not c.passes(tolerance=True)


@pytest.mark.parametrize("field,value", [(3, "0.5"), (4, "296"), (5, "53")])
def test_compare_pole_or_dark_mismatch_fails(tmp_path, field, value):
    t = list(LINES)
    f = t[2].split()
    f[field] = value
    t[2] = "  ".join(f)
    r = _write(tmp_path / "r", LINES)
    c = compare_out.compare_files(_write(tmp_path / "t", t), r)
    assert c.pole_mismatches == 1 and not c.passes()
    assert compare_out.main([str(tmp_path / "t"), r]) == 1


def test_compare_line_count_and_malformed(tmp_path):
    r = _write(tmp_path / "r", LINES)
    c = compare_out.compare_files(_write(tmp_path / "t", LINES[:2]), r)
    assert not c.counts_match and not c.passes()
    c = compare_out.compare_files(_write(tmp_path / "t", LINES[:2] + ["4.2 0.1"]), r)
    assert c.pole_mismatches == 1 and not c.passes()


def _boinc():
    lib = _abi.load()

    class InitData(C.Structure):
        _fields_ = [("under_client", C.c_int), ("gpu_device_num", C.c_int), ("checkpoint_period", C.c_double),
                    ("gpu_usage", C.c_double)]

    lib.ps_boinc_read_init_data.argtypes = [C.c_char_p, C.POINTER(InitData)]
    lib.ps_boinc_resolve_filename.argtypes = [C.c_char_p, C.c_char_p, C.c_int]
    lib.ps_boinc_time_to_checkpoint.argtypes = [C.POINTER(InitData)]
    lib.ps_boinc_fraction_done.restype = C.c_double
    lib.ps_boinc_fraction_done.argtypes = [C.c_double]
    lib.ps_boinc_fraction_done_get.restype = C.c_double
    lib.ps_boinc_finish.argtypes = [C.POINTER(InitData), C.c_char_p, C.c_int]
    return lib, InitData


def test_boinc_init_data_and_checkpoints(tmp_path):
    lib, InitData = _boinc()
    d = InitData()
    assert lib.ps_boinc_read_init_data(str(tmp_path).encode(), C.byref(d)) == 0
    assert (d.under_client, d.gpu_device_num, d.checkpoint_period) == (0, -1, 0.0)
    # standalone: a checkpoint after every batch, no finish file
    assert lib.ps_boinc_time_to_checkpoint(C.byref(d)) == 1
    assert lib.ps_boinc_finish(C.byref(d), str(tmp_path).encode(), 0) == 0
    assert not (tmp_path / "boinc_finish_called").exists()
    (tmp_path / "init_data.xml").write_text(
        "<app_init_data>\n<gpu_device_num>3</gpu_device_num>\n<checkpoint_period>3600.000000</checkpoint_period>\n"
        "<gpu_usage>0.500000</gpu_usage>\n</app_init_data>\n")
    assert lib.ps_boinc_read_init_data(str(tmp_path).encode(), C.byref(d)) == 0
    assert (d.under_client, d.gpu_device_num, d.checkpoint_period, d.gpu_usage) == (1, 3, 3600.0, 0.5)
    lib.ps_boinc_checkpoint_completed()
    assert lib.ps_boinc_time_to_checkpoint(C.byref(d)) == 0
    d.checkpoint_period = 0.0
    assert lib.ps_boinc_time_to_checkpoint(C.byref(d)) == 1
    assert lib.ps_boinc_finish(C.byref(d), str(tmp_path).encode(), 7) == 7
    assert (tmp_path / "boinc_finish_called").read_text().strip() == "7"


def test_boinc_soft_links_and_fraction(tmp_path):
    lib, _ = _boinc()
    real = tmp_path / "wu_12345_input"
    real.write_text("data\n")
    link = tmp_path / "period_search_in"
    link.write_text(f"<soft_link>{real}</soft_link>\n")
    buf = C.create_string_buffer(4096)
    assert lib.ps_boinc_resolve_filename(str(link).encode(), buf, 4096) == 0
    assert buf.value.decode() == str(real)
    plain = tmp_path / "plain_in"
    plain.write_text("10 1 20 1\n")
    assert lib.ps_boinc_resolve_filename(str(plain).encode(), buf, 4096) == 0
    assert buf.value.decode() == str(plain)
    assert lib.ps_boinc_resolve_filename(str(tmp_path / "missing").encode(), buf, 4096) == 0
    assert lib.ps_boinc_resolve_filename(str(plain).encode(), buf, 4) == -1
    assert lib.ps_boinc_fraction_done(1.5) == 1.0
    assert lib.ps_boinc_fraction_done(0.25) == 0.25
    assert lib.ps_boinc_fraction_done_get() == 0.25
