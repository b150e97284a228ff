"""The drop-in command line (ps_claude_opt_amd/lib/period_search) run as a
process on the GPU.

  * on the committed fixture its output passes the reference's acceptance
    rule against the oracle's period_search_out, line for line identical;
  * exit codes: a missing input gives 2, an absent device 999 (the
    reference's exit(999), start_CUDA.cu:216, which a POSIX parent reads
    modulo 256 as 231);
  * checkpoint/resume: a scan stopped after each batch and resumed until
    done writes the same bytes as one uninterrupted run, also when lines
    written after the last checkpoint are left in the output (a crash
    between the two) and when the run resumes from a checkpoint of the
    convexity-weight search;
  * BOINC slot files: the input named through a soft link, the device from
    init_data.xml, boinc_finish_called written at the end.
"""
import os
import shutil
import subprocess

import pytest

from ps_claude_opt_amd import compare_out

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "ps_claude_opt_amd", "lib", "period_search")


def _run(args, cwd, timeout=300):
    p = subprocess.run([BIN] + [str(a) for a in args], cwd=str(cwd), capture_output=True, text=True,
                       timeout=timeout)
    return p.returncode, p.stderr


def test_cli_matches_fixture(golden, tmp_path):
    rc, err = _run([golden("tiny.in"), tmp_path / "out"], tmp_path)
    assert rc == 0, err
    c = compare_out.compare_files(str(tmp_path / "out"), golden("tiny.out"))
    assert c.passes(tolerance=True), c.summary()
    assert c.identical == c.n_ref == 6
    assert (tmp_path / "out").read_bytes() == open(golden("tiny.out"), "rb").read()


def test_cli_exit_codes(golden, tmp_path):
    rc, _ = _run([tmp_path / "missing.in", tmp_path / "out"], tmp_path)
    assert rc == 2
    rc, _ = _run(["--device", "99", golden("tiny.in"), tmp_path / "out"], tmp_path)
    assert rc == 999 % 256


def test_cli_resume_is_byte_identical(golden, tmp_path):
    full = tmp_path / "full"
    full.mkdir()
    rc, err = _run(["--batch", "2", golden("tiny.in"), full / "out"], full)
    assert rc == 0, err
    ref = (full / "out").read_bytes()
    # stopped after every batch of 2 frequencies, resumed until done
    d = tmp_path / "parts"
    d.mkdir()
    states = []
    for _ in range(3):
        rc, err = _run(["--batch", "2", "--max-batches", "1", golden("tiny.in"), d / "out"], d)
        assert rc == 0, err
        states.append((d / "period_search_state").read_text().split())
    assert [s[0] for s in states] == ["2", "4", "6"]
    assert all(s[1] == "1" for s in states)
    assert (d / "out").read_bytes() == ref
    # a crash after lines were written but before the state was: the
    # resumed run cuts them off
    d2 = tmp_path / "crash"
    d2.mkdir()
    rc, err = _run(["--batch", "2", "--max-batches", "1", golden("tiny.in"), d2 / "out"], d2)
    assert rc == 0, err
    with open(d2 / "out", "a") as f:
        f.write("9.99999999  0.999999  9.999999  9.9  999  -99\n")
    rc, err = _run(["--batch", "2", golden("tiny.in"), d2 / "out"], d2)
    assert rc == 0, err
    assert (d2 / "out").read_bytes() == ref
    # resumed from the convexity-weight search's final checkpoint
    conw_r = states[0][2]
    d3 = tmp_path / "conw"
    d3.mkdir()
    (d3 / "period_search_state").write_text(f"0 1 {conw_r} 0\n")
    rc, err = _run(["--batch", "2", golden("tiny.in"), d3 / "out"], d3)
    assert rc == 0, err
    assert (d3 / "out").read_bytes() == ref


def test_cli_boinc_slot(golden, tmp_path):
    slot = tmp_path / "slot"
    slot.mkdir()
    shutil.copy(golden("tiny.in"), tmp_path / "wu_tiny_in")
    (slot / "period_search_in").write_text(f"<soft_link>{tmp_path / 'wu_tiny_in'}</soft_link>\n")
    (slot / "init_data.xml").write_text("<app_init_data>\n<gpu_device_num>0</gpu_device_num>\n"
                                        "<checkpoint_period>60</checkpoint_period>\n</app_init_data>\n")
    rc, err = _run(["--device", "99"], slot)  # init_data.xml's device wins over --device
    assert rc == 0, err
    assert (slot / "period_search_out").read_bytes() == open(golden("tiny.out"), "rb").read()
    assert (slot / "boinc_finish_called").read_text().strip() == "0"
