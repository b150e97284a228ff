"""The CPU oracle reproduces the committed fixture (regression pin)."""
import os

import pyoracle as po


def test_oracle_reproduces_tiny_fixture(golden, tmp_path):
    O = po.Problem(golden("tiny.in"))
    out = tmp_path / "tiny.out"
    O.run(str(out), threads=2)
    assert out.read_text() == open(golden("tiny.out")).read()
