"""Shared fixtures of the test suite.

Markers: ``gpu`` -- needs an MI355X (run with ``-m gpu``); everything else
runs on the CPU.  Inputs are generated on the fly by the product's seeded
generator (ps_claude_opt_amd/synth.py) into a session temp directory, or
read from the committed fixtures in tests/golden/.
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (HIP device 0)")


@pytest.fixture(scope="session")
def inputs(tmp_path_factory):
    """Generated period_search_in files keyed by name."""
    from ps_claude_opt_amd import synth

    d = tmp_path_factory.mktemp("inputs")
    out = {}

    def get(name, preset=None, seed=None, **overrides):
        key = (name, preset, seed, tuple(sorted(overrides.items())))
        if key not in out:
            path = str(d / f"{name}.in")
            truth = synth.write_asteroid(path, preset or name, seed, overrides or None)
            out[key] = (path, truth)
        return out[key]

    return get


@pytest.fixture(scope="session")
def golden():
    return lambda name: os.path.join(GOLDEN, name)
