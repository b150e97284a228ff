"""Regenerates the committed fixtures of tests/golden.

  tiny.in   seeded synthetic asteroid (ps_claude_opt_amd/synth.py, preset
            "tiny": 2 dense curves x 40 points + 40 sparse points, lmax 3,
            nrows 3, 6 trial frequencies)
  tiny.out  period_search_out of the CPU FP64 oracle for tiny.in
  tiny.truth.json  the generator's truth for tiny.in (pole, period, shape)

No reference-held golden vectors exist for this path (the reference repo
ships no tests or fixtures), so these pin the oracle and the kernels against
regressions; their correctness rests on the closed-form and
finite-difference tests of the oracle.

    python tests/golden/make_fixtures.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from ps_claude_opt_amd import synth  # noqa: E402
import pyoracle  # noqa: E402


def main():
    path = os.path.join(HERE, "tiny.in")
    truth = synth.write_asteroid(path, "tiny", synth.SEED_BASE)
    with open(os.path.join(HERE, "tiny.truth.json"), "w") as f:
        json.dump(truth, f, indent=1)
        f.write("\n")
    O = pyoracle.Problem(path)
    O.run(os.path.join(HERE, "tiny.out"), threads=4)


if __name__ == "__main__":
    main()
