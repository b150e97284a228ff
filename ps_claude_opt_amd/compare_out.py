"""Compare two period_search_out files by the reference's acceptance rule.

    python -m ps_claude_opt_amd.compare_out TEST REF [--tolerance]

Both files are read with carriage returns removed.  Lines are compared in
order (one line per trial frequency); each line holds six fields:
period [h], rms, chi^2, dark-facet % (the first line: the convexity
weight), lambda, beta.  The rule (hip/compare_out.py:17-52 of the
reference, hip/check_suite.sh:8-10):

  * the files must have the same number of lines;
  * the dark-facet, lambda and beta fields must be equal as strings on
    every line (a pole or dark mismatch fails);
  * period, rms and chi^2 may differ; their worst absolute differences are
    reported.

With --tolerance the numeric fields must also agree within the north-star
envelope: period within 1e-5 relative, rms within 5e-4, chi^2 within 0.5.
Exit status 0 when the files pass, 1 otherwise.
"""
from __future__ import annotations

import argparse
import dataclasses
import sys
from typing import List, Optional, Tuple

PERIOD_REL_TOL = 1e-5
RMS_TOL = 5e-4
CHI2_TOL = 0.5


def read_lines(path: str) -> List[str]:
    with open(path, "rb") as f:
        data = f.read().decode()
    return [line for line in data.replace("\r", "").split("\n") if line != ""]


@dataclasses.dataclass
class Comparison:
    n_test: int
    n_ref: int
    identical: int = 0
    pole_mismatches: int = 0       # lines whose dark / lambda / beta strings differ (or malformed lines)
    worst_period: float = 0.0      # max |dP| [h]
    worst_period_rel: float = 0.0  # max |dP| / P
    worst_rms: float = 0.0
    worst_chi2: float = 0.0
    differing: List[Tuple[int, str, str, bool]] = dataclasses.field(default_factory=list)

    @property
    def counts_match(self) -> bool:
        return self.n_test == self.n_ref

    def passes(self, tolerance: bool = False) -> bool:
        if not self.counts_match or self.pole_mismatches:
            return False
        if tolerance:
            return (self.worst_period_rel <= PERIOD_REL_TOL and self.worst_rms <= RMS_TOL
                    and self.worst_chi2 <= CHI2_TOL)
        return True

    def summary(self) -> str:
        if not self.counts_match:
            return f"line count mismatch: {self.n_test} vs {self.n_ref}"
        return (f"{self.n_test} lines: {self.identical} identical, {self.n_test - self.identical} differ; "
                f"worst |dP| = {self.worst_period:.3g} h ({self.worst_period_rel:.3g} relative), "
                f"|dRMS| = {self.worst_rms:.3g}, |dchi2| = {self.worst_chi2:.3g}; "
                f"pole / dark mismatches: {self.pole_mismatches}")


def compare_lines(test: List[str], ref: List[str]) -> Comparison:
    c = Comparison(len(test), len(ref))
    if not c.counts_match:
        return c
    for i, (a, b) in enumerate(zip(test, ref), start=1):
        if a == b:
            c.identical += 1
            continue
        fa, fb = a.split(), b.split()
        if len(fa) != 6 or len(fb) != 6:
            c.pole_mismatches += 1
            c.differing.append((i, a, b, True))
            continue
        bad = fa[3:] != fb[3:]
        c.pole_mismatches += bad
        c.differing.append((i, a, b, bad))
        p_t, p_r = float(fa[0]), float(fb[0])
        c.worst_period = max(c.worst_period, abs(p_t - p_r))
        if p_r != 0.0:
            c.worst_period_rel = max(c.worst_period_rel, abs(p_t - p_r) / abs(p_r))
        c.worst_rms = max(c.worst_rms, abs(float(fa[1]) - float(fb[1])))
        c.worst_chi2 = max(c.worst_chi2, abs(float(fa[2]) - float(fb[2])))
    return c


def compare_files(test_path: str, ref_path: str) -> Comparison:
    return compare_lines(read_lines(test_path), read_lines(ref_path))


def main(argv: Optional[List[str]] = None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("test")
    ap.add_argument("ref")
    ap.add_argument("--tolerance", action="store_true",
                    help="also require period 1e-5 relative, rms 5e-4 and chi2 0.5")
    ap.add_argument("--show", type=int, default=12, help="differing lines to print")
    a = ap.parse_args(argv)
    c = compare_files(a.test, a.ref)
    print(c.summary())
    for ln, x, y, bad in c.differing[:a.show]:
        print(f"  line {ln}{' POLE/DARK' if bad else ''}:\n    test {x}\n    ref  {y}")
    if len(c.differing) > a.show:
        print(f"  ... {len(c.differing) - a.show} more differing lines")
    return 0 if c.passes(a.tolerance) else 1


if __name__ == "__main__":
    sys.exit(main())
