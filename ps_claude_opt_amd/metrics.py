"""Algorithmic work model of one LM pair-fit (FP64 flops, FMA = 2).

Counts the arithmetic the period-scan algorithm needs, not what the GPU
executes: invisible (point, facet) pairs cost only the visibility test, the
shape gather counts only visible facets, and zero padding of the 16x16 MFMA
tiles is not counted.  Transcendentals (exp, exp2, sin/cos) are reported
separately and not included in the flop total.  Per mrqcof evaluation with N
data points, F facets, C shape coefficients, M free parameters, K = M + 1,
and V visible (point, facet) pairs:

  visibility           10 N F           two 3-term dot products
  visible facets       V (32 + 2 (C-1)) scattering law with its derivative
                                        sums (regrouped: 1 / (mu + mu0), its
                                        square and the products once), the
                                        shape-Jacobian gather
  per point            N (91 + (C-1) + 3M + 5 + K(K+1) + K)
                                        geometry, derivative columns, scaling,
                                        renormalisation, [J'|r]^T W [J'|r]
  curve sums           N (M + 1)
  facet areas          F (2C + 1)
  convexity rows       3 F (2C + 1)

and per LM iteration (one Gauss-Jordan solve + Iter2):

  Gauss-Jordan         2 M^2 (M - 1) + 2 M^2
  Iter2                F (2C + 1) + 6 F + 12
"""
from __future__ import annotations


def flops_per_eval(N: int, F: int, C: int, M: int, V: float) -> float:
    K = M + 1
    per_point = 91 + (C - 1) + 3 * M + 5 + K * (K + 1) + K
    return (10.0 * N * F + V * (32 + 2 * (C - 1)) + N * per_point + N * (M + 1)
            + F * (2 * C + 1) + 3 * F * (2 * C + 1))


def flops_per_iteration(F: int, C: int, M: int) -> float:
    return 2.0 * M * M * (M - 1) + 2.0 * M * M + F * (2 * C + 1) + 6 * F + 12


def flops_of_fits(N: int, F: int, C: int, M: int, total_vis: float, total_evals: float,
                  total_iters: float) -> float:
    """Total algorithmic flops of a set of pair-fits given their summed
    visible-pair count, evaluation count and iteration count."""
    V_mean = total_vis / max(total_evals, 1.0)
    return total_evals * flops_per_eval(N, F, C, M, V_mean) + total_iters * flops_per_iteration(F, C, M)


def transcendentals_per_eval(N: int, F: int) -> float:
    """exp per facet (areas) + sin/cos and exp2 per point."""
    return F + 2.0 * N
