"""The oracle's host-side tables against closed forms.

Pins trifac/areanorm/sphfunc/ellfit (reference trifac.c, areanorm.c,
sphfunc.c, ellfit.c) without the reference: facet count, a closed
polyhedron, unit normals, area convergence to 4 pi, the associated Legendre
functions against scipy, and the ellipsoid fit's normal equations.
"""
import math

import numpy as np
import pytest
import scipy.special

import pyoracle as po


def _problem(inputs, nrows, lmax, name=None):
    path, _ = inputs(name or f"g{nrows}_{lmax}", "tiny", None, nrows=nrows, lmax=lmax)
    return po.Problem(path)


@pytest.mark.parametrize("nrows", [2, 3, 6, 8])
def test_facets_close_a_polyhedron(inputs, nrows):
    P = _problem(inputs, nrows, 3)
    F = P.info["nfac"]
    assert F == 8 * nrows * nrows
    n = P.table("normal").reshape(3, F)
    a = P.table("d_area")
    assert np.allclose(np.linalg.norm(n, axis=0), 1.0, atol=1e-15)
    # sum of area * outward normal of a closed polyhedron vanishes
    assert np.abs(n @ a).max() < 1e-14
    assert (a > 0).all()
    # inscribed flat facets: total area approaches 4 pi as 1/nrows^2
    assert 0 < 4 * math.pi - a.sum() < 0.8 * 4 * math.pi / nrows ** 2


def test_area_converges_to_sphere(inputs):
    errs = [abs(_problem(inputs, r, 2).table("d_area").sum() - 4 * math.pi) for r in (2, 4, 8)]
    assert errs[0] > errs[1] > errs[2]
    assert errs[2] < 0.8 * 4 * math.pi / 64


@pytest.mark.parametrize("lmax", [3, 6])
def test_legendre_matches_scipy(inputs, lmax):
    P = _problem(inputs, 4, lmax)
    F = P.info["nfac"]
    n = P.table("normal").reshape(3, F)
    x = n[2]
    pleg = P.table("pleg").reshape(lmax + 1, lmax + 1, F)
    for m in range(lmax + 1):
        for l in range(m, lmax + 1):
            # no Condon-Shortley phase in the reference (sphfunc.c:46-47)
            ref = (-1) ** m * scipy.special.lpmv(m, l, x)
            assert np.allclose(pleg[m, l], ref, rtol=1e-12, atol=1e-12 * max(1.0, np.abs(ref).max())), (m, l)


def test_basis_ordering_and_constant_column(inputs):
    lmax = 4
    P = _problem(inputs, 4, lmax)
    F, C = P.info["nfac"], P.info["ncoef"]
    assert C == (lmax + 1) ** 2
    Y = P.table("d_sphere").reshape(C, F)
    assert np.all(Y[0] == 1.0)  # P_0^0 cos(0) == 1
    n = P.table("normal").reshape(3, F)
    th = np.arccos(n[2])
    ph = np.arctan2(n[1], n[0])
    k = 0
    for m in range(lmax + 1):
        for l in range(m, lmax + 1):
            ref = (-1) ** m * scipy.special.lpmv(m, l, np.cos(th))
            assert np.allclose(Y[k], ref * np.cos(m * ph), atol=1e-11 * max(1, np.abs(ref).max()))
            k += 1
            if m:
                assert np.allclose(Y[k], ref * np.sin(m * ph), atol=1e-11 * max(1, np.abs(ref).max()))
                k += 1
    assert k == C


def test_ellipsoid_fit_solves_its_normal_equations(inputs):
    P = _problem(inputs, 6, 6)
    F, C = P.info["nfac"], P.info["ncoef"]
    escl = P.info["escl"]
    a, b, c = 1.05 * escl, 1.00 * escl, 0.95 * escl
    n = P.table("normal").reshape(3, F)
    th = np.arccos(n[2])
    ph = np.arctan2(n[1], n[0])
    st = np.sin(th)
    er = 2 * (np.log(a * b * c) - np.log((a * st * np.cos(ph)) ** 2 + (b * st * np.sin(ph)) ** 2 +
                                          (c * np.cos(th)) ** 2))
    D = P.table("d_sphere").reshape(C, F).T
    cg = P.table("cg_first")
    grad = D.T @ (D @ cg - er)
    assert np.linalg.norm(grad) < 1e-9 * np.linalg.norm(D.T @ er)
    # independent least squares
    ref = np.linalg.lstsq(D, er, rcond=None)[0]
    assert np.allclose(cg, ref, rtol=1e-8, atol=1e-10)


def test_sphere_curvature_special_case(inputs):
    """curv with cg = (c, 0, ...) gives Area = Darea * e^c (Y_1 == 1)."""
    P = _problem(inputs, 4, 3)
    a = P.initial_params(1, 1)
    C = P.info["ncoef"]
    a[:C] = 0.0
    a[0] = 0.37
    _, _, _, area, _ = P.mrqcof(a, 0.0)
    assert np.allclose(area, P.table("d_area") * math.exp(0.37), rtol=1e-15)
