"""Checks of the oracle's NFVSE shock capturing that do not need the reference's build:
closed forms of the Persson-Peraire indicator, free-stream preservation of the pure
sub-cell finite-volume scheme on warped meshes (which the sub-cell metrics must satisfy),
conservation, and the blending-coefficient formula and propagation."""
import math

import numpy as np
import pytest

import pyoracle as O


def pulse(m, sigma=0.15, p0=1e-2):
    x = m["Elem_xGP"]
    r2 = (x ** 2).sum(-1)
    p = p0 + 0.4 * np.exp(-r2 / (2 * sigma ** 2)) / ((2 * math.pi) ** 1.5 * sigma ** 3)
    U = np.zeros(x.shape[:-1] + (5,))
    U[..., 0] = 1.0
    U[..., 4] = p / 0.4
    return U


def legendre(k, x):
    return np.polynomial.legendre.legval(x, [0] * k + [1])


@pytest.mark.parametrize("N", [2, 3, 4, 5])
def test_indicator_closed_forms(N):
    """rho*p equal to a pure degree-N Legendre mode in xi gives eta = 1; a polynomial of
    degree <= N-2 in each direction gives the epsilon floor."""
    m = O.box_mesh(1, N, x0=(-1, -1, -1), length=(2, 2, 2))
    x = m["Elem_xGP"]
    P = O.Problem(m, nfvse=True)
    U = np.zeros(x.shape[:-1] + (5,))
    U[..., 0] = 1.0
    U[..., 4] = legendre(N, x[..., 0]) / 0.4          # p = P_N(xi), rho = 1
    assert P.indicator(U)[0] == pytest.approx(1.0, rel=1e-12)
    U[..., 4] = (2.0 + x[..., 0] * x[..., 1]) / 0.4 if N >= 3 else 2.0 / 0.4
    assert P.indicator(U)[0] == np.finfo(float).eps


@pytest.mark.parametrize("N", [2, 3, 4])
def test_fv_free_stream_warped(N):
    """With alpha = 1 (pure sub-cell finite volumes) a constant state stays constant on a
    curved mesh to round-off: the sub-cell face metrics telescope to the DG metrics."""
    m = O.box_mesh(3, N, Ngeo=3, warp=True)
    U = np.empty(m["Elem_xGP"].shape[:-1] + (5,))
    U[...] = [1.2, 0.3, -0.2, 0.5, 3.0]
    P = O.Problem(m, nfvse=True, alpha_max=1.0, fixed_alpha=1.0)
    Ut = P.rhs(U)
    assert np.all(P.alpha == 1.0)
    assert np.abs(Ut).max() < 1e-11


def test_conservation_with_blending():
    """Periodic mesh: the integral of the RHS vanishes with mixed blending coefficients."""
    m = O.box_mesh(5, 3, x0=(-1.5,) * 3, length=(3.0,) * 3, Ngeo=3, warp=True)
    U = pulse(m)
    P = O.Problem(m, nfvse=True)
    Ut = P.rhs(U)
    a = P.alpha
    assert (a == 0).any() and (a == 0.5).any() and ((a > 0) & (a < 0.5)).any()
    wl = O.dg_operators(3)["w"]
    W = np.einsum("i,j,k->kji", wl, wl, wl)
    tot = np.einsum("kji,ekjiv->v", W, Ut / m["sJ"][..., None])
    scale = np.einsum("kji,ekjiv->v", W, np.abs(Ut) / m["sJ"][..., None])
    assert np.all(np.abs(tot) <= 1e-13 * scale)


def test_blending_formula_and_propagation():
    """alpha = logistic(eta) clamped to [alpha_min -> 0, alpha_max], then raised to
    SpacePropFactor * max over face neighbours, from the indicator the oracle reports."""
    m = O.box_mesh(5, 3, x0=(-1.5,) * 3, length=(3.0,) * 3, Ngeo=3, warp=True)
    U = pulse(m)
    P0 = O.Problem(m, nfvse=True, sweeps=0)
    P1 = O.Problem(m, nfvse=True, sweeps=1)
    eta = P0.indicator(U)
    T = O.nfvse_threshold(3)
    s = math.log((1 - 1e-4) / 1e-4)
    raw = 1.0 / (1.0 + np.exp(-s * (eta - T) / T))
    raw = np.where(raw < 0.01, 0.0, np.where(raw >= 0.5, 0.5, raw))
    assert np.array_equal(P0.blending(U), raw)
    s2e = m["SideToElem"]
    prop = raw.copy()
    for side in s2e:
        em, es = side[0] - 1, side[1] - 1
        mx = 0.5 * max(raw[em], raw[es])
        prop[em] = max(prop[em], mx)
        prop[es] = max(prop[es], mx)
    assert np.array_equal(P1.blending(U), prop)
    assert (prop > raw).any()
