"""The oracle's restatement of the reference's analysis (CalcErrorNorms, CalcBulk, the
Taylor-Green vortex testcase analysis) against the values the reference's own Analyze
computed for the same states (tests/golden/analyze, written by tests/golden/make_golden.py
through oracle/_ref), and against closed forms."""
import math

import numpy as np
import pytest

import pyoracle as O
from analysis_ref import NAMES, check_against_reference, load, mesh_of


def oracle_analysis(g, m):
    """Ut and gradients from the oracle's own RHS/lifting, then its analysis."""
    U = g["U"]
    t = float(g["t0"])
    if str(g["eq"]) == "ns":
        par = bool(int(g["parabolic"]))
        P = O.Problem(m, volflux=32, riemann=1, parabolic=par, mu0=float(g["mu0"]), Pr=float(g["Pr"]))
        Ut = P.rhs(U)
        grads = P.lifting(U) if par else None
        cfg = dict(exact_func=int(g["exact"]), n_analyze=int(g["nanalyze"]),
                   ref_state=(1.0, 0.0, 0.0, 0.0, 1.0 / 0.4), diss_factor=1.0)
        return O.analyze(m, U, Ut, grads, kappa=1.4, t=t, **cfg), Ut
    res = dict(mu=float(g["mu"]), eta=float(g["eta"]), Pr=float(g["Pr"]))
    P = O.MhdProblem(m, resistive=res)
    Ut = P.rhs(U, glm_ch=0.5)
    cfg = dict(exact_func=1, ref_state=(1.0, 0, 0, 0, 1.5, 0, 0, 0, 0), diss_factor=float(g["mu_0"]))
    return O.analyze(m, U, Ut, P.lifting(U), kappa=5.0 / 3.0, mu_0=float(g["mu_0"]), t=t, **cfg), Ut


@pytest.mark.parametrize("name", NAMES)
def test_oracle_analysis_matches_reference(name):
    g = load(name)
    m = mesh_of(g)
    a, Ut = oracle_analysis(g, m)
    assert np.linalg.norm((Ut - g["Ut"]).ravel()) <= 1e-13 * np.linalg.norm(g["Ut"].ravel())
    worst = check_against_reference(a, g)
    print(f"{name}: worst relative difference to the reference's Analyze {worst:.2e}")


def test_volume_bulk_and_exact_state():
    """Cartesian box [0,2]^3: Vol = 8; the exact density wave at the collocation points has zero
    L2_colloc error and an analyze-point L2 error that converges spectrally in N; the bulk of a
    constant state is Vol times the state and its error norms against itself vanish."""
    errs = []
    for N in (2, 3, 4, 5):
        m = O.box_mesh(2, N, x0=(0, 0, 0), length=(2, 2, 2), Ngeo=1, warp=False)
        U = O.exactfunc(2, m["Elem_xGP"], t=0.3)
        a = O.analyze(m, U, np.zeros_like(U), exact_func=2, t=0.3)
        assert a["vol"] == pytest.approx(8.0, rel=1e-14)
        assert np.all(a["L2colloc"] == 0.0)
        errs.append(a["L2"][0])
    assert all(errs[i + 1] < 0.5 * errs[i] for i in range(3)), errs
    m = O.box_mesh(2, 3, x0=(0, 0, 0), length=(2, 2, 2), Ngeo=3, warp=True)
    c = np.array([1.2, 0.1, -0.2, 0.3, 3.0])
    U = np.broadcast_to(c, m["Elem_xGP"].shape[:-1] + (5,)).copy()
    a = O.analyze(m, U, np.zeros_like(U), exact_func=1, ref_state=tuple(c))
    assert np.allclose(a["bulk"], a["vol"] * c, rtol=1e-14)
    # interpolating a constant reproduces it to round-off (the Vandermonde rows sum to 1)
    assert np.all(a["L2"] <= 4e-15 * np.abs(c).max()) and np.all(a["Linf"] <= 4e-15 * np.abs(c).max())


def test_taylor_green_closed_forms():
    """Taylor-Green vortex (ExactFunc 7, rho = 1) with the exact velocity gradients on a fine
    Cartesian mesh: Ekin = 1/8, enstrophy 1/2 <|omega|^2> = 3/8, <S:S> = <grad v : grad v> =
    3/4 (incompressible), div v = 0 so DR_p = 0, max |omega| = 2 up to sampling."""
    m = O.box_mesh(4, 5, x0=(-math.pi,) * 3, length=(2 * math.pi,) * 3, Ngeo=1, warp=False)
    x = m["Elem_xGP"]
    U = O.exactfunc(7, x)
    s, c = np.sin, np.cos
    X, Y, Z = x[..., 0], x[..., 1], x[..., 2]
    gx, gy, gz = (np.zeros_like(U) for _ in range(3))
    # d(u,v,w)/dx, /dy, /dz of u = sin x cos y cos z, v = -cos x sin y cos z, w = 0
    gx[..., 1], gx[..., 2] = c(X) * c(Y) * c(Z), s(X) * s(Y) * c(Z)
    gy[..., 1], gy[..., 2] = -s(X) * s(Y) * c(Z), -c(X) * c(Y) * c(Z)
    gz[..., 1], gz[..., 2] = -s(X) * c(Y) * s(Z), c(X) * s(Y) * s(Z)
    a = O.analyze(m, U, np.zeros_like(U), (gx, gy, gz), diss_factor=1.0)
    vol = (2 * math.pi) ** 3
    assert a["vol"] == pytest.approx(vol, rel=1e-13)
    assert a["Ekin"] == pytest.approx(1.0 / 8.0, rel=1e-6)
    assert a["Ekin_comp"] == pytest.approx(1.0 / 8.0, rel=1e-6)
    assert a["Enstrophy_comp"] == pytest.approx(3.0 / 8.0, rel=1e-6)
    assert a["DR_u"] == pytest.approx(3.0 / 4.0, rel=1e-6)
    assert a["DR_S"] == pytest.approx(2.0 * 3.0 / 8.0, rel=1e-6)  # <S:S> = 1/2 <grad v : grad v>
    assert abs(a["DR_p"]) <= 1e-12
    assert a["max_Vorticity"] == pytest.approx(2.0, rel=2e-2)
    assert a["uprime"] == pytest.approx(math.sqrt(1.0 / 8.0 * 2.0 / 3.0), rel=1e-6)
