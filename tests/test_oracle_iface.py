"""The CPU oracle's FillFlux, SurfInt and VolInt_NFVSE against the reference's own sub-interface
outputs (tests/golden/iface): the face fluxes Flux_master / Flux_slave of the right-hand side
(fillflux.f90:41-110), VolInt_NFVSE's blended volume integral (nfvse.f90:475-561), and SurfInt
(surfint.f90:48-264) checked by composition: -sJ (VolInt + SurfInt(Flux_master, Flux_slave)) is
the reference's right-hand side for the cases without a viscous volume term."""
import numpy as np
import pytest

import pyoracle as O
from iface_cases import engine_kwargs, load, mesh, mhd_kwargs, names, rel_l2


def problem(g, m):
    if g["kind"] == "mhd":
        return O.MhdProblem(m, **mhd_kwargs(g))
    kw = engine_kwargs(g)
    if kw.pop("nfvse", False):
        kw.update(nfvse=True)
    return O.Problem(m, volflux=int(g["volflux"]), riemann=int(g["riemann"]), parabolic=bool(g["parabolic"]), **kw)


@pytest.mark.parametrize("name", names())
def test_oracle_fillflux_matches_reference(name):
    g = load(name)
    m = mesh(g, O.box_mesh)
    P = problem(g, m)
    fm, fs = P.fillflux(g["U"])
    tol = 1e-13 if g["parabolic"] else 1e-14  # viscous part through the lifted gradients
    em, es = rel_l2(fm, g["Flux_master"]), rel_l2(fs, g["Flux_slave"])
    print(f"{name}: Flux_master {em:.2e}, Flux_slave {es:.2e}")
    assert em <= tol and es <= tol
    if g["kind"] != "mhd":
        assert np.array_equal(fm, fs)


@pytest.mark.parametrize("name", [n for n in names() if load(n)["parabolic"] == 0])
def test_oracle_surfint_composes_the_reference_rhs(name):
    g = load(name)
    m = mesh(g, O.box_mesh)
    P = problem(g, m)
    vol = g["volint_nfvse"] if g["kind"] == "nfvse" else g["volint"]
    ut = P.surfint(g["Flux_master"], g["Flux_slave"], vol) * (-m["sJ"][..., None])
    err = rel_l2(ut, g["Ut"])
    print(f"{name}: -sJ (VolInt + SurfInt) vs the reference's RHS {err:.2e}")
    assert err <= 1e-14


@pytest.mark.parametrize("name", [n for n in names() if load(n)["kind"] == "nfvse"])
def test_oracle_volint_nfvse_matches_reference(name):
    g = load(name)
    m = mesh(g, O.box_mesh)
    P = problem(g, m)
    err = rel_l2(P.volint_nfvse(g["U"]), g["volint_nfvse"])
    print(f"{name}: VolInt_NFVSE {err:.2e}")
    assert err <= 1e-14


def test_oracle_surfint_adds_in_side_order():
    """SurfInt on a closed-form input: Flux_master = 1 and Flux_slave = 1 on every side of a
    periodic box give +L_HatMinus0 on every master face node and -L_HatMinus0 on every slave
    face node, so Ut(node) = L0 (number of master faces - number of slave faces) of that node;
    on the Cartesian box every element owns its XI_PLUS, ETA_PLUS and ZETA_PLUS sides
    (master) and is the slave on the minus sides, so Ut = L0 (#plus faces - #minus faces)."""
    m = O.box_mesh(3, 3)
    P = O.Problem(m)
    fm = np.ones(P.face_shape())
    ut = P.surfint(fm, fm, np.zeros(P.shape()))
    n = 4
    L0 = 1.0 / (1.0 / 6.0)  # 1 / w_0 for N = 3
    idx = np.arange(n)
    plus = (idx == n - 1).astype(float)
    minus = (idx == 0).astype(float)
    cnt = (plus - minus)
    expect = L0 * (cnt[None, None, :] + cnt[None, :, None] + cnt[:, None, None])  # (k, j, i)
    got = ut[..., 0]
    for e in range(m["nElems"]):
        assert np.allclose(got[e], expect, rtol=0, atol=1e-13)
