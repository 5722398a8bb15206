"""The CPU oracle reproduces the reference's own outputs (golden vectors made by the
reference's compiled code, tests/golden/make_golden.py) on warped periodic meshes."""
import numpy as np
import pytest

import pyoracle as O
from golden_cases import initial_state, lifting_args, load, mesh_args, names, nfvse_args, rel_l2


@pytest.mark.parametrize("name", names())
def test_oracle_matches_reference(name):
    g = load(name)
    m = O.box_mesh(g["K"], g["N"], **mesh_args(g))
    U = initial_state(g, O.exactfunc, m)
    P = O.Problem(m, volflux=g["volflux"], riemann=g["riemann"], parabolic=bool(g["parabolic"]),
                  mu0=g["mu0"], Pr=g["Pr"], kappa=g["kappa"], **nfvse_args(g), **lifting_args(g))
    if "Mf" in g:
        assert np.abs(m["Mf"] - g["Mf"]).max() <= 1e-15 * np.abs(g["Mf"]).max()
        assert np.abs(m["sJ"] - g["sJ"]).max() <= 1e-15 * np.abs(g["sJ"]).max()
        assert np.abs(m["SurfElem"] - g["SurfElem"]).max() <= 1e-15 * np.abs(g["SurfElem"]).max()
        assert rel_l2(P.volint(U), g["volint"]) <= 1e-14
    Ut = P.rhs(U)
    assert rel_l2(Ut, g["Ut"]) <= 1e-14
    if g["nfvse"]:
        assert np.array_equal(P.alpha, g["alpha"])
    if g["parabolic"]:
        for k, gr in zip(("gradPx", "gradPy", "gradPz"), P.lifting(U)):
            assert rel_l2(gr, g[k]) <= 1e-12
    S, dS = P.entropy(U, Ut)
    assert S == pytest.approx(g["Entropy"], rel=1e-11)        # printed with 12 digits
    assert dS == pytest.approx(g["dSdU_Ut"], rel=1e-11, abs=1e-14)
    dt, err = P.calc_dt(U, g["cfl"], g["dfl"])
    assert err == g["errType0"] and dt == pytest.approx(g["dt0"], rel=1e-15)
    # trajectories where the oracle finishes in seconds (node-steps bounded)
    if g["nsteps"] <= 20 or g["K"] ** 3 * (g["N"] + 1) ** 3 * g["nsteps"] <= 2e5:
        Ue, Utmp, t = U.copy(), np.zeros_like(U), 0.0
        for _ in range(g["nsteps"]):
            d, _ = P.calc_dt(Ue, g["cfl"], g["dfl"])
            P.lserk_step(Ue, Utmp, t, d)
            t += d
        assert rel_l2(Ue, g["U_end"]) <= 1e-13
        assert t == pytest.approx(g["t_end"], rel=1e-14)
