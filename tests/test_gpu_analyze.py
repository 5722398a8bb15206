"""fx_analyze (the reference's Analyze on the device: CalcErrorNorms, volume, CalcBulk, the
Taylor-Green vortex testcase quantities) against the values the reference's own Analyze
computed (tests/golden/analyze) and against the oracle's restatement, through the C ABI.
Bar: relative 1e-12 for every reported value (bulk integrals of Ut, which are round-off of
zero, to 1e-13 of max|Ut| times the volume); against the oracle 1e-11 relative, error norms
included."""
import numpy as np
import pytest

import pyoracle as O
from analysis_ref import NAMES, check_against_reference, load, mesh_of

pytestmark = pytest.mark.gpu


def engine_for(g, m):
    from fluxo_amd.native import Engine, MhdEngine
    if str(g["eq"]) == "ns":
        par = bool(int(g["parabolic"]))
        e = Engine(m, volume_flux=32, riemann=1, parabolic=par, mu0=float(g["mu0"]), Pr=float(g["Pr"]))
        cfg = dict(exact_func=int(g["exact"]), n_analyze=int(g["nanalyze"]),
                   ref_state=(1.0, 0.0, 0.0, 0.0, 1.0 / 0.4), diss_factor=1.0)
    else:
        res = dict(mu=float(g["mu"]), eta=float(g["eta"]), Pr=float(g["Pr"]))
        e = MhdEngine(m, resistive=res)
        cfg = dict(exact_func=1, ref_state=(1.0, 0, 0, 0, 1.5, 0, 0, 0, 0), diss_factor=float(g["mu_0"]))
    e.upload(g["U"])
    e.upload_coords(m["Elem_xGP"])
    return e, cfg


@pytest.mark.parametrize("name", NAMES)
def test_analysis_matches_reference(name):
    g = load(name)
    m = mesh_of(g)
    e, cfg = engine_for(g, m)
    a = e.analyze(float(g["t0"]), **cfg)
    worst = check_against_reference(a, g)
    print(f"{name}: worst relative difference to the reference's Analyze {worst:.2e}")


def compare(a, b, nv, tol=1e-11):
    b = dict(b, maxabs_Ut=b["maxabsUt"])
    worst = 0.0
    for k, v in b.items():
        if k in ("maxabsUt", "bulk_t"):
            continue
        x, y = np.atleast_1d(a[k])[:nv], np.atleast_1d(v)[:nv]
        err = np.abs(x - y) / np.maximum(np.abs(y), 1e-300)
        # entries that are round-off of zero (bulk of a zero-mean field): absolute bar
        small = np.abs(x - y) <= 1e-15 * max(1.0, float(np.abs(y).max()))
        assert np.all((err <= tol) | small), (k, x, y)
        worst = max(worst, float(np.where(small, 0.0, err).max()))
    scale = np.asarray(b["maxabsUt"]) * b["vol"]
    assert np.all(np.abs(np.asarray(a["bulk_t"]) - b["bulk_t"]) <= 1e-13 * scale)
    return worst


@pytest.mark.parametrize("N", [2, 3, 4, 5])
def test_density_wave_error_norms_match_oracle(N):
    """Euler density wave with noise on a warped mesh, ExactFunc 2 at t = 0.37, default and
    non-default NAnalyze."""
    from fluxo_amd.native import Engine
    m = O.box_mesh(3, N, x0=(0, 0, 0), length=(2, 2, 2), Ngeo=3, warp=True)
    U = O.exactfunc(2, m["Elem_xGP"], t=0.37)
    U = U * (1.0 + 1e-4 * np.random.default_rng(N).uniform(-1, 1, U.shape))
    e = Engine(m, volume_flux=32, riemann=1)
    e.upload(U)
    e.upload_coords(m["Elem_xGP"])
    Ut = O.Problem(m).rhs(U)
    for na in (0, N + 2):
        a = e.analyze(0.37, exact_func=2, n_analyze=na)
        b = O.analyze(m, U, Ut, exact_func=2, n_analyze=na, t=0.37)
        worst = compare(a, b, 5)
        print(f"density wave N={N} NAnalyze={na or 2 * (N + 1)}: L2 rho {a['L2'][0]:.3e}, worst rel. diff {worst:.1e}")


@pytest.mark.parametrize("N", [2, 3, 4, 5])
def test_taylor_green_quantities_match_oracle(N):
    """Navier-Stokes Taylor-Green vortex (BR1): energies, enstrophy, dissipation rates from the
    engine's own lifted gradients, against the oracle with its lifting."""
    from fluxo_amd.native import Engine
    m = O.box_mesh(3, N, x0=(-np.pi,) * 3, length=(2 * np.pi,) * 3, Ngeo=3, warp=True)
    U = O.exactfunc(7, m["Elem_xGP"])
    mu0 = 1.0 / 1600
    P = O.Problem(m, volflux=32, riemann=1, parabolic=True, mu0=mu0, Pr=0.72)
    e = Engine(m, volume_flux=32, riemann=1, parabolic=True, mu0=mu0, Pr=0.72)
    e.upload(U)
    a = e.analyze(0.0, diss_factor=mu0)
    b = O.analyze(m, U, P.rhs(U), P.lifting(U), diss_factor=mu0)
    worst = compare(a, b, 5)
    print(f"TGV N={N}: Ekin {a['Ekin']:.6f} enstrophy {a['Enstrophy_comp']:.6f}, worst rel. diff {worst:.1e}")


def test_mhd_analysis_matches_oracle():
    from fluxo_amd.native import MhdEngine
    m = O.box_mesh(3, 4, x0=(0, 0, 0), length=(1, 1, 1), Ngeo=3, warp=True)
    res = dict(mu=0.01, eta=0.02, Pr=0.72)
    P = O.MhdProblem(m, resistive=res)
    U = P.exact_orszag_tang()
    U[..., 8] = 1e-2 * np.random.default_rng(4).uniform(-1, 1, U[..., 8].shape)
    e = MhdEngine(m, resistive=res)
    e.upload(U)
    e.upload_coords(m["Elem_xGP"])
    ref = (1.0, 0, 0, 0, 1.5, 0, 0, 0, 0)
    a = e.analyze(0.0, exact_func=1, ref_state=ref)
    b = O.analyze(m, U, P.rhs(U, glm_ch=0.5), P.lifting(U), kappa=5.0 / 3.0, exact_func=1, ref_state=ref)
    worst = compare(a, b, 9)
    print(f"MHD N=4: Emag {a['Emag_comp']:.6f} max current {a['max_Current']:.4f}, worst rel. diff {worst:.1e}")


def test_rejects_bad_options():
    from fluxo_amd.native import Engine, FxError, MhdEngine
    m = O.box_mesh(2, 3)
    e = Engine(m)
    e.upload(O.exactfunc(2, m["Elem_xGP"]))
    with pytest.raises(FxError, match="fx_elem_xgp_upload"):
        e.analyze(exact_func=2)
    e.upload_coords(m["Elem_xGP"])
    with pytest.raises(FxError, match="above 16"):
        e.analyze(exact_func=2, n_analyze=17)
    mm_mesh = O.box_mesh(2, 3, x0=(0, 0, 0), length=(1, 1, 1))
    mm = MhdEngine(mm_mesh)
    mm.upload(O.MhdProblem(mm_mesh).exact_orszag_tang())
    mm.upload_coords(mm_mesh["Elem_xGP"])
    with pytest.raises(FxError, match="AnalyzeExactFunc"):
        mm.analyze(exact_func=2)
