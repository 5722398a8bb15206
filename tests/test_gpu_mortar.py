"""The engine on the reference's non-conforming HOPR meshes (mortars; read by fluxo_amd.io,
metrics from the Fortran host) against the reference's code on the same meshes and states
(tests/golden/mortar: U_Mortar, FillFlux, Flux_Mortar, SurfInt, for Navier-Stokes and
resistive GLM-MHD the BR1 and BR2 lifting with Flux_Cons_Mortar and U_Mortar of the gradients, for
GLM-MHD the Powell terms and GLM_ch, for NFVSE the blending coefficient's propagation across
mortars, and the LSERK trajectory of the reference, run by its harness), through the C ABI.  FP64 bar: relative L2 <= 1e-12 per
RHS and face flux, <= 1e-10 after the stored RK steps (100 on three cases).  Properties the
reference's own free-stream mortar tests check (tests/runs/navst/freestream/
nonconforming_projmortar, nonconforming_collmortar: L_inf <= 1e-11) and conservation across
the mortars close the tests."""
import glob
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
IO_DIR = os.path.join(HERE, "golden", "io")
MORTAR_DIR = os.path.join(HERE, "golden", "mortar")
NAMES = sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(MORTAR_DIR, "*.npz")))
DEFORMED = "DeformedBoxMortarPeriodic__Ngeo_2_Level_01_mesh.h5"


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def load(name):
    with np.load(os.path.join(MORTAR_DIR, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def engine(g, **kw):
    from fluxo_amd.mesh import hopr_mesh
    from fluxo_amd.native import Engine, MhdEngine
    m = hopr_mesh(os.path.join(IO_DIR, str(g["mesh_file"])), int(g["N"]))
    if "glm_scale" in g:  # GLM-MHD
        res = dict(mu=float(g["mu"]), eta=float(g["eta"]), Pr=float(g["Pr"])) if int(g["resistive"]) else None
        e = MhdEngine(m, kappa=float(g["kappa"]), mu_0=float(g["mu_0"]), glm_scale=float(g["glm_scale"]),
                      riemann=int(g["riemann"]), resistive=res, which_mortar=int(g["whichMortar"]),
                      lifting=int(g.get("lifting", 1)), etaBR2=float(g.get("etaBR2", 2.0)), **kw)
        return e, m
    par = bool(int(g.get("parabolic", 0)))
    if "nfvse" in g:
        kw.update(nfvse=True, alpha_max=float(g["alpha_max"]), alpha_min=float(g["alpha_min"]),
                  sweeps=int(g["sweeps"]), prop_factor=float(g["prop_factor"]))
    e = Engine(m, volume_flux=int(g["volflux"]), riemann=int(g["riemann"]), which_mortar=int(g["whichMortar"]),
               parabolic=par, mu0=float(g.get("mu0", 0.0)), Pr=float(g.get("Pr", 0.72)),
               lifting=int(g.get("lifting", 1)), etaBR2=float(g.get("etaBR2", 2.0)), **kw)
    return e, m


@pytest.mark.parametrize("name", NAMES)
def test_rhs_and_face_fluxes_match_reference(name):
    g = load(name)
    print(f"case {name}", flush=True)
    e, m = engine(g)
    e.upload(g["U"])
    if "gradPx" in g:  # BR1 lifting: Flux_Cons_Mortar of the lifting fluxes on the big sides
        gerr = max(rel_l2(a, g[k]) for a, k in zip(e.lifting(), ("gradPx", "gradPy", "gradPz")))
        print(f"{name}: gradients relL2 vs reference {gerr:.3e}")
        assert gerr <= 1e-12
    err = rel_l2(e.rhs(), g["Ut"])
    fm, fs = e.fillflux()
    nm = m["nMortarSides"]
    ferr = rel_l2(fm, g["Flux_master"])
    serr = rel_l2(fs[nm:], g["Flux_slave"])
    big = rel_l2(fm[:nm], g["Flux_master"][:nm])
    print(f"{name}: RHS {err:.3e}  Flux_master {ferr:.3e} (big sides {big:.3e})  Flux_slave {serr:.3e}")
    assert err <= 1e-12 and ferr <= 1e-12 and serr <= 1e-12 and big <= 1e-12
    if "alpha" in g:  # NFVSE: blending coefficient propagated across the mortars, VolInt_NFVSE
        aerr = float(np.abs(e.blending() - g["alpha"]).max())
        verr = rel_l2(e.volint_nfvse(), g["volint_nfvse"])
        print(f"{name}: alpha max diff {aerr:.2e}, VolInt_NFVSE relL2 {verr:.3e}")
        assert aerr <= 1e-13 and verr <= 1e-12
    if "gradPx" in g or "glm_scale" in g or "alpha" in g:
        return
    # Euler: SurfInt alone on the reference's fluxes and volume integral gives its RHS
    fsp = np.concatenate([np.zeros_like(fm[:nm]), g["Flux_slave"]])
    ut = e.surfint(g["Flux_master"], fsp, g["volint"]) * (-m["sJ"][..., None])
    assert rel_l2(ut, g["Ut"]) <= 1e-12


@pytest.mark.parametrize("name", NAMES)
def test_trajectory_matches_reference(name):
    g = load(name)
    print(f"case {name}", flush=True)
    e, _ = engine(g)
    e.upload(g["U"])
    dfl = float(g.get("dfl", 0.0)) if int(g.get("parabolic", g.get("resistive", 0))) else 0.0
    t, _, errt = e.run(int(g["nsteps"]), float(g["cfl"]), dfl)
    traj = rel_l2(e.download(), g["U_end"])
    print(f"{name}: {int(g['nsteps'])} steps relL2 vs reference {traj:.3e}")
    assert errt == 0 and traj <= 1e-10
    assert t == pytest.approx(float(g["t_end"]), rel=1e-12)


@pytest.mark.parametrize("which_mortar", [0, 1])
@pytest.mark.parametrize("volume_flux,riemann,mu0", [(32, 1, None), (1, 1, None), (10, 32, None), (1, 1, 1e-6),
                                                    (32, 1, 1e-3)])
def test_free_stream_on_curved_mortars(which_mortar, volume_flux, riemann, mu0):
    """The reference's nonconforming free-stream tests (navst/freestream/nonconforming_projmortar
    and _collmortar): DeformedBoxMortarPeriodic (Ngeo 2, mortar types 1, 2, 3), N=4, RefState
    (1, 0.1, 0.2, -0.3, 20), projection and collocation mortars, VolumeFlux 1, Riemann 1,
    mu0 1e-6 (BR1), CFL 0.9, DFL 0.5; also Euler and other flux pairs.  The RHS vanishes and
    100 steps keep the state to L_inf 1e-11 (the reference test's bound) for VolumeFlux 1 /
    Riemann 1, to 1e-12 of the state's magnitude for the other pairs."""
    from fluxo_amd.mesh import hopr_mesh
    from fluxo_amd.native import Engine
    print(f"free stream whichMortar={which_mortar} VF={volume_flux} RS={riemann} mu0={mu0}", flush=True)
    m = hopr_mesh(os.path.join(IO_DIR, DEFORMED), 4)
    cons = np.array([1.0, 0.1, 0.2, -0.3, 20.0 / 0.4 + 0.5 * (0.01 + 0.04 + 0.09)])  # kappa 1.4
    U = np.broadcast_to(cons, (m["nElems"], 5, 5, 5, 5)).copy()
    e = Engine(m, volume_flux=volume_flux, riemann=riemann, which_mortar=which_mortar, parabolic=mu0 is not None,
               mu0=mu0 or 0.0)
    e.upload(U)
    ut = e.rhs()
    assert np.abs(ut).max() <= 1e-10, np.abs(ut).max()
    _, _, errt = e.run(100, 0.9, 0.5)
    assert errt == 0
    err = np.abs(e.download() - U).max()
    print(f"  RHS max {np.abs(ut).max():.2e}, L_inf after 100 steps {err:.2e}")
    assert err <= (1e-11 if (volume_flux, riemann) == (1, 1) else 1e-12 * np.abs(U).max())


@pytest.mark.parametrize("which_mortar", [0, 1])
@pytest.mark.parametrize("mesh_file", ["CartBoxMortarPeriodic_Ngeo1_4to1_mesh.h5", DEFORMED])
def test_mortars_conserve(which_mortar, mesh_file):
    """sum over nodes of w_i w_j w_k J Ut = 0: the big side's projected flux carries exactly
    what its small sides carry (both mortar projections keep the mean)."""
    from fluxo_amd.mesh import hopr_mesh
    from fluxo_amd.native import Engine
    import pyoracle as O
    N = 4
    m = hopr_mesh(os.path.join(IO_DIR, mesh_file), N)
    U = O.exactfunc(2, m["Elem_xGP"])
    U = U * (1.0 + 0.05 * np.random.default_rng(7).uniform(-1.0, 1.0, U.shape))
    e = Engine(m, volume_flux=32, riemann=1, which_mortar=which_mortar)
    e.upload(U)
    ut = e.rhs()
    c = np.zeros(N + 1)
    c[N] = 1.0
    x = np.concatenate(([-1.0], np.sort(np.polynomial.legendre.legroots(np.polynomial.legendre.legder(c))), [1.0]))
    w = 2.0 / (N * (N + 1) * np.polynomial.legendre.legval(x, c) ** 2)
    W = np.einsum("i,j,k->kji", w, w, w)[None] / m["sJ"]
    tot = np.einsum("eijk,eijkv->v", W, ut)
    scale = np.einsum("eijk,eijkv->v", W, np.abs(ut))
    print(f"{mesh_file} whichMortar={which_mortar}: conservation {np.abs(tot / scale).max():.2e}")
    assert np.all(np.abs(tot) <= 1e-13 * scale)


def test_fp32_split_on_mortars():
    """The FP32-split volume integral on a mortar mesh, against FP64 (the mortar kernels are
    FP64 in both): relative L2 within the conforming FP32-split tests' 3e-5, L_inf within
    3e-4 (measured 2.2e-5 and 1.3e-4: the refined elements' large sJ scale the FP32
    rounding of the flux differencing)."""
    from fluxo_amd.native import FX_PREC_FP32_SPLIT
    g = load("mortar_deformed_n4_ranocha_llf_proj")
    e, _ = engine(g)
    e.upload(g["U"])
    ref = e.rhs()
    e32, _ = engine(g, precision=FX_PREC_FP32_SPLIT)
    e32.upload(g["U"])
    got = e32.rhs()
    err = rel_l2(got, ref)
    li = float(np.abs(got - ref).max() / np.abs(ref).max())
    print(f"FP32-split on mortars: relL2 vs FP64 {err:.2e}, relLinf {li:.2e}")
    assert err <= 3e-5 and li <= 3e-4


def test_mortar_options_are_checked():
    from fluxo_amd.native import Engine, FxError
    from fluxo_amd.mesh import hopr_mesh
    m = hopr_mesh(os.path.join(IO_DIR, DEFORMED), 3)
    with pytest.raises(FxError, match="whichMortar"):
        Engine(m, which_mortar=-1)
    with pytest.raises(FxError, match="whichMortar"):
        Engine(m, which_mortar=2)
