"""The engine's sub-interfaces through the C ABI against the reference's own outputs
(tests/golden/iface): fx_fillflux (FillFlux, fillflux.f90:41-110), fx_surfint (SurfInt,
surfint.f90:48-264) by composition with the reference's VolInt into its right-hand side, and
fx_volint_nfvse (VolInt_NFVSE, nfvse.f90:475-561).  FP64 bar: relative L2 <= 1e-12.  Also the
end-time loop fx_run_to (timedisc.f90:313-317) against the oracle stepped with the same rule."""
import numpy as np
import pytest

import pyoracle as O
from iface_cases import engine_kwargs, load, mesh, mhd_kwargs, names, rel_l2

pytestmark = pytest.mark.gpu


def engine(g, m):
    from fluxo_amd.native import Engine, MhdEngine
    if g["kind"] == "mhd":
        return MhdEngine(m, **mhd_kwargs(g))
    return Engine(m, volume_flux=int(g["volflux"]), riemann=int(g["riemann"]), parabolic=bool(g["parabolic"]),
                  **engine_kwargs(g))


@pytest.mark.parametrize("name", names())
def test_fillflux_matches_reference(name):
    from fluxo_amd.mesh import box_mesh
    g = load(name)
    m = mesh(g, box_mesh)
    e = engine(g, m)
    e.upload(g["U"])
    fm, fs = e.fillflux()
    em, es = rel_l2(fm, g["Flux_master"]), rel_l2(fs, g["Flux_slave"])
    print(f"{name}: fx_fillflux Flux_master {em:.2e}, Flux_slave {es:.2e}")
    assert em <= 1e-12 and es <= 1e-12


@pytest.mark.parametrize("name", [n for n in names() if load(n)["parabolic"] == 0])
def test_surfint_composes_the_reference_rhs(name):
    from fluxo_amd.mesh import box_mesh
    g = load(name)
    m = mesh(g, box_mesh)
    e = engine(g, m)
    vol = g["volint_nfvse"] if g["kind"] == "nfvse" else g["volint"]
    ut = e.surfint(g["Flux_master"], g["Flux_slave"], vol) * (-m["sJ"][..., None])
    err = rel_l2(ut, g["Ut"])
    print(f"{name}: -sJ (VolInt + fx_surfint) vs the reference's RHS {err:.2e}")
    assert err <= 1e-13


@pytest.mark.parametrize("name", names())
def test_surfint_matches_oracle(name):
    """Arbitrary fluxes (independent master and slave values): the engine's SurfInt and the
    oracle's agree to rounding (the engine may fuse the multiply-add)."""
    from fluxo_amd.mesh import box_mesh
    g = load(name)
    m = mesh(g, box_mesh)
    e = engine(g, m)
    rng = np.random.default_rng(3)
    fm = rng.standard_normal(g["Flux_master"].shape)
    fs = rng.standard_normal(g["Flux_master"].shape)
    ut0 = rng.standard_normal(g["U"].shape)
    got = e.surfint(fm, fs, ut0)
    P = O.MhdProblem(m) if g["kind"] == "mhd" else O.Problem(m)
    ref = P.surfint(fm, fs, ut0)
    assert np.max(np.abs(got - ref)) <= 1e-14 * np.max(np.abs(ref))


@pytest.mark.parametrize("name", [n for n in names() if load(n)["kind"] == "nfvse"])
def test_volint_nfvse_matches_reference(name):
    from fluxo_amd.mesh import box_mesh
    g = load(name)
    m = mesh(g, box_mesh)
    e = engine(g, m)
    e.upload(g["U"])
    err = rel_l2(e.volint_nfvse(), g["volint_nfvse"])
    print(f"{name}: fx_volint_nfvse {err:.2e}")
    assert err <= 1e-12
    # fx_volint stays the unblended split form
    assert rel_l2(e.volint(), g["volint"]) <= 1e-12


@pytest.mark.parametrize("parabolic", [False, True])
def test_run_to_end_time(parabolic):
    """fx_run_to stops exactly where the reference's TimeDisc does: CFL steps until the step
    that reaches tEnd within 1e-4 of dt, which is shortened to end at tEnd; the oracle stepped
    with the same rule gives the same steps, times and state."""
    from fluxo_amd.native import Engine
    m = O.box_mesh(3, 3, x0=(-np.pi,) * 3, length=(2 * np.pi,) * 3, Ngeo=3, warp=True)
    U = O.exactfunc(7, m["Elem_xGP"])
    kw = dict(parabolic=True, mu0=1.0 / 1600, Pr=0.72) if parabolic else {}
    P = O.Problem(m, volflux=32, riemann=1, **kw)
    t_end = 0.2345
    Ue, Utmp, t, steps = U.copy(), np.zeros_like(U), 0.0, 0
    while True:
        dt, err = P.calc_dt(Ue, 0.9, 0.9)
        last = t_end - t <= dt * (1.0 + 1.0e-4)
        if last:
            dt = t_end - t
        P.lserk_step(Ue, Utmp, t, dt)
        t += dt
        steps += 1
        if last:
            break
    e = Engine(m, volume_flux=32, riemann=1, **kw)
    e.upload(U)
    tg, dtg, sg, errg = e.run_to(t_end, 0.9, 0.9)
    print(f"parabolic={parabolic}: {steps} steps to t={t!r} (engine {sg} steps, t={tg!r}, last dt {dtg:.3e})")
    assert errg == 0 and sg == steps and tg == t
    assert rel_l2(e.download(), Ue) <= 1e-12
    # a step limit below the steps needed stops there; the next call continues to tEnd
    e2 = Engine(m, volume_flux=32, riemann=1, **kw)
    e2.upload(U)
    t1, _, s1, _ = e2.run_to(t_end, 0.9, 0.9, max_steps=3)
    assert s1 == 3 and t1 < t_end
    t2, _, s2, _ = e2.run_to(t_end, 0.9, 0.9, t0=t1)
    assert s1 + s2 == steps and t2 == pytest.approx(t_end, rel=1e-15)
