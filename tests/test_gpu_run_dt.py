"""The step graph of fx_run / fx_run_to forms a Navier-Stokes step's CALCTIMESTEP inside the
step's first lifting kernel (from the U and metrics it reads) instead of a pass of its own.
The trajectory, the time steps and the error codes must equal, bit for bit, those of the
step loop fx_calc_dt + fx_rk_step, whose CALCTIMESTEP is the separate kernel: every degree,
every volume-kernel tile, both precisions, BR1 and BR2, NFVSE and a mesh with mortars (and
Euler, whose graph keeps the CALCTIMESTEP pass)."""
import math

import numpy as np
import pytest

import pyoracle as O

pytestmark = pytest.mark.gpu


def engine(m, **kw):
    from fluxo_amd.native import Engine
    return Engine(m, **kw)


def tgv(m):
    return O.exactfunc(7, m["Elem_xGP"])


def step_loop(e, U0, nsteps, cfl, dfl):
    """fx_calc_dt + fx_rk_step; returns (U, t, dt of the last step, first error code, the
    step whose CALCTIMESTEP reported it or -1)."""
    e.upload(U0)
    t, dt = 0.0, 0.0
    for s in range(nsteps):
        dt, err = e.calc_dt(cfl, dfl)
        if err:
            return e.download(), t, dt, err, s
        e.rk_step(t, dt)
        t += dt
    return e.download(), t, dt, 0, -1


NS = dict(volume_flux=32, riemann=1, parabolic=True, mu0=1.0 / 1600, Pr=0.72)


@pytest.mark.parametrize("N", [2, 3, 4, 5])
@pytest.mark.parametrize("precision", [0, 1])
def test_run_equals_step_loop_every_tile(N, precision):
    m = O.box_mesh(3, N, x0=(-np.pi,) * 3, length=(2 * np.pi,) * 3, Ngeo=3, warp=True)
    U0 = tgv(m)
    e = engine(m, precision=precision, **NS)
    e.upload(U0)
    best, times = e.tune_volume(reps=1)
    for epb in times:
        e.use_volume_tile(epb)
        U_ref, t_ref, dt_ref, err_ref, _ = step_loop(e, U0, 6, 0.9, 0.9)
        e.upload(U0)
        t, dt, err = e.run(6, 0.9, 0.9)
        assert err == err_ref == 0
        assert t == t_ref and dt == dt_ref, epb
        assert np.array_equal(e.download(), U_ref), epb


@pytest.mark.parametrize("N", [2, 3, 4, 5])
def test_run_equals_step_loop_euler(N):
    m = O.box_mesh(3, N, Ngeo=3, warp=True)
    U0 = O.exactfunc(2, m["Elem_xGP"])
    e = engine(m, volume_flux=32, riemann=1)
    U_ref, t_ref, dt_ref, err_ref, _ = step_loop(e, U0, 5, 0.9, 0.0)
    e.upload(U0)
    t, dt, err = e.run(5, 0.9)
    assert err == err_ref == 0 and t == t_ref and dt == dt_ref
    assert np.array_equal(e.download(), U_ref)


@pytest.mark.parametrize("N", [2, 3, 4, 5])
def test_run_equals_step_loop_br2(N):
    m = O.box_mesh(3, N, x0=(-np.pi,) * 3, length=(2 * np.pi,) * 3, Ngeo=3, warp=True)
    U0 = tgv(m)
    e = engine(m, lifting=2, **NS)
    U_ref, t_ref, dt_ref, err_ref, _ = step_loop(e, U0, 4, 0.9, 0.9)
    e.upload(U0)
    t, dt, err = e.run(4, 0.9, 0.9)
    assert err == err_ref == 0 and t == t_ref and dt == dt_ref
    assert np.array_equal(e.download(), U_ref)


def test_run_equals_step_loop_nfvse():
    m = O.box_mesh(4, 3, x0=(-1.5,) * 3, length=(3.0,) * 3, Ngeo=3, warp=True)
    x = m["Elem_xGP"]
    sigma, p0 = 0.15, 1e-2
    p = p0 + 0.4 * np.exp(-(x ** 2).sum(-1) / (2 * sigma ** 2)) / ((2 * math.pi) ** 1.5 * sigma ** 3)
    U0 = np.zeros(x.shape[:-1] + (5,))
    U0[..., 0] = 1.0
    U0[..., 4] = p / 0.4
    e = engine(m, volume_flux=32, riemann=1, nfvse=True, parabolic=True, mu0=1e-3, Pr=0.72)
    U_ref, t_ref, dt_ref, err_ref, _ = step_loop(e, U0, 5, 0.5, 0.0)
    e.upload(U0)
    t, dt, err = e.run(5, 0.5)
    assert err == err_ref == 0 and t == t_ref and dt == dt_ref
    assert np.array_equal(e.download(), U_ref)


def test_run_to_equals_step_loop_on_a_mortar_mesh():
    from fluxo_amd.mesh import hopr_mesh
    from pathlib import Path
    path = Path(__file__).parent / "golden" / "io" / "DeformedBoxMortarPeriodic__Ngeo_2_Level_01_mesh.h5"
    m = hopr_mesh(str(path), 3)
    U0 = O.exactfunc(2, m["Elem_xGP"])
    e = engine(m, **NS)
    U_ref, t_ref, _, err_ref, _ = step_loop(e, U0, 4, 0.9, 0.9)
    e.upload(U0)
    t, _, steps, err = e.run_to(1.0e9, 0.9, 0.9, max_steps=4)
    assert err == err_ref == 0 and steps == 4 and t == t_ref
    assert np.array_equal(e.download(), U_ref)


def test_run_reports_an_error_at_the_same_step_as_the_step_loop():
    """A CFL number far above stability: the state blows up within a few steps.  The step
    whose CALCTIMESTEP first sees it (NaN density or non-positive pressure) and its error code
    are those of the step loop."""
    m = O.box_mesh(3, 3, x0=(-np.pi,) * 3, length=(2 * np.pi,) * 3, Ngeo=3, warp=True)
    U0 = tgv(m)
    e = engine(m, **NS)
    _, _, _, err_ref, s_err = step_loop(e, U0, 60, 40.0, 40.0)
    assert err_ref != 0 and s_err >= 1
    e.upload(U0)
    _, _, err = e.run(s_err, 40.0, 40.0)
    assert err == 0
    e.upload(U0)
    _, _, err = e.run(s_err + 1, 40.0, 40.0)
    assert err == err_ref
