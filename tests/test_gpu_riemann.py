"""Surface fluxes and the Ismail-Roe volume flux on the GPU against the oracle: the
entropy-stable solvers (Riemann 4, 33, 44, in the Cartesian frame) and EC-KEP + face LLF (16,
a face-wide maximum wave speed), on warped meshes with perturbed states, N = 2..5 (relative
L2 <= 1e-12 per RHS), Navier-Stokes with 33, a 10-step trajectory (<= 1e-10), VolumeFlux
10, and the other surface fluxes 2, 3, 5-8, 10-15, 17-24.  The reference's own outputs are checked in test_gpu_golden (euler_*_es*,
*_eckep_llf16_*, ns_*_es33_*, *_ismailroe_*)."""
import math

import numpy as np
import pytest

import pyoracle as O

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def state(m, seed=9):
    U = O.exactfunc(2, m["Elem_xGP"])
    rng = np.random.default_rng(seed)
    return U * (1.0 + 0.05 * rng.uniform(-1.0, 1.0, U.shape))


@pytest.mark.parametrize("rs", [4, 16, 33, 44])
@pytest.mark.parametrize("N", [2, 3, 4, 5])
def test_surface_flux_rhs(rs, N):
    from fluxo_amd.native import Engine
    m = O.box_mesh(3, N, x0=(0, 0, 0), length=(2, 2, 2), Ngeo=3, warp=True)
    U = state(m)
    e = Engine(m, volume_flux=32, riemann=rs)
    e.upload(U)
    err = rel_l2(e.rhs(), O.Problem(m, volflux=32, riemann=rs).rhs(U))
    print(f"Riemann {rs} N={N}: RHS relL2 {err:.3e}")
    assert err <= 1e-12


NEW_RS = [2, 3, 5, 6, 7, 8, 10, 11, 12, 13, 14, 15, 17, 18, 19, 20, 21, 22, 23, 24]


@pytest.mark.parametrize("vf", [2, 3, 4])
@pytest.mark.parametrize("N", [2, 5])
def test_roe_with_split_form_central_part(vf, N):
    """Roe (3) whose central part follows VolumeFlux 2 / 3 / 4 (riemann.f90:836-856)."""
    from fluxo_amd.native import Engine
    m = O.box_mesh(3, N, x0=(0, 0, 0), length=(2, 2, 2), Ngeo=3, warp=True)
    U = state(m)
    e = Engine(m, volume_flux=vf, riemann=3)
    e.upload(U)
    err = rel_l2(e.rhs(), O.Problem(m, volflux=vf, riemann=3).rhs(U))
    print(f"Roe VolumeFlux {vf} N={N}: RHS relL2 {err:.3e}")
    assert err <= 1e-12


@pytest.mark.parametrize("rs", NEW_RS)
@pytest.mark.parametrize("N", [2, 5])
def test_other_surface_fluxes_rhs(rs, N):
    """Riemann 2-24 beyond 4, 9, 16 (HLLC, Roe, HLL, HLLE, HLLEM, volume-flux averages in
    the normal direction, their LLF forms with the face's largest wave speed, Ismail-Roe +
    entropy-Jacobian LLF) against the oracle; N = 3, 4 are in the golden vectors
    (euler_*_riemann*)."""
    from fluxo_amd.native import Engine
    m = O.box_mesh(3, N, x0=(0, 0, 0), length=(2, 2, 2), Ngeo=3, warp=True)
    U = state(m)
    vf = 10 if rs in (5, 13) else 32
    e = Engine(m, volume_flux=vf, riemann=rs)
    e.upload(U)
    err = rel_l2(e.rhs(), O.Problem(m, volflux=vf, riemann=rs).rhs(U))
    print(f"Riemann {rs} N={N}: RHS relL2 {err:.3e}")
    assert err <= 1e-12


@pytest.mark.parametrize("rs", [13, 17])
def test_other_surface_fluxes_navier_stokes(rs):
    from fluxo_amd.native import Engine
    m = O.box_mesh(3, 3, x0=(-math.pi,) * 3, length=(2 * math.pi,) * 3, Ngeo=3, warp=True)
    U = O.exactfunc(7, m["Elem_xGP"])
    kw = dict(parabolic=True, mu0=1e-3, Pr=0.72)
    e = Engine(m, volume_flux=32, riemann=rs, **kw)
    e.upload(U)
    err = rel_l2(e.rhs(), O.Problem(m, volflux=32, riemann=rs, **kw).rhs(U))
    print(f"Navier-Stokes Riemann {rs}: RHS relL2 {err:.3e}")
    assert err <= 1e-12


@pytest.mark.parametrize("rs", [16, 33])
def test_surface_flux_trajectory(rs):
    from fluxo_amd.native import Engine
    m = O.box_mesh(3, 3, x0=(0, 0, 0), length=(2, 2, 2), Ngeo=3, warp=True)
    U = state(m)
    P = O.Problem(m, volflux=32, riemann=rs)
    Ue, Utmp, t = U.copy(), np.zeros_like(U), 0.0
    for _ in range(10):
        d, _ = P.calc_dt(Ue, 0.9)
        P.lserk_step(Ue, Utmp, t, d)
        t += d
    e = Engine(m, volume_flux=32, riemann=rs)
    e.upload(U)
    _, _, err = e.run(10, 0.9)
    r = rel_l2(e.download(), Ue)
    print(f"Riemann {rs}: 10 steps relL2 {r:.3e}")
    assert err == 0 and r <= 1e-10


def test_navier_stokes_es33():
    from fluxo_amd.native import Engine
    m = O.box_mesh(3, 3, x0=(-math.pi,) * 3, length=(2 * math.pi,) * 3, Ngeo=3, warp=True)
    U = O.exactfunc(7, m["Elem_xGP"])
    kw = dict(parabolic=True, mu0=0.01, Pr=0.72)
    e = Engine(m, volume_flux=32, riemann=33, **kw)
    e.upload(U)
    err = rel_l2(e.rhs(), O.Problem(m, volflux=32, riemann=33, **kw).rhs(U))
    print(f"NS Riemann 33: RHS relL2 {err:.3e}")
    assert err <= 1e-12


@pytest.mark.parametrize("N", [2, 3, 4, 5])
def test_ismail_roe_volume_flux(N):
    """VolumeFlux 10 (Ismail-Roe two-point flux) with LLF and with the Ismail-Roe ES surface
    flux (Riemann 4) against the oracle; the FP32-split variant's error stated at N=3."""
    from fluxo_amd.native import Engine, FX_PREC_FP32_SPLIT
    m = O.box_mesh(3, N, x0=(0, 0, 0), length=(2, 2, 2), Ngeo=3, warp=True)
    U = state(m)
    for rs in (1, 4):
        e = Engine(m, volume_flux=10, riemann=rs)
        e.upload(U)
        ref = O.Problem(m, volflux=10, riemann=rs).rhs(U)
        err = rel_l2(e.rhs(), ref)
        print(f"VolumeFlux 10 Riemann {rs} N={N}: RHS relL2 {err:.3e}")
        assert err <= 1e-12
    if N == 3:
        e = Engine(m, volume_flux=10, riemann=1, precision=FX_PREC_FP32_SPLIT)
        e.upload(U)
        err = rel_l2(e.rhs(), O.Problem(m, volflux=10, riemann=1).rhs(U))
        print(f"VolumeFlux 10 FP32-split N=3: RHS relL2 {err:.3e}")
        assert err <= 1e-4


@pytest.mark.parametrize("sign", [1.0, -1.0])
@pytest.mark.parametrize("rs", [2, 3, 22, 23, 24])
def test_hll_family_and_roe_supersonic(rs, sign):
    """HLLC, Roe, HLL, HLLE and HLLEM where the flow crosses every face supersonically (density
    wave advected with velocity 3 (1,1,1) or -3 (1,1,1), sound speed about 1.2): the upwind
    branches (all waves to one side: the left state's flux for +, the right state's for -) are
    taken at every face point; RHS against the oracle, whose upwind branches are pinned by the
    physical flux in test_oracle_riemann."""
    from fluxo_amd.native import Engine
    m = O.box_mesh(3, 3, x0=(0, 0, 0), length=(2, 2, 2), Ngeo=3, warp=True)
    U = O.exactfunc(2, m["Elem_xGP"], advvel=(3.0 * sign,) * 3)
    e = Engine(m, volume_flux=32, riemann=rs)
    e.upload(U)
    err = rel_l2(e.rhs(), O.Problem(m, volflux=32, riemann=rs).rhs(U))
    print(f"Riemann {rs} supersonic sign {sign:+.0f}: RHS relL2 {err:.3e}")
    assert err <= 1e-12
