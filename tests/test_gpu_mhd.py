"""GLM-MHD (ideal and resistive) on the GPU against the CPU oracle and the reference's own
outputs (tests/golden/mhd), through the C ABI.  FP64 bar: relative L2 <= 1e-12 per RHS and
for the lifted gradients, <= 1e-10 after RK steps; dt and the cleaning speed GLM_ch to
1e-14.  The FP32-split variant states its error against the FP64 oracle."""
import glob
import os

import numpy as np
import pytest

import pyoracle as O

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "mhd")
NAMES = sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLD, "*.npz")))


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


RES = dict(mu=0.01, eta=0.02, Pr=0.72)


def state(m, seed=3, resistive=None):
    P = O.MhdProblem(m, resistive=resistive)
    U = P.exact_orszag_tang()
    rng = np.random.default_rng(seed)
    U[..., 0] *= 1.0 + 1e-3 * rng.uniform(-1, 1, U[..., 0].shape)
    U[..., 5:8] *= 1.0 + 1e-3 * rng.uniform(-1, 1, U[..., 5:8].shape)
    U[..., 8] = 1e-2 * rng.uniform(-1, 1, U[..., 8].shape)
    return P, U


@pytest.mark.parametrize("N", [2, 3, 4, 5])
def test_rhs_matches_oracle(N):
    from fluxo_amd.native import MhdEngine
    m = O.box_mesh(3, N, x0=(0, 0, 0), length=(1, 1, 1), Ngeo=3, warp=True)
    P, U = state(m)
    e = MhdEngine(m)
    e.upload(U)
    assert e.glm()[0] == 0.5
    err = rel_l2(e.rhs(), P.rhs(U, glm_ch=0.5))
    print(f"MHD N={N}: RHS relL2 {err:.3e}")
    assert err <= 1e-12


@pytest.mark.parametrize("riemann", [0, 10, 11])
@pytest.mark.parametrize("N", [2, 3, 4, 5])
def test_surface_fluxes_match_oracle(N, riemann):
    """Central, Derigs + entropy-stable LLF and Derigs EC surface fluxes: the kernel forms
    them in the Cartesian frame, the oracle on rotated states like the reference."""
    from fluxo_amd.native import MhdEngine
    m = O.box_mesh(3, N, x0=(0, 0, 0), length=(1, 1, 1), Ngeo=3, warp=True)
    _, U = state(m)
    U[..., 0] *= 1.0 + 5e-2 * np.random.default_rng(N).uniform(-1, 1, U[..., 0].shape)
    P = O.MhdProblem(m, riemann=riemann)
    e = MhdEngine(m, riemann=riemann)
    e.upload(U)
    err = rel_l2(e.rhs(), P.rhs(U, glm_ch=0.5))
    print(f"MHD N={N} Riemann={riemann}: RHS relL2 {err:.3e}")
    assert err <= 1e-12


@pytest.mark.parametrize("riemann", [11, 10])
def test_entropy_balance_ec_and_es_surface_fluxes(riemann):
    """The reference's softBlast entropyCons / entropyStab checks (tests/job_def.py:519-548)
    on the engine's RHS: psi = 0, jumps at every face; Riemann 11 conserves the discrete
    entropy to round-off, Riemann 10 dissipates it."""
    from fluxo_amd.native import MhdEngine
    m = O.box_mesh(3, 3, x0=(0, 0, 0), length=(1, 1, 1), Ngeo=3, warp=True)
    U = O.MhdProblem(m).exact_orszag_tang()
    U[..., 8] = 0.0
    rng = np.random.default_rng(6)
    U[..., 0] *= 1.0 + 1e-2 * rng.uniform(-1, 1, U[..., 0].shape)
    U[..., 5:8] *= 1.0 + 1e-2 * rng.uniform(-1, 1, U[..., 5:8].shape)
    e = MhdEngine(m, riemann=riemann)
    e.upload(U)
    Ut = e.rhs()
    kappa = 5.0 / 3.0
    rho, v, B = U[..., 0], U[..., 1:4] / U[..., :1], U[..., 5:8]
    p = (kappa - 1) * (U[..., 4] - 0.5 * rho * (v * v).sum(-1) - 0.5 * (B * B).sum(-1))
    beta = 0.5 * rho / p
    w = np.zeros_like(U)
    w[..., 0] = (kappa - (np.log(p) - kappa * np.log(rho))) / (kappa - 1) - beta * (v * v).sum(-1)
    w[..., 1:4] = 2 * beta[..., None] * v
    w[..., 4] = -2 * beta
    w[..., 5:8] = 2 * beta[..., None] * B
    wl = O.dg_operators(3)["w"]
    W = np.einsum("i,j,k->kji", wl, wl, wl)[None] / m["sJ"]
    prod = float((W * (w * Ut).sum(-1)).sum())
    scale = float((W * np.abs(w * Ut).sum(-1)).sum())
    print(f"MHD Riemann={riemann}: entropy production {prod:.3e} (scale {scale:.3e})")
    if riemann == 11:
        assert abs(prod) <= 1e-13 * scale
    else:
        assert prod < -1e-6 * scale


@pytest.mark.parametrize("N", [2, 3, 4, 5])
def test_resistive_br2_matches_oracle(N):
    """BR2 lifting (etaBR2 = 2.5): volume gradients and RHS against the oracle."""
    from fluxo_amd.native import MhdEngine
    m = O.box_mesh(3, N, x0=(0, 0, 0), length=(1, 1, 1), Ngeo=3, warp=True)
    _, U = state(m, resistive=RES)
    P = O.MhdProblem(m, resistive=RES, lifting=2, etaBR2=2.5)
    e = MhdEngine(m, resistive=RES, lifting=2, etaBR2=2.5)
    e.upload(U)
    gerr = max(rel_l2(a, b) for a, b in zip(e.lifting(), P.lifting(U)))
    err = rel_l2(e.rhs(), P.rhs(U, glm_ch=0.5))
    print(f"resistive MHD BR2 N={N}: gradients relL2 {gerr:.3e}, RHS relL2 {err:.3e}")
    assert gerr <= 1e-12 and err <= 1e-12


@pytest.mark.parametrize("N", [2, 3, 4, 5])
def test_resistive_rhs_and_gradients_match_oracle(N):
    from fluxo_amd.native import MhdEngine
    m = O.box_mesh(3, N, x0=(0, 0, 0), length=(1, 1, 1), Ngeo=3, warp=True)
    P, U = state(m, resistive=RES)
    e = MhdEngine(m, resistive=RES)
    e.upload(U)
    gerr = max(rel_l2(a, b) for a, b in zip(e.lifting(), P.lifting(U)))
    err = rel_l2(e.rhs(), P.rhs(U, glm_ch=0.5))
    print(f"resistive MHD N={N}: gradients relL2 {gerr:.3e}, RHS relL2 {err:.3e}")
    assert gerr <= 1e-12 and err <= 1e-12


@pytest.mark.parametrize("name", NAMES)
def test_engine_matches_reference_mhd(name):
    from fluxo_amd.mesh import box_mesh
    from fluxo_amd.native import MhdEngine
    with np.load(os.path.join(GOLD, name + ".npz"), allow_pickle=False) as z:
        g = {k: z[k] for k in z.files}
    m = box_mesh(int(g["K"]), int(g["N"]), x0=tuple(g["x0"]), length=tuple(g["L"]), Ngeo=int(g["Ngeo"]),
                 warp=bool(g["warp"]))
    res = dict(mu=float(g["mu"]), eta=float(g["eta"]), Pr=float(g["Pr"])) if "mu" in g else None
    dfl = float(g["dfl"]) if res else 0.0
    kw = dict(kappa=float(g["kappa"]), mu_0=float(g["mu_0"]), glm_scale=float(g["glm_scale"]), resistive=res,
              riemann=int(g["riemann"]) if "riemann" in g else 1,
              lifting=int(g["lifting"]) if "lifting" in g else 1,
              etaBR2=float(g["etaBR2"]) if "etaBR2" in g else 2.0)
    e = MhdEngine(m, **kw)
    e.upload(g["U"])
    if res:
        gerr = max(rel_l2(a, g[k]) for a, k in zip(e.lifting(), ("gradPx", "gradPy", "gradPz")))
        print(f"{name}: gradients relL2 vs reference {gerr:.3e}")
        assert gerr <= 1e-12
    err = rel_l2(e.rhs(), g["Ut"])
    print(f"{name}: RHS relL2 vs reference {err:.3e}")
    assert err <= 1e-12
    dt, errt = e.calc_dt(float(g["cfl"]), dfl)
    ch, dtch1 = e.glm()
    assert errt == int(g["errType0"]) and dt == pytest.approx(float(g["dt0"]), rel=1e-14)
    assert dtch1 == pytest.approx(float(g["glm_dtch1"]), rel=1e-14)
    assert ch == pytest.approx(float(g["glm_ch0"]), rel=1e-14)
    e2 = MhdEngine(m, **kw)
    e2.upload(g["U"])
    t, _, errt = e2.run(int(g["nsteps"]), float(g["cfl"]), dfl)
    assert errt == 0
    traj = rel_l2(e2.download(), g["U_end"])
    print(f"{name}: {int(g['nsteps'])} steps relL2 vs reference {traj:.3e}")
    assert traj <= 1e-10
    assert t == pytest.approx(float(g["t_end"]), rel=1e-12)


def test_free_stream_warped():
    from fluxo_amd.native import MhdEngine
    m = O.box_mesh(3, 4, Ngeo=3, warp=True)
    e = MhdEngine(m)
    U = np.zeros(e.shape())
    U[...] = [1.1, 0.2, -0.1, 0.3, 4.0, 0.5, -0.4, 0.2, 0.0]
    e.upload(U)
    assert np.abs(e.rhs()).max() < 1e-11


def test_resistive_trajectory_matches_oracle():
    from fluxo_amd.native import MhdEngine
    m = O.box_mesh(3, 4, x0=(0, 0, 0), length=(1, 1, 1), Ngeo=3, warp=True)
    P, U = state(m, resistive=RES)
    Ue, Utmp = U.copy(), np.zeros_like(U)
    for _ in range(10):
        d, _ = P.calc_dt(Ue, 0.9, 0.9)
        P.lserk_step(Ue, Utmp, d)
    e = MhdEngine(m, resistive=RES)
    e.upload(U)
    _, _, errt = e.run(10, 0.9, 0.9)
    err = rel_l2(e.download(), Ue)
    print(f"resistive MHD N=4: 10 steps relL2 {err:.3e}")
    assert errt == 0 and err <= 1e-10


@pytest.mark.parametrize("N", [3, 4])
def test_fp32split_error_vs_fp64(N):
    """FP32 flux differencing for GLM-MHD: RHS and a 5-step trajectory against the FP64
    oracle (relative L2 / Linf printed; bounds about 3x the measured values)."""
    from fluxo_amd.native import MhdEngine, FX_PREC_FP32_SPLIT
    m = O.box_mesh(3, N, x0=(0, 0, 0), length=(1, 1, 1), Ngeo=3, warp=True)
    P, U = state(m)
    e = MhdEngine(m, precision=FX_PREC_FP32_SPLIT)
    e.upload(U)
    ref = P.rhs(U, glm_ch=0.5)
    got = e.rhs()
    l2 = rel_l2(got, ref)
    li = float(np.max(np.abs(got - ref)) / np.max(np.abs(ref)))
    Ue, Utmp = U.copy(), np.zeros_like(U)
    for _ in range(5):
        d, _ = P.calc_dt(Ue, 0.9)
        P.lserk_step(Ue, Utmp, d)
    e2 = MhdEngine(m, precision=FX_PREC_FP32_SPLIT)
    e2.upload(U)
    e2.run(5, 0.9)
    tl2 = rel_l2(e2.download(), Ue)
    print(f"MHD FP32-split N={N}: RHS relL2 {l2:.3e} relLinf {li:.3e}; 5 steps relL2 {tl2:.3e}")
    assert l2 <= 1e-4 and tl2 <= 1e-5


def test_resistive_fp32split_error_vs_fp64():
    """FP32-split resistive GLM-MHD at N=4 (the 24^3 configuration's degree): RHS and a
    5-step trajectory against the FP64 oracle (bounds about 3x the measured values)."""
    from fluxo_amd.native import MhdEngine, FX_PREC_FP32_SPLIT
    m = O.box_mesh(3, 4, x0=(0, 0, 0), length=(1, 1, 1), Ngeo=3, warp=True)
    P, U = state(m, resistive=RES)
    e = MhdEngine(m, precision=FX_PREC_FP32_SPLIT, resistive=RES)
    e.upload(U)
    l2 = rel_l2(e.rhs(), P.rhs(U, glm_ch=0.5))
    Ue, Utmp = U.copy(), np.zeros_like(U)
    for _ in range(5):
        d, _ = P.calc_dt(Ue, 0.9, 0.9)
        P.lserk_step(Ue, Utmp, d)
    e2 = MhdEngine(m, precision=FX_PREC_FP32_SPLIT, resistive=RES)
    e2.upload(U)
    e2.run(5, 0.9, 0.9)
    tl2 = rel_l2(e2.download(), Ue)
    print(f"resistive MHD FP32-split N=4: RHS relL2 {l2:.3e}; 5 steps relL2 {tl2:.3e}")
    assert l2 <= 1e-4 and tl2 <= 1e-5
