"""BR2 lifting on the GPU (lifting=2) against the oracle: volume gradients and the RHS
(FP64 bar: relative L2 <= 1e-12, the low-Mach gradients at the golden cases' 1e-12),
a trajectory (<= 1e-10), and the FP32-split variant's stated error.  The reference's own BR2
outputs are checked in test_gpu_golden (ns_*_br2_* cases)."""
import math

import numpy as np
import pytest

import pyoracle as O

pytestmark = pytest.mark.gpu

KW = dict(parabolic=True, mu0=0.01, Pr=0.72)


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def setup(N, K=3, eta=2.0):
    m = O.box_mesh(K, N, x0=(-math.pi,) * 3, length=(2 * math.pi,) * 3, Ngeo=3, warp=True)
    U = O.exactfunc(7, m["Elem_xGP"])
    rng = np.random.default_rng(11)
    U = U * (1.0 + 1e-2 * rng.uniform(-1.0, 1.0, U.shape))
    P = O.Problem(m, volflux=32, riemann=1, lifting=2, etaBR2=eta, **KW)
    return m, U, P


@pytest.mark.parametrize("N", [2, 3, 4, 5])
def test_br2_rhs_and_gradients(N):
    from fluxo_amd.native import Engine
    m, U, P = setup(N, eta=2.0 if N != 4 else 3.5)
    e = Engine(m, volume_flux=32, riemann=1, lifting=2, etaBR2=2.0 if N != 4 else 3.5, **KW)
    e.upload(U)
    g = max(rel_l2(a, b) for a, b in zip(e.lifting(), P.lifting(U)))
    r = rel_l2(e.rhs(), P.rhs(U))
    print(f"BR2 N={N}: gradients relL2 {g:.3e}, RHS relL2 {r:.3e}")
    assert g <= 1e-12 and r <= 1e-12


def test_br2_trajectory():
    from fluxo_amd.native import Engine
    m, U, P = setup(3)
    Ue, Utmp, t = U.copy(), np.zeros_like(U), 0.0
    for _ in range(10):
        d, _ = P.calc_dt(Ue, 0.9, 0.9)
        P.lserk_step(Ue, Utmp, t, d)
        t += d
    e = Engine(m, volume_flux=32, riemann=1, lifting=2, etaBR2=2.0, **KW)
    e.upload(U)
    te, _, err = e.run(10, 0.9, 0.9)
    r = rel_l2(e.download(), Ue)
    print(f"BR2 N=3: 10 steps relL2 {r:.3e}")
    assert err == 0 and r <= 1e-10 and te == pytest.approx(t, rel=1e-13)


def test_br2_fp32split_error():
    """FP32-split with BR2 (the lifting stays FP64): RHS and 5 steps against the FP64
    oracle; bounds about 3x the FP32-split Navier-Stokes errors."""
    from fluxo_amd.native import Engine, FX_PREC_FP32_SPLIT
    m, U, P = setup(3)
    e = Engine(m, volume_flux=32, riemann=1, lifting=2, etaBR2=2.0, precision=FX_PREC_FP32_SPLIT, **KW)
    e.upload(U)
    r = rel_l2(e.rhs(), P.rhs(U))
    Ue, Utmp, t = U.copy(), np.zeros_like(U), 0.0
    for _ in range(5):
        d, _ = P.calc_dt(Ue, 0.9, 0.9)
        P.lserk_step(Ue, Utmp, t, d)
        t += d
    e.upload(U)
    e.run(5, 0.9, 0.9)
    tr = rel_l2(e.download(), Ue)
    print(f"BR2 FP32-split N=3: RHS relL2 {r:.3e}; 5 steps relL2 {tr:.3e}")
    assert r <= 1e-4 and tr <= 1e-5
