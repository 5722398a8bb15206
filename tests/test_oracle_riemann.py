"""Surface fluxes of the oracle beyond the golden vectors (euler_*_es*, *_eckep_llf16_*):
for random state pairs and normals, the entropy-stable solvers (Riemann 4, 33, 44) and the
EC-KEP + LLF solver (16) dissipate entropy, [v].F <= [psi] with psi = rho v.n (Tadmor's
condition holds with equality for the EC fluxes 9 and 32), and for equal states every solver
returns the physical Euler flux."""
import numpy as np
import pytest

import pyoracle as O

KAPPA = 1.4


def random_state(rng):
    rho, p = rng.uniform(0.5, 2.0), rng.uniform(0.5, 2.0)
    v = rng.uniform(-1, 1, 3)
    return np.array([rho, *(rho * v), p / (KAPPA - 1) + 0.5 * rho * v @ v])


def frame(rng):
    n = rng.normal(size=3)
    n /= np.linalg.norm(n)
    a = rng.normal(size=3)
    t1 = a - (a @ n) * n
    t1 /= np.linalg.norm(t1)
    return n, t1, np.cross(n, t1)


def entropy_vars(U):
    w = np.zeros(5)
    O.lib().ora_cons_to_entropy(KAPPA, np.ascontiguousarray(U), w)
    return w


def flux(rs, UL, UR, n, t1, t2, lam=None):
    F = np.zeros(5)
    L = O.lib()
    if lam is None:
        L.ora_riemann(rs, KAPPA, UL, UR, n, t1, t2, F)
    else:
        L.ora_riemann_lam(rs, KAPPA, UL, UR, n, t1, t2, lam, F)
    return F


@pytest.mark.parametrize("rs", [4, 5, 9, 13, 16, 32, 33, 44])
def test_entropy_condition_and_consistency(rs):
    rng = np.random.default_rng(rs)
    for _ in range(30):
        UL, UR = random_state(rng), random_state(rng)
        n, t1, t2 = frame(rng)
        F = flux(rs, UL, UR, n, t1, t2, lam=2.5 if rs in (13, 16) else None)
        dv = entropy_vars(UR) - entropy_vars(UL)
        dpsi = UR[1:4] @ n - UL[1:4] @ n
        prod = dv @ F - dpsi
        if rs in (5, 9, 32):
            assert abs(prod) <= 1e-12 * (1 + abs(dpsi))
        else:
            assert prod < -1e-8
        # equal states: the physical flux
        Fe = flux(rs, UL, UL, n, t1, t2, lam=2.5 if rs in (13, 16) else None)
        v = UL[1:4] / UL[0]
        p = (KAPPA - 1) * (UL[4] - 0.5 * UL[0] * v @ v)
        vn = v @ n
        Fp = np.array([UL[0] * vn, *(UL[1:4] * vn + p * n), (UL[4] + p) * vn])
        assert np.allclose(Fe, Fp, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("vf", [5, 10, 32])
def test_volume_flux_tadmor_and_consistency(vf):
    """Two-point volume fluxes (Chandrashekar 5, Ismail-Roe 10, Ranocha 32) with a metric
    m (the same on both sides): Tadmor's condition [v].F#(m) = [rho v.m], symmetry, and
    F#(U,U) = f(U).m."""
    rng = np.random.default_rng(vf)
    for _ in range(30):
        UL, UR = random_state(rng), random_state(rng)
        m = rng.uniform(-1, 1, 3)
        F, Fs = O.volflux(vf, UL, UR, m, m, KAPPA), O.volflux(vf, UR, UL, m, m, KAPPA)
        dv = entropy_vars(UR) - entropy_vars(UL)
        dpsi = UR[1:4] @ m - UL[1:4] @ m
        assert abs(dv @ F - dpsi) <= 1e-12 * (1 + abs(dpsi))
        assert np.allclose(F, Fs, rtol=1e-13, atol=1e-13)
        Fe = O.volflux(vf, UL, UL, m, m, KAPPA)
        v = UL[1:4] / UL[0]
        p = (KAPPA - 1) * (UL[4] - 0.5 * UL[0] * v @ v)
        vm = v @ m
        assert np.allclose(Fe, [UL[0] * vm, *(UL[1:4] * vm + p * m), (UL[4] + p) * vm], rtol=1e-12, atol=1e-12)


KEP = {2: True, 3: False, 4: True, 6: True, 7: True, 8: True, 9: False, 1: False}


@pytest.mark.parametrize("vf", [1, 2, 3, 4, 6, 7, 8, 9])
def test_other_volume_fluxes_consistency_symmetry_kep(vf):
    """The other two-point volume fluxes (golden vectors euler_*_vf*_llf_*): symmetric under
    exchanging the sides with their metrics, F#(U,U,mL,mR) = f(U).{{m}}, and for the kinetic
    energy preserving ones (Kennedy-Gruber, Morinishi, EC with arithmetic means, gg) the
    momentum flux is F#_rho {{v}} + p~ {{m}} with one pressure average p~ for all three
    components."""
    rng = np.random.default_rng(100 + vf)
    for _ in range(30):
        UL, UR = random_state(rng), random_state(rng)
        mL, mR = rng.uniform(-1, 1, 3), rng.uniform(-1, 1, 3)
        m = 0.5 * (mL + mR)
        F = O.volflux(vf, UL, UR, mL, mR, KAPPA)
        assert np.allclose(F, O.volflux(vf, UR, UL, mR, mL, KAPPA), rtol=1e-13, atol=1e-13)
        Fe = O.volflux(vf, UL, UL, mL, mR, KAPPA)
        v = UL[1:4] / UL[0]
        p = (KAPPA - 1) * (UL[4] - 0.5 * UL[0] * v @ v)
        vm = v @ m
        assert np.allclose(Fe, [UL[0] * vm, *(UL[1:4] * vm + p * m), (UL[4] + p) * vm], rtol=1e-12, atol=1e-12)
        if KEP[vf]:
            va = 0.5 * (UL[1:4] / UL[0] + UR[1:4] / UR[0])
            pt = (F[1:4] - F[0] * va) / m
            assert np.ptp(pt) <= 1e-12 * np.abs(pt).max(), pt


LLF_OF = {14: 6, 15: 7, 17: 11, 18: 12, 19: 8, 21: 20}


@pytest.mark.parametrize("rs", [6, 7, 8, 10, 11, 12, 20, 14, 15, 17, 18, 19, 21])
def test_volume_flux_average_solvers(rs):
    """Riemann 6-8, 10-12, 20 (a two-point flux in the normal direction) and their LLF forms
    14, 15, 17-19, 21: the physical flux for equal states, independent of the choice of the
    tangents, symmetric for the pure averages (F(L,R;n) = F(R,L;n)), and the LLF form is the
    average minus lam/2 [U] with the face's lam."""
    rng = np.random.default_rng(200 + rs)
    for _ in range(20):
        UL, UR = random_state(rng), random_state(rng)
        n, t1, t2 = frame(rng)
        F = flux(rs, UL, UR, n, t1, t2, lam=1.7)
        _, s1, s2 = frame(rng)
        s1 = s1 - (s1 @ n) * n
        s1 /= np.linalg.norm(s1)
        assert np.allclose(F, flux(rs, UL, UR, n, s1, np.cross(n, s1), lam=1.7), rtol=1e-12, atol=1e-12)
        if rs in LLF_OF:
            Fa = flux(LLF_OF[rs], UL, UR, n, t1, t2)
            assert np.allclose(F, Fa - 0.85 * (UR - UL), rtol=1e-13, atol=1e-13)
        else:
            assert np.allclose(F, flux(rs, UR, UL, n, t1, t2), rtol=1e-13, atol=1e-13)
        Fe = flux(rs, UL, UL, n, t1, t2, lam=1.7)
        v = UL[1:4] / UL[0]
        p = (KAPPA - 1) * (UL[4] - 0.5 * UL[0] * v @ v)
        vn = v @ n
        assert np.allclose(Fe, [UL[0] * vn, *(UL[1:4] * vn + p * n), (UL[4] + p) * vn], rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("rs", [2, 3, 22, 23, 24])
def test_hll_and_roe_solvers(rs):
    """HLLC (2), Roe (3), HLL (22), HLLE (23), HLLEM (24): the physical flux for equal states,
    independence of the tangent choice, and for the upwind cases (supersonic flow through the
    face) the upwind side's physical flux for the HLL family."""
    rng = np.random.default_rng(300 + rs)
    for _ in range(20):
        UL, UR = random_state(rng), random_state(rng)
        n, t1, t2 = frame(rng)
        F = flux(rs, UL, UR, n, t1, t2)
        _, s1, _ = frame(rng)
        s1 = s1 - (s1 @ n) * n
        s1 /= np.linalg.norm(s1)
        assert np.allclose(F, flux(rs, UL, UR, n, s1, np.cross(n, s1)), rtol=1e-12, atol=1e-12)
        Fe = flux(rs, UL, UL, n, t1, t2)
        v = UL[1:4] / UL[0]
        p = (KAPPA - 1) * (UL[4] - 0.5 * UL[0] * v @ v)
        vn = v @ n
        assert np.allclose(Fe, [UL[0] * vn, *(UL[1:4] * vn + p * n), (UL[4] + p) * vn], rtol=1e-12, atol=1e-12)
    if rs != 3:  # supersonic along n: every wave moves right, the flux is the left state's
        UL = np.array([1.0, 5.0 * n[0], 5.0 * n[1], 5.0 * n[2], 1.0 / (KAPPA - 1) + 12.5])
        UR = np.array([1.1, 5.2 * n[0], 5.2 * n[1], 5.2 * n[2], 1.1 / (KAPPA - 1) + 0.55 * 27.04])
        vn = 5.0
        FL = [UL[0] * vn, *(UL[1:4] * vn + 1.0 * n), (UL[4] + 1.0) * vn]
        assert np.allclose(flux(rs, UL, UR, n, t1, t2), FL, rtol=1e-12, atol=1e-12)
        # supersonic against n: every wave moves left, the flux is the right state's
        UL = np.array([1.1, -5.2 * n[0], -5.2 * n[1], -5.2 * n[2], 1.1 / (KAPPA - 1) + 0.55 * 27.04])
        UR = np.array([1.0, -5.0 * n[0], -5.0 * n[1], -5.0 * n[2], 1.0 / (KAPPA - 1) + 12.5])
        vn = -5.0
        FR = [UR[0] * vn, *(UR[1:4] * vn + 1.0 * n), (UR[4] + 1.0) * vn]
        assert np.allclose(flux(rs, UL, UR, n, t1, t2), FR, rtol=1e-12, atol=1e-12)
