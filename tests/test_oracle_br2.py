"""BR2 lifting in the oracle (the reference's PP_Lifting=2 path), beyond its golden vectors
(tests/golden/ns_*_br2_*): the BR2 volume gradient is the BR1 gradient with sJ distributed
over the terms; the right-hand side is affine in the penalty etaBR2 (the face gradients are,
and the viscous fluxes are linear in the gradients at a fixed state); a constant state stays
constant on a warped mesh."""
import math

import numpy as np

import pyoracle as O


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def noisy_tgv(m, seed=4):
    U = O.exactfunc(7, m["Elem_xGP"])
    rng = np.random.default_rng(seed)
    return U * (1.0 + 1e-2 * rng.uniform(-1.0, 1.0, U.shape))


KW = dict(volflux=32, riemann=1, parabolic=True, mu0=0.01, Pr=0.72)


def mesh(N=3, K=3):
    return O.box_mesh(K, N, x0=(-math.pi,) * 3, length=(2 * math.pi,) * 3, Ngeo=3, warp=True)


def test_br2_volume_gradient_is_br1_gradient():
    m = mesh()
    U = noisy_tgv(m)
    g1 = O.Problem(m, lifting=1, **KW).lifting(U)
    g2 = O.Problem(m, lifting=2, etaBR2=2.0, **KW).lifting(U)
    for a, b in zip(g2, g1):
        assert rel_l2(a, b) <= 1e-13


def test_br2_rhs_affine_in_eta_and_distinct_from_br1():
    m = mesh(N=4, K=2)
    U = noisy_tgv(m)
    R = {eta: O.Problem(m, lifting=2, etaBR2=eta, **KW).rhs(U) for eta in (1.0, 2.0, 3.0)}
    d1, d2 = R[2.0] - R[1.0], R[3.0] - R[2.0]
    assert rel_l2(d2, d1) <= 1e-9
    assert np.linalg.norm(d1) > 1e-6 * np.linalg.norm(R[2.0])
    assert rel_l2(O.Problem(m, lifting=1, **KW).rhs(U), R[2.0]) > 1e-4


def test_br2_free_stream_warped():
    m = mesh()
    P = O.Problem(m, lifting=2, etaBR2=2.0, **KW)
    U = np.zeros(P.shape())
    U[...] = [1.2, 0.3, -0.2, 0.1, 3.0]
    assert np.abs(P.rhs(U)).max() < 1e-11
