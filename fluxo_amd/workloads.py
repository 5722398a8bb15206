"""Synthetic workloads: initial states and the mixed serving request stream.

Initial states follow the reference's ExactFunc (navierstokes/equation.f90):
  2 density wave      (:632-651)  rho = 1 + A sin(pi (x + y + z - |AdvVel|_1 t)), v = AdvVel, p = 1
  7 Taylor-Green      (:802-811)  Ma = 0.1, rho = 1
Arrays are numpy (nElems, n, n, n, 5), the memory layout of the reference's
U(5,0:N,0:N,0:N,nElems).

The serving stream: request s draws, from a generator seeded with (seed, s),
  N uniform in {2,3,4,5}; K = round(exp(U(ln 4, ln 24))) elements per direction;
  s even: density wave on [0,2]^3 (Euler);
  s odd : Taylor-Green vortex on [-pi,pi]^3, inviscid for s = 1 mod 4, Navier-Stokes with
          BR1 (Re 1600, Pr 0.72) for s = 3 mod 4;
  periodic Cartesian K^3 mesh (Ngeo 1); nSteps time steps each.
"""
import math

import numpy as np

from .mesh import box_mesh


def density_wave(xgp, kappa=1.4, advvel=(1.0, 1.0, 1.0), amp=0.1, t=0.0):
    x = xgp[..., 0] - advvel[0] * t
    y = xgp[..., 1] - advvel[1] * t
    z = xgp[..., 2] - advvel[2] * t
    rho = 1.0 * (1.0 + amp * np.sin(np.pi * (1.0 * x + 1.0 * y + 1.0 * z)))
    U = np.empty(xgp.shape[:-1] + (5,))
    U[..., 0] = rho
    U[..., 1], U[..., 2], U[..., 3] = rho * advvel[0], rho * advvel[1], rho * advvel[2]
    U[..., 4] = 1.0 / (kappa - 1.0) + 0.5 * rho * (advvel[0] ** 2 + advvel[1] ** 2 + advvel[2] ** 2)
    return U


def taylor_green(xgp, kappa=1.4):
    x, y, z = xgp[..., 0], xgp[..., 1], xgp[..., 2]
    A, Ms = 1.0, 0.1
    rho = np.ones_like(x)
    u = A * np.sin(x) * np.cos(y) * np.cos(z)
    v = -A * np.cos(x) * np.sin(y) * np.cos(z)
    w = np.zeros_like(x)
    p = (A / Ms * A / Ms / kappa * rho)
    p = p + 1.0 / 16.0 * A * A * rho * (np.cos(2 * x) * np.cos(2.0 * z) + 2.0 * np.cos(2.0 * y)
                                        + 2.0 * np.cos(2.0 * x) + np.cos(2 * y) * np.cos(2.0 * z))
    U = np.empty(x.shape + (5,))
    U[..., 0] = rho
    U[..., 1], U[..., 2], U[..., 3] = rho * u, rho * v, rho * w
    U[..., 4] = p / (kappa - 1.0) + 0.5 * (U[..., 1] * u + U[..., 2] * v + U[..., 3] * w)
    return U


DOMAINS = {"wave": ((0.0,) * 3, (2.0,) * 3), "tgv": ((-math.pi,) * 3, (2 * math.pi,) * 3)}


def request_params(seed, s):
    """(N, K, ic, parabolic) of request s of the stream seeded with `seed`."""
    rng = np.random.default_rng([seed, s])
    N = int(rng.integers(2, 6))
    K = int(round(math.exp(rng.uniform(math.log(4.0), math.log(24.0)))))
    if s % 2 == 0:
        return N, K, "wave", False
    return N, K, "tgv", s % 4 == 3


class MeshCache:
    """Periodic Cartesian box meshes (metrics from the Fortran host), shared by every
    request with the same (K, N, domain)."""

    def __init__(self):
        self._m = {}

    def get(self, K, N, ic):
        key = (K, N, ic)
        if key not in self._m:
            x0, L = DOMAINS[ic]
            self._m[key] = box_mesh(K, N, x0=x0, length=L, Ngeo=1, warp=False)
        return self._m[key]


def make_request(seed, s, cache, nsteps=50):
    """Inputs of request s: dict(N, K, ic, parabolic, mesh, U0, nsteps, nodes)."""
    N, K, ic, par = request_params(seed, s)
    mesh = cache.get(K, N, ic)
    U0 = density_wave(mesh["Elem_xGP"]) if ic == "wave" else taylor_green(mesh["Elem_xGP"])
    return dict(N=N, K=K, ic=ic, parabolic=par, mesh=mesh, U0=U0, nsteps=nsteps,
                nodes=mesh["nElems"] * (N + 1) ** 3)


SEDOV_SIGMA = 0.06
SEDOV_E0 = 1.0


def sedov_blast(xgp, kappa=1.4, sigma=SEDOV_SIGMA, E0=SEDOV_E0, p0=1.0e-5, center=(0.0, 0.0, 0.0)):
    """3D Sedov-type blast: rho = 1, v = 0, p = p0 + (kappa-1) E0 G(r) with the normalised
    Gaussian G(r) = exp(-r^2/(2 sigma^2)) / ((2 pi)^(3/2) sigma^3), so the pulse carries
    total internal energy E0 (sigma = 0.06, E0 = 1, p0 = 1e-5 on [-1.5,1.5]^3)."""
    r2 = sum((xgp[..., d] - center[d]) ** 2 for d in range(3))
    g = np.exp(-r2 / (2.0 * sigma * sigma)) / ((2.0 * np.pi) ** 1.5 * sigma ** 3)
    p = p0 + (kappa - 1.0) * E0 * g
    U = np.zeros(xgp.shape[:-1] + (5,))
    U[..., 0] = 1.0
    U[..., 4] = p / (kappa - 1.0)
    return U


def orszag_tang(xgp, kappa=5.0 / 3.0, mu_0=1.0):
    """3D Orszag-Tang-type vortex of the reference's ExactFunc 312 (mhd/equation.f90:1257-1266),
    converted with PrimToCons (mhd/equation_vars.f90:258-278): rho = 1, p = 1/kappa,
    v = (-(1 + 0.2 sin 2 pi z) sin 2 pi y, (1 + 0.2 sin 2 pi z) sin 2 pi x, 0.2 sin 2 pi z),
    B = (-sin 2 pi y, sin 4 pi x, 0) / kappa, psi = 0.  Returns (..., 9) conservative."""
    x, y, z = xgp[..., 0], xgp[..., 1], xgp[..., 2]
    tp = 2.0 * np.pi
    a = 1.0 + 0.2 * np.sin(tp * z)
    v = np.stack([-a * np.sin(tp * y), a * np.sin(tp * x), 0.2 * np.sin(tp * z)], -1)
    B = np.stack([-np.sin(tp * y) / kappa, np.sin(2.0 * tp * x) / kappa, np.zeros_like(x)], -1)
    U = np.zeros(xgp.shape[:-1] + (9,))
    U[..., 0] = 1.0
    U[..., 1:4] = v
    U[..., 5:8] = B
    U[..., 4] = (1.0 / (kappa - 1.0)) * (1.0 / kappa) + 0.5 * (v * v).sum(-1) + 0.5 * (1.0 / mu_0) * (B * B).sum(-1)
    return U


def perturbed_orszag_tang(xgp, seed, amp=1.0e-3, kappa=5.0 / 3.0, mu_0=1.0):
    """orszag_tang with rho and B multiplied by 1 + amp*uniform(-1,1) (numpy default_rng(seed)),
    the total energy recomputed for the unchanged pressure and velocity."""
    U = orszag_tang(xgp, kappa, mu_0)
    rng = np.random.default_rng(seed)
    v = U[..., 1:4] / U[..., :1]
    p = (kappa - 1.0) * (U[..., 4] - 0.5 * U[..., 0] * (v * v).sum(-1) - 0.5 / mu_0 * (U[..., 5:8] ** 2).sum(-1))
    U[..., 0] *= 1.0 + amp * rng.uniform(-1, 1, U[..., 0].shape)
    U[..., 5:8] *= 1.0 + amp * rng.uniform(-1, 1, U[..., 5:8].shape)
    U[..., 1:4] = U[..., :1] * v
    U[..., 4] = p / (kappa - 1.0) + 0.5 * U[..., 0] * (v * v).sum(-1) + 0.5 / mu_0 * (U[..., 5:8] ** 2).sum(-1)
    return U
