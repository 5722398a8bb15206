"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: the oracle is a plain-C restatement of the reference
(project-fluxo/fluxo) split-form DGSEM path.  Only tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg may import this module, and only as the checker or
the reported CPU baseline.

Array shapes use C order with the reference's Fortran index order reversed, so the
memory layout equals the reference's column-major arrays:
  U(nVar,0:N,0:N,0:N,nElems)  <->  numpy (nElems, n, n, n, nVar)
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.ora_ln_mean.restype = C.c_double
        L.ora_ln_mean.argtypes = [C.c_double, C.c_double]
        L.ora_box_connectivity.restype = C.c_int
        L.ora_box_connectivity.argtypes = [C.c_int, C.c_int, C.c_int, _ip, _ip]
        L.ora_box_nodecoords.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, C.c_int, C.c_int,
                                         C.c_double, _dp]
        L.ora_calc_metrics.restype = C.c_int
        L.ora_calc_metrics.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _dp, _ip] + [_dp] * 9
        L.ora_calc_timestep.restype = C.c_double
        L.ora_calc_timestep.argtypes = [C.c_void_p, _dp, C.c_double, C.c_double,
                                        C.POINTER(C.c_int)]
        L.ora_dg_time_derivative.argtypes = [C.c_void_p, _dp, C.c_double, _dp]
        L.ora_volint_splitform.argtypes = [C.c_void_p, _dp, _dp]
        L.ora_lifting.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
        L.ora_lserk_step.argtypes = [C.c_void_p, _dp, _dp, C.c_double, C.c_double]
        L.ora_calc_entropy.argtypes = [C.c_void_p, _dp, _dp, C.POINTER(C.c_double),
                                       C.POINTER(C.c_double)]
        L.ora_exactfunc.argtypes = [C.c_int, C.c_double, _dp, C.c_double, C.c_double, C.c_int,
                                    _dp, _dp]
        L.ora_volflux.argtypes = [C.c_int, C.c_double] + [_dp] * 7
        L.ora_riemann.argtypes = [C.c_int, C.c_double] + [_dp] * 6
        L.ora_riemann_lam.argtypes = [C.c_int, C.c_double] + [_dp] * 5 + [C.c_double, _dp]
        L.ora_cons_to_entropy.argtypes = [C.c_double, _dp, _dp]
        L.ora_rk_coefs.argtypes = [_dp, _dp, _dp]
        L.ora_lgl_nodes.argtypes = [C.c_int, _dp, _dp]
        L.ora_dg_operators.argtypes = [C.c_int] + [_dp] * 7
        L.ora_mhd_rhs.argtypes = [C.c_void_p, _dp, C.c_double, _dp]
        L.ora_mhd_glm_dtch1.restype = C.c_double
        L.ora_mhd_glm_dtch1.argtypes = [C.c_void_p, C.c_double]
        L.ora_mhd_calc_timestep.restype = C.c_double
        L.ora_mhd_calc_timestep.argtypes = [C.c_void_p, _dp, C.c_double, C.c_double, C.c_double,
                                            C.POINTER(C.c_double), C.POINTER(C.c_int)]
        L.ora_mhd_lifting.argtypes = [C.c_void_p] + [_dp] * 4
        L.ora_mhd_lserk_step.argtypes = [C.c_void_p, _dp, _dp, C.c_double, C.c_double]
        L.ora_mhd_orszag_tang.argtypes = [C.c_void_p, C.c_int, _dp, _dp]
        L.ora_mhd_derigs.argtypes = [C.c_void_p, C.c_double] + [_dp] * 5
        L.ora_mhd_riemann.argtypes = [C.c_void_p, C.c_double] + [_dp] * 6
        L.ora_nfvse_threshold.restype = C.c_double
        L.ora_nfvse_threshold.argtypes = [C.c_int]
        L.ora_subcell_metrics.argtypes = [C.c_int, _dp, _dp, _dp, _dp]
        L.ora_persson_peraire.argtypes = [C.c_void_p, _dp, _dp]
        L.ora_blending_coefficient.argtypes = [C.c_void_p, _dp, _dp]
        L.ora_volint_nfvse.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.ora_fill_flux.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.ora_surf_int.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp]
        L.ora_mhd_fill_flux.argtypes = [C.c_void_p, _dp, C.c_double, _dp, _dp]
        L.ora_analyze.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp, C.c_void_p, C.c_void_p,
                                  C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_void_p, C.c_void_p]
        _LIB = L
    return _LIB


class _Problem(C.Structure):
    _fields_ = [("N", C.c_int), ("nVar", C.c_int), ("nElems", C.c_int), ("nSides", C.c_int),
                ("volflux", C.c_int), ("riemann", C.c_int), ("parabolic", C.c_int),
                ("lifting", C.c_int),
                ("kappa", C.c_double), ("mu0", C.c_double), ("Pr", C.c_double),
                ("etaBR2", C.c_double),
                ("ElemToSide", C.c_void_p), ("SideToElem", C.c_void_p),
                ("Mf", C.c_void_p), ("Mg", C.c_void_p), ("Mh", C.c_void_p), ("sJ", C.c_void_p),
                ("NormVec", C.c_void_p), ("TangVec1", C.c_void_p), ("TangVec2", C.c_void_p),
                ("SurfElem", C.c_void_p),
                ("nfvse", C.c_int), ("spacePropSweeps", C.c_int), ("alpha_max", C.c_double),
                ("alpha_min", C.c_double), ("spacePropFactor", C.c_double), ("alpha", C.c_void_p),
                ("computeAlpha", C.c_int), ("shockBlendCoef", C.c_double)]


def box_mesh(K, N, x0=(0.0, 0.0, 0.0), length=(2.0, 2.0, 2.0), Ngeo=1, warp=False, amp=0.1):
    """Periodic box mesh: connectivity, node coordinates and FLUXO metrics."""
    L = lib()
    Kx, Ky, Kz = (K, K, K) if np.isscalar(K) else K
    E = Kx * Ky * Kz
    S = 3 * E
    ng = Ngeo + 1
    e2s = np.zeros((E, 6, 2), np.int32)
    s2e = np.zeros((S, 5), np.int32)
    assert L.ora_box_connectivity(Kx, Ky, Kz, e2s, s2e) == E
    X = np.zeros((E, ng, ng, ng, 3))
    L.ora_box_nodecoords(Kx, Ky, Kz, np.array(x0, float), np.array(length, float), Ngeo,
                         int(bool(warp)), amp, X)
    m = mesh_from(N, Ngeo, e2s, s2e, X)
    m["K"] = (Kx, Ky, Kz)
    return m


def mesh_from(N, Ngeo, e2s, s2e, X):
    """FLUXO metrics (the oracle's CalcMetrics) of a conforming mesh given by its connectivity
    and equidistant node coordinates, e.g. a HOPR file's."""
    L = lib()
    E, S, n = e2s.shape[0], s2e.shape[0], N + 1
    e2s, s2e = np.ascontiguousarray(e2s, np.int32), np.ascontiguousarray(s2e, np.int32)
    X = np.ascontiguousarray(X, np.float64)
    m = dict(N=N, Ngeo=Ngeo, nElems=E, nSides=S, ElemToSide=e2s, SideToElem=s2e, NodeCoords=X)
    m["Mf"] = np.zeros((E, n, n, n, 3)); m["Mg"] = np.zeros((E, n, n, n, 3))
    m["Mh"] = np.zeros((E, n, n, n, 3)); m["sJ"] = np.zeros((E, n, n, n))
    for k in ("NormVec", "TangVec1", "TangVec2"):
        m[k] = np.zeros((S, n, n, 3))
    m["SurfElem"] = np.zeros((S, n, n))
    m["Elem_xGP"] = np.zeros((E, n, n, n, 3))
    rc = L.ora_calc_metrics(N, Ngeo, E, S, X, e2s, m["Mf"], m["Mg"], m["Mh"], m["sJ"], m["NormVec"],
                            m["TangVec1"], m["TangVec2"], m["SurfElem"], m["Elem_xGP"])
    if rc != 0:
        raise ValueError("mesh has non-positive or badly scaled Jacobians")
    return m


class Problem:
    """An equation/mesh setup; keeps references to every array the C struct points to."""

    def __init__(self, mesh, volflux=32, riemann=1, parabolic=False, kappa=1.4, mu0=0.0, Pr=0.72,
                 lifting=1, nfvse=False, alpha_max=0.5, alpha_min=0.01, sweeps=1, prop_factor=0.5,
                 fixed_alpha=None, etaBR2=2.0):
        self.mesh = mesh
        self.N = mesh["N"]
        self.nElems = mesh["nElems"]
        self.kappa = kappa
        self.parabolic = bool(parabolic)
        self.mu0 = mu0
        self.Pr = Pr
        self.volflux, self.riemann = volflux, riemann
        p = _Problem()
        p.N, p.nVar, p.nElems, p.nSides = self.N, 5, mesh["nElems"], mesh["nSides"]
        p.volflux, p.riemann, p.parabolic, p.lifting = volflux, riemann, int(bool(parabolic)), lifting
        p.kappa, p.mu0, p.Pr, p.etaBR2 = kappa, mu0, Pr, etaBR2
        for k in ("ElemToSide", "SideToElem", "Mf", "Mg", "Mh", "sJ", "NormVec", "TangVec1",
                  "TangVec2", "SurfElem"):
            setattr(p, k, mesh[k].ctypes.data)
        p.nfvse, p.spacePropSweeps = int(bool(nfvse)), sweeps
        p.alpha_max, p.alpha_min, p.spacePropFactor = alpha_max, alpha_min, prop_factor
        p.computeAlpha = 1 if fixed_alpha is None else 3
        p.shockBlendCoef = 0.0 if fixed_alpha is None else fixed_alpha
        self.alpha = np.zeros(mesh["nElems"])
        p.alpha = self.alpha.ctypes.data
        self._p = p

    @property
    def ptr(self):
        return C.byref(self._p)

    def shape(self, nvar=5):
        n = self.N + 1
        return (self.nElems, n, n, n, nvar)

    def rhs(self, U, t=0.0):
        Ut = np.zeros(self.shape())
        lib().ora_dg_time_derivative(self.ptr, np.ascontiguousarray(U), t, Ut)
        return Ut

    def volint(self, U):
        Ut = np.zeros(self.shape())
        lib().ora_volint_splitform(self.ptr, np.ascontiguousarray(U), Ut)
        return Ut

    def lifting(self, U):
        g = [np.zeros(self.shape()) for _ in range(3)]
        lib().ora_lifting(self.ptr, np.ascontiguousarray(U), *g)
        return g

    def face_shape(self, nvar=5):
        n = self.N + 1
        return (self.mesh["nSides"], n, n, nvar)

    def fillflux(self, U):
        """FillFlux: (Flux_master, Flux_slave), each (nSides, n, n, 5)."""
        fm, fs = np.zeros(self.face_shape()), np.zeros(self.face_shape())
        lib().ora_fill_flux(self.ptr, np.ascontiguousarray(U), fm, fs)
        return fm, fs

    def surfint(self, fm, fs, Ut, nvar=5):
        """SurfInt: Ut + the surface integral of the fluxes (a new array)."""
        out = np.array(Ut, dtype=np.float64, order="C", copy=True)
        lib().ora_surf_int(self.ptr, nvar, np.ascontiguousarray(fm, dtype=np.float64),
                           np.ascontiguousarray(fs, dtype=np.float64), out)
        return out

    def volint_nfvse(self, U):
        """VolInt then VolInt_NFVSE with the blending coefficient of U."""
        U = np.ascontiguousarray(U)
        Ut = self.volint(U)
        lib().ora_volint_nfvse(self.ptr, U, self.blending(U), Ut)
        return Ut

    def lserk_step(self, U, Ut_tmp, t, dt):
        lib().ora_lserk_step(self.ptr, U, Ut_tmp, t, dt)

    def indicator(self, U):
        eta = np.zeros(self.nElems)
        lib().ora_persson_peraire(self.ptr, np.ascontiguousarray(U), eta)
        return eta

    def blending(self, U):
        a = np.zeros(self.nElems)
        lib().ora_blending_coefficient(self.ptr, np.ascontiguousarray(U), a)
        return a

    def calc_dt(self, U, cfl, dfl=0.0):
        err = C.c_int(0)
        dt = lib().ora_calc_timestep(self.ptr, np.ascontiguousarray(U), cfl, dfl, C.byref(err))
        return dt, err.value

    def entropy(self, U, Ut):
        s, d = C.c_double(), C.c_double()
        lib().ora_calc_entropy(self.ptr, np.ascontiguousarray(U), np.ascontiguousarray(Ut),
                               C.byref(s), C.byref(d))
        return s.value, d.value


class AnalyzeCfg(C.Structure):
    _fields_ = [("exactFunc", C.c_int), ("nAnalyze", C.c_int), ("refState", C.c_double * 9),
                ("advVel", C.c_double * 3), ("waveNumber", C.c_double * 3), ("amplitude", C.c_double),
                ("R", C.c_double), ("dissFactor", C.c_double)]


ANALYSIS_SCALARS = ("Ekin", "Ekin_comp", "Emag_comp", "Enstrophy_comp", "Mag_Enstrophy", "Entropy_comp",
                    "DR_u", "DR_S", "DR_Sd", "DR_p", "max_Vorticity", "max_Current", "mean_temperature",
                    "uprime")
ANALYSIS_VECTORS = ("L2", "L2colloc", "Linf", "bulk", "bulk_t", "maxabsUt")


class Analysis(C.Structure):
    _fields_ = ([("vol", C.c_double)] + [(k, C.c_double * 9) for k in ANALYSIS_VECTORS] +
                [(k, C.c_double) for k in ANALYSIS_SCALARS])


def analyze_cfg(exact_func=0, n_analyze=0, ref_state=None, adv_vel=(1.0, 1.0, 1.0), wave_number=(1.0, 1.0, 1.0),
                amplitude=0.1, R=287.058, diss_factor=1.0):
    c = AnalyzeCfg()
    c.exactFunc, c.nAnalyze = exact_func, n_analyze
    if ref_state is not None:
        for i, v in enumerate(ref_state):
            c.refState[i] = v
    for i in range(3):
        c.advVel[i], c.waveNumber[i] = adv_vel[i], wave_number[i]
    c.amplitude, c.R, c.dissFactor = amplitude, R, diss_factor
    return c


def analysis_dict(a, nVar):
    out = {"vol": a.vol}
    for k in ANALYSIS_VECTORS:
        out[k] = np.array(getattr(a, k)[:nVar])
    for k in ANALYSIS_SCALARS:
        out[k] = getattr(a, k)
    return out


def analyze(mesh, U, Ut, grads=None, kappa=1.4, mu_0=1.0, t=0.0, **cfg):
    """The reference's analysis of state U (with its RHS Ut and lifted primitive gradients
    grads = (gx, gy, gz) or None): error norms against cfg's exact function, bulk integrals,
    the Taylor-Green vortex quantities.  nVar from U's last axis (5 or 9)."""
    U, Ut = np.ascontiguousarray(U), np.ascontiguousarray(Ut)
    nVar = U.shape[-1]
    keep = [np.ascontiguousarray(g) for g in grads] if grads is not None else None
    gp = [g.ctypes.data for g in keep] if keep else [None, None, None]
    c, a = analyze_cfg(**cfg), Analysis()
    lib().ora_analyze(mesh["N"], nVar, mesh["nElems"], np.ascontiguousarray(mesh["Elem_xGP"]),
                      np.ascontiguousarray(mesh["sJ"]), U, Ut, gp[0], gp[1], gp[2], kappa, mu_0, t,
                      C.byref(c), C.byref(a))
    return analysis_dict(a, nVar)


def exactfunc(which, x, kappa=1.4, advvel=(1.0, 1.0, 1.0), amp=0.1, t=0.0):
    x = np.ascontiguousarray(x, dtype=np.float64)
    npts = x.size // 3
    U = np.zeros(x.shape[:-1] + (5,))
    lib().ora_exactfunc(which, kappa, np.array(advvel, float), amp, t, npts, x, U)
    return U


def ln_mean(a, b):
    return lib().ora_ln_mean(a, b)


def volflux(which, UL, UR, mL, mR, kappa=1.4):
    UL, UR = np.asarray(UL, float), np.asarray(UR, float)
    aL, aR = np.zeros(6), np.zeros(6)
    lib().ora_eval_uaux.argtypes = [C.c_double, _dp, _dp]
    lib().ora_eval_uaux(kappa, UL, aL)
    lib().ora_eval_uaux(kappa, UR, aR)
    F = np.zeros(5)
    lib().ora_volflux(which, kappa, UL, UR, aL, aR, np.asarray(mL, float), np.asarray(mR, float), F)
    return F


def riemann(which, UL, UR, nv, t1, t2, kappa=1.4):
    F = np.zeros(5)
    lib().ora_riemann(which, kappa, *(np.asarray(a, float) for a in (UL, UR, nv, t1, t2)), F)
    return F


def cons_to_entropy(U, kappa=1.4):
    w = np.zeros(5)
    lib().ora_cons_to_entropy(kappa, np.asarray(U, float), w)
    return w


def rk_coefs():
    A, b, c = np.zeros(5), np.zeros(5), np.zeros(5)
    lib().ora_rk_coefs(A, b, c)
    return A, b, c


def dg_operators(N):
    n = N + 1
    x, w = np.zeros(n), np.zeros(n)
    D, Dh, Ds = np.zeros(n * n), np.zeros(n * n), np.zeros(n * n)
    Lm, Lp = np.zeros(n), np.zeros(n)
    lib().ora_dg_operators(N, x, w, D, Dh, Ds, Lm, Lp)
    f = lambda a: a.reshape(n, n).T.copy()  # column-major -> A[r, c]
    return dict(x=x, w=w, D=f(D), D_hat=f(Dh), Dvolsurf=f(Ds), L_hatminus=Lm, L_hatplus=Lp)


def subcell_metrics(mesh, e):
    """Sub-cell face metrics of element e: (3, N, n, n, 10) = [d][f][b][a][nv t1 t2 norm]."""
    N = mesh["N"]
    n = N + 1
    out = np.zeros((3, N, n, n, 10))
    lib().ora_subcell_metrics(N, np.ascontiguousarray(mesh["Mf"][e]), np.ascontiguousarray(mesh["Mg"][e]),
                              np.ascontiguousarray(mesh["Mh"][e]), out)
    return out


def nfvse_threshold(N):
    return lib().ora_nfvse_threshold(N)


class _Mhd(C.Structure):
    _fields_ = [("P", C.c_void_p), ("kappa", C.c_double), ("mu_0", C.c_double), ("glm_scale", C.c_double),
                ("parabolic", C.c_int), ("riemann", C.c_int), ("mu", C.c_double), ("eta", C.c_double),
                ("Pr", C.c_double), ("lifting", C.c_int), ("reserved2", C.c_int), ("etaBR2", C.c_double)]


class MhdProblem:
    """GLM-MHD setup (9 variables) on a mesh: Derigs volume flux, surface flux `riemann`
    (0 central, 1 Rusanov, 10 Derigs + entropy-stable LLF, 11 Derigs), Powell
    non-conservative terms; resistive (BR1 lifting of the entropy variables, viscosity
    mu, resistivity eta, heat conduction mu kappa/Pr; lifting 2 = BR2 with etaBR2) when
    `resistive` is given as
    dict(mu=, eta=, Pr=).  The cleaning speed GLM_ch is explicit state: 0.5 before the first
    time step (the reference's initial value), then glm_scale * dtch1 / dt from every
    CALCTIMESTEP."""

    def __init__(self, mesh, kappa=5.0 / 3.0, mu_0=1.0, glm_scale=0.5, resistive=None, riemann=1, lifting=1,
                 etaBR2=2.0):
        self.base = Problem(mesh, volflux=10, riemann=1, kappa=kappa)
        self.mesh, self.N, self.nElems = mesh, mesh["N"], mesh["nElems"]
        m = _Mhd()
        m.P = C.cast(C.byref(self.base._p), C.c_void_p).value
        m.kappa, m.mu_0, m.glm_scale = kappa, mu_0, glm_scale
        m.riemann = riemann
        m.lifting, m.etaBR2 = lifting, etaBR2
        if resistive is not None:
            m.parabolic = 1
            m.mu, m.eta, m.Pr = float(resistive["mu"]), float(resistive["eta"]), float(resistive.get("Pr", 0.72))
        self.parabolic = resistive is not None
        self._m = m
        self.glm_ch = 0.5
        self.glm_dtch1 = None

    @property
    def ptr(self):
        return C.byref(self._m)

    def shape(self):
        n = self.N + 1
        return (self.nElems, n, n, n, 9)

    def exact_orszag_tang(self):
        x = np.ascontiguousarray(self.mesh["Elem_xGP"])
        U = np.zeros(self.shape())
        lib().ora_mhd_orszag_tang(self.ptr, x.size // 3, x, U)
        return U

    def rhs(self, U, glm_ch=None):
        Ut = np.zeros(self.shape())
        lib().ora_mhd_rhs(self.ptr, np.ascontiguousarray(U), self.glm_ch if glm_ch is None else glm_ch, Ut)
        return Ut

    def fillflux(self, U, glm_ch=None):
        """FillFlux: (Flux_master, Flux_slave), each (nSides, n, n, 9)."""
        n = self.N + 1
        shp = (self.mesh["nSides"], n, n, 9)
        fm, fs = np.zeros(shp), np.zeros(shp)
        lib().ora_mhd_fill_flux(self.ptr, np.ascontiguousarray(U), self.glm_ch if glm_ch is None else glm_ch,
                                fm, fs)
        return fm, fs

    def surfint(self, fm, fs, Ut):
        return self.base.surfint(fm, fs, Ut, nvar=9)

    def lifting(self, U):
        """Primitive gradients (rho, v, p, B, psi) x, y, z of the resistive case."""
        g = [np.zeros(self.shape()) for _ in range(3)]
        lib().ora_mhd_lifting(self.ptr, np.ascontiguousarray(U), *g)
        return g

    def calc_dt(self, U, cfl, dfl=0.9):
        """CALCTIMESTEP; updates glm_ch (and dtch1 on the first call) like the reference."""
        if self.glm_dtch1 is None:
            self.glm_dtch1 = lib().ora_mhd_glm_dtch1(self.ptr, cfl)
        ch, err = C.c_double(0.0), C.c_int(0)
        dt = lib().ora_mhd_calc_timestep(self.ptr, np.ascontiguousarray(U), cfl, dfl, self.glm_dtch1,
                                         C.byref(ch), C.byref(err))
        self.glm_ch = ch.value
        return dt, err.value

    def lserk_step(self, U, Ut_tmp, dt):
        lib().ora_mhd_lserk_step(self.ptr, U, Ut_tmp, self.glm_ch, dt)
