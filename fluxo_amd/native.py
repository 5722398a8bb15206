"""ctypes binding of the C ABI in include/fluxo_gpu.h (libfluxo_gpu.so).

This is the Python-side host of the engine: it mirrors the reference's
DGTimeDerivative / TimeStepByLSERKW2 / CALCTIMESTEP / CalcEntropy entry points with the
same argument meaning.  Arrays use numpy C order with the Fortran index order reversed,
so their memory equals the reference's column-major arrays:
  U(nVar,0:N,0:N,0:N,nElems)  <->  numpy (nElems, N+1, N+1, N+1, nVar)

There is no CPU fallback: every call runs the HIP kernels, and loading fails loudly if
the native library is missing.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfluxo_gpu.so")
_LIB = None

FX_PREC_FP64 = 0
FX_PREC_FP32_SPLIT = 1


class FxError(RuntimeError):
    pass


class fx_config(C.Structure):
    _fields_ = [("N", C.c_int), ("nVar", C.c_int), ("nElems", C.c_int), ("nSides", C.c_int),
                ("volumeFlux", C.c_int), ("riemann", C.c_int), ("parabolic", C.c_int),
                ("lifting", C.c_int), ("precision", C.c_int), ("whichMortar", C.c_int),
                ("kappa", C.c_double), ("mu0", C.c_double), ("Pr", C.c_double),
                ("etaBR2", C.c_double),
                ("nfvse", C.c_int), ("computeAlpha", C.c_int), ("spacePropSweeps", C.c_int),
                ("reserved3", C.c_int), ("alpha_max", C.c_double), ("alpha_min", C.c_double),
                ("spacePropFactor", C.c_double), ("shockBlendCoef", C.c_double),
                ("mu_0", C.c_double), ("GLM_scale", C.c_double), ("eta", C.c_double), ("reserved4", C.c_double)]


class fx_mesh(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in ("ElemToSide", "SideToElem", "Metrics_fTilde",
                                          "Metrics_gTilde", "Metrics_hTilde", "sJ", "NormVec",
                                          "TangVec1", "TangVec2", "SurfElem", "MortarType", "MortarInfo")] + [
        ("nMortarSides", C.c_int), ("reserved_m", C.c_int)]


class fx_server_config(C.Structure):
    _fields_ = [("volumeFlux", C.c_int), ("riemann", C.c_int), ("precision", C.c_int),
                ("lifting", C.c_int), ("kappa", C.c_double), ("mu0", C.c_double),
                ("Pr", C.c_double), ("CFLscale", C.c_double), ("DFLscale", C.c_double),
                ("hbmBudget", C.c_longlong), ("maxRequests", C.c_int), ("admitKNodes", C.c_int)]


class fx_request(C.Structure):
    _fields_ = [("N", C.c_int), ("parabolic", C.c_int), ("nElems", C.c_int), ("nSides", C.c_int),
                ("nSteps", C.c_int), ("reserved", C.c_int), ("mesh", fx_mesh),
                ("U0", C.c_void_p), ("U_out", C.c_void_p)]


class fx_request_status(C.Structure):
    _fields_ = [("state", C.c_int), ("errType", C.c_int), ("steps", C.c_int),
                ("reserved", C.c_int), ("t_end", C.c_double), ("dt_last", C.c_double),
                ("t_submit", C.c_double), ("t_admit", C.c_double), ("t_done", C.c_double)]


class fx_server_stats(C.Structure):
    _fields_ = [("steps_launched", C.c_longlong), ("graph_captures", C.c_longlong),
                ("bytes_in_use", C.c_longlong), ("bytes_peak", C.c_longlong),
                ("bytes_allocated", C.c_longlong), ("active", C.c_int), ("queued", C.c_int),
                ("done", C.c_int), ("failed", C.c_int), ("graph_updates", C.c_longlong)]


class fx_analyze_config(C.Structure):
    _fields_ = [("exactFunc", C.c_int), ("nAnalyze", C.c_int), ("refState", C.c_double * 9),
                ("advVel", C.c_double * 3), ("waveNumber", C.c_double * 3), ("amplitude", C.c_double),
                ("R", C.c_double), ("dissFactor", C.c_double)]


ANALYSIS_VECTORS = ("L2", "L2colloc", "Linf", "bulk", "bulk_t", "maxabs_Ut")
ANALYSIS_SCALARS = ("Ekin", "Ekin_comp", "Emag_comp", "Enstrophy_comp", "Mag_Enstrophy", "Entropy_comp",
                    "DR_u", "DR_S", "DR_Sd", "DR_p", "max_Vorticity", "max_Current", "mean_temperature",
                    "uprime")


class fx_analysis(C.Structure):
    _fields_ = ([("vol", C.c_double)] + [(k, C.c_double * 9) for k in ANALYSIS_VECTORS] +
                [(k, C.c_double) for k in ANALYSIS_SCALARS])


FX_REQ_QUEUED, FX_REQ_RUNNING, FX_REQ_DONE, FX_REQ_FAILED = range(4)

EXPORTS = ["fx_ctx_create", "fx_ctx_destroy", "fx_last_error", "fx_mesh_upload",
           "fx_state_upload", "fx_state_download", "fx_ut_download", "fx_rhs", "fx_volint",
           "fx_lifting", "fx_calc_dt", "fx_rk_step", "fx_run", "fx_entropy", "fx_synchronize",
           "fx_ctx_bytes", "fx_profile_kernel", "fx_server_create", "fx_server_submit",
           "fx_server_status", "fx_server_wait", "fx_server_get_stats", "fx_server_destroy",
           "fx_request_bytes", "fx_server_last_error", "fx_tune_volume", "fx_volume_tile",
           "fx_use_volume_tile", "fx_blending", "fx_glm", "fx_elem_xgp_upload", "fx_analyze",
           "fx_rhs_noalloc", "fx_rhs_noalloc_bytes", "fx_fillflux", "fx_surfint", "fx_volint_nfvse",
           "fx_run_to"]

(FX_KERNEL_VOLUME_RK, FX_KERNEL_SURFACE, FX_KERNEL_LIFTING, FX_KERNEL_DT, FX_KERNEL_VOLUME_RHS,
 FX_KERNEL_NFVSE_ALPHA, FX_KERNEL_NFVSE_FV) = range(7)


def lib():
    """Load libfluxo_gpu.so (raises if it was not built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise FxError(f"native engine not built: {LIB_PATH} missing (run __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)
    L.fx_ctx_create.argtypes = [C.POINTER(fx_config), C.POINTER(vp)]
    L.fx_ctx_destroy.argtypes = [vp]
    L.fx_ctx_destroy.restype = None
    L.fx_last_error.restype = C.c_char_p
    L.fx_mesh_upload.argtypes = [vp, C.POINTER(fx_mesh)]
    for f in ("fx_state_upload", "fx_state_download", "fx_ut_download", "fx_volint"):
        getattr(L, f).argtypes = [vp, vp]
    L.fx_rhs.argtypes = [vp, C.c_double, vp]
    L.fx_lifting.argtypes = [vp, vp, vp, vp]
    L.fx_blending.argtypes = [vp, vp]
    L.fx_fillflux.argtypes = [vp, vp, vp]
    L.fx_surfint.argtypes = [vp, vp, vp, vp]
    L.fx_volint_nfvse.argtypes = [vp, vp]
    L.fx_run_to.argtypes = [vp, C.c_int, C.c_double, C.c_double, C.c_double, dp, dp, ip, ip]
    L.fx_glm.argtypes = [vp, dp, dp]
    L.fx_calc_dt.argtypes = [vp, C.c_double, C.c_double, dp, ip]
    L.fx_rk_step.argtypes = [vp, C.c_double, C.c_double]
    L.fx_run.argtypes = [vp, C.c_int, C.c_double, C.c_double, dp, dp, ip]
    L.fx_entropy.argtypes = [vp, dp, dp]
    L.fx_synchronize.argtypes = [vp]
    L.fx_ctx_bytes.argtypes = [C.POINTER(fx_config)]
    L.fx_ctx_bytes.restype = C.c_longlong
    L.fx_profile_kernel.argtypes = [vp, C.c_int, C.c_int, dp]
    L.fx_server_create.argtypes = [C.POINTER(fx_server_config), C.POINTER(vp)]
    L.fx_server_submit.argtypes = [vp, C.POINTER(fx_request), ip]
    L.fx_server_status.argtypes = [vp, C.c_int, C.POINTER(fx_request_status)]
    L.fx_server_wait.argtypes = [vp, C.c_int, C.c_double]
    L.fx_server_get_stats.argtypes = [vp, C.POINTER(fx_server_stats)]
    L.fx_server_destroy.argtypes = [vp]
    L.fx_server_destroy.restype = None
    L.fx_request_bytes.argtypes = [C.POINTER(fx_server_config), C.c_int, C.c_int, C.c_int, C.c_int]
    L.fx_request_bytes.restype = C.c_longlong
    L.fx_tune_volume.argtypes = [vp, C.c_int, ip, ip, dp, ip]
    L.fx_volume_tile.argtypes = [C.c_int, C.c_int, C.c_int, ip]
    L.fx_use_volume_tile.argtypes = [vp, C.c_int]
    L.fx_elem_xgp_upload.argtypes = [vp, vp]
    L.fx_analyze.argtypes = [vp, C.c_double, C.POINTER(fx_analyze_config), C.POINTER(fx_analysis)]
    L.fx_server_last_error.argtypes = [vp]
    L.fx_server_last_error.restype = C.c_char_p
    L.fx_rhs_noalloc.argtypes = [C.POINTER(fx_config), C.POINTER(fx_mesh), vp, vp, vp]
    L.fx_rhs_noalloc_bytes.argtypes = [C.POINTER(fx_config)]
    L.fx_rhs_noalloc_bytes.restype = C.c_longlong
    _LIB = L
    return L


def _chk(rc):
    if rc != 0:
        raise FxError(f"fluxo_gpu error {rc}: {lib().fx_last_error().decode()}")


def _ptr(a):
    return a.ctypes.data


def make_config(N, nElems, nSides, volume_flux=32, riemann=1, parabolic=False, kappa=1.4,
                mu0=0.0, Pr=0.72, precision=FX_PREC_FP64, nfvse=False, alpha_max=0.5, alpha_min=0.01,
                sweeps=1, prop_factor=0.5, fixed_alpha=None, lifting=1, etaBR2=2.0, which_mortar=0):
    """Engine options; the NFVSE ones follow the reference's parameter names (alpha_max,
    alpha_min, SpacePropSweeps, SpacePropFactor; fixed_alpha = ComputeAlpha 3 with that
    ShockBlendCoef, None = Persson-Peraire indicator); lifting 1 BR1, 2 BR2 (penalty
    etaBR2); which_mortar = whichMortar on meshes with mortars (0 projection, 1 collocation)."""
    c = fx_config()
    c.mu_0, c.GLM_scale = 1.0, 0.5
    c.nfvse, c.spacePropSweeps = int(bool(nfvse)), sweeps
    c.computeAlpha = 1 if fixed_alpha is None else 3
    c.alpha_max, c.alpha_min, c.spacePropFactor = alpha_max, alpha_min, prop_factor
    c.shockBlendCoef = 0.0 if fixed_alpha is None else fixed_alpha
    c.N, c.nVar, c.nElems, c.nSides = N, 5, nElems, nSides
    c.volumeFlux, c.riemann, c.parabolic, c.lifting = volume_flux, riemann, int(bool(parabolic)), lifting
    c.etaBR2 = etaBR2
    c.precision = precision
    c.whichMortar = which_mortar
    c.kappa, c.mu0, c.Pr = kappa, mu0, Pr
    return c


def ctx_bytes(cfg):
    return lib().fx_ctx_bytes(C.byref(cfg))


_MESH_KEYS = (("ElemToSide", "ElemToSide"), ("SideToElem", "SideToElem"),
              ("Metrics_fTilde", "Mf"), ("Metrics_gTilde", "Mg"), ("Metrics_hTilde", "Mh"),
              ("sJ", "sJ"), ("NormVec", "NormVec"), ("TangVec1", "TangVec1"),
              ("TangVec2", "TangVec2"), ("SurfElem", "SurfElem"))


def _mesh_struct(mesh):
    """fx_mesh pointing at contiguous copies of the mesh arrays (returned to keep alive), with
    the mortar arrays of a mesh that has them (nMortarSides, MortarType, MortarInfo)."""
    m = fx_mesh()
    keep = {}
    for field, key in _MESH_KEYS:
        dt = np.int32 if key in ("ElemToSide", "SideToElem") else np.float64
        a = np.ascontiguousarray(mesh[key], dtype=dt)
        keep[key] = a
        setattr(m, field, _ptr(a))
    nm = int(mesh.get("nMortarSides", 0))
    if nm:
        for key in ("MortarType", "MortarInfo"):
            keep[key] = np.ascontiguousarray(mesh[key], dtype=np.int32)
            setattr(m, key, _ptr(keep[key]))
        m.nMortarSides = nm
    return m, keep


class Engine:
    """One device context: mesh, state and the RHS/RK entry points.  `options` are the
    further make_config arguments (NFVSE shock capturing, lifting/etaBR2)."""

    def __init__(self, mesh, volume_flux=32, riemann=1, parabolic=False, kappa=1.4, mu0=0.0,
                 Pr=0.72, precision=FX_PREC_FP64, **options):
        L = lib()
        self.N = int(mesh["N"])
        self.nElems = int(mesh["nElems"])
        self.nSides = int(mesh["nSides"])
        self.cfg = make_config(self.N, self.nElems, self.nSides, volume_flux, riemann, parabolic,
                               kappa, mu0, Pr, precision, **options)
        self._ctx = C.c_void_p()
        _chk(L.fx_ctx_create(C.byref(self.cfg), C.byref(self._ctx)))
        m, _keep = _mesh_struct(mesh)
        _chk(L.fx_mesh_upload(self._ctx, C.byref(m)))

    def close(self):
        if self._ctx:
            lib().fx_ctx_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def shape(self):
        n = self.N + 1
        return (self.nElems, n, n, n, int(self.cfg.nVar))

    def upload(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64)
        assert U.shape == self.shape()
        _chk(lib().fx_state_upload(self._ctx, _ptr(U)))

    def download(self):
        U = np.empty(self.shape())
        _chk(lib().fx_state_download(self._ctx, _ptr(U)))
        return U

    def rhs(self, t=0.0, copy=True):
        Ut = np.empty(self.shape()) if copy else None
        _chk(lib().fx_rhs(self._ctx, t, _ptr(Ut) if copy else None))
        return Ut

    def volint(self):
        Ut = np.empty(self.shape())
        _chk(lib().fx_volint(self._ctx, _ptr(Ut)))
        return Ut

    def lifting(self):
        g = [np.empty(self.shape()) for _ in range(3)]
        _chk(lib().fx_lifting(self._ctx, *(_ptr(a) for a in g)))
        return g

    def glm(self):
        """(GLM_ch, GLM_dtch1) of a GLM-MHD context."""
        ch, d1 = C.c_double(0.0), C.c_double(0.0)
        _chk(lib().fx_glm(self._ctx, C.byref(ch), C.byref(d1)))
        return ch.value, d1.value

    def blending(self):
        """NFVSE blending coefficient of the current state (CalcBlendingCoefficient)."""
        a = np.empty(self.nElems)
        _chk(lib().fx_blending(self._ctx, _ptr(a)))
        return a

    def fillflux(self):
        """FillFlux on the device state: (Flux_master, Flux_slave), each (nSides, n, n, nVar)."""
        n = self.N + 1
        fm = np.empty((self.nSides, n, n, int(self.cfg.nVar)))
        fs = np.empty_like(fm)
        _chk(lib().fx_fillflux(self._ctx, _ptr(fm), _ptr(fs)))
        return fm, fs

    def surfint(self, flux_master, flux_slave, Ut):
        """SurfInt: Ut + the surface integral of the given fluxes (a new array)."""
        fm = np.ascontiguousarray(flux_master, dtype=np.float64)
        fs = np.ascontiguousarray(flux_slave, dtype=np.float64)
        out = np.array(Ut, dtype=np.float64, order="C", copy=True)
        assert fm.shape == fs.shape == (self.nSides, self.N + 1, self.N + 1, int(self.cfg.nVar))
        assert out.shape == self.shape()
        _chk(lib().fx_surfint(self._ctx, _ptr(fm), _ptr(fs), _ptr(out)))
        return out

    def volint_nfvse(self):
        """VolInt followed by VolInt_NFVSE: the blended advective volume integral (before -sJ)."""
        Ut = np.empty(self.shape())
        _chk(lib().fx_volint_nfvse(self._ctx, _ptr(Ut)))
        return Ut

    def run_to(self, t_end, cfl, dfl=0.0, t0=0.0, max_steps=-1):
        """TimeDisc up to t_end (the last step shortened to end there); returns
        (t, dt_last, steps, errType)."""
        t, dt, st, err = C.c_double(t0), C.c_double(), C.c_int(), C.c_int()
        _chk(lib().fx_run_to(self._ctx, max_steps, cfl, dfl, t_end, C.byref(t), C.byref(dt), C.byref(st),
                             C.byref(err)))
        return t.value, dt.value, st.value, err.value

    def calc_dt(self, cfl, dfl=0.0):
        dt, err = C.c_double(), C.c_int()
        _chk(lib().fx_calc_dt(self._ctx, cfl, dfl, C.byref(dt), C.byref(err)))
        return dt.value, err.value

    def rk_step(self, t, dt):
        _chk(lib().fx_rk_step(self._ctx, t, dt))

    def run(self, nsteps, cfl, dfl=0.0, t0=0.0):
        t, dt, err = C.c_double(t0), C.c_double(), C.c_int()
        _chk(lib().fx_run(self._ctx, nsteps, cfl, dfl, C.byref(t), C.byref(dt), C.byref(err)))
        return t.value, dt.value, err.value

    def entropy(self):
        s, d = C.c_double(), C.c_double()
        _chk(lib().fx_entropy(self._ctx, C.byref(s), C.byref(d)))
        return s.value, d.value

    def upload_coords(self, Elem_xGP):
        """Elem_xGP (nElems, n, n, n, 3), for the exact functions of analyze()."""
        x = np.ascontiguousarray(Elem_xGP, dtype=np.float64)
        _chk(lib().fx_elem_xgp_upload(self._ctx, _ptr(x)))

    def analyze(self, t=0.0, exact_func=0, n_analyze=0, ref_state=None, adv_vel=(1.0, 1.0, 1.0),
                wave_number=(1.0, 1.0, 1.0), amplitude=0.1, R=287.058, diss_factor=1.0):
        """The reference's Analyze of the device state at time t (fx_analyze): error norms
        against exact_func (needs upload_coords), volume, bulk integrals of U and of its RHS,
        Taylor-Green vortex quantities.  Returns a dict of the fx_analysis fields."""
        c = fx_analyze_config()
        c.exactFunc, c.nAnalyze = exact_func, n_analyze
        for i, v in enumerate(ref_state or ()):
            c.refState[i] = v
        for i in range(3):
            c.advVel[i], c.waveNumber[i] = adv_vel[i], wave_number[i]
        c.amplitude, c.R, c.dissFactor = amplitude, R, diss_factor
        a = fx_analysis()
        _chk(lib().fx_analyze(self._ctx, t, C.byref(c), C.byref(a)))
        nv = self.shape()[-1]
        out = {"vol": a.vol}
        for k in ANALYSIS_VECTORS:
            out[k] = np.array(getattr(a, k)[:nv])
        for k in ANALYSIS_SCALARS:
            out[k] = getattr(a, k)
        return out

    def profile_kernel(self, which, reps=20):
        """Average ms per launch of one engine kernel (HIP events on the engine stream)."""
        ms = C.c_double()
        _chk(lib().fx_profile_kernel(self._ctx, which, reps, C.byref(ms)))
        return ms.value

    def tune_volume(self, reps=20):
        """Time every tile of the LSERK volume kernel on this mesh/state (restored), keep the
        fastest; returns (best elements per workgroup, {tile: ms})."""
        best, n = C.c_int(0), C.c_int(0)
        tiles, ms = (C.c_int * 4)(), (C.c_double * 4)()
        _chk(lib().fx_tune_volume(self._ctx, reps, C.byref(best), tiles, ms, C.byref(n)))
        return best.value, {tiles[i]: ms[i] for i in range(n.value)}

    def use_volume_tile(self, epb):
        _chk(lib().fx_use_volume_tile(self._ctx, epb))

    def sync(self):
        _chk(lib().fx_synchronize(self._ctx))


def server_config(volume_flux=32, riemann=1, precision=FX_PREC_FP64, kappa=1.4, mu0=0.0, Pr=0.72,
                  cfl=0.9, dfl=0.9, hbm_budget=64 << 30, max_requests=4096, admit_knodes=0):
    """admit_knodes > 0 caps the admitted work at that many thousand volume nodes, with strict
    FIFO order at the cap (fx_server_config.admitKNodes)."""
    c = fx_server_config()
    c.volumeFlux, c.riemann, c.precision, c.lifting = volume_flux, riemann, precision, 1
    c.kappa, c.mu0, c.Pr, c.CFLscale, c.DFLscale = kappa, mu0, Pr, cfl, dfl
    c.hbmBudget, c.maxRequests, c.admitKNodes = hbm_budget, max_requests, admit_knodes
    return c


def volume_tile(N, parabolic=False, precision=FX_PREC_FP64):
    """Elements per workgroup of the LSERK volume kernel (process-wide choice)."""
    e = C.c_int(0)
    _chk(lib().fx_volume_tile(N, int(bool(parabolic)), precision, C.byref(e)))
    return e.value


def request_bytes(cfg, N, parabolic, nElems, nSides):
    return lib().fx_request_bytes(C.byref(cfg), N, int(bool(parabolic)), nElems, nSides)


class Server:
    """Batched server: submit() queues a whole run (nSteps LSERK steps with CALCTIMESTEP) on
    one mesh; a worker thread in the library batches the running requests per
    (N, parabolic) pool and admits new ones between steps.  Arrays handed to submit() are
    kept referenced here until the request completes."""

    def __init__(self, **kw):
        L = lib()
        self.cfg = server_config(**kw)
        self._srv = C.c_void_p()
        _chk(L.fx_server_create(C.byref(self.cfg), C.byref(self._srv)))
        self._keep = {}

    def submit(self, mesh, U0, nsteps, parabolic=False):
        """Queue one request; returns (id, U_out) where U_out receives the final state."""
        m, keep = _mesh_struct(mesh)
        U0 = np.ascontiguousarray(U0, dtype=np.float64)
        out = np.empty_like(U0)
        r = fx_request()
        r.N, r.parabolic = int(mesh["N"]), int(bool(parabolic))
        r.nElems, r.nSides, r.nSteps = int(mesh["nElems"]), int(mesh["nSides"]), int(nsteps)
        r.mesh = m
        r.U0, r.U_out = _ptr(U0), _ptr(out)
        i = C.c_int(-1)
        _chk(lib().fx_server_submit(self._srv, C.byref(r), C.byref(i)))
        self._keep[i.value] = (keep, U0, out)
        return i.value, out

    def status(self, rid):
        st = fx_request_status()
        _chk(lib().fx_server_status(self._srv, rid, C.byref(st)))
        return {k: getattr(st, k) for k, _ in fx_request_status._fields_ if k != "reserved"}

    def wait(self, rid=-1, timeout=600.0):
        _chk(lib().fx_server_wait(self._srv, rid, timeout))

    def last_error(self):
        return lib().fx_server_last_error(self._srv).decode()

    def stats(self):
        st = fx_server_stats()
        _chk(lib().fx_server_get_stats(self._srv, C.byref(st)))
        return {k: getattr(st, k) for k, _ in fx_server_stats._fields_}

    def close(self):
        if self._srv:
            lib().fx_server_destroy(self._srv)
            self._srv = C.c_void_p()
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MhdEngine(Engine):
    """A GLM-MHD context (9 variables; Derigs volume flux, Powell terms; surface flux
    `riemann` 0 central, 1 Rusanov, 10 Derigs + entropy-stable LLF, 11 Derigs): U is
    (nElems, n, n, n, 9).  resistive=dict(mu=, eta=, Pr=) selects the resistive equations
    (lifting of the entropy variables: 1 BR1, 2 BR2 with etaBR2)."""

    def __init__(self, mesh, kappa=5.0 / 3.0, mu_0=1.0, glm_scale=0.5, precision=FX_PREC_FP64, resistive=None,
                 riemann=1, lifting=1, etaBR2=2.0, which_mortar=0):
        L = lib()
        self.N = int(mesh["N"])
        self.nElems = int(mesh["nElems"])
        self.nSides = int(mesh["nSides"])
        res = resistive or {}
        self.cfg = make_config(self.N, self.nElems, self.nSides, 10, riemann, resistive is not None, kappa,
                               float(res.get("mu", 0.0)), float(res.get("Pr", 0.72)), precision)
        self.cfg.nVar, self.cfg.mu_0, self.cfg.GLM_scale = 9, mu_0, glm_scale
        self.cfg.eta = float(res.get("eta", 0.0))
        self.cfg.lifting, self.cfg.etaBR2 = lifting, etaBR2
        self.cfg.whichMortar = which_mortar
        self._ctx = C.c_void_p()
        _chk(L.fx_ctx_create(C.byref(self.cfg), C.byref(self._ctx)))
        m, _keep = _mesh_struct(mesh)
        _chk(L.fx_mesh_upload(self._ctx, C.byref(m)))
