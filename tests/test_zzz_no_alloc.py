"""fx_rhs_noalloc (no device allocation: the library's static scratch array) against fx_rhs on
a context, through the C ABI.  Both run the same kernels on the same data, so the right-hand
sides must agree bitwise.  Two calls in a row, the second on a smaller problem, so that the
second call reuses scratch the first one filled."""
import ctypes as C

import numpy as np
import pytest

from fluxo_amd import native as NT

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    """The HIP runtime the engine library itself loaded (device pointers must come from it)."""
    NT.lib()
    path = None
    with open("/proc/self/maps") as f:
        for line in f:
            if "libamdhip64.so" in line and "/torch/" not in line:
                path = line.split()[-1]
                break
    h = C.CDLL(path or "libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipFree.argtypes = [C.c_void_p]
    return h


class DeviceProblem:
    """Device copies of a mesh and state, allocated by the caller (this test)."""

    def __init__(self, hip, mesh, U):
        self.hip, self.ptrs = hip, []
        self.mesh = NT.fx_mesh()
        for field, key in NT._MESH_KEYS:
            dt = np.int32 if key in ("ElemToSide", "SideToElem") else np.float64
            setattr(self.mesh, field, self.put(np.ascontiguousarray(mesh[key], dtype=dt)))
        self.U = self.put(np.ascontiguousarray(U, dtype=np.float64))
        self.shape = U.shape
        self.Ut = self.alloc(U.nbytes)

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), nbytes) == 0
        self.ptrs.append(p)
        return p.value

    def put(self, a):
        p = self.alloc(a.nbytes)
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        return p

    def ut(self):
        out = np.empty(self.shape)
        assert self.hip.hipMemcpy(out.ctypes.data, self.Ut, out.nbytes, D2H) == 0
        return out

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def _case(kind):
    import math

    import pyoracle as O
    from fluxo_amd.mesh import box_mesh
    from fluxo_amd.workloads import sedov_blast
    if kind == "ns_tgv":
        m = box_mesh(4, 3, x0=(-math.pi,) * 3, length=(2 * math.pi,) * 3, Ngeo=3, warp=True)
        U = O.exactfunc(7, m["Elem_xGP"])
        opts = dict(volume_flux=32, riemann=1, parabolic=True, mu0=1.0 / 1600, Pr=0.72)
    else:  # Euler with NFVSE shock capturing, smaller and of another degree
        m = box_mesh(2, 2, x0=(-1.5,) * 3, length=(3.0,) * 3, Ngeo=3, warp=True)
        U = sedov_blast(m["Elem_xGP"], sigma=0.2, p0=1e-2)
        opts = dict(volume_flux=32, riemann=1, nfvse=True, alpha_max=0.5, alpha_min=0.01, sweeps=1,
                    prop_factor=0.5)
    return m, U, opts


def test_noalloc_rhs_equals_context_rhs_twice_second_smaller():
    hip = _hip()
    L = NT.lib()
    for kind in ("ns_tgv", "euler_nfvse"):
        m, U, opts = _case(kind)
        e = NT.Engine(m, **opts)
        e.upload(U)
        ref = e.rhs()
        e.close()
        cfg = NT.make_config(int(m["N"]), int(m["nElems"]), int(m["nSides"]), **opts)
        need = L.fx_rhs_noalloc_bytes(C.byref(cfg))
        print(f"fx_rhs_noalloc {kind}: N={m['N']} nElems={m['nElems']} scratch {need} bytes", flush=True)
        dp = DeviceProblem(hip, m, U)
        try:
            rc = L.fx_rhs_noalloc(C.byref(cfg), C.byref(dp.mesh), dp.U, dp.Ut, None)
            assert rc == 0, L.fx_last_error().decode()
            assert hip.hipDeviceSynchronize() == 0
            got = dp.ut()
        finally:
            dp.free()
        assert np.array_equal(got, ref), float(np.max(np.abs(got - ref)))


def test_noalloc_rejects_invalid_connectivity_before_any_compute():
    """A SideToElem entry out of range is caught by the device-side check (which reads only
    in-bounds positions) and reported as FX_EINVAL; no compute kernel runs."""
    hip = _hip()
    L = NT.lib()
    m, U, opts = _case("euler_nfvse")
    bad = dict(m)
    s2e = np.array(m["SideToElem"], dtype=np.int32, copy=True)
    s2e.reshape(-1)[0] = int(m["nElems"]) + 5
    bad["SideToElem"] = s2e
    cfg = NT.make_config(int(m["N"]), int(m["nElems"]), int(m["nSides"]), **opts)
    dp = DeviceProblem(hip, bad, U)
    try:
        rc = L.fx_rhs_noalloc(C.byref(cfg), C.byref(dp.mesh), dp.U, dp.Ut, None)
    finally:
        dp.free()
    assert rc == -1 and b"SideToElem" in L.fx_last_error()
