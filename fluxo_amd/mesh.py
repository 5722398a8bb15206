"""Meshes for the engine: periodic box meshes built by the Fortran host (libfluxo_host.so,
fxh_box_mesh), and HOPR mesh files (fluxo_amd.io) with the host's metrics (fxh_metrics).

Returns the arrays the engine's fx_mesh_upload takes, in the reference layout (numpy C
order with the Fortran index order reversed), plus NodeCoords and the volume node
coordinates Elem_xGP.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
HOST_LIB = os.path.join(_HERE, "libfluxo_host.so")
_LIB = None


def _lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(HOST_LIB):
            raise RuntimeError(f"Fortran host not built: {HOST_LIB} missing")
        L = C.CDLL(HOST_LIB)
        dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        ci, cd = C.c_int, C.c_double
        L.fxh_box_mesh.argtypes = [ci, ci, ci, ci, ci, dp, dp, ci, cd, ip, ip] + [dp] * 10 + [
            C.POINTER(ci)]
        L.fxh_box_mesh.restype = None
        L.fxh_metrics.argtypes = [ci, ci, ci, ci, ip, dp] + [dp] * 9 + [C.POINTER(ci)]
        L.fxh_metrics.restype = None
        L.fxh_metrics_nc.argtypes = [ci, ci, ci, ci, ci, ci, ip, ip, ip, dp] + [dp] * 9 + [C.POINTER(ci)]
        L.fxh_metrics_nc.restype = None
        _LIB = L
    return _LIB


def box_mesh(K, N, x0=(0.0, 0.0, 0.0), length=(2.0, 2.0, 2.0), Ngeo=1, warp=False, amp=0.1):
    Kx, Ky, Kz = (K, K, K) if np.isscalar(K) else K
    E, n, ng = Kx * Ky * Kz, N + 1, Ngeo + 1
    S = 3 * E
    m = dict(N=N, Ngeo=Ngeo, nElems=E, nSides=S, K=(Kx, Ky, Kz))
    m["ElemToSide"] = np.zeros((E, 6, 2), np.int32)
    m["SideToElem"] = np.zeros((S, 5), np.int32)
    m["NodeCoords"] = np.zeros((E, ng, ng, ng, 3))
    for k in ("Mf", "Mg", "Mh", "Elem_xGP"):
        m[k] = np.zeros((E, n, n, n, 3))
    m["sJ"] = np.zeros((E, n, n, n))
    for k in ("NormVec", "TangVec1", "TangVec2"):
        m[k] = np.zeros((S, n, n, 3))
    m["SurfElem"] = np.zeros((S, n, n))
    ierr = C.c_int(0)
    _lib().fxh_box_mesh(Kx, Ky, Kz, N, Ngeo, np.array(x0, float), np.array(length, float),
                        int(bool(warp)), amp, m["ElemToSide"], m["SideToElem"], m["NodeCoords"],
                        m["Mf"], m["Mg"], m["Mh"], m["sJ"], m["NormVec"], m["TangVec1"],
                        m["TangVec2"], m["SurfElem"], m["Elem_xGP"], C.byref(ierr))
    if ierr.value != 0:
        raise ValueError(f"box mesh failed (ierr={ierr.value}): non-positive or badly scaled Jacobian")
    return m


def hopr_mesh(path, N):
    """A periodic HOPR mesh file (the reference's MeshFile), conforming or with mortars, for
    polynomial degree N: the reader's connectivity and node coordinates, the host's
    CalcMetrics/CalcSurfMetrics (with Mortar_CalcSurfMetrics on the small sides of a mortar).
    With mortars the dict also holds nMortarSides, MortarType (nSides, 2) and MortarInfo
    (nMortarSides, 5, 2)."""
    from fluxo_amd import io
    r = io.read_mesh(path)
    E, S, Ngeo, n = r["nElems"], r["nSides"], r["Ngeo"], N + 1
    m = dict(N=N, Ngeo=Ngeo, nElems=E, nSides=S, MeshFile=os.fspath(path))
    m.update(ElemToSide=r["ElemToSide"], SideToElem=r["SideToElem"], NodeCoords=r["NodeCoords"])
    for k in ("Mf", "Mg", "Mh", "Elem_xGP"):
        m[k] = np.zeros((E, n, n, n, 3))
    m["sJ"] = np.zeros((E, n, n, n))
    for k in ("NormVec", "TangVec1", "TangVec2"):
        m[k] = np.zeros((S, n, n, 3))
    m["SurfElem"] = np.zeros((S, n, n))
    ierr = C.c_int(0)
    nm = r["nMortarSides"]
    if nm == 0:
        _lib().fxh_metrics(N, Ngeo, E, int(r["nodes_gl"]), m["ElemToSide"], np.ascontiguousarray(m["NodeCoords"]),
                           m["Mf"], m["Mg"], m["Mh"], m["sJ"], m["NormVec"], m["TangVec1"], m["TangVec2"],
                           m["SurfElem"], m["Elem_xGP"], C.byref(ierr))
    else:
        m.update(nMortarSides=nm, MortarType=r["MortarType"], MortarInfo=r["MortarInfo"])
        _lib().fxh_metrics_nc(N, Ngeo, E, S, nm, int(r["nodes_gl"]), m["ElemToSide"], r["MortarType"],
                              np.ascontiguousarray(r["MortarInfo"]), np.ascontiguousarray(m["NodeCoords"]),
                              m["Mf"], m["Mg"], m["Mh"], m["sJ"], m["NormVec"], m["TangVec1"], m["TangVec2"],
                              m["SurfElem"], m["Elem_xGP"], C.byref(ierr))
    if ierr.value != 0:
        raise ValueError(f"{path}: metrics failed (ierr={ierr.value}): non-positive or badly scaled Jacobian")
    return m
