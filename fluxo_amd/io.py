"""ctypes binding of libfluxo_io.so (include/fluxo_io.h): HOPR mesh files and state files in
the reference's HDF5 formats.  Arrays use numpy C order with the Fortran index order
reversed (ElemToSide (nElems, 6, 2), SideToElem (nSides, 5), NodeCoords (nElems, ng, ng,
ng, 3), U (nElems, n, n, n, nVar)), so their memory equals the reference's arrays."""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
IO_LIB = os.path.join(_HERE, "libfluxo_io.so")
_LIB = None

VAR_NAMES = {5: ("Density", "MomentumX", "MomentumY", "MomentumZ", "EnergyStagnationDensity"),
             9: ("Density", "MomentumX", "MomentumY", "MomentumZ", "TotalEnergy", "MagneticFieldX",
                 "MagneticFieldY", "MagneticFieldZ", "Psi")}


class IOError_(RuntimeError):
    pass


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(IO_LIB):
            raise IOError_(f"{IO_LIB} not built (run __graft_entry__.build())")
        L = C.CDLL(IO_LIB)
        ip, dp = C.POINTER(C.c_int), C.POINTER(C.c_double)
        L.fxio_last_error.restype = C.c_char_p
        L.fxio_mesh_info.argtypes = [C.c_char_p, ip, ip, ip, ip]
        L.fxio_mesh_read.argtypes = [C.c_char_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.fxio_mesh_mortars.argtypes = [C.c_char_p, ip, C.c_void_p, C.c_void_p]
        L.fxio_state_write.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_double,
                                       C.POINTER(C.c_char_p), C.c_void_p]
        L.fxio_state_info.argtypes = [C.c_char_p, ip, ip, ip, dp]
        L.fxio_state_read.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_void_p]
        _LIB = L
    return _LIB


def _chk(rc):
    if rc != 0:
        raise IOError_(f"fluxo_io error {rc}: {lib().fxio_last_error().decode()}")


def mesh_info(path):
    """(nElems, nSides, Ngeo, nodes_gl) of a HOPR mesh file."""
    v = [C.c_int() for _ in range(4)]
    _chk(lib().fxio_mesh_info(os.fsencode(path), *(C.byref(x) for x in v)))
    return tuple(x.value for x in v[:3]) + (bool(v[3].value),)


def read_mesh(path):
    """Connectivity and node coordinates of a periodic HOPR mesh, conforming or with mortars:
    MortarType (nSides, 2) and MortarInfo (nMortarSides, 5, 2) as fxio_mesh_mortars gives
    them (big mortar sides are sides 1..nMortarSides)."""
    E, S, Ngeo, gl = mesh_info(path)
    ng = Ngeo + 1
    e2s = np.zeros((E, 6, 2), np.int32)
    s2e = np.zeros((S, 5), np.int32)
    X = np.zeros((E, ng, ng, ng, 3))
    _chk(lib().fxio_mesh_read(os.fsencode(path), e2s.ctypes.data, s2e.ctypes.data, X.ctypes.data))
    nm = C.c_int()
    _chk(lib().fxio_mesh_mortars(os.fsencode(path), C.byref(nm), None, None))
    mt = np.zeros((S, 2), np.int32)
    mi = np.zeros((nm.value, 5, 2), np.int32)
    _chk(lib().fxio_mesh_mortars(os.fsencode(path), C.byref(nm), mt.ctypes.data, mi.ctypes.data))
    return dict(nElems=E, nSides=S, Ngeo=Ngeo, nodes_gl=gl, ElemToSide=e2s, SideToElem=s2e, NodeCoords=X,
                nMortarSides=nm.value, MortarType=mt, MortarInfo=mi)


def write_state(path, U, time, mesh_file="", project="FLUXO_AMD"):
    """WriteState: U (nElems, n, n, n, nVar) at `time` in the reference's state file format."""
    U = np.ascontiguousarray(U, dtype=np.float64)
    E, n, nVar = U.shape[0], U.shape[1], U.shape[-1]
    names = (C.c_char_p * nVar)(*(s.encode() for s in VAR_NAMES.get(nVar, [f"Var{i + 1}" for i in range(nVar)])))
    _chk(lib().fxio_state_write(os.fsencode(path), project.encode(), mesh_file.encode(), n - 1, nVar, E, time,
                                names, U.ctypes.data))


def state_info(path):
    """(N, nVar, nElems, Time) of a state file."""
    n, nv, ne, t = C.c_int(), C.c_int(), C.c_int(), C.c_double()
    _chk(lib().fxio_state_info(os.fsencode(path), C.byref(n), C.byref(nv), C.byref(ne), C.byref(t)))
    return n.value, nv.value, ne.value, t.value


def read_state(path):
    """Restart read: (U (nElems, n, n, n, nVar), Time)."""
    N, nVar, E, t = state_info(path)
    n = N + 1
    U = np.zeros((E, n, n, n, nVar))
    _chk(lib().fxio_state_read(os.fsencode(path), N, nVar, E, U.ctypes.data))
    return U, t
