"""Algorithmic flop counts of the benchmark configurations from the counting build of the CPU
oracle (oracle/flopcount/libflopcount.so; test infrastructure only).  Writes
profiles/flops.json: floating-point operations (+ - * / sqrt log exp pow, each 1) per
DOF-update -- one node through one RK stage, counted over a whole TimeStepByLSERKW2 step with
the CALCTIMESTEP before it, divided by 5 x nodes -- and per single flux evaluation.  Counts per
node do not depend on the mesh size on periodic meshes (3 sides per element), so small meshes
are used.  Run:  make -C oracle/flopcount && python oracle/flopcount/flopcount.py"""
import ctypes as C
import json
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, ROOT)
import pyoracle as O  # noqa: E402

L = C.CDLL(os.path.join(HERE, "libflopcount.so"))
L.flopcount_ns_step.restype = C.c_longlong
L.flopcount_ns_step.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_double]
L.flopcount_mhd_step.restype = C.c_longlong
L.flopcount_mhd_step.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_double]
L.flopcount_volflux.restype = C.c_longlong
L.flopcount_volflux.argtypes = [C.c_int, C.c_double] + [C.c_void_p] * 4
L.flopcount_riemann.restype = C.c_longlong
L.flopcount_riemann.argtypes = [C.c_int, C.c_double] + [C.c_void_p] * 5
L.flopcount_derigs.restype = C.c_longlong
L.flopcount_derigs.argtypes = [C.c_void_p] + [C.c_void_p] * 4


def per_dof(flops, mesh):
    return flops / (5.0 * mesh["nElems"] * (mesh["N"] + 1) ** 3)


def ns_case(K, N, x0, L_, ic, parabolic, cfl=0.9, dfl=0.9, **kw):
    m = O.box_mesh(K, N, x0=x0, length=L_, Ngeo=3, warp=True)
    P = O.Problem(m, parabolic=parabolic, mu0=1.0 / 1600 if parabolic else 0.0, Pr=0.72, **kw)
    U = np.ascontiguousarray(ic(m))
    Ut = np.zeros_like(U)
    f = L.flopcount_ns_step(C.addressof(P._p), U.ctypes.data, Ut.ctypes.data, cfl, dfl)
    return per_dof(f, m), P


def main():
    from fluxo_amd.workloads import sedov_blast, perturbed_orszag_tang
    out = {"unit": "flop per DOF-update (one node, all variables, one RK stage; the step's CALCTIMESTEP "
                   "included); per_flux: flop per evaluation",
           "method": "the oracle's C restatement compiled with a counting scalar (oracle/flopcount); "
                     "+ - * / sqrt log exp pow count 1, fabs/fmin/fmax/comparisons 0, a*b+c counts 2",
           "per_dof_update": {}, "per_flux": {}}
    tgv = lambda m: O.exactfunc(7, m["Elem_xGP"])
    dw = lambda m: O.exactfunc(2, m["Elem_xGP"])
    pi = math.pi
    d, _ = ns_case(2, 3, (-pi,) * 3, (2 * pi,) * 3, tgv, True, volflux=32, riemann=1)
    out["per_dof_update"]["navier_stokes_tgv_n3_br1_ranocha_llf"] = d
    d, _ = ns_case(2, 3, (0.0,) * 3, (2.0,) * 3, dw, False, volflux=32, riemann=1)
    out["per_dof_update"]["euler_density_wave_n3_ranocha_llf"] = d
    # shock capturing on the Sedov-type state (alpha > 0 only near the pulse: data dependent)
    m = O.box_mesh(12, 3, x0=(-1.5,) * 3, length=(3.0,) * 3)
    P = O.Problem(m, volflux=32, riemann=1, nfvse=True, alpha_max=0.5, alpha_min=0.01, sweeps=1, prop_factor=0.5)
    U = np.ascontiguousarray(sedov_blast(m["Elem_xGP"]))
    blended = int((P.blending(U) > 0).sum())
    Ut = np.zeros_like(U)
    f = L.flopcount_ns_step(C.addressof(P._p), U.ctypes.data, Ut.ctypes.data, 0.5, 0.5)
    out["per_dof_update"]["euler_nfvse_sedov_n3_k12_initial_state"] = per_dof(f, m)
    out["nfvse_blended_elements"] = f"{blended} of {m['nElems']} at the initial state"
    # resistive GLM-MHD
    m = O.box_mesh(2, 4, x0=(0, 0, 0), length=(1, 1, 1), Ngeo=3, warp=True)
    Mp = O.MhdProblem(m, resistive=dict(mu=0.0, eta=1e-3, Pr=0.72))
    U = np.ascontiguousarray(perturbed_orszag_tang(m["Elem_xGP"], 2026))
    Ut = np.zeros_like(U)
    f = L.flopcount_mhd_step(C.addressof(Mp._m), U.ctypes.data, Ut.ctypes.data, 0.5, 0.5)
    out["per_dof_update"]["mhd_resistive_n4_derigs_rusanov_br1"] = per_dof(f, m)
    # single fluxes
    UL = np.array([1.1, 0.2, -0.1, 0.3, 2.7]); UR = np.array([0.9, -0.1, 0.2, 0.1, 2.4])
    mL = np.array([1.0, 0.1, 0.05]); mR = np.array([0.95, 0.12, 0.04])
    for vf in (32, 5, 10, 0):
        out["per_flux"][f"volume_flux_{vf}"] = L.flopcount_volflux(vf, 1.4, UL.ctypes.data, UR.ctypes.data,
                                                                  mL.ctypes.data, mR.ctypes.data)
    nv = np.array([0.6, 0.8, 0.0]); t1 = np.array([-0.8, 0.6, 0.0]); t2 = np.array([0.0, 0.0, 1.0])
    for rs in (1, 33):
        out["per_flux"][f"riemann_{rs}"] = L.flopcount_riemann(rs, 1.4, UL.ctypes.data, UR.ctypes.data, nv.ctypes.data,
                                                               t1.ctypes.data, t2.ctypes.data)
    V9L = np.array([1.1, 0.2, -0.1, 0.3, 3.7, 0.3, -0.2, 0.1, 0.01])
    V9R = np.array([0.9, -0.1, 0.2, 0.1, 3.4, 0.25, -0.1, 0.2, -0.02])
    out["per_flux"]["mhd_derigs"] = L.flopcount_derigs(C.addressof(Mp._m), V9L.ctypes.data, V9R.ctypes.data,
                                                       mL.ctypes.data, mR.ctypes.data)
    path = os.path.join(ROOT, "profiles", "flops.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
