"""The reference's file formats on the host (fluxo_amd.io over libfluxo_io.so): the HOPR mesh
reader on the reference's own mesh files (tests/golden/io, copied data), and state files.

Pinning of the reader: its connectivity, given to the reference's own CalcMetrics and
DGTimeDerivative (oracle/_ref harness, tests/golden/hopr), preserves the free stream and
conserves entropy with EC fluxes; here, on the CPU: every side has one master and one slave
whose face nodes coincide up to a periodic translation (any wrong flip breaks this), the
oracle preserves the reference's free-stream state file on its mesh, and the oracle
reproduces the reference's metrics, RHS and trajectories on these meshes."""
import ctypes as C
import glob
import os
import shutil
import subprocess

import numpy as np
import pytest

import pyoracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
IO_DIR = os.path.join(HERE, "golden", "io")
HOPR_DIR = os.path.join(HERE, "golden", "hopr")
PERIODIC = ["ConformBoxUnstructPeriodicHeavilyWarped_Ngeo2_mesh.h5", "ConformBoxHeavilyWarped_Ngeo3__03_03_03_mesh.h5",
            "TGV_004_004_004_mesh.h5"]
STATE = os.path.join(IO_DIR, "NAVIERSTOKES_FREESTREAM_State_0000000.000000000.h5")
HOPR_NAMES = sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(HOPR_DIR, "*.npz")))


def io():
    from fluxo_amd import io as fio
    return fio


def side_to_vol(N, p, q, flip, loc):
    ijk = (C.c_int * 3)()
    O.lib().ora_side_to_vol.argtypes = [C.c_int] * 6 + [C.c_int * 3]
    O.lib().ora_side_to_vol(N, 0, p, q, flip, loc, ijk)
    return tuple(ijk)


@pytest.mark.parametrize("mesh_file", PERIODIC)
def test_reader_connectivity_is_consistent(mesh_file):
    r = io().read_mesh(os.path.join(IO_DIR, mesh_file))
    E, S = r["nElems"], r["nSides"]
    e2s, s2e = r["ElemToSide"], r["SideToElem"]
    assert S == 3 * E
    assert np.all(s2e[:, 0] >= 1) and np.all(s2e[:, 1] >= 1) and np.all((s2e[:, 4] >= 1) & (s2e[:, 4] <= 4))
    for s in range(S):
        em, es, lm, ls, fl = s2e[s]
        assert tuple(e2s[em - 1, lm - 1]) == (s + 1, 0)
        assert tuple(e2s[es - 1, ls - 1]) == (s + 1, fl)
    # face nodes of master and slave coincide up to one periodic translation per side
    N = 3
    m = O.mesh_from(N, r["Ngeo"], e2s, s2e, r["NodeCoords"])
    x = m["Elem_xGP"]
    lo, hi = x.reshape(-1, 3).min(0), x.reshape(-1, 3).max(0)
    for s in range(S):
        em, es, lm, ls, fl = s2e[s]
        d = np.array([x[em - 1][side_to_vol(N, p, q, 0, lm)[::-1]] - x[es - 1][side_to_vol(N, p, q, fl, ls)[::-1]]
                      for q in range(N + 1) for p in range(N + 1)])
        assert np.ptp(d, axis=0).max() <= 1e-12, (mesh_file, s)
        shift = d[0]
        assert np.all(np.isclose(shift, 0.0, atol=1e-12) | np.isclose(np.abs(shift), hi - lo, rtol=1e-12)), shift


@pytest.mark.parametrize("mesh_file", PERIODIC)
def test_host_metrics_equal_oracle_and_free_stream(mesh_file):
    """fluxo_amd.mesh.hopr_mesh (reader + the Fortran host's CalcMetrics) equals the oracle's
    metrics bitwise; the reference's free-stream state (its tests/runs/navst/freestream state
    file, a constant RefState) gives a zero RHS to round-off on every mesh."""
    from fluxo_amd.mesh import hopr_mesh
    m = hopr_mesh(os.path.join(IO_DIR, mesh_file), 3)
    o = O.mesh_from(3, m["Ngeo"], m["ElemToSide"], m["SideToElem"], m["NodeCoords"])
    for k in ("Mf", "Mg", "Mh", "sJ", "NormVec", "TangVec1", "TangVec2", "SurfElem", "Elem_xGP"):
        assert np.array_equal(m[k], o[k]), k
    U0, _ = io().read_state(STATE)
    U = np.broadcast_to(U0[0, 0, 0, 0], o["Elem_xGP"].shape[:-1] + (5,)).copy()
    for vf, rs in ((1, 1), (32, 1), (5, 9), (10, 4)):
        Ut = O.Problem(o, volflux=vf, riemann=rs).rhs(U)
        assert np.abs(Ut).max() <= 1e-12 * np.abs(U).max() * np.abs(o["sJ"]).max(), (vf, rs)


@pytest.mark.parametrize("name", HOPR_NAMES)
def test_oracle_matches_reference_on_hopr_meshes(name):
    with np.load(os.path.join(HOPR_DIR, name + ".npz"), allow_pickle=False) as z:
        g = {k: z[k] for k in z.files}
    r = io().read_mesh(os.path.join(IO_DIR, str(g["mesh_file"])))
    N = int(g["N"])
    m = O.mesh_from(N, r["Ngeo"], r["ElemToSide"], r["SideToElem"], r["NodeCoords"])
    for k in ("sJ", "Mf", "NormVec", "SurfElem"):
        assert np.abs(m[k] - g[k]).max() <= 1e-13 * np.abs(g[k]).max(), k
    par = bool(int(g["parabolic"]))
    P = O.Problem(m, volflux=int(g["volflux"]), riemann=int(g["riemann"]), parabolic=par, mu0=float(g["mu0"]),
                  Pr=float(g["Pr"]))
    U = g["U"]
    rel = np.linalg.norm(P.rhs(U) - g["Ut"]) / np.linalg.norm(g["Ut"])
    assert rel <= 1e-13, rel
    dt, err = P.calc_dt(U, float(g["cfl"]), float(g["dfl"]))
    assert err == int(g["errType0"]) and dt == pytest.approx(float(g["dt0"]), rel=1e-14)
    Ue, Utmp, t = U.copy(), np.zeros_like(U), 0.0
    for _ in range(int(g["nsteps"])):
        d, _ = P.calc_dt(Ue, float(g["cfl"]), float(g["dfl"]))
        P.lserk_step(Ue, Utmp, t, d)
        t += d
    assert np.linalg.norm(Ue - g["U_end"]) <= 1e-13 * np.linalg.norm(g["U_end"])
    assert t == pytest.approx(float(g["t_end"]), rel=1e-13)
    if int(g["riemann"]) in (9, 32):  # EC volume + surface flux: the reference's dS/dU.Ut is round-off
        assert abs(float(g["dSdU_Ut"])) <= 1e-12


def test_reference_state_file_and_roundtrip(tmp_path):
    """The reference's state file reads as N=3, 27 elements, Time 0 and the constant RefState
    (1, 0.1, 0.2, -0.3, p=20) of its parameter file; a written state reads back bitwise with
    the reference's attributes and dataset."""
    N, nVar, E, t = io().state_info(STATE)
    assert (N, nVar, E, t) == (3, 5, 27, 0.0)
    U, _ = io().read_state(STATE)
    assert np.allclose(U, [1.0, 0.1, 0.2, -0.3, 20.0 / 0.4 + 0.5 * (0.01 + 0.04 + 0.09)], rtol=0, atol=1e-14)
    V = np.random.default_rng(1).uniform(0.5, 2.0, (7, 4, 4, 4, 5))
    f = tmp_path / "PROJ_State_0000000.250000000.h5"
    io().write_state(f, V, 0.25, mesh_file="mesh.h5", project="PROJ")
    W, tw = io().read_state(f)
    assert tw == 0.25 and np.array_equal(V, W)
    h5dump = shutil.which("h5dump") or "/opt/conda/bin/h5dump"
    if os.path.exists(h5dump):
        hdr = subprocess.run([h5dump, "-H", str(f)], capture_output=True, text=True).stdout
        for key in ("File_Type", "MeshFile", "NodeType", "VarNames", "NComputation", "Program", "Project_Name",
                    "File_Version", '"N"', '"Time"', "DG_Solution", "( 7, 4, 4, 4, 5 )"):
            assert key in hdr, key


def test_reader_rejects_what_the_engine_cannot_run():
    with pytest.raises(io().IOError_, match="periodic"):
        io().read_mesh(os.path.join(IO_DIR, "ConformBoxUnstruct2_Ngeo2_mesh.h5"))
    with pytest.raises(io().IOError_, match="cannot open"):
        io().read_mesh(os.path.join(IO_DIR, "missing.h5"))
    V = np.zeros((27, 5, 5, 5, 5))
    with pytest.raises(io().IOError_, match="N=3, computation N=4"):
        io()._chk(io().lib().fxio_state_read(os.fsencode(STATE), 4, 5, 27, V.ctypes.data))
