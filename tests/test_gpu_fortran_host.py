"""The Fortran host (fluxo_amd/fluxo_gpu: a parameter.ini driver with the reference's routine
names, DGTimeDerivative_weakForm / TimeDisc over the C ABI) on the GPU, run as a child process:
the device-resident form (whole steps on the device) and the drop-in form (host RK loop
calling DGTimeDerivative_weakForm once per stage), BR1 and BR2 lifting, against the oracle's
trajectory from the same initial state (relative L2 <= 1e-10 after 5 steps)."""
import math
import os
import subprocess

import numpy as np
import pytest

import pyoracle as O

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "fluxo_amd", "fluxo_gpu")
K, N, STEPS = 4, 3, 5


@pytest.mark.parametrize("resident", [True, False])
@pytest.mark.parametrize("lifting", [1, 2])
def test_fortran_host_matches_oracle(tmp_path, resident, lifting):
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} not built (run __graft_entry__.build())")
    sfile = tmp_path / "state.bin"
    ini = tmp_path / "parameter.ini"
    ini.write_text(
        f"N = {N}\nVolumeFlux = 32\nRiemann = 1\nParabolic = T\nLifting = {lifting}\netaBR2 = 2.5\n"
        "mu0 = 0.01\nPr = 0.72\nkappa = 1.4\nIniExactFunc = 7\nCFLscale = 0.9\nDFLscale = 0.9\n"
        f"maxIter = {STEPS}\nTEnd = 1.0e3\nDeviceResident = {'T' if resident else 'F'}\n"
        f"BoxElems = (/{K},{K},{K}/)\nBoxCorner = (/{-math.pi!r},{-math.pi!r},{-math.pi!r}/)\n"
        f"BoxLength = (/{2 * math.pi!r},{2 * math.pi!r},{2 * math.pi!r}/)\n"
        f"Ngeo = 3\nmeshdeform = T\nmeshdeformAmp = 0.1\nStateFile = {sfile}\n")
    print(f"fluxo_gpu lifting={lifting} resident={resident}", flush=True)
    p = subprocess.run([EXE, str(ini)], capture_output=True, text=True, timeout=120, cwd=tmp_path)
    assert p.returncode == 0 and "FLUXO_GPU FINISHED" in p.stdout, p.stdout[-2000:] + p.stderr[-2000:]
    n = N + 1
    U = np.fromfile(sfile, np.float64).reshape(K ** 3, n, n, n, 5)
    m = O.box_mesh(K, N, x0=(-math.pi,) * 3, length=(2 * math.pi,) * 3, Ngeo=3, warp=True, amp=0.1)
    P = O.Problem(m, volflux=32, riemann=1, parabolic=True, mu0=0.01, Pr=0.72, lifting=lifting, etaBR2=2.5)
    Ue = O.exactfunc(7, m["Elem_xGP"])
    Utmp, t = np.zeros_like(Ue), 0.0
    for _ in range(STEPS):
        d, _ = P.calc_dt(Ue, 0.9, 0.9)
        P.lserk_step(Ue, Utmp, t, d)
        t += d
    err = float(np.linalg.norm(U - Ue) / np.linalg.norm(Ue))
    print(f"  {STEPS} steps relL2 vs oracle {err:.3e}")
    assert err <= 1e-10
    assert f"iter={STEPS}" in p.stdout.replace(" ", "")


def test_fortran_host_error_norms(tmp_path):
    """Euler density wave (IniExactFunc 2) for 10 steps: the driver prints the reference's
    CalcErrorNorms (L2 on NAnalyze points, L2 on the nodes, Linf; fx_analyze on the device)
    of the final state; they match the oracle's analysis of its own trajectory to 1e-9."""
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} not built (run __graft_entry__.build())")
    Kd, Nd, steps = 3, 4, 10
    ini = tmp_path / "parameter.ini"
    ini.write_text(
        f"N = {Nd}\nVolumeFlux = 32\nRiemann = 1\nkappa = 1.4\nIniExactFunc = 2\nAdvVel = (/1.,1.,1./)\n"
        f"IniAmplitude = 0.1\nCFLscale = 0.9\nmaxIter = {steps}\nTEnd = 1.0e3\nBoxElems = (/{Kd},{Kd},{Kd}/)\n"
        "BoxCorner = (/0.,0.,0./)\nBoxLength = (/2.,2.,2./)\nNgeo = 3\nmeshdeform = T\nmeshdeformAmp = 0.1\n")
    print("fluxo_gpu density wave error norms", flush=True)
    p = subprocess.run([EXE, str(ini)], capture_output=True, text=True, timeout=120, cwd=tmp_path)
    assert p.returncode == 0 and "FLUXO_GPU FINISHED" in p.stdout, p.stdout[-2000:] + p.stderr[-2000:]
    got = {}
    for line in p.stdout.splitlines():
        for key in ("L_2_colloc", "L_2", "L_inf"):
            if line.strip().startswith(key + " "):
                got[key] = np.array([float(x) for x in line.split(":")[1].split()])
    m = O.box_mesh(Kd, Nd, x0=(0, 0, 0), length=(2, 2, 2), Ngeo=3, warp=True, amp=0.1)
    P = O.Problem(m, volflux=32, riemann=1)
    Ue = O.exactfunc(2, m["Elem_xGP"])
    Utmp, t = np.zeros_like(Ue), 0.0
    for _ in range(steps):
        d, _ = P.calc_dt(Ue, 0.9, 0.9)
        P.lserk_step(Ue, Utmp, t, d)
        t += d
    a = O.analyze(m, Ue, P.rhs(Ue), exact_func=2, t=t)
    for key, field in (("L_2", "L2"), ("L_2_colloc", "L2colloc"), ("L_inf", "Linf")):
        err = np.abs(got[key] - a[field]) / np.abs(a[field])
        print(f"  {key}: {got[key][0]:.6e} (oracle {a[field][0]:.6e}), max rel diff {err.max():.1e}")
        assert np.all(err <= 1e-9)


def test_fortran_host_runs_the_reference_free_stream_input(tmp_path):
    """The reference's navst freestream/conforming input for the driver: its HOPR mesh
    (MeshFile), its restart state (RestartFile), RefState and tend 20; the driver writes the
    final state in the reference's state file format (ProjectName) and prints CalcErrorNorms,
    whose L_inf must be below the reference test's 1e-11.  The equation keys are the
    reference's parameter.ini's: Riemann 1, VolumeFlux 1 (standard DG with averaged
    metrics), mu0 1e-6, CFL 0.9, DFL 0.5."""
    from fluxo_amd import io
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} not built (run __graft_entry__.build())")
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "io")
    ini = tmp_path / "parameter.ini"
    ini.write_text(
        "ProjectName = NAVIERSTOKES_FREESTREAM\nN = 3\nRiemann = 1\nVolumeFlux = 1\nParabolic = T\n"
        f"MeshFile = {gold}/ConformBoxHeavilyWarped_Ngeo3__03_03_03_mesh.h5\n"
        f"RestartFile = {gold}/NAVIERSTOKES_FREESTREAM_State_0000000.000000000.h5\n"
        "IniExactFunc = 1\nIniRefState = 1\nRefState = (/1.0, 0.1, 0.2, -0.3, 20. /)\nmu0 = 1.0E-06\n"
        "kappa = 1.4\nPr = 0.72\ntEnd = 20.0\nmaxIter = 100000\nCFLscale = 0.9\nDFLscale = 0.5\n")
    print("fluxo_gpu reference freestream input", flush=True)
    p = subprocess.run([EXE, str(ini)], capture_output=True, text=True, timeout=120, cwd=tmp_path)
    assert p.returncode == 0 and "FLUXO_GPU FINISHED" in p.stdout, p.stdout[-2000:] + p.stderr[-2000:]
    linf = None
    for line in p.stdout.splitlines():
        if line.strip().startswith("L_inf "):
            linf = np.array([float(x) for x in line.split(":")[1].split()])
    print(f"  L_inf {linf}")
    assert linf is not None and np.all(linf <= 1e-11)
    # the run ends at tEnd exactly as the reference's does (the last step shortened), and the
    # state file carries that time in its name
    files = sorted(tmp_path.glob("NAVIERSTOKES_FREESTREAM_State_*.h5"))
    assert [f.name for f in files] == ["NAVIERSTOKES_FREESTREAM_State_0000020.000000000.h5"]
    U, t = io.read_state(files[0])
    assert t == pytest.approx(20.0, rel=1e-12) and U.shape == (27, 4, 4, 4, 5)
    assert np.abs(U - [1.0, 0.1, 0.2, -0.3, 50.07]).max() <= 1e-11


@pytest.mark.parametrize("parabolic", [False, True])
@pytest.mark.parametrize("which_mortar", [0, 1])
def test_fortran_host_free_stream_on_a_mortar_mesh(tmp_path, which_mortar, parabolic):
    """The reference's nonconforming free-stream input (tests/runs/navst/freestream/
    nonconforming_projmortar and _collmortar: DeformedBoxMortarPeriodic Ngeo 2, N=4, Riemann 1,
    VolumeFlux 1, whichMortar 0 / 1, RefState (1, 0.1, 0.2, -0.3, 20), mu0 1e-6, CFL 0.9,
    DFL 0.5; and the same for Euler) to tEnd 1 instead of 5: L_inf of CalcErrorNorms below the
    reference test's 1e-11."""
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} not built (run __graft_entry__.build())")
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "io")
    ini = tmp_path / "parameter.ini"
    ini.write_text(
        f"N = 4\nRiemann = 1\nVolumeFlux = 1\nwhichMortar = {which_mortar}\n"
        f"MeshFile = {gold}/DeformedBoxMortarPeriodic__Ngeo_2_Level_01_mesh.h5\n"
        "IniExactFunc = 1\nIniRefState = 1\nRefState = (/1.0, 0.1, 0.2, -0.3, 20. /)\n"
        f"Parabolic = {'T' if parabolic else 'F'}\nmu0 = 1.0E-06\n"
        "kappa = 1.4\ntEnd = 1.0\nCFLscale = 0.9\nDFLscale = 0.5\n")
    print(f"fluxo_gpu mortar free stream whichMortar={which_mortar} parabolic={parabolic}", flush=True)
    p = subprocess.run([EXE, str(ini)], capture_output=True, text=True, timeout=120, cwd=tmp_path)
    assert p.returncode == 0 and "FLUXO_GPU FINISHED" in p.stdout, p.stdout[-2000:] + p.stderr[-2000:]
    linf = None
    for line in p.stdout.splitlines():
        if line.strip().startswith("L_inf "):
            linf = np.array([float(x) for x in line.split(":")[1].split()])
    print(f"  L_inf {linf}")
    assert linf is not None and np.all(linf <= 1e-11)


@pytest.mark.parametrize("which_mortar", [0, 1])
def test_fortran_host_mhd_free_stream_on_a_mortar_mesh(tmp_path, which_mortar):
    """The reference's GLM-MHD nonconforming free-stream input (tests/runs/mhd/freestream/
    nonconforming_projmortar and _collmortar: DeformedBoxMortarPeriodic Ngeo 2, N=4, Riemann 1,
    RefState (1, 0.1, 0.2, -0.3, 20, 0.5, -0.03, 0.02), eta 1e-5, mu 1e-6, CFL 0.9, DFL 0.5)
    with the Derigs volume flux (VolumeFlux 10; the reference's input uses 1, which is not
    built for GLM-MHD here), to tEnd 0.5 instead of 3: L_inf below 1e-11."""
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} not built (run __graft_entry__.build())")
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "io")
    ini = tmp_path / "parameter.ini"
    ini.write_text(
        f"EquationSystem = mhd\nN = 4\nRiemann = 1\nVolumeFlux = 10\nwhichMortar = {which_mortar}\n"
        f"MeshFile = {gold}/DeformedBoxMortarPeriodic__Ngeo_2_Level_01_mesh.h5\n"
        "IniExactFunc = 1\nIniRefState = 1\nRefState = (/1.0, 0.1, 0.2, -0.3, 20., 0.5, -0.03, 0.02 /)\n"
        "Parabolic = T\neta = 1.0E-05\nmu = 1.0E-06\n"
        "tEnd = 0.5\nCFLscale = 0.9\nDFLscale = 0.5\n")
    print(f"fluxo_gpu MHD mortar free stream whichMortar={which_mortar}", flush=True)
    p = subprocess.run([EXE, str(ini)], capture_output=True, text=True, timeout=120, cwd=tmp_path)
    assert p.returncode == 0 and "FLUXO_GPU FINISHED" in p.stdout, p.stdout[-2000:] + p.stderr[-2000:]
    linf = None
    for line in p.stdout.splitlines():
        if line.strip().startswith("L_inf "):
            linf = np.array([float(x) for x in line.split(":")[1].split()])
    print(f"  L_inf {linf}")
    assert linf is not None and np.all(linf <= 1e-11)


GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("resident", [True, False])
def test_fortran_host_mhd_against_reference(tmp_path, resident):
    """Resistive GLM-MHD through the driver's parameter.ini (EquationSystem = mhd and the
    reference's mu_0, GLM_scale, eta, mu, Pr keys): the GLM-MHD benchmark's perturbed
    Orszag-Tang state (RestartFile written from fluxo_amd.workloads.perturbed_orszag_tang) on the
    warped 4^3 mesh, N=4, 100 steps at CFL/DFL 0.5, against the reference's own 100-step run
    (golden mhd_resistive_n4_eta1e-3_100steps): relative L2 <= 1e-10."""
    from fluxo_amd import io
    from fluxo_amd.workloads import perturbed_orszag_tang
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} not built (run __graft_entry__.build())")
    with np.load(os.path.join(GOLD, "mhd", "mhd_resistive_n4_eta1e-3_100steps.npz"), allow_pickle=False) as z:
        g = {k: z[k] for k in z.files}
    K, Nm, steps = int(g["K"]), int(g["N"]), int(g["nsteps"])
    m = O.box_mesh(K, Nm, x0=(0, 0, 0), length=(1, 1, 1), Ngeo=3, warp=True)
    U0 = perturbed_orszag_tang(m["Elem_xGP"], 2026)
    assert np.array_equal(U0, g["U"])
    io.write_state(tmp_path / "ot_start.h5", U0, 0.0)
    sfile = tmp_path / "state.bin"
    (tmp_path / "parameter.ini").write_text(
        f"EquationSystem = mhd\nN = {Nm}\nVolumeFlux = 10\nRiemann = 1\nParabolic = T\nmu_0 = 1.\n"
        f"GLM_scale = 0.5\neta = {float(g['eta'])!r}\nmu = {float(g['mu'])!r}\nPr = {float(g['Pr'])!r}\n"
        "kappa = 1.6666666666666667\nIniExactFunc = 312\nRestartFile = ot_start.h5\n"
        f"CFLscale = 0.5\nDFLscale = 0.5\nmaxIter = {steps}\nTEnd = 1.0e3\n"
        f"DeviceResident = {'T' if resident else 'F'}\nBoxElems = (/{K},{K},{K}/)\nBoxCorner = (/0.,0.,0./)\n"
        f"BoxLength = (/1.,1.,1./)\nNgeo = 3\nmeshdeform = T\nmeshdeformAmp = 0.1\nStateFile = {sfile}\n")
    print(f"fluxo_gpu GLM-MHD resident={resident}", flush=True)
    p = subprocess.run([EXE, "parameter.ini"], capture_output=True, text=True, timeout=120, cwd=tmp_path)
    assert p.returncode == 0 and "FLUXO_GPU FINISHED" in p.stdout, p.stdout[-2000:] + p.stderr[-2000:]
    n = Nm + 1
    U = np.fromfile(sfile, np.float64).reshape(K ** 3, n, n, n, 9)
    err = float(np.linalg.norm(U - g["U_end"]) / np.linalg.norm(g["U_end"]))
    print(f"  {steps} steps relL2 vs the reference {err:.3e}")
    assert err <= 1e-10


def test_fortran_host_mhd_orszag_tang_exactfunc(tmp_path):
    """IniExactFunc 312 (the reference's 3D Orszag-Tang vortex, mhd/equation.f90:1257-1266) in
    the driver's own ExactFunc: 3 steps of ideal GLM-MHD (Parabolic = F) against the oracle from
    the oracle's ExactFunc 312 state."""
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} not built (run __graft_entry__.build())")
    K, Nm, steps = 3, 3, 3
    sfile = tmp_path / "state.bin"
    (tmp_path / "parameter.ini").write_text(
        f"EquationSystem = mhd\nN = {Nm}\nRiemann = 10\nIniExactFunc = 312\nCFLscale = 0.5\n"
        f"maxIter = {steps}\nTEnd = 1.0e3\nBoxElems = (/{K},{K},{K}/)\nBoxCorner = (/0.,0.,0./)\n"
        f"BoxLength = (/1.,1.,1./)\nNgeo = 3\nmeshdeform = T\nStateFile = {sfile}\n")
    p = subprocess.run([EXE, "parameter.ini"], capture_output=True, text=True, timeout=120, cwd=tmp_path)
    assert p.returncode == 0 and "FLUXO_GPU FINISHED" in p.stdout, p.stdout[-2000:] + p.stderr[-2000:]
    m = O.box_mesh(K, Nm, x0=(0, 0, 0), length=(1, 1, 1), Ngeo=3, warp=True)
    P = O.MhdProblem(m, riemann=10)
    Ue = P.exact_orszag_tang()
    Utmp = np.zeros_like(Ue)
    for _ in range(steps):
        d, _ = P.calc_dt(Ue, 0.5, 0.0)
        P.lserk_step(Ue, Utmp, d)
    n = Nm + 1
    U = np.fromfile(sfile, np.float64).reshape(K ** 3, n, n, n, 9)
    err = float(np.linalg.norm(U - Ue) / np.linalg.norm(Ue))
    print(f"  IniExactFunc 312, {steps} steps relL2 vs oracle {err:.3e}")
    assert err <= 1e-10


@pytest.mark.parametrize("resident", [True, False])
def test_fortran_host_shock_capturing_against_reference(tmp_path, resident):
    """Euler with NFVSE shock capturing through the driver's parameter.ini (ShockCapturing = T
    and the reference's InitNFVSE keys): the shock-capturing benchmark's own Sedov-type state
    (RestartFile written from fluxo_amd.workloads.sedov_blast) on the Cartesian 12^3 mesh of
    [-1.5,1.5]^3, N=3, 100 steps at CFL 0.5, against the reference's own 100-step run (golden
    euler_nfvse_c4_sedov_k12_100steps): relative L2 <= 1e-10."""
    from fluxo_amd import io
    from fluxo_amd.workloads import sedov_blast
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} not built (run __graft_entry__.build())")
    with np.load(os.path.join(GOLD, "euler_nfvse_c4_sedov_k12_100steps.npz"), allow_pickle=False) as z:
        g = {k: z[k] for k in z.files}
    K, Nn, steps = int(g["K"]), int(g["N"]), int(g["nsteps"])
    m = O.box_mesh(K, Nn, x0=(-1.5,) * 3, length=(3.0,) * 3)
    io.write_state(tmp_path / "sedov_start.h5", sedov_blast(m["Elem_xGP"]), 0.0)
    sfile = tmp_path / "state.bin"
    (tmp_path / "parameter.ini").write_text(
        f"N = {Nn}\nVolumeFlux = 32\nRiemann = 1\nkappa = 1.4\nShockCapturing = T\nComputeAlpha = 1\n"
        "alpha_max = 0.5\nalpha_min = 0.01\nSpacePropSweeps = 1\nSpacePropFactor = 0.5\nSubFVMethod = 1\n"
        "ShockIndicatorQuantity = DensityTimesPressure\nIniExactFunc = 1\nRestartFile = sedov_start.h5\n"
        f"CFLscale = 0.5\nmaxIter = {steps}\nTEnd = 1.0e3\nDeviceResident = {'T' if resident else 'F'}\n"
        f"BoxElems = (/{K},{K},{K}/)\nBoxCorner = (/-1.5,-1.5,-1.5/)\nBoxLength = (/3.,3.,3./)\nNgeo = 1\n"
        f"StateFile = {sfile}\n")
    print(f"fluxo_gpu Sedov NFVSE resident={resident}", flush=True)
    p = subprocess.run([EXE, "parameter.ini"], capture_output=True, text=True, timeout=120, cwd=tmp_path)
    assert p.returncode == 0 and "FLUXO_GPU FINISHED" in p.stdout, p.stdout[-2000:] + p.stderr[-2000:]
    n = Nn + 1
    U = np.fromfile(sfile, np.float64).reshape(K ** 3, n, n, n, 5)
    err = float(np.linalg.norm(U - g["U_end"]) / np.linalg.norm(g["U_end"]))
    print(f"  {steps} steps relL2 vs the reference {err:.3e}")
    assert err <= 1e-10
    assert f"iter={steps}" in p.stdout.replace(" ", "")


def _restart_run(tmp_path, Ur, K, N, Ngeo, warp, x0=0.0, L=2.0):
    """Run the driver for no steps from a restart file holding Ur (another N); returns the
    state it starts from (StateFile)."""
    from fluxo_amd import io
    io.write_state(tmp_path / "start.h5", Ur, 0.0)
    sfile = tmp_path / "state.bin"
    (tmp_path / "parameter.ini").write_text(
        f"N = {N}\nVolumeFlux = 32\nRiemann = 1\nIniExactFunc = 1\nRestartFile = start.h5\nCFLscale = 0.9\n"
        f"maxIter = 0\nTEnd = 1.0\nBoxElems = (/{K},{K},{K}/)\nBoxCorner = (/{x0},{x0},{x0}/)\n"
        f"BoxLength = (/{L},{L},{L}/)\nNgeo = {Ngeo}\nmeshdeform = {'T' if warp else 'F'}\nStateFile = {sfile}\n")
    p = subprocess.run([EXE, "parameter.ini"], capture_output=True, text=True, timeout=120, cwd=tmp_path)
    assert p.returncode == 0 and "FLUXO_GPU FINISHED" in p.stdout, p.stdout[-2000:] + p.stderr[-2000:]
    assert "interpolating the restart state" in p.stdout
    n = N + 1
    return np.fromfile(sfile, np.float64).reshape(K ** 3, n, n, n, 5)


def _poly_state(x, deg):
    """A state whose variables are polynomials of total degree <= deg per direction."""
    X, Y, Z = x[..., 0], x[..., 1], x[..., 2]
    U = np.empty(x.shape[:-1] + (5,))
    U[..., 0] = 2.0 + 0.1 * X + 0.05 * X * Y ** (deg - 1) + 0.02 * Z ** deg
    U[..., 1] = 0.3 - 0.1 * Y ** deg
    U[..., 2] = 0.1 * X * Z
    U[..., 3] = -0.2 + 0.03 * (X * Y * Z) ** 1
    U[..., 4] = 6.0 + 0.1 * X ** deg * Y
    return U


@pytest.mark.parametrize("Nr,N", [(5, 3), (2, 4)])
def test_fortran_host_restart_change_of_n_polynomial(tmp_path, Nr, N):
    """Restart with a change of N (restart.f90:183-232) on a Cartesian mesh: a state polynomial
    of degree <= min(Nr, N) per direction comes back exactly at the new nodes (interpolation for
    Nr < N; for Nr > N the modal projection keeps every mode up to N), and for Nr > N the
    element integrals of a non-polynomial state are kept (projection of J U with constant J)."""
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} not built (run __graft_entry__.build())")
    K = 2
    mr = O.box_mesh(K, Nr, x0=(0, 0, 0), length=(2, 2, 2))
    m = O.box_mesh(K, N, x0=(0, 0, 0), length=(2, 2, 2))
    deg = min(Nr, N)
    U = _restart_run(tmp_path, _poly_state(mr["Elem_xGP"], deg), K, N, 1, False)
    err = np.max(np.abs(U - _poly_state(m["Elem_xGP"], deg)))
    print(f"restart N={Nr} -> N={N}: polynomial of degree {deg}, max error {err:.2e}")
    assert err <= 1e-12
    if Nr > N:
        Ur = O.exactfunc(7, mr["Elem_xGP"])
        U = _restart_run(tmp_path, Ur, K, N, 1, False)
        wr, w = O.dg_operators(Nr)["w"], O.dg_operators(N)["w"]
        Wr = wr[:, None, None] * wr[None, :, None] * wr[None, None, :]
        W = w[:, None, None] * w[None, :, None] * w[None, None, :]
        ir = np.einsum("kji,ekjiv->ev", Wr, Ur / mr["sJ"][..., None])
        iN = np.einsum("kji,ekjiv->ev", W, U / m["sJ"][..., None])
        d = np.max(np.abs(iN - ir)) / np.max(np.abs(ir))
        print(f"  element integrals of the Taylor-Green state kept to {d:.2e}")
        assert d <= 1e-13


def test_fortran_host_restart_change_of_n_free_stream_warped(tmp_path):
    """On a curved mesh (Ngeo 3, warped) the conservative projection from N=5 to N=3 multiplies
    by J interpolated to N=5 and divides by J at N=3; both come from the same 3 Ngeo Jacobian by
    modal truncation, so a constant state stays constant."""
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} not built (run __graft_entry__.build())")
    K, Nr, N = 2, 5, 3
    c = np.array([1.2, 0.3, -0.2, 0.1, 2.5])
    Ur = np.broadcast_to(c, (K ** 3, Nr + 1, Nr + 1, Nr + 1, 5)).copy()
    U = _restart_run(tmp_path, Ur, K, N, 3, True)
    err = np.max(np.abs(U - c))
    print(f"warped free stream restart N={Nr} -> N={N}: max error {err:.2e}")
    assert err <= 1e-13
