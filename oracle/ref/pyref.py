"""Driver for oracle/_ref/fluxo_ref_* (the reference's own code behind a harness).

TEST INFRASTRUCTURE ONLY.  Runs in the container that holds /root/reference (the binaries
are built there by oracle/ref/build_ref.sh); used to generate golden vectors under
tests/golden/ and to pin the C oracle.  Array shapes follow oracle/pyoracle.py.
"""
import os
import re
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(os.path.dirname(HERE), "_ref")


def binary(parabolic, volflux, nfvse=False, mhd=False, lifting=1, fast=False):
    """Path of the harness build for the configuration; fast: the -O3 x86-64-v3 build (made only
    for Navier-Stokes BR1 with VolumeFlux 32, the CPU baseline's configuration)."""
    if mhd:
        name = f"fluxo_ref_mhd{f'_br{lifting}' if parabolic else ''}_vf{volflux}"
    else:
        ns = f"ns_br{lifting}"
        name = f"fluxo_ref_{ns if parabolic else 'euler'}{'_nfvse' if nfvse else ''}_vf{volflux}"
    if fast:
        name += "_fast"
    path = os.path.join(REF_DIR, name)
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} not built (oracle/ref/build_ref.sh)")
    return path


def analyze_lines(analyze):
    """parameter.ini lines of the reference's Analyze: bulk integrals and AnalyzeToFile on,
    plus the given options (AnalyzeExactFunc, NAnalyze, AdvVel, IniAmplitude, ...)."""
    lines = "CalcBulk = T\nAnalyzeToFile = T\n"
    for k, v in (analyze or {}).items():
        if isinstance(v, (tuple, list)):
            v = "(/" + ",".join(repr(float(x)) for x in v) + "/)"
        lines += f"{k} = {v}\n"
    return lines


def read_analysis(d):
    """The 'name value' lines the harness writes from the reference's AnalyzeToFile data."""
    out = {}
    path = os.path.join(d, "analysis.txt")
    if os.path.exists(path):
        with open(path) as f:
            for line in f:
                k, v = line.rsplit(None, 1)
                out[k.strip().strip('"')] = float(v)
    return out


def _mortar_input(mesh, f):
    """MortarType / MortarInfo of a mesh with non-conforming interfaces after the connectivity."""
    if mesh.get("nMortarSides", 0):
        np.ascontiguousarray(mesh["MortarType"], np.int32).tofile(f)
        np.ascontiguousarray(mesh["MortarInfo"], np.int32).tofile(f)


def run(mesh, U, volflux=32, riemann=1, parabolic=False, kappa=1.4, mu0=0.0, Pr=0.72,
        cfl=0.9, dfl=0.9, nsteps=0, t0=0.0, timeout=600, nfvse=None, lifting=1, etaBR2=2.0, analyze=None,
        fast=False, whichMortar=0):
    """Metrics, VolInt, RHS (+ gradients), the face fluxes of FillFlux, entropy, dt and an
    nsteps trajectory computed by the reference's code on the given mesh (connectivity +
    NodeCoords) and state.
    nfvse: None, or dict of NFVSE parameters (ComputeAlpha, ShockBlendCoef, alpha_max,
    alpha_min, SpacePropSweeps, SpacePropFactor) for the shock-capturing build, which also
    returns the blending coefficient alpha of the RHS evaluation and VolInt_NFVSE's blended
    volume integral of the state (volint_nfvse).  lifting 2: the BR2 build
    with penalty etaBR2 (gradients returned are the volume gradients).  out["analysis"]
    holds what the reference's Analyze collects for this state at time t0 (error norms
    against AnalyzeExactFunc, bulk integrals, entropy); `analyze` adds parameters to it.
    A mesh with mortars (nMortarSides > 0) runs with the reference's whichMortar (0 L2
    projection, 1 collocation); its Flux_slave covers the inner sides only."""
    N, Ngeo, E, S = mesh["N"], mesh["Ngeo"], mesh["nElems"], mesh["nSides"]
    nm = int(mesh.get("nMortarSides", 0))
    n = N + 1
    exe = binary(parabolic, volflux, nfvse is not None, lifting=lifting, fast=fast)
    with tempfile.TemporaryDirectory() as d:
        par = os.path.join(d, "parameter.ini")
        with open(par, "w") as f:
            analyze = dict(analyze or {})
            ini = analyze.pop("IniExactFunc", 1)
            f.write(f"N = {N}\nRiemann = {riemann}\nVolumeFlux = {volflux}\nIniExactFunc = {ini}\n"
                    f"IniRefState = 1\nRefState = (/1.,0.,0.,0.,1./)\nkappa = {kappa!r}\n"
                    f"mu0 = {mu0!r}\nPr = {Pr!r}\nCalcEntropy = T\nwhichMortar = {int(whichMortar)}\n")
            f.write(analyze_lines(analyze))
            if lifting == 2:
                f.write(f"etaBR2 = {etaBR2!r}\n")
            if nfvse is not None:
                f.write("ShockIndicator = 1\nShockIndicatorQuantity = DensityTimesPressure\n"
                        "SubFVMethod = 1\nModalThreshold = 1\nTimeRelFactor = 0.0\n")
                for k, v in nfvse.items():
                    f.write(f"{k} = {v!r}\n")
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(fin, "wb") as f:
            np.array([N, Ngeo, E, S, nsteps, int(parabolic), nm, 0], np.int32).tofile(f)
            np.array([cfl, dfl, t0, 0.0], np.float64).tofile(f)
            np.ascontiguousarray(mesh["NodeCoords"], np.float64).tofile(f)
            np.ascontiguousarray(mesh["ElemToSide"], np.int32).tofile(f)
            np.ascontiguousarray(mesh["SideToElem"], np.int32).tofile(f)
            _mortar_input(mesh, f)
            np.ascontiguousarray(U, np.float64).tofile(f)
        p = subprocess.run([exe, par, fin, fout], cwd=d, capture_output=True, text=True, errors="replace",
                           timeout=timeout)
        if p.returncode != 0 or not os.path.exists(fout):
            raise RuntimeError(f"reference harness failed ({p.returncode}):\n{p.stdout[-2000:]}\n{p.stderr[-2000:]}")
        raw = np.fromfile(fout, np.float64)
        analysis = read_analysis(d)
    out, o = {"analysis": analysis}, 0

    def take(name, shape):
        nonlocal o
        cnt = int(np.prod(shape))
        out[name] = raw[o:o + cnt].reshape(shape)
        o += cnt

    for k in ("Mf", "Mg", "Mh"):
        take(k, (E, n, n, n, 3))
    take("sJ", (E, n, n, n))
    for k in ("NormVec", "TangVec1", "TangVec2"):
        take(k, (S, n, n, 3))
    take("SurfElem", (S, n, n))
    take("volint", (E, n, n, n, 5))
    take("Ut", (E, n, n, n, 5))
    if parabolic:
        for k in ("gradPx", "gradPy", "gradPz"):
            take(k, (E, n, n, n, 5))
    if nfvse is not None:
        take("alpha", (E,))
    take("Flux_master", (S, n, n, 5))
    take("Flux_slave", (S - nm, n, n, 5))
    if nfvse is not None:
        take("volint_nfvse", (E, n, n, n, 5))
    take("dt0", (2,))
    out["errType0"] = int(out["dt0"][1])
    out["dt0"] = float(out["dt0"][0])
    take("U_end", (E, n, n, n, 5))
    take("tail", (2,))
    out["t_end"], out["dt_last"] = float(out["tail"][0]), float(out["tail"][1])
    del out["tail"]
    m = re.search(r"Entropy\s*:\s*([-+0-9.Ee]+)", p.stdout)
    out["Entropy"] = float(m.group(1)) if m else None
    m = re.search(r"dSdU\*Ut\s*:\s*([-+0-9.Ee]+)", p.stdout)
    out["dSdU_Ut"] = float(m.group(1)) if m else None
    return out


def run_mhd(mesh, U, kappa=5.0 / 3.0, mu_0=1.0, glm_scale=0.5, cfl=0.9, nsteps=0, t0=0.0, timeout=600,
            resistive=None, dfl=0.9, riemann=1, analyze=None, tgv=False, lifting=1, etaBR2=2.0, whichMortar=0):
    """GLM-MHD (Derigs volume flux, surface flux `riemann` -- 0 central, 1 Rusanov, 10 Derigs +
    entropy-stable LLF, 11 Derigs -- Powell non-conservative terms) with
    the reference's code: metrics, VolInt, RHS (GLM_ch = 0.5, its value before the first
    CALCTIMESTEP), dt with GLM_ch and GLM_dtch1 after it, and an nsteps trajectory.
    resistive: None (ideal), or dict(mu=, eta=, Pr=) for the BR1 build with viscosity,
    resistivity and heat conduction, which also returns the primitive gradients.
    U: (nElems, n, n, n, 9).  out["analysis"] as in run(); tgv=True runs the build with the
    Taylor-Green vortex testcase (resistive only), whose AnalyzeTestcase adds energies,
    enstrophies and dissipation rates.  Meshes with mortars as in run()."""
    N, Ngeo, E, S = mesh["N"], mesh["Ngeo"], mesh["nElems"], mesh["nSides"]
    nm = int(mesh.get("nMortarSides", 0))
    n = N + 1
    par_on = resistive is not None
    exe = binary(par_on, 10, mhd=True, lifting=lifting)
    if tgv:
        exe = exe + "_tgv"
    with tempfile.TemporaryDirectory() as d:
        par = os.path.join(d, "parameter.ini")
        with open(par, "w") as f:
            analyze = dict(analyze or {})
            ini = analyze.pop("IniExactFunc", 1)
            f.write(f"N = {N}\nRiemann = {int(riemann)}\nVolumeFlux = 10\nIniExactFunc = {ini}\nIniRefState = 1\n"
                    f"RefState = (/1.,0.,0.,0.,1.,0.,0.,0./)\nkappa = {kappa!r}\nmu_0 = {mu_0!r}\n"
                    f"GLM_scale = {glm_scale!r}\nwhichMortar = {int(whichMortar)}\n")
            f.write(analyze_lines(analyze))
            if tgv:
                f.write("doTCanalyze = T\n")
            if par_on and lifting == 2:
                f.write(f"etaBR2 = {etaBR2!r}\n")
            if par_on:
                f.write(f"mu = {float(resistive['mu'])!r}\neta = {float(resistive['eta'])!r}\n"
                        f"Pr = {float(resistive.get('Pr', 0.72))!r}\n")
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(fin, "wb") as f:
            np.array([N, Ngeo, E, S, nsteps, int(par_on), nm, 0], np.int32).tofile(f)
            np.array([cfl, dfl if par_on else 0.0, t0, 0.0], np.float64).tofile(f)
            np.ascontiguousarray(mesh["NodeCoords"], np.float64).tofile(f)
            np.ascontiguousarray(mesh["ElemToSide"], np.int32).tofile(f)
            np.ascontiguousarray(mesh["SideToElem"], np.int32).tofile(f)
            _mortar_input(mesh, f)
            np.ascontiguousarray(U, np.float64).tofile(f)
        p = subprocess.run([exe, par, fin, fout], cwd=d, capture_output=True, text=True, errors="replace",
                           timeout=timeout)
        if p.returncode != 0 or not os.path.exists(fout):
            raise RuntimeError(f"reference harness failed ({p.returncode}):\n{p.stdout[-2000:]}\n{p.stderr[-2000:]}")
        raw = np.fromfile(fout, np.float64)
        analysis = read_analysis(d)
    out, o = {"analysis": analysis}, 0

    def take(name, shape):
        nonlocal o
        cnt = int(np.prod(shape))
        out[name] = raw[o:o + cnt].reshape(shape)
        o += cnt

    for k in ("Mf", "Mg", "Mh"):
        take(k, (E, n, n, n, 3))
    take("sJ", (E, n, n, n))
    for k in ("NormVec", "TangVec1", "TangVec2"):
        take(k, (S, n, n, 3))
    take("SurfElem", (S, n, n))
    take("volint", (E, n, n, n, 9))
    take("Ut", (E, n, n, n, 9))
    if par_on:
        for k in ("gradPx", "gradPy", "gradPz"):
            take(k, (E, n, n, n, 9))
    take("Flux_master", (S, n, n, 9))
    take("Flux_slave", (S - nm, n, n, 9))
    take("dt0", (2,))
    out["errType0"] = int(out["dt0"][1])
    out["dt0"] = float(out["dt0"][0])
    take("glm", (2,))
    out["glm_ch0"], out["glm_dtch1"] = float(out["glm"][0]), float(out["glm"][1])
    del out["glm"]
    take("U_end", (E, n, n, n, 9))
    take("tail", (2,))
    out["t_end"], out["dt_last"] = float(out["tail"][0]), float(out["tail"][1])
    del out["tail"]
    out["stdout"] = p.stdout
    return out
