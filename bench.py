#!/usr/bin/env python3
"""Benchmark of the MI355X split-form DGSEM engine on one GPU.

Workload (default): compressible Navier-Stokes Taylor-Green vortex, FP64, Re=1600, Ma=0.1,
Pr=0.72, on [-pi,pi]^3 periodic with 32^3 hexahedra whose Ngeo=3 geometry is warped by
0.1*prod sin(x_k+pi); N=3 Legendre-Gauss-Lobatto (2,097,152 nodes x 5 variables); BR1
lifting with entropy variables; Ranocha EC volume flux (VolumeFlux=32) + LLF surface flux
(Riemann=1); Carpenter RK4-5 with CALCTIMESTEP (CFL 0.9, DFL 0.9) before every step.

A step is one full RK time step (5 right-hand-side evaluations and stage updates); the
metric counts DOF-updates: one node (all 5 variables) advanced through one RK stage.
The mesh and initial state are synthetic (generated here); U is resident in HBM before
the timed region and stays there.

The same workload is then run by the split-precision engine (FP32 flux differencing and
accumulation, FP64 RK/CFL/surface/lifting): its throughput and the relative L2/Linf
difference of its final state from the FP64 run are reported under "fp32_split".

Serving: a stream of independent requests (fluxo_amd.workloads, 512 by default, 50 steps
each) goes through the batched server, first all at once (offline throughput), then as a
Poisson stream at 0.7x that request rate (p50/p99 latency from submission to result on the
host), in FP64 and FP32-split, with the admission cap SERVE_ADMIT_KNODES (and, for FP64,
without it); every request also runs alone through its own context, back to back (the plain
per-request path), for the batched/plain throughput ratio and the FIFO latency a
one-at-a-time server would give at the same arrivals.

Prints one JSON line.  Only one GPU is used: --gpus > 1 exits with a message.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "DOF-updates/s on 1 GPU (FP64 vs FP32-split) + p50/p99 serving latency"
HBM_PEAK_GBS = 8000.0
CPU_SHARE = 16
FP64_PEAK_TFLOPS = 78.6


def tgv_state(xgp, kappa=1.4):
    """Taylor-Green vortex, Ma=0.1 (the reference's ExactFunc 7, equation.f90:802-811)."""
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


def volume_kernel_bytes(N, nElems, parabolic, metric_bytes=8):
    """Algorithmic HBM bytes of one launch of the volume kernel in its RK-stage form:
    U read + write, Ut_tmp read + write, metrics (9 per node, FP64 or the FP32 copies of
    the split-precision path), sJ (1), the face fluxes of the element's 3 owned sides
    (each side's flux is read once per side), and for Navier-Stokes the 4 components of
    the weak-form viscous volume term per node (formed by the lifting kernel)."""
    n = N + 1
    n3, n2 = n ** 3, n ** 2
    per_elem = (8 * (5 * n3 * 4 + n3 + 3 * n2 * 5 + (4 * n3 if parabolic else 0))
                + metric_bytes * 9 * n3)
    return per_elem * nElems


def lifting_kernel_bytes(N, nElems):
    """Algorithmic HBM bytes of one launch of the Navier-Stokes lifting kernel (viscous
    form): U, metrics and sJ of every node; at each of the element's 6 faces the
    neighbour's face-node state (5) and the face normal with SurfElem (4), located through the
    face-point table (two 4-byte integers per face point); written: the 4-component viscous
    volume term per node and the 4-component normal viscous flux at each of the 6 faces."""
    n = N + 1
    n3, n2 = n ** 3, n ** 2
    return (8 * (5 * n3 + 9 * n3 + n3 + 6 * n2 * 5 + 6 * n2 * 4 + 4 * n3 + 6 * n2 * 4) + 8 * 6 * n2) * nElems


def surface_kernel_bytes(N, nSides, parabolic):
    """Algorithmic HBM bytes of one launch of the surface kernel: per face point both
    sides' states, normal + SurfElem, (Navier-Stokes) both sides' normal viscous fluxes,
    the numerical flux written; SideToElem per side."""
    n2 = (N + 1) ** 2
    return (8 * n2 * (2 * 5 + 4 + (8 if parabolic else 0) + 5) + 20) * nSides


def mhd_volume_bytes(N, nElems, parabolic, metric_bytes=8):
    """Algorithmic HBM bytes of one launch of the GLM-MHD volume kernel in its RK-stage form:
    U (9 variables) read + write, Ut_tmp read + write, metrics, sJ, the master- and
    slave-side face fluxes (2 x 9 values per face point, 3 sides per element) and, in the
    resistive case, the 7 components of the weak-form viscous volume term per node."""
    n = N + 1
    n3, n2 = n ** 3, n ** 2
    per_elem = (8 * (9 * n3 * 4 + n3 + 3 * n2 * 9 * 2 + (7 * n3 if parabolic else 0))
                + metric_bytes * 9 * n3)
    return per_elem * nElems


def mhd_leg(K=24, N=4, steps=100, warmup=5, cfl=0.5, dfl=0.5, eta=1.0e-3, mu=0.0, seed=2026):
    """Resistive GLM-MHD (9 variables): the 3D Orszag-Tang vortex (ExactFunc 312) with a seeded
    1e-3 uniform perturbation of rho and B (fluxo_amd.workloads.perturbed_orszag_tang) on a
    periodic K^3 mesh of warped hexes (Ngeo=3) of [0,1]^3, N=4, Derigs EC volume flux +
    Rusanov with Powell non-conservative terms, GLM_scale 0.5, resistivity eta, viscosity mu,
    BR1 entropy-variable lifting.  FP64 and FP32-split each run `warmup` + `steps` LSERK steps
    (CFL/DFL before each) from the same state; the `steps` are timed; the FP32-split state is
    compared with the FP64 one at the end.  CFLscale 0.5: at 0.9 the grid-scale perturbation
    grows until the density is NaN (the reference's own code does the same on a 12^3 mesh
    after ~35 steps, as does the oracle; the 24^3 mesh fails within 10 steps at 0.9 and runs
    past 200 steps at 0.5)."""
    from fluxo_amd.mesh import box_mesh
    from fluxo_amd.native import (MhdEngine, FX_KERNEL_VOLUME_RK, FX_KERNEL_SURFACE, FX_KERNEL_LIFTING,
                                  FX_KERNEL_DT, FX_PREC_FP64, FX_PREC_FP32_SPLIT)
    from fluxo_amd.workloads import perturbed_orszag_tang
    m = box_mesh(K, N, x0=(0.0,) * 3, length=(1.0,) * 3, Ngeo=3, warp=True)
    U0 = perturbed_orszag_tang(m["Elem_xGP"], seed)
    res = dict(mu=mu, eta=eta, Pr=0.72)
    nodes = m["nElems"] * (N + 1) ** 3
    out = {}
    for name, prec in (("fp64", FX_PREC_FP64), ("fp32_split", FX_PREC_FP32_SPLIT)):
        e = MhdEngine(m, precision=prec, resistive=res)
        e.upload(U0)
        e.run(warmup, cfl, dfl)
        e.sync()
        t0 = time.perf_counter()
        t_end, dt_last, err = e.run(steps, cfl, dfl)
        e.sync()
        el = time.perf_counter() - t0
        U = e.download()
        reps = 10
        ms = {"volume": e.profile_kernel(FX_KERNEL_VOLUME_RK, reps),
              "surface": e.profile_kernel(FX_KERNEL_SURFACE, reps),
              "lifting": e.profile_kernel(FX_KERNEL_LIFTING, reps),
              "calc_dt": e.profile_kernel(FX_KERNEL_DT, reps)}
        ch, _ = e.glm()
        e.close()
        vb = mhd_volume_bytes(N, m["nElems"], True, 8 if prec == FX_PREC_FP64 else 4)
        # a DOF-update is one node (all 9 variables) through one of the 5 RK stages
        out[name] = {"value": nodes * 5 * steps / el, "unit": "DOF-updates/s", "ms_per_step": el / steps * 1e3,
                     "errType": err, "t_end": t_end, "dt_last": dt_last, "GLM_ch": ch,
                     "kernel_ms": {k: round(v, 4) for k, v in ms.items()},
                     "volume_kernel_GBps": round(vb / (ms["volume"] * 1e-3) / 1e9, 1), "U": U}
    U64, U32 = out["fp64"].pop("U"), out["fp32_split"].pop("U")
    d = U32 - U64
    out["fp32_split"]["speedup_vs_fp64"] = out["fp32_split"]["value"] / out["fp64"]["value"]
    out["fp32_split"]["relL2_vs_fp64"] = float(np.linalg.norm(d) / np.linalg.norm(U64))
    out["fp32_split"]["relLinf_vs_fp64"] = float(np.max(np.abs(d)) / np.max(np.abs(U64)))
    out["workload"] = (f"resistive GLM-MHD, perturbed 3D Orszag-Tang, {K}^3 warped periodic hexes of [0,1]^3, "
                       f"N={N}, eta={eta}, mu={mu}, Derigs EC + Rusanov + Powell, BR1, {warmup}+{steps} steps "
                       f"CFL/DFL {cfl}/{dfl}")
    out["nElems"], out["nodes"] = int(m["nElems"]), int(nodes)
    return out


# CFLscale/DFLscale of the serving stream: at 0.9 the N=5 density wave (Ranocha + LLF,
# Carpenter RK4-5) loses positivity after ~33 steps, in the CPU oracle as on the GPU.
SERVE_CFL = 0.5


def serving_stream(n_req, nsteps, seed):
    """The mixed request stream (fluxo_amd.workloads): N in {2..5}, K^3 periodic meshes with K
    log-uniform in [4, 24], density wave / inviscid TGV / Navier-Stokes TGV."""
    from fluxo_amd import workloads as W
    cache = W.MeshCache()
    return [W.make_request(seed, s, cache, nsteps) for s in range(n_req)]


# admission cap of the serving legs: the admitted requests hold at most this many thousand
# volume nodes, strict FIFO at the cap (fx_server_config.admitKNodes); one N=5 24^3 request is
# 2,986 thousand nodes
SERVE_ADMIT_KNODES = 4096


def run_server(reqs, precision, arrivals=None, admit_knodes=SERVE_ADMIT_KNODES):
    """Submit every request (at the given arrival offsets in seconds, else all at once) to one
    server; returns makespan, per-request latencies (s) and the DOF-updates done."""
    from fluxo_amd.native import Server, FX_REQ_DONE
    srv = Server(precision=precision, mu0=1.0 / 1600, Pr=0.72, cfl=SERVE_CFL, dfl=SERVE_CFL,
                 hbm_budget=96 << 30, max_requests=len(reqs) + 1, admit_knodes=admit_knodes)
    ids, outs = [], []
    t0 = time.perf_counter()
    for i, r in enumerate(reqs):
        if arrivals is not None:
            wait = arrivals[i] - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(wait)
        rid, out = srv.submit(r["mesh"], r["U0"], r["nsteps"], parabolic=r["parabolic"])
        ids.append(rid)
        outs.append(out)
    srv.wait(-1, timeout=1200)
    st = [srv.status(i) for i in ids]
    stats = srv.stats()
    err = srv.last_error()
    srv.close()
    bad = [(i, s) for i, s in enumerate(st) if s["state"] != FX_REQ_DONE]
    if bad:
        for i, s in bad[:10]:
            r = reqs[i]
            print(f"serving: request {i} (N={r['N']} K={r['K']} {r['ic']} parabolic={r['parabolic']}) "
                  f"failed: {s}", file=sys.stderr)
        raise RuntimeError(f"serving: {len(bad)} requests failed; worker error: {err!r}; stats {stats}")
    start = min(s["t_submit"] for s in st)
    makespan = max(s["t_done"] for s in st) - start
    lat = np.array([s["t_done"] - s["t_submit"] for s in st])
    dofs = sum(r["nodes"] * 5 * r["nsteps"] for r in reqs)
    return dict(makespan=makespan, lat=lat, dofs=dofs, stats=stats, results=outs)


def run_plain(reqs, precision):
    """The same requests one at a time, each through its own context (create, upload, run,
    download, destroy): the plain per-request GPU path.  Returns each request's service
    time (s)."""
    from fluxo_amd.native import Engine
    svc = []
    for r in reqs:
        t0 = time.perf_counter()
        e = Engine(r["mesh"], volume_flux=32, riemann=1, parabolic=r["parabolic"], mu0=1.0 / 1600,
                   Pr=0.72, precision=precision)
        e.upload(r["U0"])
        e.run(r["nsteps"], SERVE_CFL, SERVE_CFL)
        e.download()
        e.close()
        svc.append(time.perf_counter() - t0)
    return np.array(svc)


def tune_serving_tiles():
    """Run the volume-kernel tile tuner for every pool kind of the stream (N 2..5, Euler and
    Navier-Stokes, FP64 and FP32-split) on a 12^3 Taylor-Green mesh; the choices are
    process-wide, so the server pools created afterwards use them."""
    from fluxo_amd.mesh import box_mesh
    from fluxo_amd.native import Engine, FX_PREC_FP64, FX_PREC_FP32_SPLIT
    from fluxo_amd.workloads import taylor_green
    chosen = {}
    for N in (2, 3, 4, 5):
        m = box_mesh(12, N, x0=(-math.pi,) * 3, length=(2 * math.pi,) * 3)
        U0 = taylor_green(m["Elem_xGP"])
        for par in (False, True):
            for prec in (FX_PREC_FP64, FX_PREC_FP32_SPLIT):
                e = Engine(m, volume_flux=32, riemann=1, parabolic=par, mu0=1.0 / 1600, Pr=0.72,
                           precision=prec)
                e.upload(U0)
                best, _ = e.tune_volume(reps=10)
                e.close()
                chosen[f"N{N}{'_ns' if par else '_euler'}{'_fp32' if prec else '_fp64'}"] = best
    return chosen


def fifo_latency(arrivals, service):
    """Latencies of a single FIFO server (one request on the GPU at a time) with the given
    arrival times and measured service times."""
    lat, free = np.empty(len(service)), 0.0
    for i, (a, sv) in enumerate(zip(arrivals, service)):
        start = max(a, free)
        free = start + sv
        lat[i] = free - a
    return lat


def serving_leg(n_req, nsteps, seed, load):
    """Offline batch throughput, then a Poisson stream at `load` x that request rate with
    p50/p99 latency, for FP64 and FP32-split; the plain per-request path (every request run
    alone, back to back) for its throughput and, at the same arrivals, FIFO latency."""
    from fluxo_amd.native import FX_PREC_FP64, FX_PREC_FP32_SPLIT
    reqs = serving_stream(n_req, nsteps, seed)
    tiles = tune_serving_tiles()
    dofs = sum(r["nodes"] * 5 * r["nsteps"] for r in reqs)
    out = {"requests": n_req, "steps_per_request": nsteps, "seed": seed, "DOF_updates": dofs,
           "mix": "N uniform {2,3,4,5}; K^3 periodic Cartesian, K log-uniform [4,24]; density wave "
                  "(Euler) / TGV inviscid / TGV Navier-Stokes BR1; Ranocha + LLF; CFL/DFL 0.5",
           "latency": "submission to result in host memory", "volume_tiles": tiles}
    rng = np.random.default_rng(seed + 1)
    arrivals = None
    results = {}
    for name, prec, cap in (("fp64", FX_PREC_FP64, SERVE_ADMIT_KNODES), ("fp32_split", FX_PREC_FP32_SPLIT,
                                                                          SERVE_ADMIT_KNODES),
                            ("fp64_no_admission_cap", FX_PREC_FP64, 0)):
        off = run_server(reqs, prec, admit_knodes=cap)
        res = off.pop("results")
        if name not in results:
            results[name] = res
        if arrivals is None:
            rate = load * n_req / off["makespan"]
            arrivals = np.cumsum(rng.exponential(1.0 / rate, n_req))
            arrivals -= arrivals[0]
        live = run_server(reqs, prec, arrivals, admit_knodes=cap)
        out[name] = {
            "admit_knodes": cap,
            "offline_DOFps": off["dofs"] / off["makespan"],
            "offline_makespan_s": off["makespan"],
            "poisson_rate_per_s": rate,
            "p50_ms": float(np.percentile(live["lat"], 50) * 1e3),
            "p99_ms": float(np.percentile(live["lat"], 99) * 1e3),
            "stream_DOFps": live["dofs"] / live["makespan"],
            "graph_instantiations": live["stats"]["graph_captures"],
            "graph_updates": live["stats"]["graph_updates"],
            "peak_request_bytes": live["stats"]["bytes_peak"],
        }
    # the split-precision results against the FP64 ones, request by request (final states)
    l2 = np.array([np.linalg.norm(a - b) / np.linalg.norm(b) for a, b in zip(results["fp32_split"], results["fp64"])])
    linf = np.array([np.max(np.abs(a - b)) / np.max(np.abs(b)) for a, b in zip(results["fp32_split"], results["fp64"])])
    out["fp32_split"].update(relL2_vs_fp64_max=float(l2.max()), relL2_vs_fp64_median=float(np.median(l2)),
                             relLinf_vs_fp64_max=float(linf.max()))
    svc = run_plain(reqs, FX_PREC_FP64)
    lat = fifo_latency(arrivals, svc)
    out["plain_fp64"] = {
        "DOFps": dofs / svc.sum(),
        "fifo_p50_ms": float(np.percentile(lat, 50) * 1e3),
        "fifo_p99_ms": float(np.percentile(lat, 99) * 1e3),
        "note": "every request alone through its own context, back to back; FIFO latency at the "
                "batched run's Poisson arrivals from the measured service times"}
    out["batched_vs_plain_throughput"] = out["fp64"]["offline_DOFps"] / out["plain_fp64"]["DOFps"]
    # the same comparison per request size: each bucket's requests alone through the batched
    # server (offline) against the same requests through the plain path
    buckets = {}
    for lo, hi in ((0, 1e5), (1e5, 1e6), (1e6, 1e12)):
        sel = [i for i, r in enumerate(reqs) if lo <= r["nodes"] < hi]
        if not sel:
            continue
        sub = [reqs[i] for i in sel]
        b = run_server(sub, FX_PREC_FP64)
        pdofs = sum(r["nodes"] * 5 * r["nsteps"] for r in sub)
        plain = pdofs / svc[sel].sum()
        buckets[f"nodes_{int(lo)}_{int(hi) if hi < 1e12 else 'max'}"] = {
            "requests": len(sel), "batched_DOFps": b["dofs"] / b["makespan"], "plain_DOFps": plain,
            "batched_vs_plain": b["dofs"] / b["makespan"] / plain}
    out["by_size_fp64"] = buckets
    return out


def sedov_leg(K=40, N=3, steps=300, cfl=0.5):
    """Euler with NFVSE shock capturing: Gaussian pressure pulse (fluxo_amd.workloads.sedov_blast:
    sigma 0.06, E0 1, rho 1, p0 1e-5) on a periodic Cartesian K^3 mesh of [-1.5,1.5]^3, N=3,
    Ranocha + LLF, Persson-Peraire on rho*p, alpha_max 0.5, alpha_min 0.01, one propagation
    sweep with factor 0.5, first-order sub-cell FV; `steps` LSERK steps at CFL `cfl` timed
    from the initial state (a separate short run warms the kernels and the graph up)."""
    from fluxo_amd.mesh import box_mesh
    from fluxo_amd.native import Engine, FX_KERNEL_NFVSE_ALPHA, FX_KERNEL_NFVSE_FV, FX_KERNEL_VOLUME_RK
    from fluxo_amd.workloads import sedov_blast
    m = box_mesh(K, N, x0=(-1.5,) * 3, length=(3.0,) * 3)
    U0 = sedov_blast(m["Elem_xGP"])
    from fluxo_amd.native import FX_PREC_FP64, FX_PREC_FP32_SPLIT
    kw = dict(volume_flux=32, riemann=1, nfvse=True, alpha_max=0.5, alpha_min=0.01, sweeps=1,
              prop_factor=0.5)
    nodes = m["nElems"] * (N + 1) ** 3
    res = {}
    for name, prec in (("fp64", FX_PREC_FP64), ("fp32_split", FX_PREC_FP32_SPLIT)):
        warm = Engine(m, precision=prec, **kw)
        warm.upload(U0)
        warm.run(3, cfl)
        warm.close()
        e = Engine(m, precision=prec, **kw)
        e.upload(U0)
        e.sync()
        t0 = time.perf_counter()
        t_end, dt_last, err = e.run(steps, cfl)
        e.sync()
        el = time.perf_counter() - t0
        a = e.blending()
        U = e.download()
        ms = {"nfvse_alpha": e.profile_kernel(FX_KERNEL_NFVSE_ALPHA, 10),
              "nfvse_fv": e.profile_kernel(FX_KERNEL_NFVSE_FV, 10),
              "volume": e.profile_kernel(FX_KERNEL_VOLUME_RK, 10)}
        e.close()
        p = 0.4 * (U[..., 4] - 0.5 * (U[..., 1:4] ** 2).sum(-1) / U[..., 0])
        res[name] = {"value": nodes * 5 * steps / el, "unit": "DOF-updates/s", "ms_per_step": el / steps * 1e3,
                     "errType": err, "t_end": t_end, "rho_min": float(U[..., 0].min()), "p_min": float(p.min()),
                     "elements_blended_at_end": int((a > 0).sum()),
                     "kernel_ms": {k: round(v, 4) for k, v in ms.items()}, "U": U}
    U64, U32 = res["fp64"].pop("U"), res["fp32_split"].pop("U")
    d = U32 - U64
    out = dict(res["fp64"])
    out["workload"] = (f"Euler + NFVSE Sedov-type pulse, {K}^3 periodic Cartesian hexes of [-1.5,1.5]^3, N={N}, "
                       f"{steps} steps CFL {cfl}")
    out["elements"] = int(m["nElems"])
    out["fp32_split"] = dict(res["fp32_split"], speedup_vs_fp64=res["fp32_split"]["value"] / res["fp64"]["value"],
                             relL2_vs_fp64=float(np.linalg.norm(d) / np.linalg.norm(U64)),
                             relLinf_vs_fp64=float(np.max(np.abs(d)) / np.max(np.abs(U64))))
    return out


def _sample_mesh_and_state(K):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    m = O.box_mesh(K, 3, x0=(-math.pi,) * 3, length=(2 * math.pi,) * 3, Ngeo=3, warp=True)
    return m, tgv_state(m["Elem_xGP"])


def host_cpus():
    """CPUs this process may use: the affinity mask, the cgroup CPU quota, the model name."""
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, per = f.read().split()[:2]
            if q != "max":
                quota = float(q) / float(per)
    except (OSError, ValueError):
        pass
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    usable = aff if quota is None else max(1, min(aff, int(quota)))
    # at most CPU_SHARE processes: a one-GPU job's share of a multi-GPU node's host CPUs
    return dict(nproc=aff, cgroup_cpu_quota=quota, model=model, usable=min(usable, CPU_SHARE))


def cpu_baseline(steps, procs):
    """CPU baseline on a bounded sample of the same workload (8^3 warped TGV elements, N=3,
    NS + BR1, LSERK steps with CALCTIMESTEP).  Preferred: the reference's own Fortran behind
    the harness in oracle/_ref (kind "reference"; the -O3 x86-64-v3 build), run as `procs`
    independent single-core processes at once (the reference scales as independent MPI ranks);
    the per-process setup (metrics, initial RHS) is timed separately with nsteps=0 and
    subtracted.  Without that binary: the plain-C oracle on one core (kind "port")."""
    K = 8
    m, U = _sample_mesh_and_state(K)
    dofs = K ** 3 * 64 * 5 * steps
    cpus = host_cpus()
    sys.path.insert(0, os.path.join(ROOT, "oracle", "ref"))
    try:
        import pyref as R
        R.binary(True, 32, fast=True)
        have_ref = True
    except Exception:
        have_ref = False
    if have_ref:
        from concurrent.futures import ThreadPoolExecutor

        def one(n):
            t0 = time.perf_counter()
            R.run(m, U, volflux=32, riemann=1, parabolic=True, mu0=1.0 / 1600, Pr=0.72, nsteps=n,
                  timeout=900, fast=True)
            return time.perf_counter() - t0

        with ThreadPoolExecutor(procs) as ex:
            setup = max(ex.map(one, [0] * procs))
            t0 = time.perf_counter()
            list(ex.map(one, [steps] * procs))
            el = time.perf_counter() - t0 - setup
        return {"value": dofs * procs / el, "unit": "DOF-updates/s", "cores": procs,
                "kind": "reference", "nproc": cpus["nproc"], "cpu_share": CPU_SHARE, "cgroup_cpu_quota": cpus["cgroup_cpu_quota"],
                "cpu_model": cpus["model"],
                "sample": f"reference Fortran (oracle/_ref harness, flang -O3 -march=x86-64-v3), {procs} "
                          f"concurrent single-core runs of NS TGV N=3 BR1 on 8^3 warped elements, {steps} RK "
                          f"steps each ({el:.1f} s after {setup:.1f} s setup)"}
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    P = O.Problem(m, volflux=32, riemann=1, parabolic=True, mu0=1.0 / 1600, Pr=0.72)
    Ut = np.zeros_like(U)
    t0 = time.perf_counter()
    t = 0.0
    for _ in range(steps):
        dt, _ = P.calc_dt(U, 0.9, 0.9)
        P.lserk_step(U, Ut, t, dt)
        t += dt
    el = time.perf_counter() - t0
    return {"value": dofs / el, "unit": "DOF-updates/s", "cores": 1, "kind": "port",
            "nproc": cpus["nproc"], "cgroup_cpu_quota": cpus["cgroup_cpu_quota"], "cpu_model": cpus["model"],
            "sample": f"CPU oracle, NS TGV N=3 BR1 on 8^3 warped elements, {steps} RK steps "
                      f"({el:.1f} s)"}


def copy_calibration():
    """Device-to-device copies of known size (256 MiB and 1 GiB, 3 times each) for the HBM
    traffic counters: each copy reads and writes its size once."""
    import torch
    for mib in (256, 1024):
        n = mib * (1 << 20) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        c = torch.empty_like(a)
        for _ in range(3):
            c.copy_(a)
        torch.cuda.synchronize()
        print(f"copied {mib} MiB x 3", flush=True)
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--elems", type=int, default=32, help="elements per direction")
    ap.add_argument("--cpu-steps", type=int, default=200)
    ap.add_argument("--cpu-procs", type=int, default=0,
                    help="concurrent CPU baseline processes (0: every CPU this process may use)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=512)
    ap.add_argument("--serve-steps", type=int, default=50)
    ap.add_argument("--serve-load", type=float, default=0.7,
                    help="Poisson arrival rate as a fraction of the offline request throughput")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-sedov", action="store_true")
    ap.add_argument("--no-mhd", action="store_true")
    ap.add_argument("--mhd-steps", type=int, default=100)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only device-to-device copies of 256 MiB and 1 GiB (known bytes, for calibrating "
                         "the FETCH_SIZE / WRITE_SIZE counters), then exit")
    args = ap.parse_args()
    if args.copy_calibration:
        return copy_calibration()
    if args.gpus != 1:
        print("bench.py: this engine runs on one GPU; partitioning the elements over several GPUs "
              "is not built. Run with --gpus 1.", file=sys.stderr)
        return 2

    from fluxo_amd.mesh import box_mesh
    from fluxo_amd.native import (Engine, FX_KERNEL_VOLUME_RK, FX_KERNEL_SURFACE, FX_KERNEL_LIFTING,
                                  FX_KERNEL_DT, FX_PREC_FP64, FX_PREC_FP32_SPLIT)

    K, N = args.elems, 3
    mesh = box_mesh(K, N, x0=(-math.pi,) * 3, length=(2 * math.pi,) * 3, Ngeo=3, warp=True)
    U0 = tgv_state(mesh["Elem_xGP"])
    nodes = mesh["nElems"] * (N + 1) ** 3

    def leg(precision):
        """Warmup + exactly `steps` timed RK steps (CALCTIMESTEP before each) of one engine;
        returns the timing, the final state and per-kernel durations."""
        eng = Engine(mesh, volume_flux=32, riemann=1, parabolic=True, mu0=1.0 / 1600, Pr=0.72,
                     precision=precision)
        eng.upload(U0)
        tile, tile_ms = eng.tune_volume()
        if args.warmup > 0:
            eng.run(args.warmup, 0.9, 0.9)
        eng.sync()
        t0 = time.perf_counter()
        t_end, dt_last, err = eng.run(args.steps, 0.9, 0.9)
        eng.sync()
        el = time.perf_counter() - t0
        if err != 0:
            raise RuntimeError(f"CALCTIMESTEP reported errType={err}")
        U_end = eng.download()
        # the reference's TGV analysis of the final state on the device (outside the timed steps)
        ta = time.perf_counter()
        an = eng.analyze(t_end, diss_factor=1.0 / 1600)
        eng.sync()
        an_ms = (time.perf_counter() - ta) * 1e3
        analysis = {k: an[k] for k in ("Ekin", "Ekin_comp", "Enstrophy_comp", "DR_S", "DR_Sd", "DR_p",
                                       "max_Vorticity", "Entropy_comp")}
        analysis.update(vol=an["vol"], wall_ms=round(an_ms, 3),
                        note="fx_analyze at t_end: RHS + lifting + testcase integrals on the device, "
                             "dissipation rates with mu0")
        reps = 20
        ms = {"volume": eng.profile_kernel(FX_KERNEL_VOLUME_RK, reps),
              "surface": eng.profile_kernel(FX_KERNEL_SURFACE, reps),
              "lifting": eng.profile_kernel(FX_KERNEL_LIFTING, reps),
              "calc_dt": eng.profile_kernel(FX_KERNEL_DT, reps)}
        eng.close()
        return dict(el=el, t_end=t_end, dt_last=dt_last, U=U_end, ms=ms, analysis=analysis,
                    value=nodes * 5 * args.steps / el,
                    tiles={"chosen_elems_per_workgroup": tile,
                           "ms_by_tile": {str(k): round(v, 4) for k, v in tile_ms.items()}})

    r64 = leg(FX_PREC_FP64)
    r32 = leg(FX_PREC_FP32_SPLIT)
    value, el, ms = r64["value"], r64["el"], r64["ms"]
    d = r32["U"] - r64["U"]
    fp32 = {"value": r32["value"], "unit": "DOF-updates/s", "ms_per_step": r32["el"] / args.steps * 1e3,
            "speedup_vs_fp64": r32["value"] / value,
            "relL2_vs_fp64": float(np.linalg.norm(d) / np.linalg.norm(r64["U"])),
            "relLinf_vs_fp64": float(np.max(np.abs(d)) / np.max(np.abs(r64["U"]))),
            "t_end": r32["t_end"],
            "kernel_ms": {k: round(v, 4) for k, v in r32["ms"].items()},
            "volume_tile": r32["tiles"],
            "volume_kernel_GBps": round(volume_kernel_bytes(N, mesh["nElems"], True, 4)
                                        / (r32["ms"]["volume"] * 1e-3) / 1e9, 1),
            "note": "state after warmup+steps RK steps vs the FP64 run from the same initial state"}

    # per-kernel rooflines (durations from HIP events on the engine stream); the dominant
    # kernel is the one with the longest launch; HBM traffic from the committed PMC summary
    # of the same kernel and workload (profiles/pmc_traffic.json), when there is one
    traffic = {}
    try:
        with open(os.path.join(ROOT, "profiles", "pmc_traffic.json")) as f:
            traffic = json.load(f)
    except (OSError, ValueError):
        pass
    tile = r64["tiles"]["chosen_elems_per_workgroup"]
    kern = {"volume": ("k_volume<3, 32, true, 1, double, %d>" % tile, volume_kernel_bytes(N, mesh["nElems"], True)),
            "lifting": ("k_lifting<3, 1, fx::NsLift, 1, false>", lifting_kernel_bytes(N, mesh["nElems"])),
            "surface": ("k_surface<3, 1, true>", surface_kernel_bytes(N, mesh["nSides"], True))}
    roofs = {}
    for k, (kname, nbytes) in kern.items():
        ach = nbytes / (ms[k] * 1e-3) / 1e9
        tr = traffic.get(kname, {}).get("hbm_bytes_per_launch")
        roofs[k] = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": tr, "kernel": kname,
                    "launch_ms": round(ms[k], 4), "bytes_per_launch": nbytes}
    dom = max(roofs, key=lambda k: roofs[k]["launch_ms"])
    roof = dict(roofs[dom])
    if roof["traffic"] is not None:
        roof["traffic_source"] = "profiles/pmc_traffic.json (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE; calibrated)"
    # the whole path against both peaks: algorithmic flops per DOF-update counted from the CPU
    # restatement (profiles/flops.json) and the compulsory bytes per DOF-update of a single
    # fused pass, 8 (4V + 10) + 8 (2 3V) for the BR1 gradients = 480 B for V = 5
    try:
        with open(os.path.join(ROOT, "profiles", "flops.json")) as f:
            fl = json.load(f)["per_dof_update"]["navier_stokes_tgv_n3_br1_ranocha_llf"]
    except (OSError, ValueError, KeyError):
        fl = None
    compulsory = 8 * (4 * 5 + 10) + 8 * 2 * 3 * 5
    roof["path_hbm_frac"] = round(value * compulsory / (HBM_PEAK_GBS * 1e9), 4)
    roof["path_compulsory_bytes_per_dof_update"] = compulsory
    if fl is not None:
        roof["path_flops_per_dof_update"] = round(fl, 1)
        roof["path_TFLOPs"] = round(value * fl / 1e12, 3)
        roof["path_flops_frac"] = round(value * fl / (FP64_PEAK_TFLOPS * 1e12), 4)
    t_end, dt_last = r64["t_end"], r64["dt_last"]
    sedov = None if args.no_sedov else sedov_leg()
    mhd = None if args.no_mhd else mhd_leg(steps=args.mhd_steps, warmup=min(5, args.mhd_steps))
    serving = None if args.no_serving else serving_leg(args.serve_requests, args.serve_steps, 2026,
                                                          args.serve_load)
    cpu = None if args.no_cpu_baseline else cpu_baseline(args.cpu_steps, args.cpu_procs or host_cpus()["usable"])
    out = {
        "metric": METRIC, "value": value, "unit": "DOF-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": el / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"Navier-Stokes TGV Re=1600 Ma=0.1, [-pi,pi]^3 periodic, {K}^3 warped "
                               f"hexes (Ngeo=3), N=3 LGL, BR1 entropy-variable lifting, Ranocha EC "
                               f"volume flux + LLF, Carpenter RK4-5 CFL/DFL 0.9",
                   "nElems": mesh["nElems"], "nodes": nodes, "N": N, "precision": "FP64",
                   "t_end": t_end, "dt_last": dt_last},
        "kernel_ms": {k: round(v, 4) for k, v in ms.items()},
        "tgv_analysis": r64["analysis"],
        "volume_tile": r64["tiles"],
        "pid_s_per_dof_timestep": 5.0 / value,
        "fp32_split": fp32,
        "serving": serving,
        "shock_capturing": sedov,
        "mhd_resistive": mhd,
        "roofline": roof,
        "roofline_per_kernel": roofs,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
