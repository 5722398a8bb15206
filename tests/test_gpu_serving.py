"""Batched server (fx_server_*) through the C ABI: every request's result in a mixed,
continuously admitted batch is bitwise the same as the same run alone (fx_ctx + fx_run),
and the run alone matches the CPU oracle.
"""
import time

import numpy as np
import pytest

import pyoracle as O
from fluxo_amd import workloads as W

pytestmark = pytest.mark.gpu

NS = dict(mu0=1.0 / 1600, Pr=0.72)


def alone(req, precision=0):
    from fluxo_amd.native import Engine
    e = Engine(req["mesh"], volume_flux=32, riemann=1, parabolic=req["parabolic"], precision=precision,
               **NS)
    e.upload(req["U0"])
    t, dt, err = e.run(req["nsteps"], 0.9, 0.9)
    U = e.download()
    e.close()
    return U, t, dt, err


def mixed_requests():
    """Requests of every N, both initial states, Euler and Navier-Stokes, uneven sizes and
    step counts."""
    cache = W.MeshCache()
    spec = [(2, 3, "wave", False, 7), (3, 2, "tgv", True, 5), (4, 2, "tgv", False, 4),
            (5, 2, "wave", False, 3), (3, 4, "wave", False, 9), (2, 2, "tgv", True, 6),
            (3, 3, "tgv", True, 4), (4, 3, "wave", False, 5), (3, 2, "wave", False, 8),
            (5, 2, "tgv", True, 3)]
    out = []
    for N, K, ic, par, nsteps in spec:
        m = cache.get(K, N, ic)
        U0 = W.density_wave(m["Elem_xGP"]) if ic == "wave" else W.taylor_green(m["Elem_xGP"])
        out.append(dict(N=N, K=K, ic=ic, parabolic=par, mesh=m, U0=U0, nsteps=nsteps))
    return out


@pytest.mark.parametrize("precision", [0, 1])
def test_results_same_alone_and_in_the_mix(precision):
    from fluxo_amd.native import Server, FX_REQ_DONE
    reqs = mixed_requests()
    srv = Server(precision=precision, **NS)
    ids = []
    for i, r in enumerate(reqs):
        ids.append(srv.submit(r["mesh"], r["U0"], r["nsteps"], parabolic=r["parabolic"]))
        if i % 3 == 2:
            time.sleep(0.01)   # later requests arrive while earlier ones run
    srv.wait(-1, timeout=120)
    st = srv.stats()
    for (rid, out), r in zip(ids, reqs):
        s = srv.status(rid)
        assert s["state"] == FX_REQ_DONE and s["errType"] == 0 and s["steps"] == r["nsteps"]
        U, t, dt, err = alone(r, precision)
        assert err == 0
        print(f"N={r['N']} K={r['K']} {r['ic']} par={r['parabolic']}: max|mix-alone|="
              f"{np.max(np.abs(out - U)):.1e} t={s['t_end']:.6f}")
        assert np.array_equal(out, U)
        assert s["t_end"] == t and s["dt_last"] == dt
    srv.close()
    assert st["done"] == len(reqs) and st["failed"] == 0 and st["steps_launched"] > 0


def test_alone_matches_oracle():
    """One request of the mix (Navier-Stokes TGV, N=3) against the CPU oracle."""
    r = mixed_requests()[1]
    P = O.Problem(r["mesh"], volflux=32, riemann=1, parabolic=True, **NS)
    U = r["U0"].copy()
    Ut_tmp = np.zeros_like(U)
    t = 0.0
    for _ in range(r["nsteps"]):
        dt, _ = P.calc_dt(U, 0.9, 0.9)
        P.lserk_step(U, Ut_tmp, t, dt)
        t += dt
    got, tg, _, err = alone(r)
    assert err == 0
    assert np.linalg.norm(got - U) / np.linalg.norm(U) <= 1e-10


def test_hbm_budget_serialises_admission():
    """With a budget that holds one request, the second is admitted only after the first
    completes, and both results are unchanged."""
    from fluxo_amd.native import Server, FX_REQ_DONE, request_bytes
    reqs = [r for r in mixed_requests() if r["N"] == 3 and not r["parabolic"]][:2]
    srv0 = Server(**NS)
    need = max(request_bytes(srv0.cfg, r["N"], r["parabolic"], r["mesh"]["nElems"], r["mesh"]["nSides"])
               for r in reqs)
    srv0.close()
    srv = Server(hbm_budget=need, **NS)
    ids = [srv.submit(r["mesh"], r["U0"], r["nsteps"], parabolic=r["parabolic"]) for r in reqs]
    srv.wait(-1, timeout=120)
    s0, s1 = srv.status(ids[0][0]), srv.status(ids[1][0])
    assert s0["state"] == s1["state"] == FX_REQ_DONE
    assert s1["t_admit"] >= s0["t_done"]
    for (rid, out), r in zip(ids, reqs):
        assert np.array_equal(out, alone(r)[0])
    st = srv.stats()
    assert st["bytes_peak"] <= need
    srv.close()


def test_admission_cap_keeps_fifo_order_and_results():
    """admitKNodes: with a cap below two requests' nodes, a request beyond the cap is admitted
    only after the running ones completed, and a later (smaller) request does not pass it;
    results are bitwise those of the runs alone."""
    from fluxo_amd.native import Server, FX_REQ_DONE
    cache = W.MeshCache()
    spec = [(3, 4, "wave", 6), (3, 4, "wave", 5), (2, 2, "wave", 3)]  # 4096, 4096 and 216 nodes
    reqs = []
    for N, K, ic, nsteps in spec:
        m = cache.get(K, N, ic)
        reqs.append(dict(N=N, K=K, ic=ic, parabolic=False, mesh=m, U0=W.density_wave(m["Elem_xGP"]), nsteps=nsteps))
    srv = Server(admit_knodes=5, **NS)  # 5,000 nodes: the first request alone fits, not two
    ids = [srv.submit(r["mesh"], r["U0"], r["nsteps"], parabolic=r["parabolic"]) for r in reqs]
    srv.wait(-1, timeout=120)
    st = [srv.status(i) for i, _ in ids]
    assert all(x["state"] == FX_REQ_DONE for x in st)
    assert st[1]["t_admit"] >= st[0]["t_done"]   # beyond the cap: after the first completed
    assert st[2]["t_admit"] >= st[1]["t_admit"]  # the small one did not pass it
    for (rid, out), r in zip(ids, reqs):
        assert np.array_equal(out, alone(r)[0])
    srv.close()


def test_failed_request_does_not_disturb_others():
    """A request whose state has a non-positive pressure is flagged by CALCTIMESTEP and
    fails; the request batched with it completes unchanged."""
    from fluxo_amd.native import Server, FX_REQ_DONE, FX_REQ_FAILED
    good, bad = [r for r in mixed_requests() if r["N"] == 3 and not r["parabolic"]][:2]
    bad = dict(bad)
    bad["U0"] = bad["U0"].copy()
    bad["U0"][0, 1, 1, 1, 4] = 0.0   # total energy 0 < kinetic energy -> p < 0
    srv = Server(**NS)
    (ig, og), (ib, _) = (srv.submit(good["mesh"], good["U0"], good["nsteps"]),
                         srv.submit(bad["mesh"], bad["U0"], bad["nsteps"]))
    srv.wait(-1, timeout=120)
    assert srv.status(ib)["state"] == FX_REQ_FAILED and srv.status(ib)["errType"] == 2
    assert srv.status(ig)["state"] == FX_REQ_DONE
    assert np.array_equal(og, alone(good)[0])
    srv.close()


def test_submit_rejects_bad_requests():
    from fluxo_amd.native import Server, FxError
    r = mixed_requests()[0]
    srv = Server(hbm_budget=1 << 18, **NS)
    with pytest.raises(FxError):
        srv.submit(r["mesh"], r["U0"], 0)                  # no steps
    big = mixed_requests()[4]
    with pytest.raises(FxError):
        srv.submit(big["mesh"], big["U0"], 3)              # larger than the budget
    bad = dict(r["mesh"])
    bad["SideToElem"] = r["mesh"]["SideToElem"].copy()
    bad["SideToElem"][0, 0] = r["mesh"]["nElems"] + 5       # element id out of range
    with pytest.raises(FxError):
        srv.submit(bad, r["U0"], 3)
    srv.close()


def test_fp32split_serving_error_vs_fp64():
    """The split-precision server against the FP64 server on the same mixed requests: every
    request's final state within the stated error (relative L2 / Linf printed per request;
    measured at most 2.2e-7 / 1.6e-6 over these requests, the bounds about 4x that)."""
    from fluxo_amd.native import Server, FX_PREC_FP64, FX_PREC_FP32_SPLIT
    reqs = mixed_requests()
    res = {}
    for prec in (FX_PREC_FP64, FX_PREC_FP32_SPLIT):
        srv = Server(precision=prec, **NS)
        outs = [srv.submit(r["mesh"], r["U0"], r["nsteps"], parabolic=r["parabolic"])[1] for r in reqs]
        srv.wait(-1, timeout=120)
        srv.close()
        res[prec] = outs
    worst = 0.0
    for r, a, b in zip(reqs, res[FX_PREC_FP32_SPLIT], res[FX_PREC_FP64]):
        l2 = float(np.linalg.norm(a - b) / np.linalg.norm(b))
        linf = float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
        print(f"N={r['N']} K={r['K']} {r['ic']} par={r['parabolic']} {r['nsteps']} steps: "
              f"FP32-split vs FP64 relL2 {l2:.2e} relLinf {linf:.2e}")
        worst = max(worst, l2)
        assert l2 <= 1e-6 and linf <= 5e-6  # measured at most 2.2e-7 / 1.6e-6
    print(f"largest relative L2 {worst:.2e}")
