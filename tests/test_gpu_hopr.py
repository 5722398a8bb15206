"""The engine on the reference's own HOPR meshes (read by fluxo_amd.io, metrics from the
Fortran host) against the reference's code on the same meshes (tests/golden/hopr), and the
reference's Navier-Stokes free-stream test (tests/runs/navst/freestream/conforming: its
mesh, its restart state file, tend 20, L_inf error <= 1e-11 from CalcErrorNorms) through the
C ABI.  FP64 bar: relative L2 <= 1e-12 per RHS and for the lifted gradients, <= 1e-10
after RK steps."""
import glob
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
IO_DIR = os.path.join(HERE, "golden", "io")
HOPR_DIR = os.path.join(HERE, "golden", "hopr")
HOPR_NAMES = sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(HOPR_DIR, "*.npz")))
STATE = os.path.join(IO_DIR, "NAVIERSTOKES_FREESTREAM_State_0000000.000000000.h5")


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


@pytest.mark.parametrize("name", HOPR_NAMES)
def test_engine_matches_reference_on_hopr_meshes(name):
    from fluxo_amd.mesh import hopr_mesh
    from fluxo_amd.native import Engine
    with np.load(os.path.join(HOPR_DIR, name + ".npz"), allow_pickle=False) as z:
        g = {k: z[k] for k in z.files}
    m = hopr_mesh(os.path.join(IO_DIR, str(g["mesh_file"])), int(g["N"]))
    par = bool(int(g["parabolic"]))
    kw = dict(volume_flux=int(g["volflux"]), riemann=int(g["riemann"]), parabolic=par, mu0=float(g["mu0"]),
              Pr=float(g["Pr"]))
    e = Engine(m, **kw)
    e.upload(g["U"])
    if par:
        gerr = max(rel_l2(a, g[k]) for a, k in zip(e.lifting(), ("gradPx", "gradPy", "gradPz")))
        print(f"{name}: gradients relL2 vs reference {gerr:.3e}")
        assert gerr <= 1e-12
    err = rel_l2(e.rhs(), g["Ut"])
    print(f"{name}: RHS relL2 vs reference {err:.3e}")
    assert err <= 1e-12
    e2 = Engine(m, **kw)
    e2.upload(g["U"])
    t, _, errt = e2.run(int(g["nsteps"]), float(g["cfl"]), float(g["dfl"]))
    assert errt == 0
    traj = rel_l2(e2.download(), g["U_end"])
    print(f"{name}: {int(g['nsteps'])} steps relL2 vs reference {traj:.3e}")
    assert traj <= 1e-10 and t == pytest.approx(float(g["t_end"]), rel=1e-12)


@pytest.mark.parametrize("volume_flux", [1, 32])
def test_reference_navier_stokes_free_stream_test(volume_flux):
    """The reference's navst freestream/conforming test: restart from its state file on its
    heavily warped Ngeo=3 mesh, N=3, mu0 = 1e-6, CFL 0.9, DFL 0.5, run to t >= 20, then the
    L_inf error of CalcErrorNorms against the constant RefState must stay below 1e-11 (its
    err_tol).  VolumeFlux 1 (standard DG with averaged metrics, Riemann 1) is the reference
    run's setting; the Ranocha split form preserves the free stream the same way."""
    from fluxo_amd import io
    from fluxo_amd.mesh import hopr_mesh
    from fluxo_amd.native import Engine
    U0, t0 = io.read_state(STATE)
    m = hopr_mesh(os.path.join(IO_DIR, "ConformBoxHeavilyWarped_Ngeo3__03_03_03_mesh.h5"), 3)
    e = Engine(m, volume_flux=volume_flux, riemann=1, parabolic=True, mu0=1e-6, Pr=0.72)
    e.upload(U0)
    e.upload_coords(m["Elem_xGP"])
    t, steps = t0, 0
    while t < 20.0 and steps < 20000:
        t, _, err = e.run(500, 0.9, 0.5, t0=t)
        assert err == 0
        steps += 500
    a = e.analyze(t, exact_func=1, ref_state=tuple(U0[0, 0, 0, 0]))
    print(f"free stream VolumeFlux {volume_flux}: t = {t:.3f} after {steps} steps, L_inf {a['Linf']}")
    assert t >= 20.0
    assert np.all(a["Linf"] <= 1e-11)


def test_state_file_roundtrip_through_the_engine(tmp_path):
    """WriteState of the device state and Restart from that file give bitwise the same state
    and the same right-hand side."""
    from fluxo_amd import io
    from fluxo_amd.mesh import hopr_mesh
    from fluxo_amd.native import Engine
    import pyoracle as O
    m = hopr_mesh(os.path.join(IO_DIR, "ConformBoxUnstructPeriodicHeavilyWarped_Ngeo2_mesh.h5"), 4)
    U = O.exactfunc(2, m["Elem_xGP"])
    e = Engine(m, volume_flux=32, riemann=1)
    e.upload(U)
    t, _, err = e.run(4, 0.9, 0.9)
    assert err == 0
    f = tmp_path / "RUN_State.h5"
    io.write_state(f, e.download(), t, mesh_file="ConformBoxUnstructPeriodicHeavilyWarped_Ngeo2_mesh.h5")
    V, tv = io.read_state(f)
    assert tv == t and np.array_equal(V, e.download())
    e2 = Engine(m, volume_flux=32, riemann=1)
    e2.upload(V)
    assert np.array_equal(e2.rhs(), e.rhs())
