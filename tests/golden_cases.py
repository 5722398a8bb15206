"""Loader for the golden vectors in tests/golden (made by the reference's own code with
tests/golden/make_golden.py)."""
import glob
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def names():
    return sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "*.npz")))


def load(name):
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        d = {k: z[k] for k in z.files}
    for k in ("K", "N", "Ngeo", "warp", "volflux", "riemann", "parabolic", "nsteps", "errType0"):
        d[k] = int(d[k])
    for k in ("mu0", "Pr", "kappa", "cfl", "dfl", "dt0", "t_end", "Entropy", "dSdU_Ut"):
        d[k] = float(d[k])
    d["nfvse"] = int(d.get("nfvse", 0))
    d["lifting"] = int(d.get("lifting", 1))
    d["etaBR2"] = float(d.get("etaBR2", 2.0))
    if d["nfvse"]:
        d["sweeps"] = int(d["sweeps"])
        for k in ("alpha_max", "alpha_min", "prop_factor", "fixed_alpha"):
            d[k] = float(d[k])
    return d


def nfvse_args(g):
    """Problem/Engine keyword arguments of a shock-capturing case (empty otherwise)."""
    if not g["nfvse"]:
        return {}
    return dict(nfvse=True, alpha_max=g["alpha_max"], alpha_min=g["alpha_min"], sweeps=g["sweeps"],
                prop_factor=g["prop_factor"], fixed_alpha=None if g["fixed_alpha"] < 0 else g["fixed_alpha"])


def lifting_args(g):
    """Problem/Engine keyword arguments of the lifting (BR1 unless the case says BR2)."""
    return dict(lifting=g["lifting"], etaBR2=g["etaBR2"]) if g["parabolic"] else {}


def mesh_args(g):
    return dict(x0=tuple(g["x0"]), length=tuple(g["L"]), Ngeo=g["Ngeo"], warp=bool(g["warp"]))


def initial_state(g, exactfunc, mesh):
    """The stored U, or for the large cases the state it was made from: the Sedov-type pulse
    (fluxo_amd.workloads.sedov_blast with the stored sigma, E0, p0) or the noise-free
    ExactFunc 2 density wave."""
    if "U" in g:
        return g["U"]
    if str(g.get("ic", "")) == "sedov":
        from fluxo_amd.workloads import sedov_blast
        return sedov_blast(mesh["Elem_xGP"], sigma=float(g["sigma"]), E0=float(g["E0"]), p0=float(g["p0"]))
    return exactfunc(2, mesh["Elem_xGP"])


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))
