"""Loader for the sub-interface golden vectors in tests/golden/iface (FillFlux, SurfInt,
VolInt_NFVSE of the reference's own code, made by tests/golden/make_golden.py)."""
import glob
import os

import numpy as np

IFACE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "iface")


def names():
    return sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(IFACE, "*.npz")))


def load(name):
    with np.load(os.path.join(IFACE, name + ".npz"), allow_pickle=False) as z:
        g = {k: z[k] for k in z.files}
    g["kind"] = str(g["kind"])
    for k in ("K", "N", "Ngeo", "warp", "parabolic"):
        g[k] = int(g[k])
    return g


def mesh(g, box_mesh):
    return box_mesh(g["K"], g["N"], x0=tuple(g["x0"]), length=tuple(g["L"]), Ngeo=g["Ngeo"], warp=bool(g["warp"]))


def engine_kwargs(g):
    """Engine / Problem keyword arguments of a case (Euler / Navier-Stokes / NFVSE)."""
    kw = dict(kappa=float(g["kappa"]))
    if g["kind"] == "nfvse":
        kw.update(nfvse=True, alpha_max=0.5, alpha_min=0.01, sweeps=1, prop_factor=0.5)
    if g["parabolic"]:
        kw.update(mu0=float(g["mu0"]), Pr=float(g["Pr"]), lifting=int(g["lifting"]), etaBR2=float(g["etaBR2"]))
    return kw


def mhd_kwargs(g):
    res = dict(mu=float(g["mu"]), eta=float(g["eta"]), Pr=float(g["Pr"])) if g["parabolic"] else None
    return dict(kappa=float(g["kappa"]), mu_0=float(g["mu_0"]), glm_scale=float(g["glm_scale"]), resistive=res,
                riemann=int(g["riemann"]))


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))
