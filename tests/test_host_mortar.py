"""Non-conforming HOPR meshes on the host (CPU only): the reader's mortar connectivity
(fxio_mesh_mortars) and the Fortran host's mortar surface metrics, on the reference's own
mortar meshes (tests/golden/io, copied data).

Pinning: the reference's metrics of the same meshes (its CalcMetrics and
Mortar_CalcSurfMetrics run by the reference harness, tests/golden/mortar) are reproduced
exactly; independently of both codes, every small side's face points, interpolated from its
big side's face with Lagrange polynomials written out here, coincide with the small
element's own face points up to one periodic translation, which fixes the mortar type, the
order of the small sides and every flip."""
import ctypes as C
import glob
import os

import numpy as np
import pytest

import pyoracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
IO_DIR = os.path.join(HERE, "golden", "io")
MORTAR_DIR = os.path.join(HERE, "golden", "mortar")
MESHES = {"CartBoxMortarPeriodic_Ngeo1_2to1_mesh.h5": (36, 162, 36, {2: 22, 3: 14}),
          "CartBoxMortarPeriodic_Ngeo1_4to1_mesh.h5": (36, 168, 24, {1: 24}),
          "DeformedBoxMortarPeriodic__Ngeo_2_Level_01_mesh.h5": (18, 71, 10, {1: 2, 2: 6, 3: 2})}
NAMES = sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(MORTAR_DIR, "*.npz")))


def side_to_vol(N, p, q, flip, loc):
    ijk = (C.c_int * 3)()
    O.lib().ora_side_to_vol.argtypes = [C.c_int] * 6 + [C.c_int * 3]
    O.lib().ora_side_to_vol(N, 0, p, q, flip, loc, ijk)
    return tuple(ijk)


def lgl_nodes(N):
    """Legendre-Gauss-Lobatto nodes: +-1 and the roots of P_N'."""
    c = np.zeros(N + 1)
    c[N] = 1.0
    inner = np.polynomial.legendre.legroots(np.polynomial.legendre.legder(c))
    return np.concatenate(([-1.0], np.sort(inner), [1.0]))


def lagrange(x, y):
    """L[p, l] = l_l(y_p), Lagrange polynomials of the nodes x."""
    L = np.ones((len(y), len(x)))
    for l in range(len(x)):
        for k in range(len(x)):
            if k != l:
                L[:, l] *= (y - x[k]) / (x[l] - x[k])
    return L


@pytest.mark.parametrize("mesh_file", sorted(MESHES))
def test_reader_mortar_connectivity(mesh_file):
    from fluxo_amd import io
    r = io.read_mesh(os.path.join(IO_DIR, mesh_file))
    E, S, nm, types = MESHES[mesh_file]
    assert (r["nElems"], r["nSides"], r["nMortarSides"]) == (E, S, nm)
    mt, mi, s2e, e2s = r["MortarType"], r["MortarInfo"], r["SideToElem"], r["ElemToSide"]
    assert dict(zip(*np.unique(mt[:nm, 0], return_counts=True))) == types
    assert np.array_equal(mt[:nm, 1], np.arange(1, nm + 1))
    small = []
    for m in range(nm):
        ns = 4 if mt[m, 0] == 1 else 2
        assert tuple(mi[m, 0]) == (m + 1, 0)
        assert np.all(mi[m, 1:ns + 1, 1] == 0) and np.all(mi[m, ns + 1:] == -1)
        small += list(mi[m, 1:ns + 1, 0])
        em, es, lm = s2e[m, 0], s2e[m, 1], s2e[m, 2]
        assert es == -1 and tuple(e2s[em - 1, lm - 1]) == (m + 1, 0)
    assert len(set(small)) == len(small) and min(small) > nm
    for s in range(nm, S):
        em, es, lm, ls, fl = s2e[s]
        assert 1 <= es <= E and 1 <= fl <= 4 and tuple(e2s[es - 1, ls - 1]) == (s + 1, fl)
        if s + 1 in small:
            assert em == -1 and mt[s, 0] == 0
        else:
            assert tuple(e2s[em - 1, lm - 1]) == (s + 1, 0) and mt[s, 0] == -1


@pytest.mark.parametrize("mesh_file", sorted(MESHES))
def test_small_sides_lie_on_their_big_side(mesh_file):
    from fluxo_amd.mesh import hopr_mesh
    N = 4
    m = hopr_mesh(os.path.join(IO_DIR, mesh_file), N)
    x, s2e, mt, mi = m["Elem_xGP"], m["SideToElem"], m["MortarType"], m["MortarInfo"]
    lo, hi = x.reshape(-1, 3).min(0), x.reshape(-1, 3).max(0)
    xi = lgl_nodes(N)
    M = [lagrange(xi, 0.5 * (xi - 1.0)), lagrange(xi, 0.5 * (xi + 1.0))]  # halves [-1,0], [0,1]
    for mm in range(m["nMortarSides"]):
        eb, lb = s2e[mm, 0] - 1, s2e[mm, 2]
        big = np.array([[x[eb][side_to_vol(N, p, q, 0, lb)[::-1]] for p in range(N + 1)] for q in range(N + 1)])
        ns = 4 if mt[mm, 0] == 1 else 2
        for im in range(ns):
            if mt[mm, 0] == 1:
                Mx, My = M[im % 2], M[im // 2]
            elif mt[mm, 0] == 2:  # halves in eta
                Mx, My = np.eye(N + 1), M[im]
            else:  # halves in xi
                Mx, My = M[im], np.eye(N + 1)
            sm = np.einsum("qb,pa,bad->qpd", My, Mx, big)  # big[q=b][p=a]
            s = mi[mm, im + 1, 0] - 1
            es, ls, fl = s2e[s, 1] - 1, s2e[s, 3], s2e[s, 4]
            own = np.array([[x[es][side_to_vol(N, p, q, fl, ls)[::-1]] for p in range(N + 1)] for q in range(N + 1)])
            d = (sm - own).reshape(-1, 3)
            assert np.ptp(d, axis=0).max() <= 1e-12, (mesh_file, mm, im)
            shift = d[0]
            assert np.all(np.isclose(shift, 0.0, atol=1e-12) | np.isclose(np.abs(shift), hi - lo, rtol=1e-12)), shift


@pytest.mark.parametrize("name", NAMES)
def test_host_mortar_metrics_match_reference(name):
    from fluxo_amd.mesh import hopr_mesh
    with np.load(os.path.join(MORTAR_DIR, name + ".npz"), allow_pickle=False) as z:
        g = {k: z[k] for k in z.files}
    if "Mf" not in g:
        pytest.skip("case stores no metrics (the GLM-MHD cases run on meshes the others cover)")
    m = hopr_mesh(os.path.join(IO_DIR, str(g["mesh_file"])), int(g["N"]))
    for k in ("Mf", "Mg", "Mh", "sJ", "NormVec", "TangVec1", "TangVec2", "SurfElem"):
        assert np.abs(m[k] - g[k]).max() <= 1e-14 * np.abs(g[k]).max(), k
    # the small sides' areas (LGL quadrature of SurfElem) add up to the big side's
    nm, mt, mi, N = m["nMortarSides"], m["MortarType"], m["MortarInfo"], int(g["N"])
    xi = lgl_nodes(N)
    w = np.array([2.0 / (N * (N + 1) * np.polynomial.legendre.legval(x, np.eye(N + 1)[N]) ** 2) for x in xi])
    ww = np.outer(w, w)
    for mm in range(nm):
        ns = 4 if mt[mm, 0] == 1 else 2
        area = sum(np.sum(ww * m["SurfElem"][mi[mm, im, 0] - 1]) for im in range(1, ns + 1))
        assert area == pytest.approx(np.sum(ww * m["SurfElem"][mm]), rel=1e-3)
