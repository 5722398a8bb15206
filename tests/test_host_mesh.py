"""The Fortran host's periodic box mesh and CalcMetrics agree with the oracle (CPU only)."""
import numpy as np
import pytest

import pyoracle as O
from fluxo_amd.mesh import box_mesh


@pytest.mark.parametrize("K,N,Ngeo,warp,x0,L", [
    (3, 3, 1, False, (0, 0, 0), (2, 2, 2)),
    (3, 3, 3, True, (0, 0, 0), (2, 2, 2)),
    ((2, 3, 4), 4, 2, True, (-np.pi,) * 3, (2 * np.pi,) * 3),
    (2, 5, 3, True, (-1.5,) * 3, (3.0,) * 3),
    (2, 2, 4, True, (0, 0, 0), (1, 1, 1)),
])
def test_host_mesh_matches_oracle(K, N, Ngeo, warp, x0, L):
    h = box_mesh(K, N, x0=x0, length=L, Ngeo=Ngeo, warp=warp)
    o = O.box_mesh(K, N, x0=x0, length=L, Ngeo=Ngeo, warp=warp)
    assert np.array_equal(h["ElemToSide"], o["ElemToSide"])
    assert np.array_equal(h["SideToElem"], o["SideToElem"])
    assert np.array_equal(h["NodeCoords"], o["NodeCoords"])
    for k in ("Mf", "Mg", "Mh", "sJ", "NormVec", "TangVec1", "TangVec2", "SurfElem", "Elem_xGP"):
        scale = np.abs(o[k]).max()
        assert np.abs(h[k] - o[k]).max() <= 1e-14 * scale, k


@pytest.mark.parametrize("name", [n for n in __import__("golden_cases").names()])
def test_host_metrics_match_reference(name):
    from golden_cases import load, mesh_args
    g = load(name)
    if "Mf" not in g:
        pytest.skip("case stores no metrics")
    h = box_mesh(g["K"], g["N"], **mesh_args(g))
    for k in ("Mf", "sJ", "SurfElem"):
        assert np.abs(h[k] - g[k]).max() <= 1e-14 * np.abs(g[k]).max(), k
