"""The reference's Analyze values held by tests/golden/analyze (names as its AnalyzeToFile
writes them) and their comparison with an analysis dict of the oracle or the engine."""
import glob
import os

import numpy as np

import pyoracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "analyze")
NAMES = sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLD, "*.npz")))

# reference AnalyzeToFile names of the testcase quantities -> oracle fields
TC_NAMES = {"Ekin incomp": "Ekin", "Ekin comp": "Ekin_comp", "Enstrophy comp": "Enstrophy_comp",
            "DR_u": "DR_u", "DR_S": "DR_S", "DR_Sd": "DR_Sd", "DR_p": "DR_p",
            "Maximum Vorticity": "max_Vorticity", "Mean Temperature": "mean_temperature",
            "uprime": "uprime", "Entropy comp": "Entropy_comp", "Emag comp": "Emag_comp",
            "Maximum Current": "max_Current", "Magnetic Enstrophy": "Mag_Enstrophy"}
VEC_PREFIX = (("L2_colloc_", "L2colloc"), ("L2_", "L2"), ("Linf_", "Linf"), ("Bulk_t_", "bulk_t"),
              ("Bulk_", "bulk"), ("maxabs_", "maxabsUt"))


def load(name):
    with np.load(os.path.join(GOLD, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def reference_values(g):
    """{(field, index or None): value} from the golden names/values."""
    out, counters = {}, {}
    for name, val in zip(g["names"], g["values"]):
        name = str(name)
        if name in TC_NAMES:
            out[(TC_NAMES[name], None)] = float(val)
            continue
        for pre, field in VEC_PREFIX:
            if name.startswith(pre):
                i = counters.get(field, 0)
                counters[field] = i + 1
                out[(field, i)] = float(val)
                break
    return out


def mesh_of(g):
    return O.box_mesh(int(g["K"]), int(g["N"]), x0=tuple(g["x0"]), length=tuple(g["L"]), Ngeo=int(g["Ngeo"]),
                      warp=bool(g["warp"]))


def check_against_reference(a, g, tol=1e-12):
    """Every value the reference reported, to `tol` relative, or to 1e-15 absolute for integrals
    that vanish up to round-off; the bulk integrals of Ut to 1e-13 of max|Ut| times the volume.
    Returns the largest relative difference among the values held to `tol`."""
    if "maxabs_Ut" in a:  # the engine's field name
        a = dict(a, maxabsUt=a["maxabs_Ut"])
    ref = reference_values(g)
    assert len(ref) >= 30
    worst = 0.0
    for (field, i), rv in ref.items():
        v = a[field] if i is None else a[field][i]
        if field == "bulk_t":
            scale = a["maxabsUt"][i] * a["vol"]
            assert abs(v - rv) <= 1e-13 * scale, (field, i, v, rv)
            continue
        err = abs(v - rv) / max(abs(rv), 1e-300)
        if abs(v - rv) <= 1e-15:  # round-off of a vanishing integral (e.g. zero-mean momentum)
            continue
        worst = max(worst, err)
        assert err <= tol, (field, i, v, rv)
    return worst
