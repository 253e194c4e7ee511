"""Precision modes on near-degenerate tets (edge ratios 1e3..1e6, condition numbers up to ~1e9)
against the oracle evaluated with a 64-bit significand.

For such tets FP64 itself is the limit: the volume of a sliver piece carries an error of about
eps64 * cond relative to the tet, so every FP64 evaluation - the reference algorithm in FP64
included - splits a tet's mass among cells with absolute errors up to ~1e-6 of the tet mass.  The
bars here are therefore absolute, in units of the tet mass m: FP64 forms within 2e-6 m per cell,
L1 and total-mass drift within 1e-9, and no worse than the FP64 reference algorithm.  The printed
line is the precision study's table on these inputs.
"""

import json

import numpy as np
import pytest

import adeco_amd as A
from adeco_amd.lattice import make_degenerate_tets
from tests import _oracle as O

pytestmark = pytest.mark.gpu

M = 1.0 / 6.0


def metrics(g, ref):
    g = np.asarray(g, np.float64)
    tot = float(np.abs(ref).sum())
    return {"max_abs_over_m": float(np.max(np.abs(g - ref)) / M),
            "max_rel_floor_1e-8m": float(np.max(np.abs(g - ref) / np.maximum(np.abs(ref), 1e-8 * M))),
            "l1_rel": float(np.abs(g - ref).sum() / tot), "mass_drift": float(abs(g.sum() - ref.sum()) / tot)}


def test_degenerate_tets_precision_modes():
    tets = make_degenerate_tets(50000, seed=4, box=25.0, cell=100.0 / 128)
    box = np.array([0, 25, 0, 25, 0, 25], np.float32)
    dims = 32
    with O.cell_scaling(O.EXACT):
        ld = O.deposit_soup_ld(tets, box, dims, M)
        ref64 = O.deposit_soup_f64(tets, box, dims, M)
    rows = {"fp64 reference algorithm (oracle)": metrics(ref64, ld)}
    for name, mode, prec in (("fp64 batched", A.MODE_BATCHED, A.PREC_FP64), ("fp64 plain", A.MODE_PLAIN, A.PREC_FP64),
                             ("mixed batched", A.MODE_BATCHED, A.PREC_MIXED), ("fp32 batched", A.MODE_BATCHED, A.PREC_FP32)):
        print(f"[precision] {name}", flush=True)
        g = A.deposit_soup(tets, box, dims, M, mode=mode, precision=prec)
        rows[name] = metrics(g, ld)
        if mode == A.MODE_BATCHED:
            tl = A.Octree.last_timings()
            rows[name]["fp64_recompute_fraction"] = tl["fallbacks"] / max(tl["partials"], 1)
    print(json.dumps({"precision_study_degenerate_50k": rows}), flush=True)
    ref_err = rows["fp64 reference algorithm (oracle)"]
    for name in ("fp64 batched", "fp64 plain", "mixed batched"):
        r = rows[name]
        assert r["max_abs_over_m"] <= 2e-6, name
        assert r["l1_rel"] <= 1e-9 and r["mass_drift"] <= 1e-9, name
        assert r["max_rel_floor_1e-8m"] <= ref_err["max_rel_floor_1e-8m"], name
    assert np.isfinite(rows["fp32 batched"]["l1_rel"])
