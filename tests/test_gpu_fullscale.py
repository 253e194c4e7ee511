"""Parity at the full size of the precision-study and serving workloads.

* Precision study inputs (the 128^3 lattice at rms 3 spacings, plus 1,000,000 seeded
  near-degenerate tets with edge ratios 1e3..1e6), each deposited onto one 128^3 grid over the
  100 Mpc/h box, every cell against the oracle evaluated with a 64-bit significand:
  - the lattice's 12,582,912 tets, batched FP64 form: per cell |gpu - oracle| <= 1e-6 *
    max(|oracle|, 1e-8 m), total mass to 1e-9;
  - the degenerate tets, batched and plain FP64 forms: FP64 itself is the limit on tets this
    ill-conditioned (a sliver piece carries ~eps64 * cond of the tet's volume, and at a million
    tets the condition numbers reach ~1e12), so the yardstick is the reference algorithm evaluated
    in FP64: the largest per-cell error (in tet masses, 2.0e-4 m for it here) and the per-cell
    relative error (floor 1e-8 m) no worse than its, L1 error <= 1e-8 and total-mass drift <= 1e-9
    of the total.  The mixed and FP32 clips are measured and printed (the study's table), not held
    to the bar.
* Serving on the resident 256^3 lattice (rms 3 spacings, seed 5): requests at 256^3 and 512^3 cells
  (the latter's FP64 accumulator is 1 GiB), each checked on a block of 24^3 cells inside the box
  against the oracle's deposit of the request's tets onto that block's cells (the same cell
  faces), as FP32 density: <= 1e-6 * max(|oracle|, floor) + one FP32 rounding per cell.
"""

import json

import numpy as np
import pytest

import adeco_amd as A
from adeco_amd.lattice import make_degenerate_tets, make_lattice
from tests import _oracle as O

pytestmark = pytest.mark.gpu

M = 1.0 / 6.0
BOX = np.array([0, 100, 0, 100, 0, 100], np.float32)


def metrics(g, ref):
    g = np.asarray(g, np.float64)
    tot = float(np.abs(ref).sum())
    return {"max_abs_over_m": float(np.max(np.abs(g - ref)) / M),
            "max_rel_floor_1e-8m": float(np.max(np.abs(g - ref) / np.maximum(np.abs(ref), 1e-8 * M))),
            "l1_rel": float(np.abs(g - ref).sum() / tot), "mass_drift": float(abs(g.sum() - ref.sum()) / tot)}


def test_precision_study_inputs_at_full_size():
    lat = make_lattice(128, 3.0, seed=2)
    print("[fullscale] 128^3 lattice onto one 128^3 grid", flush=True)
    with O.cell_scaling(O.EXACT):
        want = O.deposit_lattice(lat, BOX.astype(np.float64), 128, decomp=A.DECOMP_KUHN, tet_mass=M, precision=3,
                                 nthreads=16)
        with A.ResidentLattice.from_lattice(lat) as rl:
            got = rl.deposit_grid(BOX, 128, decomposition=A.DECOMP_KUHN, tet_mass=M, mode=A.MODE_BATCHED,
                                  precision=A.PREC_FP64)
        assert O.max_rel_err(got, want, tet_mass=M) <= 1e-6
        assert abs(got.sum() - want.sum()) <= 1e-9 * want.sum()
        tets = make_degenerate_tets(1_000_000, seed=4)
        print("[fullscale] 1M near-degenerate tets onto one 128^3 grid", flush=True)
        ld = O.deposit_soup_ld(tets, BOX, 128, M)
        ref64 = O.deposit_soup_f64(tets, BOX, 128, M)
        rows = {"fp64 reference algorithm (oracle)": metrics(ref64, ld)}
        for name, mode, prec in (("fp64 batched", A.MODE_BATCHED, A.PREC_FP64), ("fp64 plain", A.MODE_PLAIN, A.PREC_FP64),
                                 ("mixed batched", A.MODE_BATCHED, A.PREC_MIXED),
                                 ("fp32 batched", A.MODE_BATCHED, A.PREC_FP32)):
            print(f"[fullscale] degenerate tets, {name}", flush=True)
            g = A.deposit_soup(tets, BOX, 128, M, mode=mode, precision=prec)
            rows[name] = metrics(g, ld)
    print(json.dumps({"precision_study_degenerate_1M": rows}), flush=True)
    ref_err = rows["fp64 reference algorithm (oracle)"]
    for name in ("fp64 batched", "fp64 plain"):
        r = rows[name]
        assert r["max_abs_over_m"] <= ref_err["max_abs_over_m"], name
        assert r["max_rel_floor_1e-8m"] <= ref_err["max_rel_floor_1e-8m"], name
        assert r["l1_rel"] <= 1e-8 and r["mass_drift"] <= 1e-9, name
    assert np.isfinite(rows["mixed batched"]["l1_rel"]) and np.isfinite(rows["fp32 batched"]["l1_rel"])


SERVE_LATTICE = (256, 3.0, 5)
# (box, cells per axis, first cell of the checked 24^3 block per axis)
LARGE_REQUESTS = [([30.0, 70.0, 25.0, 65.0, 40.0, 80.0], 256, (116, 116, 116)),
                  ([41.0, 59.0, 45.0, 63.0, 38.0, 56.0], 512, (244, 244, 244))]
BLOCK = 24


def test_large_requests_on_the_resident_256_lattice():
    n, rms, seed = SERVE_LATTICE
    lat = make_lattice(n, rms, seed=seed)
    tets = O.lattice_tets(lat, A.DECOMP_KUHN)
    mn, mx = tets.min(1), tets.max(1)
    with O.cell_scaling(O.EXACT), A.ResidentLattice.from_lattice(lat) as rl:
        with A.Server(rl, A.DECOMP_KUHN, M, max_inflight=4) as srv:
            ids = [srv.submit(roi, dims) for roi, dims, _ in LARGE_REQUESTS]
            res = []
            for (roi, dims, _), rid in zip(LARGE_REQUESTS, ids):
                print(f"[fullscale] request {roi} at {dims}^3 cells", flush=True)
                res.append(srv.wait(rid, timeout=240))
        for (roi, dims, lo), (got, info) in zip(LARGE_REQUESTS, res):
            roi = np.asarray(roi, np.float64)
            rf = roi.astype(np.float32)
            sel = np.all((mn < rf[1::2]) & (rf[0::2] < mx), 1)
            assert info["n_tets"] == int(sel.sum())
            delta = (roi[1::2] - roi[0::2]) / dims
            sub = np.empty(6)
            sub[0::2] = roi[0::2] + np.asarray(lo) * delta
            sub[1::2] = roi[0::2] + (np.asarray(lo) + BLOCK) * delta
            # the request's tets that can reach the block
            near = sel & np.all((mn < sub[1::2]) & (sub[0::2] < mx), 1)
            want = O.deposit_lattice(lat, sub, BLOCK, decomp=A.DECOMP_KUHN, tet_mass=M,
                                     tet_ids=np.flatnonzero(near).astype(np.int32), precision=3, nthreads=16)
            cell_vol = float(np.prod(delta))
            want = (want / cell_vol).reshape(BLOCK, BLOCK, BLOCK)
            blk = got[lo[2]:lo[2] + BLOCK, lo[1]:lo[1] + BLOCK, lo[0]:lo[0] + BLOCK].astype(np.float64)
            floor = 1e-8 * M / cell_vol
            err = float(np.max(np.abs(blk - want) / np.maximum(np.abs(want), floor)))
            print(f"  {dims}^3: {info['n_tets']} tets, {info['n_candidates']} candidates, block max rel err {err:.2e}, "
                  f"latency {info['latency']:.3f} s", flush=True)
            assert err <= 1e-6 + 1.2e-7
            assert np.count_nonzero(want) > BLOCK ** 3 // 2
