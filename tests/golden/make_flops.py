"""Writes the algorithmic FP64 operation counts of the adaptive deposit (the roofline's numerator
in bench.py).

The count is the reference algorithm's: every floating-point add, subtract, multiply and divide
that the oracle's restatement of deposit_tet_2_grid (AdDepositTets2Grid.h:509-665, with the clip
of AdTetIntersection.h:230-476) executes in its counting instantiation, over every (leaf, tet)
pair of the reference octree (max depth 6, split above 50,000 estimated tets) whose tet's float
AABB strictly overlaps the leaf's float box (Kuhn decomposition).  Beside the FLOPs it records the
nonzero (tet, cell) contributions, which set the algorithmic bytes (8 B of FP64 accumulation per
contribution).

  c2_flops.json  128^3 lattice, rms 3 spacings, seed 2, 16^3-cell leaves: every leaf counted.
  c3_flops.json  256^3 lattice, rms 4 spacings + 64 infall centres, seed 3, 16^3- and 32^3-cell
                 leaves: the pairs of every leaf are counted exactly; the operations on a seeded
                 sample of leaves stratified by octree level (SAMPLE_FRAC of each level, at least
                 MIN_PER_LEVEL), scaled per level by that level's pairs (ratio estimator; its
                 standard error is recorded).
  c5_flops.json  the serving study's 60 region-of-interest requests (bench.py serve_requests,
                 seed 5) on its 256^3 lattice (rms 3 spacings, seed 5): per request the tets
                 selected exactly, the operations on a seeded sample of them sized to about
                 2e7 candidate cells, scaled by the selected count.

Run from the repository root:  python tests/golden/make_flops.py [c2|c3|c5]
"""

import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from adeco_amd.lattice import META, make_lattice, meta_cell_dims  # noqa: E402
from tests import _oracle as O  # noqa: E402
from tests._octree_io import oracle_octree  # noqa: E402

MAX_TETS, MAX_LEVEL, TET_MASS, KUHN = 50000, 6, 1.0 / 6.0, 2
SAMPLE_FRAC, MIN_PER_LEVEL, SAMPLE_SEED = 0.08, 24, 11
HERE = os.path.dirname(os.path.abspath(__file__))


def leaf_box(tree, node):
    ll, ext = tree["ll"][node], tree["ext"][node]
    return np.array([ll[0], ll[0] + ext, ll[1], ll[1] + ext, ll[2], ll[2] + ext], np.float32)


def leaf_tet_ids(lat, tree, leaf, mn, mx):
    """Tets whose float AABB strictly overlaps the leaf's float box.  Candidates are the tets of
    the leaf's meta-cells (those whose AABB strictly meets the leaf, AdOctreeGeneration.h:225-258),
    which hold every tet of the leaf."""
    n = lat.n
    md = meta_cell_dims(lat.dims)
    ids = []
    for m in tree["mc_ids"][tree["mc_offsets"][leaf]:tree["mc_offsets"][leaf + 1]]:
        c = (m % md[0], (m // md[0]) % md[1], m // (md[0] * md[1]))
        r = [np.arange(META * c[a], min(META * c[a] + META, n)) for a in range(3)]
        k, j, i = np.meshgrid(r[2], r[1], r[0], indexing="ij")
        cube = (i + n * (j + n * k)).ravel().astype(np.int64)
        ids.append((6 * cube[:, None] + np.arange(6)[None, :]).ravel())
    ids = np.unique(np.concatenate(ids)) if ids else np.zeros(0, np.int64)
    box = leaf_box(tree, tree["dfs"][leaf])
    keep = np.all((mn[ids] < box[1::2]) & (box[0::2] < mx[ids]), 1)
    return ids[keep].astype(np.int32)


def count(lat, box, dims, ids):
    _, st = O.deposit_lattice(lat, box, dims, decomp=KUHN, tet_mass=TET_MASS, tet_ids=ids, precision=2,
                              nthreads=os.cpu_count() or 1, stats=True)
    return int(st[12]), int(st[8])


def c2():
    t0 = time.time()
    lat = make_lattice(128, 3.0, seed=2)
    tree = oracle_octree(lat, MAX_TETS, MAX_LEVEL)
    tets = O.lattice_tets(lat, KUHN)
    mn, mx = tets.min(1), tets.max(1)
    flops, pairs, nonzero = 0, 0, 0
    for node in tree["dfs"]:
        box = leaf_box(tree, node)
        ids = np.nonzero(np.all((mn < box[1::2]) & (box[0::2] < mx), 1))[0].astype(np.int32)
        f, nz = count(lat, box, 16, ids)
        flops, pairs, nonzero = flops + f, pairs + int(ids.size), nonzero + nz
    out = {"workload": "128^3 lattice, k^-2 field rms 3 spacings seed 2, Kuhn tets, octree depth<=6 split>50000, "
                       "16^3-cell leaves", "leaves": int(len(tree["dfs"])), "pairs": pairs, "flops": flops,
           "flops_per_pair": flops / pairs, "nonzero_contributions": nonzero,
           "counted_by": "oracle Counted<double> (add/sub/mul/div)"}
    with open(os.path.join(HERE, "c2_flops.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out), f"({time.time() - t0:.0f} s)")


def c3():
    t0 = time.time()
    lat = make_lattice(256, 4.0, seed=3, infall_centres=64)
    tree = oracle_octree(lat, MAX_TETS, MAX_LEVEL)
    tets = O.lattice_tets(lat, KUHN)
    mn, mx = tets.min(1), tets.max(1)
    del tets
    nl = len(tree["dfs"])
    level = tree["level"][tree["dfs"]]
    print(f"[c3] lattice + octree ({nl} leaves) in {time.time() - t0:.0f} s", flush=True)
    ids_of = {}
    pairs = np.zeros(nl, np.int64)
    for leaf in range(nl):
        ids = leaf_tet_ids(lat, tree, leaf, mn, mx)
        pairs[leaf] = ids.size
        ids_of[leaf] = ids
    print(f"[c3] {pairs.sum()} pairs binned in {time.time() - t0:.0f} s", flush=True)
    rng = np.random.default_rng(SAMPLE_SEED)
    sample = []
    for lv in np.unique(level):
        members = np.flatnonzero(level == lv)
        k = min(members.size, max(MIN_PER_LEVEL, int(round(SAMPLE_FRAC * members.size))))
        sample.extend(int(x) for x in np.sort(rng.choice(members, k, replace=False)))
    out = {"workload": "256^3 lattice, k^-2 field rms 4 spacings + 64 infall centres seed 3, Kuhn tets, octree "
                       "depth<=6 split>50000", "leaves": nl, "pairs": int(pairs.sum()),
           "pairs_per_level": {str(int(lv)): int(pairs[level == lv].sum()) for lv in np.unique(level)},
           "leaves_per_level": {str(int(lv)): int(np.sum(level == lv)) for lv in np.unique(level)},
           "sample": {"seed": SAMPLE_SEED, "fraction_per_level": SAMPLE_FRAC, "min_per_level": MIN_PER_LEVEL,
                      "leaves": len(sample), "leaf_dfs_indices": sample},
           "counted_by": "oracle Counted<double> (add/sub/mul/div), ratio estimator per octree level",
           "dims": {}}
    for dims in (16, 32):
        f_s = np.zeros(nl)
        nz_s = np.zeros(nl)
        for leaf in sample:
            f_s[leaf], nz_s[leaf] = count(lat, leaf_box(tree, tree["dfs"][leaf]), dims, ids_of[leaf])
        est_f, est_nz, var_f = 0.0, 0.0, 0.0
        per_level = {}
        for lv in np.unique(level):
            s = [l for l in sample if level[l] == lv]
            ratio = f_s[s].sum() / pairs[s].sum()
            ratio_nz = nz_s[s].sum() / pairs[s].sum()
            p_lv = pairs[level == lv].sum()
            n_lv = int(np.sum(level == lv))
            est_f += ratio * p_lv
            est_nz += ratio_nz * p_lv
            # ratio-estimator variance with the finite-population correction
            resid = f_s[s] - ratio * pairs[s]
            mean_p = pairs[s].mean()
            var_r = (1 - len(s) / n_lv) * resid.var(ddof=1) / (len(s) * mean_p ** 2)
            var_f += var_r * p_lv ** 2
            per_level[str(int(lv))] = {"flops_per_pair": ratio, "nonzero_per_pair": ratio_nz}
        out["dims"][str(dims)] = {"flops": float(est_f), "flops_rel_std_err": float(np.sqrt(var_f) / est_f),
                                  "flops_per_pair": float(est_f / pairs.sum()), "nonzero_contributions": float(est_nz),
                                  "per_level": per_level,
                                  "sample_flops": int(f_s.sum()), "sample_pairs": int(pairs[sample].sum())}
        print(f"[c3] {dims}^3: {json.dumps(out['dims'][str(dims)])} ({time.time() - t0:.0f} s)", flush=True)
    with open(os.path.join(HERE, "c3_flops.json"), "w") as f:
        json.dump(out, f, indent=1)


def c5():
    sys.path.insert(0, ROOT)
    from bench import SERVE_N, serve_requests
    t0 = time.time()
    lat = make_lattice(SERVE_N, 3.0, seed=5)
    tets = O.lattice_tets(lat, KUHN)
    mn, mx = tets.min(1), tets.max(1)
    del tets
    ext = np.stack([mx - mn], 0)[0].astype(np.float64)
    _, rois, dims = serve_requests(60, 5)
    rng = np.random.default_rng(SAMPLE_SEED)
    out = {"workload": "serving study: 60 ROI requests (side U[5,50] Mpc/h, 64^3..512^3 cells, seed 5) on the 256^3 "
                       "lattice at rms 3 spacings (seed 5), Kuhn tets", "requests": []}
    total = 0.0
    for i in range(len(dims)):
        roi = np.asarray(rois[i], np.float64)
        rf = roi.astype(np.float32)
        sel = np.flatnonzero(np.all((mn < rf[1::2]) & (rf[0::2] < mx), 1))
        d = int(dims[i])
        cell = (roi[1::2] - roi[0::2]) / d
        per_tet = float(np.mean(np.prod(ext[sel] / cell + 1.0, axis=1))) if sel.size else 0.0
        k = int(min(sel.size, max(200, 2e7 / max(per_tet, 1.0))))
        sample = np.sort(rng.choice(sel, k, replace=False)).astype(np.int32) if k else sel.astype(np.int32)
        f = 0
        if sample.size:
            _, st = O.deposit_lattice(lat, roi, d, decomp=KUHN, tet_mass=TET_MASS, tet_ids=sample, precision=2,
                                      nthreads=os.cpu_count() or 1, stats=True)
            f = int(st[12])
        est = f * sel.size / max(sample.size, 1)
        total += est
        out["requests"].append({"dims": d, "tets": int(sel.size), "sampled": int(sample.size), "flops": est})
        print(f"[c5] request {i}: {d}^3, {sel.size} tets, sample {sample.size}, {est:.3e} flops ({time.time() - t0:.0f} s)",
              flush=True)
    out["flops"] = total
    out["counted_by"] = "oracle Counted<double> (add/sub/mul/div), per-request tet sample scaled by the selected tets"
    with open(os.path.join(HERE, "c5_flops.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    which = sys.argv[1] if len(sys.argv) > 1 else "c2"
    {"c2": c2, "c3": c3, "c5": c5}[which]()
