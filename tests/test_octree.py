"""Octree structure and refinement counts: bit-exact against the reference.

CPU: the oracle's restatement vs the compiled reference (oracle/_ref) and vs the committed
golden fixtures (tests/golden/ref_octree_<n>.npz).  GPU: the HIP octree (meta-cell AABBs and
per-node estimate sums on the device) vs the oracle, bit for bit: node table (lower-left and
extension bits, level, parent, first child), uint64 counts, DFS leaf list, processing order and
per-leaf meta-cell lists.  Plus the invariants of a refinement: leaves tile the root, every refined
node exceeded the threshold below max level, every leaf did not.
"""

import hashlib
import os

import numpy as np
import pytest

import adeco_amd as A
from adeco_amd.lattice import make_lattice
from tests._octree_io import KEYS, assert_same_octree, oracle_octree, ref_octree

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
CASES = [(32, 3.0, 2, 3000, 4), (48, 4.0, 6, 4000, 5), (128, 3.0, 2, 50000, 6)]


def _lattice_sha(lat):
    return hashlib.sha256(np.concatenate([lat.x, lat.y, lat.z]).tobytes()).hexdigest()


@pytest.mark.parametrize("case", CASES[:2])
def test_oracle_octree_bit_exact_vs_reference(ref, case):
    n, rms, seed, mt, ml = case
    lat = make_lattice(n, rms, seed=seed)
    a = oracle_octree(lat, mt, ml)
    b = ref_octree(lat, mt, ml)
    assert_same_octree(a, b)
    assert np.array_equal(b["count"], b["local_count"])


def test_oracle_octree_bit_exact_with_roi(ref):
    lat = make_lattice(32, 3.0, seed=2)
    roi = np.array([20, 60, 10, 50, 30, 70], np.float32)
    assert_same_octree(oracle_octree(lat, 2000, 5, roi=roi), ref_octree(lat, 2000, 5, roi=roi))
    bad = np.array([50, 10, 0, 1, 0, 1], np.float32)  # invalid ROI -> whole domain
    assert_same_octree(oracle_octree(lat, 2000, 5, roi=bad), oracle_octree(lat, 2000, 5))


@pytest.mark.parametrize("case", CASES)
def test_oracle_octree_matches_golden(case):
    n = case[0]
    path = os.path.join(GOLDEN, f"ref_octree_{n}.npz")
    if not os.path.exists(path):
        pytest.skip("golden octree not generated")
    g = np.load(path, allow_pickle=False)
    nn, rms, seed, mt, ml = g["params"]
    lat = make_lattice(int(nn), float(rms), seed=int(seed))
    assert _lattice_sha(lat) == str(g["lattice_sha"])
    assert_same_octree(oracle_octree(lat, int(mt), int(ml)), {k: g[k] for k in KEYS})


def check_refinement_invariants(d, max_tets, max_level, num_tets):
    thr = min(num_tets, max_tets)
    fc = d["first_child"]
    for i in range(fc.size):
        if fc[i] >= 0:
            assert d["count"][i] > thr and d["level"][i] < max_level
            kids = np.arange(fc[i], fc[i] + 8)
            assert np.all(d["parent"][kids] == i) and np.all(d["level"][kids] == d["level"][i] + 1)
        else:
            assert d["count"][i] <= thr or d["level"][i] >= max_level
    # leaves tile the root: volumes (in units of the root) sum to 1 and leaf levels are consistent
    leaves = d["dfs"]
    assert np.isclose(np.sum(8.0 ** -d["level"][leaves].astype(np.float64)), 1.0)
    assert set(leaves.tolist()) == set(np.flatnonzero(fc < 0).tolist())


def test_oracle_octree_invariants():
    for n, rms, seed, mt, ml in CASES[:2]:
        lat = make_lattice(n, rms, seed=seed)
        check_refinement_invariants(oracle_octree(lat, mt, ml), mt, ml, lat.num_tets)


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES)
def test_gpu_octree_bit_exact_vs_oracle(case):
    n, rms, seed, mt, ml = case
    print(f"[gpu octree] n={n} max_tets={mt} max_level={ml}", flush=True)
    lat = make_lattice(n, rms, seed=seed)
    want = oracle_octree(lat, mt, ml)
    with A.ResidentLattice.from_lattice(lat) as rl, A.Octree(rl, mt, ml) as tree:
        got = dict(tree.nodes())
        got.update(tree.leaves())
        got["bbox"] = tree.bbox()
        got["mc_bbox"] = tree.meta_bboxes()
    assert_same_octree(got, want)


@pytest.mark.gpu
def test_gpu_octree_roi_and_leaf_binning():
    lat = make_lattice(24, 3.0, seed=4)
    roi = np.array([10, 50, 20, 60, 0, 40], np.float32)
    want = oracle_octree(lat, 1500, 4, roi=roi)
    with A.ResidentLattice.from_lattice(lat) as rl, A.Octree(rl, 1500, 4, region_of_interest=roi) as tree:
        got = dict(tree.nodes())
        got.update(tree.leaves())
        got["bbox"] = tree.bbox()
        got["mc_bbox"] = tree.meta_bboxes()
        assert_same_octree(got, want)
        # binning: every leaf holds exactly the tets whose float AABB strictly overlaps it
        counts = tree.bin(A.DECOMP_KUHN)
        boxes = tree.leaf_boxes()
    o = np.zeros(24, np.int32)
    from tests import _oracle as O
    O.load().oracle_decomposition(2, O.P(o))
    corners = o.reshape(6, 4)
    n = lat.n
    npv = n + 1
    P = np.stack([lat.x, lat.y, lat.z], 1)
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    base = (i + npv * (j + npv * k)).ravel()
    tb = []
    for t in range(6):
        vs = []
        for c in corners[t]:
            vs.append(P[base + (c & 1) + npv * (((c >> 1) & 1) + npv * ((c >> 2) & 1))])
        vs = np.stack(vs, 1)  # [cubes, 4, 3]
        tb.append(np.concatenate([vs.min(1), vs.max(1)], 1))
    tb = np.stack(tb, 1).reshape(-1, 6)  # tet order: cube-major, 6 per cube
    # binning is conservative: every tet strictly overlapping the leaf's float box, and none beyond
    # the box widened by two float ulps per face
    for l in range(boxes.shape[0]):
        b = boxes[l]
        hit = np.all((tb[:, 0:3] < b[1::2]) & (b[0::2] < tb[:, 3:6]), axis=1)
        lo = np.nextafter(np.nextafter(b[0::2], -np.inf), -np.inf)
        hi = np.nextafter(np.nextafter(b[1::2], np.inf), np.inf)
        wide = np.all((tb[:, 0:3] < hi) & (lo < tb[:, 3:6]), axis=1)
        assert hit.sum() <= counts[l] <= wide.sum()
