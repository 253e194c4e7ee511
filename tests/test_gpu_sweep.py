"""Seeded random sweeps against the extended-precision oracle, plain and batched FP64.

* Tet soups: 80 cases over grid sizes 1..64 cells per axis, cubic and non-cubic boxes, both cell
  scalings, and tets from 0.02 to 30 cells across (thin ones down to a tenth of their length in two
  directions), placed in and around the box so that many are cut by its faces.
* Octree leaf deposits: 24 lattices (8^3..32^3 particles, rms 0.3-5 spacings, with and without
  infall centres), octrees split above 300..20000 tets at depths 1-6, 4..32 cells per leaf, all
  three tet decompositions and both cell scalings; up to 24 leaves each (the 6 heaviest and a seeded
  spread), the oracle given every tet whose AABB meets the leaf box (its own selection, not the
  library's binning).
* ROI requests: 30 random boxes on three random lattices, submitted together to servers with 1-8
  slots (test_random_requests_match_oracle).
Per cell |gpu - oracle| <= 1e-6 * max(|oracle|, 1e-8 * tet_mass); grid totals (soups) and the sum
over all leaves against the lattice's mass inside the root (exact scaling) agree to 1e-9.
"""

import numpy as np
import pytest

import adeco_amd as A
from adeco_amd.lattice import make_lattice
from tests import _oracle as O

pytestmark = pytest.mark.gpu

TET_MASS = 1.0 / 6.0
DIMS = [1, 2, 3, 5, 8, 13, 16, 24, 32, 48, 64]
CASES = 80


def rotation(rng):
    q = rng.standard_normal(4)
    w, x, y, z = q / np.linalg.norm(q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def make_case(seed):
    """(tets float32 [n,4,3], bbox float32 [6], dims, cell scaling) for one seed."""
    rng = np.random.default_rng(1000 + seed)
    dims = int(rng.choice(DIMS))
    lo = rng.uniform(-10, 10, 3)
    if rng.random() < 0.5:
        ext = np.exp(rng.uniform(np.log(0.5), np.log(20), 3))
    else:
        ext = np.full(3, np.exp(rng.uniform(np.log(0.5), np.log(20))))
    bbox = np.empty(6, np.float32)
    bbox[0::2] = lo
    bbox[1::2] = lo + ext
    cell = (bbox[1::2] - bbox[0::2]) / dims
    size_cells = float(np.exp(rng.uniform(np.log(0.02), np.log(30))))
    n = int(min(4000, max(50, 1.5e6 / max(1.0, size_cells) ** 3)))
    centres = lo + ext * rng.uniform(-0.2, 1.2, (n, 3))
    base = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float) - 0.25
    tets = np.empty((n, 4, 3))
    for t in range(n):
        shape = np.diag(np.exp(rng.uniform(np.log(0.1), 0.0, 3)))
        tets[t] = centres[t] + (base @ shape @ rotation(rng).T) * size_cells * cell.mean() * 1.5
    scaling = O.EXACT if rng.random() < 0.6 else O.REFERENCE
    return tets.astype(np.float32), bbox, dims, scaling


@pytest.mark.parametrize("block", range(4))
def test_random_soups_match_oracle(block):
    for seed in range(block * CASES // 4, (block + 1) * CASES // 4):
        tets, bbox, dims, scaling = make_case(seed)
        print(f"[sweep] seed {seed}: {len(tets)} tets, {dims}^3 cells, scaling {scaling}", flush=True)
        with O.cell_scaling(scaling):
            want = O.deposit_soup_ld(tets, bbox, dims, TET_MASS)
            total = max(abs(want.sum()), 1e-30)
            for name, mode in (("plain", A.MODE_PLAIN), ("batched", A.MODE_BATCHED)):
                got = A.deposit_soup(tets, bbox, dims, TET_MASS, mode=mode, precision=A.PREC_FP64)
                assert O.max_rel_err(got, want, tet_mass=TET_MASS) <= 1e-6, (seed, name)
                assert abs(got.sum() - want.sum()) <= 1e-9 * total, (seed, name)


def test_sweep_covers_its_regimes():
    """The seeds span what the sweep claims: every grid size class, both scalings, non-cubic boxes,
    tets much smaller and much larger than a cell."""
    cases = [make_case(s) for s in range(CASES)]
    assert {c[3] for c in cases} == {O.EXACT, O.REFERENCE}
    assert len({c[2] for c in cases}) >= 8
    ext = [c[1][1::2] - c[1][0::2] for c in cases]
    assert any(np.ptp(e) > 0.1 * e.max() for e in ext)
    # a tet's extent in cells (its first tet; a case's tets share the scale)
    sizes = [np.ptp(c[0][0], axis=0).max() / ((c[1][1::2] - c[1][0::2]) / c[2]).mean() for c in cases]
    assert min(sizes) < 0.1 and max(sizes) > 10


LEAF_CASES = 24


def make_leaf_case(seed):
    """(lattice n, rms, infall centres, split threshold, depth, cells per leaf, decomposition, scaling)."""
    rng = np.random.default_rng(5000 + seed)
    n = int(rng.integers(8, 33))
    rms = float(rng.uniform(0.3, 5.0))
    infall = int(rng.choice([0, 0, 4, 12]))
    max_tets = int(rng.choice([300, 1000, 3000, 20000]))
    depth = int(rng.integers(2, 7))
    dims = int(rng.choice([4, 8, 12, 16, 24, 32]))
    # a tet spans at most ~6 cells of its deepest leaves per axis (dims 2^depth / n): the plain form
    # walks a tet's cells on one thread and the oracle clips every one of them on the CPU
    while dims > 4 and dims * 2 ** depth / n > 6:
        dims //= 2
    while depth > 1 and dims * 2 ** depth / n > 6:
        depth -= 1
    decomp = int(rng.choice([A.DECOMP_REF_EFFECTIVE, A.DECOMP_REF_INTENDED, A.DECOMP_KUHN]))
    scaling = O.EXACT if rng.random() < 0.7 else O.REFERENCE
    return n, rms, infall, max_tets, depth, dims, decomp, scaling, rng


@pytest.mark.parametrize("block", range(3))
def test_random_octree_leaves_match_oracle(block):
    for seed in range(block * LEAF_CASES // 3, (block + 1) * LEAF_CASES // 3):
        n, rms, infall, max_tets, depth, dims, decomp, scaling, rng = make_leaf_case(seed)
        print(f"[sweep] leaves seed {seed}: {n}^3 rms {rms:.1f} infall {infall}, split {max_tets} depth {depth}, "
              f"{dims}^3 cells, decomposition {decomp}, scaling {scaling}", flush=True)
        lat = make_lattice(n, rms, seed=seed + 100, infall_centres=infall)
        with O.cell_scaling(scaling), A.ResidentLattice.from_lattice(lat) as rl, A.Octree(rl, max_tets, depth) as tree:
            counts = tree.bin(decomp)
            nl = tree.num_leaves
            boxes = tree.leaf_boxes_f64(exact=(scaling == O.EXACT))
            pick = np.unique(np.concatenate([np.argsort(-counts)[:6], rng.choice(nl, min(nl, 18), replace=False)]))
            tets = O.lattice_tets(lat, decomp)
            mn, mx = tets.min(1).astype(np.float64), tets.max(1).astype(np.float64)
            want = []
            for l in pick:
                b = boxes[l]
                pad = 1e-6 * (b[1::2] - b[0::2])
                ids = np.flatnonzero(np.all((mn < b[1::2] + pad) & (b[0::2] - pad < mx), 1)).astype(np.int32)
                want.append(O.deposit_lattice(lat, b, dims, decomp=decomp, tet_mass=TET_MASS, tet_ids=ids, precision=3,
                                              nthreads=16))
            want = np.array(want)
            inside = None
            if scaling == O.EXACT:
                inside = O.deposit_lattice(lat, tree.root_box_f64(), 1, decomp=decomp, tet_mass=TET_MASS).sum()
            for name, mode in (("plain", A.MODE_PLAIN), ("batched", A.MODE_BATCHED)):
                tree.deposit(dims, decomposition=decomp, tet_mass=TET_MASS, mode=mode, precision=A.PREC_FP64)
                g = tree.leaf_grids()
                assert O.max_rel_err(g[pick], want, tet_mass=TET_MASS) <= 1e-6, (seed, name)
                if inside is not None:
                    assert abs(g.sum() - inside) <= 1e-9 * inside, (seed, name)


SERVE_BLOCKS = 3
SERVE_REQUESTS = 10


@pytest.mark.parametrize("block", range(SERVE_BLOCKS))
def test_random_requests_match_oracle(block):
    """Seeded random ROI requests, submitted together to a server with a random number of slots, on a
    random lattice: boxes 1-60 Mpc/h per side (non-cubic, some partly or wholly outside the
    lattice), 4..64 cells per axis (at most ~8 cells across a tet), FP32 density vs the oracle's
    deposit of the tets whose float AABB strictly overlaps the box (its own selection), per cell to
    1e-6 plus one FP32 rounding, and the selected tet count exactly."""
    rng = np.random.default_rng(9000 + block)
    n = int(rng.integers(10, 25))
    lat = make_lattice(n, float(rng.uniform(0.5, 4.0)), seed=200 + block, infall_centres=int(rng.choice([0, 6])))
    tets = O.lattice_tets(lat, A.DECOMP_KUHN)
    mn, mx = tets.min(1), tets.max(1)
    tet_size = float(np.median(np.max(mx - mn, axis=1)))
    reqs = []
    for _ in range(SERVE_REQUESTS):
        side = np.exp(rng.uniform(np.log(1.0), np.log(60.0), 3))
        if rng.random() < 0.5:
            side[:] = side[0]
        lo = rng.uniform(-10.0, 105.0, 3) - side / 2
        roi = np.empty(6)
        roi[0::2] = lo
        roi[1::2] = lo + side
        cap = max(4, int(8 * float(np.min(side)) / tet_size))
        dims = int(min(int(rng.choice([4, 8, 13, 16, 24, 32, 48, 64])), cap))
        reqs.append((roi, dims))
    slots = int(rng.integers(1, 9))
    with O.cell_scaling(O.EXACT), A.ResidentLattice.from_lattice(lat) as rl, \
            A.Server(rl, A.DECOMP_KUHN, TET_MASS, max_inflight=slots) as srv:
        ids = [srv.submit(roi, dims) for roi, dims in reqs]
        for (roi, dims), rid in zip(reqs, ids):
            print(f"[sweep] serve block {block}: {n}^3, {slots} slots, box {np.round(roi, 2)}, {dims}^3", flush=True)
            got, info = srv.wait(rid, timeout=120)
            rf = roi.astype(np.float32)
            sel = np.flatnonzero(np.all((mn < rf[1::2]) & (rf[0::2] < mx), 1)).astype(np.int32)
            assert info["n_tets"] == sel.size
            cell_vol = float(np.prod((roi[1::2] - roi[0::2]) / dims))
            if sel.size:
                want = O.deposit_lattice(lat, roi, dims, decomp=A.DECOMP_KUHN, tet_mass=TET_MASS, tet_ids=sel,
                                         precision=3, nthreads=16) / cell_vol
            else:
                want = np.zeros(dims ** 3)
            g = np.asarray(got, np.float64).ravel()
            floor = 1e-8 * TET_MASS / cell_vol
            assert float(np.max(np.abs(g - want) / np.maximum(np.abs(want), floor))) <= 1e-6 + 1.2e-7
