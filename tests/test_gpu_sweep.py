"""Seeded random sweep of the tet-soup deposit, plain and batched FP64, against the extended-precision
oracle: 80 cases over grid sizes 1..64 cells per axis, cubic and non-cubic boxes, both cell scalings,
and tets from 0.02 to 30 cells across (thin ones down to a tenth of their length in two directions),
placed in and around the box so that many are cut by its faces.  Per cell
|gpu - oracle| <= 1e-6 * max(|oracle|, 1e-8 * tet_mass); the grid totals agree to 1e-9.
"""

import numpy as np
import pytest

import adeco_amd as A
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
