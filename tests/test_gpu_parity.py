"""GPU parity of the deposit kernels against the oracle.

The oracle here is the reference's algorithm evaluated with a 64-bit significand (x87 extended
precision, rounded to double at the end): the FP64 evaluation of the same algorithm loses ~1e-13 of
a tet mass to cancellation in sliver cells, which is above the bar in cells holding ~1e-8 of a tet.

Tolerances (stated per test): per-cell relative error <= 1e-6, computed as
|gpu - oracle| / max(|oracle|, 1e-8 * tet_mass); total mass conserved to 1e-9 relative when every
tet lies inside the grid.
"""

import numpy as np
import pytest

import adeco_amd as A
from adeco_amd.lattice import make_lattice, meta_patches
from tests import _oracle as O

pytestmark = pytest.mark.gpu

REL = 1e-6
MASS_REL = 1e-9


@pytest.fixture(autouse=True)
def exact_scaling():
    """Library and oracle both in per-axis cell scaling (the library default)."""
    with O.cell_scaling(O.EXACT):
        yield


def test_boundary_reference_scaling_mode():
    lat = make_lattice(10, 2.0, seed=3)
    pat = meta_patches(lat)
    bbox = A.cubic_root_box(lat.bbox())
    with O.cell_scaling(O.REFERENCE):
        data = np.zeros(16 ** 3, np.float32)
        A.DepositWrapper.deposit_tets(True, 0, 0, pat.size // 1536, pat, pat.size, 7, bbox, 16, data)
        want = O.deposit_patches_ld(pat, bbox, 16)
    assert O.max_rel_err(data, want) <= REL
    # the two scalings differ only through the box's float extents (one ulp of ~100 Mpc/h shifts
    # a cell face by ~1e-7 of a cell; low-mass cells next to such a face see up to ~1e-5)
    exact = O.deposit_patches_ld(pat, bbox, 16)
    assert 0 < O.max_rel_err(exact, want) <= 1e-4


def padded_box(bbox):
    """A cube strictly containing bbox with integer corners and a power-of-two edge, so its three
    cell sizes are identical in float (mass conservation then holds to rounding)."""
    b = np.asarray(bbox, np.float64)
    lo = np.floor(b[0::2]) - 1
    ext = 2.0 ** np.ceil(np.log2(np.max(b[1::2] - lo) + 1))
    out = np.empty(6, np.float32)
    out[0::2] = lo
    out[1::2] = lo + ext
    return out


@pytest.mark.parametrize("n,rms,dims", [(8, 1.0, 16), (12, 3.0, 24)])
def test_boundary_deposit_tets_matches_fp64_oracle(n, rms, dims):
    print(f"[boundary] n={n} rms={rms} dims={dims}", flush=True)
    lat = make_lattice(n, rms, seed=1)
    pat = meta_patches(lat)
    bbox = A.cubic_root_box(lat.bbox())
    data = np.zeros(dims ** 3, np.float32)
    A.DepositWrapper.deposit_tets(True, 0, 0, pat.size // 1536, pat, pat.size, 7, bbox, dims, data)
    want = O.deposit_patches_ld(pat, bbox, dims)
    # the boundary returns float (the reference's grid type): one float rounding per cell
    assert O.max_rel_err(data, want) <= REL
    # accumulation contract: a second call adds
    A.DepositWrapper.deposit_tets(True, 0, 0, pat.size // 1536, pat, pat.size, 7, bbox, dims, data)
    assert O.max_rel_err(data, 2 * want) <= REL


def test_boundary_patch_subrange_and_multipass():
    print("[boundary] subrange + multipass", flush=True)
    lat = make_lattice(16, 2.0, seed=7)
    pat = meta_patches(lat)
    npatch = pat.size // 1536
    bbox = A.cubic_root_box(lat.bbox())
    dims = 20
    off, cnt = 2, npatch - 3
    want = O.deposit_patches_ld(pat, bbox, dims, patch_offset=off, num_patches=cnt)
    one = np.zeros(dims ** 3, np.float32)
    A.DepositWrapper.deposit_tets(True, 0, off, cnt, pat, pat.size, 7, bbox, dims, one)
    assert O.max_rel_err(one, want) <= REL
    # a budget that holds the grid plus 2 patches forces several passes; result unchanged
    A.set_memory_budget(dims ** 3 * 8 + 2 * 1536 * 4 + 16)
    try:
        multi = np.zeros(dims ** 3, np.float32)
        A.DepositWrapper.deposit_tets(True, 0, off, cnt, pat, pat.size, 7, bbox, dims, multi)
    finally:
        A.set_memory_budget(0)
    assert O.max_rel_err(multi, want) <= REL
    # use_specific_gpu = False splits over all visible GPUs (one here)
    spread = np.zeros(dims ** 3, np.float32)
    A.DepositWrapper.deposit_tets(False, 0, off, cnt, pat, pat.size, 7, bbox, dims, spread)
    assert O.max_rel_err(spread, want) <= REL


def test_boundary_errors_on_gpu():
    lat = make_lattice(8, 1.0, seed=1)
    pat = meta_patches(lat)
    bbox = A.cubic_root_box(lat.bbox())
    data = np.zeros(8 ** 3, np.float32)
    with pytest.raises(A.AdRuntimeException):
        A.DepositWrapper.deposit_tets(True, 0, 1, pat.size // 1536, pat, pat.size, 7, bbox, 8, data)
    with pytest.raises(A.AdRuntimeException):
        A.DepositWrapper.deposit_tets(True, 99, 0, 1, pat, pat.size, 7, bbox, 8, data)
    # empty positions: no-op (AdDeposit2Grid.cpp:86-87)
    A.DepositWrapper.deposit_tets(True, 0, 0, 0, pat, 0, 7, bbox, 8, data)
    assert not data.any()


@pytest.mark.parametrize("decomp", [A.DECOMP_REF_EFFECTIVE, A.DECOMP_REF_INTENDED, A.DECOMP_KUHN])
@pytest.mark.parametrize("n,rms,dims", [(10, 1.0, 8), (10, 3.0, 32), (8, 0.3, 4)])
def test_lattice_plain_matches_fp64_oracle(decomp, n, rms, dims):
    print(f"[lattice plain] decomp={decomp} n={n} rms={rms} dims={dims}", flush=True)
    lat = make_lattice(n, rms, seed=3)
    box = padded_box(lat.bbox())
    with A.ResidentLattice.from_lattice(lat) as rl:
        got, st = rl.deposit_grid(box, dims, decomposition=decomp, tet_mass=1.0 / 6, stats=True)
    want = O.deposit_lattice(lat, box, dims, decomp=decomp, tet_mass=1.0 / 6, precision=3)
    assert O.max_rel_err(got, want, tet_mass=1.0 / 6) <= REL
    total = lat.n ** 3  # 6 n^3 tets of mass 1/6, all inside the padded box
    assert abs(got.sum() - total) <= MASS_REL * total
    assert int(st[0]) == lat.num_tets


def test_lattice_subset_and_bbox():
    lat = make_lattice(9, 2.0, seed=11)
    with A.ResidentLattice.from_lattice(lat) as rl:
        assert np.array_equal(rl.bbox(), lat.bbox())
        ids = np.arange(0, lat.num_tets, 7, dtype=np.int32)
        box = A.cubic_root_box(lat.bbox())
        got = rl.deposit_grid(box, 16, tet_ids=ids)
    want = O.deposit_lattice(lat, box, 16, tet_ids=ids, precision=3)
    assert O.max_rel_err(got, want) <= REL


def test_classification_regimes():
    """Tets much smaller than cells (tet-in-cell) and much larger (cell-in-tet) both match."""
    lat = make_lattice(6, 0.5, seed=5)
    box = padded_box(lat.bbox())
    with A.ResidentLattice.from_lattice(lat) as rl:
        coarse, st_c = rl.deposit_grid(box, 2, decomposition=A.DECOMP_KUHN, stats=True)
        fine, st_f = rl.deposit_grid(box, 96, decomposition=A.DECOMP_KUHN, stats=True)
    assert st_c[5] > 0 and st_f[6] > 0
    assert O.max_rel_err(coarse, O.deposit_lattice(lat, box, 2, decomp=2, precision=3)) <= REL
    assert O.max_rel_err(fine, O.deposit_lattice(lat, box, 96, decomp=2, precision=3)) <= REL


@pytest.mark.parametrize("dims", [24, 96])
def test_batched_column_segments_on_fine_grids(dims):
    """The batched FP64 form clips runs of up to 8 cells along each tet's longest axis with one
    x/y clip.  Tets spanning 4 (dims 24) to 20 (dims 96) cells per axis need several segments per
    column, full-inside segments and cells in both classification extremes; per cell against the
    extended-precision oracle, and the same grid as the plain form."""
    lat = make_lattice(6, 0.5, seed=5)
    box = padded_box(lat.bbox())
    tets = O.lattice_tets(lat, A.DECOMP_KUHN)
    print(f"[columns] dims={dims}", flush=True)
    got = A.deposit_soup(tets, box, dims, 1.0 / 6, mode=A.MODE_BATCHED, precision=A.PREC_FP64)
    plain = A.deposit_soup(tets, box, dims, 1.0 / 6, mode=A.MODE_PLAIN, precision=A.PREC_FP64)
    with O.cell_scaling(O.EXACT):
        want = O.deposit_soup_ld(tets, box, dims, 1.0 / 6)
    assert O.max_rel_err(got, want, tet_mass=1.0 / 6) <= REL
    assert O.max_rel_err(plain, want, tet_mass=1.0 / 6) <= REL
    assert abs(got.sum() - want.sum()) <= 1e-9 * want.sum()


@pytest.mark.parametrize("precision", [A.PREC_FP64, A.PREC_MIXED])
def test_boundary_batched_form_matches_oracle(precision):
    """DepositWrapper.deposit_tets in the batched form (the patches expanded into a tet soup on the
    device, then the batched persistent deposit): the oracle's grid to 1e-6, also over many passes
    of a small memory budget, and with another decomposition and tet mass."""
    print(f"[boundary] batched form, precision {precision}", flush=True)
    lat = make_lattice(16, 2.0, seed=7)
    pat = meta_patches(lat)
    npatch = pat.size // 1536
    bbox = A.cubic_root_box(lat.bbox())
    dims = 24
    want = O.deposit_patches_ld(pat, bbox, dims)
    want_kuhn = O.deposit_lattice(lat, bbox, dims, decomp=A.DECOMP_KUHN, tet_mass=1.0 / 6.0, precision=3)
    prev = A.set_boundary_form(mode=A.MODE_BATCHED, precision=precision)
    try:
        got = np.zeros(dims ** 3, np.float32)
        A.DepositWrapper.deposit_tets(True, 0, 0, npatch, pat, pat.size, 7, bbox, dims, got)
        assert O.max_rel_err(got, want) <= REL
        # about two patches per pass (each holds its patch, soup and candidate space)
        A.set_memory_budget(dims ** 3 * 8 + 1_200_000)
        multi = np.zeros(dims ** 3, np.float32)
        A.DepositWrapper.deposit_tets(True, 0, 0, npatch, pat, pat.size, 7, bbox, dims, multi)
        assert O.max_rel_err(multi, want) <= REL
        A.set_memory_budget(0)
        A.set_boundary_form(decomposition=A.DECOMP_KUHN, tet_mass=1.0 / 6.0)
        kuhn = np.zeros(dims ** 3, np.float32)
        A.DepositWrapper.deposit_tets(True, 0, 0, npatch, pat, pat.size, 7, bbox, dims, kuhn)
        assert O.max_rel_err(kuhn, want_kuhn, tet_mass=1.0 / 6.0) <= REL
        # the plain form with the same decomposition and mass
        A.set_boundary_form(mode=A.MODE_PLAIN, precision=A.PREC_FP64)
        plain = np.zeros(dims ** 3, np.float32)
        A.DepositWrapper.deposit_tets(True, 0, 0, npatch, pat, pat.size, 7, bbox, dims, plain)
        assert O.max_rel_err(plain, want_kuhn, tet_mass=1.0 / 6.0) <= REL
        # precision modes other than FP64 need the batched form
        A.set_boundary_form(precision=A.PREC_MIXED)
        with pytest.raises(A.AdRuntimeException, match="batched form"):
            A.DepositWrapper.deposit_tets(True, 0, 0, npatch, pat, pat.size, 7, bbox, dims, plain)
    finally:
        A.set_memory_budget(0)
        A.set_boundary_form(**prev)
