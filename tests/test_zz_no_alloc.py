"""The allocation-free boundary entry point, adeco_deposit_patches_noalloc, called through the C ABI.

It keeps its device memory in the library's fixed scratch array (no hipMalloc), so it is held to
the boundary call it mirrors, adeco_deposit_patches in the plain FP64 form on one device: two calls
in a row on the same device, the first at 128^3 cells with 4913 meta-patches (two passes
through the scratch array) adding into a float grid that already holds values, the second smaller
(16^3 cells, a patch sub-range) into a zeroed grid.  Both results match the boundary call's to
1e-6 per cell (floor 1e-8 of a tet mass), and the second matches the extended-precision oracle as
well (1e-6 per cell, total mass to 1e-6).
"""

import numpy as np
import pytest

import adeco_amd as A
from adeco_amd import _lib
from adeco_amd.lattice import make_lattice, meta_patches
from tests import _oracle as O

pytestmark = pytest.mark.gpu


def _noalloc():
    return _lib.load().adeco_deposit_patches_noalloc  # (signature from adeco_amd._lib, as include/adeco.h)


def _rel(a, b, floor):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), floor)))


def test_noalloc_matches_boundary_call_on_two_calls():
    f = _noalloc()
    prev = A.set_boundary_form(decomposition=A.DECOMP_REF_EFFECTIVE, tet_mass=1.0, mode=A.MODE_PLAIN,
                               precision=A.PREC_FP64)
    try:
        rng = np.random.default_rng(5)
        # (lattice cubes, rms, seed, grid dims, first patch, patches or None for the rest)
        for n, rms, seed, dims, first, count in ((119, 2.0, 7, 128, 0, None), (12, 3.0, 2, 16, 2, 5)):
            lat = make_lattice(n, rms, seed=seed)
            pat = meta_patches(lat)
            total = pat.size // 1536
            count = total - first if count is None else count
            box = np.asarray(A.cubic_root_box(lat.bbox()), np.float32)
            init = rng.uniform(0.0, 2.0, dims ** 3).astype(np.float32) if dims == 128 else np.zeros(dims ** 3, np.float32)
            print(f"[noalloc] n={n} dims={dims} patches {first}..{first + count} of {total}", flush=True)
            want = init.copy()
            A.DepositWrapper.deposit_tets(True, 0, first, count, pat, pat.size, 7, box, dims, want)
            got = init.copy()
            rc = f(0, first, count, pat.ctypes.data, pat.size, 7, box.ctypes.data, dims, got.ctypes.data)
            assert rc == 0, _lib.load().adeco_last_error()
            assert _rel(got, want, 1e-8) <= 1e-6
            assert float(np.abs(got - init).sum()) > 0.0
            if dims == 16:
                sub = pat[first * 1536:(first + count) * 1536]
                with O.cell_scaling(O.EXACT):
                    oracle = O.deposit_patches_ld(sub, box, dims)
                assert _rel(got, oracle, 1e-8) <= 1e-6
                assert abs(float(got.astype(np.float64).sum()) - oracle.sum()) <= 1e-6 * oracle.sum()
    finally:
        A.set_boundary_form(**prev)


def test_noalloc_rejects_a_grid_beyond_its_scratch():
    f = _noalloc()
    lat = make_lattice(6, 1.0, seed=1)
    pat = meta_patches(lat)
    box = np.asarray(A.cubic_root_box(lat.bbox()), np.float32)
    grid = np.zeros(256 ** 3, np.float32)
    rc = f(0, 0, pat.size // 1536, pat.ctypes.data, pat.size, 7, box.ctypes.data, 256, grid.ctypes.data)
    assert rc == A.ADECO_EOVERFLOW
    assert b"scratch" in _lib.load().adeco_last_error()
    assert not grid.any()
    assert _lib.load().adeco_noalloc_scratch_bytes() == 48 << 20
