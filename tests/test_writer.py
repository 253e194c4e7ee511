"""Leaf files (AdOctreeGridWriter::write_patch, AdOctreeGridWriter.h:67-144) through the C ABI.

adeco_write_patch is host code (no GPU): given the reference's own float leaf grids (rebuilt bit
for bit by the float-faithful oracle, tests/golden/ref_leaves_c2.npz), it must write the
reference's files - same name, same sha256 of the float32 densities - and return the reference's
double mass sum.
"""

import ctypes
import hashlib
import os

import numpy as np
import pytest

import adeco_amd as A
from adeco_amd._lib import load, ptr
from tests._golden_leaves import patch_density, patch_file_name, reference_leaf_grids

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "ref_leaves_c2.npz")


def write_patch(out_dir, level, ll, ext, root_ll, dims, grid):
    m = ctypes.c_double()
    st = load().adeco_write_patch(os.fsencode(out_dir), int(level), ptr(np.ascontiguousarray(ll, np.float32)),
                                  float(ext), ptr(np.ascontiguousarray(root_ll, np.float32)), int(dims),
                                  ptr(np.ascontiguousarray(grid, np.float32)), ctypes.byref(m))
    return st, m.value


def test_write_patch_reproduces_reference_files(tmp_path):
    g = np.load(GOLDEN, allow_pickle=False)
    dims = int(g["dims"])
    _, d, grids = reference_leaf_grids(g)
    for i, (node, grid) in enumerate(zip(g["node"], grids)):
        st, mass = write_patch(tmp_path, d["level"][node], d["ll"][node], d["ext"][node], d["ll"][0], dims, grid)
        assert st == A.ADECO_OK, load().adeco_last_error()
        # the file name the compiled reference wrote this leaf under (make_golden.py read it back)
        path = tmp_path / patch_file_name(d, node, dims)
        dens = np.fromfile(path, np.float32)
        assert dens.size == dims ** 3
        assert hashlib.sha256(dens.tobytes()).hexdigest() == str(g["density_sha"][i]), int(node)
        assert mass == float(g["mass"][i]), int(node)
        want, want_mass = patch_density(d, node, dims, grid)
        assert np.array_equal(dens, want) and mass == want_mass
        os.remove(path)


def test_write_patch_closed_forms(tmp_path):
    """A dyadic box: cell volume (ext/dims)^3 is exact, so density = mass * dims^3 / ext^3 exactly;
    offsets are (ll - root_ll) / (ext / dims) cells."""
    dims, ext = 8, 0.5
    root = np.array([-1.0, 2.0, 0.25], np.float32)
    ll = root + np.array([1.5, 0.0, 0.5], np.float32)
    grid = np.arange(dims ** 3, dtype=np.float32) * np.float32(0.125)
    st, mass = write_patch(tmp_path, 3, ll, ext, root, dims, grid)
    assert st == A.ADECO_OK
    dens = np.fromfile(tmp_path / "node_3_24_0_8", np.float32)
    assert np.array_equal(dens, grid * np.float32(dims ** 3 / ext ** 3))
    assert mass == float(np.sum(grid.astype(np.float64)))


def test_write_patch_errors(tmp_path):
    grid = np.zeros(8, np.float32)
    st, _ = write_patch(tmp_path / "missing" / "dir", 0, [0, 0, 0], 1.0, [0, 0, 0], 2, grid)
    assert st == A.ADECO_EIO
    assert b"failed to write node data" in load().adeco_last_error()
    assert write_patch(tmp_path, 0, [0, 0, 0], 1.0, [0, 0, 0], 0, grid)[0] == A.ADECO_EINVAL
    assert write_patch(tmp_path, 0, [0, 0, 0], 0.0, [0, 0, 0], 2, grid)[0] == A.ADECO_EINVAL
    assert write_patch(tmp_path, -1, [0, 0, 0], 1.0, [0, 0, 0], 2, grid)[0] == A.ADECO_EINVAL
    with pytest.raises(A.AdRuntimeException):
        A.AdOctreeGridWriter().write_patch(str(tmp_path), 0, None, 4, grid)
