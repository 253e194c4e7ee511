"""ctypes bindings of the oracle (oracle/liboracle.so) and of the compiled reference harness
(oracle/_ref/libadeco_ref.so).  Test infrastructure: used only as the checker."""

from __future__ import annotations

import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle", "liboracle.so")
REF = os.path.join(ROOT, "oracle", "_ref", "libadeco_ref.so")

vp = ctypes.c_void_p
sz = ctypes.c_size_t
i64 = ctypes.c_int64
u64 = ctypes.c_uint64
dbl = ctypes.c_double
cint = ctypes.c_int

_ORACLE_SIGS = {
    "oracle_deposit_patches_f32": (cint, [sz, sz, vp, sz, cint, vp, cint, vp]),
    "oracle_deposit_patches_f64": (cint, [sz, sz, vp, sz, cint, vp, cint, dbl, vp, vp]),
    "oracle_deposit_lattice": (cint, [vp, vp, vp, vp, cint, vp, i64, vp, cint, dbl, cint, vp, vp, cint, vp]),
    "oracle_deposit_lattice_d": (cint, [vp, vp, vp, vp, cint, vp, i64, vp, cint, dbl, cint, vp, vp, cint, vp]),
    "oracle_deposit_soup_f64": (cint, [vp, i64, vp, cint, dbl, vp, vp]),
    "oracle_deposit_soup_ld": (cint, [vp, i64, vp, cint, dbl, vp]),
    "oracle_deposit_patches_ld": (cint, [sz, sz, vp, sz, cint, vp, cint, dbl, vp]),
    "oracle_clip_unit_cube_f32": (ctypes.c_float, [vp]),
    "oracle_clip_unit_cube_f64": (dbl, [vp, vp]),
    "oracle_clip_unit_cube_ld": (dbl, [vp]),
    "oracle_intersection_test_f64": (cint, [vp]),
    "oracle_vol6_f64": (dbl, [vp]),
    "oracle_decomposition": (None, [cint, vp]),
    "oracle_octree_build": (vp, [vp, vp, vp, vp, vp, u64, cint]),
    "oracle_octree_build_in": (vp, [vp, vp, vp, vp, vp, vp, u64, cint]),
    "oracle_periodic_unwrap": (None, [vp, vp, vp, vp, vp, vp]),
    "oracle_octree_free": (None, [vp]),
    "oracle_octree_num_nodes": (i64, [vp]),
    "oracle_octree_num_leaves": (i64, [vp]),
    "oracle_octree_num_meta_cells": (i64, [vp]),
    "oracle_octree_bbox": (None, [vp, vp]),
    "oracle_octree_nodes": (None, [vp, vp, vp, vp, vp, vp, vp, vp]),
    "oracle_octree_leaves": (i64, [vp, vp, vp, vp, vp]),
    "oracle_meta_cell_bboxes": (None, [vp, vp]),
    "oracle_estimate": (u64, [vp, vp, i64]),
    "oracle_set_cell_scaling": (None, [cint]),
    "oracle_get_cell_scaling": (cint, []),
}

EXACT = 0  # per-axis cell scaling (library default)
REFERENCE = 1  # the reference's 1/dx on every axis (oracle default)


class cell_scaling:
    """Context manager: sets the oracle's FP64 cell scaling (and the library's, when loaded)."""

    def __init__(self, mode, library=True):
        self.mode = mode
        self.library = library

    def __enter__(self):
        o = load()
        self.prev = o.oracle_get_cell_scaling()
        o.oracle_set_cell_scaling(self.mode)
        if self.library:
            import adeco_amd as A
            self.prev_lib = A.get_cell_scaling()
            A.set_cell_scaling(self.mode)
        return self

    def __exit__(self, *exc):
        load().oracle_set_cell_scaling(self.prev)
        if self.library:
            import adeco_amd as A
            A.set_cell_scaling(self.prev_lib)

_REF_SIGS = {
    "ref_last_error": (ctypes.c_char_p, []),
    "ref_deposit_tets": (cint, [sz, sz, vp, sz, cint, vp, cint, vp]),
    "ref_octree_build": (cint, [vp, vp, vp, cint, cint, cint, vp, u64, cint]),
    "ref_octree_bbox": (None, [vp]),
    "ref_octree_num_nodes": (i64, []),
    "ref_octree_num_leaves": (i64, []),
    "ref_octree_num_meta_cells": (i64, []),
    "ref_octree_nodes": (None, [vp, vp, vp, vp, vp, vp, vp]),
    "ref_octree_leaves": (i64, [vp, vp, vp, vp]),
    "ref_meta_cell_bboxes": (None, [vp]),
    "ref_estimate": (u64, [vp, i64]),
    "ref_octree_deposit_leaf": (cint, [ctypes.c_int32, cint, vp, vp, vp, ctypes.c_char_p]),
    "ref_cli_main": (cint, [ctypes.c_char_p]),
    "ref_parse_parameter_file": (cint, [ctypes.c_char_p, vp, cint, vp, vp]),
    "ref_apply_boundary_corrections": (cint, [vp, vp, vp, cint, cint, cint, vp]),
    "ref_sorted_by_id_load": (i64, [ctypes.c_char_p, i64, vp, vp, vp, vp, vp]),
    "ref_dark_sky_load": (i64, [ctypes.c_char_p, cint, i64, vp, vp, vp, vp, vp]),
    "ref_dark_sky_header_size": (i64, [ctypes.c_char_p]),
}

_cache = {}


def _bind(path, sigs):
    lib = ctypes.CDLL(path)
    for name, (res, args) in sigs.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


def load():
    if "o" not in _cache:
        if not os.path.exists(ORACLE):
            raise FileNotFoundError(f"{ORACLE} missing: run `make -C oracle`")
        _cache["o"] = _bind(ORACLE, _ORACLE_SIGS)
    return _cache["o"]


def load_ref():
    if "r" not in _cache:
        _cache["r"] = _bind(REF, _REF_SIGS) if os.path.exists(REF) else None
    return _cache["r"]


def P(a):
    return None if a is None else vp(a.ctypes.data)


def deposit_patches_f64(patches, bbox, dims, tet_mass=1.0, patch_offset=0, num_patches=None, stats=False):
    o = load()
    patches = np.ascontiguousarray(patches, np.float32)
    bbox = np.ascontiguousarray(bbox, np.float32)
    n = patches.size // 1536 if num_patches is None else num_patches
    g = np.zeros(dims ** 3, np.float64)
    st = np.zeros(16, np.uint64)
    rc = o.oracle_deposit_patches_f64(patch_offset, n, P(patches), patches.size, 7, P(bbox), dims, tet_mass,
                                      P(g), P(st))
    assert rc == 0
    return (g, st) if stats else g


def deposit_patches_ld(patches, bbox, dims, tet_mass=1.0, patch_offset=0, num_patches=None):
    """deposit_patches_f64 in x87 extended precision (64-bit significand), rounded to double."""
    o = load()
    patches = np.ascontiguousarray(patches, np.float32)
    bbox = np.ascontiguousarray(bbox, np.float32)
    n = patches.size // 1536 if num_patches is None else num_patches
    g = np.zeros(dims ** 3, np.float64)
    assert o.oracle_deposit_patches_ld(patch_offset, n, P(patches), patches.size, 7, P(bbox), dims, tet_mass,
                                       P(g)) == 0
    return g


def deposit_patches_f32(patches, bbox, dims, patch_offset=0, num_patches=None):
    o = load()
    patches = np.ascontiguousarray(patches, np.float32)
    bbox = np.ascontiguousarray(bbox, np.float32)
    n = patches.size // 1536 if num_patches is None else num_patches
    g = np.zeros(dims ** 3, np.float32)
    assert o.oracle_deposit_patches_f32(patch_offset, n, P(patches), patches.size, 7, P(bbox), dims, P(g)) == 0
    return g


def deposit_lattice(lat, bbox, dims, decomp=0, tet_mass=1.0, tet_ids=None, precision=1, nthreads=1, stats=False):
    """precision 0 = float-faithful (float grid), 1 = FP64, 2 = FP64 with op counting,
    3 = x87 extended precision rounded to double.
    A float64 bbox is passed to the oracle as FP64 (exact dyadic leaf boxes); otherwise float32."""
    o = load()
    np_ = np.array(lat.dims, np.int32)
    fn = o.oracle_deposit_lattice_d if np.asarray(bbox).dtype == np.float64 else o.oracle_deposit_lattice
    bbox = np.ascontiguousarray(bbox, np.float64 if fn is o.oracle_deposit_lattice_d else np.float32)
    ids = None if tet_ids is None else np.ascontiguousarray(tet_ids, np.int32)
    n = lat.num_tets if ids is None else ids.size
    gf = np.zeros(dims ** 3, np.float32) if precision == 0 else None
    gd = np.zeros(dims ** 3, np.float64) if precision != 0 else None
    st = np.zeros(16, np.uint64)
    fn(P(lat.x), P(lat.y), P(lat.z), P(np_), decomp, P(ids), n, P(bbox), dims, tet_mass, precision, P(gf), P(gd),
       nthreads, P(st))
    g = gf if precision == 0 else gd
    return (g, st) if stats else g


def deposit_soup_f64(tets, bbox, dims, tet_mass=1.0):
    o = load()
    tets = np.ascontiguousarray(tets, np.float32).reshape(-1, 12)
    bbox = np.ascontiguousarray(bbox, np.float32)
    g = np.zeros(dims ** 3, np.float64)
    o.oracle_deposit_soup_f64(P(tets), tets.shape[0], P(bbox), dims, tet_mass, P(g), None)
    return g


def lattice_tets(lat, decomp):
    """All tets of a lattice as a float32 soup [6 n^3, 4, 3], tet id = 6 * cube + local tet, cube
    index i fastest (the oracle's and the library's numbering)."""
    c = np.zeros(24, np.int32)
    load().oracle_decomposition(decomp, P(c))
    n = lat.n
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    xyz = np.stack([lat.x, lat.y, lat.z], 1)
    out = np.empty((n ** 3, 6, 4, 3), np.float32)
    for t in range(6):
        for q in range(4):
            code = int(c[4 * t + q])
            vi = (i + (code & 1)) + (n + 1) * ((j + ((code >> 1) & 1)) + (n + 1) * (k + ((code >> 2) & 1)))
            out[:, t, q] = xyz[vi]
    return out.reshape(-1, 4, 3)


def deposit_soup_ld(tets, bbox, dims, tet_mass=1.0):
    """The reference algorithm in x87 extended precision (64-bit significand), rounded to double."""
    o = load()
    tets = np.ascontiguousarray(tets, np.float32).reshape(-1, 12)
    bbox = np.ascontiguousarray(bbox, np.float32)
    g = np.zeros(dims ** 3, np.float64)
    assert o.oracle_deposit_soup_ld(P(tets), tets.shape[0], P(bbox), dims, tet_mass, P(g)) == 0
    return g


def ref_deposit_tets(patches, bbox, dims, patch_offset=0, num_patches=None):
    r = load_ref()
    patches = np.ascontiguousarray(patches, np.float32)
    bbox = np.ascontiguousarray(bbox, np.float32)
    n = patches.size // 1536 if num_patches is None else num_patches
    g = np.zeros(dims ** 3, np.float32)
    assert r.ref_deposit_tets(patch_offset, n, P(patches), patches.size, 7, P(bbox), dims, P(g)) == 0
    return g


def max_rel_err(a, b, tet_mass=1.0, floor_frac=1e-8):
    """Per-cell |a - b| / max(|b|, floor), floor = floor_frac * tet_mass.

    Cells whose oracle mass is below 1e-8 of one tet's mass are held to an absolute error of
    1e-8 * tolerance * tet_mass (1e-14 m at the 1e-6 bar): there two FP64 evaluations of a sliver
    already differ by cancellation at the 1e-16 absolute level.  Every other cell is compared
    relatively."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    denom = np.maximum(np.abs(b), floor_frac * tet_mass)
    return float(np.max(np.abs(a - b) / denom)) if a.size else 0.0
