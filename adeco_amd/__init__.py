"""adeco_amd — MI355X-native exact phase-space-tetrahedron mass deposit (ADECO's hot path).

Python host mirror of the reference's C++ call surface, over the C ABI of libadeco_amd.so
(include/adeco.h).  Names, argument meaning and error behaviour follow the reference:

* ``DepositWrapper.deposit_tets`` — AdaptiveMassDeposit::DepositWrapper::deposit_tets
  (AdDeposit2Grid.h:82-91): deposits meta-patches into a float grid, accumulating (``+=``), and
  raises ``AdRuntimeException`` where the reference throws.
* ``ResidentLattice`` — the lattice kept in HBM for repeated deposits (octree leaves, ROI requests).

All compute runs in HIP kernels on the GPU; the library fails loudly when it or the GPU is missing.
"""

from __future__ import annotations

import os

import numpy as np

from . import _lib
from ._lib import load, ptr

ADECO_OK = 0
ADECO_EINVAL = 1
ADECO_EOVERFLOW = 2
ADECO_ENOMEM = 3
ADECO_EDEVICE = 4
ADECO_EINTERNAL = 5
ADECO_EIO = 6

DECOMP_REF_EFFECTIVE = 0
DECOMP_REF_INTENDED = 1
DECOMP_KUHN = 2

MODE_PLAIN = 0
MODE_BATCHED = 1

PREC_FP64 = 0
PREC_MIXED = 1
PREC_FP32 = 2

NUM_STATS = 11
STAT_NAMES = ("tets", "tets_active", "candidates", "aabb_pass", "none", "tet_in_cell", "cell_in_tet",
              "partial", "nonzero", "clip_stages", "clip_subtets")


class AdRuntimeException(RuntimeError):
    """Raised where the reference throws AdaptiveMassDeposit::AdRuntimeException (AdException.h:87-95)."""

    def __init__(self, status: int, message: str):
        super().__init__(message)
        self.status = status


def _check(status: int, what: str = ""):
    if status != ADECO_OK:
        msg = load().adeco_last_error().decode(errors="replace")
        raise AdRuntimeException(status, f"ERROR: {what}: {msg}" if what else f"ERROR: {msg}")


def device_count() -> int:
    return load().adeco_device_count()


def last_kernel_ms() -> float:
    """HIP-event time of the calling thread's last deposit launch(es), in ms."""
    return load().adeco_last_kernel_ms()


SCALING_EXACT = 0
SCALING_REFERENCE = 1


def set_cell_scaling(mode: int):
    """SCALING_EXACT (default): per-axis cell scaling, exact partition and mass conservation;
    SCALING_REFERENCE: the reference's 1/dx on every axis (AdDepositTets2Grid.h:569)."""
    _check(load().adeco_set_cell_scaling(int(mode)), "set_cell_scaling()")


def get_cell_scaling() -> int:
    return load().adeco_get_cell_scaling()


def set_boundary_form(decomposition=None, tet_mass=None, mode=None, precision=None):
    """Library-wide form of DepositWrapper.deposit_tets (the boundary call): the tets of a cube
    (default DECOMP_REF_EFFECTIVE), the mass of one tet (default 1), the kernel form (default
    MODE_PLAIN; MODE_BATCHED runs the batched persistent deposit on the patches' tets) and the
    precision (default PREC_FP64; the other modes need MODE_BATCHED).  Returns the previous form."""
    lib = load()
    prev = {"decomposition": lib.adeco_get_tet_decomposition(), "tet_mass": lib.adeco_get_tet_mass(),
            "mode": lib.adeco_get_deposit_mode(), "precision": lib.adeco_get_precision_mode()}
    if decomposition is not None:
        _check(lib.adeco_set_tet_decomposition(int(decomposition)), "set_tet_decomposition()")
    if tet_mass is not None:
        _check(lib.adeco_set_tet_mass(float(tet_mass)), "set_tet_mass()")
    if mode is not None:
        _check(lib.adeco_set_deposit_mode(int(mode)), "set_deposit_mode()")
    if precision is not None:
        _check(lib.adeco_set_precision_mode(int(precision)), "set_precision_mode()")
    return prev


def set_memory_budget(nbytes: int):
    """Per-call device memory budget (0 = 75% of device memory); lower values force multi-pass."""
    _check(load().adeco_set_memory_budget(int(nbytes)), "set_memory_budget()")


class DepositWrapper:
    """Mirror of AdaptiveMassDeposit::DepositWrapper (AdDeposit2Grid.h:59-93)."""

    h_grid_mass_type = np.float32
    vertex_type = np.float32
    mass_type = np.float32

    @staticmethod
    def get_cuda_grid_size() -> int:
        return load().adeco_get_grid_size()

    @staticmethod
    def get_cuda_block_size() -> int:
        return load().adeco_get_block_size()

    @staticmethod
    def deposit_tets(use_specific_gpu: bool, device_id_to_use: int, tet_offset: int, num_tets_2_process: int,
                     positions: np.ndarray, positions_size: int, meta_cells_per_patch: int, bbox,
                     dims: int, data: np.ndarray) -> None:
        """Deposits patches [tet_offset, tet_offset + num_tets_2_process) into ``data`` (+=).

        As in the reference, the two size arguments count meta-patches, not tets
        (AdDeposit2Grid.cpp:69-70).  ``positions`` is float32 AoS meta-patch data, ``bbox`` is
        x0,x1,y0,y1,z0,z1 and ``data`` a float32 array of dims^3 cells (i fastest), updated in place.
        """
        positions = np.ascontiguousarray(positions, dtype=np.float32)
        bbox = np.ascontiguousarray(bbox, dtype=np.float32)
        if data.dtype != np.float32 or not data.flags.c_contiguous:
            raise AdRuntimeException(ADECO_EINVAL, "ERROR: deposit_tets(): data must be a contiguous float32 array")
        if data.size != int(dims) ** 3 or bbox.size != 6:
            raise AdRuntimeException(ADECO_EINVAL, "ERROR: deposit_tets(): inconsistent grid or bbox size")
        if int(positions_size) > positions.size:
            raise AdRuntimeException(ADECO_EINVAL, "ERROR: deposit_tets(): positions_size exceeds the positions array")
        st = load().adeco_deposit_patches(int(bool(use_specific_gpu)), int(device_id_to_use), int(tet_offset),
                                          int(num_tets_2_process), ptr(positions), int(positions_size),
                                          int(meta_cells_per_patch), ptr(bbox), int(dims), ptr(data))
        _check(st, "deposit_tets()")


class AdOctreeGridWriter:
    """Mirror of AdaptiveMassDeposit::AdOctreeGridWriter (AdOctreeGridWriter.h:54-144): one raw
    float32 density file per leaf, ``<output_path>/node_<level>_<ix>_<iy>_<iz>``."""

    def write_meta_data(self, output_directory: str, filename: str, linear_patch_resolution: int, octree) -> None:
        """The reference's base writer stores no meta data (AdOctreeGridWriter.h:58-64)."""
        print("INFO: AdOctreeGridWriter::store_meta_data(): called for base class. No meta data will be written.")

    def write_patch(self, output_path: str, node: int, octree: "Octree", grid_dims: int, orig_data) -> float:
        """Writes node ``node`` of ``octree`` from its float mass grid (dims^3, i fastest) as
        densities and returns the grid's mass (double sum in index order)."""
        data = np.ascontiguousarray(orig_data, dtype=np.float32)
        if data.size != int(grid_dims) ** 3:
            raise AdRuntimeException(ADECO_EINVAL, "ERROR: write_patch(): grid size does not match grid_dims")
        nd = octree.nodes()
        ll = np.ascontiguousarray(nd["ll"][int(node)], np.float32)
        root = np.ascontiguousarray(nd["ll"][0], np.float32)
        mass = _lib.ctypes.c_double()
        _check(load().adeco_write_patch(os.fsencode(output_path), int(nd["level"][int(node)]), ptr(ll),
                                        float(nd["ext"][int(node)]), ptr(root), int(grid_dims), ptr(data),
                                        _lib.ctypes.byref(mass)), "write_patch()")
        return mass.value


class ResidentLattice:
    """A vertex lattice uploaded once to one GPU (SoA float32, i fastest)."""

    def __init__(self, x, y, z, dims, device: int = 0):
        lib = load()
        self._x = np.ascontiguousarray(x, dtype=np.float32)
        self._y = np.ascontiguousarray(y, dtype=np.float32)
        self._z = np.ascontiguousarray(z, dtype=np.float32)
        self.dims = tuple(int(d) for d in dims)
        self.device = device
        h = _lib.c_vp()
        _check(lib.adeco_lattice_create(device, ptr(self._x), ptr(self._y), ptr(self._z), *self.dims,
                                        _lib.ctypes.byref(h)), "lattice_create()")
        self.handle = h

    @classmethod
    def from_lattice(cls, lat, device: int = 0):
        return cls(lat.x, lat.y, lat.z, lat.dims, device)

    @classmethod
    def from_dark_sky(cls, filename: str, stride: int = 1, device: int = 0):
        """A DARK_SKY particle file (AdDarkSkyDataReader, AdPointDataReaders.h:361-767) read straight
        into device memory: each record's Morton id places its position on the lattice."""
        self = cls.__new__(cls)
        h = _lib.c_vp()
        _check(load().adeco_lattice_create_dark_sky(int(device), os.fsencode(filename), int(stride),
                                                    _lib.ctypes.byref(h)), "lattice_create_dark_sky()")
        self.handle = h
        self.device = device
        n = _lib.ctypes.c_int64()
        d = np.zeros(1, np.int32)
        _check(load().adeco_dark_sky_info(os.fsencode(filename), None, _lib.ctypes.byref(n), ptr(d)), "dark_sky_info()")
        side = int(d[0]) // int(stride)
        self.dims = (side, side, side)
        return self

    @property
    def num_tets(self) -> int:
        return 6 * (self.dims[0] - 1) * (self.dims[1] - 1) * (self.dims[2] - 1)

    def bbox(self) -> np.ndarray:
        out = np.zeros(6, np.float32)
        _check(load().adeco_lattice_bbox(self.handle, ptr(out)), "lattice_bbox()")
        return out

    def update(self, x, y, z):
        """Overwrites the resident positions with a new state of the same lattice (float32 SoA)."""
        n = self.dims[0] * self.dims[1] * self.dims[2]
        arrs = [np.ascontiguousarray(a, np.float32).ravel() for a in (x, y, z)]
        if any(a.size != n for a in arrs):
            raise AdRuntimeException(ADECO_EINVAL, "ERROR: lattice_update(): positions do not match the lattice")
        _check(load().adeco_lattice_update(self.handle, ptr(arrs[0]), ptr(arrs[1]), ptr(arrs[2])), "lattice_update()")

    def positions(self):
        """The resident vertex coordinates (x, y, z), float32 SoA."""
        n = self.dims[0] * self.dims[1] * self.dims[2]
        x, y, z = (np.empty(n, np.float32) for _ in range(3))
        _check(load().adeco_lattice_positions(self.handle, ptr(x), ptr(y), ptr(z)), "lattice_positions()")
        return x, y, z

    def periodic_unwrap(self, bbox, period=None):
        """apply_boundary_corrections (AdPointDataReaders.h:156-206) on the device with the global
        bbox ``bbox``.  Without ``period`` the shift is the bbox extent and ``bbox`` stays the root
        box of later octrees, as in the reference; with ``period`` (3 floats) vertices move by the
        period and later octrees use the corrected vertices' bbox."""
        b = np.ascontiguousarray(bbox, np.float32)
        per = None if period is None else np.ascontiguousarray(period, np.float32)
        if b.size != 6 or (per is not None and per.size != 3):
            raise AdRuntimeException(ADECO_EINVAL, "ERROR: periodic_unwrap(): bbox needs 6 and period 3 values")
        _check(load().adeco_lattice_periodic_unwrap(self.handle, ptr(b), ptr(per)), "lattice_periodic_unwrap()")

    def deposit_grid(self, bbox, dims: int, decomposition: int = DECOMP_REF_EFFECTIVE, tet_mass: float = 1.0,
                     tet_ids=None, mode: int = MODE_PLAIN, precision: int = PREC_FP64, stats: bool = False):
        """Deposits all (or the listed) lattice tets onto one dims^3 FP64 grid; returns it (and stats)."""
        bbox = np.ascontiguousarray(bbox, dtype=np.float32)
        grid = np.zeros(int(dims) ** 3, np.float64)
        ids = None if tet_ids is None else np.ascontiguousarray(tet_ids, dtype=np.int32)
        st = np.zeros(NUM_STATS, np.uint64) if stats else None
        _check(load().adeco_lattice_deposit_grid(self.handle, int(decomposition), float(tet_mass), ptr(ids),
                                                 0 if ids is None else ids.size, ptr(bbox), int(dims), int(mode),
                                                 int(precision), ptr(grid), ptr(st)), "lattice_deposit_grid()")
        return (grid, st) if stats else grid

    def close(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            load().adeco_lattice_destroy(self.handle)
            self.handle = _lib.c_vp()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:  # interpreter shutdown
            pass


class Octree:
    """The reference's one-rank octree over a resident lattice (MetaMesh + generate_tree + Octree).

    Mirrors AMD_Parameters' octree keys: ``max_tets_per_octree_node``, ``max_refinement_level``
    and the optional ``region_of_interest`` (x0,x1,y0,y1,z0,z1).  Node ids, counts, the DFS leaf
    list and the leaf processing order are the reference's.
    """

    def __init__(self, lattice: ResidentLattice, max_tets_per_octree_node: int, max_refinement_level: int,
                 region_of_interest=None):
        self.lattice = lattice
        roi = None if region_of_interest is None else np.ascontiguousarray(region_of_interest, np.float32)
        h = _lib.c_vp()
        _check(load().adeco_octree_build(lattice.handle, ptr(roi), int(max_tets_per_octree_node),
                                         int(max_refinement_level), _lib.ctypes.byref(h)), "octree_build()")
        self.handle = h
        self.dims = None

    @property
    def num_nodes(self) -> int:
        return load().adeco_octree_num_nodes(self.handle)

    @property
    def num_leaves(self) -> int:
        return load().adeco_octree_num_leaves(self.handle)

    @property
    def num_meta_cells(self) -> int:
        return load().adeco_octree_num_meta_cells(self.handle)

    def bbox(self) -> np.ndarray:
        out = np.zeros(6, np.float32)
        _check(load().adeco_octree_bbox(self.handle, ptr(out)), "octree_bbox()")
        return out

    def nodes(self) -> dict:
        n = self.num_nodes
        d = {"ll": np.zeros((n, 3), np.float32), "ext": np.zeros(n, np.float32), "level": np.zeros(n, np.int32),
             "parent": np.zeros(n, np.int32), "first_child": np.zeros(n, np.int32), "count": np.zeros(n, np.uint64)}
        _check(load().adeco_octree_nodes(self.handle, ptr(d["ll"]), ptr(d["ext"]), ptr(d["level"]), ptr(d["parent"]),
                                         ptr(d["first_child"]), ptr(d["count"])), "octree_nodes()")
        return d

    def leaves(self) -> dict:
        lib = load()
        nl = self.num_leaves
        total = lib.adeco_octree_leaves(self.handle, None, None, None, None)
        d = {"dfs": np.zeros(nl, np.int32), "order": np.zeros(nl, np.int32),
             "mc_offsets": np.zeros(nl + 1, np.int64), "mc_ids": np.zeros(max(total, 1), np.int32)}
        lib.adeco_octree_leaves(self.handle, ptr(d["dfs"]), ptr(d["order"]), ptr(d["mc_offsets"]), ptr(d["mc_ids"]))
        d["mc_ids"] = d["mc_ids"][:total]
        return d

    def meta_bboxes(self) -> np.ndarray:
        out = np.zeros((self.num_meta_cells, 6), np.float32)
        _check(load().adeco_octree_meta_bboxes(self.handle, ptr(out)), "octree_meta_bboxes()")
        return out

    def leaf_boxes(self) -> np.ndarray:
        """Float (ll, ll + ext) boxes of the leaves in DFS order, as the reference forms them."""
        nd = self.nodes()
        dfs = self.leaves()["dfs"]
        out = np.empty((dfs.size, 6), np.float32)
        out[:, 0::2] = nd["ll"][dfs]
        out[:, 1::2] = (nd["ll"][dfs] + nd["ext"][dfs][:, None]).astype(np.float32)
        return out

    def leaf_boxes_f64(self, exact: bool = True) -> np.ndarray:
        """Leaf boxes [num_leaves, 6] (x0,x1,y0,y1,z0,z1) as the deposit uses them: the exact dyadic
        FP64 subdivision of the root (SCALING_EXACT) or the float reference boxes."""
        out = np.zeros((self.num_leaves, 6), np.float64)
        _check(load().adeco_octree_leaf_boxes(self.handle, int(bool(exact)), ptr(out)), "octree_leaf_boxes()")
        return out

    def root_box_f64(self) -> np.ndarray:
        """The root box (ll, ll + ext) in FP64 (the region the exact leaves tile)."""
        nd = self.nodes()
        ll, ext = nd["ll"][0].astype(np.float64), float(nd["ext"][0])
        out = np.empty(6)
        out[0::2] = ll
        out[1::2] = ll + ext
        return out

    def bin(self, decomposition: int = DECOMP_REF_EFFECTIVE) -> np.ndarray:
        counts = np.zeros(self.num_leaves, np.int64)
        _check(load().adeco_octree_bin(self.handle, int(decomposition), ptr(counts)), "octree_bin()")
        return counts

    def deposit(self, dims: int, decomposition: int = DECOMP_REF_EFFECTIVE, tet_mass: float = 1.0,
                mode: int = MODE_PLAIN, precision: int = PREC_FP64) -> float:
        """deposit_leaf_nodes on one rank; returns the kernel time in ms (grids stay on the GPU)."""
        _check(load().adeco_deposit_leaves(self.handle, int(dims), int(decomposition), float(tet_mass), int(mode),
                                           int(precision)), "deposit_leaves()")
        self.dims = int(dims)
        return last_kernel_ms()

    def leaf_tets(self, leaf: int) -> np.ndarray:
        """Binned lattice tet ids of one leaf (DFS index)."""
        lib = load()
        n = lib.adeco_octree_leaf_tets(self.handle, int(leaf), None)
        if n < 0:
            raise AdRuntimeException(ADECO_EINVAL, "ERROR: leaf_tets(): not binned or bad leaf")
        out = np.zeros(max(n, 1), np.int32)
        lib.adeco_octree_leaf_tets(self.handle, int(leaf), ptr(out))
        return out[:n]

    def autotune(self, dims: int, decomposition: int = DECOMP_REF_EFFECTIVE, tet_mass: float = 1.0,
                 precision: int = PREC_FP64, reps: int = 2) -> dict:
        """Picks the batched kernel's launch shape (tiles per work-queue pop, workgroups per CU) by
        timing this octree's leaf deposit; the choice stays in effect for later deposits."""
        c, b, ms = _lib.ctypes.c_int(), _lib.ctypes.c_int(), _lib.ctypes.c_double()
        _check(load().adeco_autotune_batched(self.handle, int(dims), int(decomposition), float(tet_mass), int(precision),
                                             int(reps), _lib.ctypes.byref(c), _lib.ctypes.byref(b), _lib.ctypes.byref(ms)),
               "autotune_batched()")
        self.dims = int(dims)
        return {"chunk": c.value, "blocks_per_cu": b.value, "kernel_ms": ms.value}

    @staticmethod
    def last_timings() -> dict:
        out = np.zeros(5, np.float64)
        load().adeco_last_timings(ptr(out))
        return {"total_ms": out[0], "main_ms": out[1], "candidates": int(out[2]), "partials": int(out[3]),
                "fallbacks": int(out[4])}

    def leaf_grids(self, first: int = 0, count=None) -> np.ndarray:
        """FP64 mass grids [count, dims^3] of leaves first..first+count (DFS order)."""
        count = self.num_leaves - first if count is None else count
        out = np.zeros((count, self.dims ** 3), np.float64)
        _check(load().adeco_leaf_grids(self.handle, int(first), int(count), ptr(out)), "leaf_grids()")
        return out

    def write_leaves(self, output_path: str):
        """Writes every leaf of the last deposit as a ``node_<level>_<ix>_<iy>_<iz>`` density file
        (AdOctreeGridWriter::write_patch, AdOctreeGridWriter.h:67-144).  Returns (per-leaf masses in
        DFS order, total mass summed in the leaf processing order)."""
        masses = np.zeros(self.num_leaves, np.float64)
        total = _lib.ctypes.c_double()
        _check(load().adeco_write_leaves(self.handle, os.fsencode(output_path), ptr(masses), _lib.ctypes.byref(total)),
               "write_leaves()")
        return masses, total.value

    def close(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            load().adeco_octree_destroy(self.handle)
            self.handle = _lib.c_vp()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DepositGraph:
    """The batched leaf deposit of an octree captured as a hipGraph and replayed by ``launch()``;
    ``leaf_grids()`` reads the graph's own FP64 leaf grids (DFS leaf order).

    ``step=False``: grid clear, work-queue reset and persistent deposit of the candidate space built
    at creation.  ``step=True``: the whole step (binning of the lattice's current positions, pair
    setup, scans, tile map, deposit) with device-side sizes, so replays follow
    ``ResidentLattice.update``; buffers hold ``slack`` times the capture's work, beyond which a
    replay raises ADECO_EOVERFLOW."""

    def __init__(self, tree: "Octree", dims: int, decomposition: int = DECOMP_REF_EFFECTIVE, tet_mass: float = 1.0,
                 precision: int = PREC_FP64, step: bool = False, slack: float = 1.25):
        self.tree = tree
        self.dims = int(dims)
        h = _lib.c_vp()
        if step:
            _check(load().adeco_deposit_graph_create_step(tree.handle, int(dims), int(decomposition), float(tet_mass),
                                                          int(precision), float(slack), _lib.ctypes.byref(h)),
                   "deposit_graph_create_step()")
        else:
            _check(load().adeco_deposit_graph_create(tree.handle, int(dims), int(decomposition), float(tet_mass),
                                                     int(precision), _lib.ctypes.byref(h)), "deposit_graph_create()")
        self.handle = h

    def launch(self) -> float:
        _check(load().adeco_deposit_graph_launch(self.handle), "deposit_graph_launch()")
        return last_kernel_ms()

    def leaf_grids(self, first: int = 0, count=None) -> np.ndarray:
        n = self.tree.num_leaves - first if count is None else int(count)
        out = np.empty((n, self.dims ** 3), np.float64)
        _check(load().adeco_deposit_graph_grids(self.handle, int(first), n, ptr(out)), "deposit_graph_grids()")
        return out

    def close(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            load().adeco_deposit_graph_destroy(self.handle)
            self.handle = _lib.c_vp()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:  # interpreter shutdown
            pass


class Server:
    """Continuous-batching server of region-of-interest density requests over a resident lattice.

    ``submit(roi, dims)`` queues a request for the FP32 density (mass / cell volume) of every tet
    whose AABB strictly overlaps the box ``roi`` (x0,x1,y0,y1,z0,z1) on a dims^3 grid over it;
    ``wait(id)`` returns the density as a (dims, dims, dims) array indexed [k, j, i] with the
    request's timing.  Requests are admitted while a slot is free and the memory budget allows,
    and every step launch shares its tiles among all admitted requests.
    """

    def __init__(self, lattice: ResidentLattice, decomposition: int = DECOMP_KUHN, tet_mass: float = 1.0 / 6.0,
                 precision: int = PREC_FP64, max_inflight: int = 16, step_tiles: int = 0, memory_budget: int = 0):
        self.lattice = lattice
        cfg = _lib.ServeConfig(int(decomposition), int(precision), float(tet_mass), int(max_inflight), 0,
                               int(step_tiles), int(memory_budget))
        h = _lib.c_vp()
        _check(load().adeco_serve_create(lattice.handle, _lib.ctypes.byref(cfg), _lib.ctypes.byref(h)), "serve_create()")
        self.handle = h
        self._out = {}

    def submit(self, roi, dims: int) -> int:
        roi = np.ascontiguousarray(roi, np.float64)
        out = np.empty(int(dims) ** 3, np.float32)
        rid = _lib.ctypes.c_uint64()
        _check(load().adeco_serve_submit(self.handle, ptr(roi), int(dims), ptr(out), _lib.ctypes.byref(rid)),
               "serve_submit()")
        self._out[rid.value] = out
        return rid.value

    def wait(self, request_id: int, timeout: float = -1.0):
        info = _lib.RequestInfo()
        info.status = -1  # left unset by a wait that timed out: the request still owns its buffer
        st = load().adeco_serve_wait(self.handle, int(request_id), float(timeout), _lib.ctypes.byref(info))
        out = self._out.pop(int(request_id), None) if info.status != -1 else None
        _check(st, "serve_wait()")
        d = int(info.dims)
        return out.reshape(d, d, d), {"n_tets": int(info.n_tets), "n_candidates": int(info.n_candidates),
                                      "t_submit": info.t_submit, "t_admit": info.t_admit, "t_done": info.t_done,
                                      "latency": info.t_done - info.t_submit}

    def stats(self) -> dict:
        st = _lib.ServeStats()
        _check(load().adeco_serve_stats_get(self.handle, _lib.ctypes.byref(st)), "serve_stats()")
        return {f: getattr(st, f) for f, _ in st._fields_}

    def close(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            load().adeco_serve_destroy(self.handle)
            self.handle = _lib.c_vp()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:  # interpreter shutdown
            pass


def set_launch_shape(chunk: int, blocks_per_cu: int):
    """Batched kernel launch shape: tiles of 64 candidates per work-queue pop, workgroups per CU."""
    _check(load().adeco_set_launch_shape(int(chunk), int(blocks_per_cu)), "set_launch_shape()")


def get_launch_shape() -> tuple:
    c, b = _lib.ctypes.c_int(), _lib.ctypes.c_int()
    load().adeco_get_launch_shape(_lib.ctypes.byref(c), _lib.ctypes.byref(b))
    return c.value, b.value


def roi_deposit_plain(lattice: ResidentLattice, roi, dims: int, decomposition: int = DECOMP_KUHN,
                      tet_mass: float = 1.0 / 6.0):
    """The serial baseline of Server: one request, plain one-tet-per-thread kernel, synchronous.
    Returns the (dims, dims, dims) FP32 density and the number of selected tets."""
    roi = np.ascontiguousarray(roi, np.float64)
    out = np.empty(int(dims) ** 3, np.float32)
    n = _lib.ctypes.c_int64()
    _check(load().adeco_roi_deposit_plain(lattice.handle, int(decomposition), float(tet_mass), ptr(roi), int(dims),
                                          ptr(out), _lib.ctypes.byref(n)), "roi_deposit_plain()")
    return out.reshape(int(dims), int(dims), int(dims)), int(n.value)


def deposit_soup(tets, bbox, dims: int, tet_mass: float = 1.0, device: int = 0, mode: int = MODE_PLAIN,
                 precision: int = PREC_FP64, stats: bool = False):
    """Deposits a float32 tet soup [n, 4, 3] onto one dims^3 FP64 grid."""
    tets = np.ascontiguousarray(tets, dtype=np.float32).reshape(-1, 12)
    bbox = np.ascontiguousarray(bbox, dtype=np.float32)
    grid = np.zeros(int(dims) ** 3, np.float64)
    st = np.zeros(NUM_STATS, np.uint64) if stats else None
    _check(load().adeco_deposit_soup(device, ptr(tets), tets.shape[0], ptr(bbox), int(dims), float(tet_mass),
                                     int(mode), int(precision), ptr(grid), ptr(st)), "deposit_soup()")
    return (grid, st) if stats else grid


def cubic_root_box(bbox) -> np.ndarray:
    """The octree root box of a bbox: (min, min + max extent) per axis in float32 (AdOctreeGeneration.h:98)."""
    b = np.asarray(bbox, dtype=np.float32)
    ext = max(b[1] - b[0], b[3] - b[2], b[5] - b[4])
    mn = b[0::2]
    out = np.empty(6, np.float32)
    out[0::2] = mn
    out[1::2] = (mn + np.float32(ext)).astype(np.float32)
    return out
