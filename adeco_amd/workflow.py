"""The reference's single-rank program on one GPU, over the C ABI (include/adeco.h).

Mirrors, by name and argument meaning:

* ``AMD_Parameters`` and ``parse_parameter_file`` (AdTypeDefinitions.h:1218-1247, main.cpp:83-183);
* ``get_reader``, ``AdSortedByIdsPointDataReader`` and ``AdDarkSkyDataReader``
  (AdPointDataReaders.h:361-930);
* ``deposit_workflow`` (AdDepositWorkFlow.h:60-237): read, global bbox, periodic correction,
  octree, leaf deposit, one ``node_<level>_<ix>_<iy>_<iz>`` density file per leaf.

Errors raise ``AdRuntimeException`` with the reference's message where it throws one.  The
command-line program is ``adeco_amd/adeco_resample PARAMETER_FILE`` (adeco_amd/cpp/adeco_resample.cpp).
"""

from __future__ import annotations

import ctypes
import dataclasses
import os

import numpy as np

from . import (DECOMP_REF_EFFECTIVE, MODE_BATCHED, PREC_FP64, AdRuntimeException, ADECO_EINVAL, _check, _lib)
from ._lib import load, ptr


@dataclasses.dataclass
class AMD_Parameters:
    """AdaptiveMassDeposit::AMD_Parameters with the defaults its constructor and the parser set."""

    input_file: str = ""
    output_directory: str = ""
    id_type: str = ""
    use_regions_of_interest: bool = False
    region_of_interest: tuple = (0.0, 1.0, 0.0, 1.0, 0.0, 1.0)  # x0, x1, y0, y1, z0, z1
    max_refinement_level: int = 5
    stride: int = 1
    max_tets_per_octree_node: int = 2 ** 64 - 1
    linear_patch_resolution: int = 64
    num_pthreads: int = 1
    num_mpi_ranks_per_node: int = 1
    restart_path: str = ""
    periodic_boundaries: bool = True

    @classmethod
    def _from_c(cls, c: _lib.Parameters) -> "AMD_Parameters":
        dec = lambda b: b.decode(errors="surrogateescape")  # noqa: E731
        return cls(input_file=dec(c.input_file), output_directory=dec(c.output_directory), id_type=dec(c.id_type),
                   use_regions_of_interest=bool(c.use_regions_of_interest),
                   region_of_interest=tuple(float(v) for v in c.region_of_interest),
                   max_refinement_level=c.max_refinement_level, stride=c.stride,
                   max_tets_per_octree_node=int(c.max_tets_per_octree_node),
                   linear_patch_resolution=c.linear_patch_resolution, num_pthreads=c.num_pthreads,
                   num_mpi_ranks_per_node=c.num_mpi_ranks_per_node, restart_path=dec(c.restart_path),
                   periodic_boundaries=bool(c.periodic_boundaries))

    def _to_c(self) -> _lib.Parameters:
        c = _lib.Parameters()
        load().adeco_parameters_init(ctypes.byref(c))
        enc = lambda s: os.fsencode(s)  # noqa: E731
        for name, cap in (("input_file", 4096), ("output_directory", 4096), ("id_type", 256), ("restart_path", 4096)):
            b = enc(getattr(self, name))
            if len(b) >= cap:
                raise AdRuntimeException(ADECO_EINVAL, f"ERROR: {name} is too long")
            setattr(c, name, b)
        c.use_regions_of_interest = int(self.use_regions_of_interest)
        c.max_refinement_level = int(self.max_refinement_level)
        c.stride = int(self.stride)
        c.max_tets_per_octree_node = int(self.max_tets_per_octree_node)
        c.linear_patch_resolution = int(self.linear_patch_resolution)
        c.num_pthreads = int(self.num_pthreads)
        c.num_mpi_ranks_per_node = int(self.num_mpi_ranks_per_node)
        c.periodic_boundaries = int(self.periodic_boundaries)
        for i, v in enumerate(self.region_of_interest):
            c.region_of_interest[i] = float(v)
        return c


def parse_parameter_file(filename: str) -> AMD_Parameters:
    """parse_parameter_file (main.cpp:83-183)."""
    c = _lib.Parameters()
    _check(load().adeco_parse_parameter_file(os.fsencode(filename), ctypes.byref(c)), "parse_parameter_file()")
    return AMD_Parameters._from_c(c)


class AdSortedByIdsPointDataReader:
    """SORTED_BY_ID particles (AdPointDataReaders.h:770-911): raw float32 x,y,z per particle in id
    order, a cubic lattice of n^3 particles."""

    def load_data(self, filename: str, stride: int = 1):
        """Returns (x, y, z, dims): float32 SoA coordinates in ColumnMajorOrder and (n, n, n)."""
        lib = load()
        n = ctypes.c_int64()
        dims = np.zeros(3, np.int32)
        _check(lib.adeco_read_sorted_by_id(os.fsencode(filename), int(stride), None, None, None, 0, ctypes.byref(n),
                                           ptr(dims)), "load_data()")
        x, y, z = (np.empty(n.value, np.float32) for _ in range(3))
        _check(lib.adeco_read_sorted_by_id(os.fsencode(filename), int(stride), ptr(x), ptr(y), ptr(z), n.value,
                                           ctypes.byref(n), ptr(dims)), "load_data()")
        return x, y, z, tuple(int(d) for d in dims)


class AdDarkSkyDataReader:
    """DARK_SKY particles (AdPointDataReaders.h:361-767): a text header, then 32-byte records
    {float x, y, z, vx, vy, vz; int64 id} whose Morton ids place them on a cubic lattice.  The
    records are decoded and placed on the GPU."""

    def info(self, filename: str) -> dict:
        """Header bytes, particle records and particles per axis (host-only checks)."""
        h, n = ctypes.c_int64(), ctypes.c_int64()
        d = np.zeros(1, np.int32)
        _check(load().adeco_dark_sky_info(os.fsencode(filename), ctypes.byref(h), ctypes.byref(n), ptr(d)),
               "dark_sky_info()")
        return {"header_bytes": h.value, "num_particles": n.value, "dims": int(d[0])}

    def load_lattice(self, filename: str, stride: int = 1, device: int = 0):
        """The particles as a ResidentLattice on the GPU."""
        from . import ResidentLattice
        return ResidentLattice.from_dark_sky(filename, stride, device)

    def load_data(self, filename: str, stride: int = 1, device: int = 0):
        """Returns (x, y, z, dims) as AdSortedByIdsPointDataReader.load_data does."""
        with self.load_lattice(filename, stride, device) as rl:
            x, y, z = rl.positions()
            return x, y, z, rl.dims


def get_reader(format: str):
    """get_reader (AdPointDataReaders.h:914-930): SORTED_BY_ID, then DARK_SKY, by substring."""
    if "SORTED_BY_ID" in format:
        return AdSortedByIdsPointDataReader()
    if "DARK_SKY" in format:
        return AdDarkSkyDataReader()
    raise AdRuntimeException(ADECO_EINVAL, f"ERROR: id_type {format} not supported.")


def deposit_workflow(parameters: AMD_Parameters, device: int = 0, mode: int = MODE_BATCHED,
                     precision: int = PREC_FP64, decomposition: int = DECOMP_REF_EFFECTIVE,
                     tet_mass: float = 1.0) -> dict:
    """deposit_workflow (AdDepositWorkFlow.h:60-237) on one rank and one GPU.  Writes the leaf files
    into parameters.output_directory and returns the run's counts, total mass and step times."""
    opt = _lib.WorkflowOptions()
    load().adeco_workflow_options_init(ctypes.byref(opt))
    opt.device, opt.mode, opt.precision = int(device), int(mode), int(precision)
    opt.decomposition, opt.tet_mass = int(decomposition), float(tet_mass)
    res = _lib.WorkflowResult()
    c = parameters._to_c()
    _check(load().adeco_deposit_workflow(ctypes.byref(c), ctypes.byref(opt), ctypes.byref(res)), "deposit_workflow()")
    return {"num_particles": res.num_particles, "particle_dims": tuple(res.particle_dims),
            "global_bbox": np.array(res.global_bbox[:], np.float32), "num_nodes": res.num_nodes,
            "num_leaves": res.num_leaves, "total_mass": res.total_mass, "load_s": res.load_s,
            "octree_s": res.octree_s, "deposit_s": res.deposit_s, "write_s": res.write_s, "kernel_ms": res.kernel_ms}
