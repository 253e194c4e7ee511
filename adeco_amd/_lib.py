"""ctypes loader for libadeco_amd.so (built in-tree by adeco_amd/build.py).

The library must exist: there is no fallback.  Every compute entry point runs HIP kernels and
returns ADECO_EDEVICE when no GPU is usable.
"""

from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libadeco_amd.so")

c_f32p = ctypes.POINTER(ctypes.c_float)
c_f64p = ctypes.POINTER(ctypes.c_double)
c_i32p = ctypes.POINTER(ctypes.c_int32)
c_i64p = ctypes.POINTER(ctypes.c_int64)
c_u64p = ctypes.POINTER(ctypes.c_uint64)
c_vp = ctypes.c_void_p

# name -> (restype, argtypes); mirrors include/adeco.h
SIGNATURES = {
    "adeco_last_error": (ctypes.c_char_p, []),
    "adeco_version": (ctypes.c_int, []),
    "adeco_device_count": (ctypes.c_int, []),
    "adeco_set_cell_scaling": (ctypes.c_int, [ctypes.c_int]),
    "adeco_set_tet_decomposition": (ctypes.c_int, [ctypes.c_int]),
    "adeco_get_tet_decomposition": (ctypes.c_int, []),
    "adeco_set_tet_mass": (ctypes.c_int, [ctypes.c_double]),
    "adeco_get_tet_mass": (ctypes.c_double, []),
    "adeco_set_deposit_mode": (ctypes.c_int, [ctypes.c_int]),
    "adeco_get_deposit_mode": (ctypes.c_int, []),
    "adeco_set_precision_mode": (ctypes.c_int, [ctypes.c_int]),
    "adeco_get_precision_mode": (ctypes.c_int, []),
    "adeco_deposit_leaves_batched": (ctypes.c_int, [c_vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int]),
    "adeco_get_cell_scaling": (ctypes.c_int, []),
    "adeco_deposit_patches": (ctypes.c_int, [ctypes.c_int, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_size_t,
                                             c_vp, ctypes.c_size_t, ctypes.c_int, c_vp, ctypes.c_int, c_vp]),
    "adeco_deposit_patches_noalloc": (ctypes.c_int, [ctypes.c_int, ctypes.c_size_t, ctypes.c_size_t, c_vp,
                                                     ctypes.c_size_t, ctypes.c_int, c_vp, ctypes.c_int, c_vp]),
    "adeco_noalloc_scratch_bytes": (ctypes.c_size_t, []),
    "adeco_get_grid_size": (ctypes.c_int, []),
    "adeco_get_block_size": (ctypes.c_int, []),
    "adeco_set_memory_budget": (ctypes.c_int, [ctypes.c_size_t]),
    "adeco_lattice_create": (ctypes.c_int, [ctypes.c_int, c_vp, c_vp, c_vp, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_int, ctypes.POINTER(c_vp)]),
    "adeco_lattice_destroy": (ctypes.c_int, [c_vp]),
    "adeco_lattice_update": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp]),
    "adeco_lattice_bbox": (ctypes.c_int, [c_vp, c_vp]),
    "adeco_lattice_periodic_unwrap": (ctypes.c_int, [c_vp, c_vp, c_vp]),
    "adeco_lattice_positions": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp]),
    "adeco_lattice_deposit_grid": (ctypes.c_int, [c_vp, ctypes.c_int, ctypes.c_double, c_vp, ctypes.c_int64,
                                                  c_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, c_vp, c_vp]),
    "adeco_deposit_soup": (ctypes.c_int, [ctypes.c_int, c_vp, ctypes.c_int64, c_vp, ctypes.c_int,
                                          ctypes.c_double, ctypes.c_int, ctypes.c_int, c_vp, c_vp]),
    "adeco_last_kernel_ms": (ctypes.c_double, []),
    "adeco_octree_build": (ctypes.c_int, [c_vp, c_vp, ctypes.c_uint64, ctypes.c_int, ctypes.POINTER(c_vp)]),
    "adeco_octree_destroy": (ctypes.c_int, [c_vp]),
    "adeco_octree_num_nodes": (ctypes.c_int64, [c_vp]),
    "adeco_octree_num_leaves": (ctypes.c_int64, [c_vp]),
    "adeco_octree_num_meta_cells": (ctypes.c_int64, [c_vp]),
    "adeco_octree_bbox": (ctypes.c_int, [c_vp, c_vp]),
    "adeco_octree_nodes": (ctypes.c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "adeco_octree_leaves": (ctypes.c_int64, [c_vp, c_vp, c_vp, c_vp, c_vp]),
    "adeco_octree_meta_bboxes": (ctypes.c_int, [c_vp, c_vp]),
    "adeco_octree_bin": (ctypes.c_int, [c_vp, ctypes.c_int, c_vp]),
    "adeco_deposit_leaves": (ctypes.c_int, [c_vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                            ctypes.c_int]),
    "adeco_leaf_grids": (ctypes.c_int, [c_vp, ctypes.c_int64, ctypes.c_int64, c_vp]),
    "adeco_octree_leaf_tets": (ctypes.c_int64, [c_vp, ctypes.c_int64, c_vp]),
    "adeco_octree_leaf_boxes": (ctypes.c_int, [c_vp, ctypes.c_int, c_vp]),
    "adeco_write_patch": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, c_vp, ctypes.c_float, c_vp, ctypes.c_int,
                                         c_vp, c_vp]),
    "adeco_write_leaves": (ctypes.c_int, [c_vp, ctypes.c_char_p, c_vp, c_vp]),
    "adeco_parameters_init": (None, [c_vp]),
    "adeco_parse_parameter_file": (ctypes.c_int, [ctypes.c_char_p, c_vp]),
    "adeco_read_sorted_by_id": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, c_vp, c_vp, c_vp, ctypes.c_int64,
                                               c_i64p, c_vp]),
    "adeco_lattice_create_dark_sky": (ctypes.c_int, [ctypes.c_int, ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(c_vp)]),
    "adeco_dark_sky_info": (ctypes.c_int, [ctypes.c_char_p, c_i64p, c_i64p, c_vp]),
    "adeco_selftest_dark_sky_site": (None, [ctypes.c_longlong, c_vp]),
    "adeco_workflow_options_init": (None, [c_vp]),
    "adeco_deposit_workflow": (ctypes.c_int, [c_vp, c_vp, c_vp]),
    "adeco_last_timings": (None, [c_vp]),
    "adeco_set_launch_shape": (ctypes.c_int, [ctypes.c_int, ctypes.c_int]),
    "adeco_get_launch_shape": (ctypes.c_int, [c_vp, c_vp]),
    "adeco_autotune_batched": (ctypes.c_int, [c_vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                              ctypes.c_int, c_vp, c_vp, c_vp]),
    "adeco_deposit_graph_create": (ctypes.c_int, [c_vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                                  ctypes.POINTER(c_vp)]),
    "adeco_deposit_graph_create_step": (ctypes.c_int, [c_vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                                       ctypes.c_double, ctypes.POINTER(c_vp)]),
    "adeco_deposit_graph_launch": (ctypes.c_int, [c_vp]),
    "adeco_deposit_graph_grids": (ctypes.c_int, [c_vp, ctypes.c_int64, ctypes.c_int64, c_vp]),
    "adeco_deposit_graph_destroy": (ctypes.c_int, [c_vp]),
    "adeco_serve_create": (ctypes.c_int, [c_vp, c_vp, ctypes.POINTER(c_vp)]),
    "adeco_serve_submit": (ctypes.c_int, [c_vp, c_vp, ctypes.c_int, c_vp, c_u64p]),
    "adeco_serve_wait": (ctypes.c_int, [c_vp, ctypes.c_uint64, ctypes.c_double, c_vp]),
    "adeco_serve_stats_get": (ctypes.c_int, [c_vp, c_vp]),
    "adeco_serve_destroy": (ctypes.c_int, [c_vp]),
    "adeco_roi_deposit_plain": (ctypes.c_int, [c_vp, ctypes.c_int, ctypes.c_double, c_vp, ctypes.c_int, c_vp,
                                               c_i64p]),
    "adeco_selftest_clip_vol6_f64": (ctypes.c_double, [c_vp]),
    "adeco_selftest_clip_vol6_f32": (ctypes.c_float, [c_vp]),
    "adeco_selftest_classify_f64": (ctypes.c_int, [c_vp]),
    "adeco_selftest_beyond_fraction_f64": (ctypes.c_double, [c_vp]),
    "adeco_selftest_column_f64": (None, [c_vp, ctypes.c_int, ctypes.c_int, c_vp, c_vp]),
    "adeco_selftest_decomposition": (None, [ctypes.c_int, c_vp]),
    "adeco_selftest_deposit_soup_f64": (ctypes.c_int, [c_vp, ctypes.c_int64, c_vp, ctypes.c_int, ctypes.c_double,
                                                       c_vp]),
}


class Parameters(ctypes.Structure):
    """adeco_parameters"""
    _fields_ = [("input_file", ctypes.c_char * 4096), ("output_directory", ctypes.c_char * 4096),
                ("id_type", ctypes.c_char * 256), ("restart_path", ctypes.c_char * 4096),
                ("use_regions_of_interest", ctypes.c_int), ("max_refinement_level", ctypes.c_int),
                ("stride", ctypes.c_int), ("linear_patch_resolution", ctypes.c_int), ("num_pthreads", ctypes.c_int),
                ("num_mpi_ranks_per_node", ctypes.c_int), ("periodic_boundaries", ctypes.c_int), ("pad_", ctypes.c_int),
                ("max_tets_per_octree_node", ctypes.c_uint64), ("region_of_interest", ctypes.c_float * 6)]


class WorkflowOptions(ctypes.Structure):
    """adeco_workflow_options"""
    _fields_ = [("device", ctypes.c_int), ("mode", ctypes.c_int), ("precision", ctypes.c_int),
                ("decomposition", ctypes.c_int), ("tet_mass", ctypes.c_double)]


class WorkflowResult(ctypes.Structure):
    """adeco_workflow_result"""
    _fields_ = [("num_particles", ctypes.c_int64), ("particle_dims", ctypes.c_int * 3),
                ("global_bbox", ctypes.c_float * 6), ("num_nodes", ctypes.c_int64), ("num_leaves", ctypes.c_int64),
                ("total_mass", ctypes.c_double), ("load_s", ctypes.c_double), ("octree_s", ctypes.c_double),
                ("deposit_s", ctypes.c_double), ("write_s", ctypes.c_double), ("kernel_ms", ctypes.c_double)]


class ServeConfig(ctypes.Structure):
    """adeco_serve_config"""
    _fields_ = [("decomposition", ctypes.c_int), ("precision", ctypes.c_int), ("tet_mass", ctypes.c_double),
                ("max_inflight", ctypes.c_int), ("pad_", ctypes.c_int), ("step_tiles", ctypes.c_int64),
                ("memory_budget", ctypes.c_uint64)]


class RequestInfo(ctypes.Structure):
    """adeco_request_info"""
    _fields_ = [("status", ctypes.c_int), ("dims", ctypes.c_int), ("n_tets", ctypes.c_int64),
                ("n_candidates", ctypes.c_int64), ("t_submit", ctypes.c_double), ("t_admit", ctypes.c_double),
                ("t_done", ctypes.c_double)]


class ServeStats(ctypes.Structure):
    """adeco_serve_stats"""
    _fields_ = [("requests_done", ctypes.c_uint64), ("steps", ctypes.c_uint64), ("tiles", ctypes.c_int64),
                ("tets", ctypes.c_int64), ("busy_s", ctypes.c_double), ("device_bytes", ctypes.c_uint64),
                ("pending", ctypes.c_int64), ("active", ctypes.c_int64), ("step_kernel_s", ctypes.c_double)]


_lib = None


def load() -> ctypes.CDLL:
    """Loads the library (once) and declares the C-ABI signatures."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python adeco_amd/build.py` "
                          "(there is no CPU fallback for the deposit)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def ptr(a):
    """Data pointer of a numpy array (or None)."""
    return None if a is None else ctypes.c_void_p(a.ctypes.data)
