"""The oracle's chain for a whole program run (checker side): SORTED_BY_ID snapshot, global
bbox, periodic correction (oracle_periodic_unwrap), octree rooted at the global bbox, the leaf
deposit of each leaf's meta-patches in the reference's gather order, and write_patch's
densities.  With the float-faithful deposit it rebuilds the reference program's leaf files bit for
bit (tests/test_program.py); with the FP64 deposit it is the checker of the GPU program.
"""

from __future__ import annotations

import numpy as np

from adeco_amd.lattice import Lattice, meta_patches
from tests import _oracle as O
from tests._golden_leaves import patch_density, patch_file_name
from tests._octree_io import oracle_octree
from tests._program_cases import PROGRAM_CASES, case_lattice


def unwrapped(x, y, z, dims, bbox, period=None):
    x, y, z = x.copy(), y.copy(), z.copy()
    npd = np.array(dims, np.int32)
    per = None if period is None else np.asarray(period, np.float32)
    O.load().oracle_periodic_unwrap(O.P(x), O.P(y), O.P(z), O.P(npd), O.P(np.asarray(bbox, np.float32)), O.P(per))
    return x, y, z


def oracle_program(case_name, fp64=False):
    """The oracle chain for a program case: reader, global bbox, periodic correction, octree
    rooted at the global bbox, the float-faithful leaf deposit in the reference's gather order
    and write_patch.  Returns ({file name: float32 densities}, total mass in processing order,
    the corrected lattice, the octree).  fp64: the FP64 deposit in the reference's geometry (float
    leaf boxes, 1/dx scaling) and FP64 densities instead."""
    case = PROGRAM_CASES[case_name]
    lat = case_lattice(case)
    p = case["params"]
    gb = lat.bbox()
    if p["periodic_boundaries"] == "true":
        x, y, z = unwrapped(lat.x, lat.y, lat.z, lat.dims, gb)
        lat = Lattice(n=lat.n, box=lat.box, x=x, y=y, z=z)
    roi = np.array([float(v) for v in p["region_of_interest"].split()], np.float32) \
        if p["use_region_of_interest"] == "true" else gb
    d = oracle_octree(lat, int(p["max_tets_per_octree_node"]), int(p["max_refinement_level"]), roi=roi, root_box=gb)
    dims = int(p["linear_patch_resolution"])
    files, masses = {}, {}
    for leaf, node in enumerate(d["dfs"]):
        ids = d["mc_ids"][d["mc_offsets"][leaf]:d["mc_offsets"][leaf + 1]]
        ll, ext = d["ll"][node], np.float32(d["ext"][node])
        bbox = np.array([ll[0], ll[0] + ext, ll[1], ll[1] + ext, ll[2], ll[2] + ext], np.float32)
        if fp64:
            g = O.deposit_patches_f64(meta_patches(lat, ids), bbox, dims) if ids.size else np.zeros(dims ** 3)
            dens, masses[int(node)] = patch_density(d, node, dims, g, exact=True)
        else:
            g = O.deposit_patches_f32(meta_patches(lat, ids), bbox, dims) if ids.size else np.zeros(dims ** 3, np.float32)
            dens, masses[int(node)] = patch_density(d, node, dims, g)
        files[patch_file_name(d, node, dims)] = dens
    total = 0.0
    for node in d["order"]:
        total += masses[int(node)]
    return files, total, lat, d
