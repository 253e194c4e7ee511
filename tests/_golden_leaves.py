"""The reference's own leaf grids of the 128^3 adaptive workload (checker side).

tests/golden/ref_leaves_c2.npz holds, for 12 leaves of the 128^3 lattice (rms 3 spacings, seed 2;
octree depth <= 6, split above 50,000 estimated tets; 16^3-cell leaves), the sha256 of the float
grid that the compiled reference deposits for the leaf (its leaf gather and
DepositWrapper::deposit_tets, AdDynamicLoadBalancing.h:983-1323), the leaf's float box and the
grid total.  The float-faithful oracle rebuilds those grids from each leaf's meta-patches, in the
reference's gather order, bit for bit.
"""

import numpy as np

from adeco_amd.lattice import make_lattice, meta_patches
from tests import _oracle as O
from tests._octree_io import oracle_octree


def reference_leaf_grids(g):
    """Returns (lattice, oracle octree, [float grid per fixture leaf])."""
    lat = make_lattice(128, 3.0, seed=2)
    d = oracle_octree(lat, 50000, 6)
    dims = int(g["dims"])
    out = []
    for node, bbox in zip(g["node"], g["bbox"]):
        leaf = int(np.flatnonzero(d["dfs"] == node)[0])
        ids = d["mc_ids"][d["mc_offsets"][leaf]:d["mc_offsets"][leaf + 1]]
        out.append(O.deposit_patches_f32(meta_patches(lat, ids), bbox, dims))
    return lat, d, out
