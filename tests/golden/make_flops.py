"""Writes c2_flops.json: the algorithmic FP64 operation count of one step of the 128^3 adaptive
deposit (the roofline's numerator in bench.py).

The count is the reference algorithm's: every floating-point add, subtract, multiply and divide
that the oracle's restatement of deposit_tet_2_grid (AdDepositTets2Grid.h:509-665, with the clip
of AdTetIntersection.h:230-476) executes in its counting instantiation, summed over every leaf of
the reference octree (max depth 6, split above 50,000 estimated tets) and every tet (Kuhn
decomposition) whose float AABB strictly overlaps the leaf, on 16^3-cell leaf grids.

Run from the repository root:  python tests/golden/make_flops.py
"""

import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from adeco_amd.lattice import make_lattice  # noqa: E402
from tests import _oracle as O  # noqa: E402
from tests._octree_io import oracle_octree  # noqa: E402

N, RMS, SEED, MAX_TETS, MAX_LEVEL, DIMS, TET_MASS, KUHN = 128, 3.0, 2, 50000, 6, 16, 1.0 / 6.0, 2


def main():
    t0 = time.time()
    lat = make_lattice(N, RMS, seed=SEED)
    tree = oracle_octree(lat, MAX_TETS, MAX_LEVEL)
    tets = O.lattice_tets(lat, KUHN)
    mn, mx = tets.min(1), tets.max(1)
    flops, pairs = 0, 0
    for node in tree["dfs"]:
        ll, ext = tree["ll"][node], tree["ext"][node]
        box = np.array([ll[0], ll[0] + ext, ll[1], ll[1] + ext, ll[2], ll[2] + ext], np.float32)
        ids = np.nonzero(np.all((mn < box[1::2]) & (box[0::2] < mx), 1))[0].astype(np.int32)
        _, st = O.deposit_lattice(lat, box, DIMS, decomp=KUHN, tet_mass=TET_MASS, tet_ids=ids, precision=2,
                                  nthreads=os.cpu_count() or 1, stats=True)
        flops += int(st[12])
        pairs += int(ids.size)
    out = {"workload": "128^3 lattice, k^-2 field rms 3 spacings seed 2, Kuhn tets, octree depth<=6 split>50000, "
                       "16^3-cell leaves", "leaves": int(len(tree["dfs"])), "pairs": pairs, "flops": flops,
           "flops_per_pair": flops / pairs, "counted_by": "oracle Counted<double> (add/sub/mul/div)"}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c2_flops.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out), f"({time.time() - t0:.0f} s)")


if __name__ == "__main__":
    main()
