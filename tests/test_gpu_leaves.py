"""GPU leaf deposit (deposit_leaf_nodes on one rank) against the oracle (the reference algorithm
evaluated with a 64-bit significand, rounded to double).

For a refined octree, every leaf grid of the plain per-leaf launches, of the batched persistent
launch in FP64 and of the mixed-precision form is compared with the oracle's deposit of the same
leaf, per cell: |gpu - oracle| <= 1e-6 * max(|oracle|, 1e-8 * tet_mass).

Exact geometry (the default): leaves are the dyadic FP64 subdivision of the root and cells are
scaled per axis, so the leaves partition the root exactly and the total over all leaves equals the
oracle's mass of the lattice inside the root box to 1e-9.  Reference geometry: the reference's
float leaf boxes and 1/dx scaling, compared with the oracle in the same mode.
"""

import numpy as np
import pytest

import adeco_amd as A
from adeco_amd.lattice import make_lattice
from tests import _oracle as O

pytestmark = pytest.mark.gpu

REL = 1e-6
TET_MASS = 1.0 / 6.0


def oracle_leaves(lat, tree, decomp, dims, leaves, exact=True):
    boxes = tree.leaf_boxes_f64(exact=exact)
    out = []
    for l in leaves:
        ids = tree.leaf_tets(int(l))
        box = boxes[l] if exact else boxes[l].astype(np.float32)
        out.append(O.deposit_lattice(lat, box, dims, decomp=decomp, tet_mass=TET_MASS, tet_ids=ids, precision=3))
    return np.array(out)


@pytest.mark.parametrize("decomp", [A.DECOMP_KUHN, A.DECOMP_REF_EFFECTIVE])
def test_leaf_deposit_forms_match_oracle(decomp):
    lat = make_lattice(24, 3.0, seed=4)
    dims = 16
    print(f"[leaves] decomp={decomp}", flush=True)
    with O.cell_scaling(O.EXACT), A.ResidentLattice.from_lattice(lat) as rl, A.Octree(rl, 1500, 4) as tree:
        nl = tree.num_leaves
        assert nl > 8
        counts = tree.bin(decomp)
        pick = np.unique(np.concatenate([np.argsort(-counts)[:4], np.linspace(0, nl - 1, 6).astype(int)]))
        want = oracle_leaves(lat, tree, decomp, dims, pick)
        inside = O.deposit_lattice(lat, tree.root_box_f64(), 1, decomp=decomp, tet_mass=TET_MASS).sum()
        for name, mode, prec in (("plain", A.MODE_PLAIN, A.PREC_FP64), ("batched", A.MODE_BATCHED, A.PREC_FP64),
                                 ("mixed", A.MODE_BATCHED, A.PREC_MIXED)):
            tree.deposit(dims, decomposition=decomp, tet_mass=TET_MASS, mode=mode, precision=prec)
            g = tree.leaf_grids()
            err = O.max_rel_err(g[pick], want, tet_mass=TET_MASS)
            drift = abs(g.sum() - inside) / inside
            print(f"  {name}: max rel err {err:.2e}, mass drift {drift:.2e}", flush=True)
            assert err <= REL, name
            assert drift <= 1e-9, name
        tl = A.Octree.last_timings()
        assert tl["partials"] > 0 and 0 <= tl["fallbacks"] <= tl["partials"]


def test_reference_geometry_mode_matches_oracle():
    lat = make_lattice(16, 2.0, seed=9)
    dims = 8
    with O.cell_scaling(O.REFERENCE), A.ResidentLattice.from_lattice(lat) as rl, A.Octree(rl, 2000, 3) as tree:
        tree.bin(A.DECOMP_KUHN)
        pick = np.arange(min(tree.num_leaves, 12))
        want = oracle_leaves(lat, tree, A.DECOMP_KUHN, dims, pick, exact=False)
        for mode in (A.MODE_PLAIN, A.MODE_BATCHED):
            tree.deposit(dims, decomposition=A.DECOMP_KUHN, tet_mass=TET_MASS, mode=mode)
            assert O.max_rel_err(tree.leaf_grids()[pick], want, tet_mass=TET_MASS) <= REL


def test_batched_is_repeatable_and_dims_change():
    lat = make_lattice(20, 2.5, seed=12)
    with A.ResidentLattice.from_lattice(lat) as rl, A.Octree(rl, 1200, 4) as tree:
        tree.deposit(8, decomposition=A.DECOMP_KUHN, tet_mass=TET_MASS, mode=A.MODE_BATCHED)
        a = tree.leaf_grids()
        tree.deposit(12, decomposition=A.DECOMP_KUHN, tet_mass=TET_MASS, mode=A.MODE_BATCHED)
        b = tree.leaf_grids()
        tree.deposit(8, decomposition=A.DECOMP_KUHN, tet_mass=TET_MASS, mode=A.MODE_BATCHED)
        c = tree.leaf_grids()
    assert O.max_rel_err(c, a, tet_mass=TET_MASS) <= 1e-12
    assert abs(a.sum() - b.sum()) <= 1e-9 * a.sum()


def test_graph_replays_match_oracle():
    """The captured graph (grid clear, work-queue reset, persistent launch) replayed several times,
    with eager deposits of other sizes in between: every replay matches the oracle per leaf to 1e-6
    (floor 1e-8 of a tet mass) and the eager batched result to 1e-12."""
    lat = make_lattice(20, 2.5, seed=12)
    with O.cell_scaling(O.EXACT), A.ResidentLattice.from_lattice(lat) as rl, A.Octree(rl, 1200, 4) as tree:
        tree.bin(A.DECOMP_KUHN)
        want = oracle_leaves(lat, tree, A.DECOMP_KUHN, 12, range(tree.num_leaves))
        for prec in (A.PREC_FP64, A.PREC_MIXED):
            tree.deposit(12, decomposition=A.DECOMP_KUHN, tet_mass=TET_MASS, mode=A.MODE_BATCHED, precision=prec)
            eager = tree.leaf_grids()
            with A.DepositGraph(tree, 12, A.DECOMP_KUHN, TET_MASS, prec) as g:
                for rep in range(3):
                    print(f"[graph] precision={prec} replay {rep}", flush=True)
                    g.launch()
                    got = g.leaf_grids()
                    assert O.max_rel_err(got, want, tet_mass=TET_MASS) <= REL
                    assert O.max_rel_err(got, eager, tet_mass=TET_MASS) <= 1e-12
                    assert A.Octree.last_timings()["partials"] > 0
                    tree.deposit(8 + 2 * rep, decomposition=A.DECOMP_KUHN, tet_mass=TET_MASS, mode=A.MODE_BATCHED)


def _moved(lat, scale=1.0, shift=0.0, noise=0.0, seed=7):
    """The lattice's vertices scaled about their mean, shifted and jittered (a later snapshot)."""
    from adeco_amd.lattice import Lattice
    rng = np.random.default_rng(seed)
    out = []
    for v in (lat.x, lat.y, lat.z):
        c = float(v.mean())
        out.append(((v - c) * scale + c + shift + noise * rng.standard_normal(v.size)).astype(np.float32))
    return Lattice(n=lat.n, box=lat.box, x=out[0], y=out[1], z=out[2])


def _oracle_all_leaves(lat, boxes, dims):
    """The oracle's deposit of every lattice tet onto every leaf grid (tets outside a leaf add
    nothing, so no binning is assumed)."""
    return np.array([O.deposit_lattice(lat, b, dims, decomp=A.DECOMP_KUHN, tet_mass=TET_MASS, precision=3, nthreads=8)
                     for b in boxes])


def test_step_graph_follows_the_lattice():
    """The whole step as one graph (binning of the current positions, pair setup, scans, tile map,
    deposit, all sizes on the device), created on one lattice state and replayed on others after
    ResidentLattice.update: every replay matches the oracle's deposit of that state per leaf to 1e-6
    (floor 1e-8 of a tet mass).  The states: the capture's, shifted by 0.3 Mpc/h, jittered by 1 Mpc/h
    (3 % more (leaf, tet) pairs and 9 % more tiles than the capture, within the graph's 1.25x), and
    the capture's again (equal to the first replay to 1e-12).  A state jittered by 2 Mpc/h needs 12 %
    more pairs and 30 % more tiles: a graph with no slack reports ADECO_EOVERFLOW on it and deposits
    nothing, and a later replay on the capture's state works."""
    lat = make_lattice(20, 2.5, seed=12)
    dims = 8
    states = [("capture", lat), ("shifted", _moved(lat, shift=0.3)), ("jittered", _moved(lat, noise=1.0)),
              ("capture again", lat)]
    with O.cell_scaling(O.EXACT), A.ResidentLattice.from_lattice(lat) as rl, A.Octree(rl, 1200, 4) as tree:
        boxes = tree.leaf_boxes_f64(exact=True)
        with A.DepositGraph(tree, dims, A.DECOMP_KUHN, TET_MASS, A.PREC_FP64, step=True, slack=1.25) as g:
            first = None
            for name, st in states:
                print(f"[step graph] {name}", flush=True)
                rl.update(st.x, st.y, st.z)
                g.launch()
                got = g.leaf_grids()
                want = _oracle_all_leaves(st, boxes, dims)
                assert O.max_rel_err(got, want, tet_mass=TET_MASS) <= REL, name
                if first is None:
                    first = got
                elif name == "capture again":
                    assert O.max_rel_err(got, first, tet_mass=TET_MASS) <= 1e-12
        rl.update(lat.x, lat.y, lat.z)
        with A.DepositGraph(tree, dims, A.DECOMP_KUHN, TET_MASS, A.PREC_FP64, step=True, slack=1.0) as g:
            jit = _moved(lat, noise=2.0)
            rl.update(jit.x, jit.y, jit.z)
            with pytest.raises(A.AdRuntimeException) as ei:
                g.launch()
            assert ei.value.status == A.ADECO_EOVERFLOW
            assert not g.leaf_grids().any()
            rl.update(lat.x, lat.y, lat.z)
            g.launch()
            assert O.max_rel_err(g.leaf_grids(), first, tet_mass=TET_MASS) <= 1e-12


def test_autotune_keeps_results():
    lat = make_lattice(16, 2.0, seed=6)
    with A.ResidentLattice.from_lattice(lat) as rl, A.Octree(rl, 1500, 3) as tree:
        tree.deposit(12, decomposition=A.DECOMP_KUHN, tet_mass=TET_MASS, mode=A.MODE_BATCHED)
        ref = tree.leaf_grids()
        saved = A.get_launch_shape()
        try:
            pick = tree.autotune(12, A.DECOMP_KUHN, TET_MASS, A.PREC_FP64, reps=1)
            assert pick["chunk"] in (1, 2, 4, 8, 16, 32) and pick["blocks_per_cu"] in (1, 2, 3, 4)
            assert A.get_launch_shape() == (pick["chunk"], pick["blocks_per_cu"])
            tree.deposit(12, decomposition=A.DECOMP_KUHN, tet_mass=TET_MASS, mode=A.MODE_BATCHED)
            assert O.max_rel_err(tree.leaf_grids(), ref, tet_mass=TET_MASS) <= 1e-12
        finally:
            A.set_launch_shape(*saved)


def test_c1_single_grid_matches_oracle():
    """The CPU-runnable configuration: 32^3 lattice at rms 1 spacing, 6 Kuhn tets per cube, tet
    mass 1/6, one 64^3 grid over the root (an octree that does not refine)."""
    lat = make_lattice(32, 1.0, seed=1)
    with O.cell_scaling(O.EXACT), A.ResidentLattice.from_lattice(lat) as rl, A.Octree(rl, 10 ** 9, 6) as tree:
        assert tree.num_leaves == 1
        want = O.deposit_lattice(lat, tree.root_box_f64(), 64, decomp=A.DECOMP_KUHN, tet_mass=TET_MASS, precision=3,
                                 nthreads=8)
        for mode in (A.MODE_PLAIN, A.MODE_BATCHED):
            tree.deposit(64, decomposition=A.DECOMP_KUHN, tet_mass=TET_MASS, mode=mode)
            got = tree.leaf_grids()[0]
            assert O.max_rel_err(got, want, tet_mass=TET_MASS) <= REL
            assert abs(got.sum() - 32 ** 3) <= 1e-9 * 32 ** 3


def test_leaf_grids_match_reference_leaf_grids():
    """The 128^3 adaptive workload in the reference's own geometry (float leaf boxes, 1/dx scaling,
    its 6-tet decomposition, tet mass 1): for 12 leaves whose float grids the compiled reference
    produced (tests/golden/ref_leaves_c2.npz, rebuilt bit for bit by the float-faithful oracle),
    the GPU leaf grids - binned per tet, deposited plain and batched - match the FP64 oracle's
    deposit of the leaf's meta-patches to 1e-6 per cell, and the reference's float grids to its
    own float rounding (the reference differs from the FP64 oracle on these leaves by up to 1.5e-3
    relative per cell, with a floor of 1e-3 of a tet mass, and 1.2e-6 on leaf totals; bars 2e-3
    and 2e-6).  The leaf files written from the batched grids are write_patch's files of the
    float-rounded grids, bit for bit, and match the FP64 oracle's densities to 1e-6 per cell."""
    import os
    import tempfile
    from adeco_amd.lattice import meta_patches
    from tests._golden_leaves import patch_density, patch_file_name, reference_leaf_grids
    path = os.path.join(os.path.dirname(__file__), "golden", "ref_leaves_c2.npz")
    g = np.load(path, allow_pickle=False)
    dims = int(g["dims"])
    lat, d, ref = reference_leaf_grids(g)
    f64 = []
    with O.cell_scaling(O.REFERENCE):
        for node, bbox in zip(g["node"], g["bbox"]):
            leaf = int(np.flatnonzero(d["dfs"] == node)[0])
            ids = d["mc_ids"][d["mc_offsets"][leaf]:d["mc_offsets"][leaf + 1]]
            f64.append(O.deposit_patches_f64(meta_patches(lat, ids), bbox, dims))
        with A.ResidentLattice.from_lattice(lat) as rl, A.Octree(rl, 50000, 6) as tree:
            tree.bin(A.DECOMP_REF_EFFECTIVE)
            dfs = tree.leaves()["dfs"]
            boxes = tree.leaf_boxes_f64(exact=False)
            leaves = [int(np.flatnonzero(dfs == node)[0]) for node in g["node"]]
            for l, bbox in zip(leaves, g["bbox"]):
                assert np.array_equal(boxes[l].astype(np.float32), bbox)
            for name, mode in (("plain", A.MODE_PLAIN), ("batched", A.MODE_BATCHED)):
                print(f"[reference leaves] {name}", flush=True)
                tree.deposit(dims, decomposition=A.DECOMP_REF_EFFECTIVE, tet_mass=1.0, mode=mode)
                grids = tree.leaf_grids()
                for i, l in enumerate(leaves):
                    ours = grids[l]
                    assert O.max_rel_err(ours, f64[i], tet_mass=1.0) <= REL, (name, int(g["node"][i]))
                    r = ref[i].astype(np.float64)
                    assert abs(ours.sum() - r.sum()) <= 2e-6 * r.sum(), (name, int(g["node"][i]))
                    assert np.max(np.abs(ours - r) / np.maximum(r, 1e-3)) <= 2e-3, (name, int(g["node"][i]))
            # leaf files of the batched deposit (AdOctreeGridWriter::write_patch): every leaf's file
            # is the float rounding of its FP64 grid in the reference's file format, and on the golden
            # leaves it matches the FP64 oracle's densities to 1e-6 per cell
            nd = tree.nodes()
            with tempfile.TemporaryDirectory() as tmp:
                masses, total = tree.write_leaves(tmp)
                assert len(os.listdir(tmp)) == tree.num_leaves
                order = tree.leaves()["order"]
                want_total = 0.0
                for node in order:
                    want_total += masses[int(np.flatnonzero(dfs == node)[0])]
                assert total == want_total
                for l in list(range(0, tree.num_leaves, 37)) + leaves:
                    node = int(dfs[l])
                    name = patch_file_name(nd, node, dims)
                    got = np.fromfile(os.path.join(tmp, name), np.float32)
                    dens, mass = patch_density(nd, node, dims, grids[l].astype(np.float32))
                    assert np.array_equal(got, dens), name
                    assert masses[l] == mass, name
                for i, l in enumerate(leaves):
                    node = int(dfs[l])
                    got = np.fromfile(os.path.join(tmp, patch_file_name(nd, node, dims)), np.float32).astype(np.float64)
                    want, _ = patch_density(nd, node, dims, f64[i], exact=True)
                    assert O.max_rel_err(got, want, tet_mass=float(want.max())) <= REL, node
