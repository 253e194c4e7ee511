"""Octree extraction helpers for the oracle and the compiled reference (checker side)."""

from __future__ import annotations

import numpy as np

from tests import _oracle as O


def oracle_octree(lat, max_tets, max_level, roi=None, root_box=None):
    """The oracle's octree; root_box (6 floats) roots it at a given global bbox instead of the
    vertices' own (the periodic workflow keeps the bbox from before the correction)."""
    o = O.load()
    npd = np.array(lat.dims, np.int32)
    roi = None if roi is None else np.ascontiguousarray(roi, np.float32)
    box = None if root_box is None else np.ascontiguousarray(root_box, np.float32)
    h = o.oracle_octree_build_in(O.P(lat.x), O.P(lat.y), O.P(lat.z), O.P(npd), O.P(box), O.P(roi), max_tets,
                                 max_level)
    try:
        n = o.oracle_octree_num_nodes(h)
        nl = o.oracle_octree_num_leaves(h)
        m = o.oracle_octree_num_meta_cells(h)
        d = {"ll": np.zeros((n, 3), np.float32), "ext": np.zeros(n, np.float32), "level": np.zeros(n, np.int32),
             "parent": np.zeros(n, np.int32), "first_child": np.zeros(n, np.int32), "count": np.zeros(n, np.uint64),
             "local_count": np.zeros(n, np.uint64), "bbox": np.zeros(6, np.float32),
             "mc_bbox": np.zeros((m, 6), np.float32)}
        o.oracle_octree_nodes(h, O.P(d["ll"]), O.P(d["ext"]), O.P(d["level"]), O.P(d["parent"]),
                              O.P(d["first_child"]), O.P(d["count"]), O.P(d["local_count"]))
        o.oracle_octree_bbox(h, O.P(d["bbox"]))
        o.oracle_meta_cell_bboxes(h, O.P(d["mc_bbox"]))
        total = o.oracle_octree_leaves(h, None, None, None, None)
        d["dfs"] = np.zeros(nl, np.int32)
        d["order"] = np.zeros(nl, np.int32)
        d["mc_offsets"] = np.zeros(nl + 1, np.int64)
        ids = np.zeros(max(total, 1), np.int64)
        o.oracle_octree_leaves(h, O.P(d["dfs"]), O.P(d["order"]), O.P(d["mc_offsets"]), O.P(ids))
        d["mc_ids"] = ids[:total].astype(np.int32)
    finally:
        o.oracle_octree_free(h)
    return d


def ref_octree(lat, max_tets, max_level, roi=None):
    r = O.load_ref()
    roi = None if roi is None else np.ascontiguousarray(roi, np.float32)
    assert r.ref_octree_build(O.P(lat.x), O.P(lat.y), O.P(lat.z), *lat.dims, O.P(roi), max_tets, max_level) == 0, \
        r.ref_last_error()
    n = r.ref_octree_num_nodes()
    nl = r.ref_octree_num_leaves()
    m = r.ref_octree_num_meta_cells()
    d = {"ll": np.zeros((n, 3), np.float32), "ext": np.zeros(n, np.float32), "level": np.zeros(n, np.int32),
         "parent": np.zeros(n, np.int32), "first_child": np.zeros(n, np.int32), "count": np.zeros(n, np.uint64),
         "local_count": np.zeros(n, np.uint64), "bbox": np.zeros(6, np.float32), "mc_bbox": np.zeros((m, 6), np.float32)}
    r.ref_octree_nodes(O.P(d["ll"]), O.P(d["ext"]), O.P(d["level"]), O.P(d["parent"]), O.P(d["first_child"]),
                       O.P(d["count"]), O.P(d["local_count"]))
    r.ref_octree_bbox(O.P(d["bbox"]))
    r.ref_meta_cell_bboxes(O.P(d["mc_bbox"]))
    total = r.ref_octree_leaves(None, None, None, None)
    d["dfs"] = np.zeros(nl, np.int32)
    d["order"] = np.zeros(nl, np.int32)
    d["mc_offsets"] = np.zeros(nl + 1, np.int64)
    ids = np.zeros(max(total, 1), np.int64)
    r.ref_octree_leaves(O.P(d["dfs"]), O.P(d["order"]), O.P(d["mc_offsets"]), O.P(ids))
    d["mc_ids"] = ids[:total].astype(np.int32)
    return d


KEYS = ("ll", "ext", "level", "parent", "first_child", "count", "bbox", "mc_bbox", "dfs", "order", "mc_offsets", "mc_ids")


def assert_same_octree(a, b, keys=KEYS):
    for k in keys:
        x, y = np.asarray(a[k]), np.asarray(b[k])
        assert x.shape == y.shape, (k, x.shape, y.shape)
        if x.dtype.kind == "f":
            assert np.array_equal(x.view(np.uint32), y.astype(x.dtype).view(np.uint32)), k
        else:
            assert np.array_equal(x, y.astype(x.dtype)), k
