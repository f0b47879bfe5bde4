"""ENCORE's GPU texture encoding of the kd tree (Kdtree::buildGPU,
Kdtree.cpp:876-1000), exported from the engine's host tree, against the
oracle's restatement over its own tree, bit for bit, and the encoding's
rules read off the source."""
import numpy as np
import pytest

import oracle
from encore_amd import rq, scenes


def _scenes():
    yield "cornell", scenes.cornell()
    yield "sphere", scenes.displaced_sphere(level=3)
    yield "photon box", scenes.config4_scene(tess=4)


@pytest.mark.parametrize("name,scene", list(_scenes()), ids=lambda v: v if isinstance(v, str) else "")
def test_texture_export_matches_oracle(name, scene):
    xyz, idx = scene
    got = rq.kd_gpu_textures(xyz, idx)
    ref = oracle.KdTree(xyz, idx).gpu_textures()
    for k in ("cell_data0", "vertex0", "vertex1", "vertex2", "ref_prim"):
        assert got[k].shape == ref[k].shape, k
        assert np.array_equal(got[k].view(np.uint32), ref[k].view(np.uint32)), k
    c = got["cell_data0"]
    leaf = c[:, 3] >= 0
    # inner nodes: w = -(axis + 2); the split float carries the axis in its low bits
    assert np.all(np.isin(c[~leaf, 3], [-2.0, -3.0, -4.0]))
    axis = (-c[~leaf, 3] - 2).astype(np.uint32)
    assert np.array_equal(c[~leaf, 1].view(np.uint32) & 3, axis)
    # leaves: x = running count of the triangles of earlier leaves, node order
    counts = c[leaf, 3].astype(np.int64)
    assert np.array_equal(c[leaf, 0].astype(np.int64), np.concatenate([[0], np.cumsum(counts)[:-1]]))
    assert counts.sum() == len(got["ref_prim"])
    # vertex arrays: the referenced triangle's corners, w = 0
    tri = xyz[idx[got["ref_prim"]]]
    for k in range(3):
        assert np.array_equal(got["vertex%d" % k][:, :3], tri[:, k, :])
        assert np.all(got["vertex%d" % k][:, 3] == 0)


def test_texture_export_refuses_bvh():
    xyz, idx = scenes.cornell()
    import ctypes
    L = rq.load()
    d = rq._desc("bvh", "encore", 0, None)
    h = ctypes.c_void_p()
    rq.check(L.rq_host_tree_build(np.ascontiguousarray(xyz, np.float32).ctypes.data, len(xyz),
                                  np.ascontiguousarray(idx, np.uint32).ctypes.data, len(idx), ctypes.byref(d),
                                  ctypes.byref(h)))
    try:
        st = L.rq_host_tree_export_textures(h, None, 0, None, None, None, None, None, None)
        assert st == 4  # RQ_ERR_UNSUPPORTED
    finally:
        L.rq_host_tree_destroy(h)
