"""Hand-worked cases that pin the oracle (and the engine's host builders)
where only the reference's source text fixes the answer: the kd split snap,
the strict distribution with "overlaps neither -> both", the N = 5000/5001
leaf-size switch with the 3-level stall rule, and the photon balance's
bookkeeping-box axis below the root.  The expected values are data in
tests/golden/oracle_pins.json, worked out by hand in make_pins.py from the
reference's source lines cited there; each case is built so that one
plausible misreading of the source gives a different answer."""
import json
import os

import numpy as np
import pytest

import oracle
from encore_amd import rq

PINS = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "oracle_pins.json")))


def _oracle_nodes(kd):
    e = kd.export()
    nodes = []
    for w, c in zip(e["words"], e["childs"]):
        if (w & 3) == 3:
            nodes.append(("leaf", e["leaf_prims"][e["leaf_off"][c]:e["leaf_off"][c + 1]].tolist()))
        else:
            nodes.append(("inner", int(w & 3), int(c)))
    return nodes


def _engine_nodes(xyz, idx):
    t = rq.host_tree(xyz, idx, "kd")
    nodes = []
    for w, c in zip(t["words"], t["childs"]):
        if (w & 3) == 3:
            nodes.append(("leaf", t["leaf_prims"][c:c + (int(w) >> 2)].tolist()))
        else:
            nodes.append(("inner", int(w & 3), int(c)))
    return nodes


def _builders(xyz, idx):
    return {"oracle": _oracle_nodes(oracle.KdTree(xyz, idx)), "engine": _engine_nodes(xyz, idx)}


def _depths(nodes):
    depth = {0: 0}
    for i, n in enumerate(nodes):
        if n[0] == "inner" and i in depth:
            depth[n[2]] = depth[n[2] + 1] = depth[i] + 1
    return depth


def _check_small_tree(nodes, want):
    assert len(nodes) == len(want), nodes
    for got, (kind, val) in zip(nodes, want):
        assert got[0] == kind, (got, kind)
        if kind == "inner":
            assert got[1] == val
        else:
            assert got[1] == val


@pytest.mark.parametrize("case", ["kd_snap", "kd_distribution"])
def test_kd_small_trees(case):
    c = PINS[case]
    xyz, idx = np.array(c["xyz"], np.float32), np.array(c["idx"], np.uint32)
    for name, nodes in _builders(xyz, idx).items():
        _check_small_tree(nodes, c["nodes"])
    # the root's split word is the split with the axis in its low two bits
    e = oracle.KdTree(xyz, idx).export()
    s = np.float32(c["root_split"]).view(np.uint32)
    assert e["words"][0] == (s & ~np.uint32(3)) | np.uint32(0)


def test_kd_split_snap_rays():
    """Kdtree.cpp:711: the ray starting at x = 0.0003 (between the snapped
    split 0 and the stored split ~0.0005) goes to the right child only and
    misses the wall at x = 0.0004 that BasicAS finds."""
    c = PINS["kd_snap"]
    xyz, idx = np.array(c["xyz"], np.float32), np.array(c["idx"], np.uint32)
    rays = np.array(c["rays"], np.float32)
    got = oracle.KdTree(xyz, idx).intersect(rays)
    brute = oracle.brute_closest(xyz, idx, rays)
    assert got[:, 3].view(np.uint32).tolist() == c["kd_prim"]
    assert brute[:, 3].view(np.uint32).tolist() == c["brute_prim"]
    # each hit carries the predicate's own (t, alpha, beta) for its triangle
    tris = np.ascontiguousarray(xyz[idx[c["kd_prim"]]].reshape(-1, 9))
    h = oracle.tri_intersect_pairs(rays, tris)
    assert np.array_equal(got[:, :3].view(np.uint32), h[:, :3].view(np.uint32))


@pytest.mark.parametrize("which", [0, 1], ids=["n5000", "n5001"])
def test_kd_leaf_size_switch_and_stall_rule(which):
    c = PINS["kd_leafsize"]
    case = c["cases"][which]
    g = c["grid"]
    ax = g["start"] + g["step"] * np.arange(g["n_per_axis"])
    pts = np.stack(np.meshgrid(ax, ax, ax, indexing="ij"), -1).reshape(-1, 3)[: case["r1"]]
    pts = np.concatenate([pts, np.tile(c["satellite_at"], (3, 1)), np.tile(c["cluster_at"], (10, 1))])
    tiny = np.array(c["tiny"])
    xyz = (pts[:, None, :] + tiny[None]).reshape(-1, 3).astype(np.float32)
    idx = np.arange(len(xyz), dtype=np.uint32).reshape(-1, 3)
    assert len(idx) == case["n"]
    sats = list(range(case["r1"], case["r1"] + 3))
    cluster = list(range(case["r1"] + 3, case["r1"] + 13))
    assert oracle.KdTree(xyz, idx).stats()["max_item"] == case["max_item"]
    for name, nodes in _builders(xyz, idx).items():
        depth = _depths(nodes)
        leaves = {tuple(n[1]): i for i, n in enumerate(nodes) if n[0] == "leaf" and n[1]}
        assert tuple(cluster) in leaves, name
        assert depth[leaves[tuple(cluster)]] == case["cluster_depth"], name
        assert tuple(sats) in leaves, name
        assert depth[leaves[tuple(sats)]] == case["satellite_depth"], name


def test_photon_axis_below_root():
    c = PINS["photon_axis"]
    pos = np.array(c["pos"], np.float32)
    perm, axis = oracle.PhotonMap(pos, np.tile(np.float32([0, 1, 0]), (len(pos), 1))).order()
    assert perm.tolist() == c["perm"] and axis.tolist() == c["axis"]
    perm, axis = rq.photon_balance_host(pos, "encore")
    assert perm.tolist() == c["perm"] and axis.tolist() == c["axis"]
