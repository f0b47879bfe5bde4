"""C ABI surface and host-side builders (no GPU needed)."""
import os
import re

import numpy as np
import pytest

import oracle
from encore_amd import rq, scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    inc = os.path.join(ROOT, "include")
    for f in os.listdir(inc):
        if f.endswith(".h"):
            text = open(os.path.join(inc, f)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(rq_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_loads_and_exports_every_declared_symbol():
    L = rq.load()
    names = _declared()
    assert len(names) >= 12
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert "gfx950" in rq.version()


def test_invalid_arguments_are_reported():
    with pytest.raises(rq.RQError) as e:
        rq.host_tree(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint32), "kd")
    assert e.value.status == 1
    with pytest.raises(rq.RQError) as e:
        rq.host_tree(np.zeros((3, 3), np.float32), np.array([[0, 1, 5]], np.uint32), "bvh")
    assert e.value.status == 1 and "range" in str(e.value)


def test_counts_beyond_2_30_are_refused_before_any_read():
    """Work counters and sort sizes are 32-bit: a photon map of more than
    2^30 photons is refused with RQ_ERR_CAPACITY before the arrays are read
    (the buffers here hold one photon)."""
    pos = np.zeros(3, np.float32)
    perm = np.zeros(1, np.uint32)
    axis = np.zeros(1, np.uint8)
    L = rq.load()
    st = L.rq_photons_balance_host(pos.ctypes.data, (1 << 30) + 1, 1, perm.ctypes.data, axis.ctypes.data)
    assert st == 5 and b"2^30" in L.rq_last_error()
    assert L.rq_photons_balance_host(pos.ctypes.data, 1, 1, perm.ctypes.data, axis.ctypes.data) == 0


@pytest.mark.parametrize("level", [None, 3, 5])
def test_engine_kd_tree_is_encore_tree(level):
    """The engine's kd builder reproduces ENCORE's tree node for node."""
    if level is None:
        xyz, idx = scenes.cornell()
    else:
        xyz, idx = scenes.displaced_sphere(level=level)
    e = oracle.KdTree(xyz, idx).export()
    t = rq.host_tree(xyz, idx, "kd")
    assert np.array_equal(e["words"], t["words"])
    inner = (e["words"] & 3) != 3
    assert np.array_equal(e["childs"][inner], t["childs"][inner])
    lo = [e["leaf_prims"][e["leaf_off"][l]:e["leaf_off"][l + 1]] for l in e["childs"][~inner]]
    lt = [t["leaf_prims"][c:c + (w >> 2)] for c, w in zip(t["childs"][~inner], t["words"][~inner])]
    assert all(np.array_equal(a, b) for a, b in zip(lo, lt))
    assert np.array_equal(e["root_box"].view(np.uint32), t["root_box"].view(np.uint32))


def test_engine_kd_tree_honours_model_bounds():
    xyz, idx = scenes.cornell()
    bmin = np.array([-60, -50, -50], np.float32)
    bmax = np.array([50, 70, 50], np.float32)
    e = oracle.KdTree(xyz, idx, bmin, bmax).export()
    t = rq.host_tree(xyz, idx, "kd", bounds=(bmin, bmax))
    assert np.array_equal(e["words"], t["words"])


def _children(nodes, ni, width):
    """(ref, box lo, box hi) of each child slot of node ni."""
    out = []
    for c in range(width):
        ref = int(nodes["ref"][ni, c])
        if width == 2:
            blo = np.array([nodes["x"][ni, 2 * c], nodes["y"][ni, 2 * c], nodes["z"][ni, 2 * c]])
            bhi = np.array([nodes["x"][ni, 2 * c + 1], nodes["y"][ni, 2 * c + 1], nodes["z"][ni, 2 * c + 1]])
        else:
            blo = np.array([nodes["lox"][ni, c], nodes["loy"][ni, c], nodes["loz"][ni, c]])
            bhi = np.array([nodes["hix"][ni, c], nodes["hiy"][ni, c], nodes["hiz"][ni, c]])
        out.append((ref, blo, bhi))
    return out


@pytest.mark.parametrize("width", [2, 4])
@pytest.mark.parametrize("leaf", [1, 4, 8])
def test_bvh_structure(leaf, width):
    """Every triangle sits in exactly one leaf of at most `leaf` triangles;
    every child box contains every triangle of its subtree; inner nodes are
    referenced exactly once (a tree, parents before children); BVH4 empty
    slots carry an inverted box that no slab test can enter."""
    xyz, idx = scenes.displaced_sphere(level=4)
    t = rq.host_tree(xyz, idx, "bvh", leaf_size=leaf, width=width)
    assert t["info"]["bvh_width"] == width
    nodes, order = t["nodes"], t["tri_order"]
    n = len(idx)
    assert sorted(order.tolist()) == list(range(n))
    tri = xyz[idx[order]]  # leaf order
    lo, hi = tri.min(1), tri.max(1)
    seen = np.zeros(n, np.int64)
    parents = np.zeros(len(nodes), np.int64)

    def subtree(ni, depth):
        """Leaf-order triangle slots under node ni."""
        assert depth <= t["info"]["max_depth"] + 1
        slots = []
        for ref, blo, bhi in _children(nodes, ni, width):
            if ref == rq.REF_EMPTY:
                assert width == 4 and np.all(blo == np.inf) and np.all(bhi == -np.inf)
                continue
            if ref & 0x80000000:
                first, cnt = ref & 0x07FFFFFF, ((ref >> 27) & 15) + 1
                assert cnt <= leaf
                seen[first:first + cnt] += 1
                sub = list(range(first, first + cnt))
            else:
                assert ni < ref < len(nodes)
                parents[ref] += 1
                sub = subtree(ref, depth + 1)
            sub = np.array(sub)
            assert np.all(lo[sub] >= blo) and np.all(hi[sub] <= bhi)
            slots += sub.tolist()
        return slots

    assert sorted(subtree(0, 1)) == list(range(n))
    assert np.all(seen == 1)
    assert np.all(parents[1:] == 1) and parents[0] == 0
    if width == 4:
        # The collapse fills nodes: few inner nodes keep fewer than 3 children.
        used = (nodes["ref"] != rq.REF_EMPTY).sum(1)
        assert np.mean(used) > 3.0
        assert t["info"]["stack_bound"] >= 3 * t["info"]["max_depth"]


def test_bvh4_single_leaf_scene():
    xyz = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    t = rq.host_tree(xyz, np.array([[0, 1, 2]], np.uint32), "bvh", width=4)
    ch = _children(t["nodes"], 0, 4)
    assert ch[0][0] & 0x80000000 and all(c[0] == rq.REF_EMPTY for c in ch[1:])


REF_SRC = "/root/reference/ENCORE/src"


@pytest.mark.skipif(not os.path.isdir(REF_SRC), reason="ENCORE sources are only present in the dev container")
@pytest.mark.parametrize("src", ["HipAccel.cpp", "HipPhotonMap.cpp"])
def test_drop_in_adapter_compiles_against_encore_headers(src):
    """The drop-in adapters (encore_amd/adapter) build inside ENCORE's tree
    against its own AccelerationStructure.h, IModel.h, TrianglePrim.h,
    Photon.h and PhotonMap.h."""
    import subprocess

    inc = ["-I" + os.path.join(REF_SRC, d) for d in ("math", "scene", "misc", "accelerators", "renderers")]
    cmd = ["g++", "-std=c++11", "-fsyntax-only", "-Wall", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           "-I" + os.path.join(ROOT, "include")] + inc + [os.path.join(ROOT, "encore_amd", "adapter", src)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
