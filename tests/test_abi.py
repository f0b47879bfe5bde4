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


@pytest.mark.parametrize("leaf", [1, 4, 8])
def test_bvh_structure(leaf):
    xyz, idx = scenes.displaced_sphere(level=4)
    t = rq.host_tree(xyz, idx, "bvh", leaf_size=leaf)
    nodes, order = t["nodes"], t["tri_order"]
    n = len(idx)
    assert sorted(order.tolist()) == list(range(n))
    tri = xyz[idx[order]]  # leaf order
    lo, hi = tri.min(1), tri.max(1)
    seen = np.zeros(n, np.int64)
    for ni in range(len(nodes)):
        for c in range(2):
            ref = int(nodes["ref"][ni, c])
            blo = np.array([nodes["x"][ni, 2 * c], nodes["y"][ni, 2 * c], nodes["z"][ni, 2 * c]])
            bhi = np.array([nodes["x"][ni, 2 * c + 1], nodes["y"][ni, 2 * c + 1], nodes["z"][ni, 2 * c + 1]])
            if ref & 0x80000000:
                first, cnt = ref & 0x07FFFFFF, ((ref >> 27) & 15) + 1
                assert cnt <= leaf
                seen[first:first + cnt] += 1
                assert np.all(lo[first:first + cnt] >= blo) and np.all(hi[first:first + cnt] <= bhi)
            else:
                assert ni < ref < len(nodes)
    assert np.all(seen == 1)


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
