"""Kd scenes over a kd-tree built elsewhere (rq_scene_create_kd): the
oracle's restatement of ENCORE's Kdtree, imported as ENCORE's own arrays
would be, answers every query bit for bit as the oracle traverses it; the
engine's host tree imported the same way equals the engine's kd scene;
malformed trees are refused; an imported scene rebuilds on refit."""
import numpy as np
import pytest

import oracle
from encore_amd import rq, scenes
from helpers import log
from test_gpu_refit import _frame
from test_gpu_trace import _trace

pytestmark = pytest.mark.gpu


def _offset_form(e):
    """The oracle's export (a leaf's child = leaf-list id) in the form the
    engine takes: leaves laid out in node order, a leaf's child = offset."""
    words, childs = e["words"].copy(), e["childs"].copy()
    prims = []
    for i in np.nonzero((words & 3) == 3)[0]:
        lid = int(childs[i])
        lst = e["leaf_prims"][e["leaf_off"][lid]:e["leaf_off"][lid + 1]]
        assert len(lst) == words[i] >> 2
        childs[i] = sum(len(p) for p in prims)
        prims.append(lst)
    lp = np.concatenate(prims).astype(np.uint32) if prims else np.zeros(0, np.uint32)
    return words, childs, lp


@pytest.fixture(scope="module")
def sphere():
    xyz, idx = scenes.displaced_sphere(level=5)
    prim = scenes.config2_primary(cols=160, rows=90)
    kd = oracle.KdTree(xyz, idx)
    ph = kd.intersect(prim)
    ok, shadow = scenes.config2_shadow(xyz, idx, prim, ph)
    bounce, _ = scenes.config3_rays(xyz, idx, prim, ph, count=20000, nbatches=2)
    return xyz, idx, kd, np.concatenate([prim, bounce]), shadow


@pytest.mark.parametrize("name", ["cornell", "sphere"])
def test_oracle_tree_imported_is_bit_exact(device, sphere, name):
    if name == "cornell":
        xyz, idx, rays = scenes.config1(res=96)
        kd = oracle.KdTree(xyz, idx)
        shadow = rays[::3].copy()
        shadow[:, 7] = 300.0
    else:
        xyz, idx, kd, rays, shadow = sphere
    words, childs, lp = _offset_form(kd.export())
    box = kd.export()["root_box"]
    sc = rq.Scene.from_kd(device, xyz, idx, words, childs, lp, box)
    st = sc.stats()
    log("case: import oracle tree %s: %d nodes, %d leaves, %d refs" % (name, st["nnodes"], st["nleaves"], st["nrefs"]))
    assert st["accel"] == rq.ACCEL["kd"] and st["nrefs"] == len(lp)
    got = _trace(sc, rays)
    assert np.array_equal(got.view(np.uint32), kd.intersect(rays).view(np.uint32))
    occ = _trace(sc, shadow, "any")
    assert np.array_equal(occ.astype(np.uint32), kd.occluded(shadow))
    sc.destroy()


def test_engine_host_tree_imported_equals_kd_scene(device, sphere):
    xyz, idx, _, rays, _ = sphere
    t = rq.host_tree(xyz, idx, "kd")
    imp = rq.Scene.from_kd(device, xyz, idx, t["words"], t["childs"], t["leaf_prims"], t["root_box"])
    own = rq.Scene(device, xyz, idx, accel="kd")
    assert np.array_equal(_trace(imp, rays).view(np.uint32), _trace(own, rays).view(np.uint32))
    for pred in ("watertight",):
        a = rq.Scene.from_kd(device, xyz, idx, t["words"], t["childs"], t["leaf_prims"], t["root_box"], pred)
        b = rq.Scene(device, xyz, idx, accel="kd", predicate=pred)
        assert np.array_equal(_trace(a, rays).view(np.uint32), _trace(b, rays).view(np.uint32))
        a.destroy()
        b.destroy()
    imp.destroy()
    own.destroy()


def test_malformed_trees_are_refused(device):
    xyz, idx = scenes.cornell()
    t = rq.host_tree(xyz, idx, "kd")
    words, childs, lp, box = t["words"], t["childs"], t["leaf_prims"], t["root_box"]
    inner = np.nonzero((words & 3) != 3)[0]
    leaf = np.nonzero((words & 3) == 3)[0]
    bad = []
    c = childs.copy()
    c[inner[0]] = len(words) - 1  # right child past the end
    bad.append(("child range", words, c, lp))
    c = childs.copy()
    c[inner[-1]] = childs[inner[0]]  # two parents share a pair
    bad.append(("not a tree", words, c, lp))
    c = childs.copy()
    c[leaf[-1]] = len(lp)  # leaf triangles past the list
    bad.append(("leaf range", words, c, lp) if words[leaf[-1]] >> 2 else ("leaf range", words, c, lp[:0]))
    p = lp.copy()
    p[0] = len(idx)
    bad.append(("prim range", words, childs, p))
    for what, w, c, l in bad:
        with pytest.raises(rq.RQError) as e:
            rq.Scene.from_kd(device, xyz, idx, w, c, l, box)
        log("refused (%s): %s" % (what, e.value))
        assert e.value.status == 1


def test_imported_scene_rebuilds_on_refit(device, sphere):
    xyz, idx, kd, rays, _ = sphere
    words, childs, lp = _offset_form(kd.export())
    sc = rq.Scene.from_kd(device, xyz, idx, words, childs, lp, kd.export()["root_box"])
    xf = _frame(xyz, 2)
    sc.refit(xf)
    assert np.array_equal(_trace(sc, rays).view(np.uint32), oracle.KdTree(xf, idx).intersect(rays).view(np.uint32))
    sc.destroy()
