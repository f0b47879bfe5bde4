"""GPU-built scenes (builder "lbvh"): the linear BVH answers every query as
the oracle kd-tree does, up to the BVH classes (ties, reference phantoms
outside their triangle's box); degenerate inputs (one triangle, leaf-sized
scenes, thousands of identical triangles whose Morton keys all collide) and
refits of GPU-built scenes."""
import time

import numpy as np
import pytest

import oracle
from encore_amd import rq, scenes
from helpers import brute_occluded, log
from test_gpu_refit import _frame
from test_gpu_trace import _check_vs_oracle, _trace

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sphere():
    xyz, idx = scenes.displaced_sphere(level=5)
    prim = scenes.config2_primary(cols=200, rows=120)
    kd = oracle.KdTree(xyz, idx)
    bounce, _ = scenes.config3_rays(xyz, idx, prim, kd.intersect(prim), count=30000, nbatches=2)
    rays = np.concatenate([prim, bounce])
    return xyz, idx, rays, kd.intersect(rays)


@pytest.mark.parametrize("leaf", [1, 4, 16])
def test_lbvh_closest_matches_oracle(device, sphere, leaf):
    xyz, idx, rays, ref = sphere
    t0 = time.perf_counter()
    sc = rq.Scene(device, xyz, idx, accel="bvh", builder="lbvh", leaf_size=leaf)
    st = sc.stats()
    log("case: lbvh leaf %d: %d tris built in %.1f ms, %d leaves, depth %d" % (
        leaf, len(idx), (time.perf_counter() - t0) * 1e3, st["nleaves"], st["max_depth"]))
    assert st["bvh_width"] == 2 and st["ntris"] == len(idx)
    assert len(idx) / leaf <= st["nleaves"] <= len(idx)
    got = _trace(sc, rays)
    c = _check_vs_oracle(xyz, idx, rays, got, ref)
    log({k: v for k, v in c.items() if k != "bad_idx"})
    assert c["bad"] == 0
    assert c["exact"] >= 0.99 * len(rays)
    occ_rays = rays[::5].copy()
    occ_rays[:, 7] = 60.0
    occ = _trace(sc, occ_rays, "any")
    assert (occ.astype(np.uint32) == brute_occluded(xyz, idx, occ_rays)).mean() > 0.999
    sc.destroy()


def test_lbvh_cornell_and_watertight(device):
    xyz, idx, rays = scenes.config1(res=128)
    ref = oracle.KdTree(xyz, idx).intersect(rays)
    for pred in ("encore", "watertight"):
        sc = rq.Scene(device, xyz, idx, accel="bvh", builder="lbvh", predicate=pred)
        got = _trace(sc, rays)
        c = _check_vs_oracle(xyz, idx, rays, got, ref, bit_exact=(pred == "encore"))
        log("cornell lbvh %s: %s" % (pred, {k: v for k, v in c.items() if k != "bad_idx"}))
        assert c["bad"] == 0
        sc.destroy()


def test_lbvh_degenerate_inputs(device):
    ray = np.array([[0.25, 0.25, 5.0, 0.0, 0.0, 0.0, -1.0, np.inf]], np.float32)
    one = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    sc = rq.Scene(device, one, np.array([[0, 1, 2]], np.uint32), builder="lbvh")
    h = _trace(sc, ray)
    assert h[0, 3].view(np.uint32) == 0 and h[0, 0] == np.float32(5.0)
    sc.destroy()
    # 3000 copies of one triangle (every Morton key equal) and a floor below
    xyz = np.concatenate([np.tile(one, (3000, 1)), [[-5, -1, -5], [5, -1, -5], [5, -1, 5], [-5, -1, 5]]]).astype(
        np.float32)
    idx = np.concatenate([np.arange(9000, dtype=np.uint32).reshape(3000, 3),
                          np.array([[9000, 9002, 9001], [9000, 9003, 9002]], np.uint32)])
    rays = np.zeros((4, 8), np.float32)
    rays[:, 0:3] = [[0.25, 0.25, 5], [0.25, 0.25, -5], [3, 0.5, 3], [0.9, 0.9, 1]]
    rays[:, 4:7] = [[0, 0, -1], [0, 0, 1], [0, -1, 0], [0, 0, -1]]
    rays[:, 7] = np.inf
    for leaf in (1, 4):
        sc = rq.Scene(device, xyz, idx, builder="lbvh", leaf_size=leaf)
        h = _trace(sc, rays)
        p = h[:, 3].view(np.uint32)
        assert p[0] < 3000 and h[0, 0] == np.float32(5.0)  # any copy: a tie
        assert p[1] < 3000 and h[1, 0] == np.float32(5.0)
        assert p[2] in (3000, 3001) and np.isclose(h[2, 0], 1.5)
        assert p[3] == oracle.NO_HIT
        sc.destroy()
    # scenes no larger than one leaf
    for n in (2, 3, 4):
        sc = rq.Scene(device, xyz[:3 * n], idx[:n], builder="lbvh")
        assert _trace(sc, ray)[0, 3].view(np.uint32) < n
        sc.destroy()


def test_lbvh_refit(device, sphere):
    xyz, idx, rays, _ = sphere
    sc = rq.Scene(device, xyz, idx, accel="bvh", builder="lbvh")
    before = _trace(sc, rays)
    xf = _frame(xyz, 2)
    sc.refit(xf)
    got = _trace(sc, rays)
    c = _check_vs_oracle(xf, idx, rays, got, oracle.KdTree(xf, idx).intersect(rays))
    log("lbvh refit: %s" % {k: v for k, v in c.items() if k != "bad_idx"})
    assert c["bad"] == 0
    sc.refit(xyz)
    assert np.array_equal(_trace(sc, rays).view(np.uint32), before.view(np.uint32))
    sc.destroy()


def test_lbvh_rejects_width_4(device):
    xyz, idx = scenes.cornell()
    with pytest.raises(rq.RQError) as e:
        rq.Scene(device, xyz, idx, accel="bvh", builder="lbvh", width=4)
    assert e.value.status == 4
