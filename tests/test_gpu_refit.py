"""Scene update for moved vertices (rq_scene_refit).  A refitted BVH keeps
its tree and recomputes every box from the new vertices, so its closest hits
on the moved geometry are the oracle kd-tree's (built on the new vertices)
up to the BVH classes (ties, phantoms outside their box); an update with the
same vertices changes no result bit; a kd scene is built again and stays
bit-exact."""
import math

import numpy as np
import pytest
import torch

import oracle
from encore_amd import rq, scenes
from helpers import brute_occluded, log
from test_gpu_trace import _check_vs_oracle, _trace

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def anim():
    xyz, idx = scenes.displaced_sphere(level=4)
    prim = scenes.config2_primary(cols=160, rows=90)
    kd = oracle.KdTree(xyz, idx)
    bounce, _ = scenes.config3_rays(xyz, idx, prim, kd.intersect(prim), count=20000, nbatches=2)
    rays = np.concatenate([prim, bounce])
    return xyz, idx, rays


def _frame(xyz, f):
    """Frame f of an animation: a turn about y, a squash along y and a
    travelling wobble; f = 0 is the rest pose."""
    p = xyz.astype(np.float64)
    a = 0.3 * f
    c, s = math.cos(a), math.sin(a)
    x = c * p[:, 0] + s * p[:, 2]
    z = -s * p[:, 0] + c * p[:, 2]
    y = p[:, 1] * (1.0 - 0.1 * f)
    w = 1.0 + 0.04 * f * np.sin(p[:, 1] / 12.0 + f)
    return np.stack([x * w, y, z * w], axis=1).astype(np.float32)


@pytest.mark.parametrize("width,pred", [(4, "encore"), (2, "encore"), (4, "watertight")])
def test_refit_with_same_vertices_changes_nothing(device, anim, width, pred):
    xyz, idx, rays = anim
    log("case: refit identity bvh%d %s" % (width, pred))
    sc = rq.Scene(device, xyz, idx, accel="bvh", width=width, predicate=pred)
    before = _trace(sc, rays)
    occ_rays = rays.copy()
    occ_rays[:, 7] = 30.0
    occ0 = _trace(sc, occ_rays, "any")
    sc.refit(xyz)
    after = _trace(sc, rays)
    assert np.array_equal(before.view(np.uint32), after.view(np.uint32))
    assert np.array_equal(occ0, _trace(sc, occ_rays, "any"))
    sc.destroy()


@pytest.mark.parametrize("width", [4, 2])
def test_refit_animation_matches_rebuilt_oracle(device, anim, width):
    xyz, idx, rays = anim
    sc = rq.Scene(device, xyz, idx, accel="bvh", width=width)
    for f in (1, 2, 4):
        xf = _frame(xyz, f)
        log("case: refit bvh%d frame %d" % (width, f))
        sc.refit(xf)
        got = _trace(sc, rays)
        ref = oracle.KdTree(xf, idx).intersect(rays)
        c = _check_vs_oracle(xf, idx, rays, got, ref)
        log("frame %d: %s" % (f, {k: v for k, v in c.items() if k != "bad_idx"}))
        assert c["bad"] == 0
        assert c["exact"] >= 0.99 * len(rays)
        # any hit on the moved geometry against brute force
        occ_rays = rays[::7].copy()
        occ_rays[:, 7] = 40.0
        occ = _trace(sc, occ_rays, "any")
        want = brute_occluded(xf, idx, occ_rays)
        assert (occ.astype(np.uint32) == want).mean() > 0.999
    # back to the rest pose: the original results, bit for bit
    sc.refit(xyz)
    fresh = rq.Scene(device, xyz, idx, accel="bvh", width=width)
    assert np.array_equal(_trace(sc, rays).view(np.uint32), _trace(fresh, rays).view(np.uint32))
    sc.destroy()
    fresh.destroy()


def test_refit_kd_scene_is_rebuilt_bit_exact(device, anim):
    xyz, idx, rays = anim
    sc = rq.Scene(device, xyz, idx, accel="kd")
    for f in (1, 3):
        xf = _frame(xyz, f)
        log("case: refit kd frame %d" % f)
        sc.refit(xf)
        got = _trace(sc, rays)
        ref = oracle.KdTree(xf, idx).intersect(rays)
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    sc.destroy()


def test_refit_rejects_bad_arguments(device, anim, tmp_path):
    xyz, idx, rays = anim
    sc = rq.Scene(device, xyz, idx, accel="bvh")
    with pytest.raises(rq.RQError) as e:
        sc.refit(xyz[:-1])
    assert e.value.status == 1
    sc.save(tmp_path / "s.rqs")
    ld = rq.Scene.load(device, tmp_path / "s.rqs")
    with pytest.raises(rq.RQError) as e:
        ld.refit(xyz)
    assert e.value.status == 1
    ld.destroy()
    sc.destroy()
