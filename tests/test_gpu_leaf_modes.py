"""The BVH4 kernel's optional modes give the default kernel's results bit
for bit: the wave-shared leaf tests (RQ_BVH_LEAF_MODE=wave), where the lanes
waiting at a leaf hand its triangles to the whole wave and keep the least
(t, leaf position) of their hits -- also with leaves of 16 triangles, so
that more than 64 triangles wait in one round -- and batched launches with
each LDS stack ring size.  Closest and any hit, both predicates."""
import numpy as np
import pytest
import torch

import oracle
from encore_amd import rq, scenes
from helpers import log

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rays_and_scene():
    xyz, idx = scenes.displaced_sphere(level=5)
    prim = scenes.config2_primary(cols=200, rows=120)
    kd = oracle.KdTree(xyz, idx)
    ph = kd.intersect(prim)
    bounce, _ = scenes.config3_rays(xyz, idx, prim, ph, count=30000, nbatches=2)
    ok, shadow = scenes.config2_shadow(xyz, idx, prim, ph)
    return xyz, idx, np.concatenate([prim, bounce]), shadow


def _run(sc, rays, any_hit):
    d = torch.from_numpy(np.ascontiguousarray(rays)).cuda()
    if any_hit:
        out = torch.full((len(rays),), 7, dtype=torch.int32, device="cuda")
        sc.occluded(d, out)
    else:
        out = torch.full((len(rays), 4), -7.0, device="cuda")
        sc.intersect(d, out)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def _batched(device, items, any_hit, stack=0):
    """items: (scene, rays); one persistent launch; returns the outputs."""
    b = rq.Batch(device, any_hit=any_hit)
    tensors = []
    for sc, r in items:
        d = torch.from_numpy(np.ascontiguousarray(r)).cuda()
        o = (torch.full((len(r),), 7, dtype=torch.int32, device="cuda") if any_hit
             else torch.full((len(r), 4), -7.0, device="cuda"))
        tensors.append((sc, d, o))
    b.set(tensors)
    if stack:
        b.set_config(stack=stack, blocks_per_cu=0, fetch_threshold=44)
    b.launch(graph=False)
    torch.cuda.synchronize()
    out = [o.cpu().numpy() for _, _, o in tensors]
    b.destroy()
    return out


def _same(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint32), np.ascontiguousarray(b).view(np.uint32))


@pytest.mark.parametrize("predicate", ["encore", "watertight"])
@pytest.mark.parametrize("leaf", [1, 4, 16])
def test_wave_leaf_mode_equals_lane_mode(device, rays_and_scene, predicate, leaf, monkeypatch):
    xyz, idx, rays, shadow = rays_and_scene
    sc = rq.Scene(device, xyz, idx, accel="bvh", width=4, predicate=predicate, leaf_size=leaf)
    lane = [_run(sc, rays, False), _run(sc, shadow, True)]
    lane_b = _batched(device, [(sc, rays[:20000]), (sc, rays[20000:])], False)
    monkeypatch.setenv("RQ_BVH_LEAF_MODE", "wave")
    wave = [_run(sc, rays, False), _run(sc, shadow, True)]
    wave_b = _batched(device, [(sc, rays[:20000]), (sc, rays[20000:])], False)
    monkeypatch.delenv("RQ_BVH_LEAF_MODE")
    log("case: leaf mode wave, %s, leaf %d: %d rays, %d hits, %d shadow rays occluded" % (
        predicate, leaf, len(rays), (lane[0][:, 3].view(np.uint32) != 0xFFFFFFFF).sum(), lane[1].sum()))
    assert _same(wave[0], lane[0]) and _same(wave[1], lane[1])
    assert _same(wave_b[0], lane[0][:20000]) and _same(wave_b[1], lane[0][20000:])
    assert _same(lane_b[0], lane[0][:20000]) and _same(lane_b[1], lane[0][20000:])
    sc.destroy()


@pytest.mark.parametrize("any_hit", [False, True])
def test_stack_sizes_equal_default(device, rays_and_scene, any_hit):
    xyz, idx, rays, shadow = rays_and_scene
    r = shadow if any_hit else rays
    sc = rq.Scene(device, xyz, idx, accel="bvh", width=4)
    ref = _run(sc, r, any_hit)
    half = len(r) // 2
    for stack in (8, 16, 32):
        got = _batched(device, [(sc, r[:half]), (sc, r[half:])], any_hit, stack=stack)
        log("case: batched stack %d any_hit=%s" % (stack, any_hit))
        assert _same(got[0], ref[:half]) and _same(got[1], ref[half:])
    sc.destroy()
