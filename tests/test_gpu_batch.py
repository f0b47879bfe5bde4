"""Batched persistent launches (descriptor table, hipGraph replay, autotuner)
give exactly the per-batch launch results."""
import numpy as np
import pytest
import torch

import oracle
from encore_amd import rq, scenes
from helpers import log

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=[2, 4], ids=["bvh2", "bvh4"])
def two_scenes(device, request):
    a = scenes.displaced_sphere(level=4)
    b = scenes.cornell()
    sa = rq.Scene(device, *a, accel="bvh", width=request.param)
    sb = rq.Scene(device, *b, accel="bvh", width=request.param)
    kda = oracle.KdTree(*a)
    prim = scenes.config2_primary(cols=160, rows=90)
    ray_a, _ = scenes.config3_rays(a[0], a[1], prim, kda.intersect(prim), count=30000, nbatches=2)
    _, _, ray_b = scenes.config1(res=128)
    yield sa, sb, ray_a, ray_b
    sa.destroy()
    sb.destroy()


def _plain(scene, rays):
    d = torch.from_numpy(rays).cuda()
    out = torch.empty((len(rays), 4), dtype=torch.float32, device="cuda")
    scene.intersect(d, out)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def test_batched_launch_matches_plain(device, two_scenes):
    sa, sb, ray_a, ray_b = two_scenes
    log("case: batched mixed scenes")
    cuts_a = [0, 1, 4097, 4097, 17000, len(ray_a)]  # includes empty and 1-ray items
    cuts_b = [0, 5000, len(ray_b)]
    items, refs = [], []
    for s, rays, cuts in ((sa, ray_a, cuts_a), (sb, ray_b, cuts_b)):
        for lo, hi in zip(cuts[:-1], cuts[1:]):
            r = np.ascontiguousarray(rays[lo:hi])
            d = torch.from_numpy(r).cuda() if len(r) else torch.empty((0, 8), device="cuda")
            o = torch.full((max(1, len(r)), 4), -5.0, device="cuda")
            items.append((s, d, o))
            refs.append(_plain(s, r) if len(r) else None)
    b = rq.Batch(device)
    b.set(items)
    for graph in (False, True, True, True):
        for _, _, o in items:
            o.fill_(-5.0)
        b.launch(graph=graph)
        torch.cuda.synchronize()
        _ = torch.randn(1000, device="cuda").sum().item()  # eager work between replays
        for (s, d, o), ref in zip(items, refs):
            if ref is None:
                continue
            got = o[: d.shape[0]].cpu().numpy()
            assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), graph
    cfg = b.autotune()
    log("autotuned:", cfg)
    assert cfg["stack"] in (8, 16, 32) and cfg["ms"] > 0
    for _, _, o in items:
        o.fill_(-5.0)
    b.launch(graph=True)
    torch.cuda.synchronize()
    for (s, d, o), ref in zip(items, refs):
        if ref is not None:
            assert np.array_equal(o[: d.shape[0]].cpu().numpy().view(np.uint32), ref.view(np.uint32))
    b.destroy()


def test_batched_any_hit(device, two_scenes):
    sa, _, ray_a, _ = two_scenes
    log("case: batched any-hit")
    r = ray_a.copy()
    r[:, 7] = 50.0
    d = torch.from_numpy(r).cuda()
    plain = torch.empty((len(r),), dtype=torch.int32, device="cuda")
    sa.occluded(d, plain)
    out = torch.full((len(r),), 9, dtype=torch.int32, device="cuda")
    b = rq.Batch(device, any_hit=True)
    b.set([(sa, d[:1000], out[:1000]), (sa, d[1000:], out[1000:])])
    b.launch(graph=True)
    torch.cuda.synchronize()
    assert torch.equal(out, plain)
    b.destroy()


def test_batch_rejects_mixed_widths(device):
    xyz, idx = scenes.cornell()
    s2 = rq.Scene(device, xyz, idx, accel="bvh", width=2)
    s4 = rq.Scene(device, xyz, idx, accel="bvh", width=4)
    b = rq.Batch(device)
    d = torch.zeros((4, 8), device="cuda")
    o = torch.zeros((4, 4), device="cuda")
    with pytest.raises(rq.RQError) as e:
        b.set([(s2, d, o), (s4, d, o)])
    assert e.value.status == 4
    b.destroy()
    s2.destroy()
    s4.destroy()


def test_batch_rejects_mixed_kinds(device):
    xyz, idx = scenes.cornell()
    kd = rq.Scene(device, xyz, idx, accel="kd")
    bvh = rq.Scene(device, xyz, idx, accel="bvh")
    b = rq.Batch(device)
    d = torch.zeros((4, 8), device="cuda")
    o = torch.zeros((4, 4), device="cuda")
    with pytest.raises(rq.RQError) as e:
        b.set([(kd, d, o), (bvh, d, o)])
    assert e.value.status == 4
    b.destroy()
    kd.destroy()
    bvh.destroy()


@pytest.mark.parametrize("any_hit", [False, True], ids=["closest", "any"])
def test_batched_kd_matches_oracle(device, any_hit):
    """The exact path in one persistent launch: items on two kd scenes (a
    displaced sphere's bounce rays, the Cornell box's eye rays), of uneven
    sizes including empty and one-ray items, give ENCORE's kd-tree result bit
    for bit on every ray (Kdtree.cpp:682-784), across graph replays and after
    autotuning."""
    a = scenes.displaced_sphere(level=4)
    c = scenes.cornell()
    kda, kdc = oracle.KdTree(*a), oracle.KdTree(*c)
    prim = scenes.config2_primary(cols=160, rows=90)
    ray_a, _ = scenes.config3_rays(a[0], a[1], prim, kda.intersect(prim), count=30000, nbatches=2)
    _, _, ray_c = scenes.config1(res=128)
    if any_hit:
        ray_a = ray_a.copy()
        ray_a[:, 7] = 40.0
    sa = rq.Scene(device, *a, accel="kd")
    sc = rq.Scene(device, *c, accel="kd")
    log("case: batched kd", "any" if any_hit else "closest")
    items, refs = [], []
    for s, kd, rays, cuts in ((sa, kda, ray_a, [0, 1, 4097, 4097, 17000, len(ray_a)]),
                              (sc, kdc, ray_c, [0, 5000, 5001, len(ray_c)])):
        for lo, hi in zip(cuts[:-1], cuts[1:]):
            r = np.ascontiguousarray(rays[lo:hi])
            d = torch.from_numpy(r).cuda() if len(r) else torch.empty((0, 8), device="cuda")
            if any_hit:
                o = torch.full((max(1, len(r)),), 7, dtype=torch.int32, device="cuda")
                ref = kd.occluded(r).astype(np.int32) if len(r) else None
            else:
                o = torch.full((max(1, len(r)), 4), -5.0, device="cuda")
                ref = kd.intersect(r) if len(r) else None
            items.append((s, d, o))
            refs.append(ref)
    b = rq.Batch(device, any_hit=any_hit)
    b.set(items)

    def check(tag):
        for (s, d, o), ref in zip(items, refs):
            if ref is None:
                continue
            got = o[: d.shape[0]].cpu().numpy()
            assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), tag

    for graph in (False, True, True, True):
        for _, _, o in items:
            o.fill_(7 if any_hit else -5.0)
        b.launch(graph=graph)
        torch.cuda.synchronize()
        _ = torch.randn(1000, device="cuda").sum().item()  # eager work between replays
        check("graph=%s" % graph)
    cfg = b.autotune()
    log("autotuned:", cfg)
    assert cfg["stack"] == 8 and cfg["ms"] > 0
    for _, _, o in items:
        o.fill_(7 if any_hit else -5.0)
    b.launch(graph=True)
    torch.cuda.synchronize()
    check("after autotune")
    b.destroy()
    sa.destroy()
    sc.destroy()
