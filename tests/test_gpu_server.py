"""Serving: continuous batching returns, for every request, exactly what a
plain launch returns for it alone; admission by HBM budget queues requests
instead of failing them; cancellation, capacity and allocation errors."""
import numpy as np
import pytest
import torch

import oracle
from encore_amd import rq, scenes
from helpers import log

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def resident(device):
    out = []
    # The third scene is a BVH2: requests on it are never batched with the
    # BVH4 ones.
    for (xyz, idx), width in ((scenes.displaced_sphere(level=4), 4), (scenes.cornell(), 4),
                              (scenes.displaced_sphere(freq=30), 2)):
        sc = rq.Scene(device, xyz, idx, accel="bvh", width=width)
        prim = scenes.config2_primary(cols=200, rows=100) if len(idx) > 100 else scenes.config1(res=256)[2]
        kd = oracle.KdTree(xyz, idx)
        bounce, _ = scenes.config3_rays(xyz, idx, prim, kd.intersect(prim), count=60000, nbatches=2)
        out.append((sc, np.concatenate([prim, bounce])))
    yield out
    for sc, _ in out:
        sc.destroy()


def _alone(scene, rays, any_hit):
    d = torch.from_numpy(np.ascontiguousarray(rays)).cuda()
    if any_hit:
        o = torch.empty((len(rays),), dtype=torch.int32, device="cuda")
        scene.occluded(d, o)
    else:
        o = torch.empty((len(rays), 4), dtype=torch.float32, device="cuda")
        scene.intersect(d, o)
    torch.cuda.synchronize()
    return o.cpu().numpy()


def _requests(resident, count, seed, max_n=50000):
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(count):
        sc, pool = resident[rng.integers(len(resident))]
        n = int(rng.integers(1, max_n))
        lo = int(rng.integers(0, len(pool) - n))
        rays = pool[lo:lo + n].copy()
        any_hit = bool(rng.random() < 0.3)
        if any_hit:
            rays[:, 7] = rng.uniform(1, 200, n).astype(np.float32)
        reqs.append((sc, rays, any_hit))
    return reqs


def _run(server, reqs):
    tickets, bufs = [], []
    for sc, rays, any_hit in reqs:
        r = torch.from_numpy(rays).pin_memory()
        o = torch.empty((len(rays),) if any_hit else (len(rays), 4), dtype=torch.int32 if any_hit else torch.float32)
        o = o.pin_memory()
        tickets.append(server.submit(sc, r, o, any_hit))
        bufs.append(o)
    infos = [server.wait(t, 60000) for t in tickets]
    return [b.numpy() for b in bufs], infos


@pytest.mark.parametrize("plain", [False, True])
def test_served_results_equal_alone(device, resident, plain):
    log("case: server mix plain=%s" % plain)
    reqs = _requests(resident, 40, 1)
    srv = rq.Server(device, plain=plain)
    outs, infos = _run(srv, reqs)
    st = srv.stats()
    srv.destroy()
    for (sc, rays, any_hit), got in zip(reqs, outs):
        ref = _alone(sc, rays, any_hit)
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    mixed = max(i["requests_in_launch"] for i in infos)
    log("launches %d for %d requests; largest launch held %d requests" % (st["launches"], len(reqs), mixed))
    if plain:
        assert st["launches"] == len(reqs) and mixed == 1
    else:
        assert st["launches"] < len(reqs) and mixed > 1


def test_admission_by_small_budget_queues(device, resident):
    log("case: server small budget")
    reqs = _requests(resident, 24, 2, max_n=40000)
    budget = 2 * 1024 * 1024  # one 40k-ray request needs ~1.9 MB
    srv = rq.Server(device, hbm_budget=budget)
    outs, infos = _run(srv, reqs)
    st = srv.stats()
    srv.destroy()
    assert st["requests_done"] == len(reqs) and st["arena_used"] == 0
    for (sc, rays, any_hit), got in zip(reqs, outs):
        assert np.array_equal(got.view(np.uint32), _alone(sc, rays, any_hit).view(np.uint32))
    with pytest.raises(rq.RQError) as e:
        srv2 = rq.Server(device, hbm_budget=budget)
        try:
            big = np.zeros((60000, 8), np.float32)
            srv2.submit(resident[0][0], big, np.zeros((60000, 4), np.float32))
        finally:
            srv2.destroy()
    assert e.value.status == 5  # RQ_ERR_CAPACITY


def test_cancel_and_timeout(device, resident):
    log("case: server cancel")
    sc, pool = resident[0]
    srv = rq.Server(device, hbm_budget=1 << 20, slots=1)  # ~21k rays at a time
    reqs = [np.ascontiguousarray(pool[i * 20000:(i + 1) * 20000]) for i in range(4)]
    outs = [np.zeros((20000, 4), np.float32) for _ in reqs]
    tickets = [srv.submit(sc, r, o) for r, o in zip(reqs, outs)]
    cancelled = False
    try:
        srv.cancel(tickets[-1])
        cancelled = True
    except rq.RQError as e:
        assert e.status == 4  # already launched: allowed race
    for t in tickets[:-1]:
        srv.wait(t, 60000)
    if cancelled:
        with pytest.raises(rq.RQError) as e:
            srv.wait(tickets[-1], 1000)
        assert e.value.status == 7
    else:
        srv.wait(tickets[-1], 60000)
    srv.destroy()


def test_allocation_failure_leaves_scenes_intact(device, resident):
    log("case: injected OOM")
    sc, pool = resident[1]
    rq.fail_alloc_after(1)
    try:
        with pytest.raises(rq.RQError) as e:
            rq.Server(device)
        assert e.value.status == 2  # RQ_ERR_OOM
    finally:
        rq.fail_alloc_after(0)
    rays = np.ascontiguousarray(pool[:1000])
    ref = oracle.KdTree(*scenes.cornell()).intersect(rays)
    got = _alone(sc, rays, False)
    assert (got[:, 3].view(np.uint32) == ref[:, 3].view(np.uint32)).mean() > 0.99
    rq.fail_alloc_after(1)
    try:
        with pytest.raises(rq.RQError) as e:
            rq.Scene(device, *scenes.cornell())
        assert e.value.status == 2
    finally:
        rq.fail_alloc_after(0)
