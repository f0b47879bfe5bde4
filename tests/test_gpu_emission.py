"""GPU photon emission (rq_emit_photons) against the oracle's sequential
restatement of EmitPhotons / TracePhoton.  On ENCORE's kd-tree every ray
hits what the oracle's kd-tree hits, so all three maps come out bit for bit
the same, photon for photon; on a BVH the few rays that tie or meet the
reference's flush phantoms change their own paths only."""
import numpy as np
import pytest
import torch

import oracle
from encore_amd import rq, scenes
from helpers import log, outside_box
from test_oracle_emission import _closed_box, closed_box_expectation

pytestmark = pytest.mark.gpu

FIELDS = ("pos", "normal", "power", "dir", "shadow", "path")


def _gpu(sc, lights, cap, **kw):
    maps, paths = sc.emit_photons(lights, cap, **kw)
    torch.cuda.synchronize()
    return {m: {k: v.cpu().numpy() for k, v in maps[m].items()} for m in maps}, paths


def _same(a, b):
    for m in oracle.MAPS:
        for f in FIELDS:
            x, y = a[m][f], b[m][f]
            if x.dtype == np.float32:
                x, y = x.view(np.uint32), y.view(np.uint32)
            assert np.array_equal(x, y.astype(x.dtype)), (m, f)


@pytest.fixture(scope="module", params=[False, True], ids=["cornell", "glass"])
def pscene(request):
    xyz, idx, mats, pm, lights = scenes.photon_scene(tess=8, glass=request.param)
    kd = oracle.KdTree(xyz, idx)
    cap = (25000, 50000, 2000 if request.param else 0)  # pmConfig.txt: GlobalPhotons 25000, DirectPhotons 50000
    o_maps, o_paths = oracle.emit_photons(kd, mats, pm, lights, cap, seed=7)
    return dict(xyz=xyz, idx=idx, mats=mats, pm=pm, lights=lights, kd=kd, cap=cap, o=(o_maps, o_paths),
                glass=request.param)


def test_emission_kd_matches_oracle(device, pscene):
    f = pscene
    log("case: photon emission kd, glass=%s, capacities %s" % (f["glass"], f["cap"]))
    sc = rq.Scene(device, f["xyz"], f["idx"], accel="kd")
    sc.set_materials(f["mats"], f["pm"])
    g, paths = _gpu(sc, f["lights"], f["cap"], seed=7)
    o, o_paths = f["o"]
    log("paths gpu %d oracle %d; counts %s" % (paths, o_paths, [len(g[m]["pos"]) for m in oracle.MAPS]))
    assert paths == o_paths
    _same(g, o)
    sc.destroy()


def test_emission_is_independent_of_batch_size(device, pscene):
    f = pscene
    log("case: photon emission batch sizes, glass=%s" % f["glass"])
    sc = rq.Scene(device, f["xyz"], f["idx"], accel="kd")
    sc.set_materials(f["mats"], f["pm"])
    a, pa = _gpu(sc, f["lights"], f["cap"], seed=7, batch_paths=1024)
    b, pb = _gpu(sc, f["lights"], f["cap"], seed=7, batch_paths=3000)
    assert pa == pb == f["o"][1]
    _same(a, b)
    _same(a, f["o"][0])
    sc.destroy()


def _keys(x, shadow):
    sel = x["shadow"].astype(bool) == shadow
    return set(zip(x["path"][sel].astype(np.int64).tolist(), map(bytes, x["pos"][sel].view(np.uint32)),
                   map(bytes, x["normal"][sel].view(np.uint32))))


def _shadow_ray_ties(f, maps, paths):
    """For each path, whether its shadow-photon chain (from the direct
    photon's hit point, 2 along the photon's direction, through surfaces
    being left) meets, at some step, two or more triangles at its closest
    distance (a tie: a box standing on the floor has its bottom face in the
    floor's plane) or a reference phantom outside its triangle's box."""
    d = maps["direct"]
    first = {}
    for i in np.nonzero(d["shadow"] == 0)[0]:
        first.setdefault(int(d["path"][i]), i)
    rows = [first[p] for p in paths]
    rays = np.zeros((len(rows), 8), np.float32)
    rays[:, 0:3] = d["pos"][rows] + d["dir"][rows] * np.float32(2.0)
    rays[:, 4:7] = d["dir"][rows]
    rays[:, 7] = np.inf
    tie = np.zeros(len(rows), bool)
    live = np.arange(len(rows))
    for _ in range(32):
        if not len(live):
            break
        h = f["kd"].intersect(rays[live])
        hit = h[:, 3].view(np.uint32) != oracle.NO_HIT
        live, h = live[hit], h[hit]
        tie[live] |= oracle.count_hits_below(f["xyz"], f["idx"], rays[live], h[:, 0], 0.0) >= 2
        # a reference phantom outside its triangle's box, which no BVH reaches
        prim = h[:, 3].view(np.uint32)
        tie[live] |= np.array([outside_box(f["xyz"], f["idx"], rays[r], h[k, 0], prim[k]) > 0
                               for k, r in enumerate(live)], bool)
        _, P, N = oracle.hit_attributes(f["xyz"], f["idx"], h)
        leaving = (rays[live, 4:7] * N).sum(axis=1) >= 0
        live, P = live[leaving], P[leaving]
        rays[live, 0:3] = P + rays[live, 4:7] * np.float32(0.01)
    return tie


@pytest.mark.parametrize("width", [4, 2])
def test_emission_bvh_close_to_oracle(device, pscene, width):
    """Photons are compared by (path, position, normal).  At least 99% of
    the oracle's photons appear identically; every shadow photon found by
    only one of the two comes from a shadow ray whose closest hit is a tie
    between coplanar triangles (the BVH may return the surface being left,
    so the ray goes on and finds nothing)."""
    f = pscene
    log("case: photon emission bvh%d, glass=%s" % (width, f["glass"]))
    sc = rq.Scene(device, f["xyz"], f["idx"], accel="bvh", width=width)
    sc.set_materials(f["mats"], f["pm"])
    g, paths = _gpu(sc, f["lights"], f["cap"], seed=7)
    o, o_paths = f["o"]
    assert abs(paths - o_paths) <= 0.01 * o_paths
    for m in oracle.MAPS:
        if not len(o[m]["pos"]):
            assert not len(g[m]["pos"])
            continue
        common = len(_keys(g[m], False) & _keys(o[m], False))
        n = int((o[m]["shadow"] == 0).sum())
        log("bvh%d %s: %d of %d photons identical" % (width, m, common, n))
        assert common >= 0.99 * n
    gs, os_ = _keys(g["direct"], True), _keys(o["direct"], True)
    only = sorted({k[0] for k in gs ^ os_})
    log("bvh%d shadow photons: %d gpu, %d oracle, %d paths differ" % (width, len(gs), len(os_), len(only)))
    if only:
        assert _shadow_ray_ties(f, o, only).all()
    sc.destroy()


@pytest.mark.parametrize("cap", [(4000, 700, 0), (1, 1, 0), (0, 500, 300), (3000, 6000, 1)])
def test_emission_capacity_edges(device, cap):
    """The direct map filling first, one-photon maps, no global map (only
    the caustic loop runs), a one-photon caustic map: all as the oracle."""
    xyz, idx, mats, pm, lights = scenes.photon_scene(tess=4, glass=True)
    kd = oracle.KdTree(xyz, idx)
    o, o_paths = oracle.emit_photons(kd, mats, pm, lights, cap, seed=21, max_bounces=5)
    log("case: photon capacities %s" % (cap,))
    sc = rq.Scene(device, xyz, idx, accel="kd")
    sc.set_materials(mats, pm)
    g, paths = _gpu(sc, lights, cap, seed=21, max_bounces=5, batch_paths=2048)
    assert paths == o_paths
    _same(g, o)
    sc.destroy()


def test_emission_closed_box(device):
    xyz, idx = _closed_box()
    light = dict(center=(0.0, 0.9, 0.0), normal=(0.0, -1.0, 0.0), length=0.5, width=0.5)
    log("case: photon emission closed box")
    for accel in ("kd", "bvh"):
        sc = rq.Scene(device, xyz, idx, accel=accel)
        sc.set_materials([dict(diffuse=(1, 1, 1))], np.zeros(12, np.uint32))
        g, paths = _gpu(sc, [light], (1000, 5000, 0), seed=5)
        assert 334 <= paths <= 345
        closed_box_expectation(g, paths, 1000)
        sc.destroy()


def test_emission_bound_when_a_map_cannot_fill(device):
    xyz, idx = _closed_box()
    light = dict(center=(0.0, 0.9, 0.0), normal=(0.0, -1.0, 0.0), length=0.5, width=0.5)
    kd = oracle.KdTree(xyz, idx)
    o, o_paths = oracle.emit_photons(kd, [dict(diffuse=(1, 1, 1))], np.zeros(12, np.uint32), [light], (30, 5000, 10),
                                     seed=5, max_paths=5000)
    sc = rq.Scene(device, xyz, idx, accel="kd")
    sc.set_materials([dict(diffuse=(1, 1, 1))], np.zeros(12, np.uint32))
    g, paths = _gpu(sc, [light], (30, 5000, 10), seed=5, max_paths=5000, batch_paths=1024)
    assert paths == o_paths == 5000
    _same(g, o)
    sc.destroy()


def test_emission_rejects_bad_arguments(device):
    xyz, idx = _closed_box()
    light = dict(center=(0.0, 0.9, 0.0), normal=(0.0, -1.0, 0.0), length=0.5, width=0.5)
    sc = rq.Scene(device, xyz, idx, accel="bvh")
    with pytest.raises(rq.RQError) as e:  # no materials yet
        sc.emit_photons([light], (10, 10, 0))
    assert e.value.status == 1
    with pytest.raises(rq.RQError):
        sc.set_materials([dict(diffuse=(1, 1, 1))], np.ones(12, np.uint32))  # index out of range
    sc.set_materials([dict(diffuse=(1, 1, 1))], np.zeros(12, np.uint32))
    with pytest.raises(rq.RQError):
        sc.emit_photons([light], (10, 10, 0), max_bounces=17)
    with pytest.raises(rq.RQError):
        sc.emit_photons([dict(light, normal=(0.0, 0.0, 0.0))], (10, 10, 0))
    with pytest.raises(rq.RQError):
        sc.emit_photons([], (10, 10, 0))
    maps, paths = sc.emit_photons([light], (0, 10, 0))  # both loops end at once
    assert paths == 0 and all(len(maps[m]["pos"]) == 0 for m in maps)
    sc.destroy()


def test_emission_shadow_chain_bound_matches_oracle(device):
    """The scene and seed where shadow-photon rays loop on phantom hits: the
    bounded chain gives the oracle's maps bit for bit."""
    xyz, idx, mats, pm, lights = scenes.photon_scene(tess=32)
    kd = oracle.KdTree(xyz, idx)
    o, o_paths = oracle.emit_photons(kd, mats, pm, lights, (20000, 40000, 0), seed=7)
    log("case: photon emission shadow-chain bound (tess 32)")
    sc = rq.Scene(device, xyz, idx, accel="kd")
    sc.set_materials(mats, pm)
    g, paths = _gpu(sc, lights, (20000, 40000, 0), seed=7)
    assert paths == o_paths
    _same(g, o)
    sc.destroy()


def test_emission_several_and_tilted_lights_kd(device):
    """Three lights (the draw rand() % 3 picks one per path): the ceiling
    light, one on the back wall facing +z and one tilted (the rotation branch
    of AlignPlaneWithNormal, Sampler.cpp:174-192), with different powers."""
    xyz, idx, mats, pm, lights = scenes.photon_scene(tess=8)
    lights = lights + [dict(center=(0.5, -0.5, -2.49), normal=(0.0, 0.0, 1.0), length=0.8, width=0.6,
                            power=(0.2, 0.4, 0.9)),
                       dict(center=(-1.0, 1.5, 0.5), normal=(0.3, -0.9, 0.1), length=0.5, width=0.7,
                            power=(2.0, 1.0, 0.5))]
    kd = oracle.KdTree(xyz, idx)
    cap = (8000, 12000, 0)
    o, o_paths = oracle.emit_photons(kd, mats, pm, lights, cap, seed=99)
    log("case: photon emission, 3 lights")
    sc = rq.Scene(device, xyz, idx, accel="kd")
    sc.set_materials(mats, pm)
    g, paths = _gpu(sc, lights, cap, seed=99, batch_paths=4096)
    assert paths == o_paths
    _same(g, o)
    # every light contributes, with its own power
    pw = np.unique(o["direct"]["power"][o["direct"]["shadow"] == 0], axis=0)
    assert len(pw) == 3
    sc.destroy()
