"""Packed scene files: a saved and reloaded scene answers every query bit for
bit as the original; damaged or foreign files are refused."""
import numpy as np
import pytest
import torch

import oracle
from encore_amd import rq, scenes
from helpers import log

pytestmark = pytest.mark.gpu


def _trace(sc, rays):
    d = torch.from_numpy(rays).cuda()
    h = torch.empty((len(rays), 4), device="cuda")
    o = torch.empty((len(rays),), dtype=torch.int32, device="cuda")
    sc.intersect(d, h)
    short = rays.copy()
    short[:, 7] = 30.0
    sc.occluded(torch.from_numpy(short).cuda(), o)
    torch.cuda.synchronize()
    return h.cpu().numpy(), o.cpu().numpy()


@pytest.fixture(scope="module")
def sphere():
    xyz, idx = scenes.displaced_sphere(level=4)
    prim = scenes.config2_primary(cols=160, rows=90)
    kd = oracle.KdTree(xyz, idx)
    bounce, _ = scenes.config3_rays(xyz, idx, prim, kd.intersect(prim), count=20000, nbatches=2)
    return xyz, idx, np.concatenate([prim, bounce])


@pytest.mark.parametrize("accel,width,pred", [("bvh", 4, "encore"), ("bvh", 2, "encore"), ("bvh", 4, "watertight"),
                                              ("kd", 0, "encore")])
def test_save_load_round_trip(device, sphere, tmp_path, accel, width, pred):
    xyz, idx, rays = sphere
    log("case: scene file round trip %s%d %s" % (accel, width, pred))
    sc = rq.Scene(device, xyz, idx, accel=accel, predicate=pred, width=width)
    path = tmp_path / "scene.rqs"
    sc.save(path)
    ld = rq.Scene.load(device, path)
    a, b = sc.stats(), ld.stats()
    for k in ("accel", "predicate", "ntris", "nnodes", "nleaves", "nrefs", "max_depth", "device_bytes", "bvh_width",
              "stack_bound"):
        assert a[k] == b[k], k
    ha, oa = _trace(sc, rays)
    hb, ob = _trace(ld, rays)
    assert np.array_equal(ha.view(np.uint32), hb.view(np.uint32))
    assert np.array_equal(oa, ob)
    assert (ha[:, 3].view(np.uint32) != oracle.NO_HIT).mean() > 0.3
    sc.destroy()
    ld.destroy()


def test_loaded_scene_serves_final_gather(device, tmp_path):
    """The prim -> record map travels with the file: a final gather on a
    loaded kd scene equals one on the original."""
    xyz, idx = scenes.config4_scene(tess=8)
    kd = oracle.KdTree(xyz, idx)
    pos, nrm, pw = scenes.trace_photons(xyz, idx, kd.intersect, 20000, batch=4096)
    pm = rq.PhotonMap(device, pos, nrm, pw)
    p, n = scenes.photon_queries(xyz, idx, kd.intersect, res=16)
    dirs = scenes.hemisphere_dirs(n, 3, scenes.Stream(3))
    dif = torch.full((len(idx), 3), 0.5, device="cuda")
    sc = rq.Scene(device, xyz, idx, accel="kd")
    sc.save(tmp_path / "kd.rqs")
    ld = rq.Scene.load(device, tmp_path / "kd.rqs")
    outs = []
    for s in (sc, ld):
        rgb = torch.empty((len(p), 3), device="cuda")
        s.final_gather(pm, 20, 0.2, None, 1, 1.0, dif, torch.from_numpy(p).cuda(), torch.from_numpy(dirs).cuda(), 9,
                       rgb)
        outs.append(rgb.cpu().numpy())
    assert np.array_equal(outs[0].view(np.uint32), outs[1].view(np.uint32))
    assert (outs[0] > 0).any()
    for o in (sc, ld, pm):
        o.destroy()


def test_damaged_files_are_refused(device, tmp_path):
    xyz, idx = scenes.cornell()
    sc = rq.Scene(device, xyz, idx, accel="bvh")
    good = tmp_path / "good.rqs"
    sc.save(good)
    data = bytearray(good.read_bytes())
    cases = {}
    flipped = bytearray(data)
    flipped[-100] ^= 0x40
    cases["checksum"] = (flipped, 1)
    cases["truncated"] = (data[:-7], 1)
    cases["trailing"] = (data + b"x", 1)
    magic = bytearray(data)
    magic[0:8] = b"NOTSCENE"
    cases["magic"] = (magic, 1)
    version = bytearray(data)
    version[8:12] = (99).to_bytes(4, "little")
    cases["version"] = (version, 4)
    for name, (blob, status) in cases.items():
        path = tmp_path / (name + ".rqs")
        path.write_bytes(bytes(blob))
        with pytest.raises(rq.RQError) as e:
            rq.Scene.load(device, path)
        assert e.value.status == status, name
    with pytest.raises(rq.RQError):
        rq.Scene.load(device, tmp_path / "missing.rqs")
    ld = rq.Scene.load(device, good)  # the intact file still loads
    ld.destroy()
    sc.destroy()
