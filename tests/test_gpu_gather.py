"""GPU final gather (rq_final_gather) against the oracle's composition of
ENCORE's kd-tree, hit attributes and photon maps."""
import numpy as np
import pytest
import torch

import oracle
from encore_amd import rq, scenes
from helpers import classify_closest, log

pytestmark = pytest.mark.gpu

K_IND, R_IND = 100, 0.15  # ENCORE's NGatherPhotons 100 and MaxDistance 3, at world scale 1/20
K_DIR, R_DIR = 100, 0.10  # the direct map's GetRadianceEstimate(100, 2, ...), at 1/20


@pytest.fixture(scope="module")
def fg(device):
    xyz, idx = scenes.config4_scene()
    kd = oracle.KdTree(xyz, idx)
    # small emission batches so that every batch runs through all its bounces
    pos, nrm, pw, bounce = scenes.trace_photons(xyz, idx, kd.intersect, 200_000, batch=32768, with_bounce=True)
    assert (bounce >= 1).mean() > 0.3
    ind, dr = bounce >= 1, bounce == 0
    maps = dict(
        ind_o=oracle.PhotonMap(pos[ind], nrm[ind], pw[ind]), dir_o=oracle.PhotonMap(pos[dr], nrm[dr], pw[dr]),
        ind_g=rq.PhotonMap(device, pos[ind], nrm[ind], pw[ind]), dir_g=rq.PhotonMap(device, pos[dr], nrm[dr], pw[dr]))
    p, n = scenes.photon_queries(xyz, idx, kd.intersect, res=40)
    rng = np.random.default_rng(17)
    diffuse = rng.uniform(0.2, 0.9, (len(idx), 3)).astype(np.float32)
    yield dict(xyz=xyz, idx=idx, kd=kd, p=p[:1024], n=n[:1024], diffuse=diffuse, ind_pos=pos[ind], ind_nrm=nrm[ind],
               **maps)
    maps["ind_g"].destroy()
    maps["dir_g"].destroy()


def _engine(scene, f, p, dirs, nper, direct=True):
    rgb = torch.full((len(p), 3), -1.0, device="cuda")
    ri = torch.full((len(p),), -1.0, device="cuda")
    scene.final_gather(f["ind_g"], K_IND, R_IND, f["dir_g"] if direct else None, K_DIR, R_DIR,
                       torch.from_numpy(f["diffuse"]).cuda(), torch.from_numpy(p).cuda(),
                       torch.from_numpy(dirs).cuda(), nper, rgb, ri)
    return rgb.cpu().numpy(), ri.cpu().numpy()


@pytest.mark.parametrize("direct", [True, False])
def test_final_gather_kd_matches_oracle(device, fg, direct):
    """On ENCORE's kd-tree every sample ray hits what the reference hits, so
    the result is the reference's bit for bit except where a photon tied with
    the k-th distance is chosen differently."""
    f = fg
    ns = 4
    dirs = scenes.hemisphere_dirs(f["n"], ns, scenes.Stream(31))
    log("case: final gather kd, %d points x %d samples, direct=%s" % (len(f["p"]), ns * ns, direct))
    sc = rq.Scene(device, f["xyz"], f["idx"], accel="kd")
    g_rgb, g_ri = _engine(sc, f, f["p"], dirs, ns * ns, direct)
    o_rgb, o_ri = oracle.final_gather(f["xyz"], f["idx"], f["kd"], f["ind_o"], K_IND, R_IND,
                                      f["dir_o"] if direct else None, K_DIR, R_DIR, f["diffuse"], f["p"], dirs,
                                      ns * ns)
    same = np.all(g_rgb.view(np.uint32) == o_rgb.view(np.uint32), axis=1)
    log("bit-equal points: %d of %d" % (same.sum(), len(same)))
    assert np.array_equal(g_ri.view(np.uint32), o_ri.view(np.uint32))
    assert same.mean() >= 0.995
    assert np.allclose(g_rgb, o_rgb, rtol=2e-3, atol=1e-6)
    assert (o_rgb > 0).any(axis=1).mean() > 0.9  # the test exercises real gathers
    sc.destroy()


@pytest.mark.parametrize("width", [2, 4])
def test_final_gather_bvh_matches_oracle_up_to_ties(device, fg, width):
    """On a BVH the sample rays' closest hits equal the kd-tree's except ties
    and the reference's flush phantoms (sample rays start on the surface, a
    few hundredths from 0.03-unit triangles, where Cross()'s 1e-3 flush makes
    ENCORE's predicate hit points outside the triangle's own box).  Every
    point whose sample rays all hit what the kd-tree hits gets the oracle's
    result bit for bit; every other point has at least one sample ray in one
    of those classes."""
    f = fg
    ns = 4
    nper = ns * ns
    dirs = scenes.hemisphere_dirs(f["n"], ns, scenes.Stream(37))
    log("case: final gather bvh%d" % width)
    sc = rq.Scene(device, f["xyz"], f["idx"], accel="bvh", width=width)
    g_rgb, g_ri = _engine(sc, f, f["p"], dirs, nper)
    o_rgb, o_ri = oracle.final_gather(f["xyz"], f["idx"], f["kd"], f["ind_o"], K_IND, R_IND, f["dir_o"], K_DIR,
                                      R_DIR, f["diffuse"], f["p"], dirs, nper)
    rays = np.zeros((len(f["p"]) * nper, 8), np.float32)
    rays[:, 0:3] = np.repeat(f["p"], nper, axis=0)
    rays[:, 4:7] = dirs
    rays[:, 7] = np.inf
    d_rays = torch.from_numpy(rays).cuda()
    hb = torch.empty((len(rays), 4), device="cuda")
    sc.intersect(d_rays, hb)
    hb = hb.cpu().numpy()
    hk = f["kd"].intersect(rays)
    c = classify_closest(f["xyz"], f["idx"], rays, hb, hk)
    if c["bad"]:
        bad = np.array(c["bad_idx"])
        c2 = classify_closest(f["xyz"], f["idx"], rays[bad], hb[bad], hk[bad],
                              brute=oracle.brute_closest(f["xyz"], f["idx"], rays[bad]))
        c["bad"] = c2["bad"]
    differ = (~np.all(hb.view(np.uint32) == hk.view(np.uint32), axis=1)).reshape(-1, nper).any(axis=1)
    same = np.all(g_rgb.view(np.uint32) == o_rgb.view(np.uint32), axis=1) & (
        g_ri.view(np.uint32) == o_ri.view(np.uint32))
    log("bvh%d: %d points bit-equal, %d points with a differing sample ray; rays %s" % (
        width, same.sum(), differ.sum(), {k: v for k, v in c.items() if k != "bad_idx"}))
    assert c["bad"] == 0
    assert np.all(same[~differ])
    assert differ.mean() < 0.15

R_PRE = 0.5  # the precomputed branch's GetNearestNPhotons(1, 10, ...), at 1/20


def test_final_gather_precomputed_irradiance_kd_matches_oracle(device, fg):
    """ENCORE's default final gather (m_bPrecomputeIrr, set by the
    constructor and by pmConfig.txt; PhotonMapperCPU.cpp:1129-1140): the
    indirect term at a sample hit is the power of the nearest photon of the
    precomputed map within 10 (GetNearestNPhotons(1, 10, ...)), black when
    there is none.  For a map that is not a caustic map, GetRadianceEstimate
    with N = 1 returns that photon's power exactly (PhotonMap.cpp:121-164: one
    power times the weight 1 in double, rounded back to float, density 1), so
    rq_final_gather with k_ind = 1 on the precomputed map is that branch.  The
    map is made on the GPU (rq_photons_precompute_irradiance, ENCORE's every
    photon: PrecompIrrValue 1) and the oracle gathers over the same photons and
    powers; on the kd scene every result is the reference's bit for bit
    except where two photons tie for nearest."""
    f = fg
    made, chosen, power = f["ind_g"].precompute_irradiance(f["ind_pos"], f["ind_nrm"], 1, K_IND, R_IND,
                                                           layout="encore")
    pre_o = oracle.PhotonMap(np.ascontiguousarray(f["ind_pos"][chosen]), np.ascontiguousarray(f["ind_nrm"][chosen]),
                             power)
    ns = 4
    dirs = scenes.hemisphere_dirs(f["n"], ns, scenes.Stream(43))
    log("case: final gather, precomputed irradiance, kd, %d points x %d samples, %d photons" % (
        len(f["p"]), ns * ns, len(chosen)))
    sc = rq.Scene(device, f["xyz"], f["idx"], accel="kd")
    rgb = torch.full((len(f["p"]), 3), -1.0, device="cuda")
    ri = torch.full((len(f["p"]),), -1.0, device="cuda")
    sc.final_gather(made, 1, R_PRE, f["dir_g"], K_DIR, R_DIR, torch.from_numpy(f["diffuse"]).cuda(),
                    torch.from_numpy(f["p"]).cuda(), torch.from_numpy(dirs).cuda(), ns * ns, rgb, ri)
    g_rgb, g_ri = rgb.cpu().numpy(), ri.cpu().numpy()
    o_rgb, o_ri = oracle.final_gather(f["xyz"], f["idx"], f["kd"], pre_o, 1, R_PRE, f["dir_o"], K_DIR, R_DIR,
                                      f["diffuse"], f["p"], dirs, ns * ns)
    same = np.all(g_rgb.view(np.uint32) == o_rgb.view(np.uint32), axis=1)
    log("bit-equal points: %d of %d" % (same.sum(), len(same)))
    assert np.array_equal(g_ri.view(np.uint32), o_ri.view(np.uint32))
    assert same.mean() >= 0.995
    assert np.allclose(g_rgb, o_rgb, rtol=2e-3, atol=1e-6)
    # the indirect term alone: the nearest photon's power, bit for bit
    rgb_i = torch.full((len(f["p"]), 3), -1.0, device="cuda")
    sc.final_gather(made, 1, R_PRE, None, K_DIR, R_DIR, torch.from_numpy(f["diffuse"]).cuda(),
                    torch.from_numpy(f["p"]).cuda(), torch.from_numpy(dirs).cuda(), ns * ns, rgb_i)
    o_i, _ = oracle.final_gather(f["xyz"], f["idx"], f["kd"], pre_o, 1, R_PRE, None, K_DIR, R_DIR, f["diffuse"],
                                 f["p"], dirs, ns * ns)
    same_i = np.all(rgb_i.cpu().numpy().view(np.uint32) == o_i.view(np.uint32), axis=1)
    assert same_i.mean() >= 0.995
    assert (o_i > 0).any(axis=1).mean() > 0.9  # the test exercises real lookups
    sc.destroy()
    made.destroy()


def test_final_gather_chunking_is_invisible(device, fg):
    """A call whose sample rays span several 4M-ray chunks returns exactly what
    separate calls on the pieces return."""
    f = fg
    ns = 32  # 1024 samples per point: 4096 points per chunk
    pts = np.tile(f["p"], (5, 1))[:4500]
    nrm = np.tile(f["n"], (5, 1))[:4500]
    dirs = scenes.hemisphere_dirs(nrm, ns, scenes.Stream(41))
    log("case: final gather chunking, %d points x %d" % (len(pts), ns * ns))
    sc = rq.Scene(device, f["xyz"], f["idx"], accel="bvh")
    a_rgb, a_ri = _engine(sc, f, pts, dirs, ns * ns)
    cut = 1234
    b1 = _engine(sc, f, pts[:cut], dirs[:cut * ns * ns], ns * ns)
    b2 = _engine(sc, f, pts[cut:], dirs[cut * ns * ns:], ns * ns)
    assert np.array_equal(a_rgb.view(np.uint32), np.concatenate([b1[0], b2[0]]).view(np.uint32))
    assert np.array_equal(a_ri.view(np.uint32), np.concatenate([b1[1], b2[1]]).view(np.uint32))
    sc.destroy()


def test_final_gather_rejects_bad_arguments(device, fg):
    f = fg
    sc = rq.Scene(device, f["xyz"], f["idx"], accel="bvh")
    rgb = torch.zeros((4, 3), device="cuda")
    with pytest.raises(rq.RQError) as e:
        sc.final_gather(f["ind_g"], K_IND, R_IND, None, K_DIR, R_DIR, torch.zeros(3, device="cuda"),
                        torch.zeros((4, 3), device="cuda"), torch.zeros((4, 3), device="cuda"), 0, rgb)
    assert e.value.status == 1
    with pytest.raises(rq.RQError) as e:
        sc.final_gather(f["ind_g"], 0, R_IND, None, K_DIR, R_DIR, torch.zeros((len(f["idx"]), 3), device="cuda"),
                        torch.zeros((4, 3), device="cuda"), torch.zeros((4, 3), device="cuda"), 1, rgb)
    assert e.value.status == 1
    sc.destroy()
