"""Radiance estimates without neighbour lists at large k, when no photon
cell grid serves them, run one query per wave walking the tree
(knn_wave_kernel).  They must equal the per-lane heap kernel's estimate (the
same k-nearest set up to photons tied with the k-th distance) and the
oracle's GetRadianceEstimate.  RQ_KNN_CELLS=0 keeps the grid out for the
whole module."""
import os

import numpy as np
import pytest
import torch

import oracle
from encore_amd import rq, scenes
from helpers import log

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _tree_kernels():
    os.environ["RQ_KNN_CELLS"] = "0"
    yield
    del os.environ["RQ_KNN_CELLS"]


@pytest.fixture(scope="module")
def maps(device):
    xyz, idx = scenes.config4_scene(tess=16)
    kd = oracle.KdTree(xyz, idx)
    P, N, W = scenes.trace_photons(xyz, idx, kd.intersect, 300000, batch=65536)
    Q, QN = scenes.photon_queries(xyz, idx, kd.intersect, res=120)
    pm = rq.PhotonMap(device, P, N, W)
    yield P, N, W, Q, QN, pm
    pm.destroy()


def _radiance(pm, Q, QN, k, r, lists, caustic):
    q = torch.from_numpy(np.ascontiguousarray(Q, np.float32)).cuda()
    qn = torch.from_numpy(np.ascontiguousarray(QN, np.float32)).cuda()
    rgb = torch.full((len(Q), 3), -1.0, device="cuda")
    cnt = torch.full((len(Q),), -1, dtype=torch.int32, device="cuda")
    idx = torch.empty((len(Q), k), dtype=torch.int32, device="cuda") if lists else None
    d2 = torch.empty((len(Q), k), dtype=torch.float32, device="cuda") if lists else None
    pm.radiance(q, qn, k, r, rgb, idx, d2, cnt, caustic=caustic)
    torch.cuda.synchronize()
    return rgb.cpu().numpy(), cnt.cpu().numpy()


@pytest.mark.parametrize("k,r", [(32, 0.05), (100, 0.15), (100, 0.05), (256, 0.3)])
@pytest.mark.parametrize("caustic", [False, True])
def test_wave_radiance_equals_lane_kernel_and_oracle(maps, k, r, caustic):
    P, N, W, Q, QN, pm = maps
    log("case: wave radiance k=%d r=%g caustic=%s, %d queries" % (k, r, caustic, len(Q)))
    lane_rgb, lane_cnt = _radiance(pm, Q, QN, k, r, True, caustic)
    wave_rgb, wave_cnt = _radiance(pm, Q, QN, k, r, False, caustic)
    assert np.array_equal(wave_cnt, lane_cnt)
    assert not np.isnan(wave_rgb).any()
    same = np.all(wave_rgb.view(np.uint32) == lane_rgb.view(np.uint32), axis=1)
    log("bit-equal %d of %d, mean count %.1f" % (same.sum(), len(same), wave_cnt.mean()))
    assert same.mean() >= 0.99
    np.testing.assert_allclose(wave_rgb, lane_rgb, rtol=1e-5, atol=1e-12)
    ref = oracle.PhotonMap(P, N, W).radiance(Q[:4000], QN[:4000], k, r, caustic=caustic)
    np.testing.assert_allclose(wave_rgb[:4000], ref, rtol=1e-5, atol=1e-12)
    assert (wave_cnt > 0).mean() > 0.5


def test_wave_radiance_edge_cases(device):
    pos = np.array([[0, 0, 0], [0.01, 0, 0], [0.01, 0, 0], [0.02, 0, 0], [5, 5, 5]], np.float32)
    nrm = np.array([[0, 1, 0], [-0.0, 1, 0], [0, 1, 0], [np.nan, 1, 0], [0, 1, 0]], np.float32)
    pm = rq.PhotonMap(device, pos, nrm, np.ones_like(pos))
    Q = np.array([[0, 0, 0], [0, 0, 0], [100, 100, 100], [0.01, 0, 0]], np.float32)
    QN = np.array([[0, 1, -0.0], [1, 0, 0], [0, 1, 0], [np.nan, 1, 0]], np.float32)
    rgb, cnt = _radiance(pm, Q, QN, 64, 1.0, False, False)
    assert cnt.tolist() == [3, 0, 0, 0]
    assert np.array_equal(rgb[0], np.float32([3, 3, 3])) and np.all(rgb[1:] == 0)
    pm0 = rq.PhotonMap(device, np.zeros((0, 3), np.float32), np.zeros((0, 3), np.float32), np.zeros((0, 3), np.float32))
    rgb, cnt = _radiance(pm0, Q, QN, 64, 1.0, False, False)
    assert cnt.tolist() == [0, 0, 0, 0] and np.all(rgb == 0)
    pm.destroy()
    pm0.destroy()


def test_wave_radiance_huge_distances(device):
    """Squared distances near the top of the float range, many at or above
    2^127 (keys 0x7F......, the last bucket of the radix selection's first
    pass): photons ~1e19 from the queries, r2 = fl(1.8e19^2) = 3.24e38, some
    photons beyond float range (d2 = inf, never selected).  Counts and
    estimates equal the heap kernel's and the oracle's."""
    rng = np.random.default_rng(11)
    P = rng.uniform(-1.15e19, 1.15e19, (3000, 3)).astype(np.float32)
    N = np.tile(np.float32([0, 1, 0]), (len(P), 1))
    W = rng.uniform(0.1, 1.0, (len(P), 3)).astype(np.float32)
    Q = rng.uniform(-2e18, 2e18, (512, 3)).astype(np.float32)
    QN = np.tile(np.float32([0, 1, 0]), (len(Q), 1))
    pm = rq.PhotonMap(device, P, N, W)
    r = 1.8e19
    d2 = ((P[None, :, :].astype(np.float64) - Q[:, None, :]) ** 2).sum(-1)
    assert (d2 >= 2.0 ** 127).mean() > 0.2  # keys with top byte 0x7F are common
    for k in (32, 100):
        log("case: wave radiance at huge distances k=%d" % k)
        wave_rgb, wave_cnt = _radiance(pm, Q, QN, k, r, False, False)
        lane_rgb, lane_cnt = _radiance(pm, Q, QN, k, r, True, False)
        ref = oracle.PhotonMap(P, N, W).radiance(Q, QN, k, r)
        assert np.array_equal(wave_cnt, lane_cnt) and (wave_cnt == k).mean() > 0.9
        np.testing.assert_allclose(wave_rgb, lane_rgb, rtol=1e-5)
        np.testing.assert_allclose(wave_rgb, ref, rtol=1e-5)
    pm.destroy()
