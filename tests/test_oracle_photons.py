"""The oracle's photon map (PhotonMap.cpp:167-297) against its invariants,
brute force and closed forms; the engine's host balance against the oracle."""
import numpy as np
import pytest

import oracle
from encore_amd import rq, scenes


def _random_photons(n, seed, nclasses=3):
    rng = np.random.default_rng(seed)
    pos = rng.random((n, 3)).astype(np.float32)
    pos[: n // 5, 1] = np.float32(0.5)  # a plane: many equal coordinates
    normals = np.array([[0, 1, 0], [1, 0, 0], [0, 0, -1]], np.float32)[:nclasses]
    nrm = normals[rng.integers(0, nclasses, n)]
    pw = rng.random((n, 3)).astype(np.float32)
    return pos, nrm, pw


def brute_knn(pos, nrm, q, qn, k, r):
    r2 = np.float32(r) * np.float32(r)
    d = (q[None, :] - pos).astype(np.float32)
    d2 = ((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]).astype(np.float32)
    ok = (d2 < r2) & np.all(nrm == qn[None, :], axis=1)
    ids = np.nonzero(ok)[0]
    order = np.argsort(d2[ids], kind="stable")
    return ids[order][:k], d2[ids][order][:k]


def same_knn_set(ids_a, d2_a, ids_b, d2_b):
    """Equal counts and distances; equal ids strictly inside the k-th distance."""
    if len(ids_a) != len(ids_b):
        return False
    if len(ids_a) == 0:
        return True
    if not np.array_equal(np.sort(d2_a).view(np.uint32), np.sort(d2_b).view(np.uint32)):
        return False
    kth = max(d2_a.max(), d2_b.max())
    return set(ids_a[d2_a < kth].tolist()) == set(ids_b[d2_b < kth].tolist())


@pytest.mark.parametrize("n", [1, 2, 3, 17, 1000, 20000])
def test_balanced_tree_invariant(n):
    pos, nrm, _ = _random_photons(n, 3)
    pm = oracle.PhotonMap(pos, nrm)
    perm, axis = pm.order()
    assert sorted(perm.tolist()) == list(range(n))
    P = pos[perm]
    stack = [(0, n - 1)]
    while stack:
        lo, hi = stack.pop()
        if hi < lo:
            continue
        m = (lo + hi) // 2
        a = axis[m]
        assert np.all(P[lo:m, a] <= P[m, a]) and np.all(P[m + 1:hi + 1, a] >= P[m, a])
        stack += [(lo, m - 1), (m + 1, hi)]
    if n > 1:
        assert axis[(n - 1) // 2] == 2  # the +-1e10 bookkeeping box picks z first


@pytest.mark.parametrize("n", [1, 2, 7, 1000, 50000])
def test_engine_balance_is_encore_order(n):
    pos, nrm, _ = _random_photons(n, 5)
    p1, a1 = rq.photon_balance_host(pos, "encore")
    p2, a2 = oracle.PhotonMap(pos, nrm).order()
    assert np.array_equal(p1, p2) and np.array_equal(a1, a2)


@pytest.mark.parametrize("k,r", [(1, 0.05), (8, 0.05), (50, 0.08), (50, 0.01), (200, 2.0)])
def test_knn_equals_brute_force(k, r):
    pos, nrm, _ = _random_photons(20000, 7)
    pm = oracle.PhotonMap(pos, nrm)
    perm, _ = pm.order()
    rng = np.random.default_rng(11)
    q = rng.random((300, 3)).astype(np.float32)
    q[:50] = pos[rng.integers(0, len(pos), 50)]  # queries on photons: d2 = 0
    qn = nrm[rng.integers(0, len(pos), 300)]
    qn[::37] = [0, -1, 0]  # a normal no photon has
    slots, d2, cnt = pm.knn(q, qn, k, r)
    for i in range(len(q)):
        bi, bd = brute_knn(pos, nrm, q[i], qn[i], k, r)
        got = perm[slots[i, :cnt[i]]]
        assert same_knn_set(got, d2[i, :cnt[i]], bi, bd), i


def test_knn_normal_equality_semantics():
    pos = np.array([[0, 0, 0], [0.01, 0, 0], [0.02, 0, 0], [0.03, 0, 0]], np.float32)
    nrm = np.array([[0, 1, 0], [-0.0, 1, 0], [0, 1, 1e-7], [np.nan, 1, 0]], np.float32)
    pm = oracle.PhotonMap(pos, nrm)
    perm, _ = pm.order()
    slots, d2, cnt = pm.knn(np.zeros((1, 3), np.float32), np.array([[0, 1, -0.0]], np.float32), 4, 1.0)
    # -0 == +0 matches; 1e-7 differs; NaN never matches
    assert cnt[0] == 2 and sorted(perm[slots[0, :2]].tolist()) == [0, 1]


def test_radiance_estimate_closed_form():
    pos, nrm, pw = _random_photons(5000, 9, nclasses=1)
    pm = oracle.PhotonMap(pos, nrm, pw)
    perm, _ = pm.order()
    q = np.array([[0.5, 0.5, 0.5], [0.1, 0.9, 0.2]], np.float32)
    qn = nrm[:2]
    k, r = 20, 0.3
    slots, d2, cnt = pm.knn(q, qn, k, r)
    rgb = pm.radiance(q, qn, k, r)
    rgbc = pm.radiance(q, qn, k, r, caustic=True)
    for i in range(len(q)):
        ids = perm[slots[i, :cnt[i]]]
        dd = d2[i, :cnt[i]].astype(np.float64)
        np.testing.assert_allclose(rgb[i], pw[ids].astype(np.float64).sum(0), rtol=1e-5)
        rmax = dd.max()
        w = (rmax - dd) / dd
        ref = (pw[ids].astype(np.float64) * w[:, None]).sum(0) * (k / (np.float64(np.float32(3.14159)) * rmax * rmax))
        np.testing.assert_allclose(rgbc[i], ref, rtol=1e-5)


def test_single_photon_estimate_is_its_power():
    """ENCORE's default final gather reads the nearest photon of the
    precomputed map (GetNearestNPhotons(1, ...)[0].Power(),
    PhotonMapperCPU.cpp:1129-1140).  GetRadianceEstimate with N = 1 on a map
    that is not a caustic map adds that one power times the weight 1 in double
    and rounds it back to float, with density 1 (PhotonMap.cpp:121-164), so
    the two are the same bits; with no photon in range both are black."""
    pos, nrm, pw = _random_photons(20000, 12, nclasses=3)
    pm = oracle.PhotonMap(pos, nrm, pw)
    perm, _ = pm.order()
    rng = np.random.default_rng(13)
    q = rng.uniform(0.0, 1.0, (3000, 3)).astype(np.float32)
    qn = nrm[rng.integers(0, len(nrm), len(q))]
    for r in (0.5, 0.004):
        rgb = pm.radiance(q, qn, 1, r)
        slots, _, cnt = pm.knn(q, qn, 1, r)
        found = cnt > 0
        assert np.array_equal(rgb[found].view(np.uint32), pw[perm[slots[found, 0]]].view(np.uint32))
        assert np.all(rgb[~found] == 0.0)
        if r < 0.01:
            assert (~found).sum() > 100  # both outcomes occur


def test_photon_workload_generator():
    xyz, idx = scenes.config4_scene()
    assert len(idx) == 34816
    kd = oracle.KdTree(xyz, idx)
    P, N, W = scenes.trace_photons(xyz, idx, kd.intersect, 20000, batch=10000)
    assert P.shape == (20000, 3) and not np.isnan(N).any()
    assert np.all(np.abs(P) <= 2.5 + 1e-3)
