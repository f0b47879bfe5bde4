"""Radiance estimates without neighbour lists at k >= 32 read their
candidates from the photon cell grid (photon_grid.hip, knn_cell_kernel).
The grid must never miss a photon in range: counts equal brute force where
fewer than k photons are in range, whatever the offset of the map, and
photons that no query can select (NaN normals) stay out.  Estimates equal
the per-lane heap kernel's (RQ_KNN_CELLS=0 with lists) up to photons tied
with the k-th distance, and lists from the grid equal the heap kernel's as
tie-aware sets, also when one map is queried with
alternating radii (the grid is rebuilt) and when the radius is too large for
a grid (the tree walk serves)."""
import os

import numpy as np
import pytest
import torch

import oracle
from encore_amd import rq
from helpers import log
from test_gpu_knn_wave import _radiance
from test_gpu_photons import _gpu_knn
from test_oracle_photons import same_knn_set

pytestmark = pytest.mark.gpu


def _heap(fn, *args):
    """Runs fn with the cell grid switched off: lists come from the per-lane
    heap kernel."""
    os.environ["RQ_KNN_CELLS"] = "0"
    try:
        return fn(*args)
    finally:
        del os.environ["RQ_KNN_CELLS"]


def _brute_count(P, N, Q, QN, r):
    """Photons passing the class and distance tests, per query: float32
    (x*x + y*y) + z*z < fl(r*r); normals equal as floats (-0 matches 0,
    NaN matches nothing)."""
    r2 = np.float32(r) * np.float32(r)
    out = np.zeros(len(Q), np.int64)
    for i in range(len(Q)):
        d = (Q[i] - P).astype(np.float32)
        with np.errstate(invalid="ignore"):
            d2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
            out[i] = np.count_nonzero((d2 < r2) & np.all(N == QN[i], axis=1))
    return out


def _surface_photons(rng, n, offset):
    """Photons on three axis planes of a unit box (normals +x, +y, +z) and
    queries on the same planes, shifted by `offset`."""
    P = rng.uniform(0.0, 1.0, (n, 3)).astype(np.float32)
    N = np.zeros((n, 3), np.float32)
    ax = rng.integers(0, 3, n)
    P[np.arange(n), ax] = 0.0
    N[np.arange(n), ax] = 1.0
    W = rng.uniform(0.1, 1.0, (n, 3)).astype(np.float32)
    nq = 1500
    Q = rng.uniform(-0.05, 1.05, (nq, 3)).astype(np.float32)
    QN = np.zeros((nq, 3), np.float32)
    qa = rng.integers(0, 3, nq)
    Q[np.arange(nq), qa] = 0.0
    QN[np.arange(nq), qa] = 1.0
    off = np.float32(offset)
    return P + off, N, W, Q + off, QN


@pytest.mark.parametrize("table", ["dense", "hash"])
@pytest.mark.parametrize("offset", [0.0, -3.5, 1000.0])
def test_cells_find_every_photon_in_range(device, offset, table, monkeypatch):
    """Both cell tables: the dense array (the default for a box of at most
    256 MB of cells) and the hash table (RQ_KNN_GRID_HASH=1)."""
    if table == "hash":
        monkeypatch.setenv("RQ_KNN_GRID_HASH", "1")
    rng = np.random.default_rng(11)
    P, N, W, Q, QN = _surface_photons(rng, 40000, offset)
    # photons exactly on cell faces: copies of queries moved by r along an axis
    r = 0.03
    edge = Q[:200].copy()
    edge[:, 1] += np.float32(r)
    P = np.concatenate([P, edge, Q[200:260]])
    N = np.concatenate([N, QN[:200], QN[200:260]])
    W = np.concatenate([W, np.ones((260, 3), np.float32)])
    pm = rq.PhotonMap(device, P, N, W)
    k = 256
    rgb, cnt = _radiance(pm, Q, QN, k, r, False, False)
    want = _brute_count(P, N, Q, QN, r)
    log("case: cells offset %g, %s table: mean count %.1f, max %d" % (offset, table, want.mean(), want.max()))
    assert want.max() < k and want.mean() > 10
    assert np.array_equal(cnt, want)
    heap_rgb, heap_cnt = _heap(_radiance, pm, Q, QN, k, r, True, False)
    assert np.array_equal(cnt, heap_cnt)
    np.testing.assert_allclose(rgb, heap_rgb, rtol=1e-6, atol=0)
    pm.destroy()


@pytest.mark.parametrize("table", ["dense", "hash"])
@pytest.mark.parametrize("k", [32, 100, 256])
@pytest.mark.parametrize("caustic", [False, True])
def test_cells_saturated_equal_heap_kernel_and_oracle(device, k, caustic, table, monkeypatch):
    if table == "hash":
        monkeypatch.setenv("RQ_KNN_GRID_HASH", "1")
    rng = np.random.default_rng(12)
    P, N, W, Q, QN = _surface_photons(rng, 200000, 0.0)
    pm = rq.PhotonMap(device, P, N, W)
    r = 0.05  # about 2.5x k = 100 photons in range on a plane
    rgb, cnt = _radiance(pm, Q, QN, k, r, False, caustic)
    heap_rgb, heap_cnt = _heap(_radiance, pm, Q, QN, k, r, True, caustic)
    assert np.array_equal(cnt, heap_cnt)
    assert not np.isnan(rgb).any()
    same = np.all(rgb.view(np.uint32) == heap_rgb.view(np.uint32), axis=1)
    log("case: cells k=%d caustic=%s %s table: %.3f full, %d of %d bit-equal" % (k, caustic, table, (cnt == k).mean(),
                                                                                  same.sum(), len(same)))
    assert same.mean() >= 0.99
    np.testing.assert_allclose(rgb, heap_rgb, rtol=1e-5, atol=1e-12)
    ref = oracle.PhotonMap(P, N, W).radiance(Q[:500], QN[:500], k, r, caustic=caustic)
    np.testing.assert_allclose(rgb[:500], ref, rtol=1e-5, atol=1e-12)
    pm.destroy()


def test_cells_skip_unselectable_photons(device):
    """Photon positions are finite (ENCORE stores hit points): a balanced
    tree over NaN coordinates is not ordered, so only the normals vary here."""
    rng = np.random.default_rng(13)
    P, N, W, Q, QN = _surface_photons(rng, 20000, 0.0)
    N[2::13] = np.nan
    N[5::19] = [0.6, 0.8, 0.0]  # a normal no query has
    for far in (False, True):
        if far:
            P[3::17] = np.float32(1e30)  # 2^21 cells along an axis: no grid, the tree walk serves
        pm = rq.PhotonMap(device, P, N, W)
        rgb, cnt = _radiance(pm, Q, QN, 64, 0.04, False, False)
        assert np.array_equal(cnt, np.minimum(_brute_count(P, N, Q, QN, 0.04), 64))
        heap_rgb, _ = _heap(_radiance, pm, Q, QN, 64, 0.04, True, False)
        np.testing.assert_allclose(rgb, heap_rgb, rtol=1e-5)
        pm.destroy()


def test_cells_radius_changes_and_large_radius(device):
    rng = np.random.default_rng(14)
    P, N, W, Q, QN = _surface_photons(rng, 60000, 0.0)
    pm = rq.PhotonMap(device, P, N, W)
    for r in (0.03, 0.06, 0.03, 0.5, 2.0, 0.03):
        rgb, cnt = _radiance(pm, Q, QN, 64, r, False, False)
        heap_rgb, heap_cnt = _heap(_radiance, pm, Q, QN, 64, r, True, False)
        log("case: radius %g: mean count %.1f" % (r, cnt.mean()))
        assert np.array_equal(cnt, heap_cnt)
        same = np.all(rgb.view(np.uint32) == heap_rgb.view(np.uint32), axis=1)
        assert same.mean() >= 0.99
        np.testing.assert_allclose(rgb, heap_rgb, rtol=1e-5, atol=1e-12)
    pm.destroy()


@pytest.mark.parametrize("k", [32, 50, 256])
def test_cell_lists_equal_heap_kernel(device, k):
    """rq_knn and lists with the radiance: nearest first, ~0 / +inf padding,
    the heap kernel's set up to ties at the k-th distance."""
    rng = np.random.default_rng(15)
    P, N, W, Q, QN = _surface_photons(rng, 100000, 0.0)
    pm = rq.PhotonMap(device, P, N, W)
    r = 0.04
    gi, gd, gc = _gpu_knn(pm, Q, QN, k, r)
    hi, hd, hc = _heap(_gpu_knn, pm, Q, QN, k, r)
    assert np.array_equal(gc, hc)
    log("case: cell lists k=%d: mean count %.1f, %.3f full" % (k, gc.mean(), (gc == k).mean()))
    for i in range(len(Q)):
        c = int(gc[i])
        assert np.all(np.diff(gd[i, :c]) >= 0)
        assert np.all(gi[i, c:] == 0xFFFFFFFF) and np.all(np.isinf(gd[i, c:]))
        assert same_knn_set(gi[i, :c], gd[i, :c], hi[i, :c], hd[i, :c]), i
    gi2, gd2, gc2, rgb = _gpu_knn(pm, Q, QN, k, r, radiance=True)
    assert np.array_equal(gi2, gi) and np.array_equal(gd2.view(np.uint32), gd.view(np.uint32))
    rgb_only, cnt_only = _radiance(pm, Q, QN, k, r, False, False)
    assert np.array_equal(rgb.view(np.uint32), rgb_only.view(np.uint32)) and np.array_equal(gc2, cnt_only)
    pm.destroy()


def test_cells_degenerate_query_positions(device):
    """Queries at NaN, infinite or far positions find no photons on the grid
    path, as on the heap kernel's; lists and radiance stay consistent."""
    rng = np.random.default_rng(16)
    P, N, W, Q, QN = _surface_photons(rng, 20000, 0.0)
    pm = rq.PhotonMap(device, P, N, W)
    bad = np.array([[np.nan, 0.5, 0.0], [0.5, np.inf, 0.0], [-np.inf, 0.5, 0.0], [1e30, 0.5, 0.0],
                    [-1e30, -1e30, 0.0], [3e38, 3e38, 0.0], [0.5, 0.5, 1e-40]], np.float32)
    bad_n = np.tile(np.float32([0, 0, 1]), (len(bad), 1))
    Qm = np.concatenate([Q[:100], bad, Q[100:200]]).astype(np.float32)
    QNm = np.concatenate([QN[:100], bad_n, QN[100:200]]).astype(np.float32)
    rgb, cnt = _radiance(pm, Qm, QNm, 64, 0.04, False, False)
    heap_rgb, heap_cnt = _heap(_radiance, pm, Qm, QNm, 64, 0.04, True, False)
    assert np.array_equal(cnt, heap_cnt)
    assert cnt[100:106].tolist() == [0] * 6 and np.all(rgb[100:106] == 0)
    np.testing.assert_allclose(rgb, heap_rgb, rtol=1e-5)
    gi, gd, gc = _gpu_knn(pm, Qm, QNm, 64, 0.04)
    assert np.array_equal(gc, cnt) and np.all(gi[100:106] == 0xFFFFFFFF)
    pm.destroy()


def test_cells_counts(device):
    """rq_radiance_counts: the instrumented kernel returns the same
    estimates, one query each, and -- with fewer than k photons in range, so
    the candidate list is never cut -- exactly the brute-force number of
    photons in range as candidates, out of at least as many photons read
    from at most 128 ranges (cells or blocks) per query.  A k or radius the grid does not serve
    is refused."""
    rng = np.random.default_rng(13)
    P, N, W, Q, QN = _surface_photons(rng, 40000, 0.0)
    pm = rq.PhotonMap(device, P, N, W)
    k, r = 256, 0.03
    rgb, cnt = _radiance(pm, Q, QN, k, r, False, False)
    q = torch.from_numpy(Q).cuda()
    qn = torch.from_numpy(QN).cuda()
    rgb_c = torch.full((len(Q), 3), -1.0, device="cuda")
    work = pm.radiance_counts(q, qn, k, r, rgb_c)
    want = _brute_count(P, N, Q, QN, r)
    log("case: cell counts %s, in range %d" % (work, want.sum()))
    assert np.array_equal(rgb_c.cpu().numpy().view(np.uint32), rgb.view(np.uint32))
    assert work["queries"] == len(Q)
    assert work["candidates"] == int(want.sum()) == int(cnt.sum())
    assert work["photons"] >= work["candidates"] and 0 < work["ranges"] <= 128 * len(Q)
    with pytest.raises(rq.RQError) as e:
        pm.radiance_counts(q, qn, 8, r, rgb_c)
    assert e.value.status == 4
    pm.destroy()


@pytest.mark.parametrize("shape", ["plane", "dense_plane", "patch", "volume", "dense_volume"])
def test_cell_blocks_and_their_fallbacks(device, shape):
    """The 64-photon blocks of a dense grid and the two ways a query keeps its
    cells instead: plane -- cells of hundreds of photons, a query's cells hold
    tens of blocks (blocks ranked, sorted in one register); dense plane --
    cells of about 1,000, so that most queries keep 65 to 128 blocks (sorted
    in two registers); patch -- 400,000 photons in a 0.2-wide square, cells of
    thousands (more than 256 blocks: cells ranked); volume -- photons filling
    a cube, about 56 blocks within a query's sphere; dense volume -- three
    times as many: about 70 % of the queries rank 65 to 128 blocks, most of
    the others keep more than 128 within the radius and a few hold more than
    256 in their cells (cells ranked for those).  Counts, radiance and
    tie-aware lists equal the per-lane heap kernel's."""
    rng = np.random.default_rng(17)
    n, nq, r, k = 200000, 1200, 0.05, 64
    if shape in ("plane", "dense_plane"):
        n = 400000 if shape == "dense_plane" else n
        P = np.stack([rng.uniform(0, 1, n), rng.uniform(0, 1, n), np.zeros(n)], 1)
        Q = np.stack([rng.uniform(0.05, 0.95, nq), rng.uniform(0.05, 0.95, nq), np.zeros(nq)], 1)
    elif shape == "patch":
        n = 400000
        P = np.stack([rng.uniform(0.4, 0.6, n), rng.uniform(0.4, 0.6, n), np.zeros(n)], 1)
        # a sparse background keeps the grid's box wide enough for 4 cells
        P[:2000, 0:2] = rng.uniform(0, 1, (2000, 2))
        Q = np.stack([rng.uniform(0.38, 0.62, nq), rng.uniform(0.38, 0.62, nq), np.zeros(nq)], 1)
    else:
        r = 0.1  # cells of about 200 photons; about 56 blocks reach a query's sphere
        n = 600000 if shape == "dense_volume" else n
        P = rng.uniform(0, 1, (n, 3))
        Q = rng.uniform(0.1, 0.9, (nq, 3))
    P, Q = P.astype(np.float32), Q.astype(np.float32)
    N = np.tile(np.float32([0, 0, 1]), (len(P), 1))
    QN = np.tile(np.float32([0, 0, 1]), (nq, 1))
    W = rng.uniform(0.1, 1.0, (len(P), 3)).astype(np.float32)
    pm = rq.PhotonMap(device, P, N, W)
    rgb, cnt = _radiance(pm, Q, QN, k, r, False, False)
    work = pm.radiance_counts(torch.from_numpy(Q).cuda(), torch.from_numpy(QN).cuda(), k, r,
                              torch.empty((nq, 3), device="cuda"))
    heap_rgb, heap_cnt = _heap(_radiance, pm, Q, QN, k, r, True, False)
    log("case: blocks %s: %.1f ranges, %.0f photons read per query, mean count %.1f" % (
        shape, work["ranges"] / nq, work["photons"] / nq, cnt.mean()))
    assert np.array_equal(cnt, heap_cnt) and cnt.min() > 0
    np.testing.assert_allclose(rgb, heap_rgb, rtol=1e-5, atol=1e-12)
    if shape == "dense_plane":
        assert work["ranges"] > 64 * nq  # most queries rank more than 64 blocks
    gi, gd, gc = _gpu_knn(pm, Q[:300], QN[:300], k, r)
    hi, hd, hc = _heap(_gpu_knn, pm, Q[:300], QN[:300], k, r)
    assert np.array_equal(gc, hc)
    for i in range(300):
        c = int(gc[i])
        assert same_knn_set(gi[i, :c], gd[i, :c], hi[i, :c], hd[i, :c]), i
    pm.destroy()
