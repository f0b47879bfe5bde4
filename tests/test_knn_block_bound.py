"""The photon cell gather (encore_amd/csrc/knn.hip, knn_cell_kernel) ranks and
skips 64-photon blocks by a squared distance from the query to the block's
box (the float min and max of its photons' positions), computed in float
with the operations of a photon's own d2: per axis the difference to the
nearer face, squared, summed as (x^2 + y^2) + z^2.  The kernel's comment
argues that rounding is monotonic, so no photon of the block has a smaller
float d2; a block is skipped when that bound reaches the current k-th
distance, so a bound above a photon's d2 could drop a neighbour.  This checks
the claim in float32 over random blocks, with photons on the faces, edges and
corners nearest the query and queries within a few ulps of the faces.  The
plausible alternative -- the box distance in float64, rounded to nearest --
exceeds some photon's float d2 on the same cases."""
import numpy as np

f = np.float32


def _photon_d2(q, p):
    x, y, z = q[..., 0] - p[..., 0], q[..., 1] - p[..., 1], q[..., 2] - p[..., 2]
    return (x * x + y * y) + z * z


def _box_bound(q, lo, hi):
    d = np.maximum(np.maximum(lo - q, q - hi), f(0))
    return (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]


def _box_bound_f64_nearest(q, lo, hi):
    q64, lo64, hi64 = q.astype(np.float64), lo.astype(np.float64), hi.astype(np.float64)
    d = np.maximum(np.maximum(lo64 - q64, q64 - hi64), 0.0)
    return ((d * d).sum(-1)).astype(f)


def _cases(seed, nblocks=20000, per=16):
    """Blocks of `per` photons; queries near the boxes; photons placed where
    their float d2 comes closest to the box distance."""
    rng = np.random.default_rng(seed)
    scale = f(10.0) ** rng.uniform(-3, 2, (nblocks, 1, 1)).astype(f)
    centre = (rng.uniform(-1, 1, (nblocks, 1, 3)) * scale * 50).astype(f)
    pts = (centre + rng.uniform(-1, 1, (nblocks, per, 3)).astype(f) * scale).astype(f)
    lo, hi = pts.min(1), pts.max(1)
    # the photon nearest each face point: clamp a query direction onto the box
    side = rng.integers(0, 3, (nblocks, 3))  # below, inside, above per axis
    gap = (f(10.0) ** rng.uniform(-7, 0, (nblocks, 3)).astype(f)) * scale[:, 0]
    q = np.where(side == 0, lo - gap, np.where(side == 2, hi + gap, (lo + hi) * f(0.5))).astype(f)
    # queries a few ulps outside a face
    ulps = rng.integers(1, 4, (nblocks, 3)).astype(f)
    near = rng.random(nblocks) < 0.3
    q_ulp = np.where(side == 0, np.nextafter(lo, f(-np.inf)) - (ulps - 1) * np.spacing(lo),
                     np.where(side == 2, np.nextafter(hi, f(np.inf)) + (ulps - 1) * np.spacing(hi), q)).astype(f)
    q = np.where(near[:, None], q_ulp, q).astype(f)
    # photons on the nearest face, edge or corner of the box (moved onto it)
    closest = np.clip(q, lo, hi)
    k = rng.integers(0, per, nblocks)
    pts[np.arange(nblocks), k] = closest
    jitter = rng.integers(0, per, nblocks)
    on_face = pts[np.arange(nblocks), jitter].copy()
    axis = rng.integers(0, 3, nblocks)
    on_face[np.arange(nblocks), axis] = closest[np.arange(nblocks), axis]
    on_face = np.clip(on_face, lo, hi)
    pts[np.arange(nblocks), jitter] = on_face
    lo, hi = pts.min(1), pts.max(1)
    return q.astype(f), pts.astype(f), lo.astype(f), hi.astype(f)


def test_block_bound_never_exceeds_a_photon_d2():
    for seed in range(6):
        q, pts, lo, hi = _cases(seed)
        b = _box_bound(q, lo, hi)
        d2 = _photon_d2(q[:, None, :], pts).min(1)
        assert b.dtype == f and d2.dtype == f
        assert np.all(b <= d2), int((b > d2).sum())
        # the cases reach the bound: many blocks hold a photon at exactly it
        assert (b == d2).mean() > 0.5


def test_rounded_to_nearest_bound_fails():
    """The check has teeth: the float64 box distance rounded to nearest is
    above the float d2 of some photon of its block."""
    bad = 0
    for seed in range(6):
        q, pts, lo, hi = _cases(seed)
        b = _box_bound_f64_nearest(q, lo, hi)
        d2 = _photon_d2(q[:, None, :], pts).min(1)
        bad += int((b > d2).sum())
    assert bad > 0
