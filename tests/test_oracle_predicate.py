"""The oracle's ray/triangle predicate against ENCORE's own code and against
closed forms (TrianglePrim.cpp:204-249, Vector3.h:148-162, Vector3.cpp:105-140)."""
import os

import numpy as np
import pytest

import oracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "predicate_golden.npz")


@pytest.fixture(scope="module")
def golden():
    return np.load(GOLDEN)


def _hitpoint(verts, a, b):
    v = verts.reshape(-1, 3, 3)
    e1 = (v[:, 1] - v[:, 0]).astype(np.float32)
    e2 = (v[:, 2] - v[:, 0]).astype(np.float32)
    p = (v[:, 0] + e1 * a[:, None]).astype(np.float32)
    return (p + e2 * b[:, None]).astype(np.float32)


def test_predicate_matches_reference_golden(golden):
    r = oracle.tri_intersect_pairs(golden["tri_rays"], golden["tri_verts"])
    hit = r[:, 3] > 0
    assert np.array_equal(hit, golden["tri_hit"] > 0)
    # t is bit-identical; alpha/beta are pinned through the reference's
    # hitPoint = v0 + (v1-v0) alpha + (v2-v0) beta.
    assert np.array_equal(r[hit, 0].view(np.uint32), golden["tri_out"][hit, 0].view(np.uint32))
    hp = _hitpoint(golden["tri_verts"][hit], r[hit, 1], r[hit, 2])
    assert np.array_equal(hp.view(np.uint32), golden["tri_out"][hit, 1:4].view(np.uint32))
    # every case family contributes (flush band, edges, slivers, grazing)
    kinds = golden["tri_kind"][hit]
    for k in (0, 1, 2, 3, 6, 7):
        assert (kinds == k).sum() > 0


def test_cross_and_normalize_match_reference_golden(golden):
    c = np.array([oracle.cross(a, b) for a, b in zip(golden["cross_a"], golden["cross_b"])])
    assert np.array_equal(c.view(np.uint32), golden["cross_out"].view(np.uint32))
    n = np.array([oracle.normalize(a) for a in golden["norm_in"]])
    assert np.array_equal(n.view(np.uint32), golden["norm_out"].view(np.uint32))


def _mt_noflush(rays, verts):
    """Plain float32 Moller-Trumbore without the flush (a plausible mistake)."""
    o, d = rays[:, 0:3], rays[:, 4:7]
    v = verts.reshape(-1, 3, 3)
    e1 = (v[:, 1] - v[:, 0]).astype(np.float32)
    e2 = (v[:, 2] - v[:, 0]).astype(np.float32)
    p = np.cross(d, e2).astype(np.float32)
    inv = np.float32(1) / np.sum(e1 * p, axis=1, dtype=np.float32)
    s = (o - v[:, 0]).astype(np.float32)
    q = np.cross(s, e1).astype(np.float32)
    t = np.sum(e2 * q, axis=1, dtype=np.float32) * inv
    return t


def test_golden_vectors_reject_unflushed_predicate(golden):
    # The flush band family must expose a predicate that skips Cross()'s flush.
    k = golden["tri_kind"] == 2
    hit = golden["tri_hit"] > 0
    t = _mt_noflush(golden["tri_rays"][k & hit], golden["tri_verts"][k & hit])
    ref = golden["tri_out"][k & hit, 0]
    assert (t.view(np.uint32) != ref.view(np.uint32)).sum() > 0


def test_closed_form_axis_aligned_wall():
    # Floor quad of the Cornell box at y = -50; vertical rays hit at t = oy + 50
    v0, v1, v2 = [-50, -50, -50], [-50, -50, 50], [50, -50, 50]
    for ox, oz, oy in [(-10.0, 20.0, 0.0), (-30.5, 3.25, 12.0), (-49.0, 49.0, 100.0)]:  # z > x half
        hit, t, a, b = oracle.tri_intersect([ox, oy, oz], [0, -1, 0], v0, v1, v2)
        assert hit
        assert t == np.float32(oy + 50.0)
        # alpha weights v1 - v0 = (0,0,100), beta weights v2 - v0 = (100,0,100)
        assert np.isclose(b, (ox + 50) / 100.0, atol=1e-6)
        assert np.isclose(a + b, (oz + 50) / 100.0, atol=1e-6)


def test_special_cases():
    v0, v1, v2 = [0, 0, 0], [1, 0, 0], [0, 1, 0]
    # parallel ray: det = 0 -> inv = inf -> NaN/inf comparisons -> miss
    assert not oracle.tri_intersect([0.2, 0.2, 1], [1, 0, 0], v0, v1, v2)[0]
    # origin on the plane: t = 0 is not a hit (strict t > 0)
    assert not oracle.tri_intersect([0.2, 0.2, 0], [0, 0, -1], v0, v1, v2)[0]
    # behind the origin: miss
    assert not oracle.tri_intersect([0.2, 0.2, -1], [0, 0, -1], v0, v1, v2)[0]
    # exact vertex and edge hits are inclusive
    for p in ([0, 0], [1, 0], [0, 1], [0.5, 0.5], [0.25, 0]):
        assert oracle.tri_intersect([p[0], p[1], 2], [0, 0, -1], v0, v1, v2)[0], p
    # just outside the hypotenuse
    assert not oracle.tri_intersect([0.5, 0.5000001, 2], [0, 0, -1], v0, v1, v2)[0]


def test_flush_kills_small_triangles():
    # A 1e-4 triangle: every Cross() component is below 1e-3 and flushes to 0,
    # so the determinant is 0 and the reference reports a miss.
    s = 1e-4
    assert not oracle.tri_intersect([s / 4, s / 4, 1], [0, 0, -1], [0, 0, 0], [s, 0, 0], [0, s, 0])[0]
    # the same triangle scaled by 1e4 is hit
    assert oracle.tri_intersect([0.25, 0.25, 1], [0, 0, -1], [0, 0, 0], [1, 0, 0], [0, 1, 0])[0]


def test_cross_flush_threshold():
    # components strictly inside (-1e-3, 1e-3) flush; 1e-3 itself does not
    c = oracle.cross([1, 0, 0], [0, 0.00099999, 0])
    assert c[2] == 0
    c = oracle.cross([1, 0, 0], [0, 0.001, 0])
    assert c[2] == np.float32(0.001)
    c = oracle.cross([1, 0, 0], [0, -0.0009, 0])
    assert c[2] == 0 and not np.signbit(c[2])
