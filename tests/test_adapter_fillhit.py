"""The drop-in adapter's HitInfo arithmetic (encore_amd/adapter/HitFill.h,
used by HipAccel::fillHit) returns the reference's bits.

For every hit among the 3,600 golden vectors of the reference's own
TrianglePrim::intersect (tests/golden/predicate_golden.npz: hitPoint,
hitNormal, isEntering; TrianglePrim.cpp:226-244), the adapter's function,
compiled on ENCORE's own Point3 / Vector3 sources (oracle/_ref), is fed the
engine's (t, alpha, beta) -- the predicate's bits, which the GPU parity
tests hold equal to the oracle's -- and must reproduce hitPoint, hitNormal
and isEntering bit for bit.  A swapped alpha/beta, a dropped flush or an
unnormalized normal fails it."""
import ctypes
import os

import numpy as np
import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "libref_fillhit.so")

pytestmark = pytest.mark.skipif(not os.path.exists(LIB),
                                reason="oracle/_ref is built only where the ENCORE sources are present")


def _fill(lib, verts, rays, tuv):
    f = lib.ref_fill_hit
    p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
    f.argtypes = [p] * 6
    f.restype = None
    out = np.zeros((len(rays), 7), np.float32)
    for i in range(len(rays)):
        row = np.zeros(7, np.float32)
        f(np.ascontiguousarray(verts[i, 0:3]), np.ascontiguousarray(verts[i, 3:6]),
          np.ascontiguousarray(verts[i, 6:9]), np.ascontiguousarray(rays[i, 4:7]),
          np.ascontiguousarray(tuv[i]), row)
        out[i] = row
    return out


def test_fill_hit_equals_reference_hitinfo():
    g = np.load(os.path.join(HERE, "golden", "predicate_golden.npz"))
    rays, verts, out, hit = g["tri_rays"], g["tri_verts"], g["tri_out"], g["tri_hit"].astype(bool)
    h = oracle.tri_intersect_pairs(rays, verts)
    assert np.array_equal(h[:, 3] != 0, hit)
    sel = np.nonzero(hit)[0]
    assert len(sel) > 1000
    assert np.array_equal(h[sel, 0].view(np.uint32), out[sel, 0].view(np.uint32))  # t
    got = _fill(ctypes.CDLL(LIB), verts[sel], rays[sel], np.ascontiguousarray(h[sel, :3]))
    bad = np.nonzero(~np.all(got.view(np.uint32) == out[sel, 1:8].view(np.uint32), axis=1))[0]
    assert len(bad) == 0, (len(bad), sel[bad[:5]], got[bad[:3]], out[sel[bad[:3]], 1:8])
    # every case family contributes hits (the flush band included)
    kinds = set(g["tri_kind"][sel].tolist())
    assert {0, 1, 2, 3}.issubset(kinds)


def test_swapped_barycentrics_would_fail():
    """The comparison has teeth: feeding (t, beta, alpha) breaks hitPoint."""
    g = np.load(os.path.join(HERE, "golden", "predicate_golden.npz"))
    rays, verts, out, hit = g["tri_rays"], g["tri_verts"], g["tri_out"], g["tri_hit"].astype(bool)
    sel = np.nonzero(hit)[0][:200]
    h = oracle.tri_intersect_pairs(rays[sel], verts[sel])
    swapped = np.ascontiguousarray(h[:, [0, 2, 1]])
    got = _fill(ctypes.CDLL(LIB), verts[sel], rays[sel], swapped)
    same = np.all(got[:, 0:3].view(np.uint32) == out[sel, 1:4].view(np.uint32), axis=1)
    assert same.mean() < 0.1
