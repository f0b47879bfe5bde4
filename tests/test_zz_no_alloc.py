"""rq_intersect_closest_static (allocates nothing; its scratch is a static
device array) returns exactly what rq_intersect_closest returns.

Called through the C ABI.  Two calls in a row on each scene, the second with
fewer rays, so the second call reuses the static counter and spill stacks
the first one left behind; the hit buffer past the second call's rays must
stay untouched.  The kd scene's spill stacks need more than the static
region holds at the full grid, so its launch runs with a shrunken grid."""
import ctypes

import numpy as np
import pytest
import torch

import oracle
from encore_amd import rq, scenes
from helpers import log

pytestmark = pytest.mark.gpu


def _lib():
    L = rq.load()
    fn = L.rq_intersect_closest_static
    fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p]
    fn.restype = ctypes.c_int
    return L


@pytest.fixture(scope="module")
def sphere_rays():
    xyz, idx = scenes.displaced_sphere(level=5)
    prim = scenes.config2_primary(cols=320, rows=180)
    ph = oracle.KdTree(xyz, idx).intersect(prim)
    bounce, _ = scenes.config3_rays(xyz, idx, prim, ph, count=60000, nbatches=4)
    return xyz, idx, np.concatenate([prim[:20000], bounce])


@pytest.mark.parametrize("accel,width", [("bvh", 4), ("bvh", 2), ("kd", 0)])
def test_static_entry_matches_allocating_entry(device, sphere_rays, accel, width):
    L = _lib()
    xyz, idx, rays = sphere_rays
    kw = {"width": width} if accel == "bvh" else {}
    sc = rq.Scene(device, xyz, idx, accel=accel, **kw)
    log("case: static entry", accel, width, "rays", len(rays), "stack bound", sc.stats()["stack_bound"])
    d_rays = torch.from_numpy(np.ascontiguousarray(rays)).cuda()
    ref = torch.full((len(rays), 4), -3.0, device="cuda")
    rq.check(L.rq_intersect_closest(sc.h, d_rays.data_ptr(), len(rays), ref.data_ptr(), None))
    torch.cuda.synchronize()
    want = ref.cpu().numpy()
    for n in (len(rays), 7001):
        got = torch.full((len(rays), 4), -5.0, device="cuda")
        rq.check(L.rq_intersect_closest_static(sc.h, d_rays.data_ptr(), n, got.data_ptr(), None))
        torch.cuda.synchronize()
        g = got.cpu().numpy()
        assert np.array_equal(g[:n].view(np.uint32), want[:n].view(np.uint32)), (accel, width, n)
        assert (g[n:] == -5.0).all()
        log("  %d rays bit-equal" % n)
    sc.destroy()


def test_static_entry_argument_errors(device, sphere_rays):
    L = _lib()
    xyz, idx, _ = sphere_rays
    sc = rq.Scene(device, xyz, idx, accel="bvh")
    assert L.rq_intersect_closest_static(None, None, 0, None, None) == 1
    assert L.rq_intersect_closest_static(sc.h, None, 4, None, None) == 1
    assert L.rq_intersect_closest_static(sc.h, None, 0, None, None) == 0  # no rays: nothing to do
    sc.destroy()
