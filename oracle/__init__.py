"""ctypes front of the parity oracle (test infrastructure only).

The oracle is a CPU restatement of ENCORE's hot path (encore_oracle.cpp).
Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use it,
always as the checker; the HIP engine in encore_amd/ never imports it.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

NO_HIT = 0xFFFFFFFF

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_u32 = ctypes.c_uint32
_int = ctypes.c_int
_vp = ctypes.c_void_p


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built: run `make -C oracle`")
        L = ctypes.CDLL(path)
        sig = {
            "orc_tri_intersect": (_int, [_f32p, _f32p, _f32p, _f32p, _f32p, _f32p]),
            "orc_tri_intersect_pairs": (None, [_f32p, _f32p, _u32, _f32p]),
            "orc_cross": (None, [_f32p, _f32p, _f32p]),
            "orc_normalize": (None, [_f32p, _f32p]),
            "orc_brute_closest": (None, [_f32p, _u32, _u32p, _u32, _f32p, _u32, _f32p, _int]),
            "orc_count_hits_below": (None, [_f32p, _u32, _u32p, _u32, _f32p, _u32, _f32p, ctypes.c_float, _u32p]),
            "orc_fp64_eval": (None, [_f32p, _u32p, _f32p, _u32p, _u32, _f64p]),
            "orc_fp64_closest": (None, [_f32p, _u32, _u32p, _u32, _f32p, _u32, _f64p, _u32p, _int]),
            "orc_kd_build": (_vp, [_f32p, _u32, _u32p, _u32, _f32p, _f32p]),
            "orc_kd_free": (None, [_vp]),
            "orc_kd_stats": (None, [_vp, _u32p]),
            "orc_kd_export": (None, [_vp, _u32p, _u32p, _u32p, _u32p, _f32p]),
            "orc_kd_intersect": (None, [_vp, _f32p, _u32, _f32p, _int]),
            "orc_kd_occluded": (None, [_vp, _f32p, _u32, _u32p, _int]),
            "orc_pm_build": (_vp, [_f32p, _f32p, _vp, _u32]),
            "orc_pm_free": (None, [_vp]),
            "orc_pm_order": (None, [_vp, _u32p, _u8p]),
            "orc_pm_knn": (None, [_vp, _f32p, _f32p, _u32, _int, ctypes.c_float, _u32p, _f32p, _u32p, _int]),
            "orc_pm_radiance": (None, [_vp, _f32p, _f32p, _u32, _int, ctypes.c_float, _int, _f32p, _int]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def ref_lib():
    """The reference's own predicate sources built into oracle/_ref (None if absent)."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libref_predicate.so")
        if not os.path.exists(path):
            return None
        L = ctypes.CDLL(path)
        L.ref_tri_intersect.restype = _int
        L.ref_tri_intersect.argtypes = [_f32p] * 6
        L.ref_cross.restype = None
        L.ref_cross.argtypes = [_f32p] * 3
        L.ref_normalize.restype = None
        L.ref_normalize.argtypes = [_f32p] * 2
        _REF = L
    return _REF


def _c32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _cu32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def threads():
    return max(1, len(os.sched_getaffinity(0)))


def tri_intersect(o, d, v0, v1, v2):
    """(hit, t, alpha, beta) of ENCORE's predicate for one ray/triangle."""
    out = np.zeros(3, np.float32)
    h = lib().orc_tri_intersect(_c32(o), _c32(d), _c32(v0), _c32(v1), _c32(v2), out)
    return bool(h), out[0], out[1], out[2]


def tri_intersect_pairs(rays, tris):
    """rays (n,8), tris (n,9) -> (n,4) float32: t, alpha, beta, hit."""
    rays = _c32(rays)
    tris = _c32(tris)
    out = np.zeros((rays.shape[0], 4), np.float32)
    lib().orc_tri_intersect_pairs(rays, tris, rays.shape[0], out)
    return out


def cross(a, b):
    out = np.zeros(3, np.float32)
    lib().orc_cross(_c32(a), _c32(b), out)
    return out


def normalize(a):
    out = np.zeros(3, np.float32)
    lib().orc_normalize(_c32(a), out)
    return out


def hits_view(h):
    """(n,4) float32 hit array -> (t, u, v, prim uint32)."""
    h = np.ascontiguousarray(h, np.float32)
    return h[:, 0], h[:, 1], h[:, 2], h[:, 3].view(np.uint32)


def brute_closest(xyz, idx, rays, nthreads=None):
    xyz = _c32(xyz).reshape(-1)
    idx = _cu32(idx).reshape(-1)
    rays = _c32(rays).reshape(-1, 8)
    out = np.zeros((rays.shape[0], 4), np.float32)
    lib().orc_brute_closest(xyz, len(xyz) // 3, idx, len(idx) // 3, rays, rays.shape[0], out,
                            nthreads or threads())
    return out


def count_hits_below(xyz, idx, rays, t_limit, rel_tol):
    xyz = _c32(xyz).reshape(-1)
    idx = _cu32(idx).reshape(-1)
    rays = _c32(rays).reshape(-1, 8)
    out = np.zeros(rays.shape[0], np.uint32)
    lib().orc_count_hits_below(xyz, len(xyz) // 3, idx, len(idx) // 3, rays, rays.shape[0],
                               _c32(t_limit), float(rel_tol), out)
    return out


def fp64_eval(xyz, idx, rays, prim):
    rays = _c32(rays).reshape(-1, 8)
    out = np.zeros((rays.shape[0], 3), np.float64)
    lib().orc_fp64_eval(_c32(xyz).reshape(-1), _cu32(idx).reshape(-1), rays, _cu32(prim), rays.shape[0], out)
    return out


def fp64_closest(xyz, idx, rays, nthreads=None):
    xyz = _c32(xyz).reshape(-1)
    idx = _cu32(idx).reshape(-1)
    rays = _c32(rays).reshape(-1, 8)
    out = np.zeros((rays.shape[0], 3), np.float64)
    prim = np.zeros(rays.shape[0], np.uint32)
    lib().orc_fp64_closest(xyz, len(xyz) // 3, idx, len(idx) // 3, rays, rays.shape[0], out, prim,
                           nthreads or threads())
    return out, prim


class KdTree:
    """ENCORE's kd-tree (Kdtree.cpp) restated; bounds default to vertex bounds."""

    def __init__(self, xyz, idx, bmin=None, bmax=None):
        self.xyz = _c32(xyz).reshape(-1)
        self.idx = _cu32(idx).reshape(-1)
        v = self.xyz.reshape(-1, 3)
        bmin = _c32(v.min(axis=0) if bmin is None else bmin)
        bmax = _c32(v.max(axis=0) if bmax is None else bmax)
        self.h = lib().orc_kd_build(self.xyz, len(self.xyz) // 3, self.idx, len(self.idx) // 3, bmin, bmax)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_kd_free(self.h)
            self.h = None

    def stats(self):
        s = np.zeros(6, np.uint32)
        lib().orc_kd_stats(self.h, s)
        return dict(nodes=int(s[0]), leaves=int(s[1]), refs=int(s[2]), deepest=int(s[3]),
                    max_depth=int(s[4]), max_item=int(s[5]))

    def export(self):
        s = self.stats()
        words = np.zeros(s["nodes"], np.uint32)
        childs = np.zeros(s["nodes"], np.uint32)
        off = np.zeros(s["leaves"] + 1, np.uint32)
        prims = np.zeros(max(1, s["refs"]), np.uint32)
        box = np.zeros(6, np.float32)
        lib().orc_kd_export(self.h, words, childs, off, prims, box)
        return dict(words=words, childs=childs, leaf_off=off, leaf_prims=prims[: s["refs"]], root_box=box)

    def intersect(self, rays, nthreads=None):
        rays = _c32(rays).reshape(-1, 8)
        out = np.zeros((rays.shape[0], 4), np.float32)
        lib().orc_kd_intersect(self.h, rays, rays.shape[0], out, nthreads or threads())
        return out

    def occluded(self, rays, nthreads=None):
        rays = _c32(rays).reshape(-1, 8)
        out = np.zeros(rays.shape[0], np.uint32)
        lib().orc_kd_occluded(self.h, rays, rays.shape[0], out, nthreads or threads())
        return out


class PhotonMap:
    """ENCORE's balanced photon kd-tree (PhotonMap.cpp) restated."""

    def __init__(self, pos, nrm, power=None):
        self.pos = _c32(pos).reshape(-1, 3)
        self.nrm = _c32(nrm).reshape(-1, 3)
        self.n = self.pos.shape[0]
        pw = None if power is None else _c32(power).reshape(-1, 3)
        self._pw = pw
        self.h = lib().orc_pm_build(self.pos, self.nrm, None if pw is None else pw.ctypes.data, self.n)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_pm_free(self.h)
            self.h = None

    def order(self):
        perm = np.zeros(max(1, self.n), np.uint32)
        axis = np.zeros(max(1, self.n), np.uint8)
        lib().orc_pm_order(self.h, perm, axis)
        return perm[: self.n], axis[: self.n]

    def knn(self, qpos, qnrm, k, max_dist, nthreads=None):
        qpos = _c32(qpos).reshape(-1, 3)
        qnrm = _c32(qnrm).reshape(-1, 3)
        nq = qpos.shape[0]
        slots = np.full((nq, k), NO_HIT, np.uint32)
        d2 = np.full((nq, k), np.inf, np.float32)
        cnt = np.zeros(nq, np.uint32)
        lib().orc_pm_knn(self.h, qpos, qnrm, nq, k, float(max_dist), slots, d2, cnt, nthreads or threads())
        return slots, d2, cnt

    def radiance(self, qpos, qnrm, k, max_dist, caustic=False, nthreads=None):
        qpos = _c32(qpos).reshape(-1, 3)
        qnrm = _c32(qnrm).reshape(-1, 3)
        out = np.zeros((qpos.shape[0], 3), np.float32)
        lib().orc_pm_radiance(self.h, qpos, qnrm, qpos.shape[0], k, float(max_dist), int(caustic), out,
                              nthreads or threads())
        return out
