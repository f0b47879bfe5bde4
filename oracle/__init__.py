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
            "orc_fp64_closest_culled": (None, [_f32p, _u32, _u32p, _u32, _f32p, _u32, _f64p, _u32p, _int]),
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
            "orc_photon_dir": (None, [_f32p, ctypes.c_float, _u32p, _u32p, _u32, _f32p]),
            "orc_light_point": (None, [_f32p, _u32p, _u32p, _u32, _f32p]),
            "orc_ph_rand": (None, [ctypes.c_uint64, np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS"), _u32p,
                                   _u32, _u32p]),
            "orc_transmit_ray": (_int, [_f32p, _f32p, _f32p, _int, _f32p, _f32p]),
            "orc_kd_textures": (None, [_vp, _f32p, _f32p, _f32p, _f32p, _u32p]),
            "orc_ic_create": (_vp, [ctypes.c_float, _f32p, _f32p]),
            "orc_ic_free": (None, [_vp]),
            "orc_ic_add": (None, [_vp, _f32p, _f32p, _f32p, _f32p, _u32, _u32p]),
            "orc_ic_size": (_u32, [_vp]),
            "orc_ic_query": (None, [_vp, _f32p, _f32p, _u32, _f32p, _f32p, _u32p, _int]),
            "orc_emit_photons": (None, [_vp, _f32p, _u32p, _f32p, _u32, _u32, _u32p, _f32p, ctypes.c_uint64,
                                        ctypes.c_uint64, _vp, _u32p, ctypes.POINTER(ctypes.c_uint64)]),
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


def fp64_closest(xyz, idx, rays, nthreads=None, culled=False):
    """FP64 closest hit over every triangle (no flush): (t, alpha, beta) and
    prim per ray.  culled=True skips triangles by conservative boxes and
    returns exactly the brute-force result, fast enough for large scenes."""
    xyz = _c32(xyz).reshape(-1)
    idx = _cu32(idx).reshape(-1)
    rays = _c32(rays).reshape(-1, 8)
    out = np.zeros((rays.shape[0], 3), np.float64)
    prim = np.zeros(rays.shape[0], np.uint32)
    fn = lib().orc_fp64_closest_culled if culled else lib().orc_fp64_closest
    fn(xyz, len(xyz) // 3, idx, len(idx) // 3, rays, rays.shape[0], out, prim, nthreads or threads())
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

    def gpu_textures(self):
        """Kdtree::buildGPU's texture arrays (Kdtree.cpp:876-1000)."""
        st = self.stats()
        out = dict(cell_data0=np.zeros((st["nodes"], 4), np.float32),
                   vertex0=np.zeros((max(1, st["refs"]), 4), np.float32),
                   vertex1=np.zeros((max(1, st["refs"]), 4), np.float32),
                   vertex2=np.zeros((max(1, st["refs"]), 4), np.float32),
                   ref_prim=np.zeros(max(1, st["refs"]), np.uint32))
        lib().orc_kd_textures(self.h, out["cell_data0"], out["vertex0"], out["vertex1"], out["vertex2"],
                              out["ref_prim"])
        for k in ("vertex0", "vertex1", "vertex2", "ref_prim"):
            out[k] = out[k][: st["refs"]]
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


class IrrCache:
    """ENCORE's irradiance cache (IrrCacheOctree.h) restated: IrrCache(a,
    minPoint, size) with AddElement and the final gather's weighted lookup
    (PhotonMapperCPU.cpp:1087-1103)."""

    def __init__(self, a, min_point, size):
        self.h = lib().orc_ic_create(float(a), _c32(min_point), _c32(size))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_ic_free(self.h)
            self.h = None

    def add(self, value, pos, nrm, ri):
        """Returns per element 0 stored, 1 outside the root box, 2 refused
        (the reference's AddElement would never return)."""
        value, pos, nrm = (_c32(x).reshape(-1, 3) for x in (value, pos, nrm))
        ri = _c32(ri).reshape(-1)
        st = np.zeros(len(ri), np.uint32)
        lib().orc_ic_add(self.h, value, pos, nrm, ri, len(ri), st)
        return st

    def __len__(self):
        return int(lib().orc_ic_size(self.h))

    def query(self, qpos, qnrm, nthreads=None):
        """(rgb (n,3), weight sum (n,), count (n,)) per query point."""
        qpos, qnrm = _c32(qpos).reshape(-1, 3), _c32(qnrm).reshape(-1, 3)
        n = len(qpos)
        rgb = np.zeros((n, 3), np.float32)
        ws = np.zeros(n, np.float32)
        cnt = np.zeros(n, np.uint32)
        lib().orc_ic_query(self.h, qpos, qnrm, n, rgb, ws, cnt, nthreads or threads())
        return rgb, ws, cnt


def hit_attributes(xyz, idx, hits):
    """ENCORE's hit point v0 + (v1-v0)*alpha + (v2-v0)*beta and hit normal
    Normalize(Cross(e1, e2)) (TrianglePrim.cpp:236-239), float32 in the
    reference's order, for hit rows (prim in column 3 as uint32 bits).
    Returns (hit mask, points (n,3), normals (n,3)); miss rows are zero."""
    prim = np.ascontiguousarray(hits[:, 3]).view(np.uint32)
    ok = prim != NO_HIT
    n = len(hits)
    P = np.zeros((n, 3), np.float32)
    N = np.zeros((n, 3), np.float32)
    tri = np.asarray(idx)[prim[ok]]
    v0 = xyz[tri[:, 0]].astype(np.float32)
    e1 = (xyz[tri[:, 1]] - v0).astype(np.float32)
    e2 = (xyz[tri[:, 2]] - v0).astype(np.float32)
    a = hits[ok, 1:2].astype(np.float32)
    b = hits[ok, 2:3].astype(np.float32)
    P[ok] = ((v0 + e1 * a).astype(np.float32) + e2 * b).astype(np.float32)

    def flush(x):
        return np.where((x > -0.001) & (x < 0.001), np.float32(0), x).astype(np.float32)

    cx = flush(e1[:, 1] * e2[:, 2] - e1[:, 2] * e2[:, 1])
    cy = flush(np.float32(-1) * (e1[:, 0] * e2[:, 2] - e1[:, 2] * e2[:, 0]))
    cz = flush(e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0])
    m = np.sqrt((cx * cx + cy * cy) + cz * cz).astype(np.float32)
    with np.errstate(divide="ignore", invalid="ignore"):  # degenerate triangles give NaN, as in the reference
        N[ok] = np.stack([flush(cx / m), flush(cy / m), flush(cz / m)], axis=1)
    return ok, P, N


def final_gather(xyz, idx, kd, indirect, k_ind, r_ind, direct, k_dir, r_dir, diffuse, origin, dirs, nper,
                 nthreads=None):
    """PhotonMapperCPU::CalculateIndirectIllumination's final-gather branch
    (PhotonMapperCPU.cpp:1110-1167) for every shading point, composed from
    the restated kd-tree and photon maps: nper rays per point from its origin
    (no offset) along `dirs`; at each closest hit the radiance estimates of
    `indirect` (k_ind, r_ind) and `direct` (k_dir, r_dir; may be None), each
    times the hit triangle's diffuse colour, summed in sample order; times
    3.14159f / nper.  Returns (rgb (n,3), Ri (n,)): Ri = hits / sum(1/distance)."""
    origin = _c32(origin).reshape(-1, 3)
    dirs = _c32(dirs).reshape(-1, 3)
    npts = origin.shape[0]
    rays = np.zeros((npts * nper, 8), np.float32)
    rays[:, 0:3] = np.repeat(origin, nper, axis=0)
    rays[:, 4:7] = dirs
    rays[:, 7] = np.inf
    hits = kd.intersect(rays, nthreads=nthreads)
    ok, P, N = hit_attributes(xyz, idx, hits)
    Q = P.copy()
    Q[~ok] = np.float32(1e30)
    ind = indirect.radiance(Q, N, k_ind, r_ind, nthreads=nthreads)
    dr = direct.radiance(Q, N, k_dir, r_dir, nthreads=nthreads) if direct is not None else None
    prim = np.ascontiguousarray(hits[:, 3]).view(np.uint32)
    dif = np.asarray(diffuse, np.float32)[np.where(ok, prim, 0)]
    okm = ok.reshape(npts, nper)
    acc = np.zeros((npts, 3), np.float32)
    ri = np.zeros(npts, np.float32)
    nh = np.zeros(npts, np.float32)
    ind = ind.reshape(npts, nper, 3)
    dif = dif.reshape(npts, nper, 3)
    P = P.reshape(npts, nper, 3)
    if dr is not None:
        dr = dr.reshape(npts, nper, 3)
    for j in range(nper):
        h = okm[:, j]
        irr = (ind[:, j] * dif[:, j]).astype(np.float32)
        if dr is not None:
            irr = (irr + (dr[:, j] * dif[:, j]).astype(np.float32)).astype(np.float32)
        acc[h] = (acc[h] + irr[h]).astype(np.float32)
        v = (origin - P[:, j]).astype(np.float32)
        mag = np.sqrt((v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1]) + v[:, 2] * v[:, 2]).astype(np.float32)
        with np.errstate(divide="ignore"):
            inv = (np.float32(1) / mag).astype(np.float32)
        ri[h] = (ri[h] + inv[h]).astype(np.float32)
        nh[h] += 1
    scale = np.float32(np.float32(3.14159) / np.float32(nper))
    rgb = (acc * scale).astype(np.float32)
    with np.errstate(divide="ignore", invalid="ignore"):
        R = np.where(ri != 0, (nh / ri).astype(np.float32), np.float32(0)).astype(np.float32)
    return rgb, R


MAPS = ("global", "direct", "caustic")


def material_table(materials):
    """rq_material rows (diffuse rgb, transmittance, transparency,
    refraction index) from dicts."""
    t = np.zeros((len(materials), 6), np.float32)
    for i, m in enumerate(materials):
        t[i, 0:3] = m["diffuse"]
        t[i, 3] = m.get("transmittance", 0.0)
        t[i, 4] = m.get("transparency", 0.0)
        t[i, 5] = m.get("refraction_index", 0.0)
    return t


def light_table(lights):
    """rq_rect_light rows (center, normal, length, width, power) from dicts."""
    t = np.zeros((len(lights), 11), np.float32)
    for i, l in enumerate(lights):
        t[i, 0:3] = l["center"]
        t[i, 3:6] = l["normal"]
        t[i, 6] = l["length"]
        t[i, 7] = l["width"]
        t[i, 8:11] = l.get("power", (1.0, 1.0, 1.0))
    return t


def emit_photons(kd, materials, prim_material, lights, capacity, scale=(0.1, 0.05, 0.00005), max_bounces=3, seed=0,
                 max_paths=0):
    """ENCORE's EmitPhotons / TracePhoton (PhotonMapperCPU.cpp:503-835) on the
    oracle kd-tree, one path at a time, with the counter-based draws.
    Returns (maps, paths); maps[name] = dict(pos, normal, power, dir,
    shadow, path) of the stored photons in insertion order."""
    cap = np.asarray(capacity, np.uint32)
    sc = np.asarray(scale, np.float32)
    if not max_paths:
        max_paths = max(1 << 24, 64 * int(cap.astype(np.uint64).sum()))
    arrays = []
    ptrs = (ctypes.c_void_p * 18)()
    for m in range(3):
        c = max(1, int(cap[m]))
        a = dict(pos=np.zeros((c, 3), np.float32), normal=np.zeros((c, 3), np.float32),
                 power=np.zeros((c, 3), np.float32), dir=np.zeros((c, 3), np.float32),
                 shadow=np.zeros(c, np.uint8), path=np.zeros(c, np.uint32))
        for j, k in enumerate(("pos", "normal", "power", "dir", "shadow", "path")):
            ptrs[6 * m + j] = a[k].ctypes.data
        arrays.append(a)
    counts = np.zeros(3, np.uint32)
    paths = ctypes.c_uint64(0)
    mt = material_table(materials)
    lt = light_table(lights)
    pm = _cu32(prim_material).reshape(-1)
    lib().orc_emit_photons(kd.h, mt, pm, lt, len(lights), int(max_bounces), cap, sc, int(seed) & (2 ** 64 - 1),
                           int(max_paths), ptrs, counts, ctypes.byref(paths))
    maps = {}
    for m, name in enumerate(MAPS):
        n = int(counts[m])
        maps[name] = {k: v[:n] for k, v in arrays[m].items()}
    return maps, int(paths.value)


def photon_dir(normal, e1_lo, r1, r2):
    """randDirectionN / randDirectionSuperN about `normal` for draws r1, r2."""
    r1 = _cu32(r1).reshape(-1)
    r2 = _cu32(r2).reshape(-1)
    out = np.zeros((len(r1), 3), np.float32)
    lib().orc_photon_dir(_c32(normal).reshape(3), float(e1_lo), r1, r2, len(r1), out)
    return out


def light_point(light, r1, r2):
    """RectangleLight::EmitPhoton's origin for draws r1, r2."""
    r1 = _cu32(r1).reshape(-1)
    r2 = _cu32(r2).reshape(-1)
    out = np.zeros((len(r1), 3), np.float32)
    lib().orc_light_point(light_table([light])[0], r1, r2, len(r1), out)
    return out


def ph_rand(seed, path, slot):
    """The counter-based draw that replaces rand() for (path, slot) pairs."""
    path = np.ascontiguousarray(path, np.uint64).reshape(-1)
    slot = _cu32(slot).reshape(-1)
    out = np.zeros(len(path), np.uint32)
    lib().orc_ph_rand(int(seed) & (2 ** 64 - 1), path, slot, len(path), out)
    return out


def transmit_ray(d, P, N, entering, material):
    """TransmitRay for one hit: (origin, direction) or None when the surface
    does not refract."""
    out = np.zeros(6, np.float32)
    ok = lib().orc_transmit_ray(_c32(d).reshape(3), _c32(P).reshape(3), _c32(N).reshape(3), int(entering),
                                material_table([material])[0], out)
    return (out[:3], out[3:]) if ok else None
