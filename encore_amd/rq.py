"""ctypes front of libencore_rq.so for tests and the benchmark.

The engine itself is C++/HIP behind the C ABI in include/rq.h; this module
only loads the library, converts status codes to exceptions and passes
device pointers of torch tensors.  It raises if the library is not built:
there is no fallback path.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libencore_rq.so")

RQ_OK = 0
STATUS = {0: "RQ_OK", 1: "RQ_ERR_INVALID_ARG", 2: "RQ_ERR_OOM", 3: "RQ_ERR_DEVICE", 4: "RQ_ERR_UNSUPPORTED",
          5: "RQ_ERR_CAPACITY", 6: "RQ_ERR_TIMEOUT", 7: "RQ_ERR_CANCELLED"}
ACCEL = {"bvh": 0, "kd": 1}
PRED = {"encore": 0, "watertight": 1}
NO_HIT = 0xFFFFFFFF
PHOTON_LAYOUT = {"encore": 0, "spatial": 1}


class RQError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("%s: %s" % (STATUS.get(status, status), msg))
        self.status = status


class SceneDesc(ctypes.Structure):
    _fields_ = [("accel", ctypes.c_uint32), ("predicate", ctypes.c_uint32), ("leaf_size", ctypes.c_uint32),
                ("has_bounds", ctypes.c_uint32), ("bounds_min", ctypes.c_float * 3),
                ("bounds_max", ctypes.c_float * 3), ("bvh_width", ctypes.c_uint32), ("builder", ctypes.c_uint32)]


class SceneInfo(ctypes.Structure):
    _fields_ = [("accel", ctypes.c_uint32), ("predicate", ctypes.c_uint32), ("ntris", ctypes.c_uint32),
                ("nnodes", ctypes.c_uint32), ("nleaves", ctypes.c_uint32), ("nrefs", ctypes.c_uint32),
                ("max_depth", ctypes.c_uint32), ("kd_max_depth_limit", ctypes.c_uint32),
                ("device_bytes", ctypes.c_uint64), ("build_seconds", ctypes.c_double),
                ("bvh_width", ctypes.c_uint32), ("stack_bound", ctypes.c_uint32)]


class BatchItem(ctypes.Structure):
    _fields_ = [("scene", ctypes.c_void_p), ("d_rays", ctypes.c_void_p), ("d_out", ctypes.c_void_p),
                ("n", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


class KernelConfig(ctypes.Structure):
    _fields_ = [("stack", ctypes.c_uint32), ("blocks_per_cu", ctypes.c_uint32), ("fetch_threshold", ctypes.c_uint32),
                ("grid", ctypes.c_uint32)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("slots", ctypes.c_uint32), ("plain", ctypes.c_uint32), ("max_rays_per_launch", ctypes.c_uint32),
                ("max_requests_per_launch", ctypes.c_uint32), ("hbm_budget_bytes", ctypes.c_uint64),
                ("hbm_reserve_bytes", ctypes.c_uint64), ("stack", ctypes.c_uint32), ("blocks_per_cu", ctypes.c_uint32),
                ("fetch_threshold", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


class RequestInfo(ctypes.Structure):
    _fields_ = [("latency_ms", ctypes.c_double), ("queue_ms", ctypes.c_double), ("launch_id", ctypes.c_uint64),
                ("rays_in_launch", ctypes.c_uint32), ("requests_in_launch", ctypes.c_uint32)]


class ServerCounters(ctypes.Structure):
    _fields_ = [(f, ctypes.c_uint64) for f in ("launches", "requests_done", "rays_done", "queued", "admitted",
                                                "arena_bytes", "arena_used")]


_LIB = None
_vp = ctypes.c_void_p
_u32 = ctypes.c_uint32


class Material(ctypes.Structure):
    _fields_ = [("diffuse", ctypes.c_float * 3), ("transmittance", ctypes.c_float),
                ("transparency", ctypes.c_float), ("refraction_index", ctypes.c_float)]


class RectLight(ctypes.Structure):
    _fields_ = [("center", ctypes.c_float * 3), ("normal", ctypes.c_float * 3), ("length", ctypes.c_float),
                ("width", ctypes.c_float), ("power", ctypes.c_float * 3)]


class PhotonParams(ctypes.Structure):
    _fields_ = [("max_bounces", _u32), ("capacity", _u32 * 3), ("scale", ctypes.c_float * 3),
                ("seed", ctypes.c_uint64), ("max_paths", ctypes.c_uint64), ("batch_paths", _u32), ("pad", _u32)]


class PhotonBuffers(ctypes.Structure):
    _fields_ = [("pos", _vp), ("normal", _vp), ("power", _vp), ("dir", _vp), ("shadow", _vp), ("path", _vp)]


class PhotonResult(ctypes.Structure):
    _fields_ = [("count", _u32 * 3), ("paths", ctypes.c_uint64)]


MAPS = ("global", "direct", "caustic")


def load():
    global _LIB
    if _LIB is not None:
        return _LIB
    # RQ_LIBRARY names another build of the same library (A/B comparisons).
    path = os.environ.get("RQ_LIBRARY") or LIB_PATH
    if not os.path.exists(path):
        raise RuntimeError("%s is not built (run __graft_entry__.build()); no fallback exists" % path)
    L = ctypes.CDLL(path)
    L.rq_version.restype = ctypes.c_char_p
    L.rq_last_error.restype = ctypes.c_char_p
    L.rq_device_open.argtypes = [ctypes.c_int, ctypes.POINTER(_vp)]
    L.rq_device_close.argtypes = [_vp]
    L.rq_device_close.restype = None
    L.rq_scene_create.argtypes = [_vp, _vp, _u32, _vp, _u32, ctypes.POINTER(SceneDesc), ctypes.POINTER(_vp)]
    L.rq_scene_destroy.argtypes = [_vp]
    L.rq_scene_destroy.restype = None
    L.rq_scene_stats.argtypes = [_vp, ctypes.POINTER(SceneInfo)]
    L.rq_intersect_closest.argtypes = [_vp, _vp, _u32, _vp, _vp]
    L.rq_occluded.argtypes = [_vp, _vp, _u32, _vp, _vp]
    L.rq_trace_counts.argtypes = [_vp, _vp, _u32, _u32, _vp, _vp]
    if hasattr(L, "rq_trace_profile"):
        L.rq_trace_profile.argtypes = [_vp, _vp, _u32, _u32, _vp, _vp]
        L.rq_trace_profile.restype = ctypes.c_int
    L.rq_batch_create.argtypes = [_vp, _u32, ctypes.POINTER(_vp)]
    L.rq_batch_set.argtypes = [_vp, _vp, _u32]
    L.rq_batch_launch.argtypes = [_vp, _u32, _vp]
    L.rq_batch_get_config.argtypes = [_vp, ctypes.POINTER(KernelConfig)]
    L.rq_batch_set_config.argtypes = [_vp, ctypes.POINTER(KernelConfig)]
    L.rq_batch_autotune.argtypes = [_vp, _vp, ctypes.POINTER(KernelConfig), ctypes.POINTER(ctypes.c_double)]
    L.rq_batch_destroy.argtypes = [_vp]
    L.rq_batch_destroy.restype = None
    L.rq_server_create.argtypes = [_vp, ctypes.POINTER(ServerConfig), ctypes.POINTER(_vp)]
    L.rq_server_submit.argtypes = [_vp, _vp, _vp, _u32, _u32, _vp, ctypes.POINTER(ctypes.c_uint64)]
    L.rq_server_wait.argtypes = [_vp, ctypes.c_uint64, _u32, ctypes.POINTER(RequestInfo)]
    L.rq_server_poll.argtypes = [_vp, ctypes.c_uint64, ctypes.POINTER(ctypes.c_int)]
    L.rq_server_cancel.argtypes = [_vp, ctypes.c_uint64]
    L.rq_server_stats.argtypes = [_vp, ctypes.POINTER(ServerCounters)]
    L.rq_server_destroy.argtypes = [_vp]
    L.rq_server_destroy.restype = None
    L.rq_photons_create.argtypes = [_vp, _vp, _vp, _vp, _u32, _u32, ctypes.POINTER(_vp)]
    L.rq_photons_destroy.argtypes = [_vp]
    L.rq_photons_destroy.restype = None
    L.rq_photons_order.argtypes = [_vp, _vp, _vp, _vp]
    L.rq_photons_balance_host.argtypes = [_vp, _u32, _u32, _vp, _vp]
    L.rq_knn.argtypes = [_vp, _vp, _vp, _u32, _u32, ctypes.c_float, _vp, _vp, _vp, _vp]
    L.rq_radiance_estimate.argtypes = [_vp, _vp, _vp, _u32, _u32, ctypes.c_float, _u32, _vp, _vp, _vp, _vp, _vp]
    L.rq_radiance_counts.argtypes = [_vp, _vp, _vp, _u32, _u32, ctypes.c_float, _u32, _vp, _vp, _vp]
    L.rq_radiance_counts.restype = ctypes.c_int
    L.rq_photons_precompute_irradiance.argtypes = [_vp, _vp, _vp, _u32, _u32, _u32, ctypes.c_float, _u32, _u32, _vp,
                                                   _vp, _vp]
    L.rq_photons_precompute_irradiance.restype = ctypes.c_int
    if hasattr(L, "rq_scene_save"):
        L.rq_scene_save.argtypes = [_vp, ctypes.c_char_p]
        L.rq_scene_save.restype = ctypes.c_int
        L.rq_scene_load.argtypes = [_vp, ctypes.c_char_p, ctypes.POINTER(_vp)]
        L.rq_scene_load.restype = ctypes.c_int
    if hasattr(L, "rq_final_gather"):
        L.rq_final_gather.argtypes = [_vp, _vp, _u32, ctypes.c_float, _vp, _u32, ctypes.c_float, _vp, _vp, _u32, _vp,
                                      _u32, _vp, _vp, _vp]
        L.rq_final_gather.restype = ctypes.c_int
    if hasattr(L, "rq_scene_refit"):
        L.rq_scene_refit.argtypes = [_vp, _vp, _u32, _vp]
    if hasattr(L, "rq_scene_create_kd"):
        L.rq_scene_create_kd.argtypes = [_vp, _vp, _u32, _vp, _u32, _vp, _u32, _vp, _u32, _vp, _u32,
                                         ctypes.POINTER(_vp)]
        L.rq_scene_create_kd.restype = ctypes.c_int
        L.rq_scene_refit.restype = ctypes.c_int
    if hasattr(L, "rq_emit_photons"):
        L.rq_scene_set_materials.argtypes = [_vp, _vp, _u32, _vp]
        L.rq_scene_set_materials.restype = ctypes.c_int
        L.rq_emit_photons.argtypes = [_vp, _vp, _u32, ctypes.POINTER(PhotonParams), ctypes.POINTER(PhotonBuffers),
                                      ctypes.POINTER(PhotonResult), _vp]
        L.rq_emit_photons.restype = ctypes.c_int
    if hasattr(L, "rq_irrcache_create"):
        L.rq_irrcache_create.argtypes = [_vp, ctypes.c_float, _vp, _vp, ctypes.POINTER(_vp)]
        L.rq_irrcache_add.argtypes = [_vp, _vp, _vp, _vp, _vp, _u32, _vp]
        L.rq_irrcache_query.argtypes = [_vp, _vp, _vp, _u32, _vp, _vp, _vp, _vp]
        L.rq_irrcache_info.argtypes = [_vp, _vp, _vp, _vp, _vp]
        L.rq_irrcache_counts.argtypes = [_vp, _vp, _vp, _u32, _vp, _vp, _vp]
        L.rq_irrcache_counts.restype = ctypes.c_int
        L.rq_irrcache_destroy.argtypes = [_vp]
        L.rq_irrcache_destroy.restype = None
        for name in ("rq_irrcache_create", "rq_irrcache_add", "rq_irrcache_query", "rq_irrcache_info"):
            getattr(L, name).restype = ctypes.c_int
    L.rq_debug_fail_alloc_after.argtypes = [ctypes.c_int64]
    L.rq_debug_fail_alloc_after.restype = None
    L.rq_host_tree_build.argtypes = [_vp, _u32, _vp, _u32, ctypes.POINTER(SceneDesc), ctypes.POINTER(_vp)]
    L.rq_host_tree_info.argtypes = [_vp, ctypes.POINTER(SceneInfo)]
    L.rq_host_tree_copy.argtypes = [_vp, _vp, _vp, _vp]
    if hasattr(L, "rq_host_tree_export_textures"):
        L.rq_host_tree_export_textures.argtypes = [_vp, _vp, _u32, _vp, _vp, _vp, _vp, _vp, _vp]
        L.rq_host_tree_export_textures.restype = ctypes.c_int
    L.rq_host_tree_destroy.argtypes = [_vp]
    L.rq_host_tree_destroy.restype = None
    for name in ("rq_device_open", "rq_scene_create", "rq_scene_stats", "rq_intersect_closest", "rq_occluded",
                 "rq_host_tree_build", "rq_host_tree_info", "rq_host_tree_copy", "rq_trace_counts",
                 "rq_photons_create", "rq_photons_order", "rq_photons_balance_host", "rq_knn",
                 "rq_radiance_estimate", "rq_batch_create", "rq_batch_set", "rq_batch_launch", "rq_batch_get_config",
                 "rq_batch_set_config", "rq_batch_autotune", "rq_server_create", "rq_server_submit", "rq_server_wait",
                 "rq_server_poll", "rq_server_cancel", "rq_server_stats"):
        getattr(L, name).restype = ctypes.c_int
    _LIB = L
    return L


def check(status):
    if status != RQ_OK:
        raise RQError(status, load().rq_last_error().decode())


def _ptr(t):
    """Device pointer of a torch tensor (or an int address)."""
    if t is None:
        return None
    if isinstance(t, int):
        return t
    return t.data_ptr()


def _stream(stream):
    if stream is None:
        return None
    if isinstance(stream, int):
        return stream
    return stream.cuda_stream


class Device:
    def __init__(self, ordinal=0):
        L = load()
        h = _vp()
        check(L.rq_device_open(ordinal, ctypes.byref(h)))
        self.h = h
        self.ordinal = ordinal

    def close(self):
        if getattr(self, "h", None):
            load().rq_device_close(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # interpreter shutdown: the library may already be gone
            pass


class Scene:
    """An acceleration structure resident in HBM."""

    def __init__(self, device, xyz, idx, accel="bvh", predicate="encore", leaf_size=0, bounds=None, width=0,
                 builder="sah"):
        """width: BVH children per node, 2 or 4 (0 = 4); builder: "sah" (host)
        or "lbvh" (on the GPU, width 2)."""
        L = load()
        xyz = np.ascontiguousarray(xyz, np.float32).reshape(-1)
        idx = np.ascontiguousarray(idx, np.uint32).reshape(-1)
        d = _desc(accel, predicate, leaf_size, bounds, width, builder)
        h = _vp()
        check(L.rq_scene_create(device.h, xyz.ctypes.data, len(xyz) // 3, idx.ctypes.data, len(idx) // 3,
                                ctypes.byref(d), ctypes.byref(h)))
        self.h = h
        self.device = device
        self.accel = accel
        self.predicate = predicate

    def destroy(self):
        if getattr(self, "h", None):
            load().rq_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:  # interpreter shutdown: the library may already be gone
            pass

    def stats(self):
        info = SceneInfo()
        check(load().rq_scene_stats(self.h, ctypes.byref(info)))
        return {f: getattr(info, f) for f, _ in SceneInfo._fields_}

    def intersect(self, rays, hits, n=None, stream=None):
        """rays: (n,8) float32 device tensor; hits: (n,4) float32 device tensor."""
        n = rays.shape[0] if n is None else n
        check(load().rq_intersect_closest(self.h, _ptr(rays), n, _ptr(hits), _stream(stream)))

    def profile(self, rays, any_hit=False, n=None, stream=None):
        """SIMD-efficiency profile of the BVH kernel (synchronous): dict of
        rays, nodes, tris, node_iters, tri_iters, rounds and the node / triangle
        loop lane utilisation."""
        n = rays.shape[0] if n is None else n
        c = np.zeros(6, np.uint64)
        check(load().rq_trace_profile(self.h, _ptr(rays), n, int(any_hit), c.ctypes.data, _stream(stream)))
        d = dict(zip(("rays", "nodes", "tris", "node_iters", "tri_iters", "rounds"), (int(x) for x in c)))
        d["node_util"] = d["nodes"] / max(1, 64 * d["node_iters"])
        d["tri_util"] = d["tris"] / max(1, 64 * d["tri_iters"])
        return d

    def save(self, path):
        """Write the scene to a packed scene file."""
        check(load().rq_scene_save(self.h, os.fsencode(path)))

    @classmethod
    def from_kd(cls, device, xyz, idx, words, childs, leaf_prims, root_box, predicate="encore"):
        """A kd scene over a kd-tree built elsewhere (rq_scene_create_kd):
        node words and children in Kdtree's numbering, a leaf's child being
        the offset of its triangles in leaf_prims; root_box = pos, size."""
        xyz = np.ascontiguousarray(xyz, np.float32).reshape(-1)
        idx = np.ascontiguousarray(idx, np.uint32).reshape(-1)
        nodes = np.ascontiguousarray(np.stack([np.asarray(words, np.uint32), np.asarray(childs, np.uint32)], axis=1))
        lp = np.ascontiguousarray(leaf_prims, np.uint32).reshape(-1)
        box = np.ascontiguousarray(root_box, np.float32).reshape(6)
        h = _vp()
        check(load().rq_scene_create_kd(device.h, xyz.ctypes.data, len(xyz) // 3, idx.ctypes.data, len(idx) // 3,
                                        nodes.ctypes.data, len(nodes), lp.ctypes.data if len(lp) else None, len(lp),
                                        box.ctypes.data, PRED[predicate], ctypes.byref(h)))
        sc = cls.__new__(cls)
        sc.h = h
        sc.device = device
        sc.accel = "kd"
        sc.predicate = predicate
        return sc

    @classmethod
    def load(cls, device, path):
        """A scene read from a packed scene file (no build)."""
        h = _vp()
        check(load().rq_scene_load(device.h, os.fsencode(path), ctypes.byref(h)))
        sc = cls.__new__(cls)
        sc.h = h
        sc.device = device
        st = sc.stats()
        sc.accel = "kd" if st["accel"] == ACCEL["kd"] else "bvh"
        sc.predicate = [k for k, v in PRED.items() if v == st["predicate"]][0]
        return sc

    def final_gather(self, indirect, k_ind, r_ind, direct, k_dir, r_dir, diffuse, origin, dirs, nper, rgb, ri=None,
                     stream=None):
        """Final gather (synchronous): diffuse (ntris,3), origin (n,3),
        dirs (n*nper,3) device tensors; rgb (n,3) and ri (n,) outputs."""
        check(load().rq_final_gather(self.h, indirect.h, int(k_ind), float(r_ind), direct.h if direct else None,
                                     int(k_dir), float(r_dir), _ptr(diffuse), _ptr(origin), origin.shape[0],
                                     _ptr(dirs), int(nper), _ptr(rgb), _ptr(ri), _stream(stream)))

    def refit(self, xyz, stream=None):
        """Update for moved vertices: xyz is a (nverts, 3) float32 device
        tensor (or host array, copied up first)."""
        import torch
        t = xyz if isinstance(xyz, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(xyz, np.float32)).cuda()
        t = t.contiguous()
        check(load().rq_scene_refit(self.h, t.data_ptr(), t.shape[0], _stream(stream)))
        return t

    def set_materials(self, materials, prim_material):
        """materials: sequence of dicts (diffuse, transmittance=0,
        transparency=0, refraction_index=0); prim_material: (ntris,) ints."""
        arr = (Material * len(materials))()
        for i, m in enumerate(materials):
            arr[i].diffuse[:] = [float(x) for x in m["diffuse"]]
            arr[i].transmittance = float(m.get("transmittance", 0.0))
            arr[i].transparency = float(m.get("transparency", 0.0))
            arr[i].refraction_index = float(m.get("refraction_index", 0.0))
        pm = np.ascontiguousarray(prim_material, np.uint32)
        check(load().rq_scene_set_materials(self.h, ctypes.cast(arr, _vp), len(materials), pm.ctypes.data))

    def emit_photons(self, lights, capacity, scale=(0.1, 0.05, 0.00005), max_bounces=3, seed=0, max_paths=0,
                     batch_paths=0, extras=True, stream=None):
        """Photon emission into three maps (global, direct, caustic).  lights:
        sequence of dicts (center, normal, length, width, power).  Returns
        (maps, paths): maps[name] = dict of torch tensors (pos, normal, power,
        and with extras dir, shadow, path) cut to the stored count."""
        import torch
        la = (RectLight * len(lights))()
        for i, l in enumerate(lights):
            la[i].center[:] = [float(x) for x in l["center"]]
            la[i].normal[:] = [float(x) for x in l["normal"]]
            la[i].length = float(l["length"])
            la[i].width = float(l["width"])
            la[i].power[:] = [float(x) for x in l.get("power", (1.0, 1.0, 1.0))]
        P = PhotonParams()
        P.max_bounces = int(max_bounces)
        P.capacity[:] = [int(c) for c in capacity]
        P.scale[:] = [float(x) for x in scale]
        P.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
        P.max_paths = int(max_paths)
        P.batch_paths = int(batch_paths)
        bufs = (PhotonBuffers * 3)()
        keep = []
        for m in range(3):
            c = max(1, int(capacity[m]))
            t = dict(pos=torch.empty((c, 3), device="cuda"), normal=torch.empty((c, 3), device="cuda"),
                     power=torch.empty((c, 3), device="cuda"))
            if extras:
                t.update(dir=torch.empty((c, 3), device="cuda"), shadow=torch.empty((c,), dtype=torch.uint8,
                                                                                     device="cuda"),
                         path=torch.empty((c,), dtype=torch.int32, device="cuda"))
            for k, v in t.items():
                setattr(bufs[m], k, v.data_ptr())
            keep.append(t)
        res = PhotonResult()
        check(load().rq_emit_photons(self.h, ctypes.cast(la, _vp), len(lights), ctypes.byref(P), bufs,
                                     ctypes.byref(res), _stream(stream)))
        maps = {}
        for m, name in enumerate(MAPS):
            n = int(res.count[m])
            maps[name] = {k: v[:n] for k, v in keep[m].items()}
        return maps, int(res.paths)

    def counts(self, rays, any_hit=False, n=None, stream=None):
        """(rays, nodes visited, triangles tested) of one query (synchronous)."""
        n = rays.shape[0] if n is None else n
        c = np.zeros(3, np.uint64)
        check(load().rq_trace_counts(self.h, _ptr(rays), n, int(any_hit), c.ctypes.data, _stream(stream)))
        return int(c[0]), int(c[1]), int(c[2])

    def occluded(self, rays, out, n=None, stream=None):
        n = rays.shape[0] if n is None else n
        check(load().rq_occluded(self.h, _ptr(rays), n, _ptr(out), _stream(stream)))


class Batch:
    """Many ray batches in one persistent launch (optionally a replayed hipGraph)."""

    def __init__(self, device, any_hit=False):
        L = load()
        h = _vp()
        check(L.rq_batch_create(device.h, int(any_hit), ctypes.byref(h)))
        self.h = h
        self._keep = None

    def set(self, items):
        """items: list of (scene, rays tensor, out tensor)."""
        arr = (BatchItem * max(1, len(items)))()
        for i, (sc, rays, out) in enumerate(items):
            arr[i].scene = sc.h
            arr[i].d_rays = _ptr(rays)
            arr[i].d_out = _ptr(out)
            arr[i].n = rays.shape[0]
        self._keep = items
        check(load().rq_batch_set(self.h, arr, len(items)))

    def launch(self, graph=False, stream=None):
        check(load().rq_batch_launch(self.h, int(graph), _stream(stream)))

    def config(self):
        c = KernelConfig()
        check(load().rq_batch_get_config(self.h, ctypes.byref(c)))
        return dict(stack=c.stack, blocks_per_cu=c.blocks_per_cu, fetch_threshold=c.fetch_threshold, grid=c.grid)

    def set_config(self, stack=16, blocks_per_cu=0, fetch_threshold=20):
        c = KernelConfig(stack, blocks_per_cu, fetch_threshold, 0)
        check(load().rq_batch_set_config(self.h, ctypes.byref(c)))

    def autotune(self, stream=None):
        c = KernelConfig()
        ms = ctypes.c_double()
        check(load().rq_batch_autotune(self.h, _stream(stream), ctypes.byref(c), ctypes.byref(ms)))
        return dict(stack=c.stack, blocks_per_cu=c.blocks_per_cu, fetch_threshold=c.fetch_threshold, grid=c.grid,
                    ms=ms.value)

    def destroy(self):
        if getattr(self, "h", None):
            load().rq_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:  # interpreter shutdown: the library may already be gone
            pass


class Server:
    """Continuous-batching request server over resident scenes."""

    def __init__(self, device, slots=0, plain=False, hbm_budget=0, hbm_reserve=0, max_rays_per_launch=0,
                 max_requests_per_launch=0, stack=0, blocks_per_cu=0, fetch_threshold=0):
        c = ServerConfig(slots, int(plain), max_rays_per_launch, max_requests_per_launch, hbm_budget, hbm_reserve,
                         stack, blocks_per_cu, fetch_threshold, 0)
        h = _vp()
        check(load().rq_server_create(device.h, ctypes.byref(c), ctypes.byref(h)))
        self.h = h
        self._live = {}

    def submit(self, scene, rays, out, any_hit=False):
        """rays: host tensor/array (n,8) float32; out: host (n,4) float32 or (n,) int32.
        Both must stay alive until wait(); pinned tensors overlap copies."""
        t = ctypes.c_uint64()
        n = rays.shape[0]
        rp = rays.data_ptr() if hasattr(rays, "data_ptr") else rays.ctypes.data
        op = out.data_ptr() if hasattr(out, "data_ptr") else out.ctypes.data
        check(load().rq_server_submit(self.h, scene.h, rp, n, int(any_hit), op, ctypes.byref(t)))
        self._live[t.value] = (rays, out)
        return t.value

    def wait(self, ticket, timeout_ms=0xFFFFFFFF):
        info = RequestInfo()
        st = load().rq_server_wait(self.h, ticket, timeout_ms, ctypes.byref(info))
        if st != 6:  # a timeout keeps the ticket alive
            self._live.pop(ticket, None)
        check(st)
        return dict(latency_ms=info.latency_ms, queue_ms=info.queue_ms, launch_id=info.launch_id,
                    rays_in_launch=info.rays_in_launch, requests_in_launch=info.requests_in_launch)

    def poll(self, ticket):
        d = ctypes.c_int()
        check(load().rq_server_poll(self.h, ticket, ctypes.byref(d)))
        return bool(d.value)

    def cancel(self, ticket):
        check(load().rq_server_cancel(self.h, ticket))

    def stats(self):
        c = ServerCounters()
        check(load().rq_server_stats(self.h, ctypes.byref(c)))
        return {f: getattr(c, f) for f, _ in ServerCounters._fields_}

    def destroy(self):
        if getattr(self, "h", None):
            load().rq_server_destroy(self.h)
            self.h = None
            self._live.clear()

    def __del__(self):
        try:
            self.destroy()
        except Exception:  # interpreter shutdown: the library may already be gone
            pass


class PhotonMap:
    """ENCORE's balanced photon map resident in HBM."""

    def __init__(self, device, pos, normal, power=None, layout="spatial"):
        L = load()
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        normal = np.ascontiguousarray(normal, np.float32).reshape(-1, 3)
        pw = None if power is None else np.ascontiguousarray(power, np.float32).reshape(-1, 3)
        h = _vp()
        check(L.rq_photons_create(device.h, pos.ctypes.data, normal.ctypes.data,
                                  None if pw is None else pw.ctypes.data, len(pos), PHOTON_LAYOUT[layout],
                                  ctypes.byref(h)))
        self.h = h
        self.n = len(pos)

    def destroy(self):
        if getattr(self, "h", None):
            load().rq_photons_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:  # interpreter shutdown: the library may already be gone
            pass

    def order(self):
        perm = np.zeros(max(1, self.n), np.uint32)
        axis = np.zeros(max(1, self.n), np.uint8)
        nc = ctypes.c_uint32()
        check(load().rq_photons_order(self.h, perm.ctypes.data, axis.ctypes.data, ctypes.byref(nc)))
        return perm[: self.n], axis[: self.n], nc.value

    def knn(self, qpos, qnrm, k, max_dist, idx, d2, count=None, stream=None):
        """qpos, qnrm: (nq,3) float32 device tensors; idx (nq,k) int32, d2 (nq,k) float32, count (nq,) int32."""
        check(load().rq_knn(self.h, _ptr(qpos), _ptr(qnrm), qpos.shape[0], k, float(max_dist), _ptr(idx), _ptr(d2),
                            _ptr(count), _stream(stream)))

    def radiance(self, qpos, qnrm, k, max_dist, rgb, idx, d2, count=None, caustic=False, stream=None):
        check(load().rq_radiance_estimate(self.h, _ptr(qpos), _ptr(qnrm), qpos.shape[0], k, float(max_dist),
                                          int(caustic), _ptr(rgb), _ptr(idx), _ptr(d2), _ptr(count),
                                          _stream(stream)))

    def precompute_irradiance(self, pos, normal, increment, k, max_dist, caustic=False, layout="spatial"):
        """PhotonMap::PrecomputeIrradiance: (new PhotonMap, insertion indices
        of its photons, their powers).  pos, normal: this map's photons in
        insertion order (host)."""
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        normal = np.ascontiguousarray(normal, np.float32).reshape(-1, 3)
        m = 0 if self.n == 0 or int(increment) < 1 else (self.n - 1) // int(increment) + 1  # the library checks
        chosen = np.zeros(max(1, m), np.uint32)
        power = np.zeros((max(1, m), 3), np.float32)
        h = _vp()
        check(load().rq_photons_precompute_irradiance(self.h, pos.ctypes.data, normal.ctypes.data, len(pos),
                                                      int(increment), int(k), float(max_dist), int(caustic),
                                                      PHOTON_LAYOUT[layout], chosen.ctypes.data, power.ctypes.data,
                                                      ctypes.byref(h)))
        made = PhotonMap.__new__(PhotonMap)
        made.h = h
        made.n = m
        return made, chosen[:m], power[:m]

    def radiance_counts(self, qpos, qnrm, k, max_dist, rgb, caustic=False, stream=None):
        """Radiance estimates into rgb, with the gather's work counted:
        {"queries", "ranges" (cells or blocks ranked), "photons" (read),
        "candidates"} (cell-grid gather only)."""
        out = np.zeros(4, np.uint64)
        check(load().rq_radiance_counts(self.h, _ptr(qpos), _ptr(qnrm), qpos.shape[0], k, float(max_dist),
                                        int(caustic), _ptr(rgb), out.ctypes.data, _stream(stream)))
        return dict(zip(("queries", "ranges", "photons", "candidates"), (int(v) for v in out)))


class IrrCache:
    """The final gather's irradiance cache (IrrCache<Color>) on the GPU:
    add() cached values in order, query() the weighted averages."""

    def __init__(self, device, a, min_point=(-75.0, -75.0, -75.0), size=(150.0, 150.0, 150.0)):
        h = _vp()
        mp = np.ascontiguousarray(min_point, np.float32)
        sz = np.ascontiguousarray(size, np.float32)
        check(load().rq_irrcache_create(device.h, float(a), mp.ctypes.data, sz.ctypes.data, ctypes.byref(h)))
        self.h = h

    def add(self, value, pos, nrm, ri):
        """Host arrays (n,3), (n,3), (n,3), (n,); returns the per-value status
        (0 stored, 1 outside the root box, 2 refused); raises RQError
        (status 4) when any value is refused, storing none."""
        value, pos, nrm = (np.ascontiguousarray(x, np.float32).reshape(-1, 3) for x in (value, pos, nrm))
        ri = np.ascontiguousarray(ri, np.float32).reshape(-1)
        st = np.zeros(max(1, len(ri)), np.uint32)
        check(load().rq_irrcache_add(self.h, value.ctypes.data, pos.ctypes.data, nrm.ctypes.data, ri.ctypes.data,
                                     len(ri), st.ctypes.data))
        return st[: len(ri)]

    def query(self, qpos, qnrm, rgb, wsum=None, count=None, n=None, stream=None):
        """Device tensors: qpos, qnrm (n,3) float32 -> rgb (n,3), wsum (n,), count (n,)."""
        n = qpos.shape[0] if n is None else n
        check(load().rq_irrcache_query(self.h, _ptr(qpos), _ptr(qnrm), n, _ptr(rgb), _ptr(wsum), _ptr(count),
                                       _stream(stream)))

    def counts(self, qpos, qnrm, rgb, n=None, stream=None):
        """(lookups, cached values examined, values counted) of one lookup
        call, from the instrumented kernel (synchronous)."""
        n = qpos.shape[0] if n is None else n
        out = np.zeros(3, np.uint64)
        check(load().rq_irrcache_counts(self.h, _ptr(qpos), _ptr(qnrm), n, _ptr(rgb), out.ctypes.data,
                                        _stream(stream)))
        return tuple(int(x) for x in out)

    def info(self):
        c, g = ctypes.c_uint32(), ctypes.c_uint32()
        cells, refs = ctypes.c_uint64(), ctypes.c_uint64()
        check(load().rq_irrcache_info(self.h, ctypes.byref(c), ctypes.byref(g), ctypes.byref(cells),
                                      ctypes.byref(refs)))
        return dict(count=c.value, global_count=g.value, index_cells=cells.value, index_refs=refs.value)

    def destroy(self):
        if getattr(self, "h", None):
            load().rq_irrcache_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:  # interpreter shutdown: the library may already be gone
            pass


def photon_balance_host(pos, layout="encore"):
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
    n = len(pos)
    perm = np.zeros(max(1, n), np.uint32)
    axis = np.zeros(max(1, n), np.uint8)
    check(load().rq_photons_balance_host(pos.ctypes.data, n, PHOTON_LAYOUT[layout], perm.ctypes.data,
                                         axis.ctypes.data))
    return perm[:n], axis[:n]


BVH_NODE_DTYPE = np.dtype([("x", "<f4", 4), ("y", "<f4", 4), ("z", "<f4", 4), ("ref", "<u4", 2),
                           ("pad", "<u4", 2)])
BVH4_NODE_DTYPE = np.dtype([("lox", "<f4", 4), ("hix", "<f4", 4), ("loy", "<f4", 4), ("hiy", "<f4", 4),
                            ("loz", "<f4", 4), ("hiz", "<f4", 4), ("ref", "<u4", 4), ("pad", "<u4", 4)])
REF_EMPTY = 0xFFFFFFFF


BUILDER = {"sah": 0, "lbvh": 1}


def _desc(accel, predicate, leaf_size, bounds, width=0, builder="sah"):
    d = SceneDesc()
    d.accel = ACCEL[accel]
    d.predicate = PRED[predicate]
    d.leaf_size = leaf_size
    d.bvh_width = width
    d.builder = BUILDER[builder]
    if bounds is not None:
        d.has_bounds = 1
        for k in range(3):
            d.bounds_min[k] = float(bounds[0][k])
            d.bounds_max[k] = float(bounds[1][k])
    return d


def host_tree(xyz, idx, accel="kd", leaf_size=0, bounds=None, width=0):
    """Builds the engine's tree on the host (no GPU) and returns its arrays.
    kd: dict(words, childs, leaf_prims, root_box, info); bvh: dict(nodes,
    tri_order, info) with BVH_NODE_DTYPE (width 2) or BVH4_NODE_DTYPE
    (width 4, the default) nodes."""
    L = load()
    xyz = np.ascontiguousarray(xyz, np.float32).reshape(-1)
    idx = np.ascontiguousarray(idx, np.uint32).reshape(-1)
    d = _desc(accel, "encore", leaf_size, bounds, width)
    h = _vp()
    check(L.rq_host_tree_build(xyz.ctypes.data, len(xyz) // 3, idx.ctypes.data, len(idx) // 3, ctypes.byref(d),
                               ctypes.byref(h)))
    try:
        info = SceneInfo()
        check(L.rq_host_tree_info(h, ctypes.byref(info)))
        inf = {f: getattr(info, f) for f, _ in SceneInfo._fields_}
        box = np.zeros(6, np.float32)
        if accel == "kd":
            nodes = np.zeros((info.nnodes, 2), np.uint32)
            lp = np.zeros(max(1, info.nrefs), np.uint32)
            check(L.rq_host_tree_copy(h, nodes.ctypes.data, lp.ctypes.data, box.ctypes.data))
            return dict(words=nodes[:, 0].copy(), childs=nodes[:, 1].copy(), leaf_prims=lp[: info.nrefs],
                        root_box=box, info=inf)
        nodes = np.zeros(info.nnodes, BVH4_NODE_DTYPE if info.bvh_width == 4 else BVH_NODE_DTYPE)
        order = np.zeros(info.ntris, np.uint32)
        check(L.rq_host_tree_copy(h, nodes.ctypes.data, order.ctypes.data, box.ctypes.data))
        return dict(nodes=nodes, tri_order=order, info=inf)
    finally:
        L.rq_host_tree_destroy(h)


def kd_gpu_textures(xyz, idx, bounds=None):
    """ENCORE's GPU texture encoding (Kdtree::buildGPU) of the engine's kd
    tree over the triangles: dict(cell_data0 (nnodes, 4), vertex0/1/2
    (nrefs, 4), ref_prim (nrefs,)).  Host only."""
    L = load()
    xyz = np.ascontiguousarray(xyz, np.float32).reshape(-1)
    idx = np.ascontiguousarray(idx, np.uint32).reshape(-1)
    d = _desc("kd", "encore", 0, bounds)
    h = _vp()
    check(L.rq_host_tree_build(xyz.ctypes.data, len(xyz) // 3, idx.ctypes.data, len(idx) // 3, ctypes.byref(d),
                               ctypes.byref(h)))
    try:
        info = SceneInfo()
        check(L.rq_host_tree_info(h, ctypes.byref(info)))
        out = dict(cell_data0=np.zeros((info.nnodes, 4), np.float32),
                   vertex0=np.zeros((max(1, info.nrefs), 4), np.float32),
                   vertex1=np.zeros((max(1, info.nrefs), 4), np.float32),
                   vertex2=np.zeros((max(1, info.nrefs), 4), np.float32),
                   ref_prim=np.zeros(max(1, info.nrefs), np.uint32))
        check(L.rq_host_tree_export_textures(h, xyz.ctypes.data, len(xyz) // 3, idx.ctypes.data,
                                             out["cell_data0"].ctypes.data, out["vertex0"].ctypes.data,
                                             out["vertex1"].ctypes.data, out["vertex2"].ctypes.data,
                                             out["ref_prim"].ctypes.data))
        for k in ("vertex0", "vertex1", "vertex2", "ref_prim"):
            out[k] = out[k][: info.nrefs]
        return out
    finally:
        L.rq_host_tree_destroy(h)


def version():
    return load().rq_version().decode()


def fail_alloc_after(n):
    load().rq_debug_fail_alloc_after(int(n))
