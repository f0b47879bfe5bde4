// rq_api.cpp -- the C ABI of libencore_rq.so (declared in include/rq.h).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "rq_internal.h"

namespace rq {

namespace {
thread_local std::string g_last_error;
// Allocation fault injection for tests: the Nth dev_alloc from now fails
// (0 = off).  Set through rq_debug_fail_alloc_after.
std::atomic<int64_t> g_fail_after{0};
}  // namespace

void hip_check(hipError_t e, const char* what) {
    if (e == hipSuccess) return;
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
        throw Error(RQ_ERR_OOM, std::string(what) + ": " + hipGetErrorString(e));
    throw Error(RQ_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

void* dev_alloc(size_t bytes) {
    int64_t f = g_fail_after.load();
    if (f > 0) {
        if (g_fail_after.fetch_sub(1) == 1) throw Error(RQ_ERR_OOM, "device allocation failed (injected)");
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        throw Error(RQ_ERR_OOM, "hipMalloc(" + std::to_string(bytes) + "): " + hipGetErrorString(e));
    }
    return p;
}

void dev_free(void* p) {
    if (p) (void)hipFree(p);
}

static void set_error(const std::string& m) { g_last_error = m; }

template <class F>
static rq_status guarded(F&& f) {
    try {
        f();
        g_last_error.clear();
        return RQ_OK;
    } catch (const Error& e) {
        set_error(e.what());
        return e.status;
    } catch (const std::bad_alloc&) {
        set_error("host allocation failed");
        return RQ_ERR_OOM;
    } catch (const std::exception& e) {
        set_error(e.what());
        return RQ_ERR_DEVICE;
    } catch (...) {
        set_error("unknown error");
        return RQ_ERR_DEVICE;
    }
}

rq_status guard_call(void (*fn)(void*), void* ctx) {
    return guarded([&] { fn(ctx); });
}

static void require(bool ok, const char* msg) {
    if (!ok) throw Error(RQ_ERR_INVALID_ARG, msg);
}

// Per-stream scratch: [counter: 256 B][spill stacks].  Grows on demand.
static uint32_t* stream_scratch(rq_device* dev, hipStream_t s, size_t spill_bytes) {
    std::lock_guard<std::mutex> g(dev->mu);
    DevBuf& b = dev->scratch[s];
    const size_t need = 256 + spill_bytes;
    if (b.bytes < need) {
        // The old buffer may still be read by queued work on this stream.
        RQ_HIP(hipStreamSynchronize(s));
        b = DevBuf(need);
    }
    return b.as<uint32_t>();
}

static int grid_for(rq_device* dev, uint32_t accel, uint32_t pred, TraceKind kind, uint32_t width, uint32_t n) {
    const KernelConfig& kc = dev->kc;
    int per_cu = kc.blocks_per_cu > 0 ? kc.blocks_per_cu : occupancy_blocks_per_cu(accel, pred, kind, width, kc);
    int64_t grid = (int64_t)per_cu * dev->num_cus;
    const int64_t bt = block_threads(accel);
    // kd: up to a wave per ray (the kernel spreads a small launch over every
    // wave and tests long leaves with the idle lanes); BVH: a lane per ray.
    const int64_t per_block = accel == RQ_ACCEL_KD ? bt / 64 : bt;
    const int64_t need = ((int64_t)n + per_block - 1) / per_block;
    grid = std::max<int64_t>(1, std::min(grid, need));
    return (int)grid;
}

// n_dev: the ray count is read on the device from there (n bounds it).
// fixed: scratch of fixed_bytes given by the caller (counter, then spill
// stacks); the grid then shrinks to what it holds and nothing is allocated.
static void trace(rq_scene* sc, const rq_ray* rays, uint32_t n, rq_hit* hits, uint32_t* occ, TraceKind kind,
                  hipStream_t s, unsigned long long* counts = nullptr, const uint32_t* n_dev = nullptr,
                  uint32_t* fixed = nullptr, size_t fixed_bytes = 0) {
    if (n == 0) return;
    if (n > kMaxCount) throw Error(RQ_ERR_CAPACITY, "at most 2^30 rays per call");
    rq_device* dev = sc->dev;
    RQ_HIP(hipSetDevice(dev->ordinal));
    int grid = grid_for(dev, sc->accel, sc->predicate, kind, sc->width, n);
    const uint32_t words = scene_spill_words(sc);
    uint32_t* scratch = nullptr;
    if (fixed) {
        // A persistent launch serves every ray with fewer workgroups, only
        // more slowly: keep as many as the fixed spill region holds.
        const size_t per_block = (size_t)block_threads(sc->accel) * words * sizeof(uint32_t);
        if (fixed_bytes < 256 + per_block)
            throw Error(RQ_ERR_CAPACITY, "the scene's traversal stacks need " + std::to_string(256 + per_block) +
                                             " bytes of scratch per workgroup; the static region holds " +
                                             std::to_string(fixed_bytes));
        if (per_block) grid = (int)std::min<size_t>((size_t)grid, (fixed_bytes - 256) / per_block);
        scratch = fixed;
    } else {
        const size_t spill_bytes = (size_t)grid * block_threads(sc->accel) * words * sizeof(uint32_t);
        scratch = stream_scratch(dev, s, spill_bytes);
    }
    TraceLaunch L{};
    L.nodes = sc->nodes.ptr;
    L.tris = sc->tris.as<TriRec>();
    L.leaf_prims = sc->leaf_prims.as<uint32_t>();
    L.rays = rays;
    L.hits = hits;
    L.occluded = occ;
    L.n = n;
    L.n_dev = n_dev;
    L.counter = scratch;
    L.spill = scratch + 64;
    L.spill_depth = words;
    for (int k = 0; k < 3; ++k) {
        L.root_pos[k] = sc->root_pos[k];
        L.root_size[k] = sc->root_size[k];
    }
    L.counts = counts;
    L.fetch_threshold = (uint32_t)(sc->accel == RQ_ACCEL_KD ? dev->kc.kd_fetch_threshold : dev->kc.fetch_threshold);
    // RQ_KD_COOP_LANES / RQ_KD_COOP_MIN set the kd kernel's cooperative leaf
    // tests for sweeps (0 lanes turns them off); results do not depend on them.
    L.coop_lanes = kKdCoopLanesDefault;
    L.coop_min = kKdCoopMinDefault;
    L.leaf_batch = bvh_leaf_batch();
    if (const char* e = std::getenv("RQ_KD_COOP_LANES")) L.coop_lanes = (uint32_t)std::min(64, std::max(0, std::atoi(e)));
    if (const char* e = std::getenv("RQ_KD_COOP_MIN")) L.coop_min = (uint32_t)std::max(1, std::atoi(e));
    launch_zero_u32(scratch, 2, s);
    if (counts)
        launch_count(L, sc->accel, sc->predicate, kind, sc->width, grid, s);
    else if (sc->accel == RQ_ACCEL_KD)
        launch_kd(L, sc->predicate, kind, grid, s);
    else
        launch_bvh(L, sc->predicate, kind, sc->width, grid, dev->kc, s);
}

void trace_closest(rq_scene* sc, const rq_ray* rays, uint32_t n, rq_hit* hits, hipStream_t s, const uint32_t* n_dev) {
    trace(sc, rays, n, hits, nullptr, TraceKind::Closest, s, nullptr, n_dev);
}

void trace_closest_in(rq_scene* sc, const rq_ray* rays, uint32_t n, rq_hit* hits, hipStream_t s, uint32_t* scratch,
                      size_t bytes) {
    trace(sc, rays, n, hits, nullptr, TraceKind::Closest, s, nullptr, nullptr, scratch, bytes);
}

static void vertex_bounds(const float* xyz, uint32_t nverts, const rq_scene_desc& dd, float bmin[3], float bmax[3]);

// ENCORE's kd-tree of the triangles: nodes uploaded, the triangle records
// (one per leaf reference, in leaf order) and the prim -> record map returned.
static void kd_scene_arrays(rq_scene* sc, const float* xyz, uint32_t nverts, const uint32_t* tri_idx,
                            uint32_t ntris, const rq_scene_desc& dd, std::vector<TriRec>& tris,
                            std::vector<uint32_t>& slot) {
    float bmin[3], bmax[3];
    vertex_bounds(xyz, nverts, dd, bmin, bmax);
    KdHost h;
    build_kd(xyz, nverts, tri_idx, ntris, bmin, bmax, h);
    pack_triangles(xyz, tri_idx, h.leaf_prims.data(), (uint32_t)h.leaf_prims.size(), dd.predicate, tris);
    // A primitive can sit in several leaves; any of its records will do.
    slot.assign(ntris, 0xFFFFFFFFu);
    for (uint32_t j = 0; j < (uint32_t)h.leaf_prims.size(); ++j)
        if (slot[h.leaf_prims[j]] == 0xFFFFFFFFu) slot[h.leaf_prims[j]] = j;
    for (uint32_t& v : slot)
        if (v == 0xFFFFFFFFu) v = 0;  // in no leaf: never hit
    sc->stats = h.stats;
    for (int k = 0; k < 3; ++k) {
        sc->root_pos[k] = h.root_pos[k];
        sc->root_size[k] = h.root_size[k];
    }
    sc->nodes = DevBuf(h.nodes.size() * sizeof(KdNodeG));
    sc->nodes.upload(h.nodes.data(), sc->nodes.bytes);
}

// A kd scene rebuilt from moved vertices (same triangles and options).
static void rebuild_kd(rq_scene* sc, const float* xyz) {
    std::vector<TriRec> tris;
    std::vector<uint32_t> slot;
    rq_scene_desc dd = sc->desc;
    dd.has_bounds = 0;  // the vertices may have left the bounds given at creation
    kd_scene_arrays(sc, xyz, sc->nverts, sc->idx_host.data(), sc->ntris, dd, tris, slot);
    sc->tris = DevBuf(tris.size() * sizeof(TriRec));
    sc->tris.upload(tris.data(), sc->tris.bytes);
    sc->prim_slot.upload(slot.data(), slot.size() * 4);
    sc->device_bytes = sc->nodes.bytes + sc->tris.bytes + sc->leaf_prims.bytes + sc->prim_slot.bytes;
}

// Refit support of a BVH scene, made on its first update: each record's
// vertex indices (leaf order) and the inner nodes grouped by depth, deepest
// level first.
static void prepare_refit(rq_scene* sc) {
    const uint32_t n = sc->ntris;
    if (!sc->rec_vidx.ptr) {
        std::vector<uint32_t> slot(n), vidx(3 * (size_t)n);
        RQ_HIP(hipMemcpy(slot.data(), sc->prim_slot.ptr, (size_t)n * 4, hipMemcpyDeviceToHost));
        for (uint32_t t = 0; t < n; ++t)
            for (int j = 0; j < 3; ++j) vidx[3 * (size_t)slot[t] + j] = sc->idx_host[3 * (size_t)t + j];
        sc->rec_vidx = DevBuf(vidx.size() * 4);
        sc->rec_vidx.upload(vidx.data(), sc->rec_vidx.bytes);
    }
    // Depth of every node, from the node array itself.
    const uint32_t nn = sc->stats.nnodes;
    std::vector<uint32_t> refs(sc->width == 4 ? 4 * (size_t)nn : 2 * (size_t)nn);
    if (sc->width == 4) {
        std::vector<BvhNode4> nodes(nn);
        RQ_HIP(hipMemcpy(nodes.data(), sc->nodes.ptr, nn * sizeof(BvhNode4), hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < nn; ++i)
            for (int c = 0; c < 4; ++c) refs[4 * (size_t)i + c] = nodes[i].ref[c];
    } else {
        std::vector<BvhNode> nodes(nn);
        RQ_HIP(hipMemcpy(nodes.data(), sc->nodes.ptr, nn * sizeof(BvhNode), hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < nn; ++i)
            for (int c = 0; c < 2; ++c) refs[2 * (size_t)i + c] = nodes[i].ref[c];
    }
    // Only nodes reached from the root get a level: a GPU-built scene leaves
    // the slots of subtrees it made into leaves unused.
    const uint32_t w = sc->width == 4 ? 4 : 2;
    constexpr uint32_t kUnreached = 0xFFFFFFFFu;
    std::vector<uint32_t> depth(nn, kUnreached), queue{0};
    depth[0] = 0;
    uint32_t maxd = 0;
    for (size_t q = 0; q < queue.size(); ++q) {
        const uint32_t i = queue[q];
        for (uint32_t c = 0; c < w; ++c) {
            const uint32_t r = refs[(size_t)w * i + c];
            if (r == kRefEmpty || (r & kRefLeaf)) continue;
            if (r >= nn || depth[r] != kUnreached) throw Error(RQ_ERR_DEVICE, "BVH node array is not a tree");
            depth[r] = depth[i] + 1;
            maxd = std::max(maxd, depth[r]);
            queue.push_back(r);
        }
    }
    std::vector<uint32_t> order;
    order.reserve(nn);
    sc->refit_levels.assign(1, 0);
    for (int64_t d = maxd; d >= 0; --d) {
        for (uint32_t i = 0; i < nn; ++i)
            if (depth[i] == (uint32_t)d) order.push_back(i);
        sc->refit_levels.push_back((uint32_t)order.size());
    }
    sc->refit_order = DevBuf(order.size() * 4);
    sc->refit_order.upload(order.data(), sc->refit_order.bytes);
}

// Model-list bounds for the kd root: given, or the vertex bounds.
static void vertex_bounds(const float* xyz, uint32_t nverts, const rq_scene_desc& dd, float bmin[3], float bmax[3]) {
    if (dd.has_bounds) {
        for (int k = 0; k < 3; ++k) {
            bmin[k] = dd.bounds_min[k];
            bmax[k] = dd.bounds_max[k];
        }
        return;
    }
    for (int k = 0; k < 3; ++k) {
        bmin[k] = INFINITY;
        bmax[k] = -INFINITY;
    }
    for (uint32_t v = 0; v < nverts; ++v)
        for (int k = 0; k < 3; ++k) {
            bmin[k] = std::min(bmin[k], xyz[3 * (size_t)v + k]);
            bmax[k] = std::max(bmax[k], xyz[3 * (size_t)v + k]);
        }
}

}  // namespace rq

struct rq_host_tree {
    uint32_t accel = 0, ntris = 0;
    rq::BvhHost bvh;
    rq::KdHost kd;
};

using namespace rq;

extern "C" {

const char* rq_version(void) { return "encore-rq 0.1.0 gfx950"; }

const char* rq_last_error(void) { return g_last_error.c_str(); }

void rq_debug_fail_alloc_after(int64_t n) { g_fail_after.store(n); }

rq_status rq_device_open(int ordinal, rq_device** out) {
    return guarded([&] {
        require(out != nullptr, "out is null");
        int count = 0;
        RQ_HIP(hipGetDeviceCount(&count));
        require(ordinal >= 0 && ordinal < count, "device ordinal out of range");
        RQ_HIP(hipSetDevice(ordinal));
        hipDeviceProp_t p;
        RQ_HIP(hipGetDeviceProperties(&p, ordinal));
        auto d = std::make_unique<rq_device>();
        d->ordinal = ordinal;
        d->num_cus = p.multiProcessorCount;
        d->total_mem = p.totalGlobalMem;
        d->arch = p.gcnArchName;
        d->max_waves_per_cu = p.maxThreadsPerMultiProcessor / 64;
        if (const char* e = std::getenv("RQ_STACK")) d->kc.stack = std::atoi(e);
        if (const char* e = std::getenv("RQ_BLOCKS_PER_CU")) d->kc.blocks_per_cu = std::atoi(e);
        if (const char* e = std::getenv("RQ_FETCH_THRESHOLD")) d->kc.fetch_threshold = std::atoi(e);
        if (d->kc.stack != 8 && d->kc.stack != 16 && d->kc.stack != 32) d->kc.stack = 8;
        bind_static_scratch(d.get());  // with this device current
        *out = d.release();
    });
}

void rq_device_close(rq_device* dev) {
    if (!dev) return;
    (void)hipSetDevice(dev->ordinal);
    (void)hipDeviceSynchronize();
    delete dev;
}

rq_status rq_scene_create(rq_device* dev, const float* xyz, uint32_t nverts, const uint32_t* tri_idx,
                          uint32_t ntris, const rq_scene_desc* desc, rq_scene** out) {
    return guarded([&] {
        require(dev && xyz && tri_idx && out, "null argument");
        require(ntris > 0 && nverts > 0, "empty triangle list");
        rq_scene_desc dd{};
        if (desc) dd = *desc;
        require(dd.accel == RQ_ACCEL_BVH || dd.accel == RQ_ACCEL_KD, "unknown accel kind");
        require(dd.predicate == RQ_PRED_ENCORE || dd.predicate == RQ_PRED_WATERTIGHT, "unknown predicate");
        RQ_HIP(hipSetDevice(dev->ordinal));
        auto sc = std::make_unique<rq_scene>();
        sc->dev = dev;
        sc->accel = dd.accel;
        sc->predicate = dd.predicate;
        sc->ntris = ntris;
        std::vector<TriRec> tris;
        std::vector<uint32_t> slot;
        require(dd.builder == RQ_BUILD_HOST_SAH || dd.builder == RQ_BUILD_DEVICE_LBVH, "unknown builder");
        if (dd.accel == RQ_ACCEL_BVH && dd.builder == RQ_BUILD_DEVICE_LBVH) {
            if (dd.bvh_width == 4) throw Error(RQ_ERR_UNSUPPORTED, "the device builder makes BVH2 scenes");
            require(*std::max_element(tri_idx, tri_idx + 3 * (size_t)ntris) < nverts, "triangle index out of range");
            build_bvh_device(sc.get(), xyz, nverts, tri_idx, ntris, dd.leaf_size, nullptr);
        } else if (dd.accel == RQ_ACCEL_BVH) {
            require(dd.bvh_width == 0 || dd.bvh_width == 2 || dd.bvh_width == 4, "bvh_width must be 0, 2 or 4");
            BvhHost h;
            build_bvh(xyz, nverts, tri_idx, ntris, dd.leaf_size, dd.bvh_width ? dd.bvh_width : 4, h);
            pack_triangles(xyz, tri_idx, h.tri_order.data(), ntris, dd.predicate, tris);
            slot.assign(ntris, 0u);
            for (uint32_t i = 0; i < ntris; ++i) slot[h.tri_order[i]] = i;
            sc->stats = h.stats;
            sc->width = h.width;
            if (h.width == 4) {
                sc->nodes = DevBuf(h.nodes4.size() * sizeof(BvhNode4));
                sc->nodes.upload(h.nodes4.data(), sc->nodes.bytes);
            } else {
                sc->nodes = DevBuf(h.nodes.size() * sizeof(BvhNode));
                sc->nodes.upload(h.nodes.data(), sc->nodes.bytes);
            }
        } else {
            kd_scene_arrays(sc.get(), xyz, nverts, tri_idx, ntris, dd, tris, slot);
        }
        if (!tris.empty()) {  // host builds; the device builder has written its own arrays
            sc->tris = DevBuf(tris.size() * sizeof(TriRec));
            sc->tris.upload(tris.data(), sc->tris.bytes);
            sc->prim_slot = DevBuf(slot.size() * sizeof(uint32_t));
            sc->prim_slot.upload(slot.data(), sc->prim_slot.bytes);
        }
        sc->nverts = nverts;
        sc->idx_host.assign(tri_idx, tri_idx + 3 * (size_t)ntris);
        sc->desc = dd;
        sc->device_bytes = sc->nodes.bytes + sc->tris.bytes + sc->leaf_prims.bytes + sc->prim_slot.bytes +
                           sc->rec_vidx.bytes + sc->refit_order.bytes;
        *out = sc.release();
    });
}

rq_status rq_scene_refit(rq_scene* sc, const float* d_xyz, uint32_t nverts, void* stream) {
    return guarded([&] {
        require(sc && d_xyz, "null argument");
        require(!sc->idx_host.empty(), "scene has no triangle list (loaded from a file): rebuild it to update");
        require(nverts == sc->nverts, "vertex count differs from the scene's");
        hipStream_t s = (hipStream_t)stream;
        RQ_HIP(hipSetDevice(sc->dev->ordinal));
        if (sc->accel == RQ_ACCEL_KD) {
            // Kdtree::update builds the tree again (Kdtree.cpp:84-200): so does
            // this, on the host, from the moved vertices.
            std::vector<float> xyz(3 * (size_t)nverts);
            RQ_HIP(hipMemcpyAsync(xyz.data(), d_xyz, xyz.size() * 4, hipMemcpyDeviceToHost, s));
            RQ_HIP(hipStreamSynchronize(s));
            rebuild_kd(sc, xyz.data());
            return;
        }
        if (sc->refit_levels.empty()) prepare_refit(sc);
        RefitLaunch L{};
        L.xyz = d_xyz;
        L.rec_vidx = sc->rec_vidx.as<uint32_t>();
        L.tris = sc->tris.as<TriRec>();
        L.nrecs = sc->ntris;
        L.predicate = sc->predicate;
        L.nodes = sc->nodes.ptr;
        L.width = sc->width;
        L.order = sc->refit_order.as<uint32_t>();
        launch_refit_tris(L, s);
        for (size_t l = 0; l + 1 < sc->refit_levels.size(); ++l) {
            L.begin = sc->refit_levels[l];
            L.end = sc->refit_levels[l + 1];
            launch_refit_level(L, s);
        }
    });
}

rq_status rq_scene_create_kd(rq_device* dev, const float* xyz, uint32_t nverts, const uint32_t* tri_idx,
                             uint32_t ntris, const void* nodes, uint32_t nnodes, const uint32_t* leaf_prims,
                             uint32_t nrefs, const float* root_box, uint32_t predicate, rq_scene** out) {
    return guarded([&] {
        require(dev && xyz && tri_idx && nodes && root_box && out, "null argument");
        require(ntris > 0 && nverts > 0, "empty triangle list");
        require(nnodes > 0, "a kd-tree has at least its root");
        require(nrefs == 0 || leaf_prims, "null leaf list");
        require(predicate == RQ_PRED_ENCORE || predicate == RQ_PRED_WATERTIGHT, "unknown predicate");
        require(*std::max_element(tri_idx, tri_idx + 3 * (size_t)ntris) < nverts, "triangle index out of range");
        require(nrefs == 0 || *std::max_element(leaf_prims, leaf_prims + nrefs) < ntris, "leaf triangle out of range");
        const KdNodeG* nd = static_cast<const KdNodeG*>(nodes);
        // Walk the tree from the root: every node reached once, children and
        // leaf ranges in bounds; the depth sizes the traversal stack.
        std::vector<uint32_t> depth(nnodes, 0xFFFFFFFFu);
        std::vector<uint32_t> todo{0};
        depth[0] = 0;
        uint32_t maxd = 0, nleaves = 0;
        while (!todo.empty()) {
            const uint32_t i = todo.back();
            todo.pop_back();
            maxd = std::max(maxd, depth[i]);
            if ((nd[i].word & 3u) == 3u) {
                const uint64_t cnt = nd[i].word >> 2;
                require((uint64_t)nd[i].child + cnt <= nrefs, "leaf range outside leaf_prims");
                ++nleaves;
                continue;
            }
            const uint32_t c = nd[i].child;
            require((uint64_t)c + 1 < nnodes, "child index out of range");
            for (uint32_t k = c; k <= c + 1; ++k) {
                require(depth[k] == 0xFFFFFFFFu, "kd node array is not a tree");
                depth[k] = depth[i] + 1;
                todo.push_back(k);
            }
        }
        RQ_HIP(hipSetDevice(dev->ordinal));
        auto sc = std::make_unique<rq_scene>();
        sc->dev = dev;
        sc->accel = RQ_ACCEL_KD;
        sc->predicate = predicate;
        sc->ntris = ntris;
        std::vector<TriRec> tris;
        pack_triangles(xyz, tri_idx, leaf_prims, nrefs, predicate, tris);
        std::vector<uint32_t> slot(ntris, 0xFFFFFFFFu);
        for (uint32_t j = 0; j < nrefs; ++j)
            if (slot[leaf_prims[j]] == 0xFFFFFFFFu) slot[leaf_prims[j]] = j;
        for (uint32_t& v : slot)
            if (v == 0xFFFFFFFFu) v = 0;  // in no leaf: never hit
        for (int k = 0; k < 3; ++k) {
            sc->root_pos[k] = root_box[k];
            sc->root_size[k] = root_box[3 + k];
        }
        sc->stats = BuildStats{};
        sc->stats.nnodes = nnodes;
        sc->stats.nleaves = nleaves;
        sc->stats.nrefs = nrefs;
        sc->stats.max_depth = maxd + 1;
        sc->stats.stack_bound = maxd + 2;
        sc->nodes = DevBuf((size_t)nnodes * sizeof(KdNodeG));
        sc->nodes.upload(nodes, sc->nodes.bytes);
        sc->tris = DevBuf(std::max<size_t>(1, tris.size()) * sizeof(TriRec));
        if (!tris.empty()) sc->tris.upload(tris.data(), tris.size() * sizeof(TriRec));
        sc->prim_slot = DevBuf(slot.size() * sizeof(uint32_t));
        sc->prim_slot.upload(slot.data(), sc->prim_slot.bytes);
        sc->nverts = nverts;
        sc->idx_host.assign(tri_idx, tri_idx + 3 * (size_t)ntris);
        sc->desc = rq_scene_desc{};
        sc->desc.accel = RQ_ACCEL_KD;
        sc->desc.predicate = predicate;
        sc->device_bytes = sc->nodes.bytes + sc->tris.bytes + sc->prim_slot.bytes;
        *out = sc.release();
    });
}

void rq_scene_destroy(rq_scene* scene) {
    if (!scene) return;
    (void)hipSetDevice(scene->dev->ordinal);
    delete scene;
}

rq_status rq_scene_stats(const rq_scene* sc, rq_scene_info* info) {
    return guarded([&] {
        require(sc && info, "null argument");
        info->accel = sc->accel;
        info->predicate = sc->predicate;
        info->ntris = sc->ntris;
        info->nnodes = sc->stats.nnodes;
        info->nleaves = sc->stats.nleaves;
        info->nrefs = sc->stats.nrefs;
        info->max_depth = sc->stats.max_depth;
        info->kd_max_depth_limit = sc->stats.kd_depth_limit;
        info->device_bytes = sc->device_bytes;
        info->build_seconds = sc->stats.seconds;
        info->bvh_width = sc->accel == RQ_ACCEL_BVH ? sc->width : 0;
        info->stack_bound = sc->stats.stack_bound;
    });
}

rq_status rq_intersect_closest(rq_scene* scene, const rq_ray* d_rays, uint32_t n, rq_hit* d_hits, void* stream) {
    return guarded([&] {
        require(scene != nullptr, "scene is null");
        require(n == 0 || (d_rays && d_hits), "null ray or hit buffer");
        trace(scene, d_rays, n, d_hits, nullptr, TraceKind::Closest, (hipStream_t)stream);
    });
}

rq_status rq_occluded(rq_scene* scene, const rq_ray* d_rays, uint32_t n, uint32_t* d_occluded, void* stream) {
    return guarded([&] {
        require(scene != nullptr, "scene is null");
        require(n == 0 || (d_rays && d_occluded), "null ray or output buffer");
        trace(scene, d_rays, n, nullptr, d_occluded, TraceKind::Any, (hipStream_t)stream);
    });
}

// Runs the instrumented kernel; h[0] nodes, h[1] triangles, h[2..4] wave
// iterations of the node loop, the triangle loop and the refill round.
static void count_trace(rq_scene* scene, const rq_ray* d_rays, uint32_t n, uint32_t any_hit, void* stream,
                        unsigned long long h[5]) {
    require(n == 0 || d_rays, "null ray buffer");
    hipStream_t s = (hipStream_t)stream;
    RQ_HIP(hipSetDevice(scene->dev->ordinal));
    DevBuf out(n * (any_hit ? sizeof(uint32_t) : sizeof(rq_hit)) + 16);
    DevBuf cnt(5 * sizeof(unsigned long long));
    RQ_HIP(hipMemsetAsync(cnt.ptr, 0, cnt.bytes, s));
    trace(scene, d_rays, n, any_hit ? nullptr : out.as<rq_hit>(), any_hit ? out.as<uint32_t>() : nullptr,
          any_hit ? TraceKind::Any : TraceKind::Closest, s, cnt.as<unsigned long long>());
    RQ_HIP(hipMemcpyAsync(h, cnt.ptr, 5 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    RQ_HIP(hipStreamSynchronize(s));
}

rq_status rq_trace_counts(rq_scene* scene, const rq_ray* d_rays, uint32_t n, uint32_t any_hit, uint64_t* counts,
                          void* stream) {
    return guarded([&] {
        require(scene && counts, "null argument");
        unsigned long long h[5] = {0, 0, 0, 0, 0};
        count_trace(scene, d_rays, n, any_hit, stream, h);
        counts[0] = n;
        counts[1] = h[0];
        counts[2] = h[1];
    });
}

rq_status rq_trace_profile(rq_scene* scene, const rq_ray* d_rays, uint32_t n, uint32_t any_hit, uint64_t* counts,
                           void* stream) {
    return guarded([&] {
        require(scene && counts, "null argument");
        require(scene->accel == RQ_ACCEL_BVH, "traversal profile covers the BVH kernel");
        unsigned long long h[5] = {0, 0, 0, 0, 0};
        count_trace(scene, d_rays, n, any_hit, stream, h);
        counts[0] = n;
        for (int k = 0; k < 5; ++k) counts[1 + k] = h[k];
    });
}

rq_status rq_photons_create(rq_device* dev, const float* pos, const float* normal, const float* power, uint32_t n,
                            uint32_t layout, rq_photons** out) {
    return guarded([&] {
        require(dev && out, "null argument");
        require(n == 0 || (pos && normal), "null photon arrays");
        if (n > kMaxCount) throw Error(RQ_ERR_CAPACITY, "at most 2^30 photons per map");
        require(layout == RQ_PHOTONS_ENCORE_ORDER || layout == RQ_PHOTONS_SPATIAL, "unknown photon layout");
        RQ_HIP(hipSetDevice(dev->ordinal));
        PhotonHost h;
        build_photons(pos, normal, power, n, layout, h);
        auto p = std::make_unique<rq_photons>();
        p->dev = dev;
        p->n = n;
        p->nclasses = (uint32_t)h.normals.size();
        p->table_mask = (uint32_t)(h.table.size() / 4 - 1);
        const size_t m = std::max<size_t>(1, n);
        p->hot = DevBuf(m * sizeof(PhotonHot));
        p->power = DevBuf(m * sizeof(PhotonPower));
        p->perm = DevBuf(m * sizeof(uint32_t));
        p->table = DevBuf(h.table.size() * sizeof(uint32_t));
        if (n) {
            p->hot.upload(h.hot.data(), n * sizeof(PhotonHot));
            p->power.upload(h.power.data(), n * sizeof(PhotonPower));
            p->perm.upload(h.perm.data(), n * sizeof(uint32_t));
        }
        p->table.upload(h.table.data(), p->table.bytes);
        // Photons a query can select, and their bounds (the cell grid's).
        bool any = false;
        for (uint32_t i = 0; i < n; ++i) {
            const PhotonHot& ph = h.hot[i];
            if (!std::isfinite(ph.x) || !std::isfinite(ph.y) || !std::isfinite(ph.z) ||
                (ph.meta >> 2) == kClassNever)
                continue;
            const float c[3] = {ph.x, ph.y, ph.z};
            for (int a = 0; a < 3; ++a) {
                p->bmin[a] = any ? std::min(p->bmin[a], c[a]) : c[a];
                p->bmax[a] = any ? std::max(p->bmax[a], c[a]) : c[a];
            }
            any = true;
            ++p->nvalid;
        }
        p->perm_host = std::move(h.perm);
        p->axis_host = std::move(h.axis);
        *out = p.release();
    });
}

void rq_photons_destroy(rq_photons* p) {
    if (!p) return;
    (void)hipSetDevice(p->dev->ordinal);
    delete p;
}

rq_status rq_photons_order(const rq_photons* p, uint32_t* perm, uint8_t* axis, uint32_t* nclasses) {
    return guarded([&] {
        require(p != nullptr, "null photon map");
        if (perm && p->n) std::memcpy(perm, p->perm_host.data(), p->n * sizeof(uint32_t));
        if (axis && p->n) std::memcpy(axis, p->axis_host.data(), p->n);
        if (nclasses) *nclasses = p->nclasses;
    });
}

rq_status rq_photons_balance_host(const float* pos, uint32_t n, uint32_t layout, uint32_t* perm, uint8_t* axis) {
    return guarded([&] {
        require(n == 0 || (pos && perm && axis), "null argument");
        require(layout == RQ_PHOTONS_ENCORE_ORDER || layout == RQ_PHOTONS_SPATIAL, "unknown photon layout");
        if (n > kMaxCount) throw Error(RQ_ERR_CAPACITY, "at most 2^30 photons per map");
        std::vector<float> nrm(3 * (size_t)n, 0.0f);
        PhotonHost h;
        build_photons(pos, nrm.data(), nullptr, n, layout, h);
        if (n) {
            std::memcpy(perm, h.perm.data(), n * sizeof(uint32_t));
            std::memcpy(axis, h.axis.data(), n);
        }
    });
}

// Kernel switches, read at every call so tests and sweeps can flip them:
// RQ_KNN_LANE=1 keeps the per-lane heap kernel for every k; RQ_KNN_CELLS=0
// keeps the tree walks for large k (no cell grid); RQ_KNN_SORT=0 keeps the
// caller's query order in a cell-grid gather.
static bool env_is(const char* name, char v) {
    const char* e = std::getenv(name);
    return e && e[0] == v;
}
static bool knn_force_lane() { return env_is("RQ_KNN_LANE", '1'); }
static bool knn_cells_enabled() { return !env_is("RQ_KNN_CELLS", '0'); }
static bool knn_cells_sorted() { return !env_is("RQ_KNN_SORT", '0'); }
// Queries of a cell-grid gather are sorted by cell from this many on.
constexpr uint32_t kKnnSortMin = 16384;

// tally (device, 4 words, zeroed): counts of the cell-grid gather's
// instrumented kernel; left zero when another kernel serves the call.
static void knn_common(rq_photons* p, const float* qpos, const float* qnrm, uint32_t nq, uint32_t k, float max_dist,
                       uint32_t* idx, float* d2, uint32_t* cnt, float* rgb, uint32_t caustic, hipStream_t s,
                       unsigned long long* tally = nullptr) {
    require(p != nullptr, "null photon map");
    require(k >= 1 && k <= kKnnMaxK, "k must be 1..256");
    require(nq == 0 || (qpos && qnrm), "null query arrays");
    if (nq == 0) return;
    if (nq > kMaxCount) throw Error(RQ_ERR_CAPACITY, "at most 2^30 queries per call");
    rq_device* dev = p->dev;
    RQ_HIP(hipSetDevice(dev->ordinal));
    // Large k: one query per wave, from the photon cell grid when one suits
    // the radius; else radiance without lists walks the tree, one query per
    // wave, and lists come from the per-lane heap kernel.
    const bool wave = idx == nullptr && rgb != nullptr && k >= kKnnWaveMinK && !knn_force_lane();
    const float r2 = max_dist * max_dist;  // fl32(maxDistance * maxDistance) (PhotonMap.cpp:89)
    if (k >= kKnnWaveMinK && !knn_force_lane() && knn_cells_enabled()) {
        std::lock_guard<std::mutex> g(p->grid_mu);
        if (const PhotonGrid* pg = photon_grid_for(p, r2, s)) {
            KnnLaunch L{};
            L.power = p->power.as<PhotonPower>();
            L.perm = p->perm.as<uint32_t>();
            L.table = p->table.as<uint32_t>();
            L.table_mask = p->table_mask;
            L.n = p->n;
            L.qpos = qpos;
            L.qnrm = qnrm;
            L.nq = nq;
            L.k = k;
            L.r2 = r2;
            L.out_idx = idx;
            L.out_d2 = d2;
            L.out_cnt = cnt;
            L.out_rgb = rgb;
            L.caustic = caustic;
            L.grid = pg->view;
            L.tally = tally;
            const bool sorted = nq >= kKnnSortMin && knn_cells_sorted();
            const size_t prep_bytes = ((size_t)nq * 16 + 255) & ~(size_t)255;
            const size_t sort_bytes = sorted ? knn_cell_order_bytes(nq) : 0;
            const size_t order_bytes = sorted ? ((size_t)nq * 4 + 255) & ~(size_t)255 : 0;
            uint32_t* scratch = stream_scratch(dev, s, prep_bytes + order_bytes + sort_bytes);
            char* w = reinterpret_cast<char*>(scratch + 64);
            float4* prep = reinterpret_cast<float4*>(w);
            if (sorted) {
                uint32_t* order = reinterpret_cast<uint32_t*>(w + prep_bytes);
                knn_cell_order(L, order, w + prep_bytes + order_bytes, sort_bytes, s);
                L.order = order;
            }
            launch_knn_cells(L, prep, knn_cell_grid(nq, dev->num_cus), s);
            return;
        }
    }
    if (tally) throw Error(RQ_ERR_UNSUPPORTED, "counts come from the cell-grid gather, which does not serve this call");
    const int grid = wave ? knn_wave_grid(nq, dev->num_cus) : knn_grid(nq, dev->num_cus, k);
    const uint32_t depth = knn_spill_words_per_lane(p->n);
    const uint32_t cap = knn_wave_stack_cap(p->n);
    const size_t spill_bytes = ((wave ? (size_t)grid * 3 * cap * sizeof(uint32_t)
                                      : (size_t)grid * 64 * depth * 3 * sizeof(uint32_t)) + 255) & ~(size_t)255;
    // Many queries are walked in the Morton order of their positions (the
    // results do not depend on the order).
    const bool sorted = nq >= kKnnSortMin && knn_cells_sorted() && p->nvalid > 0;
    const size_t order_bytes = sorted ? ((size_t)nq * 4 + 255) & ~(size_t)255 : 0;
    const size_t sort_bytes = sorted ? knn_tree_order_bytes(nq) : 0;
    uint32_t* scratch = stream_scratch(dev, s, spill_bytes + order_bytes + sort_bytes);
    KnnLaunch L{};
    L.hot = p->hot.as<PhotonHot>();
    L.power = p->power.as<PhotonPower>();
    L.perm = p->perm.as<uint32_t>();
    L.table = p->table.as<uint32_t>();
    L.table_mask = p->table_mask;
    L.n = p->n;
    L.qpos = qpos;
    L.qnrm = qnrm;
    L.nq = nq;
    L.k = k;
    L.r2 = r2;
    L.out_idx = idx;
    L.out_d2 = d2;
    L.out_cnt = cnt;
    L.out_rgb = rgb;
    L.caustic = caustic;
    L.spill = scratch + 64;
    L.spill_depth = depth;
    if (sorted) {
        char* w = reinterpret_cast<char*>(scratch + 64) + spill_bytes;
        uint32_t* order = reinterpret_cast<uint32_t*>(w);
        knn_tree_order(L, p->bmin, p->bmax, order, w + order_bytes, sort_bytes, s);
        L.order = order;
    }
    if (wave)
        launch_knn_wave(L, grid, cap, s);
    else
        launch_knn(L, grid, s);
}

rq_status rq_knn(rq_photons* p, const float* d_qpos, const float* d_qnormal, uint32_t nq, uint32_t k, float max_dist,
                 uint32_t* d_idx, float* d_d2, uint32_t* d_count, void* stream) {
    return guarded([&] {
        require(nq == 0 || (d_idx && d_d2), "null output arrays");
        knn_common(p, d_qpos, d_qnormal, nq, k, max_dist, d_idx, d_d2, d_count, nullptr, 0, (hipStream_t)stream);
    });
}

rq_status rq_radiance_estimate(rq_photons* p, const float* d_qpos, const float* d_qnormal, uint32_t nq, uint32_t k,
                               float max_dist, uint32_t caustic, float* d_rgb, uint32_t* d_idx, float* d_d2,
                               uint32_t* d_count, void* stream) {
    return guarded([&] {
        require(nq == 0 || d_rgb, "null radiance output");
        require((d_idx == nullptr) == (d_d2 == nullptr), "d_idx and d_d2 go together");
        knn_common(p, d_qpos, d_qnormal, nq, k, max_dist, d_idx, d_d2, d_count, d_rgb, caustic ? 1u : 0u,
                   (hipStream_t)stream);
    });
}

rq_status rq_radiance_counts(rq_photons* p, const float* d_qpos, const float* d_qnormal, uint32_t nq, uint32_t k,
                             float max_dist, uint32_t caustic, float* d_rgb, uint64_t* counts, void* stream) {
    return guarded([&] {
        require(counts != nullptr, "null counts");
        require(nq == 0 || d_rgb, "null radiance output");
        counts[0] = counts[1] = counts[2] = counts[3] = 0;
        if (nq == 0) return;
        require(p != nullptr, "null photon map");
        RQ_HIP(hipSetDevice(p->dev->ordinal));
        DevBuf t(4 * sizeof(unsigned long long));
        launch_zero_u32(t.as<uint32_t>(), 8, (hipStream_t)stream);
        knn_common(p, d_qpos, d_qnormal, nq, k, max_dist, nullptr, nullptr, nullptr, d_rgb, caustic ? 1u : 0u,
                   (hipStream_t)stream, t.as<unsigned long long>());
        RQ_HIP(hipMemcpyAsync(counts, t.ptr, 4 * sizeof(uint64_t), hipMemcpyDeviceToHost, (hipStream_t)stream));
        RQ_HIP(hipStreamSynchronize((hipStream_t)stream));
    });
}

rq_status rq_photons_precompute_irradiance(rq_photons* src, const float* pos, const float* normal, uint32_t n,
                                           uint32_t increment, uint32_t k, float max_dist, uint32_t caustic,
                                           uint32_t layout, uint32_t* chosen, float* power_out, rq_photons** out) {
    return guarded([&] {
        require(src != nullptr && out != nullptr, "null argument");
        require(n == src->n, "n must be the map's photon count");
        require(n == 0 || (pos && normal), "null photon arrays");
        require(increment >= 1, "increment must be at least 1");
        require(layout == RQ_PHOTONS_ENCORE_ORDER || layout == RQ_PHOTONS_SPATIAL, "unknown photon layout");
        require(k >= 1 && k <= kKnnMaxK, "k must be 1..256");
        // m_photons after BalanceAndConstruct: ENCORE's balanced order, which
        // depends on the positions only (whatever layout the map uses).
        PhotonHost h;
        if (n) {
            const std::vector<float> zero(3 * (size_t)n, 0.0f);
            build_photons(pos, zero.data(), nullptr, n, RQ_PHOTONS_ENCORE_ORDER, h);
        }
        const uint32_t m = n == 0 ? 0u : (n - 1) / increment + 1;
        std::vector<uint32_t> pick(m);
        std::vector<float> qp(3 * (size_t)m), qn(3 * (size_t)m), rgb(3 * (size_t)m, 0.0f);
        for (uint32_t j = 0; j < m; ++j) {
            const uint32_t i = h.perm[(size_t)j * increment];
            pick[j] = i;
            for (int a = 0; a < 3; ++a) {
                qp[3 * (size_t)j + a] = pos[3 * (size_t)i + a];
                qn[3 * (size_t)j + a] = normal[3 * (size_t)i + a];
            }
        }
        if (m) {
            RQ_HIP(hipSetDevice(src->dev->ordinal));
            const size_t bytes = 3 * (size_t)m * sizeof(float);
            DevBuf dq(bytes), dn(bytes), drgb(bytes);
            dq.upload(qp.data(), bytes);
            dn.upload(qn.data(), bytes);
            knn_common(src, dq.as<float>(), dn.as<float>(), m, k, max_dist, nullptr, nullptr, nullptr, drgb.as<float>(),
                       caustic ? 1u : 0u, nullptr);
            RQ_HIP(hipMemcpy(rgb.data(), drgb.ptr, bytes, hipMemcpyDeviceToHost));
        }
        rq_photons* made = nullptr;
        const rq_status st = rq_photons_create(src->dev, qp.data(), qn.data(), rgb.data(), m, layout, &made);
        if (st != RQ_OK) throw Error(st, rq_last_error());
        if (chosen && m) std::memcpy(chosen, pick.data(), m * sizeof(uint32_t));
        if (power_out && m) std::memcpy(power_out, rgb.data(), 3 * (size_t)m * sizeof(float));
        *out = made;
    });
}

rq_status rq_final_gather(rq_scene* scene, rq_photons* indirect, uint32_t k_ind, float r_ind, rq_photons* direct,
                          uint32_t k_dir, float r_dir, const float* d_diffuse, const float* d_origin,
                          uint32_t npoints, const float* d_dirs, uint32_t nper, float* d_rgb, float* d_ri,
                          void* stream) {
    return guarded([&] {
        require(scene && indirect, "null scene or photon map");
        require(nper >= 1, "at least one sample per point");
        require(npoints == 0 || (d_diffuse && d_origin && d_dirs && d_rgb), "null argument");
        require(indirect->dev == scene->dev && (!direct || direct->dev == scene->dev), "objects on different devices");
        if (npoints == 0) return;
        hipStream_t s = (hipStream_t)stream;
        RQ_HIP(hipSetDevice(scene->dev->ordinal));
        // Chunks of whole shading points, at most 4M sample rays each.
        const uint64_t max_rays = std::max<uint64_t>(nper, 1u << 22);
        const uint32_t chunk_pts = (uint32_t)std::max<uint64_t>(1, max_rays / nper);
        const uint64_t cap = (uint64_t)std::min<uint64_t>(chunk_pts, npoints) * nper;
        if (cap > kMaxCount) throw Error(RQ_ERR_CAPACITY, "at most 2^30 sample rays per shading point");
        const size_t per_ray = sizeof(rq_ray) + sizeof(rq_hit) + 4 * 3 * sizeof(float);
        DevBuf work(cap * per_ray + 1024);
        char* w = static_cast<char*>(work.ptr);
        rq_ray* rays = reinterpret_cast<rq_ray*>(w);
        rq_hit* hits = reinterpret_cast<rq_hit*>(w + cap * sizeof(rq_ray));
        float* qpos = reinterpret_cast<float*>(w + cap * (sizeof(rq_ray) + sizeof(rq_hit)));
        float* qnrm = qpos + 3 * cap;
        float* rgb_ind = qnrm + 3 * cap;
        float* rgb_dir = rgb_ind + 3 * cap;
        for (uint32_t p0 = 0; p0 < npoints; p0 += chunk_pts) {
            const uint32_t np = std::min(chunk_pts, npoints - p0);
            FgLaunch L{};
            L.origin = d_origin + 3 * (size_t)p0;
            L.dirs = d_dirs + 3 * (size_t)p0 * nper;
            L.npoints = np;
            L.nper = nper;
            L.nrays = np * nper;
            L.rays = rays;
            L.hits = hits;
            L.tris = scene->tris.as<TriRec>();
            L.prim_slot = scene->prim_slot.as<uint32_t>();
            L.predicate = scene->predicate;
            L.qpos = qpos;
            L.qnrm = qnrm;
            L.rgb_ind = rgb_ind;
            L.rgb_dir = direct ? rgb_dir : nullptr;
            L.diffuse = d_diffuse;
            L.out_rgb = d_rgb + 3 * (size_t)p0;
            L.out_ri = d_ri ? d_ri + p0 : nullptr;
            launch_fg_rays(L, s);
            trace(scene, rays, L.nrays, hits, nullptr, TraceKind::Closest, s);
            launch_fg_attr(L, s);
            knn_common(indirect, qpos, qnrm, L.nrays, k_ind, r_ind, nullptr, nullptr, nullptr, rgb_ind, 0, s);
            if (direct) knn_common(direct, qpos, qnrm, L.nrays, k_dir, r_dir, nullptr, nullptr, nullptr, rgb_dir, 0, s);
            launch_fg_reduce(L, s);
        }
        // The workspace is released when this call returns.
        RQ_HIP(hipStreamSynchronize(s));
    });
}

rq_status rq_host_tree_build(const float* xyz, uint32_t nverts, const uint32_t* tri_idx, uint32_t ntris,
                             const rq_scene_desc* desc, rq_host_tree** out) {
    return guarded([&] {
        require(xyz && tri_idx && out, "null argument");
        require(ntris > 0 && nverts > 0, "empty triangle list");
        rq_scene_desc dd{};
        if (desc) dd = *desc;
        require(dd.accel == RQ_ACCEL_BVH || dd.accel == RQ_ACCEL_KD, "unknown accel kind");
        auto t = std::make_unique<rq_host_tree>();
        t->accel = dd.accel;
        t->ntris = ntris;
        if (dd.accel == RQ_ACCEL_BVH) {
            require(dd.bvh_width == 0 || dd.bvh_width == 2 || dd.bvh_width == 4, "bvh_width must be 0, 2 or 4");
            build_bvh(xyz, nverts, tri_idx, ntris, dd.leaf_size, dd.bvh_width ? dd.bvh_width : 4, t->bvh);
        } else {
            float bmin[3], bmax[3];
            vertex_bounds(xyz, nverts, dd, bmin, bmax);
            build_kd(xyz, nverts, tri_idx, ntris, bmin, bmax, t->kd);
        }
        *out = t.release();
    });
}

rq_status rq_host_tree_info(const rq_host_tree* t, rq_scene_info* info) {
    return guarded([&] {
        require(t && info, "null argument");
        const BuildStats& s = t->accel == RQ_ACCEL_BVH ? t->bvh.stats : t->kd.stats;
        *info = rq_scene_info{};
        info->accel = t->accel;
        info->ntris = t->ntris;
        info->nnodes = s.nnodes;
        info->nleaves = s.nleaves;
        info->nrefs = s.nrefs;
        info->max_depth = s.max_depth;
        info->kd_max_depth_limit = s.kd_depth_limit;
        info->build_seconds = s.seconds;
        info->bvh_width = t->accel == RQ_ACCEL_BVH ? t->bvh.width : 0;
        info->stack_bound = s.stack_bound;
    });
}

rq_status rq_host_tree_export_textures(const rq_host_tree* t, const float* xyz, uint32_t nverts,
                                       const uint32_t* tri_idx, float* cell, float* v0, float* v1, float* v2,
                                       uint32_t* ref_prim) {
    return guarded([&] {
        require(t != nullptr, "null tree");
        if (t->accel != RQ_ACCEL_KD) throw Error(RQ_ERR_UNSUPPORTED, "the texture encoding is the kd tree's");
        require((!v0 && !v1 && !v2) || (xyz && tri_idx), "vertex outputs need the triangle list");
        const KdHost& h = t->kd;
        float first = 0.0f;  // currentTriangleIndex, a float in Kdtree::buildGPU
        size_t r = 0;
        float* vs[3] = {v0, v1, v2};
        for (size_t i = 0; i < h.nodes.size(); ++i) {
            const KdNodeG& nd = h.nodes[i];
            float c[4];
            if ((nd.word & 3u) != 3u) {
                float split;
                std::memcpy(&split, &nd.word, 4);  // SplitPos(): the word, axis bits included
                c[0] = (float)nd.child;
                c[1] = split;
                c[2] = 0.0f;
                c[3] = (float)-((int)(nd.word & 3u) + 2);
            } else {
                const uint32_t count = nd.word >> 2;
                c[0] = first;
                c[1] = 0.0f;
                c[2] = 0.0f;
                c[3] = (float)(int)count;
                for (uint32_t j = 0; j < count; ++j, ++r) {
                    const uint32_t prim = h.leaf_prims[(size_t)nd.child + j];
                    first += (float)(size_t)1;  // one triangle per TrianglePrim tessellation
                    if (ref_prim) ref_prim[r] = prim;
                    for (int k = 0; k < 3; ++k) {
                        if (!vs[k]) continue;
                        const uint32_t v = tri_idx[3 * (size_t)prim + k];
                        require(v < nverts, "triangle index out of range");
                        for (int a = 0; a < 3; ++a) vs[k][4 * r + a] = xyz[3 * (size_t)v + a];
                        vs[k][4 * r + 3] = 0.0f;
                    }
                }
            }
            if (cell)
                for (int a = 0; a < 4; ++a) cell[4 * i + a] = c[a];
        }
    });
}

rq_status rq_host_tree_copy(const rq_host_tree* t, void* nodes, uint32_t* leaf_prims, float* root_box) {
    return guarded([&] {
        require(t != nullptr, "null tree");
        if (t->accel == RQ_ACCEL_BVH) {
            if (nodes && t->bvh.width == 4)
                std::memcpy(nodes, t->bvh.nodes4.data(), t->bvh.nodes4.size() * sizeof(BvhNode4));
            else if (nodes)
                std::memcpy(nodes, t->bvh.nodes.data(), t->bvh.nodes.size() * sizeof(BvhNode));
            if (leaf_prims) std::memcpy(leaf_prims, t->bvh.tri_order.data(), t->bvh.tri_order.size() * 4);
            if (root_box)
                for (int k = 0; k < 6; ++k) root_box[k] = 0.0f;
        } else {
            if (nodes) std::memcpy(nodes, t->kd.nodes.data(), t->kd.nodes.size() * sizeof(KdNodeG));
            if (leaf_prims) std::memcpy(leaf_prims, t->kd.leaf_prims.data(), t->kd.leaf_prims.size() * 4);
            if (root_box)
                for (int k = 0; k < 3; ++k) {
                    root_box[k] = t->kd.root_pos[k];
                    root_box[3 + k] = t->kd.root_size[k];
                }
        }
    });
}

void rq_host_tree_destroy(rq_host_tree* t) { delete t; }

}  // extern "C"
