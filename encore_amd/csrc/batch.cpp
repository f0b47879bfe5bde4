// batch.cpp -- many ray batches (any sizes, any resident scenes of one kind:
// BVH, or ENCORE's kd-tree for the exact path) traced by one persistent
// launch, optionally captured once into a hipGraph and replayed; plus the
// autotuner that picks the launch configuration.
//
// A batch keeps a device table of 64-byte descriptors (scene nodes and
// triangles, ray and output buffers, first global ray index, count) and, for
// kd scenes, the root box of each item's tree.  The persistent kernel's waves
// pull global ray indices from one counter and find the owning descriptor by
// binary search, so a 4k-ray batch and a 1M-ray batch share the GPU without a
// launch each.
#include <algorithm>
#include <chrono>
#include <utility>
#include <vector>

#include "rq_internal.h"

struct rq_batch {
    rq_device* dev = nullptr;
    uint32_t any_hit = 0;
    uint32_t accel = RQ_ACCEL_BVH;  // of every item
    uint32_t predicate = 0;
    uint32_t total = 0, nitems = 0, width = 2;
    uint32_t spill_words = 0;  // per-lane spill region covering every item's scene
    std::vector<rq::BatchItemDev> items;
    rq::DevBuf items_dev;
    rq::DevBuf boxes_dev;  // kd: root pos[3], pad, size[3], pad per item
    rq::DevBuf scratch;  // counter (256 B) + spill stacks; owned so a graph can bake it
    rq::KernelConfig kc;
    int grid = 0;
    hipStream_t capture = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;

    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
    ~rq_batch() {
        drop_graph();
        if (capture) (void)hipStreamDestroy(capture);
    }
};

namespace rq {

static rq::TraceKind kind_of(const rq_batch* b) { return b->any_hit ? TraceKind::Any : TraceKind::Closest; }

// Grid and scratch for the current configuration; returns true if the
// scratch buffer moved (a captured graph must then be rebuilt).
static bool prepare(rq_batch* b) {
    const bool kd = b->accel == RQ_ACCEL_KD;
    int per_cu = b->kc.blocks_per_cu;
    const int occ = kd ? occupancy_kd_batched(b->predicate, kind_of(b))
                       : occupancy_bvh_batched(b->predicate, kind_of(b), b->width, b->kc.stack);
    if (per_cu <= 0 || per_cu > occ) per_cu = occ;
    int64_t grid = (int64_t)per_cu * b->dev->num_cus;
    // BVH: a lane per ray; kd: up to a wave per ray (the kernel spreads a
    // small launch over every wave and tests long leaves with idle lanes).
    const int64_t block = kd ? kKdBlock : kBvhBlock;
    const int64_t per_block = kd ? block / 64 : block;
    grid = std::max<int64_t>(1, std::min<int64_t>(grid, ((int64_t)b->total + per_block - 1) / per_block));
    b->grid = (int)grid;
    const size_t need = 256 + (size_t)b->grid * block * b->spill_words * sizeof(uint32_t);
    if (b->scratch.bytes < need) {
        RQ_HIP(hipDeviceSynchronize());
        b->scratch = DevBuf(need);
        return true;
    }
    return false;
}

static TraceLaunch launch_params(const rq_batch* b) {
    TraceLaunch L{};
    L.n = b->total;
    L.counter = b->scratch.as<uint32_t>();
    L.spill = b->scratch.as<uint32_t>() + 64;
    L.spill_depth = b->spill_words;
    L.items = b->items_dev.as<BatchItemDev>();
    L.nitems = b->nitems;
    if (b->accel == RQ_ACCEL_KD) {
        L.fetch_threshold = (uint32_t)b->kc.kd_fetch_threshold;
        L.item_boxes = b->boxes_dev.as<float>();
        L.coop_lanes = kKdCoopLanesDefault;
        L.coop_min = kKdCoopMinDefault;
    } else {
        L.fetch_threshold = (uint32_t)b->kc.fetch_threshold;
        L.leaf_batch = bvh_leaf_batch();
    }
    return L;
}

static void enqueue(rq_batch* b, hipStream_t s) {
    const TraceLaunch L = launch_params(b);
    launch_zero_u32(L.counter, 2, s);
    if (b->accel == RQ_ACCEL_KD)
        launch_kd_batched(L, b->predicate, kind_of(b), b->grid, s);
    else
        launch_bvh_batched(L, b->predicate, kind_of(b), b->width, b->grid, b->kc, s);
}

static void launch(rq_batch* b, bool use_graph, hipStream_t s) {
    if (b->total == 0) return;
    RQ_HIP(hipSetDevice(b->dev->ordinal));
    if (prepare(b)) b->drop_graph();
    if (!use_graph) {
        enqueue(b, s);
        return;
    }
    if (!b->exec) {
        if (!b->capture) RQ_HIP(hipStreamCreateWithFlags(&b->capture, hipStreamNonBlocking));
        RQ_HIP(hipStreamBeginCapture(b->capture, hipStreamCaptureModeThreadLocal));
        try {
            enqueue(b, b->capture);
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(b->capture, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        RQ_HIP(hipStreamEndCapture(b->capture, &b->graph));
        RQ_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    RQ_HIP(hipGraphLaunch(b->exec, s));
}

}  // namespace rq

using namespace rq;

namespace {

template <class F>
rq_status run(F&& f) {
    return rq::guard_call([](void* c) { (*static_cast<F*>(c))(); }, &f);
}

void require(bool ok, const char* msg) {
    if (!ok) throw rq::Error(RQ_ERR_INVALID_ARG, msg);
}

}  // namespace

extern "C" {

rq_status rq_batch_create(rq_device* dev, uint32_t any_hit, rq_batch** out) {
    return run([&] {
        require(dev && out, "null argument");
        auto b = std::make_unique<rq_batch>();
        b->dev = dev;
        b->any_hit = any_hit ? 1u : 0u;
        b->kc = dev->kc;
        *out = b.release();
    });
}

rq_status rq_batch_set(rq_batch* b, const rq_batch_item* items, uint32_t count) {
    return run([&] {
        require(b != nullptr, "null batch");
        require(count == 0 || items, "null items");
        std::vector<BatchItemDev> v;
        std::vector<float> boxes;
        uint64_t total = 0;
        uint32_t pred = 0, words = 0, width = 2, accel = RQ_ACCEL_BVH;
        bool first = true;
        for (uint32_t i = 0; i < count; ++i) {
            const rq_batch_item& it = items[i];
            require(it.scene != nullptr, "item without scene");
            if (it.scene->dev != b->dev)
                throw Error(RQ_ERR_INVALID_ARG, "item scene lives on another device");
            if (first) {
                pred = it.scene->predicate;
                width = it.scene->width;
                accel = it.scene->accel;
                first = false;
            } else if (it.scene->accel != accel) {
                throw Error(RQ_ERR_UNSUPPORTED, "all items of a batch must use one kind of scene (BVH or kd)");
            } else if (it.scene->predicate != pred) {
                throw Error(RQ_ERR_UNSUPPORTED, "all items of a batch must use one predicate");
            } else if (accel == RQ_ACCEL_BVH && it.scene->width != width) {
                throw Error(RQ_ERR_UNSUPPORTED, "all items of a batch must use one BVH width");
            }
            if (it.n == 0) continue;
            require(it.d_rays && it.d_out, "null ray or output buffer");
            BatchItemDev d{};
            d.nodes = it.scene->nodes.ptr;
            d.tris = it.scene->tris.as<TriRec>();
            d.rays = it.d_rays;
            d.hits = b->any_hit ? nullptr : static_cast<rq_hit*>(it.d_out);
            d.occ = b->any_hit ? static_cast<uint32_t*>(it.d_out) : nullptr;
            d.begin = (uint32_t)total;
            d.n = it.n;
            total += it.n;
            if (total > 0xFFFFFFF0ull) throw Error(RQ_ERR_CAPACITY, "a batch holds at most 2^32-16 rays");
            // The kd kernel's 32-bit claim counter runs past the total by up
            // to one claim per wave.
            if (accel == RQ_ACCEL_KD && total > (1ull << 31))
                throw Error(RQ_ERR_CAPACITY, "a batch of kd scenes holds at most 2^31 rays");
            words = std::max(words, scene_spill_words(it.scene));
            v.push_back(d);
            if (accel == RQ_ACCEL_KD) {
                const float* p = it.scene->root_pos;
                const float* z = it.scene->root_size;
                const float box[8] = {p[0], p[1], p[2], 0.0f, z[0], z[1], z[2], 0.0f};
                boxes.insert(boxes.end(), box, box + 8);
            }
        }
        RQ_HIP(hipSetDevice(b->dev->ordinal));
        // Queued launches may still read the old table.
        RQ_HIP(hipDeviceSynchronize());
        if (b->items_dev.bytes < std::max<size_t>(1, v.size()) * sizeof(BatchItemDev)) {
            b->items_dev = DevBuf(std::max<size_t>(1, v.size()) * sizeof(BatchItemDev));
        }
        if (!v.empty()) b->items_dev.upload(v.data(), v.size() * sizeof(BatchItemDev));
        if (b->boxes_dev.bytes < std::max<size_t>(1, boxes.size()) * sizeof(float))
            b->boxes_dev = DevBuf(std::max<size_t>(1, boxes.size()) * sizeof(float));
        if (!boxes.empty()) b->boxes_dev.upload(boxes.data(), boxes.size() * sizeof(float));
        b->items = std::move(v);
        b->accel = accel;
        b->nitems = (uint32_t)b->items.size();
        b->total = (uint32_t)total;
        b->predicate = pred;
        b->width = width;
        b->spill_words = words;
        b->drop_graph();
    });
}

rq_status rq_batch_launch(rq_batch* b, uint32_t use_graph, void* stream) {
    return run([&] {
        require(b != nullptr, "null batch");
        launch(b, use_graph != 0, (hipStream_t)stream);
    });
}

rq_status rq_batch_get_config(const rq_batch* b, rq_kernel_config* cfg) {
    return run([&] {
        require(b && cfg, "null argument");
        const bool kd = b->accel == RQ_ACCEL_KD;
        cfg->stack = kd ? (uint32_t)kKdStackEntries : (uint32_t)b->kc.stack;
        cfg->blocks_per_cu = (uint32_t)b->kc.blocks_per_cu;
        cfg->fetch_threshold = (uint32_t)(kd ? b->kc.kd_fetch_threshold : b->kc.fetch_threshold);
        cfg->grid = (uint32_t)b->grid;
    });
}

rq_status rq_batch_set_config(rq_batch* b, const rq_kernel_config* cfg) {
    return run([&] {
        require(b && cfg, "null argument");
        require(cfg->stack == 8 || cfg->stack == 16 || cfg->stack == 32, "stack must be 8, 16 or 32");
        require(cfg->fetch_threshold <= 64, "fetch_threshold must be 0..64");
        b->kc.blocks_per_cu = (int)cfg->blocks_per_cu;
        if (b->accel == RQ_ACCEL_KD) {
            b->kc.kd_fetch_threshold = (int)cfg->fetch_threshold;  // the kd ring has a fixed depth
        } else {
            b->kc.stack = (int)cfg->stack;
            b->kc.fetch_threshold = (int)cfg->fetch_threshold;
        }
        b->drop_graph();
    });
}

rq_status rq_batch_autotune(rq_batch* b, void* stream, rq_kernel_config* chosen, double* best_ms) {
    return run([&] {
        require(b != nullptr, "null batch");
        hipStream_t s = (hipStream_t)stream;
        RQ_HIP(hipSetDevice(b->dev->ordinal));
        hipEvent_t e0, e1;
        RQ_HIP(hipEventCreate(&e0));
        RQ_HIP(hipEventCreate(&e1));
        const KernelConfig saved = b->kc;
        KernelConfig best = saved;
        float best_t = 1e30f;
        const bool kd = b->accel == RQ_ACCEL_KD;
        const std::vector<int> stacks = kd ? std::vector<int>{kKdStackEntries} : std::vector<int>{8, 16, 32};
        // kd lanes hold rays of very different cost, and refills claim rays
        // with an atomic per round: lower thresholds.
        const std::vector<int> thresholds =
            kd ? std::vector<int>{16, 32, 48, 56} : std::vector<int>{32, 44, 48, 56, 60};
        // Mean time of `reps` back-to-back launches with configuration kc.
        auto time_config = [&](const KernelConfig& kc, int reps) {
            b->kc = kc;
            launch(b, false, s);  // warm
            RQ_HIP(hipEventRecord(e0, s));
            for (int r = 0; r < reps; ++r) launch(b, false, s);
            RQ_HIP(hipEventRecord(e1, s));
            RQ_HIP(hipEventSynchronize(e1));
            float ms = 0;
            RQ_HIP(hipEventElapsedTime(&ms, e0, e1));
            return ms / (float)reps;
        };
        try {
            // One launch per configuration, then the fastest few again over
            // several launches: a single launch varies by a few percent, and
            // the minimum of one-shot times favours a lucky launch.
            std::vector<std::pair<float, KernelConfig>> sweep;
            for (int st : stacks) {
                const int occ = kd ? occupancy_kd_batched(b->predicate, kind_of(b))
                                   : occupancy_bvh_batched(b->predicate, kind_of(b), b->width, st);
                for (int bpc = 1; bpc <= occ; bpc = (bpc < 4 ? bpc + 1 : bpc + 2)) {
                    for (int th : thresholds) {
                        KernelConfig kc = saved;
                        kc.blocks_per_cu = bpc;
                        if (kd) {
                            kc.kd_fetch_threshold = th;
                        } else {
                            kc.stack = st;
                            kc.fetch_threshold = th;
                        }
                        sweep.emplace_back(time_config(kc, 1), kc);
                    }
                }
            }
            std::stable_sort(sweep.begin(), sweep.end(),
                             [](const std::pair<float, KernelConfig>& x, const std::pair<float, KernelConfig>& y) {
                                 return x.first < y.first;
                             });
            const size_t finalists = std::min<size_t>(sweep.size(), 6);
            for (size_t i = 0; i < finalists; ++i) {
                const float ms = time_config(sweep[i].second, 5);
                if (ms < best_t) {
                    best_t = ms;
                    best = sweep[i].second;
                }
            }
        } catch (...) {
            b->kc = saved;
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            throw;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        b->kc = best;
        b->drop_graph();
        prepare(b);
        if (chosen) {
            chosen->stack = kd ? (uint32_t)kKdStackEntries : (uint32_t)best.stack;
            chosen->blocks_per_cu = (uint32_t)best.blocks_per_cu;
            chosen->fetch_threshold = (uint32_t)(kd ? best.kd_fetch_threshold : best.fetch_threshold);
            chosen->grid = (uint32_t)b->grid;
        }
        if (best_ms) *best_ms = best_t;
    });
}

void rq_batch_destroy(rq_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->dev->ordinal);
    (void)hipDeviceSynchronize();
    delete b;
}

}  // extern "C"
