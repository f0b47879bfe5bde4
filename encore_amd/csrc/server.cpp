// server.cpp -- continuous-batching ray-query server.
//
// Scenes stay resident in HBM; clients submit ray batches (host buffers that
// stay valid until completion; pinned memory lets the copies overlap
// kernels).  A scheduler thread:
//   1. retires finished launches (event query): requests become done and
//      their device memory returns to the arena;
//   2. admits queued requests in arrival order while the arena -- a slice of
//      HBM sized from hipMemGetInfo at creation, or a configured budget --
//      can hold their rays and outputs (admission by free device memory; a
//      request that does not fit waits, it does not fail);
//   3. whenever one of `slots` launch slots is free, takes every admitted
//      request of the oldest one's kind and issues, on that slot's stream:
//      copy-in of each request's rays, one persistent batched launch over
//      all of them, copy-out of each output, an event.
// Requests arriving while launches run are admitted as memory frees and go
// into the next launch on another stream, so their work overlaps the
// running one.  Per-ray arithmetic never depends on the mix, so a result is
// bit-identical alone and in any batch.
//
// plain = 1 turns the server into the per-request baseline: one request per
// launch, one slot, the plain launch path (rq_intersect_closest's kernel).
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <thread>
#include <unordered_map>

#include "rq_internal.h"

namespace rq {
namespace {

using Clock = std::chrono::steady_clock;

// First-fit arena over one device allocation; offsets aligned to 256 B.
struct Arena {
    DevBuf buf;
    uint64_t size = 0, used = 0;
    std::map<uint64_t, uint64_t> free_;  // offset -> size

    void init(uint64_t bytes) {
        buf = DevBuf(bytes);
        size = bytes;
        used = 0;
        free_.clear();
        free_[0] = bytes;
    }
    static uint64_t round(uint64_t b) { return (b + 255) & ~255ull; }
    bool alloc(uint64_t bytes, uint64_t& off) {
        bytes = round(bytes);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                off = it->first;
                const uint64_t rest = it->second - bytes;
                free_.erase(it);
                if (rest) free_[off + bytes] = rest;
                used += bytes;
                return true;
            }
        }
        return false;
    }
    void release(uint64_t off, uint64_t bytes) {
        bytes = round(bytes);
        used -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    char* ptr(uint64_t off) const { return static_cast<char*>(buf.ptr) + off; }
};

enum class ReqState { Queued, Admitted, Launched, Done, Cancelled, Failed };

struct Request {
    uint64_t id = 0;
    rq_scene* scene = nullptr;
    const rq_ray* h_rays = nullptr;
    void* h_out = nullptr;
    uint32_t n = 0, any = 0;
    ReqState state = ReqState::Queued;
    Clock::time_point t_submit, t_admit, t_launch, t_done;
    uint64_t off = 0, bytes = 0;
    uint64_t launch = 0;
    uint32_t rays_in_launch = 0, reqs_in_launch = 0;
    rq_status err = RQ_OK;
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    DevBuf table;                 // BatchItemDev[cap]
    BatchItemDev* h_table = nullptr;  // pinned staging for the table
    uint32_t cap = 0;
    DevBuf scratch;
    std::vector<Request*> reqs;
    bool busy = false;
};

}  // namespace
}  // namespace rq

struct rq_server {
    rq_device* dev = nullptr;
    rq_server_config cfg{};
    rq::Arena arena;
    std::vector<rq::Slot> slots;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<rq::Request*> queue;     // not yet admitted
    std::deque<rq::Request*> admitted;  // device memory reserved, not launched
    std::unordered_map<uint64_t, std::unique_ptr<rq::Request>> reqs;
    uint64_t next_id = 1, launches = 0, rays_done = 0, reqs_done = 0;
    bool stop = false;
    std::thread worker;
    rq_status fatal = RQ_OK;
    std::string fatal_msg;
    rq::KernelConfig kc;
};

namespace rq {
namespace {

uint64_t request_bytes(uint32_t n, uint32_t any) {
    return Arena::round((uint64_t)n * sizeof(rq_ray)) + Arena::round((uint64_t)n * (any ? 4 : sizeof(rq_hit)));
}

void slot_init(rq_server* s, Slot& sl) {
    RQ_HIP(hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking));
    RQ_HIP(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
    sl.cap = std::max<uint32_t>(16, s->cfg.max_requests_per_launch);
    sl.table = DevBuf(sl.cap * sizeof(BatchItemDev));
    RQ_HIP(hipHostMalloc((void**)&sl.h_table, sl.cap * sizeof(BatchItemDev), hipHostMallocDefault));
    sl.scratch = DevBuf(256);
}

void slot_free(Slot& sl) {
    if (sl.h_table) (void)hipHostFree(sl.h_table);
    if (sl.done) (void)hipEventDestroy(sl.done);
    if (sl.stream) (void)hipStreamDestroy(sl.stream);
    sl.h_table = nullptr;
    sl.done = nullptr;
    sl.stream = nullptr;
}

// Issues one launch for the requests in sl.reqs (mutex not held).
void issue(rq_server* s, Slot& sl) {
    hipStream_t st = sl.stream;
    const bool any = sl.reqs.front()->any != 0;
    uint64_t total = 0;
    uint32_t words = 0;
    for (Request* r : sl.reqs) words = std::max(words, scene_spill_words(r->scene));
    const uint32_t pred = sl.reqs.front()->scene->predicate;
    const uint32_t width = sl.reqs.front()->scene->width;
    const TraceKind kind = any ? TraceKind::Any : TraceKind::Closest;
    int grid_max;
    if (s->cfg.plain) {
        grid_max = occupancy_blocks_per_cu(RQ_ACCEL_BVH, pred, kind, width, s->dev->kc) * s->dev->num_cus;
    } else {
        const int occ = occupancy_bvh_batched(pred, kind, width, s->kc.stack);
        grid_max = (s->kc.blocks_per_cu > 0 ? std::min(s->kc.blocks_per_cu, occ) : occ) * s->dev->num_cus;
    }
    // The slot's stream is idle here (its previous launch was retired), so
    // its scratch can be replaced.
    const size_t need = 256 + (size_t)grid_max * kBvhBlock * words * sizeof(uint32_t);
    if (sl.scratch.bytes < need) sl.scratch = DevBuf(need);
    for (Request* r : sl.reqs) {
        char* base = s->arena.ptr(r->off);
        RQ_HIP(hipMemcpyAsync(base, r->h_rays, (size_t)r->n * sizeof(rq_ray), hipMemcpyHostToDevice, st));
        total += r->n;
    }
    if (s->cfg.plain) {
        Request* r = sl.reqs.front();
        char* base = s->arena.ptr(r->off);
        TraceLaunch L{};
        L.nodes = r->scene->nodes.ptr;
        L.tris = r->scene->tris.as<TriRec>();
        L.rays = reinterpret_cast<const rq_ray*>(base);
        void* out = base + Arena::round((uint64_t)r->n * sizeof(rq_ray));
        L.hits = any ? nullptr : static_cast<rq_hit*>(out);
        L.occluded = any ? static_cast<uint32_t*>(out) : nullptr;
        L.n = r->n;
        L.counter = sl.scratch.as<uint32_t>();
        L.spill = sl.scratch.as<uint32_t>() + 64;
        L.spill_depth = words;
        L.fetch_threshold = (uint32_t)s->dev->kc.fetch_threshold;
        L.leaf_batch = bvh_leaf_batch();
        const int64_t grid = std::max<int64_t>(1, std::min<int64_t>(grid_max, ((int64_t)r->n + kBvhBlock - 1) / kBvhBlock));
        launch_zero_u32(L.counter, 2, st);
        launch_bvh(L, pred, kind, width, (int)grid, s->dev->kc, st);
    } else {
        uint32_t i = 0;
        uint64_t begin = 0;
        for (Request* r : sl.reqs) {
            char* base = s->arena.ptr(r->off);
            BatchItemDev d{};
            d.nodes = r->scene->nodes.ptr;
            d.tris = r->scene->tris.as<TriRec>();
            d.rays = reinterpret_cast<const rq_ray*>(base);
            void* out = base + Arena::round((uint64_t)r->n * sizeof(rq_ray));
            d.hits = any ? nullptr : static_cast<rq_hit*>(out);
            d.occ = any ? static_cast<uint32_t*>(out) : nullptr;
            d.begin = (uint32_t)begin;
            d.n = r->n;
            begin += r->n;
            sl.h_table[i++] = d;
        }
        RQ_HIP(hipMemcpyAsync(sl.table.ptr, sl.h_table, i * sizeof(BatchItemDev), hipMemcpyHostToDevice, st));
        TraceLaunch L{};
        L.n = (uint32_t)total;
        L.counter = sl.scratch.as<uint32_t>();
        L.spill = sl.scratch.as<uint32_t>() + 64;
        L.spill_depth = words;
        L.items = sl.table.as<BatchItemDev>();
        L.nitems = i;
        L.fetch_threshold = (uint32_t)s->kc.fetch_threshold;
        L.leaf_batch = bvh_leaf_batch();
        const int64_t grid = std::max<int64_t>(1, std::min<int64_t>(grid_max, ((int64_t)total + kBvhBlock - 1) / kBvhBlock));
        launch_zero_u32(L.counter, 2, st);
        launch_bvh_batched(L, pred, kind, width, (int)grid, s->kc, st);
    }
    for (Request* r : sl.reqs) {
        char* base = s->arena.ptr(r->off);
        const size_t ob = (size_t)r->n * (r->any ? 4 : sizeof(rq_hit));
        RQ_HIP(hipMemcpyAsync(r->h_out, base + Arena::round((uint64_t)r->n * sizeof(rq_ray)), ob,
                              hipMemcpyDeviceToHost, st));
    }
    RQ_HIP(hipEventRecord(sl.done, st));
}

void worker_loop(rq_server* s) {
    (void)hipSetDevice(s->dev->ordinal);
    std::unique_lock<std::mutex> lk(s->mu);
    while (true) {
        bool progress = false;
        // 1. retire
        for (Slot& sl : s->slots) {
            if (!sl.busy) continue;
            const hipError_t q = hipEventQuery(sl.done);
            if (q == hipErrorNotReady) continue;
            const auto now = Clock::now();
            for (Request* r : sl.reqs) {
                r->t_done = now;
                r->state = (q == hipSuccess) ? ReqState::Done : ReqState::Failed;
                if (q != hipSuccess) r->err = RQ_ERR_DEVICE;
                s->arena.release(r->off, r->bytes);
                s->rays_done += r->n;
                s->reqs_done += 1;
            }
            sl.reqs.clear();
            sl.busy = false;
            progress = true;
            s->cv_done.notify_all();
        }
        // 2. admit in arrival order
        while (!s->queue.empty()) {
            Request* r = s->queue.front();
            uint64_t off;
            if (!s->arena.alloc(r->bytes, off)) break;
            r->off = off;
            r->state = ReqState::Admitted;
            r->t_admit = Clock::now();
            s->queue.pop_front();
            s->admitted.push_back(r);
            progress = true;
        }
        // 3. launch on free slots
        for (Slot& sl : s->slots) {
            if (sl.busy || s->admitted.empty()) continue;
            const uint32_t any = s->admitted.front()->any;
            const uint32_t pred = s->admitted.front()->scene->predicate;
            const uint32_t width = s->admitted.front()->scene->width;
            uint64_t rays = 0;
            for (auto it = s->admitted.begin(); it != s->admitted.end();) {
                Request* r = *it;
                const bool fits = sl.reqs.size() < sl.cap && (sl.reqs.empty() || rays + r->n <= s->cfg.max_rays_per_launch);
                if (r->any == any && r->scene->predicate == pred && r->scene->width == width && fits) {
                    sl.reqs.push_back(r);
                    rays += r->n;
                    it = s->admitted.erase(it);
                    if (s->cfg.plain) break;
                } else {
                    ++it;
                }
            }
            if (sl.reqs.empty()) continue;
            const uint64_t lid = ++s->launches;
            const auto now = Clock::now();
            for (Request* r : sl.reqs) {
                r->state = ReqState::Launched;
                r->t_launch = now;
                r->launch = lid;
                r->rays_in_launch = (uint32_t)rays;
                r->reqs_in_launch = (uint32_t)sl.reqs.size();
            }
            sl.busy = true;
            lk.unlock();
            try {
                issue(s, sl);
            } catch (const Error& e) {
                lk.lock();
                s->fatal = e.status;
                s->fatal_msg = e.what();
                for (Request* r : sl.reqs) {
                    r->state = ReqState::Failed;
                    r->err = e.status;
                    s->arena.release(r->off, r->bytes);
                }
                sl.reqs.clear();
                sl.busy = false;
                s->cv_done.notify_all();
                lk.unlock();
            }
            lk.lock();
            progress = true;
        }
        if (s->stop) {
            bool idle = s->queue.empty() && s->admitted.empty();
            for (Slot& sl : s->slots) idle = idle && !sl.busy;
            if (idle) break;
        }
        if (!progress) {
            bool inflight = false;
            for (Slot& sl : s->slots) inflight = inflight || sl.busy;
            if (inflight)
                s->cv_work.wait_for(lk, std::chrono::microseconds(20));
            else
                s->cv_work.wait_for(lk, std::chrono::milliseconds(5));
        }
    }
}

}  // namespace
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

rq_status rq_server_create(rq_device* dev, const rq_server_config* cfg, rq_server** out) {
    return run([&] {
        require(dev && out, "null argument");
        auto s = std::make_unique<rq_server>();
        s->dev = dev;
        if (cfg) s->cfg = *cfg;
        if (s->cfg.slots == 0) s->cfg.slots = 3;
        if (s->cfg.plain) s->cfg.slots = 1;
        // 2^20 rays per launch: larger launches add their copies and
        // traversal to the latency of every request they hold (at 16M the
        // served p99 under load was a third higher for the same throughput).
        if (s->cfg.max_rays_per_launch == 0) s->cfg.max_rays_per_launch = 1u << 20;
        if (s->cfg.max_requests_per_launch == 0) s->cfg.max_requests_per_launch = 1024;
        s->kc = dev->kc;
        if (s->cfg.stack) s->kc.stack = (int)s->cfg.stack;
        if (s->cfg.blocks_per_cu) s->kc.blocks_per_cu = (int)s->cfg.blocks_per_cu;
        if (s->cfg.fetch_threshold) s->kc.fetch_threshold = (int)s->cfg.fetch_threshold;
        RQ_HIP(hipSetDevice(dev->ordinal));
        size_t free_b = 0, total_b = 0;
        RQ_HIP(hipMemGetInfo(&free_b, &total_b));
        uint64_t budget = s->cfg.hbm_budget_bytes;
        const uint64_t avail = free_b > s->cfg.hbm_reserve_bytes ? free_b - s->cfg.hbm_reserve_bytes : 0;
        if (budget == 0) budget = free_b / 4;
        if (budget > avail) budget = avail;
        if (budget < 4096) throw Error(RQ_ERR_OOM, "no device memory left for the request arena");
        s->arena.init(budget);
        s->slots.resize(s->cfg.slots);
        try {
            for (Slot& sl : s->slots) slot_init(s.get(), sl);
        } catch (...) {
            for (Slot& sl : s->slots) slot_free(sl);
            throw;
        }
        s->worker = std::thread(worker_loop, s.get());
        *out = s.release();
    });
}

rq_status rq_server_submit(rq_server* s, rq_scene* scene, const rq_ray* h_rays, uint32_t n, uint32_t any_hit,
                           void* h_out, uint64_t* ticket) {
    return run([&] {
        require(s && scene && ticket, "null argument");
        require(n == 0 || (h_rays && h_out), "null ray or output buffer");
        if (scene->dev != s->dev) throw Error(RQ_ERR_INVALID_ARG, "scene lives on another device");
        if (scene->accel != RQ_ACCEL_BVH) throw Error(RQ_ERR_UNSUPPORTED, "the server serves BVH scenes");
        const uint64_t bytes = request_bytes(n, any_hit);
        if (bytes > s->arena.size) throw Error(RQ_ERR_CAPACITY, "request larger than the server's HBM budget");
        if (n > kMaxCount) throw Error(RQ_ERR_CAPACITY, "at most 2^30 rays per request");
        auto r = std::make_unique<Request>();
        r->scene = scene;
        r->h_rays = h_rays;
        r->h_out = h_out;
        r->n = n;
        r->any = any_hit ? 1u : 0u;
        r->bytes = bytes;
        r->t_submit = Clock::now();
        std::lock_guard<std::mutex> g(s->mu);
        if (s->fatal != RQ_OK) throw Error(s->fatal, "server stopped after an error: " + s->fatal_msg);
        r->id = s->next_id++;
        *ticket = r->id;
        if (n == 0) {
            r->state = ReqState::Done;
            r->t_admit = r->t_launch = r->t_done = r->t_submit;
        } else {
            s->queue.push_back(r.get());
        }
        s->reqs[r->id] = std::move(r);
        s->cv_work.notify_one();
    });
}

rq_status rq_server_wait(rq_server* s, uint64_t ticket, uint32_t timeout_ms, rq_request_info* info) {
    rq_status st = RQ_OK;
    rq_status r = run([&] {
        require(s != nullptr, "null server");
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->reqs.find(ticket);
        if (it == s->reqs.end()) throw Error(RQ_ERR_INVALID_ARG, "unknown ticket");
        Request* q = it->second.get();
        auto pred = [&] { return q->state == ReqState::Done || q->state == ReqState::Cancelled ||
                                 q->state == ReqState::Failed; };
        if (timeout_ms == 0xFFFFFFFFu)
            s->cv_done.wait(lk, pred);
        else if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), pred))
            throw Error(RQ_ERR_TIMEOUT, "request not finished within the timeout");
        if (info) {
            auto ms = [](Clock::time_point a, Clock::time_point b) {
                return std::chrono::duration<double, std::milli>(b - a).count();
            };
            info->latency_ms = ms(q->t_submit, q->t_done);
            info->queue_ms = q->state == ReqState::Cancelled ? 0.0 : ms(q->t_submit, q->t_launch);
            info->launch_id = q->launch;
            info->rays_in_launch = q->rays_in_launch;
            info->requests_in_launch = q->reqs_in_launch;
        }
        if (q->state == ReqState::Cancelled) st = RQ_ERR_CANCELLED;
        if (q->state == ReqState::Failed) st = q->err;
        s->reqs.erase(it);
        if (st != RQ_OK) throw Error(st, st == RQ_ERR_CANCELLED ? "request was cancelled" : "request failed");
    });
    return r;
}

rq_status rq_server_poll(rq_server* s, uint64_t ticket, int* done) {
    return run([&] {
        require(s && done, "null argument");
        std::lock_guard<std::mutex> g(s->mu);
        auto it = s->reqs.find(ticket);
        if (it == s->reqs.end()) throw Error(RQ_ERR_INVALID_ARG, "unknown ticket");
        const ReqState st = it->second->state;
        *done = (st == ReqState::Done || st == ReqState::Cancelled || st == ReqState::Failed) ? 1 : 0;
    });
}

rq_status rq_server_cancel(rq_server* s, uint64_t ticket) {
    return run([&] {
        require(s != nullptr, "null server");
        std::lock_guard<std::mutex> g(s->mu);
        auto it = s->reqs.find(ticket);
        if (it == s->reqs.end()) throw Error(RQ_ERR_INVALID_ARG, "unknown ticket");
        Request* r = it->second.get();
        if (r->state == ReqState::Queued) {
            s->queue.erase(std::find(s->queue.begin(), s->queue.end(), r));
        } else if (r->state == ReqState::Admitted) {
            s->admitted.erase(std::find(s->admitted.begin(), s->admitted.end(), r));
            s->arena.release(r->off, r->bytes);
        } else {
            throw Error(RQ_ERR_UNSUPPORTED, "request already launched or finished");
        }
        r->state = ReqState::Cancelled;
        r->t_done = Clock::now();
        s->cv_done.notify_all();
        s->cv_work.notify_one();
    });
}

rq_status rq_server_stats(rq_server* s, rq_server_counters* out) {
    return run([&] {
        require(s && out, "null argument");
        std::lock_guard<std::mutex> g(s->mu);
        out->launches = s->launches;
        out->requests_done = s->reqs_done;
        out->rays_done = s->rays_done;
        out->queued = s->queue.size();
        out->admitted = s->admitted.size();
        out->arena_bytes = s->arena.size;
        out->arena_used = s->arena.used;
    });
}

void rq_server_destroy(rq_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> g(s->mu);
        for (Request* r : s->queue) r->state = ReqState::Cancelled;
        for (Request* r : s->admitted) {
            r->state = ReqState::Cancelled;
            s->arena.release(r->off, r->bytes);
        }
        s->queue.clear();
        s->admitted.clear();
        s->stop = true;
        s->cv_work.notify_all();
        s->cv_done.notify_all();
    }
    if (s->worker.joinable()) s->worker.join();
    (void)hipSetDevice(s->dev->ordinal);
    for (Slot& sl : s->slots) {
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
        slot_free(sl);
    }
    delete s;
}

}  // extern "C"
