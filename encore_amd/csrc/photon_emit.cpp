// photon_emit.cpp -- host side of photon emission (PhotonMapperCPU::EmitPhotons,
// PhotonMapperCPU.cpp:503-600): scene materials, batching of photon paths,
// the wavefront loop over the device kernels, and the emission stop rule.
//
// A batch of paths is traced completely on the device; its photons then
// enter the maps in emission order (path, then event within the path), each
// map taking photons until it is full.  Emission stops after the path that
// fills the global map or, when the caustic map has room left then, after
// the path that fills the caustic map.  Paths of the last batch beyond that
// point are discarded, so the result does not depend on the batch size.
#include <algorithm>
#include <cmath>
#include <vector>

#include "rq_internal.h"

namespace rq {

const std::vector<float>& ph_phi_table() {
    static const std::vector<float> table = [] {
        std::vector<float> t(2 * kPhTableSize);
        for (uint32_t k = 0; k < kPhTableSize; ++k) {
            // e2 = randNumber(0, 1); phi = 2 * PI * e2 with PI = 3.14159f
            const float e2 = (1.0f - 0.0f) * ((float)(int)k / (float)10000) + 0.0f;
            const float phi = (2 * 3.14159f) * e2;
            t[2 * k] = (float)std::cos((double)phi);
            t[2 * k + 1] = (float)std::sin((double)phi);
        }
        return t;
    }();
    return table;
}

namespace {

template <class F>
rq_status run(F&& f) {
    return guard_call([](void* c) { (*static_cast<F*>(c))(); }, &f);
}

void require(bool ok, const char* msg) {
    if (!ok) throw Error(RQ_ERR_INVALID_ARG, msg);
}

bool finite_all(const float* v, int n) {
    for (int i = 0; i < n; ++i)
        if (!std::isfinite(v[i])) return false;
    return true;
}

void set_materials(rq_scene* sc, const rq_material* mats, uint32_t nmats, const uint32_t* prim_mat) {
    require(mats && prim_mat && nmats > 0, "null or empty material table");
    std::vector<PhMaterial> pm(nmats);
    for (uint32_t i = 0; i < nmats; ++i) {
        const rq_material& m = mats[i];
        require(finite_all(m.diffuse, 3) && std::isfinite(m.transmittance) && std::isfinite(m.transparency) &&
                    std::isfinite(m.refraction_index),
                "material values must be finite");
        PhMaterial& p = pm[i];
        p = PhMaterial{};
        // GetAvgDiffuseReflectance (Material.h:46); pTotal = pS + pD + pT, pS = 0
        p.pd = (m.diffuse[0] + m.diffuse[1] + m.diffuse[2]) / 3.0f;
        p.pt = m.transmittance;
        p.ptotal = (0.0f + p.pd) + p.pt;
        p.transparency = m.transparency;
        p.index = m.refraction_index;
    }
    for (uint32_t i = 0; i < sc->ntris; ++i) require(prim_mat[i] < nmats, "prim material index out of range");
    RQ_HIP(hipSetDevice(sc->dev->ordinal));
    DevBuf dm(pm.size() * sizeof(PhMaterial));
    dm.upload(pm.data(), dm.bytes);
    DevBuf dp((size_t)sc->ntris * 4);
    dp.upload(prim_mat, dp.bytes);
    sc->materials = std::move(dm);
    sc->prim_mat = std::move(dp);
    sc->nmaterials = nmats;
}

// Workspace of one batch size, carved from the device's emission buffer.
struct PhWork {
    rq_ray *path_ray = nullptr, *rays[2] = {nullptr, nullptr};
    float *path_dir = nullptr, *ev_pos = nullptr, *ev_nrm = nullptr, *phi = nullptr;
    uint32_t *path_light = nullptr, *tags[2] = {nullptr, nullptr}, *boff = nullptr, *ctl = nullptr;
    uint8_t* ev_tag = nullptr;
    unsigned long long* cnt = nullptr;
    rq_hit* hits = nullptr;
    PhLight* lights = nullptr;

    PhWork(uint32_t bmax, uint32_t nslots, uint32_t nlights, DevBuf& buf) {
        const size_t B = bmax, Q = 2 * (size_t)bmax, S = B * nslots;
        const size_t nblocks = (B + kPhScanBlock - 1) / kPhScanBlock;
        size_t off = 0;
        auto take = [&](size_t bytes) {
            const size_t o = off;
            off += (bytes + 255) & ~size_t(255);
            return o;
        };
        const size_t o_pray = take(B * sizeof(rq_ray)), o_r0 = take(Q * sizeof(rq_ray)),
                     o_r1 = take(Q * sizeof(rq_ray)), o_pdir = take(B * 12), o_evp = take(S * 12),
                     o_evn = take(S * 12), o_phi = take(2 * kPhTableSize * 4), o_plight = take(B * 4),
                     o_t0 = take(Q * 4), o_t1 = take(Q * 4), o_boff = take(nblocks * 12), o_ctl = take(64),
                     o_evt = take(S), o_cnt = take(B * 8), o_hits = take(Q * sizeof(rq_hit)),
                     o_lights = take((size_t)nlights * sizeof(PhLight));
        if (buf.bytes < off) {
            buf = DevBuf();  // release the old buffer before taking the larger one
            buf = DevBuf(off);
        }
        char* b = static_cast<char*>(buf.ptr);
        path_ray = reinterpret_cast<rq_ray*>(b + o_pray);
        rays[0] = reinterpret_cast<rq_ray*>(b + o_r0);
        rays[1] = reinterpret_cast<rq_ray*>(b + o_r1);
        path_dir = reinterpret_cast<float*>(b + o_pdir);
        ev_pos = reinterpret_cast<float*>(b + o_evp);
        ev_nrm = reinterpret_cast<float*>(b + o_evn);
        phi = reinterpret_cast<float*>(b + o_phi);
        path_light = reinterpret_cast<uint32_t*>(b + o_plight);
        tags[0] = reinterpret_cast<uint32_t*>(b + o_t0);
        tags[1] = reinterpret_cast<uint32_t*>(b + o_t1);
        boff = reinterpret_cast<uint32_t*>(b + o_boff);
        ctl = reinterpret_cast<uint32_t*>(b + o_ctl);
        ev_tag = reinterpret_cast<uint8_t*>(b + o_evt);
        cnt = reinterpret_cast<unsigned long long*>(b + o_cnt);
        hits = reinterpret_cast<rq_hit*>(b + o_hits);
        lights = reinterpret_cast<PhLight*>(b + o_lights);
    }
};

uint32_t read_u32(const uint32_t* d, hipStream_t s) {
    uint32_t v = 0;
    RQ_HIP(hipMemcpyAsync(&v, d, 4, hipMemcpyDeviceToHost, s));
    RQ_HIP(hipStreamSynchronize(s));
    return v;
}

void emit(rq_scene* sc, const rq_rect_light* lights, uint32_t nlights, const rq_photon_params& P,
          const rq_photon_buffers out[3], rq_photon_result* res, hipStream_t s) {
    require(sc->nmaterials > 0, "scene has no materials (rq_scene_set_materials)");
    require(lights && nlights > 0, "at least one light");
    require(P.max_bounces <= kPhMaxBounces, "max_bounces is at most 16");
    std::vector<PhLight> hl(nlights);
    for (uint32_t i = 0; i < nlights; ++i) {
        const rq_rect_light& l = lights[i];
        require(finite_all(l.center, 3) && finite_all(l.normal, 3) && std::isfinite(l.length) &&
                    std::isfinite(l.width) && finite_all(l.power, 3),
                "light values must be finite");
        const double n2 = (double)l.normal[0] * l.normal[0] + (double)l.normal[1] * l.normal[1] +
                          (double)l.normal[2] * l.normal[2];
        require(n2 > 1e-4, "light normal must not be zero");
        PhLight& h = hl[i];
        h = PhLight{};
        for (int k = 0; k < 3; ++k) {
            h.center[k] = l.center[k];
            h.normal[k] = l.normal[k];
            h.power[k] = l.power[k];
        }
        h.length = l.length;
        h.width = l.width;
    }
    for (int m = 0; m < 3; ++m) {
        require(std::isfinite(P.scale[m]), "map scales must be finite");
        if (P.capacity[m]) require(out && out[m].pos && out[m].normal && out[m].power, "null output buffer");
        if (P.capacity[m] > kMaxCount) throw Error(RQ_ERR_CAPACITY, "at most 2^30 photons per map");
    }
    rq_photon_result r{};
    const uint32_t nslots = P.max_bounces + 2;
    uint32_t rem[3] = {P.capacity[0], P.capacity[1], P.capacity[2]};
    if (rem[RQ_MAP_GLOBAL] == 0 && rem[RQ_MAP_CAUSTIC] == 0) {  // both of EmitPhotons' loops end at once
        *res = r;
        return;
    }
    const uint64_t cap_sum = (uint64_t)P.capacity[0] + P.capacity[1] + P.capacity[2];
    const uint64_t max_paths = P.max_paths ? P.max_paths : std::max<uint64_t>(1ull << 24, 64 * cap_sum);
    const uint32_t bmax = std::max<uint32_t>(1024, std::min<uint32_t>(P.batch_paths ? P.batch_paths : 1u << 20,
                                                                      kPhMaxBatch));
    RQ_HIP(hipSetDevice(sc->dev->ordinal));
    std::lock_guard<std::mutex> hold(sc->dev->emit_mu);
    // Every call leaves the workspace idle: it ends synchronised with its
    // stream, also when it fails part way.
    struct Drain {
        hipStream_t s;
        ~Drain() { (void)hipStreamSynchronize(s); }
    } drain{s};
    PhWork w(bmax, nslots, nlights, sc->dev->emit_work);
    const std::vector<float>& phi = ph_phi_table();
    RQ_HIP(hipMemcpyAsync(w.phi, phi.data(), phi.size() * 4, hipMemcpyHostToDevice, s));
    RQ_HIP(hipMemcpyAsync(w.lights, hl.data(), hl.size() * sizeof(PhLight), hipMemcpyHostToDevice, s));

    PhLaunch L{};
    L.tris = sc->tris.as<TriRec>();
    L.prim_slot = sc->prim_slot.as<uint32_t>();
    L.predicate = sc->predicate;
    L.prim_mat = sc->prim_mat.as<uint32_t>();
    L.mats = sc->materials.as<PhMaterial>();
    L.lights = w.lights;
    L.nlights = nlights;
    L.phi_cs = w.phi;
    L.seed = P.seed;
    L.max_bounces = P.max_bounces;
    L.nslots = nslots;
    L.path_ray = w.path_ray;
    L.path_dir = w.path_dir;
    L.path_light = w.path_light;
    L.ev_pos = w.ev_pos;
    L.ev_nrm = w.ev_nrm;
    L.ev_tag = w.ev_tag;
    L.cnt = w.cnt;
    L.boff = w.boff;
    L.ctl = w.ctl;
    for (int m = 0; m < 3; ++m) {
        L.scale[m] = P.scale[m];
        L.out[m] = PhOut{out ? out[m].pos : nullptr,    out ? out[m].normal : nullptr,
                         out ? out[m].power : nullptr,  out ? out[m].dir : nullptr,
                         out ? out[m].shadow : nullptr, out ? out[m].path : nullptr};
    }
    // Live-ray counts of the two queues.  Iterations are queued in groups
    // without waiting: the traversal and the shading read the count on the
    // device, and an iteration over an empty queue does nothing.
    uint32_t* count[2] = {w.ctl + 8, w.ctl + 9};
    const uint32_t qmax = 2 * bmax;
    const int shade_blocks = 4 * sc->dev->num_cus;
    constexpr int kGroup = 4;
    // Every live ray advances a path by a bounce, a refraction step or a
    // shadow-photon pass, so a batch drains within this many iterations.
    const uint32_t max_iters = 2 * (P.max_bounces + 1) + kPhShadowPasses + 2 * kGroup;
    L.stride = bmax;
    // Photons per path seen so far, to size the next batch.
    double seen_paths = 0, seen[3] = {0, 0, 0};
    uint64_t paths = 0;
    bool done = false;
    while (!done && paths < max_paths) {
        double want = 0;
        for (int m : {RQ_MAP_GLOBAL, RQ_MAP_CAUSTIC}) {
            if (!rem[m]) continue;
            const double y = seen_paths > 0 ? seen[m] / seen_paths : 1.0;
            want = std::max(want, y > 0 ? rem[m] / y : (double)bmax);
        }
        uint64_t np = (uint64_t)std::ceil(want * 1.15) + 256;
        np = std::min<uint64_t>(std::max<uint64_t>(np, 1024), bmax);
        np = std::min<uint64_t>(np, max_paths - paths);
        L.base = paths;
        L.npaths = (uint32_t)np;
        L.nblocks = (uint32_t)((np + kPhScanBlock - 1) / kPhScanBlock);
        // Emission, then the wavefront until no ray is left.
        int cur = 0;
        launch_zero_u32(count[cur], 1, s);
        L.out_rays = w.rays[cur];
        L.out_tags = w.tags[cur];
        L.out_count = count[cur];
        launch_ph_emit(L, s);
        // Bound on the live rays, for launch sizes: a path has at most two
        // rays alive (its own and a shadow-photon ray), and an iteration at
        // most doubles the live count.
        const uint64_t per_batch = std::min<uint64_t>(qmax, 2 * np);
        uint64_t bound = np;
        for (uint32_t iters = 0;;) {
            for (int g = 0; g < kGroup; ++g, ++iters) {
                const uint32_t nb = (uint32_t)std::min<uint64_t>(bound, per_batch);
                trace_closest(sc, w.rays[cur], nb, w.hits, s, count[cur]);
                launch_zero_u32(count[1 - cur], 1, s);
                L.in_rays = w.rays[cur];
                L.in_tags = w.tags[cur];
                L.hits = w.hits;
                L.nin_dev = count[cur];
                L.nin_max = nb;
                L.out_rays = w.rays[1 - cur];
                L.out_tags = w.tags[1 - cur];
                L.out_count = count[1 - cur];
                launch_ph_shade(L, shade_blocks, s);
                cur = 1 - cur;
                bound = std::min<uint64_t>(2 * bound, per_batch);
            }
            bound = read_u32(count[cur], s);
            if (bound == 0) break;
            if (iters >= max_iters) throw Error(RQ_ERR_DEVICE, "photon tracing did not drain");
        }
        // Fill the maps in emission order.
        for (int m = 0; m < 3; ++m) {
            L.rem[m] = rem[m];
            L.acc[m] = r.count[m];
        }
        launch_ph_compact(L, s);
        uint32_t ctl[8];
        RQ_HIP(hipMemcpyAsync(ctl, w.ctl, sizeof(ctl), hipMemcpyDeviceToHost, s));
        RQ_HIP(hipStreamSynchronize(s));
        const uint32_t need_g = rem[RQ_MAP_GLOBAL] == 0 ? 0 : (ctl[0] ? ctl[0] : (uint32_t)np);
        const uint32_t need_c = rem[RQ_MAP_CAUSTIC] == 0 ? 0 : (ctl[1] ? ctl[1] : (uint32_t)np);
        for (int m = 0; m < 3; ++m) {
            r.count[m] += ctl[2 + m];
            rem[m] -= ctl[2 + m];
            seen[m] += ctl[5 + m];
        }
        seen_paths += (double)np;
        if (rem[RQ_MAP_GLOBAL] == 0 && rem[RQ_MAP_CAUSTIC] == 0) {
            paths += std::max(need_g, need_c);
            done = true;
        } else {
            paths += np;
        }
    }
    r.paths = paths;
    *res = r;
}

}  // namespace
}  // namespace rq

extern "C" {

rq_status rq_scene_set_materials(rq_scene* scene, const rq_material* materials, uint32_t nmaterials,
                                 const uint32_t* prim_material) {
    return rq::run([&] {
        if (!scene) throw rq::Error(RQ_ERR_INVALID_ARG, "null scene");
        rq::set_materials(scene, materials, nmaterials, prim_material);
    });
}

rq_status rq_emit_photons(rq_scene* scene, const rq_rect_light* lights, uint32_t nlights,
                          const rq_photon_params* params, const rq_photon_buffers out[3], rq_photon_result* result,
                          void* stream) {
    return rq::run([&] {
        if (!scene || !params || !result) throw rq::Error(RQ_ERR_INVALID_ARG, "null argument");
        rq::emit(scene, lights, nlights, *params, out, result, (hipStream_t)stream);
    });
}

}  // extern "C"
