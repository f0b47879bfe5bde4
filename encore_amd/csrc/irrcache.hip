// irrcache.hip -- the irradiance cache of ENCORE's final gather
// (IrrCache<Color>, IrrCacheOctree.h:92-326; used by
// PhotonMapperCPU::CalculateIndirectIllumination, PhotonMapperCPU.cpp:1087-1103
// and 1170-1173) with its lookup on the GPU.
//
// What the reference computes.  AddElement stores an element in the root's
// list when 2 * (a * Ri) is at most the root's x size; otherwise it descends
// into children of half the size, where the test fails again, and never
// returns.  So in every run that completes the cache is one list in insertion
// order, and a lookup at (p, n) walks all of it: an element counts when
// GetWeight = 1 / (|p - x_i| / R_i + sqrt(1 - n.n_i)) exceeds 1 / a and
// SampleInFront is false, and the caller forms sum(value_i w_i) / sum(w_i) in
// list order.  This file refuses the elements the reference would never
// return from (RQ_ERR_UNSUPPORTED) and drops those outside the root box, as
// AddElement does.
//
// How it runs here.  A weight above 1 / a needs |p - x_i| / R_i < a, so an
// element with 0 < R_i can only count inside the sphere of radius a R_i about
// x_i.  The host bins every element into the cells of a uniform grid that its
// sphere (widened by 2^-10 for the float arithmetic of the test, which
// admits at most 2^-19 beyond it) touches, in insertion order; elements whose
// reach is unbounded (R_i = 0, negative or infinite, where the distance term
// vanishes or turns negative) or that would touch more than 65,536 cells go
// to a global list.  One lane per lookup merges its cell's list with the
// global list by index and evaluates the reference's weight and front test
// with its float/double types and operation order, so the sums are the
// reference's bit for bit; an element whose squared distance alone bounds
// its weight to at most 1 / a (a per-element bound the host computes with
// upward rounding) is passed over after four products.  Lookups outside the
// grid see the global list only.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <limits>
#include <vector>

#include "rq_internal.h"

namespace rq {

struct IcElem {
    float4 pos_ri;  // x, y, z, R_i
    float4 nrm;     // n_i, 0
    float4 val;     // value (R, G, B), 0
};

struct IcLaunch {
    const float4* pos_ri;
    const float4* nrm;
    const float4* val;
    const uint32_t* cell_start;  // ncells + 1
    const uint32_t* cell_items;
    const uint32_t* glob;        // elements every lookup checks, ascending
    uint32_t nglob;
    uint32_t dims[3];
    double origin[3];
    double inv_h;
    float tol;                   // 1 / a as IrrCacheOctree.h:300 forms it
    const float* qpos;
    const float* qnrm;
    uint32_t nq;
    float* rgb;
    float* wsum;
    uint32_t* count;
    unsigned long long* tally;   // counting variant: [lookups, candidates examined, values counted]
};

namespace dev {

__device__ __forceinline__ float flush_ic(float x) { return (x > -0.001f && x < 0.001f) ? 0.0f : x; }

template <bool COUNT>
__global__ __launch_bounds__(256) void ic_query(IcLaunch L) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= L.nq) return;
    const float px = L.qpos[3 * (size_t)q], py = L.qpos[3 * (size_t)q + 1], pz = L.qpos[3 * (size_t)q + 2];
    const float nx = L.qnrm[3 * (size_t)q], ny = L.qnrm[3 * (size_t)q + 1], nz = L.qnrm[3 * (size_t)q + 2];
    // The lookup's cell (same double formula as the host's binning).
    uint32_t cb = 0, ce = 0;
    const double c0 = floor(((double)px - L.origin[0]) * L.inv_h);
    const double c1 = floor(((double)py - L.origin[1]) * L.inv_h);
    const double c2 = floor(((double)pz - L.origin[2]) * L.inv_h);
    if (c0 >= 0.0 && c1 >= 0.0 && c2 >= 0.0 && c0 < (double)L.dims[0] && c1 < (double)L.dims[1] &&
        c2 < (double)L.dims[2]) {
        const size_t cell = ((size_t)c2 * L.dims[1] + (size_t)c1) * L.dims[0] + (size_t)c0;
        cb = L.cell_start[cell];
        ce = L.cell_start[cell + 1];
    }
    float ws = 0.0f, ar = 0.0f, ag = 0.0f, ab = 0.0f;
    uint32_t cnt = 0, j = 0, examined = 0;
    const uint32_t kEnd = 0xFFFFFFFFu;
    for (;;) {
        const uint32_t a = cb < ce ? L.cell_items[cb] : kEnd;
        const uint32_t g = j < L.nglob ? L.glob[j] : kEnd;
        if (a == kEnd && g == kEnd) break;
        uint32_t e;
        if (a < g) {
            e = a;
            ++cb;
        } else {
            e = g;
            ++j;
        }
        if (COUNT) ++examined;
        const float4 pr = L.pos_ri[e];
        const float4 nn = L.nrm[e];
        // TCacheElement::GetWeight (IrrCacheOctree.h:55-77).
        const float vx = pr.x - px, vy = pr.y - py, vz = pr.z - pz;
        const float d2 = (vx * vx + vy * vy) + vz * vz;
        // nn.w: a squared distance at and beyond which the distance term
        // alone keeps the weight at or below 1 / a (set by the host, see
        // reject_d2): the element cannot count, so the rest is skipped.
        if (d2 >= nn.w) continue;
        const float distance = __builtin_sqrtf(d2);
        const float distCalc = (pr.w == 0.0f) ? 0.0f : distance / pr.w;
        const bool same = nx == nn.x && ny == nn.y && nz == nn.z;
        const float d = (nx * nn.x + ny * nn.y) + nz * nn.z;
        const float sqrtCalc = same ? 0.0f : (float)__builtin_sqrt(1.0 - (double)d);
        const float denom = sqrtCalc + distCalc;
        const float w = denom == 0.0f ? 0.0f : 1.0f / denom;
        if (!(w > L.tol)) continue;
        // SampleInFront (IrrCacheOctree.h:279-290): the lookup lies in front
        // unless dot(flush(u / |u|), g) < -0.1, u = p - x_i, g = n_i + n.
        // The flush moves each component by less than 0.001 and the
        // rounding of |u|, the divisions and the dot product moves the
        // result by less than 3e-6 more, so with |g|_1 <= 6 the test passes
        // whenever dot(u, g) >= -0.0934 |u|.  The unnormalised dot product
        // r decides that without the square root and the divisions: r >= 0,
        // or r^2 <= 0.008 d2 (d2 = |u|^2 as computed above, the same float),
        // which bounds |dot(u, g)| by 0.0895 |u| with the rounding of r, r^2
        // and 0.008 d2 included.  Otherwise the reference's own operations
        // decide.
        const float gx = nn.x + nx, gy = nn.y + ny, gz = nn.z + nz;
        float ux = px - pr.x, uy = py - pr.y, uz = pz - pr.z;
        const float r = (ux * gx + uy * gy) + uz * gz;
        const float g1 = (__builtin_fabsf(gx) + __builtin_fabsf(gy)) + __builtin_fabsf(gz);
        if (!(g1 <= 6.0f && (r >= 0.0f || r * r <= 0.008f * d2))) {
            const float m = __builtin_sqrtf((ux * ux + uy * uy) + uz * uz);
            ux = flush_ic(ux / m);
            uy = flush_ic(uy / m);
            uz = flush_ic(uz / m);
            if ((ux * gx + uy * gy) + uz * gz < -0.1f) continue;
        }
        const float4 v = L.val[e];
        ws += w;
        ar += v.x * w;
        ag += v.y * w;
        ab += v.z * w;
        ++cnt;
    }
    if (ws != 0.0f) {
        ar = ar / ws;
        ag = ag / ws;
        ab = ab / ws;
    } else {
        ar = ag = ab = 0.0f;
    }
    L.rgb[3 * (size_t)q] = ar;
    L.rgb[3 * (size_t)q + 1] = ag;
    L.rgb[3 * (size_t)q + 2] = ab;
    if (L.wsum) L.wsum[q] = ws;
    if (L.count) L.count[q] = cnt;
    if (COUNT) {
        atomicAdd(&L.tally[0], 1ull);
        atomicAdd(&L.tally[1], (unsigned long long)examined);
        atomicAdd(&L.tally[2], (unsigned long long)cnt);
    }
}

}  // namespace dev
}  // namespace rq

struct rq_irrcache {
    rq_device* dev = nullptr;
    float a = 0, minp[3] = {0, 0, 0}, size[3] = {0, 0, 0};
    std::vector<rq::IcElem> elems;  // insertion order (the reference's root list)
    bool dirty = true;
    // Device copy and index, rebuilt on the first lookup after an add.
    rq::DevBuf d_pos, d_nrm, d_val, d_start, d_items, d_glob;
    uint32_t nglob = 0, dims[3] = {0, 0, 0};
    double origin[3] = {0, 0, 0}, inv_h = 0;
    uint64_t index_cells = 0, index_refs = 0;
};

namespace rq {
namespace {

constexpr uint32_t kIcMaxDim = 128;          // cells per axis
constexpr uint64_t kIcMaxCellsPerElem = 65536;
// Cell edge: the median reach diameter over this.  A lookup examines the
// elements whose reach touches its cell; at the final gather's cache (16,364
// values) edges of 1, 1/2, 1/3, 1/4 and 1/8 diameters examine 117.5, 82.9,
// 73.6, 69.4 and 64.0 values per lookup (19.8 count) and list 0.16, 0.56,
// 1.44, 2.92 and 15.8 M references; 1/3 keeps the host's index build near
// 3.5 ms per rebuild.
constexpr double kIcCellsPerDiameter = 3.0;

template <class F>
rq_status run(F&& f) {
    return guard_call([](void* c) { (*static_cast<F*>(c))(); }, &f);
}

void require(bool ok, const char* msg) {
    if (!ok) throw Error(RQ_ERR_INVALID_ARG, msg);
}

// The least positive float T with fl(1 / T) <= tol (+inf when there is none):
// a weight 1 / denom with denom >= T is at most tol, since float division
// is monotonic.
float reject_dist(float tol) {
    if (!(tol > 0.0f) || !std::isfinite(tol)) return INFINITY;
    float t = (float)(1.0 / (double)tol);
    if (!(t > 0.0f)) t = std::numeric_limits<float>::denorm_min();
    while (!(1.0f / t <= tol)) {
        t = std::nextafter(t, INFINITY);
        if (std::isinf(t)) return INFINITY;
    }
    for (;;) {
        const float p = std::nextafter(t, 0.0f);
        if (p > 0.0f && 1.0f / p <= tol) {
            t = p;
        } else {
            break;
        }
    }
    return t;
}

// Squared-distance bound of one element (kept in its normal's w on the
// device): Z = Y^2 and Y = T R_i, each rounded up to a float.  A lookup with
// float d2 >= Z has sqrt(d2) >= Y >= T R_i, so its distance rounds to at
// least Y, distance / R_i to at least T, the denominator (the distance term
// plus a non-negative or NaN normal term) to at least T or NaN, and the
// weight to at most fl(1 / T) <= 1 / a, or NaN: GetWeight's element does
// not count, as in the reference.  +inf where R_i is not a positive finite
// float (the distance term is then 0 or negative: no bound).
float reject_d2(float T, float ri) {
    if (!(ri > 0.0f) || !std::isfinite(ri) || !std::isfinite(T)) return INFINITY;
    auto up = [](double x) {
        float f = (float)x;
        if ((double)f < x) f = std::nextafter(f, INFINITY);
        return f;
    };
    const float y = up((double)T * (double)ri);  // the double product is exact
    if (!std::isfinite(y)) return INFINITY;
    return up((double)y * (double)y);
}

// An element's reach (the sphere a lookup must enter for its weight to pass
// 1 / a, widened), or a negative value when it is not bounded.
double reach_of(const rq_irrcache* c, const IcElem& e) {
    const float ri = e.pos_ri.w;
    if (!(ri > 0.0f) || !std::isfinite(ri)) return -1.0;
    return (double)c->a * (double)ri * (1.0 + 1.0 / 1024.0);
}

void build_index(rq_irrcache* c) {
    const uint32_t n = (uint32_t)c->elems.size();
    std::vector<uint32_t> glob;
    std::vector<double> reach(n);
    std::vector<double> diam;
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = 0; i < n; ++i) {
        const IcElem& e = c->elems[i];
        reach[i] = reach_of(c, e);
        if (std::isnan(e.pos_ri.w)) continue;  // a NaN R_i gives a NaN weight: it never counts
        if (reach[i] < 0.0) continue;
        const double p[3] = {e.pos_ri.x, e.pos_ri.y, e.pos_ri.z};
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], p[k] - reach[i]);
            hi[k] = std::max(hi[k], p[k] + reach[i]);
        }
        diam.push_back(2.0 * reach[i]);
    }
    // Cell edge: a third of the median diameter, and at least 1/(kIcMaxDim -
    // 1) of the longest extent, so that floor(extent / h) + 1 <= kIcMaxDim
    // cells cover every reach: no bounded element ends beyond its axis's last
    // cell.
    double h = 1.0;
    if (!diam.empty()) {
        std::nth_element(diam.begin(), diam.begin() + diam.size() / 2, diam.end());
        const double ext = std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]});
        h = std::max(diam[diam.size() / 2] / kIcCellsPerDiameter, ext / (kIcMaxDim - 1));
        if (!(h > 0.0) || !std::isfinite(h)) h = 1.0;
    } else {
        lo[0] = lo[1] = lo[2] = 0.0;
        hi[0] = hi[1] = hi[2] = 0.0;
    }
    c->inv_h = 1.0 / h;
    for (int k = 0; k < 3; ++k) {
        c->origin[k] = lo[k];
        c->dims[k] = (uint32_t)std::floor((hi[k] - lo[k]) * c->inv_h) + 1;
        if (c->dims[k] > kIcMaxDim) throw Error(RQ_ERR_DEVICE, "irradiance cache grid sizing");
    }
    const uint64_t ncells = (uint64_t)c->dims[0] * c->dims[1] * c->dims[2];
    auto cell_range = [&](const IcElem& e, double r, int64_t b[3], int64_t t[3]) {
        const double p[3] = {e.pos_ri.x, e.pos_ri.y, e.pos_ri.z};
        uint64_t cells = 1;
        for (int k = 0; k < 3; ++k) {
            b[k] = (int64_t)std::floor((p[k] - r - c->origin[k]) * c->inv_h);
            t[k] = (int64_t)std::floor((p[k] + r - c->origin[k]) * c->inv_h);
            b[k] = std::max<int64_t>(0, b[k]);
            t[k] = std::min<int64_t>((int64_t)c->dims[k] - 1, t[k]);
            cells *= (uint64_t)std::max<int64_t>(0, t[k] - b[k] + 1);
        }
        return cells;
    };
    std::vector<uint32_t> start(ncells + 1, 0);
    std::vector<uint8_t> in_grid(n, 0);
    for (uint32_t i = 0; i < n; ++i) {
        const IcElem& e = c->elems[i];
        if (std::isnan(e.pos_ri.w)) continue;
        if (reach[i] < 0.0) {
            glob.push_back(i);
            continue;
        }
        int64_t b[3], t[3];
        const uint64_t cells = cell_range(e, reach[i], b, t);
        if (cells > kIcMaxCellsPerElem) {
            glob.push_back(i);
            continue;
        }
        in_grid[i] = 1;
        for (int64_t z = b[2]; z <= t[2]; ++z)
            for (int64_t y = b[1]; y <= t[1]; ++y)
                for (int64_t x = b[0]; x <= t[0]; ++x) ++start[((size_t)z * c->dims[1] + (size_t)y) * c->dims[0] + x + 1];
    }
    for (uint64_t i = 0; i < ncells; ++i) start[i + 1] += start[i];
    if (start[ncells] > kMaxCount) throw Error(RQ_ERR_CAPACITY, "irradiance cache index too large");
    std::vector<uint32_t> items(std::max<uint32_t>(1, start[ncells]));
    std::vector<uint32_t> fill(start.begin(), start.end() - 1);
    for (uint32_t i = 0; i < n; ++i) {
        if (!in_grid[i]) continue;
        int64_t b[3], t[3];
        cell_range(c->elems[i], reach[i], b, t);
        for (int64_t z = b[2]; z <= t[2]; ++z)
            for (int64_t y = b[1]; y <= t[1]; ++y)
                for (int64_t x = b[0]; x <= t[0]; ++x)
                    items[fill[((size_t)z * c->dims[1] + (size_t)y) * c->dims[0] + x]++] = i;
    }
    std::vector<float4> pos(std::max<uint32_t>(1, n)), nrm(std::max<uint32_t>(1, n)), val(std::max<uint32_t>(1, n));
    const float T = reject_dist(1 / c->a);  // 1 / a as the lookup forms it
    for (uint32_t i = 0; i < n; ++i) {
        pos[i] = c->elems[i].pos_ri;
        nrm[i] = c->elems[i].nrm;
        nrm[i].w = reject_d2(T, c->elems[i].pos_ri.w);
        val[i] = c->elems[i].val;
    }
    RQ_HIP(hipSetDevice(c->dev->ordinal));
    // Lookups queued on any stream may still read the old index.
    if (c->d_pos.ptr) RQ_HIP(hipDeviceSynchronize());
    DevBuf dp(pos.size() * sizeof(float4)), dn(nrm.size() * sizeof(float4)), dv(val.size() * sizeof(float4));
    DevBuf ds(start.size() * 4), di(items.size() * 4), dg(std::max<size_t>(1, glob.size()) * 4);
    dp.upload(pos.data(), dp.bytes);
    dn.upload(nrm.data(), dn.bytes);
    dv.upload(val.data(), dv.bytes);
    ds.upload(start.data(), ds.bytes);
    di.upload(items.data(), di.bytes);
    if (!glob.empty()) dg.upload(glob.data(), glob.size() * 4);
    c->d_pos = std::move(dp);
    c->d_nrm = std::move(dn);
    c->d_val = std::move(dv);
    c->d_start = std::move(ds);
    c->d_items = std::move(di);
    c->d_glob = std::move(dg);
    c->nglob = (uint32_t)glob.size();
    c->index_cells = ncells;
    c->index_refs = start[ncells];
    c->dirty = false;
}

}  // namespace
}  // namespace rq

using namespace rq;

extern "C" {

rq_status rq_irrcache_create(rq_device* dev, float a, const float* min_point, const float* size, rq_irrcache** out) {
    return run([&] {
        require(dev && min_point && size && out, "null argument");
        require(a > 0.0f && std::isfinite(a), "a must be positive and finite");
        auto c = std::make_unique<rq_irrcache>();
        c->dev = dev;
        c->a = a;
        for (int k = 0; k < 3; ++k) {
            c->minp[k] = min_point[k];
            c->size[k] = size[k];
        }
        *out = c.release();
    });
}

rq_status rq_irrcache_add(rq_irrcache* c, const float* value, const float* pos, const float* normal, const float* ri,
                          uint32_t n, uint32_t* status) {
    return run([&] {
        require(c != nullptr, "null cache");
        require(n == 0 || (value && pos && normal && ri), "null element array");
        if ((uint64_t)c->elems.size() + n > kMaxCount) throw Error(RQ_ERR_CAPACITY, "at most 2^30 cached values");
        // Check every element first, so a refused one leaves the cache unchanged.
        std::vector<uint32_t> st(n);
        float mx[3];
        for (int k = 0; k < 3; ++k) mx[k] = c->minp[k] + c->size[k];
        bool refused = false;
        for (uint32_t i = 0; i < n; ++i) {
            const float* p = pos + 3 * (size_t)i;
            bool outside = false;
            for (int k = 0; k < 3; ++k) outside = outside || p[k] < c->minp[k] || mx[k] < p[k];
            const float domain = c->a * ri[i];
            st[i] = outside ? 1u : (!(2 * domain <= c->size[0]) ? 2u : 0u);
            refused = refused || st[i] == 2u;
        }
        if (status) std::copy(st.begin(), st.end(), status);
        if (refused)
            throw Error(RQ_ERR_UNSUPPORTED,
                        "a cached value with 2 a Ri above the cache's x size: IrrCache::AddElement would not return");
        for (uint32_t i = 0; i < n; ++i) {
            if (st[i] != 0) continue;  // outside the root box: dropped, as AddElement does
            IcElem e;
            e.pos_ri = make_float4(pos[3 * (size_t)i], pos[3 * (size_t)i + 1], pos[3 * (size_t)i + 2], ri[i]);
            e.nrm = make_float4(normal[3 * (size_t)i], normal[3 * (size_t)i + 1], normal[3 * (size_t)i + 2], 0.0f);
            e.val = make_float4(value[3 * (size_t)i], value[3 * (size_t)i + 1], value[3 * (size_t)i + 2], 0.0f);
            c->elems.push_back(e);
        }
        c->dirty = true;
    });
}

}  // extern "C"

namespace {

void ic_launch(rq_irrcache* c, const float* d_qpos, const float* d_qnrm, uint32_t nq, float* d_rgb, float* d_wsum,
               uint32_t* d_count, unsigned long long* tally, hipStream_t stream) {
        if (nq > kMaxCount) throw Error(RQ_ERR_CAPACITY, "at most 2^30 lookups per call");
        if (nq == 0) return;
        RQ_HIP(hipSetDevice(c->dev->ordinal));
        if (c->dirty) build_index(c);
        IcLaunch L{};
        L.pos_ri = c->d_pos.as<float4>();
        L.nrm = c->d_nrm.as<float4>();
        L.val = c->d_val.as<float4>();
        L.cell_start = c->d_start.as<uint32_t>();
        L.cell_items = c->d_items.as<uint32_t>();
        L.glob = c->d_glob.as<uint32_t>();
        L.nglob = c->nglob;
        for (int k = 0; k < 3; ++k) {
            L.dims[k] = c->dims[k];
            L.origin[k] = c->origin[k];
        }
        L.inv_h = c->inv_h;
        L.tol = 1 / c->a;
        L.qpos = d_qpos;
        L.qnrm = d_qnrm;
        L.nq = nq;
        L.rgb = d_rgb;
        L.wsum = d_wsum;
        L.count = d_count;
        L.tally = tally;
        if (tally)
            hipLaunchKernelGGL(dev::ic_query<true>, dim3((nq + 255) / 256), dim3(256), 0, stream, L);
        else
            hipLaunchKernelGGL(dev::ic_query<false>, dim3((nq + 255) / 256), dim3(256), 0, stream, L);
        RQ_HIP(hipGetLastError());
}

}  // namespace

extern "C" {

rq_status rq_irrcache_query(rq_irrcache* c, const float* d_qpos, const float* d_qnrm, uint32_t nq, float* d_rgb,
                            float* d_wsum, uint32_t* d_count, void* stream) {
    return run([&] {
        require(c != nullptr, "null cache");
        require(nq == 0 || (d_qpos && d_qnrm && d_rgb), "null query or output buffer");
        ic_launch(c, d_qpos, d_qnrm, nq, d_rgb, d_wsum, d_count, nullptr, (hipStream_t)stream);
    });
}

rq_status rq_irrcache_counts(rq_irrcache* c, const float* d_qpos, const float* d_qnrm, uint32_t nq, float* d_rgb,
                             uint64_t* counts, void* stream) {
    return run([&] {
        require(c != nullptr && counts != nullptr, "null argument");
        require(nq == 0 || (d_qpos && d_qnrm && d_rgb), "null query or output buffer");
        counts[0] = counts[1] = counts[2] = 0;
        if (nq == 0) return;
        RQ_HIP(hipSetDevice(c->dev->ordinal));
        DevBuf t(3 * sizeof(unsigned long long));
        launch_zero_u32(t.as<uint32_t>(), 6, (hipStream_t)stream);
        ic_launch(c, d_qpos, d_qnrm, nq, d_rgb, nullptr, nullptr, t.as<unsigned long long>(), (hipStream_t)stream);
        RQ_HIP(hipMemcpyAsync(counts, t.ptr, 3 * sizeof(uint64_t), hipMemcpyDeviceToHost, (hipStream_t)stream));
        RQ_HIP(hipStreamSynchronize((hipStream_t)stream));
    });
}

rq_status rq_irrcache_info(const rq_irrcache* c, uint32_t* count, uint32_t* global_count, uint64_t* index_cells,
                           uint64_t* index_refs) {
    return run([&] {
        require(c != nullptr, "null cache");
        if (count) *count = (uint32_t)c->elems.size();
        if (global_count) *global_count = c->nglob;
        if (index_cells) *index_cells = c->index_cells;
        if (index_refs) *index_refs = c->index_refs;
    });
}

void rq_irrcache_destroy(rq_irrcache* c) {
    if (!c) return;
    (void)hipSetDevice(c->dev->ordinal);
    (void)hipDeviceSynchronize();
    delete c;
}

}  // extern "C"
