// traverse.hip -- gfx950 ray traversal kernels (BVH2 and ENCORE kd-tree).
//
// Execution model (wave64, CDNA4):
//  - Persistent waves.  A launch has about (CUs x resident blocks) workgroups;
//    each wave pulls rays from a global counter.  Lanes whose ray finished are
//    refilled together: a ballot finds them, one atomicAdd reserves a run of
//    ray indices, mbcnt gives each idle lane its slot (Aila & Laine 2009,
//    "Understanding the efficiency of ray traversal on GPUs", persistent
//    while-while with dynamic fetch).
//  - Short per-lane traversal stack in LDS, laid out [entry][lane] so the 64
//    lanes of a wave touch 64 consecutive dwords (no bank conflicts).  It is a
//    ring: when full, the oldest entry spills to a per-lane region in HBM and
//    comes back on pop, so any tree depth is handled.
//  - BVH: leaves found during node traversal are postponed until every active
//    lane has one, then all lanes test triangles together.
//  - Box tests use the near/far plane picked by the direction's sign bit, so a
//    0*inf NaN (ray in a slab plane) is dropped by max/min and the slab counts
//    as entered; the far distance is widened by 2^-21 relative to cover the
//    rounding of (b - o) * (1/d).
//  - All FP32 arithmetic is compiled without contraction (-ffp-contract=off)
//    and with IEEE division, because the ENCORE predicate and the kd-tree
//    traversal must reproduce the reference bit for bit.
#include <hip/hip_runtime.h>

#include "rq_internal.h"

namespace rq {
namespace dev {

constexpr int kWave = 64;
constexpr uint32_t kNone = 0xFFFFFFFFu;

__device__ __forceinline__ float4 ldg4(const void* p) { return *reinterpret_cast<const float4*>(p); }
__device__ __forceinline__ uint4 ldgu4(const void* p) { return *reinterpret_cast<const uint4*>(p); }

__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// ---------------------------------------------------------------------------
// Per-lane ring stack: S entries of W dwords in LDS ([entry][word][lane]),
// overflow to HBM ([depth][word][lane] per wave).
// ---------------------------------------------------------------------------
template <int S, int W>
struct RingStack {
    uint32_t* lds;      // &lds_base[0][0][lane]
    uint32_t* spill;    // &spill_base[wave][0][0][lane]
    int sp = 0, lo = 0;

    __device__ __forceinline__ void push(const uint32_t (&v)[W]) {
        if (sp - lo == S) {
            const int s = lo & (S - 1);
#pragma unroll
            for (int w = 0; w < W; ++w) spill[((size_t)lo * W + w) * kWave] = lds[(s * W + w) * kWave];
            ++lo;
        }
        const int s = sp & (S - 1);
#pragma unroll
        for (int w = 0; w < W; ++w) lds[(s * W + w) * kWave] = v[w];
        ++sp;
    }
    // Returns false when empty.
    __device__ __forceinline__ bool pop(uint32_t (&v)[W]) {
        if (sp == 0) return false;
        --sp;
        if (sp < lo) {
            lo = sp;
#pragma unroll
            for (int w = 0; w < W; ++w) v[w] = spill[((size_t)sp * W + w) * kWave];
        } else {
            const int s = sp & (S - 1);
#pragma unroll
            for (int w = 0; w < W; ++w) v[w] = lds[(s * W + w) * kWave];
        }
        return true;
    }
    __device__ __forceinline__ void reset() { sp = lo = 0; }
};

// ---------------------------------------------------------------------------
// Triangle predicates
// ---------------------------------------------------------------------------
struct RayState {
    float ox, oy, oz, dx, dy, dz;
    float ix, iy, iz;  // 1/d
    // watertight shear
    int kx, ky, kz;
    float sx, sy, sz;
};

__device__ __forceinline__ float flush_small(float x) { return (x > -0.001f && x < 0.001f) ? 0.0f : x; }

// ENCORE's Moller-Trumbore variant (TrianglePrim.cpp:204-249): edges given,
// Cross() with the component flush, no determinant test, inclusive edges.
struct PredEncore {
    __device__ __forceinline__ static void prepare(RayState&) {}
    __device__ __forceinline__ static bool test(const RayState& r, const float4 a, const float4 b, const float4 c,
                                                float tmin, float& t, float& u, float& v) {
        const float e1x = b.x, e1y = b.y, e1z = b.z;
        const float e2x = c.x, e2y = c.y, e2z = c.z;
        const float px = flush_small(r.dy * e2z - r.dz * e2y);
        const float py = flush_small(-(r.dx * e2z - r.dz * e2x));
        const float pz = flush_small(r.dx * e2y - r.dy * e2x);
        const float det = (e1x * px + e1y * py) + e1z * pz;
        const float inv = 1.0f / det;
        const float sx = r.ox - a.x, sy = r.oy - a.y, sz = r.oz - a.z;
        const float qx = flush_small(sy * e1z - sz * e1y);
        const float qy = flush_small(-(sx * e1z - sz * e1x));
        const float qz = flush_small(sx * e1y - sy * e1x);
        const float al = ((sx * px + sy * py) + sz * pz) * inv;
        const float be = ((r.dx * qx + r.dy * qy) + r.dz * qz) * inv;
        const float tt = ((e2x * qx + e2y * qy) + e2z * qz) * inv;
        const bool hit = (al >= 0.0f) & (be >= 0.0f) & (al + be <= 1.0f) & (tt > tmin);
        t = tt;
        u = al;
        v = be;
        return hit;
    }
};

__device__ __forceinline__ float comp(float x, float y, float z, int k) { return k == 0 ? x : (k == 1 ? y : z); }

// Watertight ray/triangle test (Woop, Benthin, Wald, JCGT 2(1) 2013).
struct PredWatertight {
    __device__ __forceinline__ static void prepare(RayState& r) {
        const float ax = fabsf(r.dx), ay = fabsf(r.dy), az = fabsf(r.dz);
        int kz = (ax > ay) ? ((ax > az) ? 0 : 2) : ((ay > az) ? 1 : 2);
        int kx = kz == 2 ? 0 : kz + 1;
        int ky = kx == 2 ? 0 : kx + 1;
        if (comp(r.dx, r.dy, r.dz, kz) < 0.0f) {
            const int t = kx;
            kx = ky;
            ky = t;
        }
        r.kx = kx;
        r.ky = ky;
        r.kz = kz;
        const float dz = comp(r.dx, r.dy, r.dz, kz);
        r.sx = comp(r.dx, r.dy, r.dz, kx) / dz;
        r.sy = comp(r.dx, r.dy, r.dz, ky) / dz;
        r.sz = 1.0f / dz;
    }
    __device__ __forceinline__ static bool test(const RayState& r, const float4 a, const float4 b, const float4 c,
                                                float tmin, float& t, float& u, float& v) {
        const float Ax0 = a.x - r.ox, Ay0 = a.y - r.oy, Az0 = a.z - r.oz;
        const float Bx0 = b.x - r.ox, By0 = b.y - r.oy, Bz0 = b.z - r.oz;
        const float Cx0 = c.x - r.ox, Cy0 = c.y - r.oy, Cz0 = c.z - r.oz;
        const float Akz = comp(Ax0, Ay0, Az0, r.kz), Bkz = comp(Bx0, By0, Bz0, r.kz), Ckz = comp(Cx0, Cy0, Cz0, r.kz);
        const float Ax = comp(Ax0, Ay0, Az0, r.kx) - r.sx * Akz;
        const float Ay = comp(Ax0, Ay0, Az0, r.ky) - r.sy * Akz;
        const float Bx = comp(Bx0, By0, Bz0, r.kx) - r.sx * Bkz;
        const float By = comp(Bx0, By0, Bz0, r.ky) - r.sy * Bkz;
        const float Cx = comp(Cx0, Cy0, Cz0, r.kx) - r.sx * Ckz;
        const float Cy = comp(Cx0, Cy0, Cz0, r.ky) - r.sy * Ckz;
        float U = Cx * By - Cy * Bx;
        float V = Ax * Cy - Ay * Cx;
        float W = Bx * Ay - By * Ax;
        if (U == 0.0f || V == 0.0f || W == 0.0f) {
            U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
            V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
            W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
        }
        const bool anyneg = (U < 0.0f) | (V < 0.0f) | (W < 0.0f);
        const bool anypos = (U > 0.0f) | (V > 0.0f) | (W > 0.0f);
        const float det = U + V + W;
        const float Az = r.sz * Akz, Bz = r.sz * Bkz, Cz = r.sz * Ckz;
        const float T = U * Az + V * Bz + W * Cz;
        const float inv = 1.0f / det;
        t = T * inv;
        u = V * inv;
        v = W * inv;
        return !(anyneg & anypos) & (det != 0.0f) & (t > tmin);
    }
};

__device__ __forceinline__ void init_ray(RayState& r, const rq_ray* rays, uint32_t id, float& tmin, float& tmax) {
    const float4 p = ldg4(&rays[id].ox);
    const float4 q = ldg4(&rays[id].dx);
    r.ox = p.x;
    r.oy = p.y;
    r.oz = p.z;
    tmin = p.w;
    r.dx = q.x;
    r.dy = q.y;
    r.dz = q.z;
    tmax = q.w;
    r.ix = 1.0f / r.dx;
    r.iy = 1.0f / r.dy;
    r.iz = 1.0f / r.dz;
}

// Work distribution: refill idle lanes (id == kNone) from the global counter.
// Lanes get kDone when the counter is exhausted.  Must be called by the
// whole wave.
constexpr uint32_t kDone = 0xFFFFFFFEu;

__device__ __forceinline__ uint32_t fetch_work(uint32_t cur, uint32_t* counter, uint32_t n) {
    const bool idle = (cur == kNone);
    const uint64_t m = __ballot(idle);
    if (m == 0) return cur;
    const uint32_t cnt = (uint32_t)__popcll(m);
    const int leader = __ffsll((unsigned long long)m) - 1;
    uint32_t base = 0;
    if (lane_id() == leader) base = atomicAdd(counter, cnt);
    base = __shfl(base, leader);
    if (!idle) return cur;
    const uint32_t id = base + lanes_below(m);
    return id < n ? id : kDone;
}

// ---------------------------------------------------------------------------
// BVH2 traversal
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool box2(const RayState& r, const float4 nx, const float4 ny, const float4 nz, float tmin,
                                     float tbest, float& tn0, float& tn1, bool& h1) {
    const bool sx = __builtin_signbit(r.dx), sy = __builtin_signbit(r.dy), sz = __builtin_signbit(r.dz);
    // child 0
    const float n0x = ((sx ? nx.y : nx.x) - r.ox) * r.ix;
    const float f0x = ((sx ? nx.x : nx.y) - r.ox) * r.ix;
    const float n0y = ((sy ? ny.y : ny.x) - r.oy) * r.iy;
    const float f0y = ((sy ? ny.x : ny.y) - r.oy) * r.iy;
    const float n0z = ((sz ? nz.y : nz.x) - r.oz) * r.iz;
    const float f0z = ((sz ? nz.x : nz.y) - r.oz) * r.iz;
    // child 1
    const float n1x = ((sx ? nx.w : nx.z) - r.ox) * r.ix;
    const float f1x = ((sx ? nx.z : nx.w) - r.ox) * r.ix;
    const float n1y = ((sy ? ny.w : ny.z) - r.oy) * r.iy;
    const float f1y = ((sy ? ny.z : ny.w) - r.oy) * r.iy;
    const float n1z = ((sz ? nz.w : nz.z) - r.oz) * r.iz;
    const float f1z = ((sz ? nz.z : nz.w) - r.oz) * r.iz;
    tn0 = fmaxf(fmaxf(n0x, n0y), fmaxf(n0z, tmin));
    const float tf0 = fminf(fminf(f0x, f0y), fminf(f0z, tbest)) * 1.00000048f;
    tn1 = fmaxf(fmaxf(n1x, n1y), fmaxf(n1z, tmin));
    const float tf1 = fminf(fminf(f1x, f1y), fminf(f1z, tbest)) * 1.00000048f;
    h1 = tn1 <= tf1;
    return tn0 <= tf0;
}

// Item of a batched launch whose descriptor holds global ray index `id`:
// the last item with begin <= id (items sorted by begin).
__device__ __forceinline__ uint32_t find_item(const BatchItemDev* items, uint32_t nitems, uint32_t id) {
    uint32_t lo = 0, hi = nitems;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (items[mid].begin <= id)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

template <class Pred, bool ANY, int S, bool COUNT = false, bool BATCHED = false>
__global__ __launch_bounds__(256) void bvh_trace(TraceLaunch L) {
    __shared__ uint32_t lds_stack[256 / kWave][S][kWave];
    const int lane = lane_id();
    const int wave = threadIdx.x / kWave;
    const uint32_t gwave = blockIdx.x * (blockDim.x / kWave) + wave;
    RingStack<S, 1> stk;
    stk.lds = &lds_stack[wave][0][lane];
    stk.spill = L.spill + (size_t)gwave * L.spill_depth * kWave + lane;
    // Per-lane scene and buffers: uniform in a plain launch, per item in a
    // batched one.
    const BvhNode* nodes = static_cast<const BvhNode*>(L.nodes);
    const TriRec* tris = L.tris;
    rq_hit* hout = L.hits;
    uint32_t* oout = L.occluded;
    uint32_t local = 0;

    RayState r;
    float tmin = 0, tbest = 0, bu = 0, bv = 0;
    uint32_t bprim = kNone;
    uint32_t ray = kNone;
    uint32_t cur = kRefSentinel, pend = kNone;
    const uint32_t thresh = L.fetch_threshold;
    uint32_t n_nodes = 0, n_tris = 0;  // COUNT only

    for (;;) {
        ray = fetch_work(ray, L.counter, L.n);
        if (ray != kNone && ray != kDone && cur == kRefSentinel && pend == kNone) {
            const rq_ray* rin = L.rays;
            local = ray;
            if (BATCHED) {
                const BatchItemDev* it = L.items + find_item(L.items, L.nitems, ray);
                nodes = it->nodes;
                tris = it->tris;
                rin = it->rays;
                hout = it->hits;
                oout = it->occ;
                local = ray - it->begin;
            }
            init_ray(r, rin, local, tmin, tbest);
            Pred::prepare(r);
            bu = bv = 0.0f;
            bprim = kNone;
            cur = 0;
            stk.reset();
        }
        if (__ballot(ray != kDone) == 0) break;

        // Trace until this lane's ray ends or the wave runs low on active lanes.
        while (ray != kDone && (cur != kRefSentinel || pend != kNone)) {
            // Node phase.
            while (cur != kRefSentinel && (int)cur >= 0) {
                const BvhNode* np = nodes + cur;
                if (COUNT) ++n_nodes;
                const float4 nx = ldg4(np->x), ny = ldg4(np->y), nz = ldg4(np->z);
                const uint4 meta = ldgu4(np->ref);
                float tn0, tn1;
                bool h1;
                const bool h0 = box2(r, nx, ny, nz, tmin, tbest, tn0, tn1, h1);
                if (h0 && h1) {
                    const bool first0 = tn0 <= tn1;
                    const uint32_t nearr = first0 ? meta.x : meta.y;
                    const uint32_t farr = first0 ? meta.y : meta.x;
                    uint32_t v[1] = {farr};
                    stk.push(v);
                    cur = nearr;
                } else if (h0) {
                    cur = meta.x;
                } else if (h1) {
                    cur = meta.y;
                } else {
                    uint32_t v[1];
                    cur = stk.pop(v) ? v[0] : kRefSentinel;
                }
                if ((int)cur < 0 && pend == kNone) {
                    pend = cur;
                    uint32_t v[1];
                    cur = stk.pop(v) ? v[0] : kRefSentinel;
                }
                if (__ballot(pend == kNone) == 0) break;
            }
            // Leaf phase.
            while (pend != kNone) {
                const uint32_t first = pend & 0x07FFFFFFu;
                const uint32_t count = ((pend >> 27) & 15u) + 1u;
                bool done = false;
                for (uint32_t i = 0; i < count; ++i) {
                    const TriRec* tp = tris + first + i;
                    if (COUNT) ++n_tris;
                    const float4 a = ldg4(tp->a), b = ldg4(tp->b), c = ldg4(tp->c);
                    float t, u, v;
                    if (Pred::test(r, a, b, c, tmin, t, u, v) && t < tbest) {
                        tbest = t;
                        bu = u;
                        bv = v;
                        bprim = __float_as_uint(a.w);
                        if (ANY) {
                            done = true;
                            break;
                        }
                    }
                }
                if (ANY && done) {
                    pend = kNone;
                    cur = kRefSentinel;
                    break;
                }
                if (cur != kRefSentinel && (int)cur < 0) {
                    pend = cur;
                    uint32_t v[1];
                    cur = stk.pop(v) ? v[0] : kRefSentinel;
                } else {
                    pend = kNone;
                }
            }
            if ((uint32_t)__popcll(__ballot(true)) < thresh) break;
        }
        if (ray != kNone && ray != kDone && cur == kRefSentinel && pend == kNone) {
            if (ANY) {
                oout[local] = (bprim != kNone) ? 1u : 0u;
            } else {
                const bool has = bprim != kNone;
                *reinterpret_cast<float4*>(&hout[local]) =
                    make_float4(tbest, has ? bu : 0.0f, has ? bv : 0.0f, __uint_as_float(bprim));
            }
            if (COUNT) {
                atomicAdd(&L.counts[0], (unsigned long long)n_nodes);
                atomicAdd(&L.counts[1], (unsigned long long)n_tris);
                n_nodes = n_tris = 0;
            }
            ray = kNone;
        }
    }
}

// ---------------------------------------------------------------------------
// ENCORE kd-tree traversal (Kdtree.cpp:682-784), same decisions bit for bit.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float comp3(const RayState& r, int k, bool origin) {
    return origin ? comp(r.ox, r.oy, r.oz, k) : comp(r.ix, r.iy, r.iz, k);
}

// AABB::Intersect(Ray&, float&, float&) (AABB.cpp:117-143) on the root box.
__device__ __forceinline__ bool root_clip(const RayState& r, const float* pos, const float* size, float& t0o,
                                          float& t1o) {
    float t0 = 0.0f, t1 = __int_as_float(0x7f800000);
    const float o[3] = {r.ox, r.oy, r.oz};
    const float d[3] = {r.dx, r.dy, r.dz};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float tin = (pos[k] - o[k]) / d[k];
        const float mx = pos[k] + size[k];
        float tout = (mx - o[k]) / d[k];
        if (tin > tout) {
            const float s = tin;
            tin = tout;
            tout = s;
        }
        t0 = tin > t0 ? tin : t0;
        t1 = tout < t1 ? tout : t1;
        if (t0 > t1) return false;
    }
    t0o = t0;
    t1o = t1;
    return true;
}

// Lanes run a state machine -- one node step, or up to four triangle tests of
// the current leaf, per iteration -- and idle lanes are refilled from the
// counter whenever fewer than fetch_threshold lanes of the wave are busy, so
// rays of very different cost (a few to thousands of tests) keep the SIMD
// full.  Leaf triangles are stored in leaf order (one record per reference),
// so a leaf's records are contiguous and need no index load.
template <class Pred, bool ANY, int S, bool COUNT = false>
__global__ __launch_bounds__(256) void kd_trace(TraceLaunch L) {
    __shared__ uint32_t lds_stack[256 / kWave][S * 3][kWave];
    const int lane = lane_id();
    const int wave = threadIdx.x / kWave;
    const uint32_t gwave = blockIdx.x * (blockDim.x / kWave) + wave;
    RingStack<S, 3> stk;
    stk.lds = &lds_stack[wave][0][lane];
    stk.spill = L.spill + (size_t)gwave * L.spill_depth * 3 * kWave + lane;
    const KdNodeG* nodes = static_cast<const KdNodeG*>(L.nodes);
    const TriRec* tris = L.tris;
    const float kInf = __int_as_float(0x7f800000);
    const uint32_t thresh = L.fetch_threshold;

    RayState r;
    uint32_t ray = kNone;
    bool active = false;
    float rtmax = 0, best = 0, bu = 0, bv = 0, tmin = 0, tmax = 0;
    uint32_t bprim = kNone, cur = 0, lpos = 0, lend = 0;
    uint32_t n_nodes = 0, n_tris = 0;  // COUNT only

    for (;;) {
        ray = fetch_work(ray, L.counter, L.n);
        if (ray != kNone && ray != kDone && !active) {
            float rtmin;
            init_ray(r, L.rays, ray, rtmin, rtmax);
            Pred::prepare(r);
            best = kInf;  // HitInfo.hitTime = INFINITY (Kdtree.cpp:685)
            bu = bv = 0.0f;
            bprim = kNone;
            cur = 0;
            lpos = lend = 0;
            stk.reset();
            active = root_clip(r, L.root_pos, L.root_size, tmin, tmax);
            if (!active) {  // the ray misses the scene box
                if (ANY)
                    L.occluded[ray] = 0u;
                else
                    *reinterpret_cast<float4*>(&L.hits[ray]) = make_float4(rtmax, 0.0f, 0.0f, __uint_as_float(kNone));
                ray = kNone;
            }
        }
        if (__ballot(ray != kDone) == 0) break;

        while (active) {
            bool finish = false;
            if (lpos < lend) {
                // Up to four triangles of the current leaf, in leaf order.
                const uint32_t cnt = min(4u, lend - lpos);
                float4 ta[4], tb[4], tc[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    if ((uint32_t)i < cnt) {
                        const TriRec* tp = tris + lpos + i;
                        ta[i] = ldg4(tp->a);
                        tb[i] = ldg4(tp->b);
                        tc[i] = ldg4(tp->c);
                    }
                }
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    if ((uint32_t)i < cnt && !finish) {
                        if (COUNT) ++n_tris;
                        float t, u, v;
                        // A closer hit replaces the record (strict <); an empty
                        // record (hitTime FLT_MAX) replaces the initial +inf
                        // (HitInfo.cpp:8-18, Kdtree.cpp:757-765).
                        if (Pred::test(r, ta[i], tb[i], tc[i], 0.0f, t, u, v)) {
                            if (t < best) {
                                best = t;
                                bu = u;
                                bv = v;
                                bprim = __float_as_uint(ta[i].w);
                                if (ANY && t < rtmax) finish = true;
                            }
                        } else if (best == kInf) {
                            best = 3.40282347e+38f;
                        }
                    }
                }
                lpos += cnt;
                if (!finish && lpos >= lend) {
                    uint32_t v[3];
                    if (stk.pop(v)) {
                        cur = v[0];
                        tmin = __uint_as_float(v[1]);
                        tmax = __uint_as_float(v[2]);
                    } else {
                        finish = true;
                    }
                }
            } else if (best < tmin) {
                finish = true;
            } else {
                if (COUNT) ++n_nodes;
                const uint2 nd = *reinterpret_cast<const uint2*>(&nodes[cur]);
                if ((nd.x & 3u) != 3u) {
                    const int axis = (int)(nd.x & 3u);
                    float split = __uint_as_float(nd.x);
                    split = (split > -0.001f && split < 0.001f) ? 0.0f : split;
                    const float oa = comp(r.ox, r.oy, r.oz, axis);
                    const float ia = comp(r.ix, r.iy, r.iz, axis);
                    const float tplane = (split - oa) * ia;
                    const bool below = oa <= split;
                    const uint32_t first = below ? nd.y : nd.y + 1;
                    const uint32_t second = below ? nd.y + 1 : nd.y;
                    if (tplane == 0.0f) {
                        cur = (ia > 0.0f) ? second : first;
                    } else if (tplane >= tmax || tplane < 0.0f) {
                        cur = first;
                    } else if (tplane <= tmin) {
                        cur = second;
                    } else {
                        uint32_t v[3] = {second, __float_as_uint(tplane), __float_as_uint(tmax)};
                        stk.push(v);
                        cur = first;
                        tmax = tplane;
                    }
                } else {
                    lpos = nd.y;
                    lend = nd.y + (nd.x >> 2);
                    if (lpos == lend) {  // empty leaf: straight to the next entry
                        uint32_t v[3];
                        if (stk.pop(v)) {
                            cur = v[0];
                            tmin = __uint_as_float(v[1]);
                            tmax = __uint_as_float(v[2]);
                        } else {
                            finish = true;
                        }
                    }
                }
            }
            if (finish) {
                const bool has = (bprim != kNone) && (best < rtmax);
                if (ANY) {
                    L.occluded[ray] = has ? 1u : 0u;
                } else {
                    *reinterpret_cast<float4*>(&L.hits[ray]) = make_float4(
                        has ? best : rtmax, has ? bu : 0.0f, has ? bv : 0.0f, __uint_as_float(has ? bprim : kNone));
                }
                if (COUNT) {
                    atomicAdd(&L.counts[0], (unsigned long long)n_nodes);
                    atomicAdd(&L.counts[1], (unsigned long long)n_tris);
                    n_nodes = n_tris = 0;
                }
                active = false;
                ray = kNone;
            }
            if ((uint32_t)__popcll(__ballot(true)) < thresh) break;
        }
    }
}

__global__ void zero_u32(uint32_t* p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

}  // namespace dev

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
// kd ring: 8 entries x 3 dwords x 64 lanes x 4 waves = 24 KB of LDS per block.
constexpr int kKdStack = 8;
template <class Pred, bool ANY>
static void launch_bvh_t(const TraceLaunch& L, int grid, const KernelConfig& kc, hipStream_t s) {
    switch (kc.stack) {
        case 8:
            hipLaunchKernelGGL((dev::bvh_trace<Pred, ANY, 8>), dim3(grid), dim3(256), 0, s, L);
            break;
        case 32:
            hipLaunchKernelGGL((dev::bvh_trace<Pred, ANY, 32>), dim3(grid), dim3(256), 0, s, L);
            break;
        default:
            hipLaunchKernelGGL((dev::bvh_trace<Pred, ANY, 16>), dim3(grid), dim3(256), 0, s, L);
            break;
    }
}

template <class Pred, bool ANY>
static void launch_bvh_batched_t(const TraceLaunch& L, int grid, const KernelConfig& kc, hipStream_t s) {
    switch (kc.stack) {
        case 8:
            hipLaunchKernelGGL((dev::bvh_trace<Pred, ANY, 8, false, true>), dim3(grid), dim3(256), 0, s, L);
            break;
        case 32:
            hipLaunchKernelGGL((dev::bvh_trace<Pred, ANY, 32, false, true>), dim3(grid), dim3(256), 0, s, L);
            break;
        default:
            hipLaunchKernelGGL((dev::bvh_trace<Pred, ANY, 16, false, true>), dim3(grid), dim3(256), 0, s, L);
            break;
    }
}

void launch_bvh_batched(const TraceLaunch& L, uint32_t predicate, TraceKind kind, int grid, const KernelConfig& kc,
                        hipStream_t s) {
    const bool any = kind == TraceKind::Any;
    if (predicate == RQ_PRED_WATERTIGHT) {
        if (any)
            launch_bvh_batched_t<dev::PredWatertight, true>(L, grid, kc, s);
        else
            launch_bvh_batched_t<dev::PredWatertight, false>(L, grid, kc, s);
    } else {
        if (any)
            launch_bvh_batched_t<dev::PredEncore, true>(L, grid, kc, s);
        else
            launch_bvh_batched_t<dev::PredEncore, false>(L, grid, kc, s);
    }
    RQ_HIP(hipGetLastError());
}

int occupancy_bvh_batched(uint32_t predicate, TraceKind kind, int stack) {
    const bool any = kind == TraceKind::Any;
    const void* fn;
#define RQ_PICKB(P, A)                                                                        \
    (stack == 8    ? (const void*)dev::bvh_trace<P, A, 8, false, true>                       \
     : stack == 32 ? (const void*)dev::bvh_trace<P, A, 32, false, true>                      \
                   : (const void*)dev::bvh_trace<P, A, 16, false, true>)
    if (predicate == RQ_PRED_WATERTIGHT)
        fn = any ? RQ_PICKB(dev::PredWatertight, true) : RQ_PICKB(dev::PredWatertight, false);
    else
        fn = any ? RQ_PICKB(dev::PredEncore, true) : RQ_PICKB(dev::PredEncore, false);
#undef RQ_PICKB
    int nb = 0;
    RQ_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, 256, 0));
    return nb > 0 ? nb : 1;
}

template <class Pred, bool ANY>
static void launch_kd_t(const TraceLaunch& L, int grid, hipStream_t s) {
    hipLaunchKernelGGL((dev::kd_trace<Pred, ANY, kKdStack>), dim3(grid), dim3(256), 0, s, L);
}

void launch_bvh(const TraceLaunch& L, uint32_t predicate, TraceKind kind, int grid, const KernelConfig& kc,
                hipStream_t s) {
    const bool any = kind == TraceKind::Any;
    if (predicate == RQ_PRED_WATERTIGHT) {
        if (any)
            launch_bvh_t<dev::PredWatertight, true>(L, grid, kc, s);
        else
            launch_bvh_t<dev::PredWatertight, false>(L, grid, kc, s);
    } else {
        if (any)
            launch_bvh_t<dev::PredEncore, true>(L, grid, kc, s);
        else
            launch_bvh_t<dev::PredEncore, false>(L, grid, kc, s);
    }
    RQ_HIP(hipGetLastError());
}

void launch_kd(const TraceLaunch& L, uint32_t predicate, TraceKind kind, int grid, const KernelConfig& kc,
               hipStream_t s) {
    (void)kc;
    const bool any = kind == TraceKind::Any;
    if (predicate == RQ_PRED_WATERTIGHT) {
        if (any)
            launch_kd_t<dev::PredWatertight, true>(L, grid, s);
        else
            launch_kd_t<dev::PredWatertight, false>(L, grid, s);
    } else {
        if (any)
            launch_kd_t<dev::PredEncore, true>(L, grid, s);
        else
            launch_kd_t<dev::PredEncore, false>(L, grid, s);
    }
    RQ_HIP(hipGetLastError());
}

void launch_count(const TraceLaunch& L, uint32_t accel, uint32_t predicate, TraceKind kind, int grid,
                  hipStream_t s) {
    const bool any = kind == TraceKind::Any;
#define RQ_CNT(K, P, A) hipLaunchKernelGGL((dev::K<P, A, 16, true>), dim3(grid), dim3(256), 0, s, L)
#define RQ_CNTK(K, P, A) hipLaunchKernelGGL((dev::K<P, A, kKdStack, true>), dim3(grid), dim3(256), 0, s, L)
    if (accel == RQ_ACCEL_KD) {
        if (predicate == RQ_PRED_WATERTIGHT) {
            if (any) RQ_CNTK(kd_trace, dev::PredWatertight, true); else RQ_CNTK(kd_trace, dev::PredWatertight, false);
        } else {
            if (any) RQ_CNTK(kd_trace, dev::PredEncore, true); else RQ_CNTK(kd_trace, dev::PredEncore, false);
        }
    } else {
        if (predicate == RQ_PRED_WATERTIGHT) {
            if (any) RQ_CNT(bvh_trace, dev::PredWatertight, true); else RQ_CNT(bvh_trace, dev::PredWatertight, false);
        } else {
            if (any) RQ_CNT(bvh_trace, dev::PredEncore, true); else RQ_CNT(bvh_trace, dev::PredEncore, false);
        }
    }
#undef RQ_CNT
#undef RQ_CNTK
    RQ_HIP(hipGetLastError());
}

void launch_zero_u32(uint32_t* p, uint32_t n, hipStream_t s) {
    hipLaunchKernelGGL(dev::zero_u32, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
    RQ_HIP(hipGetLastError());
}

int occupancy_blocks_per_cu(uint32_t accel, uint32_t predicate, TraceKind kind, const KernelConfig& kc) {
    int nb = 0;
    const bool any = kind == TraceKind::Any;
    const void* fn = nullptr;
    if (accel == RQ_ACCEL_KD) {
        if (predicate == RQ_PRED_WATERTIGHT)
            fn = any ? (const void*)dev::kd_trace<dev::PredWatertight, true, kKdStack>
                     : (const void*)dev::kd_trace<dev::PredWatertight, false, kKdStack>;
        else
            fn = any ? (const void*)dev::kd_trace<dev::PredEncore, true, kKdStack>
                     : (const void*)dev::kd_trace<dev::PredEncore, false, kKdStack>;
    } else {
#define RQ_PICK(P, A)                                                               \
    (kc.stack == 8    ? (const void*)dev::bvh_trace<P, A, 8>                       \
     : kc.stack == 32 ? (const void*)dev::bvh_trace<P, A, 32>                      \
                      : (const void*)dev::bvh_trace<P, A, 16>)
        if (predicate == RQ_PRED_WATERTIGHT)
            fn = any ? RQ_PICK(dev::PredWatertight, true) : RQ_PICK(dev::PredWatertight, false);
        else
            fn = any ? RQ_PICK(dev::PredEncore, true) : RQ_PICK(dev::PredEncore, false);
#undef RQ_PICK
    }
    RQ_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, 256, 0));
    return nb > 0 ? nb : 1;
}

uint32_t spill_entry_words(uint32_t accel) { return accel == RQ_ACCEL_KD ? 3u : 1u; }

}  // namespace rq
