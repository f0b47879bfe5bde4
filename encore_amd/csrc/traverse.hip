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

#include <cstdlib>
#include <cstring>

#include "rq_internal.h"

namespace rq {
namespace dev {

constexpr int kWave = 64;
constexpr uint32_t kNone = 0xFFFFFFFFu;

// Loads and stores through the global address space.  Every buffer the
// kernels touch is device memory, whose generic address equals its global
// address; going through an integer makes the compiler emit global_load /
// global_store (which wait on vmcnt alone) even for pointers read from a
// descriptor table, where it could not prove the address space and would
// emit flat accesses that also wait on lgkmcnt.
typedef float F4v __attribute__((ext_vector_type(4)));
typedef uint32_t U4v __attribute__((ext_vector_type(4)));
typedef const __attribute__((address_space(1))) F4v GF4v;
typedef const __attribute__((address_space(1))) U4v GU4v;
typedef __attribute__((address_space(1))) F4v GF4vOut;
typedef __attribute__((address_space(1))) uint32_t GU32Out;
__device__ __forceinline__ float4 ldg4(const void* p) {
    const F4v v = *(GF4v*)(uintptr_t)p;
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ uint4 ldgu4(const void* p) {
    const U4v v = *(GU4v*)(uintptr_t)p;
    return make_uint4(v.x, v.y, v.z, v.w);
}
// Streaming accesses (rays read once, hits written once): nontemporal, so
// that they do not displace the scene's nodes and triangles in the caches.
__device__ __forceinline__ float4 ldg4_stream(const void* p) {
    const F4v v = __builtin_nontemporal_load((GF4v*)(uintptr_t)p);
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void stg4_stream(void* p, float4 v) {
    const F4v o = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(o, (GF4vOut*)(uintptr_t)p);
}
__device__ __forceinline__ void stg1_stream(uint32_t* p, uint32_t v) {
    __builtin_nontemporal_store(v, (GU32Out*)(uintptr_t)p);
}

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
    // Makes room for n more entries (spilling the oldest), so that put(i)
    // for i < n and advance(n) need no check of their own.
    __device__ __forceinline__ void reserve(int n) {
        while (sp - lo + n > S) {
            const int s = lo & (S - 1);
#pragma unroll
            for (int w = 0; w < W; ++w) spill[((size_t)lo * W + w) * kWave] = lds[(s * W + w) * kWave];
            ++lo;
        }
    }
    __device__ __forceinline__ void put(int i, const uint32_t (&v)[W]) {
        const int s = (sp + i) & (S - 1);
#pragma unroll
        for (int w = 0; w < W; ++w) lds[(s * W + w) * kWave] = v[w];
    }
    __device__ __forceinline__ void advance(int n) { sp += n; }
};

// ---------------------------------------------------------------------------
// Triangle predicates
// ---------------------------------------------------------------------------
struct RayState {
    float ox, oy, oz, dx, dy, dz;
    float ix, iy, iz;  // 1/d
    // BVH slab offsets: a plane b is entered at fma(b, i, cn) or left at
    // fma(b, i, cf) (see slab_offsets)
    float cnx, cny, cnz, cfx, cfy, cfz;
    // watertight shear
    int kx, ky, kz;
    float sx, sy, sz;
};

// Conservative slab distances with one fma per plane.  The exact distance
// to plane b is (b - o) i = b i - P with P = o i.  With p = rn(P) and
// D = |p| 2^-22 (at least two ulps of P), cn = -p - D and cf = -p + D keep
// fma(b, i, cn) at or below, and fma(b, i, cf) at or above, the exact
// distance up to the final rounding of the fma (half an ulp of the result,
// covered by the relative widening applied to the far distance).
//
// A zero direction component (i infinite) would turn the slab into NaN and
// stop it culling anything, so such a ray would visit every box crossing
// its plane.  There i becomes +-2^80 (its sign kept) and D at least 2^40:
// the slab then rejects exactly the boxes whose extent on that axis misses
// o by more than 2^-40 and keeps (conservatively) those within 2^-40.
__device__ __forceinline__ void slab_offsets(float o, float& i, float& cn, float& cf) {
    float dmin = 0.0f;
    if (__builtin_isinf(i)) {
        i = __builtin_copysignf(0x1p80f, i);
        dmin = 0x1p40f;
    }
    const float p = o * i;
    const float D = fmaxf(fabsf(p) * 2.38418579e-07f, dmin);
    cn = -p - D;
    cf = -p + D;
}

__device__ __forceinline__ float flush_small(float x) { return (x > -0.001f && x < 0.001f) ? 0.0f : x; }

// ENCORE's Moller-Trumbore variant (TrianglePrim.cpp:204-249): edges given,
// Cross() with the component flush, no determinant test, inclusive edges.
struct PredEncore {
    static constexpr bool kShear = false;  // reads only the ray's origin and direction
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
    static constexpr bool kShear = true;  // also reads the shear (kx, ky, kz, sx, sy, sz)
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
    const float4 p = ldg4_stream(&rays[id].ox);
    const float4 q = ldg4_stream(&rays[id].dx);
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
    const float n0x = __builtin_fmaf(sx ? nx.y : nx.x, r.ix, r.cnx);
    const float f0x = __builtin_fmaf(sx ? nx.x : nx.y, r.ix, r.cfx);
    const float n0y = __builtin_fmaf(sy ? ny.y : ny.x, r.iy, r.cny);
    const float f0y = __builtin_fmaf(sy ? ny.x : ny.y, r.iy, r.cfy);
    const float n0z = __builtin_fmaf(sz ? nz.y : nz.x, r.iz, r.cnz);
    const float f0z = __builtin_fmaf(sz ? nz.x : nz.y, r.iz, r.cfz);
    // child 1
    const float n1x = __builtin_fmaf(sx ? nx.w : nx.z, r.ix, r.cnx);
    const float f1x = __builtin_fmaf(sx ? nx.z : nx.w, r.ix, r.cfx);
    const float n1y = __builtin_fmaf(sy ? ny.w : ny.z, r.iy, r.cny);
    const float f1y = __builtin_fmaf(sy ? ny.z : ny.w, r.iy, r.cfy);
    const float n1z = __builtin_fmaf(sz ? nz.w : nz.z, r.iz, r.cnz);
    const float f1z = __builtin_fmaf(sz ? nz.z : nz.w, r.iz, r.cfz);
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

// Compare-exchange of (distance, ref) pairs, ascending by distance.
__device__ __forceinline__ void cswap(float& ka, uint32_t& ra, float& kb, uint32_t& rb) {
    const bool sw = kb < ka;
    const float k = sw ? kb : ka;
    kb = sw ? ka : kb;
    ka = k;
    const uint32_t r = sw ? rb : ra;
    rb = sw ? ra : rb;
    ra = r;
}

// Ray staging.  A wave claims rays in chunks of 64 consecutive ids with one
// atomicAdd, loads them with every lane busy, computes 1/d (and, in a
// batched launch, the owning item and the item-local index) and parks the
// records in LDS, [field][slot].  Idle lanes are refilled from there: a
// ballot, a prefix count and a few LDS reads, with no global memory access
// and no division on the refill path, so lanes can be refilled as soon as a
// handful of them are idle.
constexpr int kStageFields = 13;  // ox oy oz tmin dx dy dz tmax ix iy iz local item
struct Stage {
    float* base;  // &lds_stage[wave][0][0]
    __device__ __forceinline__ float& at(int f, uint32_t slot) { return base[f * kWave + slot]; }
};

// Item of a batched launch whose rays include global ray index `id` (items
// sorted by begin): a wave-uniform binary search on the chunk's first id,
// then a per-lane forward scan (a 64-ray chunk spans few items).
__device__ __forceinline__ uint32_t chunk_item(const BatchItemDev* items, uint32_t nitems, uint32_t first_id,
                                               uint32_t id) {
    uint32_t lo = 0, hi = nitems;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (items[mid].begin <= first_id)
            lo = mid;
        else
            hi = mid;
    }
    while (lo + 1 < nitems && items[lo + 1].begin <= id) ++lo;
    return lo;
}

// Entry distances of a BVH4 node's four children from the node's
// sign-selected near/far rows (near x, far x, near y, far y, near z, far z)
// and its refs: k[j] is child j's entry distance when the ray enters it
// before tbest, +inf otherwise; c[j] its ref.
__device__ __forceinline__ void keys4(const RayState& r, const float4 (&d)[7], float tmin, float tbest, float (&k)[4],
                                      uint32_t (&c)[4]) {
    const float kInf = __int_as_float(0x7f800000);
    const float4 nx = d[0], fx = d[1], ny = d[2], fy = d[3], nz = d[4], fz = d[5], rf = d[6];
    c[0] = __float_as_uint(rf.x);
    c[1] = __float_as_uint(rf.y);
    c[2] = __float_as_uint(rf.z);
    c[3] = __float_as_uint(rf.w);
    const float nxa[4] = {nx.x, nx.y, nx.z, nx.w}, fxa[4] = {fx.x, fx.y, fx.z, fx.w};
    const float nya[4] = {ny.x, ny.y, ny.z, ny.w}, fya[4] = {fy.x, fy.y, fy.z, fy.w};
    const float nza[4] = {nz.x, nz.y, nz.z, nz.w}, fza[4] = {fz.x, fz.y, fz.z, fz.w};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const float tnx = __builtin_fmaf(nxa[j], r.ix, r.cnx), tfx = __builtin_fmaf(fxa[j], r.ix, r.cfx);
        const float tny = __builtin_fmaf(nya[j], r.iy, r.cny), tfy = __builtin_fmaf(fya[j], r.iy, r.cfy);
        const float tnz = __builtin_fmaf(nza[j], r.iz, r.cnz), tfz = __builtin_fmaf(fza[j], r.iz, r.cfz);
        const float tn = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, tmin));
        const float tf = fminf(fminf(tfx, tfy), fminf(tfz, tbest)) * 1.00000048f;
        k[j] = (tn <= tf && c[j] != kRefEmpty) ? tn : kInf;
    }
}

// The entered children sorted by entry distance with a 5-comparator
// network; the nearest is visited next and the others are pushed far to
// near with their entry distance.
template <int S, class Pop>
__device__ __forceinline__ void order_push4(float (&k)[4], uint32_t (&c)[4], RingStack<S, 2>& stk, uint32_t& cur,
                                            Pop&& pop_next) {
    const float kInf = __int_as_float(0x7f800000);
    cswap(k[0], c[0], k[1], c[1]);
    cswap(k[2], c[2], k[3], c[3]);
    cswap(k[0], c[0], k[2], c[2]);
    cswap(k[1], c[1], k[3], c[3]);
    cswap(k[1], c[1], k[2], c[2]);
    if (k[0] < kInf) {
        // Push children 1..nh-1 far to near: child j lands at offset
        // npush - j from the current top.
        const int npush = (k[1] < kInf) + (k[2] < kInf) + (k[3] < kInf);
        stk.reserve(npush);
        if (npush >= 3) {
            uint32_t v[2] = {c[3], __float_as_uint(k[3])};
            stk.put(0, v);
        }
        if (npush >= 2) {
            uint32_t v[2] = {c[2], __float_as_uint(k[2])};
            stk.put(npush - 2, v);
        }
        if (npush >= 1) {
            uint32_t v[2] = {c[1], __float_as_uint(k[1])};
            stk.put(npush - 1, v);
        }
        stk.advance(npush);
        cur = c[0];
    } else {
        cur = pop_next();
    }
}

// BVH4 node step on the node's sign-selected near/far rows and refs.
template <int S, class Pop>
__device__ __forceinline__ void node4_step(const RayState& r, const float4 (&d)[7], float tmin, float tbest,
                                           RingStack<S, 2>& stk, uint32_t& cur, Pop&& pop_next) {
    float k[4];
    uint32_t c[4];
    keys4(r, d, tmin, tbest, k, c);
    order_push4(k, c, stk, cur, pop_next);
}

// One traversal step per lane per iteration, through shared loads.  Every
// busy lane fetches one record with the same NL 16-byte load instructions --
// a node (BVH4: the sign-selected near/far rows and the refs, 7 loads; BVH2:
// the 64-byte node, 4 loads) or the next TPS triangles of its current leaf
// (BVH4: 2 triangles in 6 loads; BVH2: 1 in 3) -- and then runs the node or
// the triangle arithmetic.  Node and leaf lanes share the load instructions
// and the loop, so there are no separate node and leaf phases whose lanes
// wait for each other.
//
// WIDTH 2: BvhNode (both child boxes in one 64-byte node), stack entries are
// bare refs.  WIDTH 4: BvhNode4 (four child boxes, 128 bytes), the hit
// children are sorted by entry distance with a 5-comparator network, the
// nearest is visited next and the others are pushed far to near together
// with their entry distance; a popped entry whose entry distance is beyond
// the current closest hit is dropped (the box test would now fail).
//
// One wave per block.  Idle lanes are refilled from the staged chunk
// whenever fewer than fetch_threshold lanes are busy.
//
// COOP (BVH4, RQ_BVH_LEAF_MODE=wave): leaves are tested by the whole wave
// instead (see below).  It cuts VALU instructions by about a third and raises
// the active lanes per VALU instruction, but at 116 VGPRs it holds fewer waves,
// and the config-3 launch, bound by node-fetch latency, runs no faster.
template <class Pred, bool ANY, int S, bool COUNT = false, bool BATCHED = false, int WIDTH = 2, bool COOP = false>
__global__ __launch_bounds__(kWave) void bvh_trace(TraceLaunch L) {
    constexpr int SW = WIDTH == 4 ? 2 : 1;      // dwords per stack entry
    constexpr int NL = WIDTH == 4 ? 7 : 4;      // 16-byte rows a lane reads from its record
    constexpr int TPS = WIDTH == 4 ? 2 : 1;     // triangles per leaf step
    __shared__ uint32_t lds_stack[S * SW][kWave];
    __shared__ float lds_stage[kStageFields][kWave];
    __shared__ unsigned long long lds_leafkey[COOP ? kWave : 1];
    const int lane = lane_id();
    const uint32_t gwave = blockIdx.x;
    RingStack<S, SW> stk;
    stk.lds = &lds_stack[0][lane];
    stk.spill = L.spill + (size_t)gwave * L.spill_depth * kWave + lane;
    Stage stage{&lds_stage[0][0]};
    // Per-lane scene and output: uniform in a plain launch, per item in a
    // batched one.
    const void* nodes = L.nodes;
    const TriRec* tris = L.tris;
    void* out = ANY ? static_cast<void*>(L.occluded) : static_cast<void*>(L.hits);
    const uint32_t nrays = L.n_dev ? *L.n_dev : L.n;
    const float kInf = __int_as_float(0x7f800000);

    RayState r;
    float tmin = 0, tbest = 0, bu = 0, bv = 0;
    uint32_t bprim = kNone;
    uint32_t local = kNone;  // item-local ray index; kNone = idle lane
    // cur: next record (inner node, leaf ref or kRefSentinel); [lpos, lend):
    // triangles of the leaf in progress.
    uint32_t cur = kRefSentinel, lpos = 0, lend = 0;
    uint32_t n_nodes = 0, n_tris = 0;  // COUNT only
    // COUNT only: iterations, iterations with a leaf lane, refill rounds
    // (tallied by lane 0).
    uint32_t w_node = 0, w_tri = 0, w_round = 0;
    // Byte offsets of the near-plane rows picked by direction sign (BVH4):
    // x in {0, 16}, y in {32, 48}, z in {64, 80}; the far row is at offset ^ 16.
    uint32_t offx = 0, offy = 0, offz = 0;
    // Wave-uniform chunk state: the claimed id range [cl_base, cl_end) and
    // the staged chunk (ch_n rays from ch_base, ch_used handed out).
    uint32_t cl_base = 0, cl_end = 0, ch_base = 0, ch_n = 0, ch_used = 0;
    bool exhausted = false;

    // Next stacked ref, skipping (WIDTH 4) entries the closest hit has passed.
    auto pop_next = [&]() -> uint32_t {
        uint32_t v[SW];
        while (stk.pop(v)) {
            if (SW == 1 || __uint_as_float(v[SW - 1]) <= tbest * 1.00000048f) return v[0];
        }
        return kRefSentinel;
    };

    for (;;) {
        // Refill idle lanes once fewer than fetch_threshold are busy (or the
        // wave is empty), staging new chunks as needed.  Ids are claimed in
        // guided ranges -- a quarter of this wave's share of what was left at
        // its last claim, 64..2048 ids -- so the shared counter sees a few
        // atomics per wave rather than one per 64 rays.
        const uint64_t idle0 = __ballot(local == kNone);
        if (!exhausted && (idle0 == ~0ull || 64u - (uint32_t)__popcll(idle0) < L.fetch_threshold)) {
            if (COUNT && lane == 0) ++w_round;
            for (;;) {
                const uint64_t idle = __ballot(local == kNone);
                if (idle == 0) break;
                if (ch_used == ch_n) {
                    if (cl_base >= cl_end) {
                        const uint32_t left = nrays - min(nrays, cl_end);
                        const uint32_t want =
                            min(2048u, max((uint32_t)kWave, (left / (4u * gridDim.x)) & ~(uint32_t)(kWave - 1)));
                        // 64-bit counter: it runs past n by up to a claim per
                        // wave, which must not wrap for n close to 2^32.
                        unsigned long long b = 0;
                        if (lane == 0)
                            b = atomicAdd(reinterpret_cast<unsigned long long*>(L.counter), (unsigned long long)want);
                        b = ((unsigned long long)__builtin_amdgcn_readfirstlane((uint32_t)(b >> 32)) << 32) |
                            __builtin_amdgcn_readfirstlane((uint32_t)b);
                        if (b >= nrays) {
                            exhausted = true;
                            break;
                        }
                        cl_base = (uint32_t)b;
                        cl_end = (uint32_t)min((unsigned long long)nrays, b + want);
                    }
                    const uint32_t base = cl_base;
                    ch_base = base;
                    ch_n = min((uint32_t)kWave, cl_end - base);
                    cl_base += ch_n;
                    ch_used = 0;
                    if ((uint32_t)lane < ch_n) {
                        const uint32_t id = base + lane;
                        const rq_ray* rin = L.rays;
                        uint32_t loc = id, item = 0;
                        if (BATCHED) {
                            item = chunk_item(L.items, L.nitems, base, id);
                            const BatchItemDev* it = L.items + item;
                            rin = it->rays;
                            loc = id - it->begin;
                        }
                        const float4 p = ldg4_stream(&rin[loc].ox);
                        const float4 q = ldg4_stream(&rin[loc].dx);
                        stage.at(0, lane) = p.x;
                        stage.at(1, lane) = p.y;
                        stage.at(2, lane) = p.z;
                        stage.at(3, lane) = p.w;
                        stage.at(4, lane) = q.x;
                        stage.at(5, lane) = q.y;
                        stage.at(6, lane) = q.z;
                        stage.at(7, lane) = q.w;
                        stage.at(8, lane) = 1.0f / q.x;
                        stage.at(9, lane) = 1.0f / q.y;
                        stage.at(10, lane) = 1.0f / q.z;
                        stage.at(11, lane) = __uint_as_float(loc);
                        stage.at(12, lane) = __uint_as_float(item);
                    }
                }
                const uint32_t take = min((uint32_t)__popcll(idle), ch_n - ch_used);
                const uint32_t rank = lanes_below(idle);
                if (local == kNone && rank < take) {
                    const uint32_t slot = ch_used + rank;
                    r.ox = stage.at(0, slot);
                    r.oy = stage.at(1, slot);
                    r.oz = stage.at(2, slot);
                    tmin = stage.at(3, slot);
                    r.dx = stage.at(4, slot);
                    r.dy = stage.at(5, slot);
                    r.dz = stage.at(6, slot);
                    tbest = stage.at(7, slot);
                    r.ix = stage.at(8, slot);
                    r.iy = stage.at(9, slot);
                    r.iz = stage.at(10, slot);
                    slab_offsets(r.ox, r.ix, r.cnx, r.cfx);
                    slab_offsets(r.oy, r.iy, r.cny, r.cfy);
                    slab_offsets(r.oz, r.iz, r.cnz, r.cfz);
                    local = __float_as_uint(stage.at(11, slot));
                    if (BATCHED) {
                        const BatchItemDev* it = L.items + __float_as_uint(stage.at(12, slot));
                        const uint4 pp = ldgu4(&it->nodes);
                        nodes = reinterpret_cast<const void*>(((uint64_t)pp.y << 32) | pp.x);
                        tris = reinterpret_cast<const TriRec*>(((uint64_t)pp.w << 32) | pp.z);
                        const uint4 oo = ldgu4(&it->hits);
                        out = reinterpret_cast<void*>(ANY ? (((uint64_t)oo.w << 32) | oo.z)
                                                          : (((uint64_t)oo.y << 32) | oo.x));
                    }
                    Pred::prepare(r);
                    if (WIDTH == 4) {
                        offx = __builtin_signbit(r.dx) ? 16u : 0u;
                        offy = __builtin_signbit(r.dy) ? 48u : 32u;
                        offz = __builtin_signbit(r.dz) ? 80u : 64u;
                    } else {
                        offx = offy = offz = 0u;
                    }
                    bu = bv = 0.0f;
                    bprim = kNone;
                    cur = 0;
                    lpos = lend = 0;
                    stk.reset();
                }
                ch_used += take;
            }
        }
        if (__ballot(local != kNone) == 0) {
            if (exhausted) break;  // no work left for this wave
            continue;
        }
        if (COUNT && lane == 0) ++w_node;

        bool did_leaf = false;  // COUNT only

        if constexpr (COOP) {
            // Inner nodes and leaves in separate phases.  A lane at an inner
            // node takes its node step every iteration; a lane at a leaf
            // waits until leaf_batch lanes wait (or no lane is at a node),
            // and then the wave tests the waiting leaves' triangles,
            // concatenated in lane order, one per lane, 64 per round.  Lane
            // j of a round reads from LDS which leaf (owner lane) and
            // triangle hold position base + j, takes the owner's ray
            // through cross-lane reads and tests the triangle against the
            // owner's closest hit so far.  Each owner keeps the least
            // (t, position) of its hits -- the hit its own strict-< loop over
            // the leaf in order would keep -- through a 64-bit LDS minimum,
            // and pops its next entry with that hit.
            const bool at_node = local != kNone && (int)cur >= 0 && cur != kRefSentinel;
            const bool at_leaf = local != kNone && (int)cur < 0 && cur != kRefSentinel;
            // node rows first, so their loads overlap the leaf phase
            const float* np = WIDTH == 4 ? reinterpret_cast<const float*>(static_cast<const BvhNode4*>(nodes) + cur)
                                         : reinterpret_cast<const float*>(static_cast<const BvhNode*>(nodes) + cur);
            float4 d[NL];
            if (at_node) {
                if (WIDTH == 4) {
                    const char* nb = reinterpret_cast<const char*>(np);
                    d[0] = ldg4(nb + offx);
                    d[1] = ldg4(nb + (offx ^ 16u));
                    d[2] = ldg4(nb + offy);
                    d[3] = ldg4(nb + (offy ^ 16u));
                    d[4] = ldg4(nb + offz);
                    d[5] = ldg4(nb + (offz ^ 16u));
                    d[6] = ldg4(nb + 96);
                } else {
#pragma unroll
                    for (int q = 0; q < NL; ++q) d[q] = ldg4(np + 4 * q);
                }
            }
            const uint64_t lm = __ballot(at_leaf);
            const uint32_t nwait = (uint32_t)__popcll(lm);
            if (nwait != 0 && (nwait >= L.leaf_batch || __ballot(at_node) == 0)) {
                did_leaf = at_leaf;
                const uint32_t cnt = at_leaf ? ((cur >> 27) & 15u) + 1u : 0u;  // 1..16
                const uint32_t first = cur & 0x07FFFFFFu;
                // exclusive prefix of cnt over the lanes, one ballot per bit
                uint32_t start = 0, total = 0;
#pragma unroll
                for (int b = 0; b < 5; ++b) {
                    const uint64_t mb = __ballot((cnt >> b) & 1u);
                    start += lanes_below(mb) << b;
                    total += (uint32_t)__popcll(mb) << b;
                }
                if (COUNT) n_tris += cnt;
                bool hit_any = false;
                for (uint32_t base = 0; base < total; base += kWave) {
                    // (owner, triangle) of positions base .. base + 63
                    if (at_leaf) {
                        const int j0 = max(0, (int)base - (int)start);
                        const int j1 = min((int)cnt, (int)(base + kWave) - (int)start);
                        for (int j = j0; j < j1; ++j)
                            lds_leafkey[start + j - base] = ((unsigned long long)lane << 32) | (first + (uint32_t)j);
                    }
                    __builtin_amdgcn_wave_barrier();
                    const uint32_t p = base + (uint32_t)lane;
                    const unsigned long long ot = p < total ? lds_leafkey[lane] : 0ull;
                    __builtin_amdgcn_wave_barrier();
                    const int o = (int)(ot >> 32);
                    const uint32_t tri = (uint32_t)ot;
                    RayState orr;
                    orr.ox = __shfl(r.ox, o, kWave);
                    orr.oy = __shfl(r.oy, o, kWave);
                    orr.oz = __shfl(r.oz, o, kWave);
                    orr.dx = __shfl(r.dx, o, kWave);
                    orr.dy = __shfl(r.dy, o, kWave);
                    orr.dz = __shfl(r.dz, o, kWave);
                    if constexpr (Pred::kShear) {
                        orr.kx = __shfl(r.kx, o, kWave);
                        orr.ky = __shfl(r.ky, o, kWave);
                        orr.kz = __shfl(r.kz, o, kWave);
                        orr.sx = __shfl(r.sx, o, kWave);
                        orr.sy = __shfl(r.sy, o, kWave);
                        orr.sz = __shfl(r.sz, o, kWave);
                    }
                    const float otmin = __shfl(tmin, o, kWave);
                    const float otbest = __shfl(tbest, o, kWave);
                    const TriRec* otris = tris;
                    if (BATCHED) {
                        const uint64_t tp = reinterpret_cast<uint64_t>(tris);
                        const uint32_t lo32 = (uint32_t)__shfl((int)(uint32_t)tp, o, kWave);
                        const uint32_t hi32 = (uint32_t)__shfl((int)(uint32_t)(tp >> 32), o, kWave);
                        otris = reinterpret_cast<const TriRec*>(((uint64_t)hi32 << 32) | lo32);
                    }
                    unsigned long long key = ~0ull;
                    float tu = 0.0f, tv = 0.0f;
                    uint32_t tprim = 0;
                    if (p < total) {
                        const float* tp = reinterpret_cast<const float*>(otris + tri);
                        const float4 a = ldg4(tp), b = ldg4(tp + 4), c = ldg4(tp + 8);
                        float t;
                        if (Pred::test(orr, a, b, c, otmin, t, tu, tv) && t < otbest)
                            key = ((unsigned long long)__float_as_uint(t) << 32) | p;
                        tprim = __float_as_uint(a.w);
                    }
                    if (at_leaf) lds_leafkey[lane] = ~0ull;
                    __builtin_amdgcn_wave_barrier();
                    if (key != ~0ull) atomicMin(&lds_leafkey[o], key);
                    __builtin_amdgcn_wave_barrier();
                    const unsigned long long best = at_leaf ? lds_leafkey[lane] : ~0ull;
                    const int wl = (int)((uint32_t)best - base);  // the winning lane, when best is a hit
                    const float wu = __shfl(tu, wl, kWave), wv = __shfl(tv, wl, kWave);
                    const uint32_t wp = (uint32_t)__shfl((int)tprim, wl, kWave);
                    if (best != ~0ull) {
                        tbest = __uint_as_float((uint32_t)(best >> 32));
                        bu = wu;
                        bv = wv;
                        bprim = wp;
                        hit_any = true;
                    }
                    __builtin_amdgcn_wave_barrier();
                }
                if (at_leaf) cur = (ANY && hit_any) ? kRefSentinel : pop_next();
            }
            // Node step (the loads above).
            if (at_node) {
                if (COUNT) ++n_nodes;
                if constexpr (WIDTH == 2) {
                    const uint4 meta = make_uint4(__float_as_uint(d[3].x), __float_as_uint(d[3].y),
                                                  __float_as_uint(d[3].z), __float_as_uint(d[3].w));
                    float tn0, tn1;
                    bool h1;
                    const bool h0 = box2(r, d[0], d[1], d[2], tmin, tbest, tn0, tn1, h1);
                    if (h0 && h1) {
                        const bool first0 = tn0 <= tn1;
                        uint32_t v[SW] = {first0 ? meta.y : meta.x};
                        stk.push(v);
                        cur = first0 ? meta.x : meta.y;
                    } else if (h0) {
                        cur = meta.x;
                    } else if (h1) {
                        cur = meta.y;
                    } else {
                        cur = pop_next();
                    }
                } else {
                    node4_step(r, d, tmin, tbest, stk, cur, pop_next);
                }
            }
            // Retire a finished ray.
            if (local != kNone && cur == kRefSentinel) {
                if (ANY) {
                    stg1_stream(static_cast<uint32_t*>(out) + local, (bprim != kNone) ? 1u : 0u);
                } else {
                    const bool has = bprim != kNone;
                    stg4_stream(static_cast<rq_hit*>(out) + local,
                         make_float4(tbest, has ? bu : 0.0f, has ? bv : 0.0f, __uint_as_float(bprim)));
                }
                if (COUNT) {
                    atomicAdd(&L.counts[0], (unsigned long long)n_nodes);
                    atomicAdd(&L.counts[1], (unsigned long long)n_tris);
                    n_nodes = n_tris = 0;
                }
                local = kNone;
            }
        } else if (local != kNone) {
            // A leaf ref in cur becomes the leaf in progress.
            if (lpos >= lend && (int)cur < 0 && cur != kRefSentinel) {
                lpos = cur & 0x07FFFFFFu;
                lend = lpos + ((cur >> 27) & 15u) + 1u;
                cur = kNone;  // popped when the leaf is done
            }
            const bool leaf = lpos < lend;
            // Record addresses: one base per lane (the node, or the leaf's
            // next TPS triangles) plus a byte offset per row (a missing
            // second triangle repeats the first).
            const float* np = WIDTH == 4
                                  ? reinterpret_cast<const float*>(static_cast<const BvhNode4*>(nodes) + cur)
                                  : reinterpret_cast<const float*>(static_cast<const BvhNode*>(nodes) + cur);
            const float* tp = reinterpret_cast<const float*>(tris + lpos);
            const bool two = TPS == 2 && lpos + 1 < lend;
            float4 d[NL];
            if (WIDTH == 4) {
                const char* base = reinterpret_cast<const char*>(leaf ? tp : np);
                const uint32_t o2 = two ? 48u : 0u;
                d[0] = ldg4(base + (leaf ? 0u : offx));
                d[1] = ldg4(base + (leaf ? 16u : offx ^ 16u));
                d[2] = ldg4(base + (leaf ? 32u : offy));
                d[3] = ldg4(base + (leaf ? o2 : offy ^ 16u));
                d[4] = ldg4(base + (leaf ? o2 + 16u : offz));
                d[5] = ldg4(base + (leaf ? o2 + 32u : offz ^ 16u));
                d[6] = ldg4(base + (leaf ? 0u : 96u));
            } else {
                d[0] = ldg4(leaf ? tp : np);
                d[1] = ldg4(leaf ? tp + 4 : np + 4);
                d[2] = ldg4(leaf ? tp + 8 : np + 8);
                d[3] = ldg4(leaf ? tp : np + 12);
            }
            bool finished = false;
            did_leaf = leaf;
            if (leaf) {
#pragma unroll
                for (int k = 0; k < TPS; ++k) {
                    if (k == 0 || two) {
                        if (COUNT) ++n_tris;
                        const float4 a = d[3 * k], b = d[3 * k + 1], c = d[3 * k + 2];
                        float t, u, v;
                        if (Pred::test(r, a, b, c, tmin, t, u, v) && t < tbest) {
                            tbest = t;
                            bu = u;
                            bv = v;
                            bprim = __float_as_uint(a.w);
                            if (ANY) finished = true;
                        }
                    }
                }
                lpos += TPS;
                if (!finished && lpos >= lend) cur = pop_next();
            } else {
                if (COUNT) ++n_nodes;
                if constexpr (WIDTH == 2) {
                    const uint4 meta = make_uint4(__float_as_uint(d[3].x), __float_as_uint(d[3].y),
                                                  __float_as_uint(d[3].z), __float_as_uint(d[3].w));
                    float tn0, tn1;
                    bool h1;
                    const bool h0 = box2(r, d[0], d[1], d[2], tmin, tbest, tn0, tn1, h1);
                    if (h0 && h1) {
                        const bool first0 = tn0 <= tn1;
                        uint32_t v[SW] = {first0 ? meta.y : meta.x};
                        stk.push(v);
                        cur = first0 ? meta.x : meta.y;
                    } else if (h0) {
                        cur = meta.x;
                    } else if (h1) {
                        cur = meta.y;
                    } else {
                        cur = pop_next();
                    }
                } else {
                    float k[4];
                    uint32_t c[4];
                    keys4(r, d, tmin, tbest, k, c);
                    order_push4(k, c, stk, cur, pop_next);
                }
            }
            if (ANY && finished) {
                cur = kRefSentinel;
                lpos = lend;
            }
            // Retire a finished ray.
            if (cur == kRefSentinel && lpos >= lend) {
                if (ANY) {
                    stg1_stream(static_cast<uint32_t*>(out) + local, (bprim != kNone) ? 1u : 0u);
                } else {
                    const bool has = bprim != kNone;
                    stg4_stream(static_cast<rq_hit*>(out) + local,
                         make_float4(tbest, has ? bu : 0.0f, has ? bv : 0.0f, __uint_as_float(bprim)));
                }
                if (COUNT) {
                    atomicAdd(&L.counts[0], (unsigned long long)n_nodes);
                    atomicAdd(&L.counts[1], (unsigned long long)n_tris);
                    n_nodes = n_tris = 0;
                }
                local = kNone;
            }
        }
        if (COUNT && __ballot(did_leaf) != 0 && lane == 0) ++w_tri;
    }
    if (COUNT) {
        if (w_node) atomicAdd(&L.counts[2], (unsigned long long)w_node);
        if (w_tri) atomicAdd(&L.counts[3], (unsigned long long)w_tri);
        if (w_round) atomicAdd(&L.counts[4], (unsigned long long)w_round);
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

// Claims for the kd kernel: idle lanes take rays only while the wave holds
// fewer than `cap` of them, and once the counter has run out every idle lane
// is done.  With cap = ceil(n / waves) a launch of fewer rays than lanes is
// spread over every wave of the grid instead of filling the first ones, and
// the idle lanes serve the cooperative leaf tests below.
__device__ __forceinline__ uint32_t fetch_capped(uint32_t cur, uint32_t* counter, uint32_t n, uint32_t cap,
                                                 bool drained) {
    const bool idle = (cur == kNone);
    const uint64_t m = __ballot(idle);
    if (m == 0) return cur;
    if (drained) return idle ? kDone : cur;
    const uint32_t busy = (uint32_t)__popcll(__ballot(cur != kNone && cur != kDone));
    const uint32_t want = min((uint32_t)__popcll(m), cap > busy ? cap - busy : 0u);
    if (want == 0) return cur;
    // the first `want` idle lanes claim
    uint64_t take = m;
    for (uint32_t k = (uint32_t)__popcll(m); k > want; --k) take &= ~(1ull << (63 - __clzll(take)));
    const int leader = __ffsll((unsigned long long)take) - 1;
    uint32_t base = 0;
    if (lane_id() == leader) base = atomicAdd(counter, want);
    base = __shfl(base, leader);
    if (base >= n) return idle ? kDone : cur;
    if (!((take >> lane_id()) & 1ull)) return cur;
    const uint32_t id = base + lanes_below(take);
    return id < n ? id : kDone;
}


// Lanes run a state machine -- one node step, or up to four triangle tests of
// the current leaf, per iteration -- and idle lanes are refilled from the
// counter whenever fewer than fetch_threshold lanes of the wave are busy, so
// rays of very different cost (a few to thousands of tests) keep the SIMD
// full.  Leaf triangles are stored in leaf order (one record per reference),
// so a leaf's records are contiguous and need no index load.
//
// BATCHED: the rays of every item of a descriptor table (kd scenes, any
// sizes) in one launch.  Lanes claim global ray indices; a lane's item gives
// its rays, tree, triangle records, root box (L.item_boxes, 8 floats per
// item) and output, so the scene pointers are per lane.
//
// Plain launches compile for 5 waves per SIMD (96 VGPRs, a few values
// spilled): config 3's rays in one launch ran 65.2 against 55.7 Mrays/s at 4
// waves (116 VGPRs), one launch per batch 26.8 against 25.5; config 2's
// coherent eye rays 161 against 165.  The batched launch keeps 4 waves (81.3
// against 74.0 Mrays/s at 5).
template <class Pred, bool ANY, int S, bool COUNT = false, bool BATCHED = false>
__global__ __launch_bounds__(256, BATCHED ? 4 : 5) void kd_trace(TraceLaunch L) {
    __shared__ uint32_t lds_stack[256 / kWave][S * 3][kWave];
    const int lane = lane_id();
    const int wave = threadIdx.x / kWave;
    const uint32_t gwave = blockIdx.x * (blockDim.x / kWave) + wave;
    RingStack<S, 3> stk;
    stk.lds = &lds_stack[wave][0][lane];
    stk.spill = L.spill + (size_t)gwave * L.spill_depth * kWave + lane;
    const KdNodeG* nodes = static_cast<const KdNodeG*>(L.nodes);
    const TriRec* tris = L.tris;
    rq_hit* hits_out = L.hits;
    uint32_t* occ_out = L.occluded;
    const float kInf = __int_as_float(0x7f800000);
    const uint32_t thresh = L.fetch_threshold;
    const uint32_t nrays = L.n_dev ? *L.n_dev : L.n;

    RayState r;
    uint32_t ray = kNone;
    uint32_t slot = 0;  // output index: the ray, or its index within its item
    bool active = false;
    float rtmax = 0, best = 0, bu = 0, bv = 0, tmin = 0, tmax = 0;
    uint32_t bprim = kNone, cur = 0, lpos = 0, lend = 0;
    uint32_t n_nodes = 0, n_tris = 0;  // COUNT only
    // Rays of ENCORE's tree differ in cost by three orders of magnitude (a
    // ray along the ground plane tests tens of thousands of triangles).  A
    // launch ends with its costliest rays; the per-wave cap and the
    // cooperative leaf tests below shorten that tail.
    bool drained = false;  // the work counter has run out (wave-uniform)
    const uint32_t waves = gridDim.x * (blockDim.x / kWave);
    const uint32_t cap = COUNT ? (uint32_t)kWave
                               : (uint32_t)min((uint64_t)kWave, max(1ull, ((uint64_t)nrays + waves - 1) / waves));

    for (;;) {
        ray = fetch_capped(ray, L.counter, nrays, cap, drained);
        if (ray != kNone && ray != kDone && !active) {
            float rtmin;
            const rq_ray* rin = L.rays;
            const float* bpos = L.root_pos;
            const float* bsize = L.root_size;
            slot = ray;
            if (BATCHED) {
                const uint32_t item = find_item(L.items, L.nitems, ray);
                const BatchItemDev* it = L.items + item;
                const uint4 pp = ldgu4(&it->nodes);
                nodes = reinterpret_cast<const KdNodeG*>(((uint64_t)pp.y << 32) | pp.x);
                tris = reinterpret_cast<const TriRec*>(((uint64_t)pp.w << 32) | pp.z);
                const uint4 oo = ldgu4(&it->hits);
                hits_out = reinterpret_cast<rq_hit*>(((uint64_t)oo.y << 32) | oo.x);
                occ_out = reinterpret_cast<uint32_t*>(((uint64_t)oo.w << 32) | oo.z);
                rin = it->rays;
                slot = ray - it->begin;
                bpos = L.item_boxes + 8 * (size_t)item;
                bsize = bpos + 4;
            }
            init_ray(r, rin, slot, rtmin, rtmax);
            Pred::prepare(r);
            best = kInf;  // HitInfo.hitTime = INFINITY (Kdtree.cpp:685)
            bu = bv = 0.0f;
            bprim = kNone;
            cur = 0;
            lpos = lend = 0;
            stk.reset();
            active = root_clip(r, bpos, bsize, tmin, tmax);
            if (!active) {  // the ray misses the scene box
                if (ANY)
                    stg1_stream(occ_out + slot, 0u);
                else
                    stg4_stream(hits_out + slot, make_float4(rtmax, 0.0f, 0.0f, __uint_as_float(kNone)));
                ray = kNone;
            }
        }
        if (__ballot(ray != kDone) == 0) break;
        if (!COUNT) {
            // Once the rays have run out (or the wave holds all it may), the
            // few in the wave can hold leaves of thousands of triangles,
            // tested four per step by one lane.  Then the wave splits into
            // one group per such ray, and
            // each group tests its ray's leaf gs triangles per step, keeping
            // the first (in leaf order) of the closest hits that beat the
            // ray's record -- the record the lane's sequential strict-< loop
            // keeps.  The leaf's last triangle is left to the lane, whose
            // state machine then pops its stack or finishes as usual.
            drained = drained || __ballot(ray == kDone) != 0;
            // The wave cannot take more rays now (the counter has run out, or
            // it holds its cap) and few lanes are busy.
            const uint32_t busy = (drained || cap < (uint32_t)kWave) ? (uint32_t)__popcll(__ballot(active)) : 0xFFu;
            const bool coop = (drained || busy >= cap) && busy <= L.coop_lanes;
            const uint64_t longm = coop ? __ballot(active && lend - lpos > L.coop_min) : 0ull;
            const uint32_t nl = (uint32_t)__popcll(longm);
            if (nl != 0 && nl <= L.coop_lanes) {
                // The wave splits into nl groups of gs lanes (gs = 64 / nl
                // rounded down to a power of two); group g serves the ray of
                // the g-th lane in longm.
                uint32_t np2 = 1;
                while (np2 < nl) np2 <<= 1;
                const uint32_t gs = kWave / np2, g = lane / gs, j = lane % gs;
                const bool in_group = g < nl;
                uint64_t rest = longm;
                for (uint32_t q = 0; q < g && rest; ++q) rest &= rest - 1;
                const int src = (in_group && rest) ? __ffsll((unsigned long long)rest) - 1 : lane;
                RayState c;
                c.ox = __shfl(r.ox, src);
                c.oy = __shfl(r.oy, src);
                c.oz = __shfl(r.oz, src);
                c.dx = __shfl(r.dx, src);
                c.dy = __shfl(r.dy, src);
                c.dz = __shfl(r.dz, src);
                c.ix = __shfl(r.ix, src);
                c.iy = __shfl(r.iy, src);
                c.iz = __shfl(r.iz, src);
                c.cnx = __shfl(r.cnx, src);
                c.cny = __shfl(r.cny, src);
                c.cnz = __shfl(r.cnz, src);
                c.cfx = __shfl(r.cfx, src);
                c.cfy = __shfl(r.cfy, src);
                c.cfz = __shfl(r.cfz, src);
                c.kx = __shfl(r.kx, src);
                c.ky = __shfl(r.ky, src);
                c.kz = __shfl(r.kz, src);
                c.sx = __shfl(r.sx, src);
                c.sy = __shfl(r.sy, src);
                c.sz = __shfl(r.sz, src);
                const TriRec* ctris = tris;
                if (BATCHED) {
                    const uint64_t tv = reinterpret_cast<uint64_t>(tris);
                    const uint32_t tlo = __shfl((uint32_t)tv, src), thi = __shfl((uint32_t)(tv >> 32), src);
                    ctris = reinterpret_cast<const TriRec*>(((uint64_t)thi << 32) | tlo);
                }
                float cbest = __shfl(best, src), cbu = __shfl(bu, src), cbv = __shfl(bv, src);
                uint32_t cprim = __shfl(bprim, src), cpos = __shfl(lpos, src);
                // every lane takes part in the shuffle (a source lane can lie
                // outside the grouped lanes); lanes outside a group get no work
                const uint32_t send = __shfl(lend, src);
                const uint32_t cend = in_group ? send : cpos;
                const uint64_t gmask = gs == kWave ? ~0ull : ((1ull << gs) - 1ull) << (g * gs);
                while (__ballot(in_group && cend - cpos > 1u) != 0) {
                    const bool work = in_group && cend - cpos > 1u;
                    const uint32_t cnt = work ? min(gs, cend - cpos - 1u) : 0u;
                    bool hit = false;
                    float t = kInf, u = 0.0f, v = 0.0f;
                    uint32_t prim = kNone;
                    if (j < cnt) {
                        const TriRec* tp = ctris + cpos + j;
                        const float4 ta = ldg4(tp->a), tb = ldg4(tp->b), tc = ldg4(tp->c);
                        hit = Pred::test(c, ta, tb, tc, 0.0f, t, u, v);
                        prim = __float_as_uint(ta.w);
                    }
                    const bool better = hit && t < cbest;
                    float mn = better ? t : kInf;
                    for (uint32_t o = gs >> 1; o > 0; o >>= 1) mn = fminf(mn, __shfl_xor(mn, (int)o, kWave));
                    // the lowest lane holding the minimum is the first in leaf order
                    const uint64_t win = __ballot(better && t == mn) & gmask;
                    const uint64_t missed = __ballot(j < cnt && !hit) & gmask;
                    const int w = win ? __ffsll((unsigned long long)win) - 1 : lane;
                    const float wu = __shfl(u, w), wv = __shfl(v, w);
                    const uint32_t wp = __shfl(prim, w);
                    if (work && win) {
                        cbest = mn;  // the first closest hit that beats the record, in leaf order
                        cbu = wu;
                        cbv = wv;
                        cprim = wp;
                    } else if (work && cbest == kInf && missed) {
                        cbest = 3.40282347e+38f;  // an empty record, as after a miss
                    }
                    cpos += cnt;
                }
                // each source lane takes its group's record from the group's first lane
                const uint32_t rank = (uint32_t)__popcll(longm & ((1ull << lane) - 1ull));
                const int from = ((longm >> lane) & 1ull) ? (int)(rank * gs) : lane;
                const float nb = __shfl(cbest, from), nu = __shfl(cbu, from), nv = __shfl(cbv, from);
                const uint32_t npr = __shfl(cprim, from), npos = __shfl(cpos, from);
                if ((longm >> lane) & 1ull) {
                    best = nb;
                    bu = nu;
                    bv = nv;
                    bprim = npr;
                    lpos = npos;
                }
            }
        }

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
                    stg1_stream(occ_out + slot, has ? 1u : 0u);
                } else {
                    stg4_stream(hits_out + slot, make_float4(has ? best : rtmax, has ? bu : 0.0f, has ? bv : 0.0f,
                                                      __uint_as_float(has ? bprim : kNone)));
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
constexpr int kKdStack = kKdStackEntries;

// RQ_BVH_LEAF_MODE=wave selects the wave-shared leaf tests for BVH4 launches
// (read at every launch; results do not depend on it).
static bool bvh_leaf_wave() {
    const char* e = std::getenv("RQ_BVH_LEAF_MODE");
    return e != nullptr && std::strcmp(e, "wave") == 0;
}

template <class Pred, bool ANY, bool COUNT, bool BATCHED, int WIDTH>
static const void* bvh_fn_s(int stack) {
    if (COUNT) return (const void*)dev::bvh_trace<Pred, ANY, 16, COUNT, BATCHED, WIDTH>;
    if constexpr (WIDTH == 4) {
        if (bvh_leaf_wave()) return (const void*)dev::bvh_trace<Pred, ANY, 8, false, BATCHED, 4, true>;
    }
    switch (stack) {
        case 8:
            return (const void*)dev::bvh_trace<Pred, ANY, 8, COUNT, BATCHED, WIDTH>;
        case 32:
            return (const void*)dev::bvh_trace<Pred, ANY, 32, COUNT, BATCHED, WIDTH>;
        default:
            return (const void*)dev::bvh_trace<Pred, ANY, 16, COUNT, BATCHED, WIDTH>;
    }
}

template <class Pred, bool ANY, bool COUNT, bool BATCHED>
static const void* bvh_fn_w(uint32_t width, int stack) {
    return width == 4 ? bvh_fn_s<Pred, ANY, COUNT, BATCHED, 4>(stack) : bvh_fn_s<Pred, ANY, COUNT, BATCHED, 2>(stack);
}

template <class Pred, bool ANY>
static const void* bvh_fn_v(bool count, bool batched, uint32_t width, int stack) {
    if (count) return bvh_fn_w<Pred, ANY, true, false>(width, stack);
    return batched ? bvh_fn_w<Pred, ANY, false, true>(width, stack) : bvh_fn_w<Pred, ANY, false, false>(width, stack);
}

// The bvh_trace instantiation for a launch.
static const void* bvh_fn(uint32_t predicate, TraceKind kind, bool count, bool batched, uint32_t width, int stack) {
    const bool any = kind == TraceKind::Any;
    if (predicate == RQ_PRED_WATERTIGHT)
        return any ? bvh_fn_v<dev::PredWatertight, true>(count, batched, width, stack)
                   : bvh_fn_v<dev::PredWatertight, false>(count, batched, width, stack);
    return any ? bvh_fn_v<dev::PredEncore, true>(count, batched, width, stack)
               : bvh_fn_v<dev::PredEncore, false>(count, batched, width, stack);
}

static const void* kd_fn(uint32_t predicate, TraceKind kind, bool count, bool batched = false) {
    const bool any = kind == TraceKind::Any;
#define RQ_KD(P, A)                                                                  \
    (count     ? (const void*)dev::kd_trace<P, A, kKdStack, true>                    \
     : batched ? (const void*)dev::kd_trace<P, A, kKdStack, false, true>             \
               : (const void*)dev::kd_trace<P, A, kKdStack>)
    if (predicate == RQ_PRED_WATERTIGHT)
        return any ? RQ_KD(dev::PredWatertight, true) : RQ_KD(dev::PredWatertight, false);
    return any ? RQ_KD(dev::PredEncore, true) : RQ_KD(dev::PredEncore, false);
#undef RQ_KD
}

static void launch_trace(const void* fn, const TraceLaunch& L, int grid, int block, hipStream_t s) {
    TraceLaunch arg = L;
    void* args[] = {&arg};
    RQ_HIP(hipLaunchKernel(fn, dim3(grid), dim3(block), args, 0, s));
}

static int occupancy_of(const void* fn, int block) {
    int nb = 0;
    RQ_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, block, 0));
    return nb > 0 ? nb : 1;
}

uint32_t bvh_leaf_batch() {
    static const uint32_t v = [] {
        uint32_t b = 16;
        if (const char* e = std::getenv("RQ_BVH_LEAF_BATCH")) b = (uint32_t)std::min(64, std::max(1, std::atoi(e)));
        return b;
    }();
    return v;
}

void launch_bvh(const TraceLaunch& L, uint32_t predicate, TraceKind kind, uint32_t width, int grid,
                const KernelConfig& kc, hipStream_t s) {
    launch_trace(bvh_fn(predicate, kind, false, false, width, kc.stack), L, grid, kBvhBlock, s);
}

void launch_bvh_batched(const TraceLaunch& L, uint32_t predicate, TraceKind kind, uint32_t width, int grid,
                        const KernelConfig& kc, hipStream_t s) {
    launch_trace(bvh_fn(predicate, kind, false, true, width, kc.stack), L, grid, kBvhBlock, s);
}

int occupancy_bvh_batched(uint32_t predicate, TraceKind kind, uint32_t width, int stack) {
    return occupancy_of(bvh_fn(predicate, kind, false, true, width, stack), kBvhBlock);
}

void launch_kd(const TraceLaunch& L, uint32_t predicate, TraceKind kind, int grid, hipStream_t s) {
    launch_trace(kd_fn(predicate, kind, false), L, grid, kKdBlock, s);
}

void launch_kd_batched(const TraceLaunch& L, uint32_t predicate, TraceKind kind, int grid, hipStream_t s) {
    launch_trace(kd_fn(predicate, kind, false, true), L, grid, kKdBlock, s);
}

int occupancy_kd_batched(uint32_t predicate, TraceKind kind) {
    return occupancy_of(kd_fn(predicate, kind, false, true), kKdBlock);
}

void launch_count(const TraceLaunch& L, uint32_t accel, uint32_t predicate, TraceKind kind, uint32_t width, int grid,
                  hipStream_t s) {
    launch_trace(accel == RQ_ACCEL_KD ? kd_fn(predicate, kind, true) : bvh_fn(predicate, kind, true, false, width, 16),
                 L, grid, block_threads(accel), s);
}

void launch_zero_u32(uint32_t* p, uint32_t n, hipStream_t s) {
    hipLaunchKernelGGL(dev::zero_u32, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
    RQ_HIP(hipGetLastError());
}

int occupancy_blocks_per_cu(uint32_t accel, uint32_t predicate, TraceKind kind, uint32_t width,
                            const KernelConfig& kc) {
    return occupancy_of(accel == RQ_ACCEL_KD ? kd_fn(predicate, kind, false)
                                             : bvh_fn(predicate, kind, false, false, width, kc.stack),
                        block_threads(accel));
}

uint32_t stack_entry_words(uint32_t accel, uint32_t width) {
    return accel == RQ_ACCEL_KD ? 3u : (width == 4 ? 2u : 1u);
}

}  // namespace rq
