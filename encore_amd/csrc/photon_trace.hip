// photon_trace.hip -- photon emission and tracing on the device
// (PhotonMapperCPU::EmitPhotons / TracePhoton / ReflectPhotonDiffusely /
// TransmitPhoton / EmitShadowPhoton / TransmitRay, PhotonMapperCPU.cpp:503-835).
//
// A batch of photon paths runs as a wavefront: ph_emit makes every path's
// first ray; then, until no ray is left, the closest-hit traversal runs over
// the live rays and ph_shade decides each one's interaction and appends the
// rays that follow (the next bounce, a shadow-photon ray, the ray inside a
// refracting object).  Stored photons land in fixed event slots of their
// path, so the order of live rays never matters.  ph_compact then walks the
// paths in emission order and fills each map until it is full, as the
// reference's sequential loop does, and finds the path after which the
// reference would have stopped emitting.
//
// Draws: ENCORE calls rand() % 10000 (randNumber, Common.cpp:17-21); here
// each draw is SplitMix64 of (seed, path, slot), reduced the same way, so a
// photon's fate does not depend on the batch it runs in.  Directions follow
// randDirectionN / randDirectionSuperN (Common.cpp:98-182) without calls to
// acos, sin and cos: cos(acos(sqrt(e1))) = sqrt(e1), sin(acos(sqrt(e1))) =
// sqrt(1 - e1), the cosine and sine of 2*PI*e2 come from a table over the
// 10000 possible e2, and the rotation matrix of GetRotationMatrix
// (TAffineMatrix.h:160-200) is applied in factored form R^T X R v, with the
// axis left unnormalised as the reference leaves it.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "rq_internal.h"

namespace rq {
namespace dev {

constexpr uint32_t kPhNone = 0xFFFFFFFFu;
constexpr uint32_t kPathMask = (1u << kPhPathBits) - 1;
constexpr uint32_t kKindPath = 0, kKindShadow = 1, kKindTransmit = 2;
constexpr uint8_t kTagGlobal = 1, kTagDirect = 2, kTagCaustic = 3, kTagShadow = 4;

struct V3 {
    float x, y, z;
};

__device__ __forceinline__ float flush(float v) { return (v > -0.001f && v < 0.001f) ? 0.0f : v; }
__device__ __forceinline__ V3 add(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 mul(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ float dot3(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
// Cross with the component flush (Vector3.h:148-162).
__device__ __forceinline__ V3 cross_flush(V3 a, V3 b) {
    return V3{flush(a.y * b.z - a.z * b.y), flush(-1.0f * (a.x * b.z - a.z * b.x)), flush(a.x * b.y - a.y * b.x)};
}
// Vector3::Normalize (Vector3.cpp:105-124).
__device__ __forceinline__ V3 normalize_flush(V3 v) {
    const float m = __builtin_sqrtf((v.x * v.x + v.y * v.y) + v.z * v.z);
    return V3{flush(v.x / m), flush(v.y / m), flush(v.z / m)};
}
__device__ __forceinline__ bool finite3(V3 v) {
    return __builtin_isfinite(v.x) && __builtin_isfinite(v.y) && __builtin_isfinite(v.z);
}

__device__ __forceinline__ uint32_t ph_rand(uint64_t seed, uint64_t path, uint32_t slot) {
    uint64_t z = seed + (path * 64ull + slot + 1ull) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (uint32_t)(z >> 33);
}

// randNumber(lo, hi) = (hi - lo) * ((rand() % 10000) / (float)10000) + lo.
__device__ __forceinline__ float rand_number(float lo, float hi, uint32_t r) {
    return (hi - lo) * ((float)(int)(r % 10000u) / 10000.0f) + lo;
}

// GetRotationMatrix(r, acos(c)) * v (TAffineMatrix.h:160-200): rows r, s, t
// with s perpendicular to r through r's smallest component, t = r x s; the
// matrix is R^T X(angle) R.
__device__ V3 rotate_about(V3 r, float c, V3 v) {
    const float ax = fabsf(r.x), ay = fabsf(r.y), az = fabsf(r.z);
    V3 s;
    if (ax <= ay && ax <= az)
        s = V3{0.0f, -r.z, r.y};
    else if (ay <= ax && ay <= az)
        s = V3{-r.z, 0.0f, r.x};
    else
        s = V3{-r.y, r.x, 0.0f};
    s = normalize_flush(s);
    const V3 t = normalize_flush(cross_flush(r, s));
    const float sa = __builtin_sqrtf((1.0f - c) * (1.0f + c));
    const float w0 = dot3(r, v), w1 = dot3(s, v), w2 = dot3(t, v);
    const float u1 = c * w1 - sa * w2, u2 = sa * w1 + c * w2;
    return V3{(r.x * w0 + s.x * u1) + t.x * u2, (r.y * w0 + s.y * u1) + t.y * u2, (r.z * w0 + s.z * u1) + t.z * u2};
}

// randDirectionN (e1_lo = 0.01) and randDirectionSuperN (e1_lo = 0.001):
// theta = acos(sqrt(e1)), phi = 2*PI*e2 about +z, turned onto n.
__device__ V3 rand_dir_about(V3 n, float e1_lo, uint32_t r1, uint32_t r2, const float* phi_cs) {
    const float e1 = rand_number(e1_lo, 1.0f, r1);
    const uint32_t k2 = r2 % 10000u;
    const float cp = phi_cs[2 * k2], sp = phi_cs[2 * k2 + 1];
    const float ct = __builtin_sqrtf(e1), st = __builtin_sqrtf(1.0f - e1);
    const V3 d{st * cp, st * sp, ct};
    n = normalize_flush(n);
    if (n.x == 0.0f && n.y == 0.0f && n.z == 1.0f) return d;
    if (0.0f == -n.x && 0.0f == -n.y && 1.0f == -n.z) return V3{-d.x, -d.y, -d.z};
    const V3 z{0.0f, 0.0f, 1.0f};
    return normalize_flush(rotate_about(cross_flush(z, n), dot3(n, z), d));
}

// RectangleLight::EmitPhoton's origin: Sampler::GetSamplePointsForRect for
// one fully jittered cell (Sampler.cpp:49-95, 154-218).
__device__ V3 light_point(const PhLight& l, uint32_t r1, uint32_t r2) {
    const float jx = rand_number(-0.5f, 0.5f, r1) * 1.0f;
    const float jz = rand_number(-0.5f, 0.5f, r2) * 1.0f;
    V3 p{0.5f + jx, 0.0f, 0.5f + jz};
    p = V3{-0.5f + p.x, 0.0f + p.y, -0.5f + p.z};
    p.x *= l.length / 1.0f;
    p.z *= l.width / 1.0f;
    const V3 n = normalize_flush(V3{l.normal[0], l.normal[1], l.normal[2]});
    if (n.x == 0.0f && n.y == 1.0f && n.z == 0.0f) {
        // already aligned with the sampler's +y plane
    } else if (0.0f == -n.x && 1.0f == -n.y && 0.0f == -n.z) {
        p = add(p, mul(p, -2.0f));
    } else {
        const V3 up{0.0f, 1.0f, 0.0f};
        const float c = dot3(up, n);
        if (c != 1.0f) p = rotate_about(normalize_flush(cross_flush(up, n)), c, p);
    }
    p = add(V3{l.center[0], l.center[1], l.center[2]}, p);
    return add(V3{l.normal[0] * 0.01f, l.normal[1] * 0.01f, l.normal[2] * 0.01f}, p);
}

// Hit point and hit normal as TrianglePrim::intersect computes them
// (TrianglePrim.cpp:236-239).
__device__ __forceinline__ void hit_attr(const PhLaunch& L, uint32_t prim, float u, float v, V3& P, V3& N) {
    const TriRec& t = L.tris[L.prim_slot[prim]];
    const float v0x = t.a[0], v0y = t.a[1], v0z = t.a[2];
    float e1x = t.b[0], e1y = t.b[1], e1z = t.b[2];
    float e2x = t.c[0], e2y = t.c[1], e2z = t.c[2];
    if (L.predicate == RQ_PRED_WATERTIGHT) {  // records hold v1, v2
        e1x -= v0x;
        e1y -= v0y;
        e1z -= v0z;
        e2x -= v0x;
        e2y -= v0y;
        e2z -= v0z;
    }
    P = V3{(v0x + e1x * u) + e2x * v, (v0y + e1y * u) + e2y * v, (v0z + e1z * u) + e2z * v};
    N = normalize_flush(cross_flush(V3{e1x, e1y, e1z}, V3{e2x, e2y, e2z}));
}

__device__ __forceinline__ rq_ray make_ray(V3 o, V3 d) {
    rq_ray r;
    r.ox = o.x;
    r.oy = o.y;
    r.oz = o.z;
    r.tmin = 0.0f;
    r.dx = d.x;
    r.dy = d.y;
    r.dz = d.z;
    r.tmax = __int_as_float(0x7f800000);
    return r;
}
__device__ __forceinline__ bool ray_finite(const rq_ray& r) {
    return finite3(V3{r.ox, r.oy, r.oz}) && finite3(V3{r.dx, r.dy, r.dz});
}

// TransmitRay (PhotonMapperCPU.cpp:800-835) for a surface that refracts:
// Snell with airIndex 0.99, origin 0.01 along the new direction.  A ray under
// total internal reflection comes out NaN, as in the reference.
__device__ rq_ray refract(V3 d, V3 P, V3 N, bool entering, const PhMaterial& m) {
    const float air = 0.99f;
    float c1, c2;
    if (entering) {
        c1 = air;
        c2 = m.index;
    } else {
        c2 = air;
        c1 = m.index;
    }
    const float cr = (c1 == 0.0f) ? (float)((double)c2 / 0.0001) : (c2 / c1);
    const float nd = dot3(N, d);
    const float th2 = __builtin_sqrtf(1.0f - (cr * cr) * (1.0f - (nd * nd)));
    const float k = cr * nd - th2;
    const V3 T = normalize_flush(add(mul(d, cr), mul(N, k)));
    return make_ray(add(P, mul(T, 0.01f)), T);
}

__device__ __forceinline__ uint32_t make_tag(uint32_t lp, uint32_t kind, uint32_t bounce, uint32_t spec) {
    return lp | (kind << 22) | (bounce << 24) | (spec << 29);
}

// Up to two rays a lane appends (held in registers: no indexed arrays).
struct Emit {
    rq_ray r0, r1;
    uint32_t t0 = 0, t1 = 0;
    uint32_t n = 0;
    __device__ void push(const rq_ray& ray, uint32_t tag) {
        if (n == 0) {
            r0 = ray;
            t0 = tag;
        } else {
            r1 = ray;
            t1 = tag;
        }
        ++n;
    }
};

constexpr uint32_t kPhWaves = kPhBlock / 64;

// Appends every thread's new rays to the output queue with one atomic per
// block (one counter shared by a million rays serialises at the L2 when
// every wave takes its own).  Every thread of the block must call it; s_w
// holds kPhWaves + 1 words of LDS.
__device__ __forceinline__ void block_append(const PhLaunch& L, const Emit& e, uint32_t* s_w) {
    const uint32_t lane = __lane_id(), wid = threadIdx.x / 64;
    uint32_t incl = e.n;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t v = __shfl_up(incl, off, 64);
        if (lane >= (uint32_t)off) incl += v;
    }
    if (lane == 63) s_w[wid] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (uint32_t w = 0; w < kPhWaves; ++w) {
            const uint32_t t = s_w[w];
            s_w[w] = run;
            run += t;
        }
        s_w[kPhWaves] = run ? atomicAdd(L.out_count, run) : 0;
    }
    __syncthreads();
    const uint32_t base = s_w[kPhWaves] + s_w[wid] + (incl - e.n);
    if (e.n > 0) {
        L.out_rays[base] = e.r0;
        L.out_tags[base] = e.t0;
    }
    if (e.n > 1) {
        L.out_rays[base + 1] = e.r1;
        L.out_tags[base + 1] = e.t1;
    }
    __syncthreads();  // s_w is reused by the next call
}

__device__ __forceinline__ void store_event(const PhLaunch& L, uint32_t lp, uint32_t slot, uint8_t tag, V3 P, V3 N) {
    const size_t e = (size_t)slot * L.stride + lp;
    L.ev_pos[3 * e] = P.x;
    L.ev_pos[3 * e + 1] = P.y;
    L.ev_pos[3 * e + 2] = P.z;
    L.ev_nrm[3 * e] = N.x;
    L.ev_nrm[3 * e + 1] = N.y;
    L.ev_nrm[3 * e + 2] = N.z;
    L.ev_tag[e] = tag;
}

__global__ __launch_bounds__(kPhBlock) void ph_emit(PhLaunch L) {
    __shared__ uint32_t s_w[kPhWaves + 1];
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    Emit e;
    if (p < L.npaths) {
        const uint64_t gp = L.base + p;
        const uint32_t li = ph_rand(L.seed, gp, 0) % L.nlights;
        const PhLight& l = L.lights[li];
        const V3 o = light_point(l, ph_rand(L.seed, gp, 1), ph_rand(L.seed, gp, 2));
        const V3 d = rand_dir_about(V3{l.normal[0], l.normal[1], l.normal[2]}, 0.01f, ph_rand(L.seed, gp, 3),
                                    ph_rand(L.seed, gp, 4), L.phi_cs);
        // TracePhoton's first ray starts 0.01 along the direction.
        const rq_ray r = make_ray(add(o, mul(d, 0.01f)), d);
        L.path_ray[p] = r;
        L.path_dir[3 * (size_t)p] = d.x;
        L.path_dir[3 * (size_t)p + 1] = d.y;
        L.path_dir[3 * (size_t)p + 2] = d.z;
        L.path_light[p] = li;
        for (uint32_t s = 0; s < L.nslots; ++s) L.ev_tag[(size_t)s * L.stride + p] = 0;
        if (ray_finite(r)) e.push(r, make_tag(p, kKindPath, 0, 0));
    }
    block_append(L, e, s_w);
}

// The photon continues with `r` at bounce nb (TracePhoton's loop condition
// bounces <= MaxBounces); a ray with a NaN component finds no hit in the
// reference's traversal, so it ends the path without being traced.
__device__ __forceinline__ void continue_path(const PhLaunch& L, Emit& e, uint32_t lp, const rq_ray& r, uint32_t nb,
                                              uint32_t spec) {
    if (nb <= L.max_bounces && ray_finite(r)) e.push(r, make_tag(lp, kKindPath, nb, spec));
}

__device__ void shade_one(const PhLaunch& L, uint32_t i, Emit& e) {
    const uint32_t tag = L.in_tags[i];
    const uint32_t lp = tag & kPathMask;
    const uint32_t kind = (tag >> 22) & 3u, b = (tag >> 24) & 31u, spec = (tag >> 29) & 1u;
    const rq_ray ray = L.in_rays[i];
    const rq_hit h = L.hits[i];
    const V3 d{ray.dx, ray.dy, ray.dz};
    if (h.prim == kPhNone) {
        // TransmitPhoton: the refracted ray hit nothing, the photon keeps its
        // ray and is now specular.  Other rays end here.
        if (kind == kKindTransmit) continue_path(L, e, lp, L.path_ray[lp], b + 1, 1);
        return;
    }
    V3 P, N;
    hit_attr(L, h.prim, h.u, h.v, P, N);
    const bool entering = dot3(d, N) < 0.0f;  // TrianglePrim.cpp:242
    if (kind == kKindShadow) {
        // EmitShadowPhoton: pass through surfaces being left, store at the
        // first one entered.  The bounce field counts the surfaces passed; the
        // chain gives up at kPhShadowPasses, where a phantom hit behind the
        // ray (Cross()'s flush) would keep the reference's loop going for ever.
        if (!entering) {
            if (b + 1 < kPhShadowPasses) e.push(make_ray(add(P, mul(d, 0.01f)), d), make_tag(lp, kKindShadow, b + 1, 0));
        } else {
            store_event(L, lp, 1, kTagDirect | kTagShadow, P, N);
        }
        return;
    }
    const PhMaterial m = L.mats[L.prim_mat[h.prim]];
    if (kind == kKindTransmit) {
        rq_ray next = L.path_ray[lp];
        if (!entering) {  // leaving the object: refract again (or keep the inner ray)
            next = (m.transparency < 0.6f) ? ray : refract(d, P, N, false, m);
            L.path_ray[lp] = next;
        }
        continue_path(L, e, lp, next, b + 1, 1);
        return;
    }
    // A path ray (TracePhoton, PhotonMapperCPU.cpp:606-655).  With
    // pS + pD + pT > 1 nothing happens at the hit and the same ray is traced
    // again until the bounces run out, so the path ends here.
    if (!(m.ptotal <= 1.0f)) return;
    const uint64_t gp = L.base + lp;
    const float E = rand_number(0.0f, 1.0f, ph_rand(L.seed, gp, 5 + 3 * b));
    if (E < m.pd || !(E < m.ptotal)) {
        // ReflectPhotonDiffusely: store (caustic after a refraction, global
        // after the first bounce, else direct with a shadow photon).
        const uint32_t slot = b == 0 ? 0 : b + 1;
        if (spec) {
            store_event(L, lp, slot, kTagCaustic, P, N);
        } else if (b > 0) {
            store_event(L, lp, slot, kTagGlobal, P, N);
        } else {
            store_event(L, lp, slot, kTagDirect, P, N);
            e.push(make_ray(add(P, mul(d, 2.0f)), d), make_tag(lp, kKindShadow, 0, 0));
        }
        if (E < m.pd) {  // reflect; otherwise absorbed after storing
            const V3 nd = rand_dir_about(N, 0.001f, ph_rand(L.seed, gp, 6 + 3 * b), ph_rand(L.seed, gp, 7 + 3 * b),
                                         L.phi_cs);
            const rq_ray nr = make_ray(add(P, mul(nd, 0.001f)), nd);
            L.path_ray[lp] = nr;
            continue_path(L, e, lp, nr, b + 1, 0);
        }
        return;
    }
    // TransmitPhoton (PhotonMapperCPU.cpp:747-773).
    if (m.transparency < 0.6f) {  // TransmitRay refuses: same ray, same state
        continue_path(L, e, lp, ray, b + 1, spec);
        return;
    }
    const rq_ray tr = refract(d, P, N, entering, m);
    if (!ray_finite(tr)) {  // no hit for a NaN ray: the photon keeps its ray
        continue_path(L, e, lp, ray, b + 1, 1);
        return;
    }
    e.push(tr, make_tag(lp, kKindTransmit, b, 1));
}

// Grid-stride over the live rays, whose count is read on the device, so the
// host can queue iterations without waiting for each one.
__global__ __launch_bounds__(kPhBlock) void ph_shade(PhLaunch L) {
    __shared__ uint32_t s_w[kPhWaves + 1];
    const uint32_t nin = *L.nin_dev;
    for (uint32_t base = blockIdx.x * kPhBlock; base < nin; base += gridDim.x * kPhBlock) {
        Emit e;
        const uint32_t i = base + threadIdx.x;
        if (i < nin) shade_one(L, i, e);
        block_append(L, e, s_w);
    }
}

// ----- compaction in emission order ----------------------------------------
__device__ __forceinline__ uint32_t field(unsigned long long c, uint32_t m) {
    return (uint32_t)((c >> (21 * m)) & 0x1FFFFFull);
}

// Exclusive scan of one value per thread over a kPhScanBlock block.
__device__ unsigned long long block_excl_scan(unsigned long long v) {
    __shared__ unsigned long long wsum[kPhScanBlock / 64];
    const uint32_t lane = __lane_id(), wid = threadIdx.x / 64;
    unsigned long long incl = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const unsigned long long u = __shfl_up(incl, off, 64);
        if (lane >= (uint32_t)off) incl += u;
    }
    if (lane == 63) wsum[wid] = incl;
    __syncthreads();
    unsigned long long pre = 0;
    for (uint32_t w = 0; w < wid; ++w) pre += wsum[w];
    __syncthreads();
    return pre + incl - v;
}

__global__ void ph_count(PhLaunch L) {
    const uint32_t p = blockIdx.x * kPhScanBlock + threadIdx.x;
    unsigned long long c = 0;
    if (p < L.npaths) {
        for (uint32_t s = 0; s < L.nslots; ++s) {
            const uint32_t t = L.ev_tag[(size_t)s * L.stride + p] & 3u;
            if (t) c += 1ull << (21 * (t - 1));
        }
        L.cnt[p] = c;
    }
    const unsigned long long ex = block_excl_scan(c);
    if (threadIdx.x == kPhScanBlock - 1) {
        const unsigned long long tot = ex + c;
        for (uint32_t m = 0; m < 3; ++m) L.boff[3 * blockIdx.x + m] = field(tot, m);
    }
}

// One block of 1024 threads: block totals -> exclusive block offsets, and the
// batch totals in ctl[5..7].
__global__ void ph_scan_blocks(PhLaunch L) {
    __shared__ uint32_t s[3][1024];
    const uint32_t tid = threadIdx.x;
    const uint32_t chunk = (L.nblocks + 1023) / 1024;
    const uint32_t b0 = min(tid * chunk, L.nblocks), b1 = min(b0 + chunk, L.nblocks);
    uint32_t sum[3] = {0, 0, 0};
    for (uint32_t b = b0; b < b1; ++b)
        for (uint32_t m = 0; m < 3; ++m) sum[m] += L.boff[3 * b + m];
    for (uint32_t m = 0; m < 3; ++m) s[m][tid] = sum[m];
    __syncthreads();
    for (uint32_t off = 1; off < 1024; off <<= 1) {
        uint32_t v[3];
        for (uint32_t m = 0; m < 3; ++m) v[m] = tid >= off ? s[m][tid - off] : 0;
        __syncthreads();
        for (uint32_t m = 0; m < 3; ++m) s[m][tid] += v[m];
        __syncthreads();
    }
    uint32_t run[3];
    for (uint32_t m = 0; m < 3; ++m) run[m] = s[m][tid] - sum[m];
    for (uint32_t b = b0; b < b1; ++b)
        for (uint32_t m = 0; m < 3; ++m) {
            const uint32_t v = L.boff[3 * b + m];
            L.boff[3 * b + m] = run[m];
            run[m] += v;
        }
    if (tid == 1023)
        for (uint32_t m = 0; m < 3; ++m) L.ctl[5 + m] = s[m][1023];
}

// The path whose photon fills the global map (EmitPhotons' first loop ends
// after it) and the one that fills the caustic map (second loop).
__global__ void ph_stops(PhLaunch L) {
    const uint32_t p = blockIdx.x * kPhScanBlock + threadIdx.x;
    const unsigned long long c = p < L.npaths ? L.cnt[p] : 0;
    const unsigned long long ex = block_excl_scan(c);
    if (p >= L.npaths) return;
    const uint32_t maps[2] = {RQ_MAP_GLOBAL, RQ_MAP_CAUSTIC};
    for (uint32_t k = 0; k < 2; ++k) {
        const uint32_t m = maps[k];
        if (L.rem[m] == 0) continue;
        const uint32_t before = L.boff[3 * blockIdx.x + m] + field(ex, m);
        if (before < L.rem[m] && before + field(c, m) >= L.rem[m]) L.ctl[k] = p + 1;
    }
}

__global__ void ph_scatter(PhLaunch L) {
    const uint32_t p = blockIdx.x * kPhScanBlock + threadIdx.x;
    const unsigned long long c = p < L.npaths ? L.cnt[p] : 0;
    const unsigned long long ex = block_excl_scan(c);
    // Paths the reference emits: up to the later of the two stops.
    const uint32_t need_g = L.rem[RQ_MAP_GLOBAL] == 0 ? 0 : (L.ctl[0] ? L.ctl[0] : L.npaths);
    const uint32_t need_c = L.rem[RQ_MAP_CAUSTIC] == 0 ? 0 : (L.ctl[1] ? L.ctl[1] : L.npaths);
    const uint32_t keep = max(need_g, need_c);
    uint32_t wrote[3] = {0, 0, 0};
    if (p < L.npaths && p < keep) {
        uint32_t k[3];
        for (uint32_t m = 0; m < 3; ++m) k[m] = L.boff[3 * blockIdx.x + m] + field(ex, m);
        const uint32_t li = L.path_light[p];
        const PhLight& l = L.lights[li];
        for (uint32_t s = 0; s < L.nslots; ++s) {
            const size_t e = (size_t)s * L.stride + p;
            const uint8_t t = L.ev_tag[e];
            if (!(t & 3u)) continue;
            const uint32_t m = (t & 3u) - 1;
            const uint32_t r = k[m]++;
            if (r >= L.rem[m]) continue;  // the map is full (AddPhoton drops it)
            const size_t o = (size_t)L.acc[m] + r;
            const PhOut& out = L.out[m];
            const bool sh = (t & kTagShadow) != 0;
            for (uint32_t a = 0; a < 3; ++a) {
                out.pos[3 * o + a] = L.ev_pos[3 * e + a];
                out.nrm[3 * o + a] = L.ev_nrm[3 * e + a];
                const float pw = sh ? -l.power[a] : l.power[a];
                out.pow[3 * o + a] = pw * L.scale[m];  // ScalePhotons
                if (out.dir) out.dir[3 * o + a] = L.path_dir[3 * (size_t)p + a];
            }
            if (out.shadow) out.shadow[o] = sh ? 1 : 0;
            if (out.path) out.path[o] = (uint32_t)(L.base + p);
            ++wrote[m];
        }
    }
    for (uint32_t m = 0; m < 3; ++m) {
        const uint32_t tot = (uint32_t)block_excl_scan(wrote[m]) + wrote[m];
        if (threadIdx.x == kPhScanBlock - 1 && tot) atomicAdd(&L.ctl[2 + m], tot);
    }
}

}  // namespace dev

void launch_ph_emit(const PhLaunch& L, hipStream_t s) {
    hipLaunchKernelGGL(dev::ph_emit, dim3((L.npaths + kPhBlock - 1) / kPhBlock), dim3(kPhBlock), 0, s, L);
    RQ_HIP(hipGetLastError());
}

void launch_ph_shade(const PhLaunch& L, int max_blocks, hipStream_t s) {
    const uint32_t need = (L.nin_max + kPhBlock - 1) / kPhBlock;
    const uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>(need, (uint32_t)max_blocks));
    hipLaunchKernelGGL(dev::ph_shade, dim3(grid), dim3(kPhBlock), 0, s, L);
    RQ_HIP(hipGetLastError());
}

void launch_ph_compact(const PhLaunch& L, hipStream_t s) {
    launch_zero_u32(L.ctl, 8, s);
    const dim3 grid(L.nblocks), block(kPhScanBlock);
    hipLaunchKernelGGL(dev::ph_count, grid, block, 0, s, L);
    RQ_HIP(hipGetLastError());
    hipLaunchKernelGGL(dev::ph_scan_blocks, dim3(1), dim3(1024), 0, s, L);
    RQ_HIP(hipGetLastError());
    hipLaunchKernelGGL(dev::ph_stops, grid, block, 0, s, L);
    RQ_HIP(hipGetLastError());
    hipLaunchKernelGGL(dev::ph_scatter, grid, block, 0, s, L);
    RQ_HIP(hipGetLastError());
}

}  // namespace rq
