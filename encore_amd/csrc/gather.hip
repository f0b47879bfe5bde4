// gather.hip -- final gather (PhotonMapperCPU::CalculateIndirectIllumination,
// PhotonMapperCPU.cpp:1049-1181, final-gather branch) around the traversal and
// k-NN kernels:
//   fg_rays    one ray per (shading point, sample direction), origin on the
//              surface with no offset (PhotonMapperCPU.cpp:1124)
//   (closest-hit traversal of the rays)
//   fg_attr    hit point and hit normal of every sample hit, as
//              TrianglePrim::intersect computes them (TrianglePrim.cpp:236-239)
//   (radiance estimates of the indirect and direct photon maps at them)
//   fg_reduce  per shading point, in sample order: the sum of the estimates
//              times the hit triangle's diffuse colour, times PI / samples,
//              and the harmonic mean hit distance (PhotonMapperCPU.cpp:1126-1167)
// The reduction runs one lane per shading point over its samples in order, so
// it rounds exactly as the reference's loop does.
#include <hip/hip_runtime.h>

#include "rq_internal.h"

namespace rq {
namespace dev {

constexpr uint32_t kFgNone = 0xFFFFFFFFu;
constexpr float kFgMissPos = 1e30f;  // far from every photon: the k-NN gathers nothing

__device__ __forceinline__ float fg_flush(float x) { return (x > -0.001f && x < 0.001f) ? 0.0f : x; }

__global__ void fg_rays(FgLaunch L) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= L.nrays) return;
    const uint32_t p = i / L.nper;
    rq_ray r;
    r.ox = L.origin[3 * (size_t)p];
    r.oy = L.origin[3 * (size_t)p + 1];
    r.oz = L.origin[3 * (size_t)p + 2];
    r.tmin = 0.0f;
    r.dx = L.dirs[3 * (size_t)i];
    r.dy = L.dirs[3 * (size_t)i + 1];
    r.dz = L.dirs[3 * (size_t)i + 2];
    r.tmax = __int_as_float(0x7f800000);
    L.rays[i] = r;
}

__global__ void fg_attr(FgLaunch L) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= L.nrays) return;
    const rq_hit h = L.hits[i];
    const uint32_t prim = h.prim;
    float* qp = L.qpos + 3 * (size_t)i;
    float* qn = L.qnrm + 3 * (size_t)i;
    if (prim == kFgNone) {
        qp[0] = qp[1] = qp[2] = kFgMissPos;
        qn[0] = qn[1] = qn[2] = 0.0f;
        return;
    }
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
    // hitPoint = v0 + (v1-v0)*alpha + (v2-v0)*beta, left to right.
    qp[0] = (v0x + e1x * h.u) + e2x * h.v;
    qp[1] = (v0y + e1y * h.u) + e2y * h.v;
    qp[2] = (v0z + e1z * h.u) + e2z * h.v;
    // Normalize(Cross(e1, e2)): Cross flushes each component, Normalize
    // divides by the correctly rounded magnitude and flushes again.
    const float cx = fg_flush(e1y * e2z - e1z * e2y);
    const float cy = fg_flush(-(e1x * e2z - e1z * e2x));
    const float cz = fg_flush(e1x * e2y - e1y * e2x);
    const float m = __builtin_sqrtf((cx * cx + cy * cy) + cz * cz);
    qn[0] = fg_flush(cx / m);
    qn[1] = fg_flush(cy / m);
    qn[2] = fg_flush(cz / m);
}

__global__ void fg_reduce(FgLaunch L) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= L.npoints) return;
    const float ox = L.origin[3 * (size_t)p], oy = L.origin[3 * (size_t)p + 1], oz = L.origin[3 * (size_t)p + 2];
    float R = 0.0f, G = 0.0f, B = 0.0f, ri = 0.0f;
    uint32_t nhit = 0;
    for (uint32_t j = 0; j < L.nper; ++j) {
        const size_t i = (size_t)p * L.nper + j;
        const uint32_t prim = L.hits[i].prim;
        if (prim == kFgNone) continue;
        const float dr = L.diffuse[3 * (size_t)prim], dg = L.diffuse[3 * (size_t)prim + 1],
                    db = L.diffuse[3 * (size_t)prim + 2];
        // irrValue = ind * diffuse; irrValue += dir * diffuse
        float ir = L.rgb_ind[3 * i] * dr, ig = L.rgb_ind[3 * i + 1] * dg, ib = L.rgb_ind[3 * i + 2] * db;
        if (L.rgb_dir) {
            ir += L.rgb_dir[3 * i] * dr;
            ig += L.rgb_dir[3 * i + 1] * dg;
            ib += L.rgb_dir[3 * i + 2] * db;
        }
        R += ir;
        G += ig;
        B += ib;
        // Ri += 1 / |hit.hitPoint - sampleHit.hitPoint|
        const float vx = ox - L.qpos[3 * i], vy = oy - L.qpos[3 * i + 1], vz = oz - L.qpos[3 * i + 2];
        ri += 1.0f / __builtin_sqrtf((vx * vx + vy * vy) + vz * vz);
        ++nhit;
    }
    const float scale = 3.14159f / (float)L.nper;  // PI / sampleRays.size() with PI = 3.14159f
    L.out_rgb[3 * (size_t)p] = R * scale;
    L.out_rgb[3 * (size_t)p + 1] = G * scale;
    L.out_rgb[3 * (size_t)p + 2] = B * scale;
    if (L.out_ri) L.out_ri[p] = (ri != 0.0f) ? (float)nhit / ri : 0.0f;
}

}  // namespace dev

static unsigned blocks_for(uint32_t n) { return (n + 255) / 256; }

void launch_fg_rays(const FgLaunch& L, hipStream_t s) {
    hipLaunchKernelGGL(dev::fg_rays, dim3(blocks_for(L.nrays)), dim3(256), 0, s, L);
    RQ_HIP(hipGetLastError());
}

void launch_fg_attr(const FgLaunch& L, hipStream_t s) {
    hipLaunchKernelGGL(dev::fg_attr, dim3(blocks_for(L.nrays)), dim3(256), 0, s, L);
    RQ_HIP(hipGetLastError());
}

void launch_fg_reduce(const FgLaunch& L, hipStream_t s) {
    hipLaunchKernelGGL(dev::fg_reduce, dim3(blocks_for(L.npoints)), dim3(256), 0, s, L);
    RQ_HIP(hipGetLastError());
}

}  // namespace rq
