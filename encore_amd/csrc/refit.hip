// refit.hip -- BVH update for moved vertices (Bvh::update, BVH.cpp:52-121):
// the tree keeps its topology and every box is recomputed from the new
// vertices.  ENCORE's update only grows the old boxes (bound += new box) and
// misses ancestors at the top of each path; here every child box becomes
// exactly the min/max of its triangles' vertices, as a fresh build would make
// it for the same topology, so queries after an update are exact closest hits
// on the moved geometry.
//
// refit_tris rewrites the triangle records (ENCORE predicate: v0 and the
// edges rounded once, as Vector3(v0, v1) rounds them; watertight: v0, v1, v2);
// refit_level recomputes the child boxes of one depth level of nodes.  Levels
// run deepest first, so an inner child's own boxes are final when its parent
// reads them.  Min and max follow std::min / std::max (the builder's): the
// first operand wins ties, so signed zeros land as the builder puts them.
#include <hip/hip_runtime.h>

#include "rq_internal.h"

namespace rq {
namespace dev {

__device__ __forceinline__ float lo_of(float a, float b) { return (b < a) ? b : a; }  // std::min(a, b)
__device__ __forceinline__ float hi_of(float a, float b) { return (a < b) ? b : a; }  // std::max(a, b)

struct Box {
    float lo[3], hi[3];
    __device__ void empty() {
        for (int k = 0; k < 3; ++k) {
            lo[k] = __int_as_float(0x7f800000);
            hi[k] = __int_as_float(0xff800000);
        }
    }
    __device__ void grow(const float* p) {
        for (int k = 0; k < 3; ++k) {
            lo[k] = lo_of(lo[k], p[k]);
            hi[k] = hi_of(hi[k], p[k]);
        }
    }
    __device__ void grow(const Box& b) {
        for (int k = 0; k < 3; ++k) {
            lo[k] = lo_of(lo[k], b.lo[k]);
            hi[k] = hi_of(hi[k], b.hi[k]);
        }
    }
};

__global__ void refit_tris(RefitLaunch L) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= L.nrecs) return;
    const float* v0 = L.xyz + 3 * (size_t)L.rec_vidx[3 * (size_t)i];
    const float* v1 = L.xyz + 3 * (size_t)L.rec_vidx[3 * (size_t)i + 1];
    const float* v2 = L.xyz + 3 * (size_t)L.rec_vidx[3 * (size_t)i + 2];
    TriRec& r = L.tris[i];
    for (int k = 0; k < 3; ++k) {
        r.a[k] = v0[k];
        if (L.predicate == RQ_PRED_ENCORE) {
            r.b[k] = v1[k] - v0[k];
            r.c[k] = v2[k] - v0[k];
        } else {
            r.b[k] = v1[k];
            r.c[k] = v2[k];
        }
    }
}

__device__ Box leaf_box(const RefitLaunch& L, uint32_t ref) {
    const uint32_t first = ref & kLeafMaxFirst, count = ((ref >> 27) & 15u) + 1;
    Box b;
    b.empty();
    for (uint32_t t = first; t < first + count; ++t) {
        Box tb;  // the triangle's own box first, as the builder forms it
        tb.empty();
        for (int j = 0; j < 3; ++j) tb.grow(L.xyz + 3 * (size_t)L.rec_vidx[3 * (size_t)t + j]);
        b.grow(tb);
    }
    return b;
}

__device__ Box node_box(const RefitLaunch& L, uint32_t node) {
    Box b;
    b.empty();
    if (L.width == 4) {
        const BvhNode4& n = static_cast<const BvhNode4*>(L.nodes)[node];
        for (int c = 0; c < 4; ++c) {
            if (n.ref[c] == kRefEmpty) continue;
            Box cb;
            cb.lo[0] = n.lox[c];
            cb.hi[0] = n.hix[c];
            cb.lo[1] = n.loy[c];
            cb.hi[1] = n.hiy[c];
            cb.lo[2] = n.loz[c];
            cb.hi[2] = n.hiz[c];
            b.grow(cb);
        }
    } else {
        const BvhNode& n = static_cast<const BvhNode*>(L.nodes)[node];
        for (int c = 0; c < 2; ++c) {
            Box cb;
            cb.lo[0] = n.x[2 * c];
            cb.hi[0] = n.x[2 * c + 1];
            cb.lo[1] = n.y[2 * c];
            cb.hi[1] = n.y[2 * c + 1];
            cb.lo[2] = n.z[2 * c];
            cb.hi[2] = n.z[2 * c + 1];
            b.grow(cb);
        }
    }
    return b;
}

__device__ __forceinline__ Box child_box(const RefitLaunch& L, uint32_t ref) {
    return (ref & kRefLeaf) ? leaf_box(L, ref) : node_box(L, ref);
}

// One thread per child slot of the level's nodes (4 per BVH4 node, 2 per
// BVH2 node): each computes one child box and writes its own columns.
__global__ void refit_level(RefitLaunch L) {
    const uint32_t w = L.width == 4 ? 4u : 2u;
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t i = L.begin + t / w;
    if (i >= L.end) return;
    const int c = (int)(t % w);
    const uint32_t node = L.order[i];
    if (L.width == 4) {
        BvhNode4& n = static_cast<BvhNode4*>(L.nodes)[node];
        if (n.ref[c] == kRefEmpty) return;
        const Box b = child_box(L, n.ref[c]);
        n.lox[c] = b.lo[0];
        n.hix[c] = b.hi[0];
        n.loy[c] = b.lo[1];
        n.hiy[c] = b.hi[1];
        n.loz[c] = b.lo[2];
        n.hiz[c] = b.hi[2];
    } else {
        BvhNode& n = static_cast<BvhNode*>(L.nodes)[node];
        const Box b = child_box(L, n.ref[c]);
        n.x[2 * c] = b.lo[0];
        n.x[2 * c + 1] = b.hi[0];
        n.y[2 * c] = b.lo[1];
        n.y[2 * c + 1] = b.hi[1];
        n.z[2 * c] = b.lo[2];
        n.z[2 * c + 1] = b.hi[2];
    }
}

}  // namespace dev

void launch_refit_tris(const RefitLaunch& L, hipStream_t s) {
    if (!L.nrecs) return;
    hipLaunchKernelGGL(dev::refit_tris, dim3((L.nrecs + 255) / 256), dim3(256), 0, s, L);
    RQ_HIP(hipGetLastError());
}

void launch_refit_level(const RefitLaunch& L, hipStream_t s) {
    if (L.end <= L.begin) return;
    const uint64_t threads = (uint64_t)(L.end - L.begin) * (L.width == 4 ? 4u : 2u);
    hipLaunchKernelGGL(dev::refit_level, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, s, L);
    RQ_HIP(hipGetLastError());
}

}  // namespace rq
