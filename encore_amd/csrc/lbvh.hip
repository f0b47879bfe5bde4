// lbvh.hip -- BVH build on the GPU for scenes admitted on the fly: a linear
// BVH (Karras, "Maximizing Parallelism in the Construction of BVHs, Octrees,
// and k-d Trees", HPG 2012) over 63-bit Morton codes of the triangle
// centroids, written in the engine's BVH2 node layout so the same traversal
// kernels serve it.
//
//   lb_boxes      triangle boxes and centroids (as the host builder forms
//                 them: min/max of the three vertices, centre 0.5 (lo + hi))
//                 and the centroids' bounds
//   lb_morton     21 bits per axis of the centroid within those bounds,
//                 interleaved; the sort key carries the triangle id as value
//   (radix sort of the keys)
//   lb_hierarchy  one thread per internal node finds its key range and split
//                 from the longest common prefixes (equal keys broken by
//                 index), so every node is built independently
//   lb_bottom_up  boxes from the leaves up: the second of two children to
//                 arrive at a parent forms its box and carries on
//   lb_emit       one BvhNode per internal node whose range holds more than
//                 leaf_size triangles; a child range of at most leaf_size
//                 triangles (contiguous in sorted order) becomes one leaf
//   lb_pack       triangle records in sorted order, prim -> record map and the
//                 record vertex indices (for later refits)
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <vector>

#include "rq_internal.h"

namespace rq {
namespace dev {

constexpr uint32_t kLbLeafBit = 0x80000000u;

struct LbLaunch {
    const float* xyz;
    const uint32_t* idx;
    uint32_t n;  // triangles
    float* tri_box;      // 6 per triangle: lo xyz, hi xyz
    float* cen;          // 3 per triangle
    uint32_t* cbounds;   // 6: order-preserving bits of centroid min xyz, max xyz
    unsigned long long* keys;
    uint32_t* ids;       // sorted: triangle id at each position
    const unsigned long long* skeys;
    uint32_t* child;     // 2 per internal node: index | kLbLeafBit for a leaf
    uint32_t* first;     // per internal node: key range [first, last]
    uint32_t* last;
    uint32_t* parent_in; // per internal node (root: ~0)
    uint32_t* parent_lf; // per leaf
    uint32_t* arrive;    // per internal node: children done
    float* node_box;     // 6 per internal node
    BvhNode* nodes;
    uint32_t leaf_size;
    uint32_t* stats;     // [0] leaf refs, [1] deepest emitted node
    TriRec* tris;
    uint32_t* prim_slot;
    uint32_t* rec_vidx;
    uint32_t predicate;
};

__device__ __forceinline__ uint32_t ord(float f) {  // order-preserving bits
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float unord(uint32_t u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}
__device__ __forceinline__ float lo_of(float a, float b) { return (b < a) ? b : a; }  // std::min(a, b)
__device__ __forceinline__ float hi_of(float a, float b) { return (a < b) ? b : a; }  // std::max(a, b)

__global__ void lb_boxes(LbLaunch L) {
    __shared__ uint32_t s_b[6];
    if (threadIdx.x < 6) s_b[threadIdx.x] = threadIdx.x < 3 ? 0xFFFFFFFFu : 0u;
    __syncthreads();
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < L.n) {
        float lo[3] = {__int_as_float(0x7f800000), __int_as_float(0x7f800000), __int_as_float(0x7f800000)};
        float hi[3] = {__int_as_float(0xff800000), __int_as_float(0xff800000), __int_as_float(0xff800000)};
        for (int j = 0; j < 3; ++j) {
            const float* v = L.xyz + 3 * (size_t)L.idx[3 * (size_t)t + j];
            for (int k = 0; k < 3; ++k) {
                lo[k] = lo_of(lo[k], v[k]);
                hi[k] = hi_of(hi[k], v[k]);
            }
        }
        for (int k = 0; k < 3; ++k) {
            L.tri_box[6 * (size_t)t + k] = lo[k];
            L.tri_box[6 * (size_t)t + 3 + k] = hi[k];
            const float c = 0.5f * (lo[k] + hi[k]);
            L.cen[3 * (size_t)t + k] = c;
            atomicMin(&s_b[k], ord(c));
            atomicMax(&s_b[3 + k], ord(c));
        }
    }
    __syncthreads();
    if (threadIdx.x < 3) atomicMin(&L.cbounds[threadIdx.x], s_b[threadIdx.x]);
    else if (threadIdx.x < 6) atomicMax(&L.cbounds[threadIdx.x], s_b[threadIdx.x]);
}

__device__ __forceinline__ unsigned long long spread21(uint32_t v) {
    unsigned long long x = v & 0x1fffffu;
    x = (x | x << 32) & 0x1f00000000ffffull;
    x = (x | x << 16) & 0x1f0000ff0000ffull;
    x = (x | x << 8) & 0x100f00f00f00f00full;
    x = (x | x << 4) & 0x10c30c30c30c30c3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}

__global__ void lb_morton(LbLaunch L) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= L.n) return;
    uint32_t q[3];
    for (int k = 0; k < 3; ++k) {
        const float lo = unord(L.cbounds[k]), hi = unord(L.cbounds[3 + k]);
        const float ext = hi - lo;
        const float u = ext > 0.0f ? (L.cen[3 * (size_t)t + k] - lo) / ext : 0.0f;
        q[k] = (uint32_t)fminf(fmaxf(u * 2097151.0f, 0.0f), 2097151.0f);
    }
    L.keys[t] = (spread21(q[0]) << 2) | (spread21(q[1]) << 1) | spread21(q[2]);
    L.ids[t] = t;
}

// Length of the common prefix of keys i and j (equal keys: 64 + that of the
// indices); -1 outside [0, n).
__device__ __forceinline__ int delta(const LbLaunch& L, int i, int j) {
    if (j < 0 || j >= (int)L.n) return -1;
    const unsigned long long a = L.skeys[i], b = L.skeys[j];
    if (a == b) return 64 + __clz((uint32_t)i ^ (uint32_t)j);
    return __clzll(a ^ b);
}

__global__ void lb_hierarchy(LbLaunch L) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int)L.n - 1) return;
    const int d = (delta(L, i, i + 1) - delta(L, i, i - 1)) >= 0 ? 1 : -1;
    const int dmin = delta(L, i, i - d);
    int lmax = 2;
    while (delta(L, i, i + lmax * d) > dmin) lmax *= 2;
    int l = 0;
    for (int t = lmax / 2; t >= 1; t /= 2)
        if (delta(L, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = delta(L, i, j);
    int s = 0, t = l;
    do {
        t = (t + 1) / 2;
        if (delta(L, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    const int gamma = i + s * d + min(d, 0);
    const int f = min(i, j), la = max(i, j);
    const uint32_t left = (f == gamma) ? ((uint32_t)gamma | kLbLeafBit) : (uint32_t)gamma;
    const uint32_t right = (la == gamma + 1) ? ((uint32_t)(gamma + 1) | kLbLeafBit) : (uint32_t)(gamma + 1);
    L.child[2 * i] = left;
    L.child[2 * i + 1] = right;
    L.first[i] = (uint32_t)f;
    L.last[i] = (uint32_t)la;
    if (left & kLbLeafBit) L.parent_lf[gamma] = (uint32_t)i; else L.parent_in[gamma] = (uint32_t)i;
    if (right & kLbLeafBit) L.parent_lf[gamma + 1] = (uint32_t)i; else L.parent_in[gamma + 1] = (uint32_t)i;
    if (i == 0) L.parent_in[0] = 0xFFFFFFFFu;
}

// Reads six floats written by another wave: agent-scope loads, so the values
// come from L2 and not from a stale line of this CU's cache.
__device__ __forceinline__ void load_box(const float* p, float* b) {
    for (int k = 0; k < 6; ++k)
        b[k] = __uint_as_float(__hip_atomic_load(reinterpret_cast<const uint32_t*>(p) + k, __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_AGENT));
}

__device__ __forceinline__ void child_box(const LbLaunch& L, uint32_t c, float* b) {
    if (c & kLbLeafBit)
        load_box(L.tri_box + 6 * (size_t)L.ids[c & ~kLbLeafBit], b);
    else
        load_box(L.node_box + 6 * (size_t)c, b);
}

__global__ void lb_bottom_up(LbLaunch L) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= L.n) return;
    uint32_t p = L.parent_lf[s];
    while (p != 0xFFFFFFFFu) {
        __threadfence();
        if (atomicAdd(&L.arrive[p], 1u) == 0) return;  // the sibling is not done yet: it carries on
        __threadfence();
        float a[6], b[6];
        child_box(L, L.child[2 * p], a);
        child_box(L, L.child[2 * p + 1], b);
        float* o = L.node_box + 6 * (size_t)p;
        for (int k = 0; k < 3; ++k) {
            o[k] = lo_of(a[k], b[k]);
            o[3 + k] = hi_of(a[3 + k], b[3 + k]);
        }
        p = L.parent_in[p];
    }
}

__device__ __forceinline__ uint32_t range_size(const LbLaunch& L, uint32_t c) {
    return (c & kLbLeafBit) ? 1u : L.last[c] - L.first[c] + 1;
}

__global__ void lb_emit(LbLaunch L) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i + 1 >= L.n) return;
    if (L.last[i] - L.first[i] + 1 <= L.leaf_size && i != 0) return;  // inside a leaf of its parent
    BvhNode g;
    for (int side = 0; side < 2; ++side) {
        const uint32_t c = L.child[2 * i + side];
        float b[6];
        child_box(L, c, b);
        g.x[2 * side] = b[0];
        g.x[2 * side + 1] = b[3];
        g.y[2 * side] = b[1];
        g.y[2 * side + 1] = b[4];
        g.z[2 * side] = b[2];
        g.z[2 * side + 1] = b[5];
        const uint32_t sz = range_size(L, c);
        if (sz <= L.leaf_size) {
            const uint32_t f = (c & kLbLeafBit) ? (c & ~kLbLeafBit) : L.first[c];
            g.ref[side] = make_leaf_ref(f, sz);
            atomicAdd(&L.stats[0], 1u);
        } else {
            g.ref[side] = c;
        }
    }
    g.pad[0] = g.pad[1] = 0;
    L.nodes[i] = g;
    // depth of this node: hops to the root
    uint32_t depth = 0;
    for (uint32_t p = L.parent_in[i]; p != 0xFFFFFFFFu; p = L.parent_in[p]) ++depth;
    atomicMax(&L.stats[1], depth);
}

__global__ void lb_pack(LbLaunch L) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= L.n) return;
    const uint32_t t = L.ids[s];
    const uint32_t i0 = L.idx[3 * (size_t)t], i1 = L.idx[3 * (size_t)t + 1], i2 = L.idx[3 * (size_t)t + 2];
    const float* v0 = L.xyz + 3 * (size_t)i0;
    const float* v1 = L.xyz + 3 * (size_t)i1;
    const float* v2 = L.xyz + 3 * (size_t)i2;
    TriRec r;
    for (int k = 0; k < 3; ++k) {
        r.a[k] = v0[k];
        if (L.predicate == RQ_PRED_ENCORE) {
            r.b[k] = v1[k] - v0[k];  // edges rounded once, as Vector3(v0, v1)
            r.c[k] = v2[k] - v0[k];
        } else {
            r.b[k] = v1[k];
            r.c[k] = v2[k];
        }
    }
    r.a[3] = __uint_as_float(t);
    r.b[3] = 0.0f;
    r.c[3] = 0.0f;
    L.tris[s] = r;
    L.prim_slot[t] = s;
    L.rec_vidx[3 * (size_t)s] = i0;
    L.rec_vidx[3 * (size_t)s + 1] = i1;
    L.rec_vidx[3 * (size_t)s + 2] = i2;
}

}  // namespace dev

static unsigned grid_of(uint64_t n, unsigned block = 256) { return (unsigned)((n + block - 1) / block); }

void build_bvh_device(rq_scene* sc, const float* xyz, uint32_t nverts, const uint32_t* idx, uint32_t ntris,
                      uint32_t leaf_size, hipStream_t s) {
    using dev::LbLaunch;
    if (leaf_size == 0) leaf_size = 4;
    if (leaf_size > kLeafMaxCount) throw Error(RQ_ERR_INVALID_ARG, "leaf_size must be 1..16");
    if (ntris > kLeafMaxFirst) throw Error(RQ_ERR_CAPACITY, "BVH supports at most 2^27-1 triangles");
    const size_t n = ntris, ni = std::max<size_t>(1, n - 1);
    DevBuf d_xyz(3 * (size_t)nverts * 4), d_idx(3 * n * 4);
    RQ_HIP(hipMemcpyAsync(d_xyz.ptr, xyz, d_xyz.bytes, hipMemcpyHostToDevice, s));
    RQ_HIP(hipMemcpyAsync(d_idx.ptr, idx, d_idx.bytes, hipMemcpyHostToDevice, s));
    DevBuf tri_box(6 * n * 4), cen(3 * n * 4), cb(64), keys(n * 8), skeys(n * 8), ids(n * 4), sids(n * 4),
        child(2 * ni * 4), first(ni * 4), last(ni * 4), parent_in(ni * 4), parent_lf(n * 4), arrive(ni * 4),
        node_box(6 * ni * 4), stats(64);
    DevBuf nodes(ni * sizeof(BvhNode)), tris(n * sizeof(TriRec)), slot(n * 4), vidx(3 * n * 4);
    LbLaunch L{};
    L.xyz = d_xyz.as<float>();
    L.idx = d_idx.as<uint32_t>();
    L.n = ntris;
    L.tri_box = tri_box.as<float>();
    L.cen = cen.as<float>();
    L.cbounds = cb.as<uint32_t>();
    L.keys = keys.as<unsigned long long>();
    L.ids = ids.as<uint32_t>();
    L.child = child.as<uint32_t>();
    L.first = first.as<uint32_t>();
    L.last = last.as<uint32_t>();
    L.parent_in = parent_in.as<uint32_t>();
    L.parent_lf = parent_lf.as<uint32_t>();
    L.arrive = arrive.as<uint32_t>();
    L.node_box = node_box.as<float>();
    L.nodes = nodes.as<BvhNode>();
    L.leaf_size = leaf_size;
    L.stats = stats.as<uint32_t>();
    L.tris = tris.as<TriRec>();
    L.prim_slot = slot.as<uint32_t>();
    L.rec_vidx = vidx.as<uint32_t>();
    L.predicate = sc->predicate;
    const uint32_t init[6] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0u, 0u, 0u};
    RQ_HIP(hipMemcpyAsync(cb.ptr, init, sizeof(init), hipMemcpyHostToDevice, s));
    launch_zero_u32(L.arrive, (uint32_t)ni, s);
    // Slots of internal nodes that end up inside a leaf are never written by
    // lb_emit: keep them zero rather than stale memory.
    launch_zero_u32(nodes.as<uint32_t>(), (uint32_t)(ni * sizeof(BvhNode) / 4), s);
    launch_zero_u32(L.stats, 2, s);
    hipLaunchKernelGGL(dev::lb_boxes, dim3(grid_of(n)), dim3(256), 0, s, L);
    RQ_HIP(hipGetLastError());
    hipLaunchKernelGGL(dev::lb_morton, dim3(grid_of(n)), dim3(256), 0, s, L);
    RQ_HIP(hipGetLastError());
    // Sort (key, triangle) pairs by key; the sort is stable, so equal keys
    // keep triangle order.
    size_t tmp_bytes = 0;
    RQ_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, L.keys, skeys.as<unsigned long long>(), L.ids,
                                              sids.as<uint32_t>(), (int)n, 0, 63, s));
    DevBuf tmp(std::max<size_t>(tmp_bytes, 16));
    RQ_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.ptr, tmp_bytes, L.keys, skeys.as<unsigned long long>(), L.ids,
                                              sids.as<uint32_t>(), (int)n, 0, 63, s));
    L.skeys = skeys.as<unsigned long long>();
    L.ids = sids.as<uint32_t>();
    uint32_t stat[2] = {0, 0};
    if (n == 1) {
        // One triangle: a root whose two slots name the same one-triangle leaf.
        BvhNode g{};
        float b[6];
        RQ_HIP(hipMemcpyAsync(b, L.tri_box, 24, hipMemcpyDeviceToHost, s));
        RQ_HIP(hipStreamSynchronize(s));
        for (int side = 0; side < 2; ++side) {
            g.x[2 * side] = b[0];
            g.x[2 * side + 1] = b[3];
            g.y[2 * side] = b[1];
            g.y[2 * side + 1] = b[4];
            g.z[2 * side] = b[2];
            g.z[2 * side + 1] = b[5];
            g.ref[side] = make_leaf_ref(0, 1);
        }
        RQ_HIP(hipMemcpyAsync(nodes.ptr, &g, sizeof(g), hipMemcpyHostToDevice, s));
        stat[0] = 1;
    } else {
        hipLaunchKernelGGL(dev::lb_hierarchy, dim3(grid_of(n - 1)), dim3(256), 0, s, L);
        RQ_HIP(hipGetLastError());
        hipLaunchKernelGGL(dev::lb_bottom_up, dim3(grid_of(n)), dim3(256), 0, s, L);
        RQ_HIP(hipGetLastError());
        hipLaunchKernelGGL(dev::lb_emit, dim3(grid_of(n - 1)), dim3(256), 0, s, L);
        RQ_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(dev::lb_pack, dim3(grid_of(n)), dim3(256), 0, s, L);
    RQ_HIP(hipGetLastError());
    if (n > 1) RQ_HIP(hipMemcpyAsync(stat, L.stats, 8, hipMemcpyDeviceToHost, s));
    RQ_HIP(hipStreamSynchronize(s));
    sc->width = 2;
    sc->nodes = std::move(nodes);
    sc->tris = std::move(tris);
    sc->prim_slot = std::move(slot);
    sc->rec_vidx = std::move(vidx);
    sc->stats = BuildStats{};
    sc->stats.nnodes = (uint32_t)ni;  // internal nodes inside leaves stay unused
    sc->stats.nleaves = stat[0];
    sc->stats.nrefs = ntris;
    sc->stats.max_depth = stat[1] + 1;
    sc->stats.stack_bound = stat[1] + 3;
}

}  // namespace rq
