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
//   lb_pack       triangle records in sorted order, prim -> record map and the
//                 record vertex indices
//   lb_emit       the refs of one BvhNode per internal node whose range holds
//                 more than leaf_size triangles (a child range of at most
//                 leaf_size triangles, contiguous in sorted order, becomes one
//                 leaf), and the node's depth, counted per level
//   lb_order      nodes grouped by depth, deepest level first
//   (refit levels) the boxes, from the leaves up, by the update kernels: one
//                 launch per level instead of a bottom-up pass whose
//                 cross-wave handshakes need device-scope fences, which cost
//                 an L2 write-back each on a GPU with one L2 per XCD
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
    uint32_t* depth;     // per internal node (emitted ones)
    uint32_t* level_cnt; // nodes per depth, then the fill cursor per depth
    uint32_t* order;     // emitted nodes, deepest level first
    BvhNode* nodes;
    uint32_t leaf_size;
    uint32_t* stats;     // [0] leaf refs, [1] deepest emitted node, [2] emitted nodes, [3] lb_order overflow
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
    if (!(left & kLbLeafBit)) L.parent_in[gamma] = (uint32_t)i;
    if (!(right & kLbLeafBit)) L.parent_in[gamma + 1] = (uint32_t)i;
    if (i == 0) L.parent_in[0] = 0xFFFFFFFFu;
}

__device__ __forceinline__ uint32_t range_size(const LbLaunch& L, uint32_t c) {
    return (c & kLbLeafBit) ? 1u : L.last[c] - L.first[c] + 1;
}

constexpr uint32_t kLbMaxLevels = 256;

__device__ __forceinline__ bool emitted(const LbLaunch& L, uint32_t i) {
    return i + 1 < L.n && (i == 0 || L.last[i] - L.first[i] + 1 > L.leaf_size);
}

// Counts go through LDS: a block's nodes add to its own histogram of depths
// and the block adds that to the global one, so a few hundred global
// atomics per level replace one per node.
__global__ __launch_bounds__(256) void lb_emit(LbLaunch L) {
    __shared__ uint32_t s_cnt[kLbMaxLevels];
    __shared__ uint32_t s_leaves, s_deep, s_emitted;
    for (uint32_t d = threadIdx.x; d < kLbMaxLevels; d += blockDim.x) s_cnt[d] = 0;
    if (threadIdx.x == 0) s_leaves = s_deep = s_emitted = 0;
    __syncthreads();
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (emitted(L, i)) {
        BvhNode g;
        uint32_t leaves = 0;
        for (int side = 0; side < 2; ++side) {
            const uint32_t c = L.child[2 * i + side];
            const uint32_t sz = range_size(L, c);
            if (sz <= L.leaf_size) {
                const uint32_t f = (c & kLbLeafBit) ? (c & ~kLbLeafBit) : L.first[c];
                g.ref[side] = make_leaf_ref(f, sz);
                ++leaves;
            } else {
                g.ref[side] = c;
            }
            g.x[2 * side] = g.x[2 * side + 1] = 0.0f;  // boxes: filled by the refit levels
            g.y[2 * side] = g.y[2 * side + 1] = 0.0f;
            g.z[2 * side] = g.z[2 * side + 1] = 0.0f;
        }
        g.pad[0] = g.pad[1] = 0;
        L.nodes[i] = g;
        uint32_t depth = 0;  // hops to the root
        for (uint32_t p = L.parent_in[i]; p != 0xFFFFFFFFu; p = L.parent_in[p]) ++depth;
        L.depth[i] = depth;
        if (leaves) atomicAdd(&s_leaves, leaves);
        atomicAdd(&s_emitted, 1u);
        atomicMax(&s_deep, depth);
        if (depth < kLbMaxLevels) atomicAdd(&s_cnt[depth], 1u);
    }
    __syncthreads();
    for (uint32_t d = threadIdx.x; d < kLbMaxLevels; d += blockDim.x)
        if (s_cnt[d]) atomicAdd(&L.level_cnt[d], s_cnt[d]);
    if (threadIdx.x == 0) {
        if (s_leaves) atomicAdd(&L.stats[0], s_leaves);
        if (s_emitted) atomicAdd(&L.stats[2], s_emitted);
        atomicMax(&L.stats[1], s_deep);
    }
}

// level_cnt[kLbMaxLevels + d] holds, on entry, the first slot of depth d in
// the deepest-first order and serves as that level's fill cursor: each block
// reserves one range per depth it holds and places its nodes inside it.  The
// host has checked the level counts against the emitted nodes (at most
// n - 1), so every slot lies in order[0, n - 1); a slot beyond it is not
// written but flagged in stats[3], which the host turns into an error.
__global__ __launch_bounds__(256) void lb_order(LbLaunch L) {
    __shared__ uint32_t s_cnt[kLbMaxLevels], s_base[kLbMaxLevels];
    for (uint32_t d = threadIdx.x; d < kLbMaxLevels; d += blockDim.x) s_cnt[d] = 0;
    __syncthreads();
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool mine = emitted(L, i);
    uint32_t depth = 0, rank = 0;
    if (mine) {
        depth = L.depth[i];
        rank = atomicAdd(&s_cnt[depth], 1u);
    }
    __syncthreads();
    for (uint32_t d = threadIdx.x; d < kLbMaxLevels; d += blockDim.x)
        if (s_cnt[d]) s_base[d] = atomicAdd(&L.level_cnt[kLbMaxLevels + d], s_cnt[d]);
    __syncthreads();
    if (mine) {
        const uint32_t at = s_base[depth] + rank;
        if (at < L.n - 1)
            L.order[at] = i;
        else
            atomicOr(&L.stats[3], 1u);
    }
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

// A piece of a larger device allocation.
struct Carved {
    void* ptr;
    size_t bytes;
    template <class T>
    T* as() const {
        return static_cast<T*>(ptr);
    }
};

void build_bvh_device(rq_scene* sc, const float* xyz, uint32_t nverts, const uint32_t* idx, uint32_t ntris,
                      uint32_t leaf_size, hipStream_t s) {
    using dev::LbLaunch;
    if (leaf_size == 0) leaf_size = 4;
    if (leaf_size > kLeafMaxCount) throw Error(RQ_ERR_INVALID_ARG, "leaf_size must be 1..16");
    if (ntris > kLeafMaxFirst) throw Error(RQ_ERR_CAPACITY, "BVH supports at most 2^27-1 triangles");
    const size_t n = ntris, ni = std::max<size_t>(1, n - 1);
    // The temporaries share one allocation (a hipMalloc and a hipFree each
    // cost tens of microseconds); the arrays the scene keeps have their own.
    size_t sort_bytes = 0;
    RQ_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, (unsigned long long*)nullptr,
                                              (unsigned long long*)nullptr, (uint32_t*)nullptr, (uint32_t*)nullptr,
                                              (int)n, 0, 63, s));
    const size_t sizes[] = {3 * (size_t)nverts * 4, 3 * n * 4, 6 * n * 4, 3 * n * 4, 64, n * 8, n * 8, n * 4, n * 4,
                            2 * ni * 4, ni * 4, ni * 4, ni * 4, ni * 4, 2 * dev::kLbMaxLevels * 4, 64,
                            std::max<size_t>(sort_bytes, 16)};
    size_t total = 0;
    for (size_t b : sizes) total += (b + 255) & ~(size_t)255;
    DevBuf work(total);
    size_t at = 0;
    auto take = [&](size_t b) {
        Carved c{static_cast<char*>(work.ptr) + at, b};
        at += (b + 255) & ~(size_t)255;
        return c;
    };
    const Carved d_xyz = take(sizes[0]), d_idx = take(sizes[1]), tri_box = take(sizes[2]), cen = take(sizes[3]),
                 cb = take(sizes[4]), keys = take(sizes[5]), skeys = take(sizes[6]), ids = take(sizes[7]),
                 sids = take(sizes[8]), child = take(sizes[9]), first = take(sizes[10]), last = take(sizes[11]),
                 parent_in = take(sizes[12]), depth = take(sizes[13]), level_cnt = take(sizes[14]),
                 stats = take(sizes[15]), tmp = take(sizes[16]);
    RQ_HIP(hipMemcpyAsync(d_xyz.ptr, xyz, d_xyz.bytes, hipMemcpyHostToDevice, s));
    RQ_HIP(hipMemcpyAsync(d_idx.ptr, idx, d_idx.bytes, hipMemcpyHostToDevice, s));
    DevBuf nodes(ni * sizeof(BvhNode)), tris(n * sizeof(TriRec)), slot(n * 4), vidx(3 * n * 4), order(ni * 4);
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
    L.depth = depth.as<uint32_t>();
    L.level_cnt = level_cnt.as<uint32_t>();
    L.order = order.as<uint32_t>();
    L.nodes = nodes.as<BvhNode>();
    L.leaf_size = leaf_size;
    L.stats = stats.as<uint32_t>();
    L.tris = tris.as<TriRec>();
    L.prim_slot = slot.as<uint32_t>();
    L.rec_vidx = vidx.as<uint32_t>();
    L.predicate = sc->predicate;
    const uint32_t init[6] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0u, 0u, 0u};
    RQ_HIP(hipMemcpyAsync(cb.ptr, init, sizeof(init), hipMemcpyHostToDevice, s));
    launch_zero_u32(L.level_cnt, dev::kLbMaxLevels, s);
    // Slots of internal nodes that end up inside a leaf are never written by
    // lb_emit: keep them zero rather than stale memory.
    launch_zero_u32(nodes.as<uint32_t>(), (uint32_t)(ni * sizeof(BvhNode) / 4), s);
    launch_zero_u32(L.stats, 4, s);
    hipLaunchKernelGGL(dev::lb_boxes, dim3(grid_of(n)), dim3(256), 0, s, L);
    RQ_HIP(hipGetLastError());
    hipLaunchKernelGGL(dev::lb_morton, dim3(grid_of(n)), dim3(256), 0, s, L);
    RQ_HIP(hipGetLastError());
    // Sort (key, triangle) pairs by key; the sort is stable, so equal keys
    // keep triangle order.
    size_t tmp_bytes = tmp.bytes;
    RQ_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.ptr, tmp_bytes, L.keys, skeys.as<unsigned long long>(), L.ids,
                                              sids.as<uint32_t>(), (int)n, 0, 63, s));
    L.skeys = skeys.as<unsigned long long>();
    L.ids = sids.as<uint32_t>();
    hipLaunchKernelGGL(dev::lb_pack, dim3(grid_of(n)), dim3(256), 0, s, L);
    RQ_HIP(hipGetLastError());
    uint32_t stat[4] = {0, 0, 0, 0};
    std::vector<uint32_t> levels;  // offsets into order, deepest level first
    // Host sources of asynchronous copies stay alive until the stream
    // synchronizes at the end of the build.
    BvhNode g{};
    const uint32_t zero = 0;
    std::vector<uint32_t> start(dev::kLbMaxLevels, 0);
    if (n == 1) {
        // One triangle: a root whose two slots name the same one-triangle leaf.
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
        RQ_HIP(hipMemcpyAsync(order.ptr, &zero, 4, hipMemcpyHostToDevice, s));
        stat[0] = 1;
        levels = {0, 1};
    } else {
        hipLaunchKernelGGL(dev::lb_hierarchy, dim3(grid_of(n - 1)), dim3(256), 0, s, L);
        RQ_HIP(hipGetLastError());
        hipLaunchKernelGGL(dev::lb_emit, dim3(grid_of(n - 1)), dim3(256), 0, s, L);
        RQ_HIP(hipGetLastError());
        std::vector<uint32_t> cnt(dev::kLbMaxLevels);
        RQ_HIP(hipMemcpyAsync(stat, L.stats, 16, hipMemcpyDeviceToHost, s));
        RQ_HIP(hipMemcpyAsync(cnt.data(), L.level_cnt, cnt.size() * 4, hipMemcpyDeviceToHost, s));
        RQ_HIP(hipStreamSynchronize(s));
        if (stat[1] >= dev::kLbMaxLevels)
            throw Error(RQ_ERR_UNSUPPORTED, "linear BVH deeper than 256 levels (use the host builder)");
        // Every count the host uses below comes from the device: check that
        // they describe a binary tree of at most n - 1 inner nodes before they
        // become offsets (levels, the slots lb_order writes).
        uint64_t counted = 0;
        for (uint32_t d = 0; d < dev::kLbMaxLevels; ++d) {
            if (d > stat[1] && cnt[d] != 0)
                throw Error(RQ_ERR_DEVICE, "linear BVH build: nodes counted below the deepest level");
            counted += cnt[d];
        }
        if (stat[2] == 0 || stat[2] > n - 1 || counted != stat[2] || stat[0] != stat[2] + 1)
            throw Error(RQ_ERR_DEVICE, "linear BVH build: inconsistent node counts (" + std::to_string(stat[2]) +
                                           " inner nodes, " + std::to_string(counted) + " by level, " +
                                           std::to_string(stat[0]) + " leaves, " + std::to_string(n) +
                                           " triangles)");
        // Deepest level first: the slot where each depth starts.
        uint32_t run = 0;
        levels.push_back(0);
        for (int64_t d = stat[1]; d >= 0; --d) {
            start[d] = run;
            run += cnt[d];
            levels.push_back(run);
        }
        RQ_HIP(hipMemcpyAsync(L.level_cnt + dev::kLbMaxLevels, start.data(), start.size() * 4, hipMemcpyHostToDevice,
                              s));
        hipLaunchKernelGGL(dev::lb_order, dim3(grid_of(n - 1)), dim3(256), 0, s, L);
        RQ_HIP(hipGetLastError());
        // Boxes from the leaves up.
        RefitLaunch R{};
        R.xyz = L.xyz;
        R.rec_vidx = L.rec_vidx;
        R.tris = L.tris;
        R.nrecs = ntris;
        R.predicate = sc->predicate;
        R.nodes = nodes.ptr;
        R.width = 2;
        R.order = order.as<uint32_t>();
        for (size_t l = 0; l + 1 < levels.size(); ++l) {
            R.begin = levels[l];
            R.end = levels[l + 1];
            launch_refit_level(R, s);
        }
        RQ_HIP(hipMemcpyAsync(&stat[3], L.stats + 3, 4, hipMemcpyDeviceToHost, s));
    }
    RQ_HIP(hipStreamSynchronize(s));
    if (stat[3] != 0) throw Error(RQ_ERR_DEVICE, "linear BVH build: a node was placed outside the level order");
    sc->width = 2;
    sc->nodes = std::move(nodes);
    sc->tris = std::move(tris);
    sc->prim_slot = std::move(slot);
    sc->rec_vidx = std::move(vidx);
    sc->refit_order = std::move(order);
    sc->refit_levels = std::move(levels);
    sc->stats = BuildStats{};
    sc->stats.nnodes = (uint32_t)ni;  // internal nodes inside leaves stay unused
    sc->stats.nleaves = stat[0];
    sc->stats.nrefs = ntris;
    sc->stats.max_depth = stat[1] + 1;
    sc->stats.stack_bound = stat[1] + 3;
}

}  // namespace rq
