// photon_grid.hip -- cell grid over a photon map for radiance queries (gfx950).
//
// A radiance estimate only ever selects photons whose squared distance is
// below fl(maxDist^2) (PhotonMap.cpp:89).  Bucketing the photons into cubic
// cells whose edge is that reach bounds a query to at most 4 x 4 x 4 cells,
// each a contiguous range of photons, so a wave reads its candidates with
// whole-wave loads instead of walking the balanced tree one node per step.
// The grid answers the same question as the tree (which photons pass the
// class and distance tests); the selecting kernel then keeps the k nearest.
//
// Build, on the device, for one squared radius:
//   pg_keys    Morton key of each photon's cell (~0 for photons no query can
//              select: a position that is not finite or no normal class);
//   sort       (key, slot) pairs by key, rocPRIM radix sort;
//   pg_starts  the first photon of each cell inserts the cell into the table;
//   pg_ends    the last photon of each cell writes its count;
//   pg_gather  positions and classes in grid order.
// The table is a dense array over the photons' bounding box when that takes
// at most kGridDenseMaxBytes (one load per cell, no hashing or probing),
// otherwise an open-addressing hash table keyed by the cell's Morton code.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstdlib>

#include "rq_internal.h"

namespace rq {
namespace dev {

struct PgBuild {
    const PhotonHot* hot;
    uint32_t n, nvalid;
    unsigned long long* keys;
    uint32_t* ids;
    const unsigned long long* skeys;
    const uint32_t* sids;
    float4* pos;
    uint32_t* slot;
    uint4* cells;
    uint32_t cell_mask;
    uint2* dense;  // null: the hash table
    uint32_t dims[3];
    double origin[3];
    double inv_cell;
};

// Dense index of photon p's cell (within the box by construction; clamped
// so that a write can never leave the table).
__device__ __forceinline__ uint64_t pg_dense_index(const PgBuild& B, const PhotonHot& p) {
    const float x[3] = {p.x, p.y, p.z};
    uint64_t c[3];
    for (int a = 0; a < 3; ++a) {
        const int64_t v = grid_coord((double)x[a], B.origin[a], B.inv_cell);
        c[a] = (uint64_t)min(max(v, (int64_t)0), (int64_t)B.dims[a] - 1);
    }
    return c[0] + (uint64_t)B.dims[0] * (c[1] + (uint64_t)B.dims[1] * c[2]);
}

__global__ void pg_clear_dense(uint2* dense, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dense[i] = make_uint2(0u, 0u);
}

constexpr unsigned long long kEmptyKey = ~0ull;

__global__ void pg_clear(uint4* cells, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) cells[i] = make_uint4(0xFFFFFFFFu, 0xFFFFFFFFu, 0u, 0u);
}

__global__ void pg_keys(PgBuild B) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B.n) return;
    const PhotonHot p = B.hot[i];
    unsigned long long key = kEmptyKey;
    if (isfinite(p.x) && isfinite(p.y) && isfinite(p.z) && (p.meta >> 2) != kClassNever) {
        // within [0, dims) by construction: the bounds come from these photons
        const int64_t cx = grid_coord((double)p.x, B.origin[0], B.inv_cell);
        const int64_t cy = grid_coord((double)p.y, B.origin[1], B.inv_cell);
        const int64_t cz = grid_coord((double)p.z, B.origin[2], B.inv_cell);
        key = cell_key((uint64_t)cx, (uint64_t)cy, (uint64_t)cz);
    }
    B.keys[i] = key;
    B.ids[i] = i;
}

__global__ void pg_starts(PgBuild B) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B.nvalid) return;
    const unsigned long long key = B.skeys[i];
    if (i > 0 && B.skeys[i - 1] == key) return;
    if (B.dense) {
        B.dense[pg_dense_index(B, B.hot[B.sids[i]])].x = i;
        return;
    }
    uint32_t h = cell_hash(key, B.cell_mask);
    // The table has at least twice as many entries as cells: a free entry
    // is always found.
    for (uint32_t probe = 0; probe <= B.cell_mask; ++probe) {
        unsigned long long* kp = reinterpret_cast<unsigned long long*>(B.cells + h);
        const unsigned long long prev = atomicCAS(kp, kEmptyKey, key);
        if (prev == kEmptyKey) {
            reinterpret_cast<uint32_t*>(B.cells + h)[2] = i;
            return;
        }
        h = (h + 1) & B.cell_mask;
    }
}

__global__ void pg_ends(PgBuild B) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B.nvalid) return;
    const unsigned long long key = B.skeys[i];
    if (i + 1 < B.nvalid && B.skeys[i + 1] == key) return;
    if (B.dense) {
        uint2* e = B.dense + pg_dense_index(B, B.hot[B.sids[i]]);
        e->y = i + 1 - e->x;
        return;
    }
    uint32_t h = cell_hash(key, B.cell_mask);
    for (uint32_t probe = 0; probe <= B.cell_mask; ++probe) {
        const uint4 e = B.cells[h];
        if (((unsigned long long)e.y << 32 | e.x) == key) {
            reinterpret_cast<uint32_t*>(B.cells + h)[3] = i + 1 - e.z;
            return;
        }
        h = (h + 1) & B.cell_mask;
    }
}

__global__ void pg_gather(PgBuild B) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B.nvalid) return;
    const uint32_t s = B.sids[i];
    const PhotonHot p = B.hot[s];
    B.pos[i] = make_float4(p.x, p.y, p.z, __uint_as_float(p.meta >> 2));
    B.slot[i] = s;
}

}  // namespace dev

static unsigned blocks_for(uint64_t n, unsigned block = 256) { return (unsigned)((n + block - 1) / block); }

// The caller holds p->grid_mu until its queries are queued.
const PhotonGrid* photon_grid_for(rq_photons* p, float r2, hipStream_t s) {
    if (!(r2 > 0.0f) || !std::isfinite(r2) || p->nvalid == 0) return nullptr;
    uint32_t bits;
    std::memcpy(&bits, &r2, 4);
    if (p->grid && p->grid->r2_bits == bits) return p->grid.get();
    // Every photon with float d2 < r2 lies within sqrt(r2) (1 + 3 2^-24) of
    // the query; the reach adds a margin far beyond that.
    const double reach = std::sqrt((double)r2) * (1.0 + 1.0 / 1024.0);
    const double cell = reach, inv_cell = 1.0 / cell;
    uint32_t dims[3];
    uint32_t widest = 0;
    for (int a = 0; a < 3; ++a) {
        const int64_t top = grid_coord((double)p->bmax[a], (double)p->bmin[a], inv_cell);
        if (top < 0 || top >= (1 << 21) - 1) return nullptr;
        dims[a] = (uint32_t)top + 1;
        widest = std::max(widest, dims[a]);
    }
    if (widest < 4) return nullptr;
    // Queries widen their reach by |q| 2^-40 for rounding; near the map that
    // must stay below a quarter cell.
    double far = 0.0;
    for (int a = 0; a < 3; ++a) far = std::max({far, std::fabs((double)p->bmin[a]), std::fabs((double)p->bmax[a])});
    if ((far + 2.0 * reach) * 0x1p-40 >= 0.25 * cell) return nullptr;
    const uint32_t n = p->n, nv = p->nvalid;
    const uint64_t ncells = (uint64_t)dims[0] * dims[1] * dims[2];
    const char* force_hash = std::getenv("RQ_KNN_GRID_HASH");  // A/B checks of the hash table
    const bool dense = ncells <= kGridDenseMaxBytes / sizeof(uint2) && !(force_hash && force_hash[0] == '1');
    uint32_t cap = dense ? 1 : 64;
    while (!dense && cap < 2 * (uint64_t)nv) cap <<= 1;
    auto grid = std::make_unique<PhotonGrid>();
    grid->r2_bits = bits;
    grid->pos = DevBuf((size_t)nv * sizeof(float4));
    grid->slot = DevBuf((size_t)nv * sizeof(uint32_t));
    grid->cells = DevBuf((size_t)cap * sizeof(uint4));
    if (dense) grid->dense = DevBuf(ncells * sizeof(uint2));
    DevBuf keys((size_t)n * 8), skeys((size_t)n * 8), ids((size_t)n * 4), sids((size_t)n * 4);
    dev::PgBuild B{};
    B.hot = p->hot.as<PhotonHot>();
    B.n = n;
    B.nvalid = nv;
    B.keys = keys.as<unsigned long long>();
    B.ids = ids.as<uint32_t>();
    B.skeys = skeys.as<unsigned long long>();
    B.sids = sids.as<uint32_t>();
    B.pos = grid->pos.as<float4>();
    B.slot = grid->slot.as<uint32_t>();
    B.cells = grid->cells.as<uint4>();
    B.cell_mask = cap - 1;
    B.dense = dense ? grid->dense.as<uint2>() : nullptr;
    for (int a = 0; a < 3; ++a) {
        B.dims[a] = dims[a];
        B.origin[a] = (double)p->bmin[a];
    }
    B.inv_cell = inv_cell;
    hipLaunchKernelGGL(dev::pg_clear, dim3(blocks_for(cap)), dim3(256), 0, s, B.cells, cap);
    RQ_HIP(hipGetLastError());
    if (dense) {
        hipLaunchKernelGGL(dev::pg_clear_dense, dim3(blocks_for(ncells)), dim3(256), 0, s, B.dense, ncells);
        RQ_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(dev::pg_keys, dim3(blocks_for(n)), dim3(256), 0, s, B);
    RQ_HIP(hipGetLastError());
    // Stable sort: photons of one cell keep their slot order.
    size_t tmp_bytes = 0;
    RQ_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, B.keys, skeys.as<unsigned long long>(), B.ids,
                                              sids.as<uint32_t>(), (int)n, 0, 64, s));
    DevBuf tmp(std::max<size_t>(tmp_bytes, 16));
    RQ_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.ptr, tmp_bytes, B.keys, skeys.as<unsigned long long>(), B.ids,
                                              sids.as<uint32_t>(), (int)n, 0, 64, s));
    hipLaunchKernelGGL(dev::pg_starts, dim3(blocks_for(nv)), dim3(256), 0, s, B);
    RQ_HIP(hipGetLastError());
    hipLaunchKernelGGL(dev::pg_ends, dim3(blocks_for(nv)), dim3(256), 0, s, B);
    RQ_HIP(hipGetLastError());
    hipLaunchKernelGGL(dev::pg_gather, dim3(blocks_for(nv)), dim3(256), 0, s, B);
    RQ_HIP(hipGetLastError());
    // The temporaries are freed on return, and other streams may query the
    // grid next: finish the build here.
    RQ_HIP(hipStreamSynchronize(s));
    PhotonGridView& v = grid->view;
    v.pos = grid->pos.as<float4>();
    v.slot = grid->slot.as<uint32_t>();
    v.cells = grid->cells.as<uint4>();
    v.cell_mask = cap - 1;
    v.dense = B.dense;
    for (int a = 0; a < 3; ++a) {
        v.dims[a] = dims[a];
        v.origin[a] = (double)p->bmin[a];
    }
    v.cell = cell;
    v.inv_cell = inv_cell;
    v.reach = reach;
    // A grid for another radius may still be read by queued queries.
    if (p->grid) RQ_HIP(hipDeviceSynchronize());
    p->grid = std::move(grid);
    return p->grid.get();
}

}  // namespace rq
