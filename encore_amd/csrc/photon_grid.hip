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
//   pg_subkeys Morton key of each photon's sub-cell (16 per axis) within its
//              cell, and a stable radix sort of the slots by it;
//   pg_keys    Morton key of each photon's cell (~0 for photons no query can
//              select: a position that is not finite or no normal class);
//   sort       (key, slot) pairs by key, stable, rocPRIM radix sort: a cell's
//              photons keep their sub-cell order;
//   pg_starts  the first photon of each cell inserts the cell into the table;
//   pg_ends    the last photon of each cell writes its count;
//   pg_gather  positions and classes in grid order;
//   blocks     (dense table only) each cell's photons cut into blocks of up
//              to 64 from its first photon, numbered by a scan, with their
//              bounding boxes (pg_bflags, scan, pg_blocks).
// The table is a dense array over the photons' bounding box when that takes
// at most kGridDenseMaxBytes (one load per cell, no hashing or probing),
// otherwise an open-addressing hash table keyed by the cell's Morton code.
//
// A query whose cells hold at most 256 blocks, at most 128 of them within the
// radius, ranks those blocks instead of the cells: along the sub-cell curve a
// block covers a small part of its cell, so once the k-th distance has shrunk
// below a block's box the block is skipped.
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
    uint32_t* subkeys;
    const uint32_t* ids1;  // slots in sub-cell order
    unsigned long long* keys;
    uint32_t* ids;
    const unsigned long long* skeys;
    const uint32_t* sids;
    uint32_t* flags;       // 1 where a block starts (grid order)
    const uint32_t* bidx;  // exclusive scan of flags: the block of a start
    uint32_t* dense_block;
    float4* blocks;
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

__device__ __forceinline__ uint32_t spread4(uint32_t v) {
    uint32_t r = 0;
    for (int b = 0; b < 4; ++b) r |= ((v >> b) & 1u) << (3 * b);
    return r;
}

// Sub-cell (16 per axis) of each photon within its cell, as a 12-bit Morton
// key; the slots are then sorted by it.
__global__ void pg_subkeys(PgBuild B) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B.n) return;
    const PhotonHot p = B.hot[i];
    uint32_t key = 0;
    if (isfinite(p.x) && isfinite(p.y) && isfinite(p.z)) {
        const float x[3] = {p.x, p.y, p.z};
        uint32_t f[3];
        for (int a = 0; a < 3; ++a) {
            const double t = ((double)x[a] - B.origin[a]) * B.inv_cell;
            const double c = floor(t);
            f[a] = (uint32_t)min(max((int64_t)floor((t - c) * 16.0), (int64_t)0), (int64_t)15);
        }
        key = (spread4(f[0]) << 2) | (spread4(f[1]) << 1) | spread4(f[2]);
    }
    B.subkeys[i] = key;
    B.ids[i] = i;
}

__global__ void pg_keys(PgBuild B) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= B.n) return;
    const uint32_t i = B.ids1[j];
    const PhotonHot p = B.hot[i];
    unsigned long long key = kEmptyKey;
    if (isfinite(p.x) && isfinite(p.y) && isfinite(p.z) && (p.meta >> 2) != kClassNever) {
        // within [0, dims) by construction: the bounds come from these photons
        const int64_t cx = grid_coord((double)p.x, B.origin[0], B.inv_cell);
        const int64_t cy = grid_coord((double)p.y, B.origin[1], B.inv_cell);
        const int64_t cz = grid_coord((double)p.z, B.origin[2], B.inv_cell);
        key = cell_key((uint64_t)cx, (uint64_t)cy, (uint64_t)cz);
    }
    B.keys[j] = key;
    B.ids[j] = i;
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

// Block starts: a cell's first photon and every 64th after it.
__global__ void pg_bflags(PgBuild B) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B.nvalid) return;
    const uint2 e = B.dense[pg_dense_index(B, B.hot[B.sids[i]])];
    B.flags[i] = ((i - e.x) & 63u) == 0u ? 1u : 0u;
}

// One thread per block start: the block's bounding box and range, and the
// cell's first block.
__global__ void pg_blocks(PgBuild B) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B.nvalid || B.flags[i] == 0u) return;
    const uint64_t c = pg_dense_index(B, B.hot[B.sids[i]]);
    const uint2 e = B.dense[c];
    const uint32_t off = i - e.x;
    const uint32_t len = min(64u, e.y - off);
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t k = 0; k < len; ++k) {
        const float4 q = B.pos[i + k];
        lo[0] = fminf(lo[0], q.x);
        lo[1] = fminf(lo[1], q.y);
        lo[2] = fminf(lo[2], q.z);
        hi[0] = fmaxf(hi[0], q.x);
        hi[1] = fmaxf(hi[1], q.y);
        hi[2] = fmaxf(hi[2], q.z);
    }
    const uint32_t b = B.bidx[i];
    B.blocks[2 * (size_t)b] = make_float4(lo[0], lo[1], lo[2], __uint_as_float(i));
    B.blocks[2 * (size_t)b + 1] = make_float4(hi[0], hi[1], hi[2], __uint_as_float(len));
    if (off == 0) B.dense_block[c] = b;
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
    DevBuf subkeys((size_t)n * 4), ssub((size_t)n * 4), ids1((size_t)n * 4);
    dev::PgBuild B{};
    B.hot = p->hot.as<PhotonHot>();
    B.n = n;
    B.nvalid = nv;
    B.subkeys = subkeys.as<uint32_t>();
    B.ids1 = ids1.as<uint32_t>();
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
    // Slots by sub-cell, then (stable) by cell: a cell's photons follow the
    // sub-cell curve.
    hipLaunchKernelGGL(dev::pg_subkeys, dim3(blocks_for(n)), dim3(256), 0, s, B);
    RQ_HIP(hipGetLastError());
    size_t tmp_sub = 0, tmp_bytes = 0;
    RQ_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_sub, B.subkeys, ssub.as<uint32_t>(), B.ids,
                                              ids1.as<uint32_t>(), (int)n, 0, 12, s));
    RQ_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, B.keys, skeys.as<unsigned long long>(), B.ids,
                                              sids.as<uint32_t>(), (int)n, 0, 64, s));
    DevBuf tmp(std::max<size_t>(std::max(tmp_bytes, tmp_sub), 16));
    RQ_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.ptr, tmp_sub, B.subkeys, ssub.as<uint32_t>(), B.ids,
                                              ids1.as<uint32_t>(), (int)n, 0, 12, s));
    hipLaunchKernelGGL(dev::pg_keys, dim3(blocks_for(n)), dim3(256), 0, s, B);
    RQ_HIP(hipGetLastError());
    RQ_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.ptr, tmp_bytes, B.keys, skeys.as<unsigned long long>(), B.ids,
                                              sids.as<uint32_t>(), (int)n, 0, 64, s));
    hipLaunchKernelGGL(dev::pg_starts, dim3(blocks_for(nv)), dim3(256), 0, s, B);
    RQ_HIP(hipGetLastError());
    hipLaunchKernelGGL(dev::pg_ends, dim3(blocks_for(nv)), dim3(256), 0, s, B);
    RQ_HIP(hipGetLastError());
    hipLaunchKernelGGL(dev::pg_gather, dim3(blocks_for(nv)), dim3(256), 0, s, B);
    RQ_HIP(hipGetLastError());
    if (dense) {
        DevBuf flags((size_t)nv * 4), bidx((size_t)nv * 4);
        B.flags = flags.as<uint32_t>();
        B.bidx = bidx.as<uint32_t>();
        hipLaunchKernelGGL(dev::pg_bflags, dim3(blocks_for(nv)), dim3(256), 0, s, B);
        RQ_HIP(hipGetLastError());
        size_t scan_bytes = 0;
        RQ_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, B.flags, bidx.as<uint32_t>(), (int)nv, s));
        DevBuf scan_tmp(std::max<size_t>(scan_bytes, 16));
        RQ_HIP(hipcub::DeviceScan::ExclusiveSum(scan_tmp.ptr, scan_bytes, B.flags, bidx.as<uint32_t>(), (int)nv, s));
        uint32_t last[2] = {0, 0};
        RQ_HIP(hipMemcpyAsync(&last[0], bidx.as<uint32_t>() + (nv - 1), 4, hipMemcpyDeviceToHost, s));
        RQ_HIP(hipMemcpyAsync(&last[1], B.flags + (nv - 1), 4, hipMemcpyDeviceToHost, s));
        RQ_HIP(hipStreamSynchronize(s));
        const uint64_t nblocks = (uint64_t)last[0] + last[1];
        // every block holds a photon: a count above nvalid is a device fault
        if (nblocks == 0 || nblocks > nv) throw Error(RQ_ERR_DEVICE, "photon grid: inconsistent block count");
        grid->nblocks = (uint32_t)nblocks;
        grid->blocks = DevBuf((size_t)nblocks * 2 * sizeof(float4));
        grid->dense_block = DevBuf(ncells * sizeof(uint32_t));
        B.blocks = grid->blocks.as<float4>();
        B.dense_block = grid->dense_block.as<uint32_t>();
        hipLaunchKernelGGL(dev::pg_blocks, dim3(blocks_for(nv)), dim3(256), 0, s, B);
        RQ_HIP(hipGetLastError());
        RQ_HIP(hipStreamSynchronize(s));
    }
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
    v.dense_block = dense ? grid->dense_block.as<uint32_t>() : nullptr;
    v.blocks = dense ? grid->blocks.as<float4>() : nullptr;
    // A grid for another radius may still be read by queued queries.
    if (p->grid) RQ_HIP(hipDeviceSynchronize());
    p->grid = std::move(grid);
    return p->grid.get();
}

}  // namespace rq
