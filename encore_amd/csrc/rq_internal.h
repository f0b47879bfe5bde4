// rq_internal.h -- shared declarations of the ray-query engine (not public).
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/rq.h"

namespace rq {

// ---------------------------------------------------------------------------
// Errors: internal code throws rq::Error; the C ABI converts it to a status
// and a thread-local message.
// ---------------------------------------------------------------------------
struct Error : std::runtime_error {
    rq_status status;
    Error(rq_status s, const std::string& m) : std::runtime_error(m), status(s) {}
};

void hip_check(hipError_t e, const char* what);
// Runs fn(ctx) under the C ABI's error convention (status + rq_last_error).
rq_status guard_call(void (*fn)(void*), void* ctx);
#define RQ_HIP(call) ::rq::hip_check((call), #call)

// ---------------------------------------------------------------------------
// Device memory: every engine allocation goes through dev_alloc so that an
// allocation failure surfaces as RQ_ERR_OOM and can be injected in tests.
// ---------------------------------------------------------------------------
void* dev_alloc(size_t bytes);
void dev_free(void* p);

struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    explicit DevBuf(size_t n) : ptr(n ? dev_alloc(n) : nullptr), bytes(n) {}
    ~DevBuf() { dev_free(ptr); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : ptr(o.ptr), bytes(o.bytes) {
        o.ptr = nullptr;
        o.bytes = 0;
    }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) {
            dev_free(ptr);
            ptr = o.ptr;
            bytes = o.bytes;
            o.ptr = nullptr;
            o.bytes = 0;
        }
        return *this;
    }
    template <class T>
    T* as() const {
        return static_cast<T*>(ptr);
    }
    void upload(const void* src, size_t n) { RQ_HIP(hipMemcpy(ptr, src, n, hipMemcpyHostToDevice)); }
};

// ---------------------------------------------------------------------------
// Layouts in HBM
// ---------------------------------------------------------------------------
// BVH2 node, 64 bytes: the boxes of both children, interleaved so that one
// float4 holds one axis of both (lo, hi) pairs, then the two child refs.
//   x = (c0.lo.x, c0.hi.x, c1.lo.x, c1.hi.x), y = ..., z = ...
//   ref: bit 31 clear -> inner node index; bit 31 set -> leaf with
//        count-1 in bits 27..30 and the first triangle in bits 0..26.
struct alignas(64) BvhNode {
    float x[4], y[4], z[4];
    uint32_t ref[2];
    uint32_t pad[2];
};
static_assert(sizeof(BvhNode) == 64, "BVH node must be one 64-byte line");

// BVH4 node, 128 bytes: four child boxes as six float4 (lo.x, hi.x, lo.y,
// hi.y, lo.z, hi.z of children 0..3), four refs, padding.  An unused slot
// has an inverted infinite box (never hit) and ref kRefEmpty.
struct alignas(128) BvhNode4 {
    float lox[4], hix[4], loy[4], hiy[4], loz[4], hiz[4];
    uint32_t ref[4];
    uint32_t pad[4];
};
static_assert(sizeof(BvhNode4) == 128, "BVH4 node must be two 64-byte lines");
constexpr uint32_t kRefEmpty = 0xFFFFFFFFu;

constexpr uint32_t kRefLeaf = 0x80000000u;
constexpr uint32_t kRefSentinel = 0x7FFFFFFFu;
constexpr uint32_t kLeafMaxCount = 16;
constexpr uint32_t kLeafMaxFirst = (1u << 27) - 1;
// Rays, queries or sample rays per call and photons per map or emission
// buffer: the work counters and radix-sort sizes are 32-bit.
constexpr uint32_t kMaxCount = 1u << 30;

__host__ __device__ inline uint32_t make_leaf_ref(uint32_t first, uint32_t count) {
    return kRefLeaf | ((count - 1) << 27) | first;
}

// Triangle record, 48 bytes: three float4.  ENCORE predicate: (v0, prim),
// (v1 - v0, 0), (v2 - v0, 0) -- the edges are rounded exactly as
// TrianglePrim::intersect rounds them (TrianglePrim.cpp:210-211).
// Watertight predicate: (v0, prim), (v1, 0), (v2, 0).
struct alignas(16) TriRec {
    float a[4], b[4], c[4];
};
static_assert(sizeof(TriRec) == 48, "triangle record is 48 bytes");

// kd-tree node, 8 bytes, in ENCORE's numbering (root 0, children of an inner
// node at child and child + 1): word = leaf ? (count << 2) | 3 :
// (split bits & ~3) | axis (Kdtree.h:27-52); child = inner ? left child :
// offset of the leaf's primitive list.
struct KdNodeG {
    uint32_t word;
    uint32_t child;
};

struct BuildStats {
    uint32_t nnodes = 0, nleaves = 0, nrefs = 0, max_depth = 0, kd_depth_limit = 0;
    uint32_t stack_bound = 0;  // traversal stack entries a ray can need
    double seconds = 0;
};

// Host-side BVH build result.
struct BvhHost {
    uint32_t width = 2;
    std::vector<BvhNode> nodes;
    std::vector<BvhNode4> nodes4;
    std::vector<uint32_t> tri_order;  // leaf order -> original triangle index
    BuildStats stats;
};

void build_bvh(const float* xyz, uint32_t nverts, const uint32_t* idx, uint32_t ntris, uint32_t leaf_size,
               uint32_t width, BvhHost& out);

// Host-side kd build result (ENCORE's tree).
struct KdHost {
    std::vector<KdNodeG> nodes;
    std::vector<uint32_t> leaf_prims;
    float root_pos[3], root_size[3];
    BuildStats stats;
};

void build_kd(const float* xyz, uint32_t nverts, const uint32_t* idx, uint32_t ntris, const float bmin[3],
              const float bmax[3], KdHost& out);

void pack_triangles(const float* xyz, const uint32_t* idx, const uint32_t* order, uint32_t ntris,
                    uint32_t predicate, std::vector<TriRec>& out);

// Photon record in balanced (slot) order, 16 bytes: position and
// meta = splitting axis (bits 0..1) | normal class << 2.
struct alignas(16) PhotonHot {
    float x, y, z;
    uint32_t meta;
};
struct alignas(16) PhotonPower {
    float c[4];
};
constexpr uint32_t kClassNever = (1u << 30) - 1;  // matches no query (NaN normal)

struct PhotonHost {
    std::vector<PhotonHot> hot;
    std::vector<PhotonPower> power;
    std::vector<uint32_t> perm;   // slot -> insertion index
    std::vector<uint8_t> axis;    // slot -> splitting axis
    std::vector<std::array<uint32_t, 3>> normals;  // class -> canonical normal bits
    std::vector<uint32_t> table;  // open addressing: (x, y, z, class) per slot, class ~0 = empty
};

void build_photons(const float* pos, const float* nrm, const float* power, uint32_t n, uint32_t layout,
                   PhotonHost& out);
// Hash of a canonical normal (bits of x, y, z with -0 folded to +0); shared
// by the host table builder and the device lookup.
__host__ __device__ inline uint32_t normal_hash(uint32_t x, uint32_t y, uint32_t z) {
    uint32_t h = x * 0x9E3779B1u;
    h ^= (y + 0x7F4A7C15u) * 0x85EBCA77u;
    h ^= (z + 0x165667B1u) * 0xC2B2AE3Du;
    h ^= h >> 15;
    h *= 0x2C1B3C6Du;
    h ^= h >> 13;
    return h;
}

// Photon cell grid for radiance queries (photon_grid.hip).  The photons of a
// map that can ever be selected (finite position, a normal class) are
// bucketed into cubic cells whose edge is the query reach, stored in Morton
// order of their cells; an open-addressing table maps a cell to its range.
// Build and queries compute a cell coordinate with the same double
// arithmetic, grid_coord, which is monotonic in x.
struct PhotonGridView {
    const float4* pos;     // grid order: x, y, z and the normal class (bits)
    const uint32_t* slot;  // grid order -> balanced-tree slot (power, perm)
    const uint4* cells;    // (key low, key high, first, count); key ~0 = empty
    uint32_t cell_mask;    // table entries - 1
    // When the grid's box is small enough: (first, count) of every cell,
    // x fastest (count 0 = empty), read instead of the hash table.
    const uint2* dense;
    uint32_t dims[3];      // cells per axis, each below 2^21
    double origin[3];      // the photons' lower bounds
    double cell, inv_cell;
    double reach;          // every photon with float d2 < r2 lies within it
    // Dense grids only (else null): a cell's photons run along a Morton
    // curve of 16^3 sub-cells and are cut into blocks of up to 64 from the
    // cell's first photon.  dense_block: a cell's first block; blocks[2b],
    // blocks[2b + 1]: block b's bounding box, (lo.xyz, first photon) and
    // (hi.xyz, photon count), the integers as float bits.
    const uint32_t* dense_block;
    const float4* blocks;
};

__host__ __device__ inline int64_t grid_coord(double x, double origin, double inv_cell) {
    return (int64_t)floor((x - origin) * inv_cell);
}

__host__ __device__ inline uint64_t grid_spread21(uint64_t v) {
    v &= 0x1FFFFFull;
    v = (v | (v << 32)) & 0x1F00000000FFFFull;
    v = (v | (v << 16)) & 0x1F0000FF0000FFull;
    v = (v | (v << 8)) & 0x100F00F00F00F00Full;
    v = (v | (v << 4)) & 0x10C30C30C30C30C3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}

__host__ __device__ inline uint64_t cell_key(uint64_t x, uint64_t y, uint64_t z) {
    return (grid_spread21(x) << 2) | (grid_spread21(y) << 1) | grid_spread21(z);
}

__host__ __device__ inline uint32_t cell_hash(uint64_t key, uint32_t mask) {
    return (uint32_t)((key * 0x9E3779B97F4A7C15ull) >> 32) & mask;
}

struct KnnLaunch {
    const PhotonHot* hot;
    const PhotonPower* power;
    const uint32_t* perm;
    const uint32_t* table;
    uint32_t table_mask;
    uint32_t n;
    const float* qpos;
    const float* qnrm;
    uint32_t nq;
    uint32_t k;
    float r2;
    uint32_t* out_idx;   // nq*k insertion indices (ascending distance), ~0 padding; null = radiance only
    float* out_d2;       // nq*k squared distances, +inf padding (with out_idx)
    uint32_t* out_cnt;   // nq
    float* out_rgb;      // radiance estimate (nq*3) or null
    uint32_t caustic;
    uint32_t* spill;     // per-lane stack spill
    uint32_t spill_depth;
    PhotonGridView grid;    // knn_cell_kernel only
    const uint32_t* order;  // queries served in this order (null: 0..nq-1)
    const float4* qprep;    // knn_cell_kernel: position and class per served query
    unsigned long long* tally;  // knn_cell_kernel counting copy: queries, cells, photons read, candidates
    uint32_t compact_slack;     // knn_cell_kernel: compact at k + this many candidates (0: 64, or 128 for k > 64)
};

void launch_knn(const KnnLaunch& L, int grid, hipStream_t s);
int knn_grid(uint32_t nq, int num_cus, uint32_t k);
size_t knn_lds_bytes(uint32_t k);
constexpr uint32_t kKnnMaxK = 256;
uint32_t knn_spill_words_per_lane(uint32_t n);
// Radiance-only gather for large k, one query per wave (knn.hip): the stack
// capacity (entries per wave; three words each) for n photons, the grid, the
// launch.  From k = kKnnWaveMinK the cell-grid kernel serves queries; this
// one serves radiance without lists when no grid suits the radius.
constexpr uint32_t kKnnWaveMinK = 32;
uint32_t knn_wave_stack_cap(uint32_t n);
// Radiance without lists from the photon cell grid, one query per wave.
int knn_cell_grid(uint32_t nq, int num_cus);
// prep: nq float4 of scratch for the prepared queries.
void launch_knn_cells(const KnnLaunch& L, float4* prep, int grid, hipStream_t s);
// Queries ordered by the Morton key of their grid cell, so that the waves
// running at one time read the same cells: scratch bytes for nq queries, and
// the sort (order receives nq query indices; scratch holds the rest).
size_t knn_cell_order_bytes(uint32_t nq);
void knn_cell_order(const KnnLaunch& L, uint32_t* order, void* scratch, size_t scratch_bytes, hipStream_t s);
// Tree walks: the queries sorted by a Morton code of their position within
// [bmin, bmax] (scratch of knn_tree_order_bytes(nq)).
size_t knn_tree_order_bytes(uint32_t nq);
void knn_tree_order(const KnnLaunch& L, const float bmin[3], const float bmax[3], uint32_t* order, void* scratch,
                    size_t scratch_bytes, hipStream_t s);
int knn_wave_grid(uint32_t nq, int num_cus);
void launch_knn_wave(const KnnLaunch& L, int grid, uint32_t stack_cap, hipStream_t s);

// ---------------------------------------------------------------------------
// Kernel launch parameters (by value).
// ---------------------------------------------------------------------------
// One item of a batched launch (device copy), 64 bytes.
// Field order keeps (nodes, tris) and (hits, occ) in aligned 16-byte pairs,
// each read by one vector load when a lane takes a ray of the item.
struct alignas(64) BatchItemDev {
    const void* nodes;  // BvhNode* or BvhNode4* (all items of a launch share the width)
    const TriRec* tris;
    rq_hit* hits;
    uint32_t* occ;
    const rq_ray* rays;
    uint32_t begin;  // global index of the item's first ray
    uint32_t n;
    uint32_t pad[4];
};
static_assert(sizeof(BatchItemDev) == 64, "batch descriptor is 64 bytes");
static_assert(offsetof(BatchItemDev, hits) == 16 && offsetof(BatchItemDev, occ) == 24, "pointer pairs aligned");

struct TraceLaunch {
    const void* nodes;       // BvhNode* or KdNodeG*
    const TriRec* tris;
    const uint32_t* leaf_prims;  // kd only
    const rq_ray* rays;
    rq_hit* hits;            // closest
    uint32_t* occluded;      // any hit
    uint32_t n;
    const uint32_t* n_dev;   // if set, the ray count is read here at kernel start (n is then a bound)
    uint32_t* counter;       // work counter, zeroed before launch
    uint32_t* spill;         // per-lane spill stacks
    uint32_t spill_depth;    // entries per lane
    unsigned long long* counts;  // counting variant: [nodes visited, triangles tested]
    const BatchItemDev* items;   // batched launch: descriptor table sorted by begin
    uint32_t nitems;
    uint32_t fetch_threshold;    // refill idle lanes when fewer than this many are busy
    float root_pos[3], root_size[3];  // kd only
    uint32_t coop_lanes, coop_min;    // kd: cooperative leaf tests (see kd_trace)
    uint32_t leaf_batch;              // BVH: run the wave's leaf phase once this many lanes wait at a leaf
    const float* item_boxes;          // batched kd launch: root pos[3], pad, size[3], pad per item
};

enum class TraceKind { Closest = 0, Any = 1 };
// kd cooperative leaf tests: at most this many busy lanes, and leaves with
// more than this many triangles left.
// kd traversal ring: entries per lane in LDS (older ones spill to HBM).
constexpr int kKdStackEntries = 8;
constexpr uint32_t kKdCoopLanesDefault = 16;
constexpr uint32_t kKdCoopMinDefault = 16;

struct KernelConfig {
    int block = 256;
    int blocks_per_cu = 0;   // 0 = occupancy maximum
    int stack = 8;           // LDS ring entries per lane
    int fetch_threshold = 48; // BVH: refill idle lanes when fewer than this many of 64 are busy
    int kd_fetch_threshold = 16;  // kd path (its refill claims rays with an atomic per round)
};

// Final gather (gather.hip).
struct FgLaunch {
    const float* origin;      // shading points, 3 floats each
    const float* dirs;        // nper directions per point, 3 floats each
    uint32_t npoints, nper, nrays;
    rq_ray* rays;             // nrays
    const rq_hit* hits;       // nrays
    const TriRec* tris;
    const uint32_t* prim_slot;
    uint32_t predicate;
    float* qpos;              // nrays x 3: hit point (1e30 for a miss)
    float* qnrm;              // nrays x 3: hit normal
    const float* rgb_ind;     // nrays x 3
    const float* rgb_dir;     // nrays x 3 or null
    const float* diffuse;     // 3 per prim
    float* out_rgb;           // npoints x 3
    float* out_ri;            // npoints or null
};
void launch_fg_rays(const FgLaunch& L, hipStream_t s);
void launch_fg_attr(const FgLaunch& L, hipStream_t s);
void launch_fg_reduce(const FgLaunch& L, hipStream_t s);

// BVH update for moved vertices (refit.hip): triangle records rewritten from
// the new vertices, then every child box recomputed level by level, deepest
// first, as the min/max of its triangles' vertices or of its child node's
// boxes.
struct RefitLaunch {
    const float* xyz;          // new vertex positions, 3 per vertex
    const uint32_t* rec_vidx;  // 3 vertex indices per triangle record
    TriRec* tris;
    uint32_t nrecs, predicate;
    void* nodes;               // BvhNode (width 2) or BvhNode4
    uint32_t width;
    const uint32_t* order;     // node ids of one level: order[begin..end)
    uint32_t begin, end;
};
void launch_refit_tris(const RefitLaunch& L, hipStream_t s);
// GPU build of a BVH2 scene (lbvh.hip): fills the scene's nodes, triangle
// records, prim map, record vertex indices and statistics.
void build_bvh_device(rq_scene* sc, const float* xyz, uint32_t nverts, const uint32_t* idx, uint32_t ntris,
                      uint32_t leaf_size, hipStream_t s);
void launch_refit_level(const RefitLaunch& L, hipStream_t s);

// Photon emission (photon_trace.hip).
// A material as the tracer uses it: pd = mean diffuse reflectance,
// pt = transmittance, ptotal = 0 + pd + pt (PhotonMapperCPU.cpp:615-618).
struct PhMaterial {
    float pd, pt, ptotal, transparency, index, pad[3];
};
static_assert(sizeof(PhMaterial) == 32, "material record is 32 bytes");
struct PhLight {
    float center[3], normal[3], length, width, power[3], pad[3];
};
// Draws per path: slot 0 light, 1-2 emission point, 3-4 emission direction,
// then per bounce b: 5+3b the interaction draw, 6+3b and 7+3b the diffuse
// direction.  Event slots per path: 0 the first stored photon, 1 its shadow
// photon, b+1 the photon stored at bounce b >= 1.
constexpr uint32_t kPhMaxBounces = 16;
constexpr uint32_t kPhShadowPasses = 8;  // surfaces a shadow-photon ray may leave before it gives up
constexpr uint32_t kPhBlock = 1024;      // threads per block of the emission and shading kernels
constexpr uint32_t kPhPathBits = 22;  // live-ray tag: local path in bits 0..21
constexpr uint32_t kPhMaxBatch = 1u << kPhPathBits;
constexpr uint32_t kPhScanBlock = 256;
constexpr uint32_t kPhTableSize = 10000;  // rand() % 10000 outcomes
struct PhOut {
    float* pos;
    float* nrm;
    float* pow;
    float* dir;
    uint8_t* shadow;
    uint32_t* path;
};
struct PhLaunch {
    const TriRec* tris;
    const uint32_t* prim_slot;
    uint32_t predicate;
    const uint32_t* prim_mat;
    const PhMaterial* mats;
    const PhLight* lights;
    uint32_t nlights;
    const float* phi_cs;  // kPhTableSize (cos, sin) pairs
    uint64_t seed, base;  // draws key; global index of the batch's first path
    uint32_t npaths, max_bounces, nslots;
    rq_ray* path_ray;     // npaths: the photon's current ray
    float* path_dir;      // npaths x 3: emission direction
    uint32_t* path_light; // npaths
    // Events, slot-major: event (slot, path) at slot * stride + path.
    uint32_t stride;
    float* ev_pos;        // 3 floats per event
    float* ev_nrm;
    uint8_t* ev_tag;      // map + 1 (0 none), bit 2 shadow
    const rq_ray* in_rays;
    const uint32_t* in_tags;
    const rq_hit* hits;
    const uint32_t* nin_dev;  // live rays in the input queue (read on the device)
    uint32_t nin_max;
    rq_ray* out_rays;
    uint32_t* out_tags;
    uint32_t* out_count;
    unsigned long long* cnt;  // npaths: packed per-map event counts (21 bits each)
    uint32_t* boff;           // nblocks x 3: exclusive block offsets
    uint32_t nblocks;
    uint32_t* ctl;            // [0..1] stop path + 1 (global, caustic), [2..4] written, [5..7] batch totals
    uint32_t rem[3], acc[3];
    float scale[3];
    PhOut out[3];
};
void launch_ph_emit(const PhLaunch& L, hipStream_t s);
void launch_ph_shade(const PhLaunch& L, int max_blocks, hipStream_t s);
void launch_ph_compact(const PhLaunch& L, hipStream_t s);
// (cos, sin) of 2 * PI * (k / 10000) for every rand() % 10000 outcome k, in
// float as Common.cpp:150-156 evaluates them (PI = 3.14159f).
const std::vector<float>& ph_phi_table();
// Closest hit of n rays on the scene (the traversal behind rq_intersect);
// with n_dev the count is read on the device and n only bounds it.
void trace_closest(rq_scene* sc, const rq_ray* rays, uint32_t n, rq_hit* hits, hipStream_t s,
                   const uint32_t* n_dev = nullptr);

// Allocation-free queries (noalloc.hip): the work counter and the spill
// stacks of rq_intersect_closest_static live in a file-scope device array of
// kStaticScratchBytes, one copy per GPU.  bind_static_scratch (called by
// rq_device_open with the device current) stores that GPU's address of the
// array in dev->static_scratch.
constexpr size_t kStaticScratchBytes = RQ_STATIC_SCRATCH_BYTES;
void bind_static_scratch(rq_device* dev);
// Closest hit with caller-provided scratch (counter in the first 256 bytes,
// spill stacks after): the grid shrinks to what `bytes` can hold, and
// RQ_ERR_CAPACITY results when not even one workgroup's spill stacks fit.
// Allocates nothing.
void trace_closest_in(rq_scene* sc, const rq_ray* rays, uint32_t n, rq_hit* hits, hipStream_t s, uint32_t* scratch,
                      size_t bytes);

// Threads per block: the BVH kernels run one wave per block (their LDS use
// per wave is large, so single-wave blocks pack a CU best); the kd kernel
// four.
constexpr int kBvhBlock = 64;
constexpr int kKdBlock = 256;
inline int block_threads(uint32_t accel) { return accel == RQ_ACCEL_KD ? kKdBlock : kBvhBlock; }

// Launchers (traverse.hip).
void launch_bvh(const TraceLaunch& L, uint32_t predicate, TraceKind kind, uint32_t width, int grid,
                const KernelConfig& kc, hipStream_t s);
void launch_bvh_batched(const TraceLaunch& L, uint32_t predicate, TraceKind kind, uint32_t width, int grid,
                        const KernelConfig& kc, hipStream_t s);
int occupancy_bvh_batched(uint32_t predicate, TraceKind kind, uint32_t width, int stack);
void launch_kd(const TraceLaunch& L, uint32_t predicate, TraceKind kind, int grid, hipStream_t s);
void launch_kd_batched(const TraceLaunch& L, uint32_t predicate, TraceKind kind, int grid, hipStream_t s);
// BVH launches: lanes waiting at a leaf before the wave tests leaves
// (RQ_BVH_LEAF_BATCH overrides the default for sweeps).
uint32_t bvh_leaf_batch();
int occupancy_kd_batched(uint32_t predicate, TraceKind kind);
void launch_count(const TraceLaunch& L, uint32_t accel, uint32_t predicate, TraceKind kind, uint32_t width, int grid,
                  hipStream_t s);
void launch_zero_u32(uint32_t* p, uint32_t n, hipStream_t s);

// The cell grid of a photon map for squared radius r2 (photon_grid.hip).
// Largest dense cell table (bytes); larger grids use the hash table only.
constexpr uint64_t kGridDenseMaxBytes = 256ull << 20;
struct PhotonGrid {
    uint32_t r2_bits = 0;
    DevBuf pos, slot, cells, dense, dense_block, blocks;
    uint32_t nblocks = 0;
    PhotonGridView view{};
};
// Makes (or keeps) the grid for r2 on stream s and returns it, or null when
// a grid does not suit this radius (not finite or not positive, more than
// 2^21 - 1 cells along an axis, fewer than 4 along every axis, or map
// coordinates near 2^38 cells from the origin).  The
// caller holds p->grid_mu until the queries that read the grid are queued.
const PhotonGrid* photon_grid_for(rq_photons* p, float r2, hipStream_t s);
int occupancy_blocks_per_cu(uint32_t accel, uint32_t predicate, TraceKind kind, uint32_t width,
                            const KernelConfig& kc);
// Dwords per traversal stack entry: BVH2 1 (ref), BVH4 2 (ref, entry
// distance), kd 3 (node, tmin, tmax).
uint32_t stack_entry_words(uint32_t accel, uint32_t width);

}  // namespace rq

// ---------------------------------------------------------------------------
// Opaque public objects
// ---------------------------------------------------------------------------
struct rq_device {
    int ordinal = 0;
    int num_cus = 0;
    int max_waves_per_cu = 0;
    size_t total_mem = 0;
    std::string arch;
    std::mutex mu;
    // Per-stream scratch: the work counter in the first 256 bytes, spill
    // stacks after it.  Launches on one stream are ordered, so they can share.
    std::map<hipStream_t, rq::DevBuf> scratch;
    // This GPU's address of the allocation-free entry point's static array.
    uint32_t* static_scratch = nullptr;
    rq::KernelConfig kc;
    // Photon-emission workspace, kept between calls (grown on demand);
    // emissions on one device take turns on it.
    std::mutex emit_mu;
    rq::DevBuf emit_work;
};

struct rq_photons {
    rq_device* dev = nullptr;
    uint32_t n = 0, nclasses = 0, table_mask = 0;
    rq::DevBuf hot, power, perm, table;
    std::vector<uint32_t> perm_host;
    std::vector<uint8_t> axis_host;
    // Photons that a query can select (finite position, a normal class) and
    // their bounds; the cell grid for the last squared radius it was built
    // for (made on first use by a radiance query).
    uint32_t nvalid = 0;
    float bmin[3] = {0, 0, 0}, bmax[3] = {0, 0, 0};
    std::mutex grid_mu;
    std::unique_ptr<rq::PhotonGrid> grid;
};

struct rq_scene {
    rq_device* dev = nullptr;
    uint32_t accel = 0, predicate = 0, ntris = 0, width = 2;
    rq::DevBuf nodes, tris, leaf_prims;
    rq::DevBuf prim_slot;  // prim id -> index of one of its triangle records
    float root_pos[3] = {0, 0, 0}, root_size[3] = {0, 0, 0};
    rq::BuildStats stats;
    uint64_t device_bytes = 0;
    // Update support: the triangle list and build options of rq_scene_create,
    // and for a BVH the vertex indices of every record (leaf order) and its
    // nodes grouped by depth, deepest level first.  Absent in scenes loaded
    // from files.
    uint32_t nverts = 0;
    std::vector<uint32_t> idx_host;
    rq_scene_desc desc{};
    rq::DevBuf rec_vidx, refit_order;
    std::vector<uint32_t> refit_levels;  // offsets into refit_order, one level per pair
    rq::DevBuf materials;  // PhMaterial per material (rq_scene_set_materials)
    rq::DevBuf prim_mat;   // prim id -> material index
    uint32_t nmaterials = 0;
};

// Per-lane spill region (dwords) that covers the deepest traversal stack of
// a scene.
inline uint32_t scene_spill_words(const rq_scene* sc) {
    return sc->stats.stack_bound * rq::stack_entry_words(sc->accel, sc->width);
}
