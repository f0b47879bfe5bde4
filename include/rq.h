/*
 * rq.h -- C ABI of the MI355X ray-query engine (libencore_rq.so).
 *
 * The engine answers the queries that ENCORE's renderers send through
 * `AccelerationStructure` (src/accelerators/AccelerationStructure.h:33-67) and
 * `PhotonMap` (src/renderers/PhotonMap.h:28-104): build from a triangle list,
 * closest hit, any hit (shadow), and photon k-nearest-neighbour gather.
 *
 * Conventions
 *  - Every entry point is `extern "C"`, takes plain pointers and sizes, and
 *    returns an `rq_status`.  No exception crosses the ABI.  On failure the
 *    message is available from `rq_last_error()` (thread-local).
 *  - Pointers named `d_*` are device memory owned by the caller (hipMalloc'd
 *    or any device allocation on the scene's GPU).  Pointers without the
 *    prefix are host memory and are copied before the call returns.
 *  - `stream` is a `hipStream_t` passed as `void*`; NULL is the default
 *    stream.  Query calls are asynchronous with respect to the host.
 *  - Scenes and photon maps change only through rq_scene_refit and
 *    rq_scene_set_materials; between such calls any number of threads may
 *    issue queries on one object concurrently on distinct streams.
 *  - One call takes at most 2^30 rays, queries or sample rays, and a photon
 *    map or emission buffer holds at most 2^30 photons (work counters and
 *    sort sizes are 32-bit); larger counts return RQ_ERR_CAPACITY.
 */
#ifndef ENCORE_RQ_H
#define ENCORE_RQ_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum rq_status {
    RQ_OK = 0,
    RQ_ERR_INVALID_ARG = 1,  /* null pointer, bad size, bad enum value */
    RQ_ERR_OOM = 2,          /* device or host allocation failed; state unchanged */
    RQ_ERR_DEVICE = 3,       /* a HIP runtime call failed */
    RQ_ERR_UNSUPPORTED = 4,  /* option not available for this object */
    RQ_ERR_CAPACITY = 5,     /* request larger than a configured limit */
    RQ_ERR_TIMEOUT = 6,      /* a wait or a kernel's iteration cap expired */
    RQ_ERR_CANCELLED = 7     /* request cancelled before completion */
} rq_status;

typedef struct rq_device rq_device;
typedef struct rq_scene rq_scene;
typedef struct rq_photons rq_photons;
typedef struct rq_batch rq_batch;
typedef struct rq_server rq_server;

/* A ray: origin, direction and the open parameter interval (tmin, tmax).
 * A triangle hit at t is accepted iff tmin < t < tmax.  ENCORE's eye, shadow
 * and bounce rays use tmin = 0 and tmax = +inf (Kdtree.cpp:682-784 clips the
 * ray to the scene box starting at t = 0 and accepts any t > 0,
 * TrianglePrim.cpp:223-224).  32 bytes. */
typedef struct rq_ray {
    float ox, oy, oz, tmin;
    float dx, dy, dz, tmax;
} rq_ray;

/* Closest-hit record, 16 bytes.  `prim` is the index of the triangle in the
 * list given to rq_scene_create (ENCORE's model-list x primitive-list
 * concatenation order, Kdtree.cpp:99-114), RQ_NO_HIT on a miss.  `u` is the
 * barycentric weight of v1 and `v` the weight of v2 (ENCORE's alpha and beta,
 * TrianglePrim.cpp:219-220,239).  On a miss t = tmax and u = v = 0. */
typedef struct rq_hit {
    float t, u, v;
    uint32_t prim;
} rq_hit;

#define RQ_NO_HIT 0xFFFFFFFFu

typedef enum rq_accel_kind {
    RQ_ACCEL_BVH = 0,  /* binned-SAH BVH: BVH4, 128-byte nodes (default), or BVH2, 64-byte nodes */
    RQ_ACCEL_KD = 1    /* ENCORE's spatial-median kd-tree, same build rules */
} rq_accel_kind;

typedef enum rq_predicate {
    /* ENCORE's Moller-Trumbore variant, bit-exact: Cross() flushes components
     * below 1e-3 to zero, no determinant test, inclusive edges, t > tmin
     * (TrianglePrim.cpp:204-249, Vector3.h:148-162, Vector3.cpp:126-140). */
    RQ_PRED_ENCORE = 0,
    /* Watertight test of Woop, Benthin and Wald (JCGT 2013): no gaps or
     * double hits along shared edges; FP64 fallback for exact-zero edge
     * functions. */
    RQ_PRED_WATERTIGHT = 1
} rq_predicate;

typedef struct rq_scene_desc {
    uint32_t accel;        /* rq_accel_kind */
    uint32_t predicate;    /* rq_predicate */
    uint32_t leaf_size;    /* BVH only: max triangles per leaf, 1..16 (0 = 4) */
    uint32_t has_bounds;   /* KD only: 1 = bounds_min/max are the model-list
                              bounds Kdtree::build reads from IModel::MinPoint
                              and MaxPoint (Kdtree.cpp:99-107); 0 = vertex
                              bounds of the triangle list */
    float bounds_min[3];
    float bounds_max[3];
    uint32_t bvh_width;    /* BVH only: children per node, 2 (64-byte nodes)
                              or 4 (128-byte nodes); 0 = 4 */
    uint32_t builder;      /* BVH only: RQ_BUILD_HOST_SAH (0) or
                              RQ_BUILD_DEVICE_LBVH (1, width 2) */
} rq_scene_desc;

enum {
    /* Binned-SAH build on the host's cores: the fastest trees. */
    RQ_BUILD_HOST_SAH = 0,
    /* Linear BVH built on the GPU (Karras 2012 over 63-bit Morton codes of
     * the triangle centroids, subtrees of up to leaf_size triangles made
     * leaves): builds in milliseconds for scenes admitted on the fly; its
     * trees trace more slowly than SAH trees. */
    RQ_BUILD_DEVICE_LBVH = 1
};

/* Scene statistics, filled by rq_scene_stats. */
typedef struct rq_scene_info {
    uint32_t accel, predicate;
    uint32_t ntris;
    uint32_t nnodes;       /* BVH: inner nodes; KD: nodes incl. leaves */
    uint32_t nleaves;
    uint32_t nrefs;        /* triangle references in leaves */
    uint32_t max_depth;
    uint32_t kd_max_depth_limit;  /* KD: 11 + floor(1.3 ln N) (Kdtree.cpp:123) */
    uint64_t device_bytes; /* bytes resident in HBM for this scene */
    double build_seconds;  /* host build time */
    uint32_t bvh_width;    /* BVH: 2 or 4; KD: 0 */
    uint32_t stack_bound;  /* traversal stack entries a ray can need */
} rq_scene_info;

/* Library version string, e.g. "encore-rq 0.1.0 gfx950". */
const char* rq_version(void);

/* Message of the last failed call on this thread ("" if none). */
const char* rq_last_error(void);

/* Open GPU `ordinal` (hipSetDevice numbering).  One rq_device per GPU and
 * process is enough; it is thread-safe. */
rq_status rq_device_open(int ordinal, rq_device** out);
void rq_device_close(rq_device* dev);

/* Build an acceleration structure over an indexed triangle list (host
 * arrays: xyz = 3*nverts floats, tri_idx = 3*ntris vertex indices) and
 * upload it.  Replaces AccelerationStructure::build
 * (AccelerationStructure.h:47; Kdtree.cpp:84-200; BVH.cpp:123-187).
 * desc may be NULL (BVH, ENCORE predicate). */
rq_status rq_scene_create(rq_device* dev, const float* xyz, uint32_t nverts,
                          const uint32_t* tri_idx, uint32_t ntris,
                          const rq_scene_desc* desc, rq_scene** out);
/* A kd scene over a kd-tree built elsewhere -- ENCORE's own Kdtree, kept
 * instead of built again (Kdtree.h:25-58, 144-176: m_pTree[0..m_NodeUsed),
 * each leaf's m_pPrimList, m_Bounds) -- in the form rq_host_tree_copy
 * writes: nnodes nodes of 8 bytes {word, child}.  A leaf's word is
 * (count << 2) | 3 and its child the offset of its triangles in leaf_prims;
 * an inner node's word is the split's float bits with the axis in the low
 * two bits and its child the left child, the right one following it.
 * leaf_prims holds nrefs triangle indices into tri_idx; root_box = pos[3],
 * size[3] of the root cell.  The tree is checked (children and leaf ranges
 * in bounds, no node reached twice) and traversed as the kd scenes
 * rq_scene_create builds.  rq_scene_refit rebuilds such a scene with the
 * engine's own builder, as Kdtree::update rebuilds. */
rq_status rq_scene_create_kd(rq_device* dev, const float* xyz, uint32_t nverts, const uint32_t* tri_idx,
                             uint32_t ntris, const void* nodes, uint32_t nnodes, const uint32_t* leaf_prims,
                             uint32_t nrefs, const float* root_box, uint32_t predicate, rq_scene** out);
void rq_scene_destroy(rq_scene* scene);
rq_status rq_scene_stats(const rq_scene* scene, rq_scene_info* info);

/* Update for moved vertices (AccelerationStructure::update,
 * AccelerationStructure.h:33-67): d_xyz (device, 3*nverts floats, the same
 * vertex count and triangles as at rq_scene_create) replaces the positions.
 * A BVH keeps its tree and refits every box on the device, exactly, from the
 * new vertices (Bvh::update, BVH.cpp:52-121, only grows its old boxes); a kd
 * scene is built again on the host, as Kdtree::update does, over the new
 * vertex bounds (bounds given at creation are not reused: the vertices may
 * have left them).  Queries on the scene must not run while it updates.
 * Scenes loaded from files cannot be updated (RQ_ERR_INVALID_ARG). */
rq_status rq_scene_refit(rq_scene* scene, const float* d_xyz, uint32_t nverts, void* stream);

/* Packed scene files.  rq_scene_save writes a resident scene (its node,
 * triangle, leaf-reference and prim-map arrays, with a versioned header and
 * an FNV-1a checksum) to `path`; rq_scene_load reads one back onto `dev`
 * without a build.  The loaded scene answers every query bit for bit as the
 * saved one.  A file from another library version, with another record
 * layout, truncated, with trailing bytes or a bad checksum is refused
 * (RQ_ERR_INVALID_ARG or RQ_ERR_UNSUPPORTED).  No counterpart in ENCORE,
 * which rebuilds its tree from the model list on every start
 * (Kdtree.cpp:84-200). */
rq_status rq_scene_save(const rq_scene* scene, const char* path);
rq_status rq_scene_load(rq_device* dev, const char* path, rq_scene** out);

/* Closest hit for n rays: d_rays[n] -> d_hits[n].  Replaces
 * AccelerationStructure::intersect (AccelerationStructure.h:63,
 * Kdtree.cpp:682-784) for a batch. */
rq_status rq_intersect_closest(rq_scene* scene, const rq_ray* d_rays, uint32_t n,
                               rq_hit* d_hits, void* stream);

/* Closest hit exactly as rq_intersect_closest, allocating nothing.  No
 * counterpart in ENCORE: it is for callers that cannot allocate device memory
 * (code under hipGraph capture, tight serving loops, harnesses that forbid
 * allocation).  The launch's work counter and traversal spill stacks live in
 * a static device array of 64 MiB per GPU (RQ_STATIC_SCRATCH_BYTES), held
 * until the process ends; the grid shrinks to the workgroups whose spill
 * stacks fit there, and RQ_ERR_CAPACITY is returned when not even one does.
 * Every call on one GPU shares that array and leaves its contents behind
 * (the counter is cleared at the start of each call): calls that overlap on
 * one GPU (other streams, other threads) race and must not be issued. */
#define RQ_STATIC_SCRATCH_BYTES (64u << 20)
rq_status rq_intersect_closest_static(rq_scene* scene, const rq_ray* d_rays, uint32_t n, rq_hit* d_hits,
                                      void* stream);

/* Any hit: d_occluded[i] = 1 iff some triangle is hit with tmin < t < tmax,
 * else 0.  With tmax = |L - p| this is PhotonMapperCPU::IsInShadow
 * (PhotonMapperCPU.cpp:1192-1207), which ENCORE answers with a full
 * closest-hit query. */
rq_status rq_occluded(rq_scene* scene, const rq_ray* d_rays, uint32_t n,
                      uint32_t* d_occluded, void* stream);

/* ---------------------------------------------------------------------
 * Batched launches.  Many ray batches of any sizes, on any resident scenes
 * of one kind (all BVH scenes of one node width, or all kd scenes -- the
 * exact path, ENCORE's own tree traversed bit for bit as Kdtree::intersect,
 * Kdtree.cpp:682-784) sharing one predicate, traced by ONE persistent
 * launch: waves pull global ray indices from one counter and find the owning
 * batch in a device descriptor table.  The launch can be captured once into
 * a hipGraph and replayed.  A kd scene rebuilt by rq_scene_refit moves its
 * arrays: call rq_batch_set again before the next launch.  No counterpart in
 * ENCORE, whose renderers trace one ray per AccelerationStructure::intersect
 * call.
 * --------------------------------------------------------------------- */
typedef struct rq_batch_item {
    rq_scene* scene;       /* RQ_ACCEL_BVH or RQ_ACCEL_KD scene */
    const rq_ray* d_rays;  /* n rays */
    void* d_out;           /* n rq_hit (closest) or n uint32_t (any hit) */
    uint32_t n;
    uint32_t reserved;
} rq_batch_item;

typedef struct rq_kernel_config {
    uint32_t stack;            /* LDS ring entries per lane: 8, 16 or 32 (kd batches: always 8) */
    uint32_t blocks_per_cu;    /* workgroups per CU (0 = occupancy limit): one wave each for BVH
                                  scenes, four for kd scenes */
    uint32_t fetch_threshold;  /* refill idle lanes when fewer than this many of 64 are busy (0..64) */
    uint32_t grid;             /* out: workgroups of the last launch */
} rq_kernel_config;

rq_status rq_batch_create(rq_device* dev, uint32_t any_hit, rq_batch** out);
/* Replace the batch's items (host array; copied).  Waits for queued work of
 * the device, uploads the descriptor table and drops a captured graph. */
rq_status rq_batch_set(rq_batch* batch, const rq_batch_item* items, uint32_t count);
/* One persistent launch over every item.  use_graph = 1 captures the launch
 * (counter reset + kernel) into a hipGraph on first use and replays it. */
rq_status rq_batch_launch(rq_batch* batch, uint32_t use_graph, void* stream);
rq_status rq_batch_get_config(const rq_batch* batch, rq_kernel_config* cfg);
rq_status rq_batch_set_config(rq_batch* batch, const rq_kernel_config* cfg);
/* Time every stack x blocks-per-CU x fetch-threshold candidate once on the
 * batch's current items (outputs are overwritten), time the six fastest
 * again over five launches each, and keep the fastest of those; best_ms is
 * its mean launch time.  kd batches sweep blocks per CU and the threshold. */
rq_status rq_batch_autotune(rq_batch* batch, void* stream, rq_kernel_config* chosen, double* best_ms);
void rq_batch_destroy(rq_batch* batch);

/* ---------------------------------------------------------------------
 * Serving.  Scenes stay resident; clients submit ray batches from host
 * memory at any time.  A scheduler thread admits requests in arrival order
 * while a device arena (HBM budget) can hold their rays and outputs, and
 * keeps `slots` launches in flight: each free slot takes every admitted
 * request (of one kind) and runs copy-in, ONE persistent batched launch and
 * copy-out on its own stream, so requests arriving meanwhile overlap the
 * running launch (continuous batching).  A result never depends on the mix.
 * plain = 1 gives the per-request baseline (one request per launch, one
 * slot, the rq_intersect_closest kernel).  No counterpart in ENCORE.
 * --------------------------------------------------------------------- */
typedef struct rq_server_config {
    uint32_t slots;                    /* launches in flight (0 = 3) */
    uint32_t plain;                    /* 1 = one request per launch */
    uint32_t max_rays_per_launch;      /* no request joins a launch that would then exceed this many
                                          rays; a larger request runs alone (0 = 2^20) */
    uint32_t max_requests_per_launch;  /* 0 = 1024 */
    uint64_t hbm_budget_bytes;         /* request arena (0 = a quarter of free HBM) */
    uint64_t hbm_reserve_bytes;        /* HBM the arena must leave free */
    uint32_t stack, blocks_per_cu, fetch_threshold;  /* kernel config (0 = device default) */
    uint32_t reserved;
} rq_server_config;

typedef struct rq_request_info {
    double latency_ms;            /* submit -> results in host memory */
    double queue_ms;              /* submit -> launch issued */
    uint64_t launch_id;
    uint32_t rays_in_launch;
    uint32_t requests_in_launch;
} rq_request_info;

typedef struct rq_server_stats {
    uint64_t launches, requests_done, rays_done, queued, admitted, arena_bytes, arena_used;
} rq_server_counters;

rq_status rq_server_create(rq_device* dev, const rq_server_config* cfg, rq_server** out);
/* Queue n rays on a resident BVH scene.  h_rays (n rq_ray) and h_out (n
 * rq_hit, or n uint32_t for any_hit) must stay valid until the request
 * completes; pinned host memory (hipHostMalloc) lets the copies overlap
 * kernels.  RQ_ERR_CAPACITY if the request can never fit the arena. */
rq_status rq_server_submit(rq_server* server, rq_scene* scene, const rq_ray* h_rays, uint32_t n,
                           uint32_t any_hit, void* h_out, uint64_t* ticket);
/* Wait for a request (timeout_ms = 0xFFFFFFFF: forever) and release its
 * ticket.  RQ_ERR_TIMEOUT leaves the ticket valid; RQ_ERR_CANCELLED after
 * rq_server_cancel. */
rq_status rq_server_wait(rq_server* server, uint64_t ticket, uint32_t timeout_ms, rq_request_info* info);
rq_status rq_server_poll(rq_server* server, uint64_t ticket, int* done);
/* Cancel a request that has not been launched yet. */
rq_status rq_server_cancel(rq_server* server, uint64_t ticket);
rq_status rq_server_stats(rq_server* server, rq_server_counters* stats);
/* Cancels queued requests, waits for launched ones, frees the arena. */
void rq_server_destroy(rq_server* server);

/* ---------------------------------------------------------------------
 * Photon maps (PhotonMap.h:28-104).
 * --------------------------------------------------------------------- */

typedef enum rq_photon_layout {
    /* ENCORE's balanced tree, photon for photon: node = middle of its index
     * range, axis from BalanceAndConstruct's bookkeeping box
     * (PhotonMap.cpp:167-221; that box picks z at nearly every node). */
    RQ_PHOTONS_ENCORE_ORDER = 0,
    /* Same implicit layout, axis = longest extent of each range's box.
     * Gathers the same sets with far fewer visits. */
    RQ_PHOTONS_SPATIAL = 1
} rq_photon_layout;

/* Build a balanced photon kd-tree over n photons given in insertion order
 * (PhotonMap::AddPhoton then BalanceAndConstruct, PhotonMap.cpp:46-53,
 * 167-221) and upload it.  Host arrays of 3*n floats: position, surface
 * normal, power (power may be NULL when only k-NN is needed).  Positions are
 * finite, as ENCORE's hit points are: a balanced tree over NaN or infinite
 * coordinates is not ordered, and gathers over it are not defined. */
rq_status rq_photons_create(rq_device* dev, const float* pos, const float* normal, const float* power,
                            uint32_t n, uint32_t layout, rq_photons** out);
void rq_photons_destroy(rq_photons* photons);
/* The balanced order: perm[slot] = insertion index and axis[slot] = the
 * splitting axis ENCORE stores in the photon (Photon::SetPlane); nclasses =
 * number of distinct normals.  Any output may be NULL. */
rq_status rq_photons_order(const rq_photons* photons, uint32_t* perm, uint8_t* axis, uint32_t* nclasses);
/* Host only (no GPU): the balanced order rq_photons_create would use for n
 * positions (3*n floats) and a layout: perm[n], axis[n]. */
rq_status rq_photons_balance_host(const float* pos, uint32_t n, uint32_t layout, uint32_t* perm, uint8_t* axis);

/* k nearest photons of each query (GetNearestNPhotons, PhotonMap.cpp:76-95,
 * 224-297): photons whose normal equals the query normal (== per component,
 * so -0 equals +0) with squared distance (x*x)+(y*y)+(z*z) strictly below
 * fl(max_dist*max_dist); the k nearest of them.  Device arrays: d_qpos and
 * d_qnormal are 3*nq floats; d_idx and d_d2 receive nq*k entries per query,
 * nearest first (insertion indices; ~0 and +inf pad short lists); d_count[nq]
 * (may be NULL) receives the count.  1 <= k <= 256.  Photons tied with the
 * k-th distance may be chosen differently from ENCORE's heap. */
rq_status rq_knn(rq_photons* photons, const float* d_qpos, const float* d_qnormal, uint32_t nq, uint32_t k,
                 float max_dist, uint32_t* d_idx, float* d_d2, uint32_t* d_count, void* stream);

/* Radiance estimate (GetRadianceEstimate, PhotonMap.cpp:121-164) fused with
 * the k-NN gather: d_rgb[3*nq] = sum of photon power over the gathered set
 * (caustic = 0), or the cone-weighted sum (r - d)/d times k / (PI r^2) with r
 * and d the squared distances of the farthest and of each photon and
 * PI = 3.14159f (caustic = 1).  The gathered set is also written to d_idx,
 * d_d2, d_count as in rq_knn; d_idx and d_d2 may both be NULL when only the
 * radiance (and count) is wanted. */
rq_status rq_radiance_estimate(rq_photons* photons, const float* d_qpos, const float* d_qnormal, uint32_t nq,
                               uint32_t k, float max_dist, uint32_t caustic, float* d_rgb, uint32_t* d_idx,
                               float* d_d2, uint32_t* d_count, void* stream);
/* The gather work of rq_radiance_estimate (no lists), counted by an
 * instrumented copy of the photon cell-grid kernel (synchronous):
 * counts[0] = queries, counts[1] = photon ranges ranked (grid cells, or the
 * 64-photon blocks of those cells when a query ranks blocks), counts[2] =
 * photons read, counts[3] = photons within the radius at the time they were
 * read (candidates); d_rgb receives the estimates.  RQ_ERR_UNSUPPORTED when k or
 * the radius sends the call to another kernel.  No counterpart in ENCORE
 * (measurement). */
rq_status rq_radiance_counts(rq_photons* photons, const float* d_qpos, const float* d_qnormal, uint32_t nq,
                             uint32_t k, float max_dist, uint32_t caustic, float* d_rgb, uint64_t* counts,
                             void* stream);

/* PhotonMap::PrecomputeIrradiance (PhotonMap.cpp:57-74): of the map's
 * photons in ENCORE's balanced order (m_photons after BalanceAndConstruct,
 * which depends on the positions only), every increment-th from the first
 * gets the radiance estimate of the map at its own position and normal
 * (GetRadianceEstimate with k, max_dist and the caustic weighting when
 * caustic != 0) as its power, and these m = ceil(n / increment) photons, in
 * that order, make a new map *out with the given layout.  pos and normal are
 * the map's photons in insertion order (host, 3*n floats each, as given to
 * rq_photons_create).  chosen[m] (host, may be NULL) receives the insertion
 * index of each new photon, power_out[3*m] (host, may be NULL) its power.
 * Synchronous. */
rq_status rq_photons_precompute_irradiance(rq_photons* photons, const float* pos, const float* normal, uint32_t n,
                                           uint32_t increment, uint32_t k, float max_dist, uint32_t caustic,
                                           uint32_t layout, uint32_t* chosen, float* power_out, rq_photons** out);

/* Final gather: PhotonMapperCPU::CalculateIndirectIllumination's
 * final-gather branch (PhotonMapperCPU.cpp:1110-1167) for npoints shading
 * points at once.  For point i (origin d_origin[3i..3i+2]) and each of its
 * nper sample directions d_dirs[3(i nper + j)..] (unit vectors, as
 * Sampler::GetSampleRaysForHemisphere returns them, Sampler.cpp:23-47, with
 * the caller's jitter), a ray from the origin with no offset is traced to its
 * closest hit; at each hit, with ENCORE's hit point and hit normal
 * (TrianglePrim.cpp:236-239), the radiance estimate of `indirect` (k_ind,
 * r_ind) and, if `direct` is not NULL, of `direct` (k_dir, r_dir) are each
 * multiplied by the hit triangle's diffuse colour d_diffuse[3 prim..] and
 * summed in sample order.  d_rgb[3i..] = that sum times 3.14159f / nper;
 * d_ri[i] (may be NULL) = hits / sum(1 / hit distance), or 0 without hits.
 * The caller applies the shading point's own diffuse colour and the 0.35
 * scale (PhotonMapperCPU.cpp:1178).  With the kd scene the result is the
 * reference's bit for bit (up to k-th-distance photon ties); with a BVH scene
 * up to closest-hit ties and flush phantoms.  Synchronous; uses a temporary
 * device workspace of up to 4M sample rays x 96 bytes.  No counterpart as a
 * single call in ENCORE (it loops one shading point at a time). */
rq_status rq_final_gather(rq_scene* scene, rq_photons* indirect, uint32_t k_ind, float r_ind, rq_photons* direct,
                          uint32_t k_dir, float r_dir, const float* d_diffuse, const float* d_origin,
                          uint32_t npoints, const float* d_dirs, uint32_t nper, float* d_rgb, float* d_ri,
                          void* stream);

/* ---------------------------------------------------------------------
 * Irradiance cache of the final gather (IrrCache<Color>,
 * IrrCacheOctree.h:92-326; PhotonMapperCPU.cpp:175, 1087-1103, 1170-1173).
 * --------------------------------------------------------------------- */
typedef struct rq_irrcache rq_irrcache;

/* IrrCache(a, minPoint, xSize, ySize, zSize): a is the final-gather
 * tolerance (FGTolerance, pmConfig 0.08); the root box is min_point[3],
 * size[3] (ENCORE: (-75, -75, -75), 150^3). */
rq_status rq_irrcache_create(rq_device* dev, float a, const float* min_point, const float* size,
                             rq_irrcache** out);
/* AddElement for n cached values in order (host arrays): value[3n] (the
 * illumination Color), pos[3n], normal[3n], ri[n] (the harmonic mean
 * distance).  status[i] (may be NULL): 0 stored, 1 outside the root box and
 * dropped (as AddElement drops it), 2 refused.  A value with 2 a Ri above
 * the root's x size is one the reference's AddElement never returns from
 * (it descends without end, IrrCacheOctree.h:229-241): the call then returns
 * RQ_ERR_UNSUPPORTED and stores none of the n values. */
rq_status rq_irrcache_add(rq_irrcache* cache, const float* value, const float* pos, const float* normal,
                          const float* ri, uint32_t n, uint32_t* status);
/* GetIrrValuesNearPoint and the caller's weighted average for nq lookups
 * (device arrays d_qpos[3nq], d_qnrm[3nq]): d_rgb[3i..] = sum(value_j w_j) /
 * sum(w_j) over the cached values with GetWeight w_j = 1 / (|p - x_j| / R_j
 * + sqrt(1 - n.n_j)) above 1 / a that are not SampleInFront, summed in
 * insertion order (0 when no value counts); d_wsum[i] (may be NULL) = the
 * weight sum, which ENCORE tests before using the cache; d_count[i] (may be
 * NULL) = the values counted.  The caller applies the hit's diffuse colour
 * and the 0.35 scale (PhotonMapperCPU.cpp:1103).  Bit for bit the reference's
 * result.  The first lookup after rq_irrcache_add builds the cache's index
 * (host, synchronous). */
rq_status rq_irrcache_query(rq_irrcache* cache, const float* d_qpos, const float* d_qnrm, uint32_t nq,
                            float* d_rgb, float* d_wsum, uint32_t* d_count, void* stream);
/* The lookup work of rq_irrcache_query, counted by an instrumented copy of
 * its kernel (synchronous): counts[0] = lookups, counts[1] = cached values
 * examined, counts[2] = values counted; d_rgb receives the results.  No
 * counterpart in ENCORE (measurement). */
rq_status rq_irrcache_counts(rq_irrcache* cache, const float* d_qpos, const float* d_qnrm, uint32_t nq,
                             float* d_rgb, uint64_t* counts, void* stream);
/* Values stored; of them, the ones every lookup checks (unbounded reach);
 * cells and cell references of the index (after a lookup).  Any may be NULL. */
rq_status rq_irrcache_info(const rq_irrcache* cache, uint32_t* count, uint32_t* global_count,
                           uint64_t* index_cells, uint64_t* index_refs);
void rq_irrcache_destroy(rq_irrcache* cache);

/* ---------------------------------------------------------------------
 * Photon emission: PhotonMapperCPU::EmitPhotons and TracePhoton
 * (PhotonMapperCPU.cpp:503-775) on the device.
 * --------------------------------------------------------------------- */

/* A surface material, as the photon tracer reads it (Material.h:46-57):
 * diffuse reflectance (its mean (r+g+b)/3 is the reflection probability),
 * transmittance (the refraction probability), transparency (a surface
 * refracts only when it is at least 0.6) and refraction index. */
typedef struct rq_material {
    float diffuse[3];
    float transmittance;
    float transparency;
    float refraction_index;
} rq_material;

/* Attach materials to a scene: prim_material[ntris] (host) gives each
 * triangle's index into materials[nmaterials].  Replaces earlier ones.
 * Materials are not stored in packed scene files. */
rq_status rq_scene_set_materials(rq_scene* scene, const rq_material* materials, uint32_t nmaterials,
                                 const uint32_t* prim_material);

/* Rectangle light (RectangleLight.cpp:14-80, 172-183): emission points
 * uniform over a length x width rectangle centred at `center` facing
 * `normal`, lifted 0.01 along the normal; cosine-weighted directions. */
typedef struct rq_rect_light {
    float center[3];
    float normal[3];
    float length, width;
    float power[3];
} rq_rect_light;

/* Photon maps, in the order of PhotonMapperCPU's members. */
enum { RQ_MAP_GLOBAL = 0, RQ_MAP_DIRECT = 1, RQ_MAP_CAUSTIC = 2 };

typedef struct rq_photon_params {
    uint32_t max_bounces;   /* MaxBounces (pmConfig: 3); at most 16 */
    uint32_t capacity[3];   /* GlobalPhotons, DirectPhotons, CausticPhotons */
    float scale[3];         /* ScalePhotons factors: PhotonScale, 0.05, 0.00005 */
    uint64_t seed;          /* keys the counter-based draws that replace rand() */
    uint64_t max_paths;     /* emission bound (0: max(2^24, 64 x capacities)) */
    uint32_t batch_paths;   /* paths per device batch (0: up to 2^20) */
    uint32_t pad;
} rq_photon_params;

/* Device output arrays of one map, each sized for its capacity.  pos,
 * normal and power (3 floats per photon) are required when the capacity is
 * non-zero; dir (incoming direction, 3 floats), shadow (1 byte: 1 for a
 * shadow photon) and path (the emission index of the photon's path) may be
 * NULL. */
typedef struct rq_photon_buffers {
    float* pos;
    float* normal;
    float* power;
    float* dir;
    uint8_t* shadow;
    uint32_t* path;
} rq_photon_buffers;

typedef struct rq_photon_result {
    uint32_t count[3];  /* photons stored per map */
    uint64_t paths;     /* photon paths emitted */
} rq_photon_result;

/* Emit photon paths from the lights until the global map is full and then
 * until the caustic map is full (EmitPhotons' two loops), tracing each path
 * through the scene (TracePhoton): at every hit a draw E in [0, 1) against
 * the hit material chooses diffuse reflection (E < mean diffuse: store, new
 * cosine-weighted direction), refraction (E < mean diffuse +
 * transmittance: TransmitPhoton) or absorption (store, stop).  A stored
 * photon goes to the caustic map after a refraction, else to the global map
 * after the first bounce, else to the direct map together with a shadow
 * photon (negative power) at the next surface entered behind it
 * (EmitShadowPhoton).  Each map accepts photons in emission order until it
 * is full; powers are multiplied by the map's scale.  ENCORE draws from
 * rand(); here every draw is a function of (seed, path, draw slot), so the
 * photons are reproducible and the same for any batch size.  The scene needs
 * materials (rq_scene_set_materials).  Synchronous; `stream` orders the
 * work. */
rq_status rq_emit_photons(rq_scene* scene, const rq_rect_light* lights, uint32_t nlights,
                          const rq_photon_params* params, const rq_photon_buffers out[3],
                          rq_photon_result* result, void* stream);

/* Traversal work of one query over n rays, counted by an instrumented copy
 * of the same kernel (synchronous): counts[0] = rays, counts[1] = nodes
 * visited (64-byte BVH2, 128-byte BVH4 or 8-byte kd nodes), counts[2] =
 * triangles tested (48-byte records, stored in leaf order on both paths).
 * These are the N_node and N_tri of the algorithmic-bytes figure used for
 * the roofline.  No counterpart in ENCORE (measurement). */
rq_status rq_trace_counts(rq_scene* scene, const rq_ray* d_rays, uint32_t n, uint32_t any_hit,
                          uint64_t* counts, void* stream);

/* SIMD efficiency of the BVH kernel on n rays (synchronous, BVH scenes):
 * counts[0..2] as rq_trace_counts, counts[3] = wave iterations of the node
 * loop, counts[4] = wave iterations of the triangle loop, counts[5] = refill
 * rounds.  Lane utilisation of the node loop is counts[1] / (64 counts[3]),
 * of the triangle loop counts[2] / (64 counts[4]).  No counterpart in ENCORE
 * (measurement). */
rq_status rq_trace_profile(rq_scene* scene, const rq_ray* d_rays, uint32_t n, uint32_t any_hit,
                           uint64_t* counts, void* stream);

/* ---------------------------------------------------------------------
 * Host-side trees.  Builds the same structure rq_scene_create uploads, on
 * the host only (no GPU needed), and copies it out: for inspection, for
 * comparison with ENCORE's own Kdtree numbering, and for storing a packed
 * scene.  KD nodes are 8 bytes {word, child} in Kdtree's numbering
 * (Kdtree.h:25-58: word = (count << 2) | 3 for a leaf, else the split's float
 * bits with the axis in the low two bits; child = left child, or the leaf's
 * offset into leaf_prims).  BVH nodes are 64 bytes: float x[4], y[4], z[4]
 * holding (lo, hi) of child 0 then child 1 per axis, then uint32 ref[2]
 * (bit 31 clear: inner node index; set: leaf, count-1 in bits 27..30, first
 * triangle in bits 0..26) and 8 bytes of padding; `leaf_prims` then receives
 * the triangle order of the leaves (leaf slot -> original triangle index).
 * --------------------------------------------------------------------- */
typedef struct rq_host_tree rq_host_tree;

rq_status rq_host_tree_build(const float* xyz, uint32_t nverts, const uint32_t* tri_idx, uint32_t ntris,
                             const rq_scene_desc* desc, rq_host_tree** out);
rq_status rq_host_tree_info(const rq_host_tree* tree, rq_scene_info* info);
/* nodes: info.nnodes * (8 (KD), 64 (BVH2) or 128 (BVH4)) bytes; leaf_prims: info.nrefs (KD) or
 * info.ntris (BVH) entries; root_box: pos[3], size[3] (KD) or zeros. */
rq_status rq_host_tree_copy(const rq_host_tree* tree, void* nodes, uint32_t* leaf_prims, float* root_box);
void rq_host_tree_destroy(rq_host_tree* tree);

/* ENCORE's GPU texture encoding of a kd tree (Kdtree::buildGPU,
 * Kdtree.cpp:876-1000; read by shaders/Kdtree_restart.cg), from a KD host
 * tree and the triangle list it was built from.  cell_data0: 4 floats per
 * node in node order -- inner (left child, split word as float, 0,
 * -(axis + 2)), leaf (index of its first triangle, 0, 0, count); vertex0/1/2:
 * 4 floats per leaf reference (x, y, z, 0) in node order; ref_prim: the
 * triangle id of each reference, for the caller's normals and materials
 * (m_pNormal*Array, m_pMaterial0Array).  Sizes: info.nnodes and info.nrefs.
 * Any output may be NULL.  Host only. */
rq_status rq_host_tree_export_textures(const rq_host_tree* tree, const float* xyz, uint32_t nverts,
                                       const uint32_t* tri_idx, float* cell_data0, float* vertex0, float* vertex1,
                                       float* vertex2, uint32_t* ref_prim);

/* Test hook: the n-th device allocation from now fails with RQ_ERR_OOM
 * (n = 0 disables).  For exercising the out-of-memory paths. */
void rq_debug_fail_alloc_after(int64_t n);

#ifdef __cplusplus
}
#endif

#endif /* ENCORE_RQ_H */
