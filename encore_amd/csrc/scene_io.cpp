// scene_io.cpp -- packed scene files: a resident scene written to disk as its
// device arrays and loaded back without a build (cache-warm restarts of a
// server, scenes shared between processes).
//
// File layout (little-endian, as the arrays sit in HBM):
//   SceneFileHeader (magic "RQSCENE\0", version, layout sizes, the scene's
//   kind and statistics, the byte size of each array, FNV-1a 64 of the
//   payload), then the arrays in order: nodes, triangle records, leaf
//   references, prim -> record map.
// A loader checks the magic, version, record sizes, that every array size
// matches the statistics, the file length and the checksum before it
// allocates anything on the device.
#include <chrono>
#include <cstdio>
#include <cstring>
#include <memory>

#include "rq_internal.h"

namespace rq {
namespace {

constexpr char kMagic[8] = {'R', 'Q', 'S', 'C', 'E', 'N', 'E', '\0'};
constexpr uint32_t kVersion = 1;

struct SceneFileHeader {
    char magic[8];
    uint32_t version, header_bytes;
    uint32_t bvh_node_bytes, bvh4_node_bytes, kd_node_bytes, tri_bytes;
    uint32_t accel, predicate, width, ntris;
    uint32_t nnodes, nleaves, nrefs, max_depth, kd_depth_limit, stack_bound;
    float root_pos[3], root_size[3];
    uint64_t nodes_bytes, tris_bytes, leaf_prims_bytes, prim_slot_bytes;
    uint64_t checksum;
};

uint64_t fnv1a(uint64_t h, const void* data, size_t n) {
    const unsigned char* p = static_cast<const unsigned char*>(data);
    for (size_t i = 0; i < n; ++i) {
        h ^= p[i];
        h *= 0x100000001b3ull;
    }
    return h;
}

constexpr uint64_t kFnvBasis = 0xcbf29ce484222325ull;

struct File {
    FILE* f = nullptr;
    ~File() {
        if (f) std::fclose(f);
    }
};

std::vector<unsigned char> download(const DevBuf& b) {
    std::vector<unsigned char> h(b.bytes);
    if (b.bytes) RQ_HIP(hipMemcpy(h.data(), b.ptr, b.bytes, hipMemcpyDeviceToHost));
    return h;
}

uint64_t node_bytes(uint32_t accel, uint32_t width) {
    if (accel == RQ_ACCEL_KD) return sizeof(KdNodeG);
    return width == 4 ? sizeof(BvhNode4) : sizeof(BvhNode);
}

void save(const rq_scene* sc, const char* path) {
    RQ_HIP(hipSetDevice(sc->dev->ordinal));
    RQ_HIP(hipDeviceSynchronize());
    const std::vector<unsigned char> nodes = download(sc->nodes), tris = download(sc->tris),
                                     leaf = download(sc->leaf_prims), slot = download(sc->prim_slot);
    SceneFileHeader h{};
    std::memcpy(h.magic, kMagic, sizeof(kMagic));
    h.version = kVersion;
    h.header_bytes = sizeof(SceneFileHeader);
    h.bvh_node_bytes = sizeof(BvhNode);
    h.bvh4_node_bytes = sizeof(BvhNode4);
    h.kd_node_bytes = sizeof(KdNodeG);
    h.tri_bytes = sizeof(TriRec);
    h.accel = sc->accel;
    h.predicate = sc->predicate;
    h.width = sc->width;
    h.ntris = sc->ntris;
    h.nnodes = sc->stats.nnodes;
    h.nleaves = sc->stats.nleaves;
    h.nrefs = sc->stats.nrefs;
    h.max_depth = sc->stats.max_depth;
    h.kd_depth_limit = sc->stats.kd_depth_limit;
    h.stack_bound = sc->stats.stack_bound;
    for (int k = 0; k < 3; ++k) {
        h.root_pos[k] = sc->root_pos[k];
        h.root_size[k] = sc->root_size[k];
    }
    h.nodes_bytes = nodes.size();
    h.tris_bytes = tris.size();
    h.leaf_prims_bytes = leaf.size();
    h.prim_slot_bytes = slot.size();
    uint64_t c = kFnvBasis;
    c = fnv1a(c, nodes.data(), nodes.size());
    c = fnv1a(c, tris.data(), tris.size());
    c = fnv1a(c, leaf.data(), leaf.size());
    c = fnv1a(c, slot.data(), slot.size());
    h.checksum = c;
    File f;
    f.f = std::fopen(path, "wb");
    if (!f.f) throw Error(RQ_ERR_INVALID_ARG, std::string("cannot open for writing: ") + path);
    bool ok = std::fwrite(&h, sizeof(h), 1, f.f) == 1;
    for (const auto* v : {&nodes, &tris, &leaf, &slot})
        ok = ok && (v->empty() || std::fwrite(v->data(), 1, v->size(), f.f) == v->size());
    ok = (std::fclose(f.f) == 0) && ok;
    f.f = nullptr;
    if (!ok) throw Error(RQ_ERR_DEVICE, std::string("write failed: ") + path);
}

void read_exact(FILE* f, void* dst, size_t n, const char* what) {
    if (n && std::fread(dst, 1, n, f) != n) throw Error(RQ_ERR_INVALID_ARG, std::string("truncated scene file: ") + what);
}

rq_scene* load(rq_device* dev, const char* path) {
    auto t0 = std::chrono::steady_clock::now();
    File f;
    f.f = std::fopen(path, "rb");
    if (!f.f) throw Error(RQ_ERR_INVALID_ARG, std::string("cannot open: ") + path);
    SceneFileHeader h{};
    read_exact(f.f, &h, sizeof(h), "header");
    if (std::memcmp(h.magic, kMagic, sizeof(kMagic)) != 0) throw Error(RQ_ERR_INVALID_ARG, "not a scene file");
    if (h.version != kVersion || h.header_bytes != sizeof(SceneFileHeader))
        throw Error(RQ_ERR_UNSUPPORTED, "scene file version not supported");
    if (h.bvh_node_bytes != sizeof(BvhNode) || h.bvh4_node_bytes != sizeof(BvhNode4) ||
        h.kd_node_bytes != sizeof(KdNodeG) || h.tri_bytes != sizeof(TriRec))
        throw Error(RQ_ERR_UNSUPPORTED, "scene file record layout differs from this library");
    if (!(h.accel == RQ_ACCEL_BVH || h.accel == RQ_ACCEL_KD) ||
        !(h.predicate == RQ_PRED_ENCORE || h.predicate == RQ_PRED_WATERTIGHT) ||
        (h.accel == RQ_ACCEL_BVH && h.width != 2 && h.width != 4) || h.ntris == 0)
        throw Error(RQ_ERR_INVALID_ARG, "scene file header is inconsistent");
    const uint64_t want_nodes = (uint64_t)h.nnodes * node_bytes(h.accel, h.width);
    const uint64_t want_tris = (uint64_t)(h.accel == RQ_ACCEL_KD ? h.nrefs : h.ntris) * sizeof(TriRec);
    if (h.nodes_bytes != want_nodes || h.tris_bytes != want_tris || h.prim_slot_bytes != (uint64_t)h.ntris * 4 ||
        h.leaf_prims_bytes % 4 != 0)
        throw Error(RQ_ERR_INVALID_ARG, "scene file array sizes do not match its header");
    std::vector<unsigned char> nodes(h.nodes_bytes), tris(h.tris_bytes), leaf(h.leaf_prims_bytes),
        slot(h.prim_slot_bytes);
    read_exact(f.f, nodes.data(), nodes.size(), "nodes");
    read_exact(f.f, tris.data(), tris.size(), "triangles");
    read_exact(f.f, leaf.data(), leaf.size(), "leaf references");
    read_exact(f.f, slot.data(), slot.size(), "prim map");
    unsigned char extra;
    if (std::fread(&extra, 1, 1, f.f) != 0) throw Error(RQ_ERR_INVALID_ARG, "scene file has trailing bytes");
    uint64_t c = kFnvBasis;
    c = fnv1a(c, nodes.data(), nodes.size());
    c = fnv1a(c, tris.data(), tris.size());
    c = fnv1a(c, leaf.data(), leaf.size());
    c = fnv1a(c, slot.data(), slot.size());
    if (c != h.checksum) throw Error(RQ_ERR_INVALID_ARG, "scene file checksum mismatch");
    // Every prim maps to a record inside the array.
    const uint32_t* s32 = reinterpret_cast<const uint32_t*>(slot.data());
    const uint64_t nrec = h.tris_bytes / sizeof(TriRec);
    for (uint32_t i = 0; i < h.ntris; ++i)
        if (s32[i] >= nrec) throw Error(RQ_ERR_INVALID_ARG, "scene file prim map out of range");

    RQ_HIP(hipSetDevice(dev->ordinal));
    auto sc = std::make_unique<rq_scene>();
    sc->dev = dev;
    sc->accel = h.accel;
    sc->predicate = h.predicate;
    sc->ntris = h.ntris;
    sc->width = h.accel == RQ_ACCEL_BVH ? h.width : 2;
    sc->stats.nnodes = h.nnodes;
    sc->stats.nleaves = h.nleaves;
    sc->stats.nrefs = h.nrefs;
    sc->stats.max_depth = h.max_depth;
    sc->stats.kd_depth_limit = h.kd_depth_limit;
    sc->stats.stack_bound = h.stack_bound;
    for (int k = 0; k < 3; ++k) {
        sc->root_pos[k] = h.root_pos[k];
        sc->root_size[k] = h.root_size[k];
    }
    sc->nodes = DevBuf(nodes.size());
    sc->nodes.upload(nodes.data(), nodes.size());
    sc->tris = DevBuf(tris.size());
    sc->tris.upload(tris.data(), tris.size());
    sc->leaf_prims = DevBuf(leaf.size());
    if (!leaf.empty()) sc->leaf_prims.upload(leaf.data(), leaf.size());
    sc->prim_slot = DevBuf(slot.size());
    sc->prim_slot.upload(slot.data(), slot.size());
    sc->device_bytes = sc->nodes.bytes + sc->tris.bytes + sc->leaf_prims.bytes + sc->prim_slot.bytes;
    sc->stats.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return sc.release();
}

template <class F>
rq_status run(F&& f) {
    return guard_call([](void* c) { (*static_cast<F*>(c))(); }, &f);
}

}  // namespace
}  // namespace rq

extern "C" {

rq_status rq_scene_save(const rq_scene* scene, const char* path) {
    return rq::run([&] {
        if (!scene || !path) throw rq::Error(RQ_ERR_INVALID_ARG, "null argument");
        rq::save(scene, path);
    });
}

rq_status rq_scene_load(rq_device* dev, const char* path, rq_scene** out) {
    return rq::run([&] {
        if (!dev || !path || !out) throw rq::Error(RQ_ERR_INVALID_ARG, "null argument");
        *out = rq::load(dev, path);
    });
}

}  // extern "C"
