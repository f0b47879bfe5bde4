// kd_build.cpp -- ENCORE's spatial-median kd-tree, built on the host for the
// GPU kd path.
//
// The tree is the one Kdtree::build produces (Kdtree.cpp:84-200, non-SAH
// branch of construct at :299-680), node for node, so that the GPU traversal
// visits the same leaves with the same primitive order:
//  - root box = model bounds -/+ 0.001f, stored as corner + extent (AABB.h);
//  - max depth 11 + (unsigned)(1.3 * logf(N)); leaf size 2 if N <= 5000
//    else 10; a node also becomes a leaf when its size changed by at most 3
//    for three consecutive levels;
//  - axis = depth % 3, split = (lo + hi) / 2 of the node box;
//  - primitives go left/right by strict box overlap, to both when neither;
//  - node numbering: children of an inner node are allocated as a pair in
//    depth-first order (m_NodeUsed), split word = float bits with the axis in
//    the two low bits.
#include <chrono>
#include <cmath>
#include <cstdlib>

#include "rq_internal.h"

namespace rq {
namespace {

struct KBox {
    float pos[3], size[3];
};

inline KBox kbox(const float lo[3], const float hi[3]) {
    KBox b;
    for (int k = 0; k < 3; ++k) {
        b.pos[k] = lo[k];
        b.size[k] = hi[k] - lo[k];
    }
    return b;
}

// Primitive box pre-resolved into the two corners the overlap test reads:
// lo = pos, hi = pos + size (each rounded as AABB computes it).
struct PBox {
    float lo[3], hi[3];
};

struct KdBuilder {
    std::vector<PBox> pbox;
    std::vector<KdNodeG> nodes;
    std::vector<uint32_t> leaf_prims;
    uint32_t max_depth = 0, max_item = 10, used = 1, deepest = 0, nleaves = 0;

    void node_at(uint32_t fn) {
        if (nodes.size() <= fn) nodes.resize((size_t)fn + 1);
    }

    void construct(uint32_t fn, const KBox& box, std::vector<uint32_t>& prims, uint32_t depth, uint32_t pre_size,
                   bool dsize, int tries) {
        const uint32_t size = (uint32_t)prims.size();
        node_at(fn);
        if (depth > deepest) deepest = depth;
        if (std::abs((int)(pre_size - size)) <= 3) {
            dsize = true;
            tries++;
        } else {
            dsize = false;
            tries = 0;
        }
        if (size <= max_item || depth >= max_depth || (dsize && tries == 3)) {
            nodes[fn].word = (size << 2) | 3u;
            nodes[fn].child = (uint32_t)leaf_prims.size();
            leaf_prims.insert(leaf_prims.end(), prims.begin(), prims.end());
            ++nleaves;
            return;
        }
        const int axis = (int)(depth % 3);
        float lo[3], hi[3];
        for (int k = 0; k < 3; ++k) {
            lo[k] = box.pos[k];
            hi[k] = box.pos[k] + box.size[k];
        }
        const float split = (lo[axis] + hi[axis]) / 2;
        used += 2;
        const uint32_t left = used - 2;
        float lmax[3] = {hi[0], hi[1], hi[2]};
        float rmin[3] = {lo[0], lo[1], lo[2]};
        lmax[axis] = split;
        rmin[axis] = split;
        const KBox lv = kbox(lo, lmax), rv = kbox(rmin, hi);
        float lv_hi[3], rv_hi[3];
        for (int k = 0; k < 3; ++k) {
            lv_hi[k] = lv.pos[k] + lv.size[k];
            rv_hi[k] = rv.pos[k] + rv.size[k];
        }
        std::vector<uint32_t> L, R;
        L.reserve(size);
        R.reserve(size);
        for (uint32_t p : prims) {
            const PBox& b = pbox[p];
            const bool inl = (lv_hi[0] > b.lo[0]) && (lv.pos[0] < b.hi[0]) && (lv_hi[1] > b.lo[1]) &&
                             (lv.pos[1] < b.hi[1]) && (lv_hi[2] > b.lo[2]) && (lv.pos[2] < b.hi[2]);
            const bool inr = (rv_hi[0] > b.lo[0]) && (rv.pos[0] < b.hi[0]) && (rv_hi[1] > b.lo[1]) &&
                             (rv.pos[1] < b.hi[1]) && (rv_hi[2] > b.lo[2]) && (rv.pos[2] < b.hi[2]);
            if (inl || !inr) L.push_back(p);
            if (inr || !inl) R.push_back(p);
        }
        uint32_t bits;
        std::memcpy(&bits, &split, 4);
        nodes[fn].word = (bits & ~3u) | (uint32_t)axis;
        nodes[fn].child = left;
        std::vector<uint32_t>().swap(prims);
        construct(left, lv, L, depth + 1, size, dsize, tries);
        construct(left + 1, rv, R, depth + 1, size, dsize, tries);
    }
};

}  // namespace

void build_kd(const float* xyz, uint32_t nverts, const uint32_t* idx, uint32_t ntris, const float bmin[3],
              const float bmax[3], KdHost& out) {
    auto t0 = std::chrono::steady_clock::now();
    if (ntris == 0) throw Error(RQ_ERR_INVALID_ARG, "scene has no triangles");
    KdBuilder B;
    B.pbox.resize(ntris);
    for (uint32_t t = 0; t < ntris; ++t) {
        const float* v[3];
        for (int j = 0; j < 3; ++j) {
            const uint32_t vi = idx[3 * (size_t)t + j];
            if (vi >= nverts) throw Error(RQ_ERR_INVALID_ARG, "triangle index out of range");
            v[j] = xyz + 3 * (size_t)vi;
        }
        // TrianglePrim::getAABB (TrianglePrim.cpp:184-202), then the corner
        // pos + size that AABB::Intersect reads back.
        float lo[3], hi[3];
        for (int k = 0; k < 3; ++k) {
            const float a = v[0][k], b = v[1][k], c = v[2][k];
            const float m = (a > b) ? b : a;
            lo[k] = (m > c) ? c : m;
            const float M = (a > b) ? a : b;
            hi[k] = (M > c) ? M : c;
        }
        const KBox tb = kbox(lo, hi);
        for (int k = 0; k < 3; ++k) {
            B.pbox[t].lo[k] = tb.pos[k];
            B.pbox[t].hi[k] = tb.pos[k] + tb.size[k];
        }
    }
    float lo[3], hi[3];
    for (int k = 0; k < 3; ++k) {
        lo[k] = bmin[k] - 0.001f;
        hi[k] = bmax[k] + 0.001f;
    }
    const KBox root = kbox(lo, hi);
    B.max_depth = 11u + (unsigned)(1.3 * (double)std::log((float)ntris));
    B.max_item = (ntris <= 5000) ? 2u : 10u;
    B.nodes.reserve(2 * (size_t)ntris);
    std::vector<uint32_t> all(ntris);
    for (uint32_t i = 0; i < ntris; ++i) all[i] = i;
    B.construct(0, root, all, 0, 0, false, 0);
    out.nodes = std::move(B.nodes);
    out.nodes.resize(B.used);
    out.leaf_prims = std::move(B.leaf_prims);
    for (int k = 0; k < 3; ++k) {
        out.root_pos[k] = root.pos[k];
        out.root_size[k] = root.size[k];
    }
    out.stats.nnodes = B.used;
    out.stats.nleaves = B.nleaves;
    out.stats.nrefs = (uint32_t)out.leaf_prims.size();
    out.stats.max_depth = B.deepest;
    out.stats.kd_depth_limit = B.max_depth;
    out.stats.stack_bound = B.deepest + 2;
    out.stats.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace rq
