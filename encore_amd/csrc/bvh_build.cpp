// bvh_build.cpp -- binned-SAH BVH2 builder (host) and triangle packing.
//
// Replaces the median-split builder of Bvh::build (BVH.cpp:123-187, 189-312)
// for the GPU path: 32-bin SAH over all three axes, leaves of up to
// `leaf_size` triangles, child boxes stored in the parent so one 64-byte node
// fetch decides both children.  Triangles are re-ordered into leaf order; the
// original index (ENCORE's primitive id) is kept in each record.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <future>
#include <numeric>

#include "rq_internal.h"

namespace rq {
namespace {

struct Aabb {
    float lo[3] = {INFINITY, INFINITY, INFINITY};
    float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    void grow(const Aabb& b) {
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], b.lo[k]);
            hi[k] = std::max(hi[k], b.hi[k]);
        }
    }
    void grow(const float* p) {
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], p[k]);
            hi[k] = std::max(hi[k], p[k]);
        }
    }
    float area() const {
        float d[3];
        for (int k = 0; k < 3; ++k) d[k] = std::max(0.0f, hi[k] - lo[k]);
        return d[0] * d[1] + d[1] * d[2] + d[2] * d[0];
    }
};

struct PrimRef {
    Aabb box;
    float c[3];
};

struct BNode {
    Aabb box;
    int32_t left = -1, right = -1;  // children (build-node index) if inner
    uint32_t first = 0, count = 0;  // leaf range if count > 0
};

constexpr int kBins = 32;
constexpr uint32_t kMedianDepth = 48;     // below this depth split at the object median
constexpr uint32_t kParallelMin = 65536;  // subtrees at least this large build on their own thread

struct Builder {
    const std::vector<PrimRef>& refs;
    std::vector<uint32_t>& ids;
    uint32_t leaf_size;

    Builder(const std::vector<PrimRef>& r, std::vector<uint32_t>& i, uint32_t ls) : refs(r), ids(i), leaf_size(ls) {}

    // Builds [b, e) into `out` (a subtree-local node list); returns the root index.
    int32_t build(uint32_t b, uint32_t e, uint32_t depth, std::vector<BNode>& out, uint32_t& max_depth,
                  int par_budget) {
        BNode node;
        Aabb cb;
        for (uint32_t i = b; i < e; ++i) {
            node.box.grow(refs[ids[i]].box);
            cb.grow(refs[ids[i]].c);
        }
        const uint32_t n = e - b;
        max_depth = std::max(max_depth, depth);
        const int32_t me = (int32_t)out.size();
        out.push_back(node);
        if (n <= 1) {
            out[me].first = b;
            out[me].count = n;
            return me;
        }
        uint32_t mid = b;
        bool found = false;
        if (depth < kMedianDepth) {
            // Binned SAH over all axes.
            float best_cost = INFINITY;
            int best_axis = -1, best_split = -1;
            for (int ax = 0; ax < 3; ++ax) {
                const float ext = cb.hi[ax] - cb.lo[ax];
                if (!(ext > 0.0f)) continue;
                const float scale = kBins / ext;
                Aabb bb[kBins];
                uint32_t bc[kBins] = {0};
                for (uint32_t i = b; i < e; ++i) {
                    const PrimRef& r = refs[ids[i]];
                    int k = std::min(kBins - 1, (int)((r.c[ax] - cb.lo[ax]) * scale));
                    bb[k].grow(r.box);
                    bc[k]++;
                }
                float ra[kBins];
                uint32_t rc[kBins];
                Aabb acc;
                uint32_t cnt = 0;
                for (int k = kBins - 1; k > 0; --k) {
                    acc.grow(bb[k]);
                    cnt += bc[k];
                    ra[k] = acc.area();
                    rc[k] = cnt;
                }
                acc = Aabb();
                cnt = 0;
                for (int k = 0; k < kBins - 1; ++k) {
                    acc.grow(bb[k]);
                    cnt += bc[k];
                    if (cnt == 0 || rc[k + 1] == 0) continue;
                    const float cost = acc.area() * cnt + ra[k + 1] * rc[k + 1];
                    if (cost < best_cost) {
                        best_cost = cost;
                        best_axis = ax;
                        best_split = k;
                    }
                }
            }
            const float parent_area = node.box.area();
            const float leaf_cost = (float)n;
            const float split_cost = 1.0f + (parent_area > 0 ? best_cost / parent_area : 0.0f);
            if (best_axis >= 0 && !(n <= leaf_size && leaf_cost <= split_cost)) {
                const float scale = kBins / (cb.hi[best_axis] - cb.lo[best_axis]);
                auto it = std::partition(ids.begin() + b, ids.begin() + e, [&](uint32_t id) {
                    const PrimRef& r = refs[id];
                    int k = std::min(kBins - 1, (int)((r.c[best_axis] - cb.lo[best_axis]) * scale));
                    return k <= best_split;
                });
                mid = (uint32_t)(it - ids.begin());
                found = (mid > b && mid < e);
            } else if (best_axis >= 0 || n <= leaf_size) {
                if (n <= leaf_size) {
                    out[me].first = b;
                    out[me].count = n;
                    return me;
                }
            }
        }
        if (!found) {
            if (n <= leaf_size) {
                out[me].first = b;
                out[me].count = n;
                return me;
            }
            // Object median on the longest centroid axis (also handles
            // coincident centroids, where any split point works).
            int ax = 0;
            float ext = -1;
            for (int k = 0; k < 3; ++k)
                if (cb.hi[k] - cb.lo[k] > ext) {
                    ext = cb.hi[k] - cb.lo[k];
                    ax = k;
                }
            mid = b + n / 2;
            std::nth_element(ids.begin() + b, ids.begin() + mid, ids.begin() + e,
                             [&](uint32_t x, uint32_t y) { return refs[x].c[ax] < refs[y].c[ax]; });
        }
        int32_t l, r;
        if (par_budget > 0 && n >= kParallelMin) {
            std::vector<BNode> rsub;
            uint32_t rdepth = 0;
            auto fut = std::async(std::launch::async, [&]() {
                return build(mid, e, depth + 1, rsub, rdepth, par_budget - 1);
            });
            l = build(b, mid, depth + 1, out, max_depth, par_budget - 1);
            const int32_t rroot = fut.get();
            // Append the right subtree, shifting its child links.
            const int32_t base = (int32_t)out.size();
            for (BNode nd : rsub) {
                if (nd.count == 0 && nd.left >= 0) {
                    nd.left += base;
                    nd.right += base;
                }
                out.push_back(nd);
            }
            r = rroot + base;
            max_depth = std::max(max_depth, rdepth);
        } else {
            l = build(b, mid, depth + 1, out, max_depth, 0);
            r = build(mid, e, depth + 1, out, max_depth, 0);
        }
        out[me].left = l;
        out[me].right = r;
        return me;
    }
};

}  // namespace

// Collapses the binary build tree into 4-wide nodes: a node's children are
// found by repeatedly opening the inner child with the largest surface area
// until four children are reached or only leaves remain.
static void flatten4(const std::vector<BNode>& bn, int32_t root, BvhHost& out, uint32_t& nleaves,
                     uint32_t& depth4) {
    out.nodes4.clear();
    out.nodes4.reserve(bn.size() / 3 + 2);
    auto empty_node = []() {
        BvhNode4 g;
        for (int c = 0; c < 4; ++c) {
            g.lox[c] = g.loy[c] = g.loz[c] = INFINITY;
            g.hix[c] = g.hiy[c] = g.hiz[c] = -INFINITY;
            g.ref[c] = kRefEmpty;
            g.pad[c] = 0;
        }
        return g;
    };
    auto set_child = [&](BvhNode4& g, int c, const BNode& n) {
        g.lox[c] = n.box.lo[0];
        g.hix[c] = n.box.hi[0];
        g.loy[c] = n.box.lo[1];
        g.hiy[c] = n.box.hi[1];
        g.loz[c] = n.box.lo[2];
        g.hiz[c] = n.box.hi[2];
    };
    out.nodes4.push_back(empty_node());
    nleaves = 0;
    depth4 = 0;
    if (bn[root].count > 0) {
        set_child(out.nodes4[0], 0, bn[root]);
        out.nodes4[0].ref[0] = make_leaf_ref(bn[root].first, bn[root].count);
        nleaves = 1;
        depth4 = 1;
        return;
    }
    struct Item {
        int32_t bi;
        uint32_t slot, depth;
    };
    std::vector<Item> todo{{root, 0, 1}};
    while (!todo.empty()) {
        const Item it = todo.back();
        todo.pop_back();
        depth4 = std::max(depth4, it.depth);
        std::vector<int32_t> kids{bn[it.bi].left, bn[it.bi].right};
        while (kids.size() < 4) {
            int best = -1;
            float best_area = -1.0f;
            for (size_t k = 0; k < kids.size(); ++k) {
                const BNode& n = bn[kids[k]];
                if (n.count == 0 && n.left >= 0 && n.box.area() > best_area) {
                    best_area = n.box.area();
                    best = (int)k;
                }
            }
            if (best < 0) break;
            const int32_t open = kids[best];
            kids[best] = bn[open].left;
            kids.insert(kids.begin() + best + 1, bn[open].right);
        }
        std::vector<Item> next;
        for (size_t k = 0; k < kids.size(); ++k) {
            const BNode& n = bn[kids[k]];
            uint32_t ref;
            if (n.count > 0) {
                ref = make_leaf_ref(n.first, n.count);
                ++nleaves;
            } else {
                ref = (uint32_t)out.nodes4.size();
                out.nodes4.push_back(empty_node());
                next.push_back({kids[k], ref, it.depth + 1});
            }
            BvhNode4& g = out.nodes4[it.slot];
            set_child(g, (int)k, n);
            g.ref[k] = ref;
        }
        for (auto r = next.rbegin(); r != next.rend(); ++r) todo.push_back(*r);
    }
}

void build_bvh(const float* xyz, uint32_t nverts, const uint32_t* idx, uint32_t ntris, uint32_t leaf_size,
               uint32_t width, BvhHost& out) {
    auto t0 = std::chrono::steady_clock::now();
    if (ntris == 0) throw Error(RQ_ERR_INVALID_ARG, "scene has no triangles");
    if (ntris > kLeafMaxFirst) throw Error(RQ_ERR_CAPACITY, "BVH supports at most 2^27-1 triangles");
    if (leaf_size == 0) leaf_size = 4;
    if (leaf_size > kLeafMaxCount) throw Error(RQ_ERR_INVALID_ARG, "leaf_size must be 1..16");
    std::vector<PrimRef> refs(ntris);
    for (uint32_t t = 0; t < ntris; ++t) {
        PrimRef& r = refs[t];
        for (int j = 0; j < 3; ++j) {
            const uint32_t v = idx[3 * (size_t)t + j];
            if (v >= nverts) throw Error(RQ_ERR_INVALID_ARG, "triangle index out of range");
            r.box.grow(xyz + 3 * (size_t)v);
        }
        for (int k = 0; k < 3; ++k) r.c[k] = 0.5f * (r.box.lo[k] + r.box.hi[k]);
    }
    std::vector<uint32_t> ids(ntris);
    std::iota(ids.begin(), ids.end(), 0u);
    std::vector<BNode> bn;
    bn.reserve(2 * (size_t)ntris / std::max(1u, leaf_size / 2) + 8);
    uint32_t max_depth = 0;
    Builder B(refs, ids, leaf_size);
    const int root = B.build(0, ntris, 0, bn, max_depth, 3);

    if (width == 4) {
        uint32_t nleaves4 = 0, depth4 = 0;
        flatten4(bn, root, out, nleaves4, depth4);
        out.width = 4;
        out.nodes.clear();
        out.tri_order = std::move(ids);
        out.stats.nnodes = (uint32_t)out.nodes4.size();
        out.stats.nleaves = nleaves4;
        out.stats.nrefs = ntris;
        out.stats.max_depth = depth4;
        out.stats.stack_bound = 3 * depth4 + 2;
        out.stats.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return;
    }
    if (width != 2) throw Error(RQ_ERR_INVALID_ARG, "bvh_width must be 2 or 4");
    out.width = 2;
    // Flatten: one BvhNode per inner build node, pre-order, left first.
    out.nodes.clear();
    out.nodes.reserve(bn.size() / 2 + 1);
    uint32_t nleaves = 0;
    std::vector<std::pair<int32_t, int32_t>> todo;  // (build node, gpu slot)
    auto child_ref = [&](int32_t c, std::vector<std::pair<int32_t, int32_t>>& q) -> uint32_t {
        const BNode& nd = bn[c];
        if (nd.count > 0) {
            ++nleaves;
            return make_leaf_ref(nd.first, nd.count);
        }
        const uint32_t slot = (uint32_t)out.nodes.size();
        out.nodes.push_back(BvhNode());
        q.push_back({c, (int32_t)slot});
        return slot;
    };
    auto fill = [&](BvhNode& g, int side, const BNode& c) {
        g.x[2 * side] = c.box.lo[0];
        g.x[2 * side + 1] = c.box.hi[0];
        g.y[2 * side] = c.box.lo[1];
        g.y[2 * side + 1] = c.box.hi[1];
        g.z[2 * side] = c.box.lo[2];
        g.z[2 * side + 1] = c.box.hi[2];
    };
    out.nodes.push_back(BvhNode());
    if (bn[root].count > 0) {
        // Single leaf: both child slots name it (a repeated test is harmless).
        BvhNode& g = out.nodes[0];
        fill(g, 0, bn[root]);
        fill(g, 1, bn[root]);
        g.ref[0] = g.ref[1] = make_leaf_ref(bn[root].first, bn[root].count);
        nleaves = 1;
    } else {
        todo.push_back({root, 0});
        while (!todo.empty()) {
            auto [bi, slot] = todo.back();
            todo.pop_back();
            const BNode nd = bn[bi];
            std::vector<std::pair<int32_t, int32_t>> kids;
            const uint32_t r0 = child_ref(nd.left, kids);
            const uint32_t r1 = child_ref(nd.right, kids);
            BvhNode& g = out.nodes[slot];
            fill(g, 0, bn[nd.left]);
            fill(g, 1, bn[nd.right]);
            g.ref[0] = r0;
            g.ref[1] = r1;
            g.pad[0] = g.pad[1] = 0;
            // Visit left before right so the left subtree follows its parent.
            for (auto it = kids.rbegin(); it != kids.rend(); ++it) todo.push_back(*it);
        }
    }
    out.tri_order = std::move(ids);
    out.stats.nnodes = (uint32_t)out.nodes.size();
    out.stats.nleaves = nleaves;
    out.stats.nrefs = ntris;
    out.stats.max_depth = max_depth;
    out.stats.stack_bound = max_depth + 2;
    out.stats.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

void pack_triangles(const float* xyz, const uint32_t* idx, const uint32_t* order, uint32_t ntris,
                    uint32_t predicate, std::vector<TriRec>& out) {
    out.resize(ntris);
    for (uint32_t i = 0; i < ntris; ++i) {
        const uint32_t t = order ? order[i] : i;
        const float* v0 = xyz + 3 * (size_t)idx[3 * (size_t)t];
        const float* v1 = xyz + 3 * (size_t)idx[3 * (size_t)t + 1];
        const float* v2 = xyz + 3 * (size_t)idx[3 * (size_t)t + 2];
        TriRec& r = out[i];
        for (int k = 0; k < 3; ++k) {
            r.a[k] = v0[k];
            if (predicate == RQ_PRED_ENCORE) {
                // Edges rounded once, as Vector3(v0, v1) rounds them.
                r.b[k] = v1[k] - v0[k];
                r.c[k] = v2[k] - v0[k];
            } else {
                r.b[k] = v1[k];
                r.c[k] = v2[k];
            }
        }
        std::memcpy(&r.a[3], &t, 4);
        r.b[3] = 0.0f;
        r.c[3] = 0.0f;
    }
}

}  // namespace rq
