// encore_oracle.cpp -- CPU restatement of ENCORE's ray-query hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the parity oracle for the HIP
// engine in encore_amd/: tests, the smoke check and the CPU-baseline leg of
// bench.py load it as the checker.  The engine never links or calls it.
//
// It restates, in this project's own code, the arithmetic and control flow of
// the reference files cited at each function, keeping ENCORE's FP32 operation
// order so that results are bit-identical to the reference built with
// g++ -O2 -ffp-contract=off on x86-64 (SSE2, no FMA).  It also keeps the
// reference's cost profile (pointer-linked primitives, a std::list traversal
// stack, per-primitive hit records) because the multi-threaded kd-tree
// traversal here is the timed CPU baseline.
//
// Pinning (what checks each function; see tests/):
//   tri_intersect      golden vectors from the reference's own TrianglePrim,
//                      Vector3 and Point3 sources (oracle/_ref), closed forms
//   kd build/intersect structural invariants; result == brute force except
//                      for ties and the split-snap defect class
//   brute force        closed forms on axis-aligned walls; FP64 brute force
//   photon balance     implicit-tree median invariant; libstdc++ nth_element
//   photon kNN         brute-force k-nearest set (tie-aware)
//   radiance estimate  closed form over the returned set
//   photon emission    the draws' uniformity; the direction samplers against
//                      Rodrigues turns and the source's matrix product in
//                      float64; the emission point's square; TransmitRay's
//                      closed forms; the emission loop on scenes with closed
//                      forms (black floor, closed white box, a caustic map
//                      that cannot fill).  Parity with ENCORE's own photons
//                      is statistical only: it draws from rand().
//   kd texture export  the encoding rules read off Kdtree::buildGPU (inner
//                      w = -(axis + 2), leaf x = running triangle count,
//                      vertex rows) and the engine's independent export
#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <list>
#include <thread>
#include <vector>

namespace {

// ---------------------------------------------------------------------------
// Vector arithmetic (Vector3.h:137-162, Vector3.cpp:16-21,105-140, Point3.cpp)
// ---------------------------------------------------------------------------
struct F3 {
    float c[3];
};

// ENC_EPSILON is the double 0.001 (Vector3.h:19); the comparison promotes
// the float component to double (Vector3.cpp:126-140).
inline void flush_small(float& x) {
    if (-0.001 < (double)x && (double)x < 0.001) x = 0.0f;
}

inline F3 sub(const F3& end, const F3& start) {  // Vector3(start, end)
    F3 r;
    r.c[0] = end.c[0] - start.c[0];
    r.c[1] = end.c[1] - start.c[1];
    r.c[2] = end.c[2] - start.c[2];
    return r;
}

inline float dot3(const F3& a, const F3& b) {
    float p = a.c[0] * b.c[0] + a.c[1] * b.c[1];
    return p + a.c[2] * b.c[2];
}

// Cross with the component flush (Vector3.h:148-162).  The y term is
// "-1 * (...)": an exact sign change.
inline F3 cross_flush(const F3& a, const F3& b) {
    F3 r;
    r.c[0] = a.c[1] * b.c[2] - a.c[2] * b.c[1];
    r.c[1] = -1.0f * (a.c[0] * b.c[2] - a.c[2] * b.c[0]);
    r.c[2] = a.c[0] * b.c[1] - a.c[1] * b.c[0];
    flush_small(r.c[0]);
    flush_small(r.c[1]);
    flush_small(r.c[2]);
    return r;
}

// ---------------------------------------------------------------------------
// Triangle predicate (TrianglePrim.cpp:204-249)
// ---------------------------------------------------------------------------
struct TriHit {
    float t, a, b;
    bool hit;
};

inline TriHit tri_test(const F3& o, const F3& d, const F3& v0, const F3& v1, const F3& v2) {
    const F3 e1 = sub(v1, v0);
    const F3 e2 = sub(v2, v0);
    const F3 p = cross_flush(d, e2);
    const float inv = 1.0f / dot3(e1, p);
    const F3 s = sub(o, v0);
    const F3 q = cross_flush(s, e1);
    TriHit h;
    h.a = dot3(s, p) * inv;
    h.b = dot3(d, q) * inv;
    h.t = dot3(e2, q) * inv;
    h.hit = (h.a >= 0.0f) && (h.b >= 0.0f) && (h.a + h.b <= 1.0f) && (h.t > 0.0f);
    return h;
}

// ---------------------------------------------------------------------------
// Scene objects with the reference's indirection (Triangle.h:33-37,
// TrianglePrim.h:22): primitive -> triangle -> three vertex pointers.
// ---------------------------------------------------------------------------
struct Vtx {
    F3 p;
};
struct TriObj {
    const Vtx* v[3];
    uint32_t id;
};

// Hit record of the size and defaults of HitInfo (HitInfo.h:14-28,
// HitInfo.cpp:8-18): hitTime is a double holding a float, default FLT_MAX.
struct HitRec {
    double t;
    float a, b;
    uint32_t prim;
    bool has;
    float pad[8];
};

inline void prim_intersect(const TriObj* tp, const F3& o, const F3& d, HitRec* out) {
    TriHit h = tri_test(o, d, tp->v[0]->p, tp->v[1]->p, tp->v[2]->p);
    if (h.hit) {
        out->t = (double)h.t;
        out->a = h.a;
        out->b = h.b;
        out->prim = tp->id;
        out->has = true;
    }
}

inline void hitrec_reset(HitRec* h) {
    h->t = (double)FLT_MAX;  // POS_INF (Common.h:28)
    h->a = h->b = 0.0f;
    h->prim = 0xFFFFFFFFu;
    h->has = false;
}

struct SceneData {
    std::vector<Vtx> verts;
    std::vector<TriObj> tris;
};

void scene_init(SceneData& s, const float* xyz, uint32_t nv, const uint32_t* idx, uint32_t nt) {
    s.verts.resize(nv);
    for (uint32_t i = 0; i < nv; ++i)
        for (int k = 0; k < 3; ++k) s.verts[i].p.c[k] = xyz[3 * (size_t)i + k];
    s.tris.resize(nt);
    for (uint32_t i = 0; i < nt; ++i) {
        for (int k = 0; k < 3; ++k) s.tris[i].v[k] = &s.verts[idx[3 * (size_t)i + k]];
        s.tris[i].id = i;
    }
}

// ---------------------------------------------------------------------------
// Axis-aligned box stored as corner + extent (AABB.h:177-180, 207-209): every
// max corner is recomputed as pos + size, with its own rounding.
// ---------------------------------------------------------------------------
struct Box {
    F3 pos, size;
};

inline Box box_from_corners(const F3& lo, const F3& hi) {
    Box b;
    b.pos = lo;
    b.size = sub(hi, lo);
    return b;
}

inline F3 box_max(const Box& b) {
    F3 m;
    for (int k = 0; k < 3; ++k) m.c[k] = b.pos.c[k] + b.size.c[k];
    return m;
}

// AABB::Intersect(AABB) with strict inequalities (AABB.cpp:9-18); `self` is
// the voxel, `other` the primitive box.
inline bool box_overlap(const Box& self, const Box& other) {
    const F3 v1 = other.pos, v2 = box_max(other);
    const F3 v3 = self.pos, v4 = box_max(self);
    return (v4.c[0] > v1.c[0]) && (v3.c[0] < v2.c[0]) && (v4.c[1] > v1.c[1]) &&
           (v3.c[1] < v2.c[1]) && (v4.c[2] > v1.c[2]) && (v3.c[2] < v2.c[2]);
}

// TrianglePrim::getAABB (TrianglePrim.cpp:184-202) with dmin/dmax
// (TrianglePrim.h:29-30).
inline Box tri_box(const TriObj& t) {
    F3 lo, hi;
    for (int k = 0; k < 3; ++k) {
        float a = t.v[0]->p.c[k], b = t.v[1]->p.c[k], c = t.v[2]->p.c[k];
        float m = (a > b) ? b : a;
        lo.c[k] = (m > c) ? c : m;
        float M = (a > b) ? a : b;
        hi.c[k] = (M > c) ? M : c;
    }
    return box_from_corners(lo, hi);
}

// AABB::Intersect(Ray&, float&, float&) (AABB.cpp:117-143).  The comparison
// structure is kept so NaN components behave as in the reference.
inline bool box_clip(const Box& b, const F3& o, const F3& d, float& t0out, float& t1out) {
    float t0 = 0.0f, t1 = INFINITY;
    F3 tin, tout;
    for (int k = 0; k < 3; ++k) tin.c[k] = (b.pos.c[k] - o.c[k]) / d.c[k];
    const F3 mx = box_max(b);
    for (int k = 0; k < 3; ++k) tout.c[k] = (mx.c[k] - o.c[k]) / d.c[k];
    for (int k = 0; k < 3; ++k) {
        if (tin.c[k] > tout.c[k]) std::swap(tin.c[k], tout.c[k]);
        t0 = tin.c[k] > t0 ? tin.c[k] : t0;
        t1 = tout.c[k] < t1 ? tout.c[k] : t1;
        if (t0 > t1) return false;
    }
    t0out = t0;
    t1out = t1;
    return true;
}

// ---------------------------------------------------------------------------
// kd-tree (Kdtree.h:25-58, Kdtree.cpp:84-200, 299-680 non-SAH branch)
// ---------------------------------------------------------------------------
// Node word: leaf = (count << 2) | 3; inner = (float bits of split & ~3) | axis
// -- the split and the flags share one union, so the stored split has its two
// low mantissa bits replaced by the axis (Kdtree.h:27-52).
struct KdNode {
    uint32_t word;
    uint32_t child;  // inner: index of left child (right = left + 1); leaf: leaf-list id
};

struct KdTree {
    SceneData scene;
    Box bounds;
    uint32_t max_depth = 0, max_item = 10;
    std::vector<KdNode> nodes;
    std::vector<std::vector<const TriObj*>> leaves;
    uint32_t used = 1;  // m_NodeUsed
    uint32_t deepest = 0;
};

inline uint32_t f2u(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}
inline float u2f(uint32_t u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

void kd_construct(KdTree& T, uint32_t fn, const Box& box, std::vector<uint32_t>& prims,
                  uint32_t depth, uint32_t pre_size, bool dsize, int tries) {
    const uint32_t size = (uint32_t)prims.size();
    if (T.nodes.size() <= fn) T.nodes.resize(fn + 1);
    if (depth > T.deepest) T.deepest = depth;
    // Size-change stall counter (Kdtree.cpp:314-323).
    if (std::abs((int)(pre_size - size)) <= 3) {
        dsize = true;
        tries++;
    } else {
        dsize = false;
        tries = 0;
    }
    if (size <= T.max_item || depth >= T.max_depth || (dsize && tries == 3)) {
        // Leaf (Kdtree.cpp:324-352): primitives keep the parent's order.
        KdNode& n = T.nodes[fn];
        n.word = (size << 2) | 3u;
        n.child = (uint32_t)T.leaves.size();
        std::vector<const TriObj*> lst;
        lst.reserve(size);
        for (uint32_t p : prims) lst.push_back(&T.scene.tris[p]);
        T.leaves.push_back(std::move(lst));
        return;
    }
    // Spatial median on axis depth % 3 (Kdtree.cpp:513-518).
    const int axis = (int)(depth % 3);
    const F3 lo = box.pos, hi = box_max(box);
    const float split = (lo.c[axis] + hi.c[axis]) / 2;
    T.used += 2;
    const uint32_t left = T.used - 2;
    // Child voxels (Kdtree.cpp:547-555).
    F3 lmax = hi;
    lmax.c[axis] = split;
    F3 rmin = lo;
    rmin.c[axis] = split;
    const Box lvox = box_from_corners(lo, lmax);
    const Box rvox = box_from_corners(rmin, hi);
    // Distribution (Kdtree.cpp:561-650): strict overlap; a primitive that
    // overlaps neither side goes to both.
    std::vector<uint32_t> L, R;
    L.reserve(size);
    R.reserve(size);
    for (uint32_t p : prims) {
        const Box pb = tri_box(T.scene.tris[p]);
        const bool inl = box_overlap(lvox, pb);
        const bool inr = box_overlap(rvox, pb);
        if (inl) L.push_back(p);
        if (inr) R.push_back(p);
        if (!inl && !inr) {
            L.push_back(p);
            R.push_back(p);
        }
    }
    {
        KdNode& n = T.nodes[fn];
        n.word = (f2u(split) & ~3u) | (uint32_t)axis;
        n.child = left;
    }
    std::vector<uint32_t>().swap(prims);
    kd_construct(T, left, lvox, L, depth + 1, size, dsize, tries);
    kd_construct(T, left + 1, rvox, R, depth + 1, size, dsize, tries);
}

void kd_build(KdTree& T, const float bmin[3], const float bmax[3]) {
    const uint32_t n = (uint32_t)T.scene.tris.size();
    // Root box: model bounds expanded by 0.001f (Kdtree.cpp:92-117).
    F3 mn, mx, diff;
    for (int k = 0; k < 3; ++k) {
        mn.c[k] = bmin[k];
        mx.c[k] = bmax[k];
        diff.c[k] = 0.001f;
    }
    F3 lo, hi;
    for (int k = 0; k < 3; ++k) {
        lo.c[k] = mn.c[k] - diff.c[k];
        hi.c[k] = mx.c[k] + diff.c[k];
    }
    T.bounds = box_from_corners(lo, hi);
    // Depth and leaf-size limits (Kdtree.cpp:17, 123-131): log of a float
    // argument resolves to the float overload.
    T.max_depth = 11u + (unsigned)(1.3 * (double)std::log((float)n));
    T.max_item = (n <= 5000) ? 2u : 10u;
    T.used = 1;
    T.nodes.clear();
    T.nodes.reserve(n > 0 ? 2 * (size_t)n : 1);
    T.leaves.clear();
    std::vector<uint32_t> all(n);
    for (uint32_t i = 0; i < n; ++i) all[i] = i;
    kd_construct(T, 0, T.bounds, all, 0, 0, false, 0);
}

struct Todo {
    const KdNode* node;
    float tmin, tmax;
};

// Kdtree::intersect (Kdtree.cpp:682-784).
void kd_intersect_one(const KdTree& T, const F3& o, const F3& d, HitRec* out) {
    out->has = false;
    out->t = (double)INFINITY;
    out->prim = 0xFFFFFFFFu;
    out->a = out->b = 0.0f;
    F3 inv;
    for (int k = 0; k < 3; ++k) inv.c[k] = 1.0f / d.c[k];
    float tmin = 0, tmax = 0;
    if (!box_clip(T.bounds, o, d, tmin, tmax)) return;
    std::list<Todo> todo;
    const KdNode* cur = &T.nodes[0];
    for (;;) {
        if (out->t < (double)tmin) break;
        if ((cur->word & 3u) != 3u) {
            const int axis = (int)(cur->word & 3u);
            float split = u2f(cur->word);
            split = ((double)split > -0.001 && (double)split < 0.001) ? 0.0f : split;
            const float tplane = (split - o.c[axis]) * inv.c[axis];
            const KdNode* lo = &T.nodes[cur->child];
            const KdNode* hi = &T.nodes[cur->child + 1];
            const bool below = o.c[axis] <= split;
            const KdNode* first = below ? lo : hi;
            const KdNode* second = below ? hi : lo;
            if (tplane == 0) {
                cur = (inv.c[axis] > 0) ? second : first;
            } else if (tplane >= tmax || tplane < 0) {
                cur = first;
            } else if (tplane <= tmin) {
                cur = second;
            } else {
                Todo t;
                t.node = second;
                t.tmax = tmax;
                t.tmin = tplane;
                todo.push_front(t);
                cur = first;
                tmax = tplane;
            }
        } else {
            const std::vector<const TriObj*>& lst = T.leaves[cur->child];
            const uint32_t cnt = cur->word >> 2;
            for (uint32_t i = 0; i < cnt; ++i) {
                HitRec h;
                hitrec_reset(&h);
                prim_intersect(lst[i], o, d, &h);
                if (h.t < out->t) *out = h;
            }
            if (!todo.empty()) {
                const Todo t = todo.front();
                cur = t.node;
                tmin = t.tmin;
                tmax = t.tmax;
                todo.pop_front();
            } else {
                break;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Photon map (PhotonMap.cpp:21-23, 76-95, 121-297; Photon.h:42-120)
// ---------------------------------------------------------------------------
struct Ph {
    float pos[3];
    float nrm[3];
    float pow[3];
    uint32_t id;
    int plane;
};

struct PhotonMapO {
    std::vector<Ph> ph;
    float bbmin[3], bbmax[3];
};

void pm_balance(PhotonMapO& M, int lo, int hi) {
    if (hi - lo < 0) return;
    // Axis from the bookkeeping box (PhotonMap.cpp:182-191); Z by default.
    const float ex = M.bbmax[0] - M.bbmin[0];
    const float ey = M.bbmax[1] - M.bbmin[1];
    const float ez = M.bbmax[2] - M.bbmin[2];
    int axis = 2;
    if (ex > ey && ey > ez)
        axis = 0;
    else if (ey > ez)
        axis = 1;
    const int med = (hi + lo) / 2;
    std::nth_element(M.ph.begin() + lo, M.ph.begin() + med, M.ph.begin() + hi + 1,
                     [axis](const Ph& a, const Ph& b) { return a.pos[axis] < b.pos[axis]; });
    M.ph[med].plane = axis;
    const float keep_lo = M.ph[lo].pos[axis];
    const float keep_hi = M.ph[hi].pos[axis];
    M.bbmin[axis] = M.ph[med].pos[axis];
    M.bbmax[axis] = M.ph[med].pos[axis];
    pm_balance(M, lo, med - 1);
    M.bbmin[axis] = M.ph[med].pos[axis];
    M.bbmax[axis] = M.ph[med].pos[axis];
    pm_balance(M, med + 1, hi);
    M.bbmin[axis] = keep_lo;
    M.bbmax[axis] = keep_hi;
}

struct DistPair {  // PhotonDistPair (Photon.h:103-120)
    const Ph* p;
    float d;
};
inline bool operator<(const DistPair& a, const DistPair& b) { return a.d < b.d; }

inline bool normal_eq(const float* a, const float* b) {  // Vector3::operator== (Vector3.cpp:89-94)
    return a[0] == b[0] && a[1] == b[1] && a[2] == b[2];
}

// Recursive gather (PhotonMap.cpp:224-297); r2 is passed by value.
void pm_gather(const PhotonMapO& M, int N, float r2, const float* q, const float* qn, int lo, int hi,
               std::vector<DistPair>& heap) {
    if (hi - lo < 0) return;
    const int node = (hi + lo) / 2;
    const Ph* p = &M.ph[node];
    const float x = q[0] - p->pos[0];
    const float y = q[1] - p->pos[1];
    const float z = q[2] - p->pos[2];
    const float d2 = (x * x) + (y * y) + (z * z);
    if (d2 < r2 && normal_eq(p->nrm, qn)) {
        if ((int)heap.size() < N - 1) {
            heap.push_back(DistPair{p, d2});
        } else if ((int)heap.size() == N - 1) {
            heap.push_back(DistPair{p, d2});
            std::make_heap(heap.begin(), heap.end());
            r2 = heap[0].d;
        } else {
            heap.push_back(DistPair{p, d2});
            std::make_heap(heap.begin(), heap.end());
            if ((int)heap.size() > N) {
                std::pop_heap(heap.begin(), heap.end());
                heap.pop_back();
            }
            r2 = heap[0].d;
        }
    }
    const int plane = p->plane;
    const float dp = q[plane] - p->pos[plane];
    if (dp * dp > r2) {
        if (dp > 0)
            pm_gather(M, N, r2, q, qn, node + 1, hi, heap);
        else
            pm_gather(M, N, r2, q, qn, lo, node - 1, heap);
    } else {
        pm_gather(M, N, r2, q, qn, lo, node - 1, heap);
        pm_gather(M, N, r2, q, qn, node + 1, hi, heap);
    }
}

template <class F>
void parallel_chunks(uint32_t n, int nthreads, uint32_t chunk, F&& fn) {
    if (nthreads <= 1) {
        fn(0u, n);
        return;
    }
    std::atomic<uint32_t> next{0};
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; ++t) {
        th.emplace_back([&]() {
            for (;;) {
                uint32_t b = next.fetch_add(chunk);
                if (b >= n) break;
                fn(b, std::min(n, b + chunk));
            }
        });
    }
    for (auto& t : th) t.join();
}

inline void load_ray(const float* r, F3& o, F3& d) {
    o.c[0] = r[0];
    o.c[1] = r[1];
    o.c[2] = r[2];
    d.c[0] = r[4];
    d.c[1] = r[5];
    d.c[2] = r[6];
}

inline void store_hit(const HitRec& h, float* out) {
    uint32_t prim = h.has ? h.prim : 0xFFFFFFFFu;
    out[0] = h.has ? (float)h.t : INFINITY;
    out[1] = h.has ? h.a : 0.0f;
    out[2] = h.has ? h.b : 0.0f;
    std::memcpy(&out[3], &prim, 4);
}

}  // namespace

extern "C" {

// Single predicate evaluation: returns hit flag, writes t, alpha, beta.
int orc_tri_intersect(const float* o, const float* d, const float* v0, const float* v1, const float* v2,
                      float* out3) {
    F3 O, D, A, B, C;
    for (int k = 0; k < 3; ++k) {
        O.c[k] = o[k];
        D.c[k] = d[k];
        A.c[k] = v0[k];
        B.c[k] = v1[k];
        C.c[k] = v2[k];
    }
    TriHit h = tri_test(O, D, A, B, C);
    out3[0] = h.t;
    out3[1] = h.a;
    out3[2] = h.b;
    return h.hit ? 1 : 0;
}

// Batched predicate: rays (n x 8 floats), tris (n x 9 floats), one pair each.
void orc_tri_intersect_pairs(const float* rays, const float* tris, uint32_t n, float* out4) {
    for (uint32_t i = 0; i < n; ++i) {
        F3 o, d, a, b, c;
        load_ray(rays + 8 * (size_t)i, o, d);
        const float* t = tris + 9 * (size_t)i;
        for (int k = 0; k < 3; ++k) {
            a.c[k] = t[k];
            b.c[k] = t[3 + k];
            c.c[k] = t[6 + k];
        }
        TriHit h = tri_test(o, d, a, b, c);
        out4[4 * i + 0] = h.t;
        out4[4 * i + 1] = h.a;
        out4[4 * i + 2] = h.b;
        out4[4 * i + 3] = h.hit ? 1.0f : 0.0f;
    }
}

// Cross with flush and Normalize with flush, for golden-vector checks.
void orc_cross(const float* a, const float* b, float* out) {
    F3 A, B;
    for (int k = 0; k < 3; ++k) {
        A.c[k] = a[k];
        B.c[k] = b[k];
    }
    F3 r = cross_flush(A, B);
    for (int k = 0; k < 3; ++k) out[k] = r.c[k];
}

void orc_normalize(const float* a, float* out) {  // Vector3::Normalize (Vector3.cpp:105-124)
    float x = a[0], y = a[1], z = a[2];
    float m2 = (x * x) + (y * y) + (z * z);
    float m = (float)std::sqrt(m2);
    float r[3] = {x / m, y / m, z / m};
    for (int k = 0; k < 3; ++k) {
        flush_small(r[k]);
        out[k] = r[k];
    }
}

// Brute force over all triangles in list order (BasicAS.cpp:167-194): keeps
// the first hit with 0 < t < best.  Unlike BasicAS, a ray that hits nothing
// reports a miss (BasicAS reports bHasInfo = true with hitTime = FLT_MAX
// because an empty HitInfo's FLT_MAX is below its INFINITY start value).
void orc_brute_closest(const float* xyz, uint32_t nv, const uint32_t* idx, uint32_t nt, const float* rays,
                       uint32_t n, float* hits, int nthreads) {
    SceneData s;
    scene_init(s, xyz, nv, idx, nt);
    parallel_chunks(n, nthreads, 256, [&](uint32_t b, uint32_t e) {
        for (uint32_t i = b; i < e; ++i) {
            F3 o, d;
            load_ray(rays + 8 * (size_t)i, o, d);
            HitRec best;
            hitrec_reset(&best);
            best.t = (double)INFINITY;
            for (const TriObj& t : s.tris) {
                HitRec h;
                hitrec_reset(&h);
                prim_intersect(&t, o, d, &h);
                if (0 < h.t && h.t < best.t && h.has) best = h;
            }
            store_hit(best, hits + 4 * (size_t)i);
        }
    });
}

// Brute-force list of every hit t per ray is handy for tie adjudication:
// returns the number of triangles hit with t < t_limit[i] + tol*t_limit[i].
void orc_count_hits_below(const float* xyz, uint32_t nv, const uint32_t* idx, uint32_t nt, const float* rays,
                          uint32_t n, const float* t_limit, float rel_tol, uint32_t* counts) {
    SceneData s;
    scene_init(s, xyz, nv, idx, nt);
    for (uint32_t i = 0; i < n; ++i) {
        F3 o, d;
        load_ray(rays + 8 * (size_t)i, o, d);
        uint32_t c = 0;
        const double lim = (double)t_limit[i] * (1.0 + (double)rel_tol);
        for (const TriObj& t : s.tris) {
            TriHit h = tri_test(o, d, t.v[0]->p, t.v[1]->p, t.v[2]->p);
            if (h.hit && (double)h.t <= lim) ++c;
        }
        counts[i] = c;
    }
}

// FP64 Moller-Trumbore without the flush, for the error report: evaluates
// triangle prim[i] for ray i; out3 = (t, alpha, beta) in double, NaN if the
// determinant is zero.
void orc_fp64_eval(const float* xyz, const uint32_t* idx, const float* rays, const uint32_t* prim, uint32_t n,
                   double* out3) {
    for (uint32_t i = 0; i < n; ++i) {
        const float* r = rays + 8 * (size_t)i;
        double o[3] = {r[0], r[1], r[2]}, d[3] = {r[4], r[5], r[6]};
        if (prim[i] == 0xFFFFFFFFu) {
            out3[3 * i] = out3[3 * i + 1] = out3[3 * i + 2] = NAN;
            continue;
        }
        double v[3][3];
        for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k) v[j][k] = xyz[3 * (size_t)idx[3 * (size_t)prim[i] + j] + k];
        double e1[3], e2[3], s[3], p[3], q[3];
        for (int k = 0; k < 3; ++k) {
            e1[k] = v[1][k] - v[0][k];
            e2[k] = v[2][k] - v[0][k];
            s[k] = o[k] - v[0][k];
        }
        p[0] = d[1] * e2[2] - d[2] * e2[1];
        p[1] = d[2] * e2[0] - d[0] * e2[2];
        p[2] = d[0] * e2[1] - d[1] * e2[0];
        q[0] = s[1] * e1[2] - s[2] * e1[1];
        q[1] = s[2] * e1[0] - s[0] * e1[2];
        q[2] = s[0] * e1[1] - s[1] * e1[0];
        double det = e1[0] * p[0] + e1[1] * p[1] + e1[2] * p[2];
        if (det == 0.0) {
            out3[3 * i] = out3[3 * i + 1] = out3[3 * i + 2] = NAN;
            continue;
        }
        double inv = 1.0 / det;
        out3[3 * i + 0] = (e2[0] * q[0] + e2[1] * q[1] + e2[2] * q[2]) * inv;
        out3[3 * i + 1] = (s[0] * p[0] + s[1] * p[1] + s[2] * p[2]) * inv;
        out3[3 * i + 2] = (d[0] * q[0] + d[1] * q[1] + d[2] * q[2]) * inv;
    }
}

// FP64 Moller-Trumbore without the flush, with early outs: true and
// (t, alpha, beta) when triangle t is crossed at t > 0.
static inline bool fp64_tri(const float* xyz, const uint32_t* idx, uint32_t t, const double o[3], const double d[3],
                     double& tt, double& a, double& bta) {
    double v[3][3];
    for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3; ++k) v[j][k] = xyz[3 * (size_t)idx[3 * (size_t)t + j] + k];
    double e1[3], e2[3], s[3], p[3], q[3];
    for (int k = 0; k < 3; ++k) {
        e1[k] = v[1][k] - v[0][k];
        e2[k] = v[2][k] - v[0][k];
        s[k] = o[k] - v[0][k];
    }
    p[0] = d[1] * e2[2] - d[2] * e2[1];
    p[1] = d[2] * e2[0] - d[0] * e2[2];
    p[2] = d[0] * e2[1] - d[1] * e2[0];
    double det = e1[0] * p[0] + e1[1] * p[1] + e1[2] * p[2];
    if (det == 0.0) return false;
    double inv = 1.0 / det;
    a = (s[0] * p[0] + s[1] * p[1] + s[2] * p[2]) * inv;
    if (a < 0.0 || a > 1.0) return false;
    q[0] = s[1] * e1[2] - s[2] * e1[1];
    q[1] = s[2] * e1[0] - s[0] * e1[2];
    q[2] = s[0] * e1[1] - s[1] * e1[0];
    bta = (d[0] * q[0] + d[1] * q[1] + d[2] * q[2]) * inv;
    if (bta < 0.0 || a + bta > 1.0) return false;
    tt = (e2[0] * q[0] + e2[1] * q[1] + e2[2] * q[2]) * inv;
    return tt > 0.0;
}

// FP64 brute-force closest hit (no flush): the independent oracle for a ray
// sample.  hits: (t, alpha, beta) double + prim.  Among equal t the lowest
// triangle index wins (list order with a strict <).
void orc_fp64_closest(const float* xyz, uint32_t nv, const uint32_t* idx, uint32_t nt, const float* rays,
                      uint32_t n, double* out3, uint32_t* prim_out, int nthreads) {
    (void)nv;
    parallel_chunks(n, nthreads, 16, [&](uint32_t b, uint32_t e) {
        for (uint32_t i = b; i < e; ++i) {
            const float* r = rays + 8 * (size_t)i;
            double o[3] = {r[0], r[1], r[2]}, d[3] = {r[4], r[5], r[6]};
            double best = INFINITY, ba = 0, bb = 0;
            uint32_t bp = 0xFFFFFFFFu;
            for (uint32_t t = 0; t < nt; ++t) {
                double tt, a, bta;
                if (fp64_tri(xyz, idx, t, o, d, tt, a, bta) && tt < best) {
                    best = tt;
                    ba = a;
                    bb = bta;
                    bp = t;
                }
            }
            out3[3 * i] = best;
            out3[3 * i + 1] = ba;
            out3[3 * i + 2] = bb;
            prim_out[i] = bp;
        }
    });
}

// The same FP64 closest hit over every triangle, with the triangles culled by
// a bounding-volume hierarchy so that 1% samples of million-triangle scenes
// finish in seconds.  Boxes are the triangles' vertex bounds in double,
// widened by 1e-6 of their magnitude and extent, so a ray that the FP64 test
// accepts always enters every box around the triangle; a box is skipped only
// when its entry distance exceeds the best t.  Ties keep the lowest triangle
// index, so the output is exactly orc_fp64_closest's (tests check this).
void orc_fp64_closest_culled(const float* xyz, uint32_t nv, const uint32_t* idx, uint32_t nt, const float* rays,
                             uint32_t n, double* out3, uint32_t* prim_out, int nthreads) {
    (void)nv;
    struct CNode {
        double lo[3], hi[3];
        uint32_t left, first, count;  // count > 0: leaf of tris[first, first + count)
    };
    std::vector<uint32_t> tris(nt);
    std::vector<double> cen(3 * (size_t)nt), blo(3 * (size_t)nt), bhi(3 * (size_t)nt);
    for (uint32_t t = 0; t < nt; ++t) {
        tris[t] = t;
        for (int k = 0; k < 3; ++k) {
            double lo = INFINITY, hi = -INFINITY;
            for (int j = 0; j < 3; ++j) {
                const double c = xyz[3 * (size_t)idx[3 * (size_t)t + j] + k];
                lo = std::min(lo, c);
                hi = std::max(hi, c);
            }
            const double pad = 1e-6 * (std::max(std::fabs(lo), std::fabs(hi)) + (hi - lo)) + 1e-30;
            blo[3 * (size_t)t + k] = lo - pad;
            bhi[3 * (size_t)t + k] = hi + pad;
            cen[3 * (size_t)t + k] = 0.5 * (lo + hi);
        }
    }
    std::vector<CNode> nodes;
    nodes.reserve(nt ? 2 * (size_t)nt / 4 + 8 : 1);
    // Median split on the longest centroid axis, leaves of up to 8.
    struct Todo {
        uint32_t node, first, count;
    };
    std::vector<Todo> todo;
    nodes.push_back(CNode{});
    todo.push_back({0, 0, nt});
    while (!todo.empty()) {
        const Todo w = todo.back();
        todo.pop_back();
        CNode nd{};
        for (int k = 0; k < 3; ++k) {
            nd.lo[k] = INFINITY;
            nd.hi[k] = -INFINITY;
        }
        double clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (uint32_t j = w.first; j < w.first + w.count; ++j) {
            const uint32_t t = tris[j];
            for (int k = 0; k < 3; ++k) {
                nd.lo[k] = std::min(nd.lo[k], blo[3 * (size_t)t + k]);
                nd.hi[k] = std::max(nd.hi[k], bhi[3 * (size_t)t + k]);
                clo[k] = std::min(clo[k], cen[3 * (size_t)t + k]);
                chi[k] = std::max(chi[k], cen[3 * (size_t)t + k]);
            }
        }
        if (w.count <= 8) {
            nd.first = w.first;
            nd.count = w.count;
            nodes[w.node] = nd;
            continue;
        }
        int ax = 0;
        for (int k = 1; k < 3; ++k)
            if (chi[k] - clo[k] > chi[ax] - clo[ax]) ax = k;
        const uint32_t half = w.count / 2;
        std::nth_element(tris.begin() + w.first, tris.begin() + w.first + half, tris.begin() + w.first + w.count,
                         [&](uint32_t a, uint32_t b) {
                             const double ca = cen[3 * (size_t)a + ax], cb = cen[3 * (size_t)b + ax];
                             return ca < cb || (ca == cb && a < b);
                         });
        nd.left = (uint32_t)nodes.size();
        nd.count = 0;
        nodes[w.node] = nd;
        nodes.push_back(CNode{});
        nodes.push_back(CNode{});
        todo.push_back({nd.left, w.first, half});
        todo.push_back({nd.left + 1, w.first + half, w.count - half});
    }
    auto enter = [](const CNode& b, const double o[3], const double d[3], double& tin) {
        double t0 = 0.0, t1 = INFINITY;
        for (int k = 0; k < 3; ++k) {
            if (d[k] == 0.0) {
                if (o[k] < b.lo[k] || o[k] > b.hi[k]) return false;
                continue;
            }
            double a = (b.lo[k] - o[k]) / d[k], c = (b.hi[k] - o[k]) / d[k];
            if (a > c) std::swap(a, c);
            t0 = std::max(t0, a);
            t1 = std::min(t1, c);
        }
        tin = t0;
        return t0 <= t1;
    };
    parallel_chunks(n, nthreads, 64, [&](uint32_t b, uint32_t e) {
        std::vector<uint32_t> stack;
        for (uint32_t i = b; i < e; ++i) {
            const float* r = rays + 8 * (size_t)i;
            double o[3] = {r[0], r[1], r[2]}, d[3] = {r[4], r[5], r[6]};
            double best = INFINITY, ba = 0, bb = 0;
            uint32_t bp = 0xFFFFFFFFu;
            stack.assign(1, 0u);
            while (nt && !stack.empty()) {
                const CNode& nd = nodes[stack.back()];
                stack.pop_back();
                double tin;
                if (!enter(nd, o, d, tin) || tin > best) continue;
                if (nd.count) {
                    for (uint32_t j = nd.first; j < nd.first + nd.count; ++j) {
                        const uint32_t t = tris[j];
                        double tt, a, bta;
                        if (fp64_tri(xyz, idx, t, o, d, tt, a, bta) && (tt < best || (tt == best && t < bp))) {
                            best = tt;
                            ba = a;
                            bb = bta;
                            bp = t;
                        }
                    }
                } else {
                    stack.push_back(nd.left);
                    stack.push_back(nd.left + 1);
                }
            }
            out3[3 * i] = best;
            out3[3 * i + 1] = ba;
            out3[3 * i + 2] = bb;
            prim_out[i] = bp;
        }
    });
}

// ----- kd-tree -------------------------------------------------------------
void* orc_kd_build(const float* xyz, uint32_t nv, const uint32_t* idx, uint32_t nt, const float* bmin,
                   const float* bmax) {
    KdTree* T = new KdTree();
    scene_init(T->scene, xyz, nv, idx, nt);
    kd_build(*T, bmin, bmax);
    return T;
}

void orc_kd_free(void* h) { delete static_cast<KdTree*>(h); }

// stats: [nodes, leaves, refs, deepest, max_depth, max_item]
void orc_kd_stats(void* h, uint32_t* out6) {
    KdTree* T = static_cast<KdTree*>(h);
    uint64_t refs = 0;
    for (auto& l : T->leaves) refs += l.size();
    out6[0] = T->used;
    out6[1] = (uint32_t)T->leaves.size();
    out6[2] = (uint32_t)refs;
    out6[3] = T->deepest;
    out6[4] = T->max_depth;
    out6[5] = T->max_item;
}

// Export in the reference's own numbering: node words and child fields
// (leaf child = leaf-list id), plus leaf lists concatenated in leaf-id order
// with offsets.  Sizes from orc_kd_stats.  root_box = pos[3], size[3].
void orc_kd_export(void* h, uint32_t* words, uint32_t* childs, uint32_t* leaf_off, uint32_t* leaf_prims,
                   float* root_box) {
    KdTree* T = static_cast<KdTree*>(h);
    for (uint32_t i = 0; i < T->used; ++i) {
        words[i] = T->nodes[i].word;
        childs[i] = T->nodes[i].child;
    }
    uint32_t off = 0;
    for (size_t l = 0; l < T->leaves.size(); ++l) {
        leaf_off[l] = off;
        for (const TriObj* t : T->leaves[l]) leaf_prims[off++] = t->id;
    }
    leaf_off[T->leaves.size()] = off;
    for (int k = 0; k < 3; ++k) {
        root_box[k] = T->bounds.pos.c[k];
        root_box[3 + k] = T->bounds.size.c[k];
    }
}

// Multi-threaded closest hit through the kd-tree; threads pull 1024-ray
// chunks (Kdtree::intersect only reads shared state).
void orc_kd_intersect(void* h, const float* rays, uint32_t n, float* hits, int nthreads) {
    const KdTree* T = static_cast<const KdTree*>(h);
    parallel_chunks(n, nthreads, 1024, [&](uint32_t b, uint32_t e) {
        for (uint32_t i = b; i < e; ++i) {
            F3 o, d;
            load_ray(rays + 8 * (size_t)i, o, d);
            HitRec r;
            kd_intersect_one(*T, o, d, &r);
            store_hit(r, hits + 4 * (size_t)i);
        }
    });
}

// Shadow query with ENCORE's semantics (PhotonMapperCPU.cpp:1192-1207):
// closest hit from the ray's origin compared with tmax (the caller sets
// tmax = |L - p|); occluded = has hit && t < tmax.
void orc_kd_occluded(void* h, const float* rays, uint32_t n, uint32_t* occ, int nthreads) {
    const KdTree* T = static_cast<const KdTree*>(h);
    parallel_chunks(n, nthreads, 1024, [&](uint32_t b, uint32_t e) {
        for (uint32_t i = b; i < e; ++i) {
            F3 o, d;
            load_ray(rays + 8 * (size_t)i, o, d);
            HitRec r;
            kd_intersect_one(*T, o, d, &r);
            occ[i] = (r.has && r.t < (double)rays[8 * (size_t)i + 7]) ? 1u : 0u;
        }
    });
}

// ----- photon map ------------------------------------------------------------
// pos/nrm/pow: n x 3 floats in insertion order (PhotonMap::AddPhoton).
void* orc_pm_build(const float* pos, const float* nrm, const float* pw, uint32_t n) {
    PhotonMapO* M = new PhotonMapO();
    M->ph.resize(n);
    for (uint32_t i = 0; i < n; ++i) {
        for (int k = 0; k < 3; ++k) {
            M->ph[i].pos[k] = pos[3 * (size_t)i + k];
            M->ph[i].nrm[k] = nrm[3 * (size_t)i + k];
            M->ph[i].pow[k] = pw ? pw[3 * (size_t)i + k] : 0.0f;
        }
        M->ph[i].id = i;
        M->ph[i].plane = -1;
    }
    // Bookkeeping box starts at +-1e10 (PhotonMap.cpp:21, 25-29).
    const float big = 1.0e10f;
    for (int k = 0; k < 3; ++k) {
        M->bbmin[k] = -big;
        M->bbmax[k] = big;
    }
    if (n > 0) pm_balance(*M, 0, (int)n - 1);  // BalanceAndConstruct (:167-175)
    return M;
}

void orc_pm_free(void* h) { delete static_cast<PhotonMapO*>(h); }

// Balanced order: perm[slot] = insertion index, axis[slot] = plane.
void orc_pm_order(void* h, uint32_t* perm, uint8_t* axis) {
    PhotonMapO* M = static_cast<PhotonMapO*>(h);
    for (size_t i = 0; i < M->ph.size(); ++i) {
        perm[i] = M->ph[i].id;
        axis[i] = (uint8_t)M->ph[i].plane;
    }
}

// GetNearestNPhotons (PhotonMap.cpp:76-95): for each query, up to k slots
// (balanced-order indices) and squared distances in the final heap order
// (front = largest), count in cnt[i].  r2 = fl32(max_dist * max_dist).
void orc_pm_knn(void* h, const float* qpos, const float* qnrm, uint32_t nq, int k, float max_dist,
                uint32_t* slots, float* d2, uint32_t* cnt, int nthreads) {
    const PhotonMapO* M = static_cast<const PhotonMapO*>(h);
    parallel_chunks(nq, nthreads, 256, [&](uint32_t b, uint32_t e) {
        std::vector<DistPair> heap;
        for (uint32_t i = b; i < e; ++i) {
            heap.clear();
            heap.reserve(k);
            if (!M->ph.empty() && k > 0) {
                const float r2 = max_dist * max_dist;
                pm_gather(*M, k, r2, qpos + 3 * (size_t)i, qnrm + 3 * (size_t)i, 0, (int)M->ph.size() - 1,
                          heap);
                std::make_heap(heap.begin(), heap.end());
            }
            cnt[i] = (uint32_t)heap.size();
            for (size_t j = 0; j < heap.size(); ++j) {
                slots[(size_t)i * k + j] = (uint32_t)(heap[j].p - M->ph.data());
                d2[(size_t)i * k + j] = heap[j].d;
            }
        }
    });
}

// GetRadianceEstimate (PhotonMap.cpp:121-164): power summed over the gathered
// set in heap order; caustic maps weight by (r - d)/d with r, d squared
// distances, then scale by N / (PI r^2) with PI = 3.14159f (Common.h:36).
void orc_pm_radiance(void* h, const float* qpos, const float* qnrm, uint32_t nq, int k, float max_dist,
                     int caustic, float* rgb, int nthreads) {
    const PhotonMapO* M = static_cast<const PhotonMapO*>(h);
    const float PI_F = 3.14159f;
    parallel_chunks(nq, nthreads, 256, [&](uint32_t b, uint32_t e) {
        std::vector<DistPair> heap;
        for (uint32_t i = b; i < e; ++i) {
            heap.clear();
            if (!M->ph.empty() && k > 0) {
                pm_gather(*M, k, max_dist * max_dist, qpos + 3 * (size_t)i, qnrm + 3 * (size_t)i, 0,
                          (int)M->ph.size() - 1, heap);
                std::make_heap(heap.begin(), heap.end());
            }
            float R = 0, G = 0, B = 0;
            if (!heap.empty()) {
                const double r = heap.front().d;
                for (size_t j = 0; j < heap.size(); ++j) {
                    const Ph* p = heap[j].p;
                    const double radius = heap[j].d;
                    const double w = caustic ? (r - radius) / radius : 1;
                    R = (float)((double)R + (double)p->pow[0] * w);
                    G = (float)((double)G + (double)p->pow[1] * w);
                    B = (float)((double)B + (double)p->pow[2] * w);
                }
                const float dens = caustic ? (float)(k / ((double)PI_F * r * r)) : 1.0f;
                R *= dens;
                G *= dens;
                B *= dens;
            }
            rgb[3 * (size_t)i] = R;
            rgb[3 * (size_t)i + 1] = G;
            rgb[3 * (size_t)i + 2] = B;
        }
    });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Photon emission (PhotonMapperCPU.cpp:503-835), one path at a time as the
// reference runs it: EmitPhotons' two loops, TracePhoton's bounce loop with
// its Full() checks, ReflectPhotonDiffusely, EmitShadowPhoton, TransmitPhoton
// and TransmitRay.  rand() is replaced by a counter-based draw of (seed, path,
// slot) reduced modulo 10000 as randNumber reduces rand() (Common.cpp:17-21):
//   slot 0 light (rand() % lights), 1-2 emission point, 3-4 emission
//   direction, 5+3b interaction draw at bounce b, 6+3b and 7+3b the diffuse
//   direction.
// The direction samplers (randDirectionN / randDirectionSuperN,
// Common.cpp:98-182) keep their formulas with cos(acos(x)) = x,
// sin(acos(x)) = sqrt(1 - x^2), and the cosine and sine of 2*PI*e2 from
// double-precision libm rounded to float.  GetRotationMatrix
// (TAffineMatrix.h:160-200) keeps its unnormalised axis.
// ---------------------------------------------------------------------------
namespace {

struct OLight {  // layout of rq_rect_light
    float center[3], normal[3], length, width, power[3];
};
struct OMat {  // layout of rq_material
    float diffuse[3], transmittance, transparency, index;
};

inline F3 f3(float x, float y, float z) {
    F3 r;
    r.c[0] = x;
    r.c[1] = y;
    r.c[2] = z;
    return r;
}
inline F3 add3(const F3& a, const F3& b) { return f3(a.c[0] + b.c[0], a.c[1] + b.c[1], a.c[2] + b.c[2]); }
inline F3 mul3(const F3& a, float s) { return f3(a.c[0] * s, a.c[1] * s, a.c[2] * s); }
inline F3 norm3(const F3& a) {
    float r[3];
    orc_normalize(a.c, r);
    return f3(r[0], r[1], r[2]);
}
inline bool eq3(const F3& a, const F3& b) { return a.c[0] == b.c[0] && a.c[1] == b.c[1] && a.c[2] == b.c[2]; }
inline F3 neg3(const F3& a) { return f3(-a.c[0], -a.c[1], -a.c[2]); }

uint32_t ph_rand(uint64_t seed, uint64_t path, uint32_t slot) {
    uint64_t z = seed + (path * 64ull + slot + 1ull) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (uint32_t)(z >> 33);
}

float rand_number(float lo, float hi, uint32_t r) {  // randNumber (Common.cpp:17-21)
    return (hi - lo) * ((float)(int)(r % 10000u) / (float)10000) + lo;
}

// TAffineMatrix::GetRotationMatrix(axis, acos(c)) (TAffineMatrix.h:160-200)
// applied to v: M = R^T X R with rows r = axis, s, t.
F3 rotate_about(const F3& r, float c, const F3& v) {
    const float ax = std::fabs(r.c[0]), ay = std::fabs(r.c[1]), az = std::fabs(r.c[2]);
    F3 s;
    if (ax <= ay && ax <= az)
        s = f3(0.0f, -r.c[2], r.c[1]);
    else if (ay <= ax && ay <= az)
        s = f3(-r.c[2], 0.0f, r.c[0]);
    else
        s = f3(-r.c[1], r.c[0], 0.0f);
    s = norm3(s);
    const F3 t = norm3(cross_flush(r, s));
    const float sa = std::sqrt((1.0f - c) * (1.0f + c));
    const float w0 = dot3(r, v), w1 = dot3(s, v), w2 = dot3(t, v);
    const float u1 = c * w1 - sa * w2, u2 = sa * w1 + c * w2;
    F3 o;
    for (int k = 0; k < 3; ++k) o.c[k] = (r.c[k] * w0 + s.c[k] * u1) + t.c[k] * u2;
    return o;
}

// randDirectionN (e1 from 0.01) / randDirectionSuperN (e1 from 0.001).
F3 rand_dir_about(F3 n, float e1_lo, uint32_t r1, uint32_t r2) {
    const float e1 = rand_number(e1_lo, 1.0f, r1);
    const float e2 = rand_number(0.0f, 1.0f, r2);
    const float phi = (2 * 3.14159f) * e2;
    const float cp = (float)std::cos((double)phi), sp = (float)std::sin((double)phi);
    const float ct = std::sqrt(e1), st = std::sqrt(1.0f - e1);
    const F3 d = f3(st * cp, st * sp, ct);
    n = norm3(n);
    const F3 z = f3(0.0f, 0.0f, 1.0f);
    if (eq3(z, n)) return d;
    if (eq3(z, neg3(n))) return neg3(d);
    return norm3(rotate_about(cross_flush(z, n), dot3(n, z), d));
}

// RectangleLight::EmitPhoton's origin (RectangleLight.cpp:172-183) through
// Sampler::GetSamplePointsForRect (Sampler.cpp:49-95, 154-218) with one cell
// and full jitter.
F3 light_point(const OLight& l, uint32_t r1, uint32_t r2) {
    const float half = 1.0f / 2.0f;
    const float jx = rand_number(-half, half, r1) * 1.0f;
    const float jz = rand_number(-half, half, r2) * 1.0f;
    F3 p = f3(half + jx, 0.0f, half + jz);
    p = add3(f3(-0.5f, 0.0f, -0.5f), p);
    p.c[0] *= l.length / 1.0f;
    p.c[2] *= l.width / 1.0f;
    const F3 n = norm3(f3(l.normal[0], l.normal[1], l.normal[2]));
    const F3 up = f3(0.0f, 1.0f, 0.0f);
    if (eq3(up, n)) {
    } else if (eq3(up, neg3(n))) {
        p = add3(p, mul3(p, -2.0f));
    } else {
        const float c = dot3(up, n);
        if (c != 1.0f) p = rotate_about(norm3(cross_flush(up, n)), c, p);  // theta = acos(c) != 0
    }
    p = add3(f3(l.center[0], l.center[1], l.center[2]), p);
    return add3(mul3(f3(l.normal[0], l.normal[1], l.normal[2]), 0.01f), p);
}

struct OHit {
    bool has;
    uint32_t prim;
    F3 P, N;
    bool entering;
};

OHit intersect(const KdTree& T, const F3& o, const F3& d) {
    HitRec h;
    kd_intersect_one(T, o, d, &h);
    OHit r;
    r.has = h.has;
    r.prim = h.prim;
    if (!h.has) return r;
    const TriObj& t = T.scene.tris[h.prim];
    const F3 v0 = t.v[0]->p, e1 = sub(t.v[1]->p, v0), e2 = sub(t.v[2]->p, v0);
    for (int k = 0; k < 3; ++k) r.P.c[k] = (v0.c[k] + e1.c[k] * h.a) + e2.c[k] * h.b;  // TrianglePrim.cpp:239
    r.N = norm3(cross_flush(e1, e2));                                                  // :236-238
    r.entering = dot3(d, r.N) < 0;                                                      // :242
    return r;
}

struct OMap {
    uint32_t cap, n = 0;
    float *pos, *nrm, *pow, *dir;
    uint8_t* shadow;
    uint32_t* path;
    bool full() const { return n == cap; }  // PhotonMap::Full (PhotonMap.h:54)
    void add(const F3& P, const F3& N, const float* power, bool neg, const F3& incoming, bool sh, uint64_t gp) {
        if (n >= cap) return;  // PhotonMap::AddPhoton (PhotonMap.cpp:46-53)
        for (int k = 0; k < 3; ++k) {
            pos[3 * (size_t)n + k] = P.c[k];
            nrm[3 * (size_t)n + k] = N.c[k];
            pow[3 * (size_t)n + k] = neg ? -power[k] : power[k];
            dir[3 * (size_t)n + k] = incoming.c[k];
        }
        shadow[n] = sh ? 1 : 0;
        path[n] = (uint32_t)gp;
        ++n;
    }
};

constexpr uint32_t kShadowPasses = 8;

struct OEmit {
    const KdTree* T;
    const OMat* mats;
    const uint32_t* prim_mat;
    const OLight* lights;
    uint32_t nlights, max_bounces;
    uint64_t seed;
    OMap* maps;  // global, direct, caustic
};

// TransmitRay (PhotonMapperCPU.cpp:800-835).
bool transmit_ray(const F3& d, const OHit& h, const OMat& m, F3& oo, F3& od) {
    if (m.transparency < 0.6) return false;  // caller keeps its ray
    float c1, c2;
    const float air = 0.99f;
    if (h.entering) {
        c1 = air;
        c2 = m.index;
    } else {
        c2 = air;
        c1 = m.index;
    }
    const float cr = (c1 == 0) ? (float)(c2 / 0.0001) : (c2 / c1);
    const float nd = dot3(h.N, d);
    const float th2 = std::sqrt(1 - (cr * cr) * (1 - (nd * nd)));
    const F3 T = norm3(add3(mul3(d, cr), mul3(h.N, cr * nd - th2)));
    od = T;
    oo = add3(h.P, mul3(T, 0.01f));
    return true;
}

void trace_path(const OEmit& E, uint64_t gp) {
    const OLight& l = E.lights[ph_rand(E.seed, gp, 0) % E.nlights];
    const F3 origin = light_point(l, ph_rand(E.seed, gp, 1), ph_rand(E.seed, gp, 2));
    const F3 n = f3(l.normal[0], l.normal[1], l.normal[2]);
    const F3 dir0 = rand_dir_about(n, 0.01f, ph_rand(E.seed, gp, 3), ph_rand(E.seed, gp, 4));
    // TracePhoton (PhotonMapperCPU.cpp:603-655)
    F3 ro = add3(origin, mul3(dir0, 0.01f)), rd = dir0;
    bool specular = false;  // LastReflection: NONE
    for (uint32_t b = 0; b <= E.max_bounces; ++b) {
        const OHit h = intersect(*E.T, ro, rd);
        if (!h.has) break;
        const OMat& m = E.mats[E.prim_mat[h.prim]];
        const float pS = 0;
        const float pD = (m.diffuse[0] + m.diffuse[1] + m.diffuse[2]) / 3.0f;
        const float pT = m.transmittance;
        const float pTotal = pS + pD + pT;
        if (!(pTotal <= 1.0f)) continue;  // nothing happens at this hit
        const float Ed = rand_number(0.0f, 1.0f, ph_rand(E.seed, gp, 5 + 3 * b));
        if (Ed < pS) {
        } else if (Ed < pS + pD || !(Ed < pTotal)) {
            // ReflectPhotonDiffusely (PhotonMapperCPU.cpp:700-745)
            if (specular) {
                if (!E.maps[2].full()) E.maps[2].add(h.P, h.N, l.power, false, dir0, false, gp);
            } else if (b > 0 && !E.maps[0].full()) {
                E.maps[0].add(h.P, h.N, l.power, false, dir0, false, gp);
            } else if (b == 0 && !E.maps[1].full()) {
                E.maps[1].add(h.P, h.N, l.power, false, dir0, false, gp);
                // EmitShadowPhoton (PhotonMapperCPU.cpp:775-797)
                // The chain gives up after kShadowPasses surfaces left: a
                // phantom hit whose hit point lies behind the ray (Cross()'s
                // flush) returns the same point for ever, and the reference's
                // loop never ends there.
                F3 so = add3(h.P, mul3(rd, 2.0f));
                OHit sh = intersect(*E.T, so, rd);
                uint32_t passes = 0;
                while (sh.has && !sh.entering) {
                    if (++passes == kShadowPasses) {
                        sh.has = false;
                        break;
                    }
                    so = add3(sh.P, mul3(rd, 0.01f));
                    sh = intersect(*E.T, so, rd);
                }
                if (sh.has) E.maps[1].add(sh.P, sh.N, l.power, true, rd, true, gp);
            }
            specular = false;
            const F3 nd = rand_dir_about(h.N, 0.001f, ph_rand(E.seed, gp, 6 + 3 * b), ph_rand(E.seed, gp, 7 + 3 * b));
            ro = add3(h.P, mul3(nd, 0.001f));
            rd = nd;
            if (!(Ed < pS + pD)) break;  // absorbed
        } else {
            // TransmitPhoton (PhotonMapperCPU.cpp:747-773)
            F3 to, td;
            if (!transmit_ray(rd, h, m, to, td)) continue;
            const OHit h2 = intersect(*E.T, to, td);
            if (h2.has && !h2.entering) {
                const OMat& m2 = E.mats[E.prim_mat[h2.prim]];
                if (!transmit_ray(td, h2, m2, ro, rd)) {
                    ro = to;
                    rd = td;
                }
            }
            specular = true;
        }
    }
}

}  // namespace

extern "C" {

// Pieces of the emission for direct checks: n draws of the direction sampler
// about `nrm` (e1 from e1_lo) and of the light's emission point, for draw
// values r1[i], r2[i]; and the counter-based draw itself.
void orc_photon_dir(const float* nrm, float e1_lo, const uint32_t* r1, const uint32_t* r2, uint32_t n, float* out) {
    for (uint32_t i = 0; i < n; ++i) {
        const F3 d = rand_dir_about(f3(nrm[0], nrm[1], nrm[2]), e1_lo, r1[i], r2[i]);
        for (int k = 0; k < 3; ++k) out[3 * (size_t)i + k] = d.c[k];
    }
}

void orc_light_point(const float* light, const uint32_t* r1, const uint32_t* r2, uint32_t n, float* out) {
    OLight l;
    std::memcpy(&l, light, sizeof(OLight));
    for (uint32_t i = 0; i < n; ++i) {
        const F3 p = light_point(l, r1[i], r2[i]);
        for (int k = 0; k < 3; ++k) out[3 * (size_t)i + k] = p.c[k];
    }
}

void orc_ph_rand(uint64_t seed, const uint64_t* path, const uint32_t* slot, uint32_t n, uint32_t* out) {
    for (uint32_t i = 0; i < n; ++i) out[i] = ph_rand(seed, path[i], slot[i]);
}

// TransmitRay for one hit: d, P, N, entering, material row; returns 0 when the
// surface does not refract, else writes origin and direction.
int orc_transmit_ray(const float* d, const float* P, const float* N, int entering, const float* mat, float* out6) {
    OHit h;
    h.has = true;
    h.prim = 0;
    h.P = f3(P[0], P[1], P[2]);
    h.N = f3(N[0], N[1], N[2]);
    h.entering = entering != 0;
    OMat m;
    std::memcpy(&m, mat, sizeof(OMat));
    F3 o, t;
    if (!transmit_ray(f3(d[0], d[1], d[2]), h, m, o, t)) return 0;
    for (int k = 0; k < 3; ++k) {
        out6[k] = o.c[k];
        out6[3 + k] = t.c[k];
    }
    return 1;
}

// maps: 3 x {pos, nrm, pow, dir (float*), shadow (uint8*), path (uint32*)}
// with capacity[3]; counts[3] and paths out.  EmitPhotons (PhotonMapperCPU.cpp:
// 537-585): paths until the global map is full, then until the caustic map is
// full (bounded by max_paths), then ScalePhotons on each map.
void orc_emit_photons(void* kd, const void* mats, const uint32_t* prim_mat, const void* lights, uint32_t nlights,
                      uint32_t max_bounces, const uint32_t* capacity, const float* scale, uint64_t seed,
                      uint64_t max_paths, void** bufs, uint32_t* counts, uint64_t* paths_out) {
    OMap maps[3];
    for (int m = 0; m < 3; ++m) {
        maps[m].cap = capacity[m];
        maps[m].pos = static_cast<float*>(bufs[6 * m]);
        maps[m].nrm = static_cast<float*>(bufs[6 * m + 1]);
        maps[m].pow = static_cast<float*>(bufs[6 * m + 2]);
        maps[m].dir = static_cast<float*>(bufs[6 * m + 3]);
        maps[m].shadow = static_cast<uint8_t*>(bufs[6 * m + 4]);
        maps[m].path = static_cast<uint32_t*>(bufs[6 * m + 5]);
    }
    OEmit E;
    E.T = static_cast<const KdTree*>(kd);
    E.mats = static_cast<const OMat*>(mats);
    E.prim_mat = prim_mat;
    E.lights = static_cast<const OLight*>(lights);
    E.nlights = nlights;
    E.max_bounces = max_bounces;
    E.seed = seed;
    E.maps = maps;
    uint64_t gp = 0;
    while (!maps[0].full() && gp < max_paths) trace_path(E, gp++);
    while (!maps[2].full() && gp < max_paths) trace_path(E, gp++);
    for (int m = 0; m < 3; ++m) {
        for (uint32_t i = 0; i < maps[m].n; ++i)
            for (int k = 0; k < 3; ++k) maps[m].pow[3 * (size_t)i + k] *= scale[m];  // ScalePhotons
        counts[m] = maps[m].n;
    }
    *paths_out = gp;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Kdtree::buildGPU (Kdtree.cpp:876-1000): the float4 texture encoding of the
// tree, node by node, and the vertex arrays of the leaves' triangles.
// ---------------------------------------------------------------------------
extern "C" void orc_kd_textures(void* h, float* cell, float* v0, float* v1, float* v2, uint32_t* ref_prim) {
    const KdTree* T = static_cast<const KdTree*>(h);
    float currentTriangleIndex = 0;
    size_t r = 0;
    float* vs[3] = {v0, v1, v2};
    for (uint32_t i = 0; i < T->used; ++i) {
        const KdNode& nd = T->nodes[i];
        float* c = cell + 4 * (size_t)i;
        if ((nd.word & 3u) != 3u) {
            c[0] = (float)nd.child;  // leftChild
            c[1] = u2f(nd.word);     // SplitPos()
            c[2] = 0;
            c[3] = (float)-((int)(nd.word & 3u) + 2);
        } else {
            const int ntri = (int)(nd.word >> 2);
            c[0] = currentTriangleIndex;
            c[1] = 0;
            c[2] = 0;
            c[3] = (float)ntri;
            const std::vector<const TriObj*>& lst = T->leaves[nd.child];
            for (int ti = 0; ti < ntri; ++ti, ++r) {
                currentTriangleIndex += (size_t)1;
                ref_prim[r] = lst[ti]->id;
                for (int k = 0; k < 3; ++k) {
                    for (int a = 0; a < 3; ++a) vs[k][4 * r + a] = lst[ti]->v[k]->p.c[a];
                    vs[k][4 * r + 3] = 0.0f;
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Irradiance cache (IrrCacheOctree.h:92-326) and the final gather's use of it
// (PhotonMapperCPU.cpp:1087-1103, 1170-1173).
//
// AddElement (IrrCacheOctree.h:212-243) drops an element outside the root
// box and stores it in the root's list when 2 * (a * Ri) <= the root's x
// size.  Otherwise it descends into a child of half the size, where the same
// test fails again, and never stops: in any run of the reference that
// completes, every element sits in the root's list, in insertion order, and
// the root stays a leaf.  Such an element is refused here (status 2).
// GetIrrValuesNearPoint (:294-326) then walks that list: an element counts
// when GetWeight (:55-77, Jensen eq. 11.5) exceeds 1 / a and SampleInFront
// (:279-290) is false; the caller sums value * weight and the weights in list
// order and divides (PhotonMapperCPU.cpp:1093-1101).  Every expression keeps
// the reference's float/double types and operation order.
// ---------------------------------------------------------------------------
namespace {

struct IrrElem {
    F3 value, pos, nrm;
    float ri;
};

struct IrrCacheO {
    float a;
    F3 minp, size;
    std::vector<IrrElem> root;
};

inline float mag3(const F3& v) {  // Vector3::Magnitude (Vector3.cpp:116-124)
    const float m2 = (v.c[0] * v.c[0]) + (v.c[1] * v.c[1]) + (v.c[2] * v.c[2]);
    return std::sqrt(m2);
}

// TCacheElement::GetWeight (IrrCacheOctree.h:55-77).
inline float irr_weight(const IrrElem& e, const F3& p, const F3& n) {
    const float distance = mag3(sub(e.pos, p));
    const float distCalc = (e.ri == 0) ? 0.0f : distance / e.ri;
    const bool same = n.c[0] == e.nrm.c[0] && n.c[1] == e.nrm.c[1] && n.c[2] == e.nrm.c[2];
    const float sqrtCalc = same ? 0.0f : (float)std::sqrt(1.0 - (double)dot3(n, e.nrm));
    const float denom = sqrtCalc + distCalc;
    return denom == 0 ? 0.0f : 1 / denom;
}

// SampleInFront (IrrCacheOctree.h:279-290): navg = sampleN + n, pVec =
// Normalize(p - sampleP) with the flush, Dot(pVec, navg) < -0.1f.
inline bool irr_in_front(const F3& p, const F3& sp, const F3& n, const F3& sn) {
    F3 navg;
    for (int k = 0; k < 3; ++k) navg.c[k] = sn.c[k] + n.c[k];
    F3 v = sub(p, sp);
    const float m = mag3(v);
    for (int k = 0; k < 3; ++k) {
        v.c[k] = v.c[k] / m;
        flush_small(v.c[k]);
    }
    return dot3(v, navg) < -0.1f;
}

}  // namespace

extern "C" {

void* orc_ic_create(float a, const float* minp, const float* size) {
    IrrCacheO* c = new IrrCacheO();
    c->a = a;
    for (int k = 0; k < 3; ++k) {
        c->minp.c[k] = minp[k];
        c->size.c[k] = size[k];
    }
    return c;
}

void orc_ic_free(void* h) { delete static_cast<IrrCacheO*>(h); }

// status[i]: 0 stored, 1 outside the root box (dropped), 2 refused (the
// reference's AddElement would not return).
void orc_ic_add(void* h, const float* value, const float* pos, const float* nrm, const float* ri, uint32_t n,
                uint32_t* status) {
    IrrCacheO* c = static_cast<IrrCacheO*>(h);
    for (uint32_t i = 0; i < n; ++i) {
        IrrElem e;
        for (int k = 0; k < 3; ++k) {
            e.value.c[k] = value[3 * (size_t)i + k];
            e.pos.c[k] = pos[3 * (size_t)i + k];
            e.nrm.c[k] = nrm[3 * (size_t)i + k];
        }
        e.ri = ri[i];
        F3 mx;
        for (int k = 0; k < 3; ++k) mx.c[k] = c->minp.c[k] + c->size.c[k];
        bool out = false;
        for (int k = 0; k < 3; ++k) out = out || e.pos.c[k] < c->minp.c[k] || mx.c[k] < e.pos.c[k];
        if (out) {
            status[i] = 1;
            continue;
        }
        const float domain = c->a * e.ri;
        if (!(2 * domain <= c->size.c[0])) {
            status[i] = 2;
            continue;
        }
        c->root.push_back(e);
        status[i] = 0;
    }
}

uint32_t orc_ic_size(void* h) { return (uint32_t)static_cast<IrrCacheO*>(h)->root.size(); }

// Per query: rgb = (sum of value * weight) / weightSum over the counted
// elements in list order (0 when weightSum is 0), wsum = weightSum, count.
void orc_ic_query(void* h, const float* qpos, const float* qnrm, uint32_t nq, float* rgb, float* wsum,
                  uint32_t* count, int nthreads) {
    const IrrCacheO* c = static_cast<const IrrCacheO*>(h);
    const double tolerance = 1 / c->a;  // IrrCacheOctree.h:300 (float division, widened)
    parallel_chunks(nq, nthreads, 64, [&](uint32_t b, uint32_t e) {
        for (uint32_t q = b; q < e; ++q) {
            F3 p, n;
            for (int k = 0; k < 3; ++k) {
                p.c[k] = qpos[3 * (size_t)q + k];
                n.c[k] = qnrm[3 * (size_t)q + k];
            }
            float ws = 0, acc[3] = {0, 0, 0};
            uint32_t cnt = 0;
            for (const IrrElem& el : c->root) {
                const float w = irr_weight(el, p, n);
                if (w > tolerance && !irr_in_front(p, el.pos, n, el.nrm)) {
                    ws += w;
                    for (int k = 0; k < 3; ++k) acc[k] += el.value.c[k] * w;
                    ++cnt;
                }
            }
            if (ws)
                for (int k = 0; k < 3; ++k) acc[k] /= ws;
            else
                acc[0] = acc[1] = acc[2] = 0;
            for (int k = 0; k < 3; ++k) rgb[3 * (size_t)q + k] = acc[k];
            wsum[q] = ws;
            count[q] = cnt;
        }
    });
}

}  // extern "C"
