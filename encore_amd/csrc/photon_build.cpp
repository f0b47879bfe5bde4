// photon_build.cpp -- ENCORE's balanced photon kd-tree, built on the host.
//
// PhotonMap::BalanceAndConstruct (PhotonMap.cpp:167-221) stores the tree
// implicitly: the node of index range [lo, hi] is the photon at (lo+hi)/2,
// put there by std::nth_element along the node's axis; the children are the
// ranges on either side.  The axis comes from a bookkeeping box that starts
// at +-1e10 (PhotonMap.cpp:21, 25-29) and is narrowed and restored around the
// recursion, so it depends on the partition order; the same standard-library
// nth_element therefore reproduces ENCORE's photon order slot for slot.
//
// Normals are mapped to class ids so the GPU compares one integer where
// ENCORE compares three floats with == (PhotonMap.cpp:244, Vector3.cpp:89-94):
// -0 and +0 share a class, and a normal with a NaN component matches nothing.
#include <algorithm>
#include <cmath>
#include <unordered_map>

#include "rq_internal.h"

namespace rq {
namespace {

struct BPhoton {
    float pos[3];
    uint32_t id;
    int axis;
};

struct Balancer {
    std::vector<BPhoton>& ph;
    float bmin[3], bmax[3];

    explicit Balancer(std::vector<BPhoton>& p) : ph(p) {
        for (int k = 0; k < 3; ++k) {
            bmin[k] = -1.0e10f;
            bmax[k] = 1.0e10f;
        }
    }

    void run(int lo, int hi) {
        if (hi - lo < 0) return;
        const float ex = bmax[0] - bmin[0], ey = bmax[1] - bmin[1], ez = bmax[2] - bmin[2];
        int axis = 2;
        if (ex > ey && ey > ez)
            axis = 0;
        else if (ey > ez)
            axis = 1;
        const int med = (hi + lo) / 2;
        std::nth_element(ph.begin() + lo, ph.begin() + med, ph.begin() + hi + 1,
                         [axis](const BPhoton& a, const BPhoton& b) { return a.pos[axis] < b.pos[axis]; });
        ph[med].axis = axis;
        const float keep_lo = ph[lo].pos[axis], keep_hi = ph[hi].pos[axis];
        const float m = ph[med].pos[axis];
        bmin[axis] = bmax[axis] = m;
        run(lo, med - 1);
        bmin[axis] = bmax[axis] = ph[med].pos[axis];
        run(med + 1, hi);
        bmin[axis] = keep_lo;
        bmax[axis] = keep_hi;
    }
};

// Balanced tree with the same implicit layout, split on the longest extent
// of each range's own bounding box: same gathered sets, tighter pruning.
void balance_spatial(std::vector<BPhoton>& ph, int lo, int hi) {
    if (hi - lo < 0) return;
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = lo; i <= hi; ++i)
        for (int k = 0; k < 3; ++k) {
            mn[k] = std::min(mn[k], ph[i].pos[k]);
            mx[k] = std::max(mx[k], ph[i].pos[k]);
        }
    int axis = 0;
    for (int k = 1; k < 3; ++k)
        if (mx[k] - mn[k] > mx[axis] - mn[axis]) axis = k;
    const int med = (hi + lo) / 2;
    std::nth_element(ph.begin() + lo, ph.begin() + med, ph.begin() + hi + 1,
                     [axis](const BPhoton& a, const BPhoton& b) { return a.pos[axis] < b.pos[axis]; });
    ph[med].axis = axis;
    balance_spatial(ph, lo, med - 1);
    balance_spatial(ph, med + 1, hi);
}

inline uint32_t canon_bits(float f) {
    if (f == 0.0f) return 0u;  // -0 == +0
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}

}  // namespace

void build_photons(const float* pos, const float* nrm, const float* power, uint32_t n, uint32_t layout,
                   PhotonHost& out) {
    std::vector<BPhoton> ph(n);
    for (uint32_t i = 0; i < n; ++i) {
        for (int k = 0; k < 3; ++k) ph[i].pos[k] = pos[3 * (size_t)i + k];
        ph[i].id = i;
        ph[i].axis = 2;
    }
    if (n > 0) {
        if (layout == RQ_PHOTONS_ENCORE_ORDER) {
            Balancer B(ph);
            B.run(0, (int)n - 1);
        } else {
            balance_spatial(ph, 0, (int)n - 1);
        }
    }
    // Normal classes in first-seen order over the insertion sequence.
    std::unordered_map<uint64_t, std::vector<std::pair<std::array<uint32_t, 3>, uint32_t>>> seen;
    std::vector<uint32_t> cls(n);
    out.normals.clear();
    for (uint32_t i = 0; i < n; ++i) {
        const float* v = nrm + 3 * (size_t)i;
        if (std::isnan(v[0]) || std::isnan(v[1]) || std::isnan(v[2])) {
            cls[i] = kClassNever;
            continue;
        }
        std::array<uint32_t, 3> key = {canon_bits(v[0]), canon_bits(v[1]), canon_bits(v[2])};
        const uint64_t h = normal_hash(key[0], key[1], key[2]);
        auto& bucket = seen[h];
        uint32_t c = kClassNever;
        for (auto& e : bucket)
            if (e.first == key) c = e.second;
        if (c == kClassNever) {
            c = (uint32_t)out.normals.size();
            if (c >= kClassNever) throw Error(RQ_ERR_CAPACITY, "too many distinct photon normals");
            out.normals.push_back(key);
            bucket.push_back({key, c});
        }
        cls[i] = c;
    }
    // Open-addressing table of the classes for the GPU query lookup.
    size_t cap = 16;
    while (cap < 2 * out.normals.size() + 1) cap <<= 1;
    out.table.assign(cap * 4, 0xFFFFFFFFu);
    for (uint32_t c = 0; c < out.normals.size(); ++c) {
        const auto& k = out.normals[c];
        size_t s = normal_hash(k[0], k[1], k[2]) & (cap - 1);
        while (out.table[4 * s + 3] != 0xFFFFFFFFu) s = (s + 1) & (cap - 1);
        out.table[4 * s + 0] = k[0];
        out.table[4 * s + 1] = k[1];
        out.table[4 * s + 2] = k[2];
        out.table[4 * s + 3] = c;
    }
    out.hot.resize(n);
    out.power.resize(n);
    out.perm.resize(n);
    out.axis.resize(n);
    for (uint32_t s = 0; s < n; ++s) {
        const BPhoton& p = ph[s];
        PhotonHot& h = out.hot[s];
        h.x = p.pos[0];
        h.y = p.pos[1];
        h.z = p.pos[2];
        h.meta = (uint32_t)p.axis | (cls[p.id] << 2);
        for (int k = 0; k < 3; ++k) out.power[s].c[k] = power ? power[3 * (size_t)p.id + k] : 0.0f;
        out.power[s].c[3] = 0.0f;
        out.perm[s] = p.id;
        out.axis[s] = (uint8_t)p.axis;
    }
}

}  // namespace rq
