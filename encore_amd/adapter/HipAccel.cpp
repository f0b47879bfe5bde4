// HipAccel.cpp -- see HipAccel.h.
#include "HipAccel.h"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>

#include "HitFill.h"
#include "IModel.h"
#include "Primitive.h"
#include "Triangle.h"
#include "TrianglePrim.h"
#include "Vertex.h"

using encore::Point3;
using encore::Vector3;

HipAccel::HipAccel()
    : m_Kind(RQ_ACCEL_BVH), m_Predicate(RQ_PRED_ENCORE), m_Width(4), m_Builder(RQ_BUILD_HOST_SAH), m_Ordinal(0),
      m_Device(0), m_Scene(0), m_dRays(0), m_dOut(0), m_dXyz(0), m_Capacity(0) {}

HipAccel::~HipAccel() {
    release();
    if (m_Device) rq_device_close(m_Device);
}

void HipAccel::usage(void) {
    printf("HipAccel: MI355X ray-query engine\n"
           "  HipAccelKind  BVH | KD      (KD reproduces Kdtree exactly)\n"
           "  HipPredicate  Encore | Watertight\n"
           "  HipDevice     GPU ordinal\n"
           "  HipBvhWidth   4 | 2\n"
           "  HipBvhBuilder SAH | LBVH  (LBVH builds on the GPU)\n");
}

void HipAccel::configure(Options* o) {
    if (!o) return;
    if (o->doesOptionExist("HipAccelKind")) m_Kind = (o->getOption("HipAccelKind") == "KD") ? RQ_ACCEL_KD : RQ_ACCEL_BVH;
    if (o->doesOptionExist("HipPredicate"))
        m_Predicate = (o->getOption("HipPredicate") == "Watertight") ? RQ_PRED_WATERTIGHT : RQ_PRED_ENCORE;
    if (o->doesOptionExist("HipDevice")) m_Ordinal = atoi(o->getOption("HipDevice").c_str());
    if (o->doesOptionExist("HipBvhWidth")) m_Width = (atoi(o->getOption("HipBvhWidth").c_str()) == 2) ? 2u : 4u;
    if (o->doesOptionExist("HipBvhBuilder"))
        m_Builder = (o->getOption("HipBvhBuilder") == "LBVH") ? RQ_BUILD_DEVICE_LBVH : RQ_BUILD_HOST_SAH;
}

bool HipAccel::ok(rq_status s) {
    if (s == RQ_OK) return true;
    m_Error = rq_last_error();
    printf("HipAccel: %s\n", m_Error.c_str());
    return false;
}

void HipAccel::release() {
    if (m_Scene) rq_scene_destroy(m_Scene);
    m_Scene = 0;
    if (m_dRays) (void)hipFree(m_dRays);
    if (m_dOut) (void)hipFree(m_dOut);
    if (m_dXyz) (void)hipFree(m_dXyz);
    m_dRays = m_dOut = m_dXyz = 0;
    m_Capacity = 0;
}

void HipAccel::gather(std::list<IModel*>& modelList, std::vector<IPrimitive*>& prims, std::vector<float>& xyz,
                      float bmin[3], float bmax[3]) {
    prims.clear();
    xyz.clear();
    // Model-list bounds, as Kdtree::build gathers them (Kdtree.cpp:92-107).
    for (int k = 0; k < 3; ++k) {
        bmin[k] = INFINITY;
        bmax[k] = -INFINITY;
    }
    for (std::list<IModel*>::iterator m = modelList.begin(); m != modelList.end(); ++m) {
        for (int k = 0; k < 3; ++k) {
            bmin[k] = std::min((*m)->MinPoint()[k], bmin[k]);
            bmax[k] = std::max((*m)->MaxPoint()[k], bmax[k]);
        }
        std::list<IPrimitive*>& pl = (*m)->getPrimitiveList();
        for (std::list<IPrimitive*>::iterator p = pl.begin(); p != pl.end(); ++p) {
            if (!dynamic_cast<TrianglePrim*>(*p)) {
                printf("HipAccel: skipping a non-triangle primitive (only TrianglePrim is supported)\n");
                continue;
            }
            std::list<Triangle*>* tl = (*p)->getNewTesselation();  // one copy of the triangle
            Triangle* t = tl->front();
            const Vertex* v[3] = {t->getVertex0(), t->getVertex1(), t->getVertex2()};
            for (int j = 0; j < 3; ++j) {
                const Point3& c = v[j]->getCoordinates();
                xyz.push_back(c.X());
                xyz.push_back(c.Y());
                xyz.push_back(c.Z());
            }
            for (std::list<Triangle*>::iterator it = tl->begin(); it != tl->end(); ++it) delete *it;
            delete tl;
            prims.push_back(*p);
        }
    }
}

void HipAccel::build(std::list<IModel*>& modelList) {
    release();
    if (!m_Device && !ok(rq_device_open(m_Ordinal, &m_Device))) return;
    float bmin[3], bmax[3];
    gather(modelList, m_Prims, m_Xyz, bmin, bmax);
    if (m_Prims.empty()) return;
    const uint32_t nt = (uint32_t)m_Prims.size();
    std::vector<uint32_t> idx(3 * (size_t)nt);
    for (size_t i = 0; i < idx.size(); ++i) idx[i] = (uint32_t)i;
    rq_scene_desc d;
    memset(&d, 0, sizeof(d));
    d.accel = m_Kind;
    d.predicate = m_Predicate;
    d.bvh_width = m_Builder == RQ_BUILD_DEVICE_LBVH ? 2 : m_Width;
    d.builder = m_Builder;
    d.has_bounds = 1;
    for (int k = 0; k < 3; ++k) {
        d.bounds_min[k] = bmin[k];
        d.bounds_max[k] = bmax[k];
    }
    ok(rq_scene_create(m_Device, m_Xyz.data(), 3 * nt, idx.data(), nt, &d, &m_Scene));
}

// AccelerationStructure::update for dynamic models (DynamicModel.cpp:85):
// the same primitives with moved vertices are refitted in place; anything
// else is built again.
void HipAccel::update(std::list<IModel*>& modelList) {
    if (!m_Scene) {
        build(modelList);
        return;
    }
    std::vector<IPrimitive*> prims;
    std::vector<float> xyz;
    float bmin[3], bmax[3];
    gather(modelList, prims, xyz, bmin, bmax);
    if (prims != m_Prims) {
        build(modelList);
        return;
    }
    if (!m_dXyz && hipMalloc(&m_dXyz, xyz.size() * sizeof(float)) != hipSuccess) {
        m_dXyz = 0;
        m_Error = "device vertex buffer allocation failed";
        return;
    }
    if (hipMemcpy(m_dXyz, xyz.data(), xyz.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) {
        m_Error = "vertex upload failed";
        return;
    }
    m_Xyz.swap(xyz);
    ok(rq_scene_refit(m_Scene, static_cast<const float*>(m_dXyz), (uint32_t)(m_Xyz.size() / 3), 0));
}

void HipAccel::ensureCapacity(size_t n) {
    if (n <= m_Capacity) return;
    if (m_dRays) (void)hipFree(m_dRays);
    if (m_dOut) (void)hipFree(m_dOut);
    m_dRays = m_dOut = 0;
    m_Capacity = 0;
    if (hipMalloc(&m_dRays, n * sizeof(rq_ray)) != hipSuccess || hipMalloc(&m_dOut, n * sizeof(rq_hit)) != hipSuccess) {
        m_Error = "device staging allocation failed";
        return;
    }
    m_Capacity = n;
}

// HitInfo from the engine's hit record, with the arithmetic of
// TrianglePrim::intersect (TrianglePrim.cpp:226-244) on ENCORE's own types.
void HipAccel::fillHit(Ray& ray, const rq_hit& h, HitInfo* out) const {
    out->bHasInfo = false;
    out->hitTime = INFINITY;  // Kdtree.cpp:685
    if (h.prim == RQ_NO_HIT || h.prim >= m_Prims.size()) return;
    const float* f = &m_Xyz[9 * (size_t)h.prim];
    const Point3 v0(f[0], f[1], f[2]), v1(f[3], f[4], f[5]), v2(f[6], f[7], f[8]);
    const EncoreHitFields hf = encore_hit_fields(v0, v1, v2, ray.Direction(), h);
    out->hitTime = h.t;
    out->hitNormal = hf.normal;
    out->hitPoint = hf.point;
    out->isEntering = hf.entering;
    out->hitObject = m_Prims[h.prim];
    out->bHasInfo = true;
}

static void to_rq(Ray& r, rq_ray& q, float tmax) {
    q.ox = r.Origin().X();
    q.oy = r.Origin().Y();
    q.oz = r.Origin().Z();
    q.tmin = 0.0f;
    q.dx = r.Direction().X();
    q.dy = r.Direction().Y();
    q.dz = r.Direction().Z();
    q.tmax = tmax;
}

bool HipAccel::intersect(Ray& l_Ray, HitInfo* pHitInfo) {
    pHitInfo->bHasInfo = false;
    pHitInfo->hitTime = INFINITY;
    if (!m_Scene) return false;
    ensureCapacity(1);
    if (!m_Capacity) return false;
    rq_ray q;
    to_rq(l_Ray, q, INFINITY);
    rq_hit h;
    if (hipMemcpy(m_dRays, &q, sizeof(q), hipMemcpyHostToDevice) != hipSuccess) return false;
    if (!ok(rq_intersect_closest(m_Scene, (const rq_ray*)m_dRays, 1, (rq_hit*)m_dOut, 0))) return false;
    if (hipMemcpy(&h, m_dOut, sizeof(h), hipMemcpyDeviceToHost) != hipSuccess) return false;
    fillHit(l_Ray, h, pHitInfo);
    return pHitInfo->bHasInfo;
}

void HipAccel::intersectBatch(std::vector<Ray>& rays, std::vector<HitInfo>& hits) {
    const size_t n = rays.size();
    hits.resize(n);
    for (size_t i = 0; i < n; ++i) {
        hits[i].bHasInfo = false;
        hits[i].hitTime = INFINITY;
    }
    if (!m_Scene || n == 0) return;
    ensureCapacity(n);
    if (m_Capacity < n) return;
    std::vector<rq_ray> q(n);
    for (size_t i = 0; i < n; ++i) to_rq(rays[i], q[i], INFINITY);
    std::vector<rq_hit> h(n);
    if (hipMemcpy(m_dRays, q.data(), n * sizeof(rq_ray), hipMemcpyHostToDevice) != hipSuccess) return;
    if (!ok(rq_intersect_closest(m_Scene, (const rq_ray*)m_dRays, (uint32_t)n, (rq_hit*)m_dOut, 0))) return;
    if (hipMemcpy(h.data(), m_dOut, n * sizeof(rq_hit), hipMemcpyDeviceToHost) != hipSuccess) return;
    for (size_t i = 0; i < n; ++i) fillHit(rays[i], h[i], &hits[i]);
}

void HipAccel::occludedBatch(std::vector<Ray>& rays, const std::vector<float>& dist, std::vector<bool>& occluded) {
    const size_t n = rays.size();
    occluded.assign(n, false);
    if (!m_Scene || n == 0 || dist.size() != n) return;
    ensureCapacity(n);
    if (m_Capacity < n) return;
    std::vector<rq_ray> q(n);
    for (size_t i = 0; i < n; ++i) to_rq(rays[i], q[i], dist[i]);
    std::vector<uint32_t> o(n);
    if (hipMemcpy(m_dRays, q.data(), n * sizeof(rq_ray), hipMemcpyHostToDevice) != hipSuccess) return;
    if (!ok(rq_occluded(m_Scene, (const rq_ray*)m_dRays, (uint32_t)n, (uint32_t*)m_dOut, 0))) return;
    if (hipMemcpy(o.data(), m_dOut, n * sizeof(uint32_t), hipMemcpyDeviceToHost) != hipSuccess) return;
    for (size_t i = 0; i < n; ++i) occluded[i] = o[i] != 0;
}
