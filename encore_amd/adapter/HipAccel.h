// HipAccel.h -- drop-in AccelerationStructure backed by the MI355X engine.
//
// Compiles inside ENCORE's source tree (it includes ENCORE's own headers:
// AccelerationStructure.h, IModel.h, Primitive.h, TrianglePrim.h) and links
// against libencore_rq.so through include/rq.h.  It replaces Kdtree, Bvh,
// BasicAS or UniformGrid behind the interface of
// src/accelerators/AccelerationStructure.h:33-67:
//
//   build(modelList)    flattens every TrianglePrim of every model, in list
//                       order, into one indexed triangle list (primitive id =
//                       position in that order, as Kdtree.cpp:99-114 and
//                       BasicAS.cpp:31-41 concatenate them) and builds the
//                       selected structure in HBM;
//   intersect(ray, hit) one-ray query; fills HitInfo exactly as
//                       TrianglePrim::intersect does (TrianglePrim.cpp:226-244)
//                       from the engine's (t, alpha, beta, prim);
//   update(modelList)   moved vertices, same primitives: the BVH is refitted
//                       on the GPU (rq_scene_refit), a KD scene rebuilt;
//                       a changed primitive list builds afresh;
//   intersectBatch /    batched extensions for renderers that trace many
//   occludedBatch       rays at once (one launch per call).
//
// Options read by configure() (EncoreApp's KEY VALUE config map):
//   HipAccelKind   BVH (default) | KD      KD reproduces Kdtree's results
//                                          bit for bit, BVH up to ties
//   HipPredicate   Encore (default) | Watertight
//   HipDevice      GPU ordinal (default 0)
//   HipBvhWidth    4 (default) | 2
//   HipBvhBuilder  SAH (default, host) | LBVH (GPU, BVH2 trees)
#pragma once

#include <cstdint>
#include <list>
#include <string>
#include <vector>

#include "AccelerationStructure.h"
#include "rq.h"

class IModel;
class IPrimitive;

class HipAccel : public AccelerationStructure {
public:
    HipAccel();
    virtual ~HipAccel();

    void usage(void);
    void configure(Options* l_pOptions);
    void build(std::list<IModel*>& modelList);
    void update(std::list<IModel*>& modelList);
#ifdef WIN32
    // ENCORE's GL-texture GPU path is not used: the engine keeps its own
    // layout in HBM.
    void buildGPU(std::list<IModel*>&, std::list<Triangle*>&, GPUAccelerationStructureData&) {}
    void setGPUParameters(CShader&, GPUAccelerationStructureData&) {}
#endif
    bool intersect(Ray& l_Ray, HitInfo* pHitInfo);
    void keyboard(unsigned char) {}

    // Batched queries (host arrays).  hits[i] follows intersect(); occluded[i]
    // is IsInShadow's answer for a ray toward a light at distance dist[i].
    void intersectBatch(std::vector<Ray>& rays, std::vector<HitInfo>& hits);
    void occludedBatch(std::vector<Ray>& rays, const std::vector<float>& dist, std::vector<bool>& occluded);

    const char* lastError() const { return m_Error.c_str(); }
    size_t primitiveCount() const { return m_Prims.size(); }

private:
    void release();
    // Triangles of the model list in primitive-id order (9 floats each) and
    // the model-list bounds.
    void gather(std::list<IModel*>& modelList, std::vector<IPrimitive*>& prims, std::vector<float>& xyz,
                float bmin[3], float bmax[3]);
    void ensureCapacity(size_t n);
    void fillHit(Ray& ray, const rq_hit& h, HitInfo* out) const;
    bool ok(rq_status s);

    uint32_t m_Kind;
    uint32_t m_Predicate;
    uint32_t m_Width;    // BVH children per node: 2 or 4
    uint32_t m_Builder;  // RQ_BUILD_HOST_SAH or RQ_BUILD_DEVICE_LBVH
    int m_Ordinal;
    rq_device* m_Device;
    rq_scene* m_Scene;
    std::vector<IPrimitive*> m_Prims;  // primitive id -> ENCORE primitive
    std::vector<float> m_Xyz;          // 9 floats per triangle (v0, v1, v2)
    void* m_dRays;                     // device staging buffers
    void* m_dOut;
    void* m_dXyz;                      // vertex positions for update()
    size_t m_Capacity;
    std::string m_Error;
};
