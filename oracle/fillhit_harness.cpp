// fillhit_harness.cpp -- drives the drop-in adapter's HitInfo arithmetic
// (encore_amd/adapter/HitFill.h) on ENCORE's own Point3 / Vector3, compiled
// from /root/reference by oracle/Makefile (target `ref`) into oracle/_ref/.
//
// TEST INFRASTRUCTURE ONLY: tests/test_adapter_fillhit.py compares its
// output with the reference's TrianglePrim::intersect golden vectors.
#include "HitFill.h"

using encore::Point3;
using encore::Vector3;

extern "C" {

// tuv = the engine's (t, alpha, beta); out7 = hitPoint.xyz, hitNormal.xyz,
// isEntering.
void ref_fill_hit(const float* v0, const float* v1, const float* v2, const float* d, const float* tuv, float* out7) {
    rq_hit h;
    h.t = tuv[0];
    h.u = tuv[1];
    h.v = tuv[2];
    h.prim = 0;
    const EncoreHitFields f = encore_hit_fields(Point3(v0[0], v0[1], v0[2]), Point3(v1[0], v1[1], v1[2]),
                                                Point3(v2[0], v2[1], v2[2]), Vector3(d[0], d[1], d[2]), h);
    out7[0] = f.point.X();
    out7[1] = f.point.Y();
    out7[2] = f.point.Z();
    out7[3] = f.normal.X();
    out7[4] = f.normal.Y();
    out7[5] = f.normal.Z();
    out7[6] = f.entering ? 1.0f : 0.0f;
}

}  // extern "C"
