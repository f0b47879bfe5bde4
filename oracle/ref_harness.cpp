// ref_harness.cpp -- drives ENCORE's own predicate sources, compiled from
// /root/reference by oracle/Makefile (target `ref`) into oracle/_ref/.
//
// TEST INFRASTRUCTURE ONLY: used by tests/golden/make_golden.py to write the
// predicate golden vectors that pin the oracle.  Nothing here is shipped.
//
// Only the part of the reference that builds from its own sources without
// stand-ins is used: math/{Vector3,Point3,Ray}.cpp and
// scene/{TrianglePrim,Triangle,Vertex}.cpp.  HitInfo's constructor lives in
// scene/HitInfo.cpp, which includes misc/Common.h and with it the Win32 DWORD
// type (Common.h:82); so the hit record is set up here field by field with
// the defaults of HitInfo.cpp:8-18 instead of through that constructor.
#include <cfloat>
#include <cstring>
#include <new>

#include "Ray.h"
#include "Triangle.h"
#include "TrianglePrim.h"
#include "Vector3.h"
#include "Vertex.h"

using encore::Point3;
using encore::Ray;
using encore::Vector3;

extern "C" {

// out8 = t, hitPoint.xyz, hitNormal.xyz, isEntering; returns bHasInfo.
int ref_tri_intersect(const float* o, const float* d, const float* v0, const float* v1, const float* v2,
                      float* out8) {
    Vertex a, b, c;
    a.setCoordinates(v0[0], v0[1], v0[2]);
    b.setCoordinates(v1[0], v1[1], v1[2]);
    c.setCoordinates(v2[0], v2[1], v2[2]);
    Triangle tri(a, b, c);
    TrianglePrim prim(tri);
    Ray ray(Point3(o[0], o[1], o[2]), Vector3(d[0], d[1], d[2]));

    alignas(HitInfo) unsigned char storage[sizeof(HitInfo)];
    std::memset(storage, 0, sizeof(storage));
    HitInfo* h = reinterpret_cast<HitInfo*>(storage);
    h->bHasInfo = false;
    h->hitTime = FLT_MAX;
    h->hitObject = nullptr;
    h->isEntering = false;

    bool hit = prim.intersect(ray, h);
    out8[0] = (float)h->hitTime;
    out8[1] = h->hitPoint.X();
    out8[2] = h->hitPoint.Y();
    out8[3] = h->hitPoint.Z();
    out8[4] = h->hitNormal.X();
    out8[5] = h->hitNormal.Y();
    out8[6] = h->hitNormal.Z();
    out8[7] = h->isEntering ? 1.0f : 0.0f;
    return hit ? 1 : 0;
}

void ref_cross(const float* a, const float* b, float* out) {
    Vector3 A(a[0], a[1], a[2]), B(b[0], b[1], b[2]);
    Vector3 r = encore::Cross(A, B);
    out[0] = r.X();
    out[1] = r.Y();
    out[2] = r.Z();
}

void ref_normalize(const float* a, float* out) {
    Vector3 v(a[0], a[1], a[2]);
    v.Normalize();
    out[0] = v.X();
    out[1] = v.Y();
    out[2] = v.Z();
}

}  // extern "C"
