// HitFill.h -- the HitInfo fields TrianglePrim::intersect derives from a hit
// (TrianglePrim.cpp:226-244), computed from the engine's rq_hit with ENCORE's
// own Point3 / Vector3 arithmetic, so a drop-in returns the reference's bits:
//   hitNormal  = Normalize(Cross(v1 - v0, v2 - v0))   (both flush below 1e-3)
//   hitPoint   = v0 + (v1 - v0) * alpha + (v2 - v0) * beta
//   isEntering = Dot(direction, hitNormal) < 0
// with alpha = h.u (the weight of v1) and beta = h.v (the weight of v2),
// TrianglePrim.cpp:219-220.  Header-only: HipAccel uses it inside ENCORE's
// tree, and a CPU test compiles it against ENCORE's math sources.
#pragma once

#include "Point3.h"
#include "Vector3.h"
#include "rq.h"

struct EncoreHitFields {
    encore::Point3 point;
    encore::Vector3 normal;
    bool entering;
};

inline EncoreHitFields encore_hit_fields(const encore::Point3& v0, const encore::Point3& v1,
                                         const encore::Point3& v2, const encore::Vector3& direction,
                                         const rq_hit& h) {
    const encore::Vector3 edge1(v0, v1), edge2(v0, v2);
    EncoreHitFields f;
    f.normal = Cross(edge1, edge2);
    f.normal.Normalize();
    f.point = v0 + (v1 - v0) * h.u + (v2 - v0) * h.v;
    f.entering = Dot(direction, f.normal) < 0;
    return f;
}
