"""Writes tests/golden/predicate_golden.npz from ENCORE's own predicate code.

Runs in the development container only: it needs oracle/_ref/
libref_predicate.so, which oracle/Makefile (target `ref`) compiles from the
reference's math/{Vector3,Point3,Ray}.cpp and scene/{TrianglePrim,Triangle,
Vertex}.cpp.  The fixture holds inputs and the reference's outputs only:

  tri_rays (n,8), tri_verts (n,9)  ->  tri_out (n,8) = t, hitPoint.xyz,
                                        hitNormal.xyz, isEntering; tri_hit (n,)
  cross_a, cross_b (m,3)            ->  cross_out (m,3)      (Vector3.h:148-162)
  norm_in (m,3)                     ->  norm_out (m,3)       (Vector3.cpp:105-114)

Case families (tri_kind): 0 random hits at scale 1, 1 at scale 100, 2 at
scale 0.05 (the Cross flush is active), 3 rays through edges and vertices,
4 rays parallel to the plane, 5 origins on the plane, 6 near-degenerate
slivers, 7 grazing rays.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402


def cases(rng):
    rays, tris, kinds = [], [], []

    def add(o, d, v, k):
        r = np.zeros(8, np.float32)
        r[0:3] = o
        r[4:7] = d
        r[7] = np.inf
        rays.append(r)
        tris.append(np.asarray(v, np.float32).reshape(9))
        kinds.append(k)

    def random_tri(scale):
        c = rng.uniform(-1, 1, 3) * scale
        return c + rng.normal(0, 1, (3, 3)) * scale * 0.5

    def aim(v, bary, scale, k, jitter=0.0):
        p = bary[0] * v[0] + bary[1] * v[1] + bary[2] * v[2]
        o = p + rng.normal(0, 1, 3) * scale * 2.0
        d = p - o + rng.normal(0, 1, 3) * jitter * scale
        d /= np.linalg.norm(d)
        add(o, d, v, k)

    for k, scale in ((0, 1.0), (1, 100.0), (2, 0.05)):
        for _ in range(700):
            v = random_tri(scale)
            b = rng.dirichlet([1, 1, 1]) * rng.uniform(0.5, 1.3)
            b[0] = 1 - b[1] - b[2]
            aim(v, b, scale, k, jitter=0.05)
    for _ in range(500):  # edges and vertices
        v = random_tri(rng.choice([0.01, 1.0, 100.0]))
        w = rng.uniform()
        b = [np.array([1 - w, w, 0]), np.array([0, 1 - w, w]), np.array([w, 0, 1 - w]),
             np.eye(3)[rng.integers(3)]][rng.integers(4)]
        aim(v, b, np.linalg.norm(v[1] - v[0]), 3)
    for _ in range(200):  # parallel to the plane
        v = random_tri(1.0)
        n = np.cross(v[1] - v[0], v[2] - v[0])
        d = np.cross(n, rng.normal(0, 1, 3))
        d /= np.linalg.norm(d)
        o = (v[0] + v[1] + v[2]) / 3 - d * 2.0
        add(o, d, v, 4)
    for _ in range(200):  # origin on the plane (exactly, for axis-aligned planes)
        z = rng.uniform(-5, 5)
        v = np.array([[rng.uniform(-3, 0), rng.uniform(-3, 0), z], [rng.uniform(0, 3), rng.uniform(-3, 0), z],
                      [rng.uniform(-1, 1), rng.uniform(0, 3), z]])
        o = np.array([rng.uniform(-1, 1), rng.uniform(-1, 1), z])
        d = rng.normal(0, 1, 3)
        d /= np.linalg.norm(d)
        add(o, d, v, 5)
    for _ in range(300):  # slivers: one tiny edge
        v = random_tri(1.0)
        v[2] = v[1] + rng.normal(0, 1, 3) * 10 ** rng.uniform(-5, -2)
        aim(v, rng.dirichlet([1, 1, 1]), 1.0, 6, jitter=0.01)
    for _ in range(300):  # grazing
        v = random_tri(1.0)
        n = np.cross(v[1] - v[0], v[2] - v[0])
        n /= np.linalg.norm(n)
        t_in = np.cross(n, rng.normal(0, 1, 3))
        t_in /= np.linalg.norm(t_in)
        d = t_in + n * 10 ** rng.uniform(-4, -1)
        d /= np.linalg.norm(d)
        p = (v[0] + v[1] + v[2]) / 3
        add(p - d * 3.0, d, v, 7)
    return np.array(rays, np.float32), np.array(tris, np.float32), np.array(kinds, np.uint8)


def main():
    ref = oracle.ref_lib()
    if ref is None:
        raise SystemExit("oracle/_ref/libref_predicate.so is missing: run `make -C oracle ref`")
    rng = np.random.default_rng(20240601)
    rays, tris, kinds = cases(rng)
    out = np.zeros((len(rays), 8), np.float32)
    hit = np.zeros(len(rays), np.uint8)
    for i in range(len(rays)):
        o8 = np.zeros(8, np.float32)
        hit[i] = ref.ref_tri_intersect(rays[i, 0:3].copy(), rays[i, 4:7].copy(), tris[i, 0:3].copy(),
                                       tris[i, 3:6].copy(), tris[i, 6:9].copy(), o8)
        out[i] = o8
    m = 2000
    ca = (rng.normal(0, 1, (m, 3)) * (10.0 ** rng.uniform(-4, 2, (m, 1)))).astype(np.float32)
    cb = (rng.normal(0, 1, (m, 3)) * (10.0 ** rng.uniform(-4, 2, (m, 1)))).astype(np.float32)
    cout = np.zeros((m, 3), np.float32)
    nin = (rng.normal(0, 1, (m, 3)) * (10.0 ** rng.uniform(-3, 3, (m, 1)))).astype(np.float32)
    nin[: m // 10, rng.integers(0, 3)] *= 1e-4  # components that flush after normalising
    nout = np.zeros((m, 3), np.float32)
    for i in range(m):
        o3 = np.zeros(3, np.float32)
        ref.ref_cross(ca[i].copy(), cb[i].copy(), o3)
        cout[i] = o3
        o3 = np.zeros(3, np.float32)
        ref.ref_normalize(nin[i].copy(), o3)
        nout[i] = o3
    path = os.path.join(HERE, "predicate_golden.npz")
    np.savez_compressed(path, tri_rays=rays, tri_verts=tris, tri_kind=kinds, tri_out=out, tri_hit=hit,
                        cross_a=ca, cross_b=cb, cross_out=cout, norm_in=nin, norm_out=nout)
    print("wrote", path, len(rays), "predicate cases,", int(hit.sum()), "hits")


if __name__ == "__main__":
    main()
