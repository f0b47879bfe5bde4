"""Writes tests/golden/oracle_pins.json: tiny cases whose answers follow from
the reference's source text, worked out by hand below (not by running the
oracle or the engine), each built so that one plausible misreading of the
source gives another answer.

  kd_snap          Kdtree::intersect's split snap, Kdtree.cpp:709-716
  kd_distribution  strict box overlap and "overlaps neither -> both",
                   Kdtree.cpp:561-650 with AABB::Intersect, AABB.cpp:9-18
  kd_leafsize      m_MaxItemPerNode = 2 for N <= 5000, else 10
                   (Kdtree.cpp:127-128, Kdtree.cpp:17) and the size-change
                   stall rule, |pre_size - size| <= 3 for 3 levels
                   (Kdtree.cpp:314-325)
  photon_axis      BalanceAndConstruct's bookkeeping box, which starts at
                   +-1e10 and is restored to the range's end photons after
                   each subtree (PhotonMap.cpp:21, 176-221)

Only float32 arithmetic that the source spells out is computed here: the
root box (model bounds -+ 0.001f, AABB corner + size, Kdtree.cpp:92-117,
AABB.h:34-37) and the median split (minp + maxp) / 2 (Kdtree.cpp:513-518).

    python tests/golden/make_pins.py
"""
import json
import os

import numpy as np

F = np.float32
EPS = F(0.001)


def root_split(lo, hi):
    """Root split on axis 0 of Kdtree::build / construct for model bounds
    [lo, hi] on that axis: AABB(min - diff, max + diff) stores the corner and
    the size, construct recomputes the max corner as pos + size and splits at
    (pos + max) / 2."""
    pos = F(F(lo) - EPS)
    size = F(F(F(hi) + EPS) - pos)
    mx = F(pos + size)
    return F(F(pos + mx) / F(2)), pos, mx


def wall(x, lo=-1000.0, hi=2000.0):
    """A triangle in the plane x = const, wide enough that no Cross()
    component of the rays below falls under the 1e-3 flush."""
    return [[x, lo, lo], [x, hi, lo], [x, lo, hi]]


def f32(v):
    return [float(F(c)) for c in v]


def kd_snap():
    # Three walls at x = -1, 0.0004 and 1.001.  Bounds x in [-1, 1.001]: the
    # root (depth 0, axis x) splits at s ~ 0.0005.  N = 3 > m_MaxItemPerNode
    # = 2, so the root is inner; the walls at -1 and 0.0004 overlap only the
    # left voxel (strictly: s > 0.0004 > -1.001), the wall at 1.001 only the
    # right one; both children are leaves (2 and 1 triangles).
    s, pos, mx = root_split(-1.0, 1.001)
    assert F(0) < s < F(0.001), s  # the traversal snaps it to 0
    tris = [wall(-1.0), wall(0.0004), wall(1.001)]
    xyz = [f32(v) for t in tris for v in t]
    idx = [[0, 1, 2], [3, 4, 5], [6, 7, 8]]
    inf = float("inf")
    rays = [
        # o.x = 0.0003 lies left of s but right of the snapped split 0:
        # below = (0.0003 <= 0) is false, so `first` is the right child, and
        # tplane = (0 - 0.0003) * 1 < 0 sends the ray to `first` only
        # (Kdtree.cpp:716-736).  The wall at 0.0004 (t = 1e-4) is never
        # tested; the wall at 1.001 (t ~ 1.0007) is the hit.
        [0.0003, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, inf],
        # o.x = -0.0003: below, tplane = 0.0003 inside (0, tmax): the left
        # child first, the wall at 0.0004 is hit at t ~ 7e-4 and the right
        # leaf's wall (t ~ 1.0013) does not beat it.
        [-0.0003, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, inf],
        # o.x = 0.0003 going -x: right child first (tmax = tplane = 3e-4; its
        # wall lies behind: a miss leaves the empty record, hitTime FLT_MAX),
        # then the left child from tmin = 3e-4: the wall at -1 (t ~ 1.0003);
        # the wall at 0.0004 lies behind.
        [0.0003, 0.0, 0.0, 0.0, -1.0, 0.0, 0.0, inf],
    ]
    return {"xyz": xyz, "idx": idx, "rays": [f32(r) for r in rays],
            "root_split": float(s),
            # (leaf?, axis or sorted prims) per node in ENCORE's numbering
            "nodes": [["inner", 0], ["leaf", [0, 1]], ["leaf", [2]]],
            "kd_prim": [2, 1, 0],       # Kdtree::intersect
            "brute_prim": [1, 1, 0]}    # BasicAS::intersect (closest over all)


def kd_distribution():
    # Bounds [-1, 1] on every axis; the root splits x at s (s ~ 0).
    #   B: x in [-1, s], touching the split plane from the left: the left
    #      voxel overlaps strictly (s > -1, pos < s); the right voxel needs
    #      s < max_x = s, false -> left only.  (A non-strict test would put
    #      it in both.)
    #   C: flat in the plane x = s: left needs s > s, right needs s < s ->
    #      neither, so it goes to both (Kdtree.cpp:579-586).  (Dropping it,
    #      or sending it left only, changes the right leaf.)
    #   D: x in [0.5, 1] -> right only.
    # N = 3 > 2: the root splits; each child holds 2 -> leaves, in primitive
    # order: left [B, C], right [C, D].
    s, _, _ = root_split(-1.0, 1.0)
    xyz = [[-1.0, 0.0, 0.0], [float(s), 0.0, 0.0], [-1.0, 1.0, 1.0],
           [float(s), -1.0, -1.0], [float(s), 1.0, -1.0], [float(s), 0.0, 1.0],
           [0.5, 0.0, 0.0], [1.0, 0.0, 0.0], [1.0, 1.0, 1.0]]
    return {"xyz": [f32(v) for v in xyz], "idx": [[0, 1, 2], [3, 4, 5], [6, 7, 8]], "root_split": float(s),
            "nodes": [["inner", 0], ["leaf", [0, 1]], ["leaf", [1, 2]]]}


def kd_leafsize():
    # R1: `r1` tiny triangles at grid points of [0.5, 95.5]^3 (ids 0..r1-1),
    # then 3 identical tiny "satellites" at (1000, 20, 50.7) and 10 identical
    # tiny "cluster" triangles at (1000, 50.3, 50.7).  A tiny triangle at p:
    # p, p + (0.01, 0, 0), p + (0, 0.01, 0.01).
    #
    # Root (N > 3 from pre_size 0: no stall), axis x, split ~500.25: R1 left,
    # the 13 far triangles right (depth 1, |N - 13| > 3: tries 0), axis y,
    # split ~48.0: satellites left, cluster right (depth 2).
    #   N = 5001 (m_MaxItemPerNode 10): 13 > 10 at depth 1 -> inner; at
    #     depth 2 both sides hold <= 10 -> the cluster leaf and the
    #     satellite leaf are at depth 2.
    #   N = 5000 (m_MaxItemPerNode 2): cluster at depth 2: |13 - 10| = 3 <= 3
    #     -> tries 1 (a "< 3" reading would give 0); depth 3 (axis x, split
    #     ~750: cluster right): tries 2; depth 4: tries 3 -> leaf at depth 4.
    #     Satellites at depth 2: |13 - 3| > 3 -> tries 0, 3 > 2 -> inner
    #     (axis z, split ~48: right); depth 3 tries 1 (axis x: right), depth 4
    #     tries 2 (axis y, split ~24.25: left), depth 5 tries 3 -> leaf.
    # Max depth 11 + floor(1.3 ln N) = 22 is never reached.
    return {"grid": {"start": 0.5, "step": 5.0, "n_per_axis": 20},
            "tiny": [[0.0, 0.0, 0.0], [0.01, 0.0, 0.0], [0.0, 0.01, 0.01]],
            "satellite_at": [1000.0, 20.0, 50.7], "cluster_at": [1000.0, 50.3, 50.7],
            "cases": [{"n": 5000, "r1": 4987, "max_item": 2, "cluster_depth": 4, "satellite_depth": 5},
                      {"n": 5001, "r1": 4988, "max_item": 10, "cluster_depth": 2, "satellite_depth": 2}]}


def photon_axis():
    # Five photons (insertion order a, b, c, d, e):
    #   a = (0.8, 0.1, 0.6), b = (0.1, 0.7, 0.2), c = (0.2, 0.9, 0.8),
    #   d = (0.5, 0.5, 0.5), e = (0.9, 0.3, 0.4).
    # Root [0, 4]: box +-1e10, equal extents -> Z (X needs x > y, Y needs
    #   y > z).  Median slot 2 = d (z 0.5); slots {0, 1} = {b, e} (z 0.2,
    #   0.4), slots {3, 4} = {a, c} (z 0.6, 0.8) whatever nth_element's order.
    #   Box z := [0.5, 0.5].
    # [0, 1]: x 2e10, y 2e10, z 0 -> Y.  nth_element over two photons sorts
    #   them (insertion sort): slot 0 = e (y 0.3), slot 1 = b (y 0.7); the
    #   median slot 0 = e gets Y; tempMin/Max = 0.3, 0.7; box y := [0.3, 0.3].
    #   [1, 1]: x 2e10, y 0, z 0 -> Z (X needs y > z, Y needs y > z).
    #   Back in [0, 1]: box y := [tempMin, tempMax] = [0.3, 0.7].
    # Root: box z := [0.5, 0.5] again.  [3, 4]: x 2e10 > y 0.4 > z 0 -> X.
    #   Sorted by x: slot 3 = c (x 0.2), slot 4 = a (x 0.8); slot 3 gets X;
    #   box x := [0.2, 0.2].  [4, 4]: x 0 > y 0.4 fails; y 0.4 > z 0 -> Y.
    # (Restoring the box's previous values instead of the range's end
    # photons would leave y at +-1e10 and give slot 3 Y and slot 4 Z.)
    pos = [[0.8, 0.1, 0.6], [0.1, 0.7, 0.2], [0.2, 0.9, 0.8], [0.5, 0.5, 0.5], [0.9, 0.3, 0.4]]
    return {"pos": [f32(p) for p in pos], "perm": [4, 1, 3, 2, 0], "axis": [1, 2, 2, 0, 1]}


def main():
    out = {"kd_snap": kd_snap(), "kd_distribution": kd_distribution(), "kd_leafsize": kd_leafsize(),
           "photon_axis": photon_axis()}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "oracle_pins.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
