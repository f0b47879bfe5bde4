"""Shared parity helpers for the tests (compare engine output with the oracle)."""
import sys

import numpy as np

import oracle

TIE_REL = 1e-6


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def split_hits(h):
    h = np.ascontiguousarray(h, np.float32)
    return h[:, 0], h[:, 1], h[:, 2], h[:, 3].view(np.uint32)


def prim_t(xyz, idx, rays, prims):
    """t, alpha, beta, hit of ENCORE's predicate for (ray i, triangle prims[i])."""
    ok = prims != oracle.NO_HIT
    tris = np.zeros((len(rays), 9), np.float32)
    tri = idx[prims[ok]]
    tris[ok] = np.concatenate([xyz[tri[:, 0]], xyz[tri[:, 1]], xyz[tri[:, 2]]], axis=1)
    r = oracle.tri_intersect_pairs(rays, tris)
    r[~ok] = np.nan
    return r


def close_rel(a, b, rel=1e-5, abs_tol=1e-6):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.abs(a - b) <= np.maximum(rel * np.abs(b), abs_tol)


EDGE_BAND = 1e-5


def _near_edge(u, v, prim):
    if prim == oracle.NO_HIT:
        return False
    w = 1.0 - float(u) - float(v)
    return min(float(u), float(v), w) < EDGE_BAND


def classify_closest(xyz, idx, rays, got, ref, brute=None, bit_exact=True):
    """Compare closest-hit outputs.  Returns dict of counts:
    exact        same prim and bit-identical t, u, v
    tie          different prims whose t agree within 1e-6 t and both are
                 genuine hits of the predicate at those t
    ref_defect   reference misses the brute-force closest hit and `got`
                 finds it (only if `brute` is given)
    edge_band    (tolerance mode only) either hit lies within 1e-5 of a
                 triangle edge in barycentric terms
    bad          anything else (must be 0)."""
    gt, gu, gv, gp = split_hits(got)
    rt, ru, rv, rp = split_hits(ref)
    same = gp == rp
    bits_ok = same & (gt.view(np.uint32) == rt.view(np.uint32)) & (gu.view(np.uint32) == ru.view(np.uint32)) & (
        gv.view(np.uint32) == rv.view(np.uint32))
    miss_both = same & (gp == oracle.NO_HIT)
    if bit_exact:
        exact = bits_ok | miss_both
    else:
        # t within 1e-5 relative; barycentrics within 1e-5 relative or 1e-3
        # absolute near zero (where ENCORE's flush shifts them).
        with np.errstate(invalid="ignore"):
            exact = miss_both | (same & close_rel(gt, rt) & close_rel(gu, ru, abs_tol=1e-3) &
                                 close_rel(gv, rv, abs_tol=1e-3))
    out = dict(n=len(rays), exact=int(exact.sum()), tie=0, ref_defect=0, phantom_outside_box=0, edge_band=0, bad=0,
               bad_idx=[])
    for i in np.nonzero(~exact)[0]:
        if not bit_exact and (_near_edge(gu[i], gv[i], gp[i]) or _near_edge(ru[i], rv[i], rp[i])):
            # A ray through a shared edge or vertex: ENCORE's predicate is not
            # watertight there (it can miss both triangles or hit both).
            out["edge_band"] += 1
            continue
        if gp[i] == rp[i]:
            out["bad"] += 1
            out["bad_idx"].append(int(i))
            continue
        if gp[i] != rp[i] and gp[i] != oracle.NO_HIT and rp[i] != oracle.NO_HIT:
            chk = prim_t(xyz, idx, rays[i:i + 1], np.array([gp[i]], np.uint32))[0]
            genuine = chk[3] > 0 and (chk[0] == gt[i] or not bit_exact)
            if genuine and abs(gt[i] - rt[i]) <= TIE_REL * abs(rt[i]):
                out["tie"] += 1
                continue
        if rp[i] != oracle.NO_HIT and (gp[i] == oracle.NO_HIT or gt[i] > rt[i]):
            # The reference's (closer) hit lies outside its own triangle's
            # bounding box: a flush phantom of ENCORE's predicate that no BVH
            # reaches, while `got` is an exact predicate hit (or none).
            if outside_box(xyz, idx, rays[i], rt[i], rp[i]) > 0 and (
                    gp[i] == oracle.NO_HIT or _exact_hit(xyz, idx, rays[i], gt[i], gp[i])):
                out["phantom_outside_box"] += 1
                continue
        if brute is not None:
            bt, _, _, bp = split_hits(brute[i:i + 1])
            g_is_brute = (gp[i] == bp[0]) or (gp[i] != oracle.NO_HIT and bp[0] != oracle.NO_HIT and
                                               abs(gt[i] - bt[0]) <= TIE_REL * abs(bt[0]))
            if g_is_brute:
                out["ref_defect"] += 1
                continue
        out["bad"] += 1
        out["bad_idx"].append(int(i))
    return out


def outside_box(xyz, idx, ray, t, prim):
    """How far (world units) the point o + t d lies outside triangle prim's
    bounding box, less a rounding allowance of 1e-4 of its magnitude."""
    tri = xyz[idx[prim]].astype(np.float64)
    P = ray[0:3].astype(np.float64) + float(t) * ray[4:7].astype(np.float64)
    out = (np.maximum(tri.min(0) - P, 0) + np.maximum(P - tri.max(0), 0)).max()
    return out - 1e-4 * max(1.0, np.abs(P).max())


def _exact_hit(xyz, idx, ray, t, prim):
    chk = prim_t(xyz, idx, ray[None, :], np.array([prim], np.uint32))[0]
    return chk[3] > 0 and np.float32(chk[0]).view(np.uint32) == np.float32(t).view(np.uint32)


def fp64_adjudicate(xyz, idx, rays, got, bad_idx):
    """Of the rays in bad_idx, those where `got` is the FP64 brute-force
    closest hit (same prim, or t within 1e-6 t): there ENCORE's float32 flush
    arithmetic, not the engine, departs from the exact answer."""
    if not len(bad_idx):
        return 0, []
    bad_idx = np.asarray(bad_idx)
    f, fp = oracle.fp64_closest(xyz, idx, rays[bad_idx])
    gt, _, _, gp = split_hits(got[bad_idx])
    agree = (gp == fp) | ((gp != oracle.NO_HIT) & (fp != oracle.NO_HIT) & (np.abs(gt - f[:, 0]) <= 1e-6 * np.abs(f[:, 0])))
    return int(agree.sum()), bad_idx[~agree].tolist()


def brute_occluded(xyz, idx, rays):
    """Any triangle hit with 0 < t < tmax (strict), by brute force."""
    lim = np.nextafter(rays[:, 7].astype(np.float32), np.float32(0))
    return (oracle.count_hits_below(xyz, idx, rays, lim, 0.0) > 0).astype(np.uint32)
