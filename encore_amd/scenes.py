"""Deterministic synthetic workloads for the ray-query engine.

Scenes are indexed triangle lists (xyz float32 (V,3), idx uint32 (T,3)); rays
are float32 (n,8) rows in the rq_ray layout (ox, oy, oz, tmin, dx, dy, dz,
tmax).  Every generator is seeded through a counter-based SplitMix64 stream so
results do not depend on the numpy version.

Ray formulas follow ENCORE where it defines them:
  eye rays        RayTracer::CalcEyeRaySamples (RayTracer.cpp:489-525) with
                  Camera::CalcCamVectors (Camera.cpp:157-166); up is always
                  (0,1,0) because Camera::set ignores its argument (:40-47)
  shadow rays     PhotonMapperCPU::IsInShadow (PhotonMapperCPU.cpp:1192-1207)
  bounce rays     cosine-weighted about the normal, origin pushed along the
                  new direction (PhotonMapperCPU.cpp:743-744)
"""
import math

import numpy as np

SEED_BASE = 0x454E434F52450000
_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(seed, counter):
    """SplitMix64 output for (seed, counter) pairs; counter is a uint64 array."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (np.asarray(counter, np.uint64) + np.uint64(1)) * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


class Stream:
    """Counter-based uniform stream: draw(n) gives the next n doubles in [0,1)."""

    def __init__(self, seed):
        self.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
        self.pos = 0

    def draw(self, n):
        c = np.arange(self.pos, self.pos + n, dtype=np.uint64)
        self.pos += n
        return (splitmix64(self.seed, c) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)

    def normal(self, n):
        u1 = np.maximum(self.draw(n), 1e-300)
        u2 = self.draw(n)
        return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


# ---------------------------------------------------------------------------
# ENCORE's float32 vector helpers (Vector3.h:137-162, Vector3.cpp:105-140)
# ---------------------------------------------------------------------------
def _flush(x):
    x = np.asarray(x, np.float32)
    return np.where((x > -0.001) & (x < 0.001), np.float32(0), x).astype(np.float32)


def cross_flush(a, b):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    x = a[..., 1] * b[..., 2] - a[..., 2] * b[..., 1]
    y = np.float32(-1) * (a[..., 0] * b[..., 2] - a[..., 2] * b[..., 0])
    z = a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]
    return np.stack([_flush(x), _flush(y), _flush(z)], axis=-1)


def normalize_flush(v):
    v = np.asarray(v, np.float32)
    m2 = (v[..., 0] * v[..., 0] + v[..., 1] * v[..., 1]) + v[..., 2] * v[..., 2]
    m = np.sqrt(m2).astype(np.float32)
    return np.stack([_flush(v[..., k] / m) for k in range(3)], axis=-1)


# ---------------------------------------------------------------------------
# Scenes
# ---------------------------------------------------------------------------
_QUAD = np.array([[0, 1, 2], [2, 3, 0]], np.uint32)
# Unit box faces in the winding of models/bwl/box.ply.
_BOX_FACES = np.array([[0, 5, 1], [0, 4, 5], [3, 2, 6], [6, 7, 3], [0, 7, 4], [0, 3, 7],
                       [2, 1, 5], [5, 6, 2], [4, 6, 5], [4, 7, 6], [0, 1, 2], [2, 3, 0]], np.uint32)
_BOX_VERTS = np.array([[0, 0, 0], [1, 0, 0], [1, 0, 1], [0, 0, 1],
                       [0, 1, 0], [1, 1, 0], [1, 1, 1], [0, 1, 1]], np.float64)


def _merge(parts):
    xyz, idx, base = [], [], 0
    for v, f in parts:
        xyz.append(np.asarray(v, np.float64))
        idx.append(np.asarray(f, np.uint32) + base)
        base += len(v)
    return np.concatenate(xyz).astype(np.float32), np.concatenate(idx).astype(np.uint32)


def cornell(seed=SEED_BASE + 1, scale=1.0):
    """34-triangle Cornell box: five walls at +-50 (2 triangles each) and two
    upright boxes (12 triangles each) standing on the floor.  The cube's edge
    is U(20,35); the 1x2x1 tall box is U(50,70) high; rotations about +y are
    U(-45,45) degrees; centres are U(-30,30) in x and z.  Geometry follows
    environment/scenes/diffuseCornell.txt:5-30.  `scale` multiplies every
    coordinate (config4_scene's photon-mapping box uses 1/20: 5 units wide)."""
    st = Stream(seed)
    h = 50.0
    walls = [
        [[-h, -h, -h], [-h, -h, h], [h, -h, h], [h, -h, -h]],  # floor
        [[-h, h, -h], [h, h, -h], [h, h, h], [-h, h, h]],      # ceiling
        [[-h, -h, -h], [h, -h, -h], [h, h, -h], [-h, h, -h]],  # back
        [[-h, -h, -h], [-h, h, -h], [-h, h, h], [-h, -h, h]],  # left
        [[h, -h, -h], [h, -h, h], [h, h, h], [h, h, -h]],      # right
    ]
    parts = [(np.array(w), _QUAD) for w in walls]
    r = st.draw(8)
    for bi, (lo, hi, tall) in enumerate([(20.0, 35.0, False), (50.0, 70.0, True)]):
        size = lo + (hi - lo) * r[4 * bi]
        ang = math.radians(-45.0 + 90.0 * r[4 * bi + 1])
        cx = -30.0 + 60.0 * r[4 * bi + 2]
        cz = -30.0 + 60.0 * r[4 * bi + 3]
        v = _BOX_VERTS.copy()
        if tall:
            v[:, 1] *= 2.0
            v *= size / 2.0
        else:
            v *= size
        v[:, 0] -= v[:, 0].mean()
        v[:, 2] -= v[:, 2].mean()
        c, s = math.cos(ang), math.sin(ang)
        x = c * v[:, 0] + s * v[:, 2]
        z = -s * v[:, 0] + c * v[:, 2]
        v = np.stack([x + cx, v[:, 1] - h, z + cz], axis=1)
        parts.append((v, _BOX_FACES))
    xyz, idx = _merge(parts)
    if scale != 1.0:
        xyz = (xyz.astype(np.float64) * scale).astype(np.float32)
    return xyz, idx


def _icosahedron():
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t],
                  [0, -1, -t], [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4],
                  [11, 10, 2], [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8],
                  [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], np.int64)
    return v / np.linalg.norm(v, axis=1, keepdims=True), f


def icosphere(level):
    """Unit icosphere: 20*4^level triangles, 10*4^level + 2 vertices."""
    v, f = _icosahedron()
    for _ in range(level):
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        e.sort(axis=1)
        key = e[:, 0] * (len(v) + 1) + e[:, 1]
        uk, inv = np.unique(key, return_inverse=True)
        a = uk // (len(v) + 1)
        b = uk % (len(v) + 1)
        mid = v[a] + v[b]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        m = (inv + len(v)).reshape(3, -1).T  # midpoints of edges 01, 12, 20
        v = np.concatenate([v, mid])
        f0, f1, f2 = f[:, 0], f[:, 1], f[:, 2]
        m01, m12, m20 = m[:, 0], m[:, 1], m[:, 2]
        f = np.concatenate([np.stack([f0, m01, m20], 1), np.stack([f1, m12, m01], 1),
                            np.stack([f2, m20, m12], 1), np.stack([m01, m12, m20], 1)])
    return v, f


def geodesic_sphere(freq):
    """Icosahedron with each face split into freq^2 triangles, on the unit
    sphere: 20*freq^2 triangles."""
    v0, f0 = _icosahedron()
    verts, faces = [], []
    # barycentric lattice per face; shared edges are welded by rounding keys.
    ii, jj = np.meshgrid(np.arange(freq + 1), np.arange(freq + 1), indexing="ij")
    mask = ii + jj <= freq
    ii, jj = ii[mask], jj[mask]
    lid = -np.ones((freq + 1, freq + 1), np.int64)
    lid[ii, jj] = np.arange(len(ii))
    # Per lattice row i: the upward triangles (i, j), (i+1, j), (i, j+1) for
    # j < freq - i, then the downward (i+1, j), (i+1, j+1), (i, j+1) for
    # j < freq - i - 1, interleaved up, down, up, ... along the row.
    rows = []
    for i in range(freq):
        j = np.arange(freq - i)
        up = np.stack([lid[i, j], lid[i + 1, j], lid[i, j + 1]], 1)
        jd = j[:-1]
        down = np.stack([lid[i + 1, jd], lid[i + 1, jd + 1], lid[i, jd + 1]], 1)
        row = np.empty((len(up) + len(down), 3), np.int64)
        row[0::2] = up
        row[1::2] = down
        rows.append(row)
    tri = np.concatenate(rows)
    base = 0
    for a, b, c in f0:
        p = (v0[a][None] * (freq - ii - jj)[:, None] + v0[b][None] * ii[:, None] + v0[c][None] * jj[:, None]) / freq
        verts.append(p)
        faces.append(tri + base)
        base += len(p)
    v = np.concatenate(verts)
    f = np.concatenate(faces)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    key = np.round(v * 1e9).astype(np.int64)
    # Unique rows in lexicographic order, each with its first occurrence (what
    # np.unique(key, axis=0, return_index=True, return_inverse=True) returns),
    # through a stable lexsort.
    order = np.lexsort((key[:, 2], key[:, 1], key[:, 0]))
    ks = key[order]
    start = np.ones(len(ks), bool)
    start[1:] = np.any(ks[1:] != ks[:-1], axis=1)
    first = order[start]
    inv = np.empty(len(ks), np.int64)
    inv[order] = np.cumsum(start) - 1
    return v[first], inv[f]


def value_noise3(p, seed, freqs=(4.0, 8.0, 16.0), amps=(1.0, 0.5, 0.25)):
    """Sum of value-noise octaves at points p (n,3), normalised to [-1,1].
    Lattice values are hashed from (seed, octave, cell) with SplitMix64 and
    interpolated trilinearly with a smoothstep fade.  The lattice values of
    an octave are hashed once over the cells the points span and gathered."""
    out = np.zeros(len(p))
    for o, (fr, am) in enumerate(zip(freqs, amps)):
        q = p * fr
        c = np.floor(q)
        w = q - c
        w = w * w * (3.0 - 2.0 * w)
        ci = c.astype(np.int64)
        lo = ci.min(axis=0)
        ext = ci.max(axis=0) - lo + 2  # lattice points per axis, corners included
        gx, gy, gz = np.meshgrid(np.arange(ext[0]) + lo[0], np.arange(ext[1]) + lo[1], np.arange(ext[2]) + lo[2],
                                 indexing="ij")
        key = ((gx & 0xFFFFF) << 40) | ((gy & 0xFFFFF) << 20) | (gz & 0xFFFFF)
        hv = splitmix64(seed + 0x1000 * (o + 1), key.reshape(-1).astype(np.uint64))
        table = (hv >> np.uint64(11)).astype(np.float64) * (2.0 / 9007199254740992.0) - 1.0
        rel = ci - lo
        flat = (rel[:, 0] * ext[1] + rel[:, 1]) * ext[2] + rel[:, 2]
        acc = np.zeros(len(p))
        for dx in (0, 1):
            for dy in (0, 1):
                for dz in (0, 1):
                    val = table[flat + (dx * ext[1] + dy) * ext[2] + dz]
                    wx = w[:, 0] if dx else 1.0 - w[:, 0]
                    wy = w[:, 1] if dy else 1.0 - w[:, 1]
                    wz = w[:, 2] if dz else 1.0 - w[:, 2]
                    acc += val * wx * wy * wz
        out += am * acc
    return out / sum(amps)


def displaced_sphere(level=None, freq=None, seed=SEED_BASE + 2, amplitude=0.1, world_scale=100.0,
                     ground=True):
    """Icosphere (level) or geodesic sphere (freq) displaced radially by
    r = 1 + amplitude * noise(p), scaled by world_scale, optionally above a
    two-triangle ground plane at y = -1.2 S spanning +-10 S."""
    if level is not None:
        v, f = icosphere(level)
    else:
        v, f = geodesic_sphere(freq)
    r = 1.0 + amplitude * value_noise3(v, seed)
    v = v * r[:, None] * world_scale
    parts = [(v, f)]
    if ground:
        S = world_scale
        g = np.array([[-10 * S, -1.2 * S, -10 * S], [-10 * S, -1.2 * S, 10 * S],
                      [10 * S, -1.2 * S, 10 * S], [10 * S, -1.2 * S, -10 * S]])
        parts.append((g, _QUAD))
    return _merge(parts)


# ---------------------------------------------------------------------------
# Rays
# ---------------------------------------------------------------------------
def camera_basis(eye, look):
    """(n, u, v) of Camera::CalcCamVectors with up = (0,1,0), float32."""
    eye = np.asarray(eye, np.float32)
    look = np.asarray(look, np.float32)
    n = (eye - look).astype(np.float32)
    up = np.array([0, 1, 0], np.float32)
    u = cross_flush(up, n)
    n = normalize_flush(n)
    u = normalize_flush(u)
    v = cross_flush(n, u)
    return n, u, v


def eye_rays(cols, rows, eye, look, fov_deg=45.0, aspect=None, near=0.1):
    """One pixel-centre ray per pixel, row-major from row 0, with the float32
    operation order of RayTracer::CalcEyeRaySamples (PI = 3.14159f)."""
    if aspect is None:
        aspect = cols / rows
    near = np.float32(near)
    H = np.float32(float(near) * math.tan(float(np.float32(fov_deg)) * (float(np.float32(3.14159)) / 360.0)))
    W = np.float32(H * np.float32(aspect))
    n, u, v = camera_basis(eye, look)
    col = np.arange(cols, dtype=np.float32)
    row = np.arange(rows, dtype=np.float32)
    x = col + np.float32(0.5)
    y = row + np.float32(0.5)
    Uc = -W + (W * (np.float32(2) * x) / np.float32(cols))
    Vr = -H + (H * (np.float32(2) * y) / np.float32(rows))
    Uc = np.broadcast_to(Uc[None, :], (rows, cols)).reshape(-1).astype(np.float32)
    Vr = np.broadcast_to(Vr[:, None], (rows, cols)).reshape(-1).astype(np.float32)
    nn = n * (-near)
    d = (nn[None, :] + u[None, :] * Uc[:, None]).astype(np.float32)
    d = (d + v[None, :] * Vr[:, None]).astype(np.float32)
    d = normalize_flush(d)
    rays = np.zeros((len(d), 8), np.float32)
    rays[:, 0:3] = np.asarray(eye, np.float32)
    rays[:, 3] = 0.0
    rays[:, 4:7] = d
    rays[:, 7] = np.inf
    return rays


def hit_points(xyz, idx, hits):
    """ENCORE's hitPoint = v0 + (v1-v0)*alpha + (v2-v0)*beta (TrianglePrim.cpp:239)
    and the normal Normalize(Cross(e1, e2)) (:236-238) for hit rows; hits is
    (n,4) float32 with the prim id in column 3 as uint32 bits."""
    prim = np.ascontiguousarray(hits[:, 3]).view(np.uint32)
    ok = prim != 0xFFFFFFFF
    tri = idx[prim[ok]]
    v0, v1, v2 = xyz[tri[:, 0]], xyz[tri[:, 1]], xyz[tri[:, 2]]
    a = hits[ok, 1:2].astype(np.float32)
    b = hits[ok, 2:3].astype(np.float32)
    e1 = (v1 - v0).astype(np.float32)
    e2 = (v2 - v0).astype(np.float32)
    p = (v0 + e1 * a).astype(np.float32)
    p = (p + e2 * b).astype(np.float32)
    nrm = normalize_flush(cross_flush(e1, e2))
    return ok, p, nrm


PI32 = np.float32(3.14159)  # ENCORE's PI (Common.h:36)


def hemisphere_dirs(normals, ns, stream, jitter=0.5):
    """Final-gather sample directions: Sampler(ns, ns, 1) with
    SetJitterPercent(jitter) and GetSampleRaysForHemisphere (Sampler.cpp:23-47,
    101-186, 217-231) for every normal, ns*ns unit vectors per normal in the
    reference's order (phi division outer, theta division inner).  theta and
    phi are cell centres of a (pi/2) x (2 pi) grid plus a jitter of up to
    `jitter` of half a cell (two draws per sample, theta first); the point is
    converted to Cartesian around +z, rotated onto the normal by
    R^T Rx(acos n_z) R (TAffineMatrix::GetRotationMatrix) with R's rows the
    rotation axis Cross(+z, n) and two vectors orthogonal to it, and
    normalised.  The reference draws the jitter with rand(); here it comes
    from `stream`.  Float32 throughout."""
    f = np.float32
    nrm = normalize_flush(np.asarray(normals, np.float32).reshape(-1, 3))
    n = len(nrm)
    m = ns * ns
    theta_div = f(f(1) / f(ns)) * f(PI32 / f(2))
    phi_div = f(f(1) / f(ns)) * f(f(2) * PI32)
    ht, hp = f(theta_div / f(2)), f(phi_div / f(2))
    row, col = np.meshgrid(np.arange(ns), np.arange(ns), indexing="ij")
    theta = (theta_div * col.reshape(-1).astype(np.float32) + ht).astype(np.float32)
    phi = (phi_div * row.reshape(-1).astype(np.float32) + hp).astype(np.float32)
    theta = np.broadcast_to(theta, (n, m)).copy()
    phi = np.broadcast_to(phi, (n, m)).copy()
    if jitter > 0:
        u = stream.draw(2 * n * m).reshape(n, m, 2).astype(np.float32)
        theta = (theta + ((u[..., 0] * f(2) - f(1)) * ht) * f(jitter)).astype(np.float32)
        phi = (phi + ((u[..., 1] * f(2) - f(1)) * hp) * f(jitter)).astype(np.float32)
    st = np.sin(theta)
    p = np.stack([st * np.cos(phi), st * np.sin(phi), np.cos(theta)], axis=-1).astype(np.float32)
    p = normalize_flush(p)
    out = p.copy()
    z = np.array([0, 0, 1], np.float32)
    same = np.all(nrm == z, axis=1)
    flip = np.all(nrm == -z, axis=1)
    out[flip] = (p[flip] + p[flip] * f(-2)).astype(np.float32)
    rot = ~(same | flip)
    ang = np.arccos(nrm[:, 2]).astype(np.float32)  # Dot(+z, n) = n_z
    rot &= ang != 0
    if rot.any():
        r = normalize_flush(cross_flush(np.broadcast_to(z, (int(rot.sum()), 3)), nrm[rot]))
        a = np.abs(r)
        s_ = np.zeros_like(r)
        xs = (a[:, 0] <= a[:, 1]) & (a[:, 0] <= a[:, 2])
        ys = ~xs & (a[:, 1] <= a[:, 0]) & (a[:, 1] <= a[:, 2])
        zs = ~xs & ~ys
        s_[xs] = np.stack([np.zeros(xs.sum(), np.float32), -r[xs, 2], r[xs, 1]], axis=1)
        s_[ys] = np.stack([-r[ys, 2], np.zeros(ys.sum(), np.float32), r[ys, 0]], axis=1)
        s_[zs] = np.stack([-r[zs, 1], r[zs, 0], np.zeros(zs.sum(), np.float32)], axis=1)
        s_ = normalize_flush(s_)
        t_ = normalize_flush(cross_flush(r, s_))
        R = np.stack([r, s_, t_], axis=1)  # rows r, s, t
        c, sn = np.cos(ang[rot]).astype(np.float32), np.sin(ang[rot]).astype(np.float32)
        X = np.zeros((len(c), 3, 3), np.float32)
        X[:, 0, 0] = 1
        X[:, 1, 1], X[:, 1, 2], X[:, 2, 1], X[:, 2, 2] = c, -sn, sn, c

        def mul(A, B):  # float32, sums in k order
            C = np.zeros_like(A)
            for k in range(3):
                C = (C + A[:, :, k:k + 1] * B[:, k:k + 1, :]).astype(np.float32)
            return C

        M = mul(mul(np.transpose(R, (0, 2, 1)), X), R)
        q = p[rot]
        out[rot] = (((M[:, None, :, 0] * q[..., 0:1] + M[:, None, :, 1] * q[..., 1:2]).astype(np.float32)
                     + M[:, None, :, 2] * q[..., 2:3])).astype(np.float32)
    return normalize_flush(out).reshape(n * m, 3)


def shadow_rays(points, light_center, light_size, stream):
    """One ray per point toward a uniform point on a square light facing -y;
    origin pushed by 0.01 along the ray; tmax = |L - p| measured from the
    un-pushed point (PhotonMapperCPU.cpp:1192-1207)."""
    n = len(points)
    r = stream.draw(2 * n).reshape(n, 2)
    L = np.empty((n, 3), np.float32)
    L[:, 0] = light_center[0] + (r[:, 0] - 0.5) * light_size
    L[:, 1] = light_center[1]
    L[:, 2] = light_center[2] + (r[:, 1] - 0.5) * light_size
    dvec = (L - points).astype(np.float32)
    dist = np.sqrt((dvec[:, 0] * dvec[:, 0] + dvec[:, 1] * dvec[:, 1]) + dvec[:, 2] * dvec[:, 2]).astype(np.float32)
    d = normalize_flush(dvec)
    o = (points + d * np.float32(0.01)).astype(np.float32)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = o
    rays[:, 4:7] = d
    rays[:, 7] = dist
    return rays


def cosine_dirs(normals, stream):
    """Cosine-weighted directions about unit normals (float64 math)."""
    n = len(normals)
    r = stream.draw(2 * n).reshape(n, 2)
    cos_t = np.sqrt(r[:, 0])
    sin_t = np.sqrt(1.0 - r[:, 0])
    phi = 2.0 * np.pi * r[:, 1]
    N = np.asarray(normals, np.float64)
    N = N / np.linalg.norm(N, axis=1, keepdims=True)
    a = np.where(np.abs(N[:, 0:1]) > 0.9, np.array([[0.0, 1.0, 0.0]]), np.array([[1.0, 0.0, 0.0]]))
    T = np.cross(a, N)
    T /= np.linalg.norm(T, axis=1, keepdims=True)
    B = np.cross(N, T)
    d = T * (sin_t * np.cos(phi))[:, None] + B * (sin_t * np.sin(phi))[:, None] + N * cos_t[:, None]
    return d / np.linalg.norm(d, axis=1, keepdims=True)


def bounce_rays(points, normals, incoming, count, stream, offset):
    """`count` diffuse-bounce rays from points sampled with replacement: the
    normal is flipped to face the incoming ray, the direction is
    cosine-weighted about it and the origin is pushed by `offset` along it."""
    pick = np.minimum((stream.draw(count) * len(points)).astype(np.int64), len(points) - 1)
    P = points[pick].astype(np.float64)
    N = normals[pick].astype(np.float64)
    D_in = incoming[pick].astype(np.float64)
    flip = np.sum(N * D_in, axis=1) > 0
    N[flip] = -N[flip]
    d = cosine_dirs(N, stream)
    rays = np.zeros((count, 8), np.float32)
    rays[:, 0:3] = (P + d * offset).astype(np.float32)
    rays[:, 4:7] = d.astype(np.float32)
    rays[:, 7] = np.inf
    return rays


def lognormal_batches(total, count, stream, median=65536, sigma=1.18, lo=4096, hi=1048576):
    """`count` batch sizes: exp(ln median + sigma z) clipped to [lo, hi], then
    rescaled (and re-clipped) until they sum to exactly `total`."""
    if not count * lo <= total <= count * hi:
        raise ValueError("%d batches of %d..%d rays cannot hold %d rays" % (count, lo, hi, total))
    z = stream.normal(count)
    s = np.clip(np.exp(np.log(median) + sigma * z), lo, hi)
    for _ in range(64):
        s = np.clip(s * (total / s.sum()), lo, hi)
    s = np.floor(s).astype(np.int64)
    diff = int(total - s.sum())
    i = 0
    while diff != 0:
        j = i % count
        step = 1 if diff > 0 else -1
        if lo <= s[j] + step <= hi:
            s[j] += step
            diff -= step
        i += 1
    return s


# ---------------------------------------------------------------------------
# Named workloads
# ---------------------------------------------------------------------------
def config1(seed=SEED_BASE + 1, res=512):
    """Cornell box (34 triangles) and res x res pixel-centre eye rays from
    (0,0,170) looking at the origin, fov 45, aspect 1, near 0.1."""
    xyz, idx = cornell(seed)
    rays = eye_rays(res, res, (0.0, 0.0, 170.0), (0.0, 0.0, 0.0), 45.0, 1.0, 0.1)
    return xyz, idx, rays


def config2_scene(level=8, seed=SEED_BASE + 2, world_scale=100.0):
    """Displaced icosphere (level 8: 1,310,720 triangles) plus a 2-triangle
    ground, world scale S = 100."""
    return displaced_sphere(level=level, seed=seed, world_scale=world_scale)


def config2_primary(cols=1920, rows=1080, world_scale=100.0):
    S = world_scale
    return eye_rays(cols, rows, (0.0, 0.3 * S, 3.5 * S), (0.0, 0.0, 0.0), 45.0, cols / rows, 0.1)


def config2_shadow(xyz, idx, primary_rays, primary_hits, seed=SEED_BASE + 2, world_scale=100.0):
    S = world_scale
    ok, p, nrm = hit_points(xyz, idx, primary_hits)
    st = Stream(seed + 0x5000)
    return ok, shadow_rays(p, (2 * S, 4 * S, 2 * S), 1.0 * S, st)


def config3_rays(xyz, idx, primary_rays, primary_hits, count=8388608, nbatches=64,
                 seed=SEED_BASE + 3, world_scale=100.0):
    """Incoherent secondary rays from primary hit points and the 64 lognormal
    batch sizes that partition them."""
    ok, p, nrm = hit_points(xyz, idx, primary_hits)
    st = Stream(seed)
    rays = bounce_rays(p, nrm, primary_rays[ok, 4:7], count, st, 1e-3 * world_scale)
    sizes = lognormal_batches(count, nbatches, Stream(seed + 0x77))
    return rays, sizes


# ---------------------------------------------------------------------------
# Photon mapping workload
# ---------------------------------------------------------------------------
def tessellate(xyz, idx, f):
    """Split every triangle into f*f triangles on a barycentric grid (vertices
    are not shared between input triangles)."""
    ii, jj = np.meshgrid(np.arange(f + 1), np.arange(f + 1), indexing="ij")
    mask = ii + jj <= f
    ii, jj = ii[mask], jj[mask]
    lid = -np.ones((f + 1, f + 1), np.int64)
    lid[ii, jj] = np.arange(len(ii))
    tri = []
    for i in range(f):
        for j in range(f - i):
            tri.append([lid[i, j], lid[i + 1, j], lid[i, j + 1]])
            if i + j + 1 < f:
                tri.append([lid[i + 1, j], lid[i + 1, j + 1], lid[i, j + 1]])
    tri = np.array(tri, np.int64)
    v = xyz.astype(np.float64)[idx]  # (T,3,3)
    w0 = (f - ii - jj)[None, :, None] / f
    w1 = ii[None, :, None] / f
    w2 = jj[None, :, None] / f
    pts = v[:, 0:1, :] * w0 + v[:, 1:2, :] * w1 + v[:, 2:3, :] * w2  # (T, P, 3)
    P = len(ii)
    out_idx = (tri[None, :, :] + (np.arange(len(idx)) * P)[:, None, None]).reshape(-1, 3)
    return pts.reshape(-1, 3).astype(np.float32), out_idx.astype(np.uint32)


def config4_scene(scale=1.0 / 20.0, tess=32, seed=SEED_BASE + 1):
    """The Cornell box of config 1 at world scale `scale` with each triangle
    split tess x tess (34 * 1024 = 34,816 triangles).  At 1/20 the box is
    5 units wide; with 2,000,000 photons a 0.05 gather radius reaches tens
    to a hundred of them, and the tessellated edges (0.04-0.3) keep the
    products of two edges mostly clear of Cross()'s 1e-3 flush
    (Vector3.h:148-160), so faces keep well-defined normals.  At 1/100 the walls'
    edges would be about 0.03 long, their products about 1e-3: normals
    would flush to zero or to one axis."""
    xyz, idx = cornell(seed, scale=scale)
    return tessellate(xyz, idx, tess)


# Materials of environment/scenes/diffuseCornell.txt:5-30 in cornell()'s
# triangle order: floor, ceiling, back (0.7 grey), left (red), right (blue),
# then the two boxes (0.7, 0.7, 0.5).
CORNELL_MATERIALS = [
    dict(diffuse=(0.7, 0.7, 0.7)),
    dict(diffuse=(0.6, 0.36, 0.36)),
    dict(diffuse=(0.36, 0.36, 0.6)),
    dict(diffuse=(0.7, 0.7, 0.5)),
]
# fgCompCornell.txt:28-33's outer walls: no diffuse reflection, refracting
# (transparency 0.8, index 0.5, transmittance 0.3).
GLASS = dict(diffuse=(0.0, 0.0, 0.0), transmittance=0.3, transparency=0.8, refraction_index=0.5)


def cornell_prim_materials(tess=1):
    """Material index of every triangle of cornell() (tessellated tess x tess
    by tessellate(), which keeps children next to their parent)."""
    ids = np.array([0, 0, 0, 0, 0, 0, 1, 1, 2, 2] + [3] * 24, np.uint32)
    return np.repeat(ids, tess * tess)


def cornell_light(scale=1.0 / 20.0, power=(1.0, 1.0, 1.0)):
    """ENCORE's RECT_LIGHT (RectangleLight.cpp:14-80): a 30 x 30 square at
    y = 49.99999 facing down, centre the mean of its corners, at world scale
    `scale`."""
    y = np.float32(49.99999)
    cy = np.float32(np.float32(np.float32(y + y) + y) + y) / np.float32(4.0)
    s = np.float32(scale)
    return dict(center=(0.0, float(np.float32(cy * s)), 0.0), normal=(0.0, -1.0, 0.0),
                length=float(np.float32(30.0) * s), width=float(np.float32(30.0) * s), power=power)


def glass_shell(scale=1.0 / 20.0, tess=1):
    """fgCompCornell's outer walls: ceiling, floor, back, right and left at
    +-55 spanning 150, normals facing the inside (the winding of the cornell_*
    PLY models).  Returns xyz, idx."""
    h, e = 55.0, 75.0
    walls = [
        [[-e, h, -e], [e, h, -e], [e, h, e], [-e, h, e]],      # ceiling (normal -y)
        [[-e, -h, -e], [-e, -h, e], [e, -h, e], [e, -h, -e]],  # floor (+y)
        [[-e, -e, -h], [e, -e, -h], [e, e, -h], [-e, e, -h]],  # back (+z)
        [[h, -e, -e], [h, -e, e], [h, e, e], [h, e, -e]],      # right (-x)
        [[-h, -e, -e], [-h, e, -e], [-h, e, e], [-h, -e, e]],  # left (+x)
    ]
    xyz, idx = _merge([(np.array(w), _QUAD) for w in walls])
    xyz = (xyz.astype(np.float64) * scale).astype(np.float32)
    return tessellate(xyz, idx, tess) if tess > 1 else (xyz, idx)


def photon_scene(tess=32, glass=False, scale=1.0 / 20.0):
    """The config-4 Cornell box with materials (and, with `glass`, the
    refracting outer shell of fgCompCornell).  Returns xyz, idx, materials,
    prim_material, lights."""
    xyz, idx = config4_scene(scale=scale, tess=tess)
    mats = list(CORNELL_MATERIALS)
    pm = cornell_prim_materials(tess)
    if glass:
        gx, gi = glass_shell(scale, tess=max(1, tess // 4))
        idx = np.concatenate([idx, gi + len(xyz)]).astype(np.uint32)
        xyz = np.concatenate([xyz, gx]).astype(np.float32)
        mats.append(GLASS)
        pm = np.concatenate([pm, np.full(len(gi), len(mats) - 1, np.uint32)])
    return xyz, idx, mats, pm, [cornell_light(scale)]


def normals_of(xyz, idx, prims):
    """ENCORE's hit normal Normalize(Cross(e1, e2)) (TrianglePrim.cpp:236-238)."""
    tri = idx[prims]
    v0, v1, v2 = xyz[tri[:, 0]], xyz[tri[:, 1]], xyz[tri[:, 2]]
    e1 = (v1 - v0).astype(np.float32)
    e2 = (v2 - v0).astype(np.float32)
    return normalize_flush(cross_flush(e1, e2))


def trace_photons(xyz, idx, trace, count, seed=SEED_BASE + 4, scale=1.0 / 20.0, light_frac=0.3, bounces=3,
                  albedo=0.7, batch=262144, with_bounce=False):
    """Photons from a square light centred under the ceiling (side light_frac of
    the box, emitting downward with cosine weighting), stored at every
    diffuse hit; a photon survives each hit with probability `albedo` and
    bounces at most `bounces` times.  `trace(rays) -> hits (n,4)` is the
    closest-hit query.  Returns positions, normals, powers (count x 3), and
    with `with_bounce` the bounce index of each photon (0 = first hit)."""
    st = Stream(seed)
    half = 50.0 * scale
    side = 2 * half * light_frac
    pos_out, nrm_out, pow_out, bnc_out = [], [], [], []
    stored = 0
    emitted = 0
    while stored < count:
        n = batch
        r = st.draw(2 * n).reshape(n, 2)
        o = np.empty((n, 3))
        o[:, 0] = (r[:, 0] - 0.5) * side
        o[:, 1] = half * 0.999
        o[:, 2] = (r[:, 1] - 0.5) * side
        d = cosine_dirs(np.tile([[0.0, -1.0, 0.0]], (n, 1)), st)
        rays = np.zeros((n, 8), np.float32)
        rays[:, 0:3] = o
        rays[:, 4:7] = d
        rays[:, 7] = np.inf
        power = np.full((n, 3), 1.0, np.float32)
        emitted += n
        for bounce in range(bounces + 1):
            hits = trace(rays)
            ok, p, nrm = hit_points(xyz, idx, hits)
            prims = np.ascontiguousarray(hits[:, 3]).view(np.uint32)[ok]
            good = ~np.isnan(nrm).any(axis=1)
            pos_out.append(p[good])
            nrm_out.append(nrm[good])
            pow_out.append(power[ok][good])
            bnc_out.append(np.full(int(good.sum()), bounce, np.uint8))
            stored += int(good.sum())
            surv = st.draw(len(p)) < albedo
            keep = good & surv
            if not keep.any() or stored >= count:
                break
            N = nrm[keep].astype(np.float64)
            D_in = rays[ok, 4:7][keep].astype(np.float64)
            flip = np.sum(N * D_in, axis=1) > 0
            N[flip] = -N[flip]
            nd = cosine_dirs(N, st)
            rays = np.zeros((len(nd), 8), np.float32)
            rays[:, 0:3] = (p[keep].astype(np.float64) + nd * 0.001).astype(np.float32)
            rays[:, 4:7] = nd.astype(np.float32)
            rays[:, 7] = np.inf
            power = (power[ok][keep] * np.float32(albedo)).astype(np.float32)
    P = np.concatenate(pos_out)[:count]
    Nn = np.concatenate(nrm_out)[:count]
    W = np.concatenate(pow_out)[:count] / np.float32(emitted / 1e5)
    if with_bounce:
        # bounce 0 = first hit from the light (ENCORE's direct map), >= 1 the
        # indirect map
        return P.astype(np.float32), Nn.astype(np.float32), W.astype(np.float32), np.concatenate(bnc_out)[:count]
    return P.astype(np.float32), Nn.astype(np.float32), W.astype(np.float32)


def photon_queries(xyz, idx, trace, res=1000, scale=1.0 / 20.0):
    """Gather points: the hits of a res x res view from the config-1 camera
    (scaled), with the hit triangle's normal."""
    rays = eye_rays(res, res, (0.0, 0.0, 170.0 * scale), (0.0, 0.0, 0.0), 45.0, 1.0, 0.1 * scale)
    hits = trace(rays)
    ok, p, nrm = hit_points(xyz, idx, hits)
    return p.astype(np.float32), nrm.astype(np.float32)
