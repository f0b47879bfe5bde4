"""The irradiance-cache lookup (encore_amd/csrc/irrcache.hip, ic_query)
decides SampleInFront (IrrCacheOctree.h:279-290) without normalising u when
|g|_1 <= 6 and the unnormalised dot product r = u.g satisfies r >= 0 or
r^2 <= 0.008 |u|^2; the kernel's comment derives that the reference's test
then always passes.  This checks the derivation in float32 over random
offsets aimed at the -0.1 boundary (normals equal, close and unrelated,
components near the 0.001 flush, lengths from 1e-4 to 1e3): wherever the
condition holds, the reference's own operations accept the lookup too.  A
condition loosened to r^2 <= 0.0125 |u|^2 (|r| up to 0.112 |u|) fails."""
import numpy as np

f = np.float32


def _flush(x):
    return np.where((x > f(-0.001)) & (x < f(0.001)), f(0), x).astype(f)


def _reference_front(u, g):
    m = np.sqrt((u[:, 0] * u[:, 0] + u[:, 1] * u[:, 1]) + u[:, 2] * u[:, 2])
    with np.errstate(all="ignore"):
        v = _flush(u / m[:, None])
    d = (v[:, 0] * g[:, 0] + v[:, 1] * g[:, 1]) + v[:, 2] * g[:, 2]
    return ~(d < f(-0.1))


def _shortcut(u, g, c):
    d2 = (u[:, 0] * u[:, 0] + u[:, 1] * u[:, 1]) + u[:, 2] * u[:, 2]
    r = (u[:, 0] * g[:, 0] + u[:, 1] * g[:, 1]) + u[:, 2] * g[:, 2]
    g1 = (np.abs(g[:, 0]) + np.abs(g[:, 1])) + np.abs(g[:, 2])
    with np.errstate(all="ignore"):
        return (g1 <= f(6)) & ((r >= f(0)) | (r * r <= f(c) * d2))


def _cases(rng, n, kind):
    def unit():
        v = rng.normal(size=(n, 3))
        return (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(f)

    ni, nq = unit(), unit()
    if kind == 1:
        nq = ni.copy()
    elif kind == 2:
        nq = ni + f(0.05) * unit()
        nq = (nq / np.linalg.norm(nq, axis=1, keepdims=True)).astype(f)
    g = (ni + nq).astype(f)
    gn = g / np.maximum(np.linalg.norm(g, axis=1, keepdims=True), 1e-12)
    t = rng.normal(size=(n, 3))
    t = t - np.sum(t * gn, axis=1, keepdims=True) * gn
    t = t / np.maximum(np.linalg.norm(t, axis=1, keepdims=True), 1e-12)
    s = rng.uniform(-0.08, 0.02, size=(n, 1))  # sine of the angle below the tangent plane
    scale = np.exp(rng.uniform(np.log(1e-4), np.log(1e3), size=(n, 1)))
    u = ((t * np.sqrt(1 - s * s) + gn * s) * scale).astype(f)
    u[::7] = _flush(u[::7] * f(1e-3))
    return u, g


def _violations(c, rounds):
    rng = np.random.default_rng(11)
    bad = acc = 0
    for it in range(rounds):
        u, g = _cases(rng, 100000, it % 3)
        sc = _shortcut(u, g, c)
        bad += int(np.sum(sc & ~_reference_front(u, g)))
        acc += int(np.sum(sc))
    return bad, acc


def test_front_shortcut_never_accepts_what_the_reference_rejects():
    bad, acc = _violations(0.008, 30)
    assert acc > 1_000_000
    assert bad == 0


def test_front_shortcut_loosened_fails():
    bad, _ = _violations(0.0125, 6)
    assert bad > 0
