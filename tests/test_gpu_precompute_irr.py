"""PhotonMap::PrecomputeIrradiance (PhotonMap.cpp:57-74) on the GPU
(rq_photons_precompute_irradiance): the photons it keeps are every
increment-th of ENCORE's balanced order (the oracle's balance, pinned by the
hand-worked axis case and the median invariants), each with the radiance
estimate of the map at its own position and normal as its power (the
oracle's estimate within 1e-5 relative, the bar of every radiance test here:
the double sums run in another order), and the new map is those photons,
added in that order and balanced."""
import numpy as np
import pytest

import oracle
from encore_amd import rq
from helpers import log

pytestmark = pytest.mark.gpu


def _photons(rng, n):
    """Photons on three axis planes of a unit box, normals +x, +y, +z."""
    P = rng.uniform(0.0, 1.0, (n, 3)).astype(np.float32)
    N = np.zeros((n, 3), np.float32)
    ax = rng.integers(0, 3, n)
    P[np.arange(n), ax] = 0.0
    N[np.arange(n), ax] = 1.0
    W = rng.uniform(0.1, 1.0, (n, 3)).astype(np.float32)
    return P, N, W


@pytest.mark.parametrize("increment,k,r,caustic,layout", [(1, 8, 0.05, False, "spatial"),
                                                          (7, 50, 0.08, False, "encore"),
                                                          (3, 100, 0.1, True, "spatial")])
def test_precompute_irradiance_matches_oracle(device, increment, k, r, caustic, layout):
    rng = np.random.default_rng(31)
    P, N, W = _photons(rng, 30000)
    pm = rq.PhotonMap(device, P, N, W, layout=layout)
    made, chosen, power = pm.precompute_irradiance(P, N, increment, k, r, caustic=caustic, layout="encore")
    ref = oracle.PhotonMap(P, N, W)
    perm, _ = ref.order()
    want = perm[::increment]
    log("case: precompute increment %d k=%d r=%g caustic=%s layout %s: %d photons" % (
        increment, k, r, caustic, layout, len(want)))
    assert len(chosen) == len(want) == (len(P) - 1) // increment + 1
    assert np.array_equal(chosen, want)
    rgb = ref.radiance(np.ascontiguousarray(P[want]), np.ascontiguousarray(N[want]), k, r, caustic=caustic)
    assert (rgb > 0).all(axis=1).mean() > 0.99  # every photon finds at least itself
    np.testing.assert_allclose(power, rgb, rtol=1e-5, atol=1e-12)
    # the new map: the kept photons in that order, balanced as BalanceAndConstruct does
    perm_made, _, _ = made.order()
    perm_ref, _ = oracle.PhotonMap(np.ascontiguousarray(P[want]), np.ascontiguousarray(N[want]), power).order()
    assert np.array_equal(perm_made, perm_ref)
    made.destroy()
    pm.destroy()


def test_precompute_irradiance_arguments(device):
    rng = np.random.default_rng(32)
    P, N, W = _photons(rng, 100)
    pm = rq.PhotonMap(device, P, N, W)
    for bad in (dict(increment=0, k=8), dict(increment=1, k=0), dict(increment=1, k=300)):
        with pytest.raises(rq.RQError):
            pm.precompute_irradiance(P, N, bad["increment"], bad["k"], 0.1)
    with pytest.raises(rq.RQError):  # the map's photons, not another count
        pm.precompute_irradiance(P[:50], N[:50], 1, 8, 0.1)
    made, chosen, _ = pm.precompute_irradiance(P, N, 1000, 8, 0.1)  # one photon: the root
    assert len(chosen) == 1 and chosen[0] == oracle.PhotonMap(P, N, W).order()[0][0]
    made.destroy()
    pm.destroy()
