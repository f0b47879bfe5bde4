"""The oracle's irradiance cache (IrrCacheOctree.h, PhotonMapperCPU.cpp:
1087-1103) against closed forms built from the source text: the weight
1 / (|p - x| / Ri + sqrt(1 - n.N)) against the strict bound 1 / a, the
Ri = 0 and n == N special cases, SampleInFront, AddElement's bounds and its
non-terminating case, and the insertion-order sums."""
import numpy as np

import oracle

A = 0.08
LO, SIZE = (-75.0, -75.0, -75.0), (150.0, 150.0, 150.0)
UP = (0.0, 1.0, 0.0)


def _cache(elems, a=A):
    c = oracle.IrrCache(a, LO, SIZE)
    st = c.add([e[0] for e in elems], [e[1] for e in elems], [e[2] for e in elems], [e[3] for e in elems])
    return c, st


def test_weight_closed_forms():
    # Ri = 16, n == N: sqrt term 0, w = Ri / d.  1 / 0.08f rounds to 12.5.
    assert np.float32(1) / np.float32(A) == np.float32(12.5)
    c, st = _cache([((1, 2, 3), (0, 0, 0), UP, 16.0)])
    assert st.tolist() == [0]
    q = np.array([[1, 0, 0], [2, 0, 0], [1.27, 0, 0], [1.28, 0, 0]], np.float32)
    rgb, ws, cnt = c.query(q, np.tile(np.float32(UP), (4, 1)))
    # d = 1: w = 16 exactly, the average is the value itself
    assert cnt[0] == 1 and ws[0] == 16.0 and rgb[0].tolist() == [1, 2, 3]
    # d = 2: w = 8 < 12.5
    assert cnt[1] == 0 and ws[1] == 0 and rgb[1].tolist() == [0, 0, 0]
    # d = 1.27: w = 12.598 > 12.5; d = 1.28f: w = 1 / (1.28f / 16) = 12.5, not above (strict >)
    assert cnt[2] == 1 and abs(ws[2] - 16 / 1.27) < 1e-5
    assert cnt[3] == 0


def test_normal_term():
    # n = (0, 0.99995, 0.01): n.N = 0.99995, sqrt term sqrt(5e-5) ~ 0.00707;
    # at d = 0.5 (Ri = 16) w = 1 / (0.03125 + 0.00707) ~ 26.1.  A missing
    # square root (1 - n.N = 5e-5) would give 31.9; a missing normal term 32.
    c, _ = _cache([((1, 1, 1), (0, 0, 0), UP, 16.0)])
    n = np.array([[0, 0.99995, 0.01]], np.float32)
    rgb, ws, cnt = c.query(np.array([[0.5, 0, 0]], np.float32), n)
    want = 1.0 / (0.5 / 16 + np.sqrt(1.0 - float(np.float32(0.99995))))
    assert cnt[0] == 1 and abs(ws[0] - want) < 1e-4 * want


def test_zero_ri_reaches_everywhere():
    # Ri == 0: the distance term is 0 (IrrCacheOctree.h:62), so only the normal
    # term decides -- at any distance; with n == N the denominator is 0 and
    # the weight 0.
    c, st = _cache([((5, 5, 5), (0, 0, 0), UP, 0.0)])
    assert st.tolist() == [0]
    q = np.array([[40, 0, 0], [40, 0, 0]], np.float32)
    n = np.array([[0, 0.99995, 0.01], UP], np.float32)
    rgb, ws, cnt = c.query(q, n)
    assert cnt.tolist() == [1, 0]
    assert rgb[0].tolist() == [5, 5, 5] and ws[1] == 0


def test_sample_in_front():
    # Below the cached sample's surface (p - x along -N): Dot(Normalize(p - x),
    # N + n) = -2 < -0.1, so it does not count although its weight (16) does.
    c, _ = _cache([((1, 1, 1), (0, 0, 0), UP, 16.0)])
    q = np.array([[0, -1, 0], [0, 1, 0]], np.float32)
    rgb, ws, cnt = c.query(q, np.tile(np.float32(UP), (2, 1)))
    assert cnt.tolist() == [0, 1]


def test_add_bounds_and_refusal():
    # Outside the root box: dropped.  2 * (a * Ri) above the x size (150):
    # AddElement would descend without end -> refused.  Ri = 937.5 gives
    # 2 * (0.08f * 937.5) = 149.99999 <= 150 -> stored at the root.
    c, st = _cache([((1, 1, 1), (80, 0, 0), UP, 1.0), ((1, 1, 1), (0, 0, 0), UP, 1000.0),
                    ((1, 1, 1), (0, 0, 0), UP, 937.5), ((1, 1, 1), (75, -75, 75), UP, 1.0),
                    ((1, 1, 1), (0, 0, 0), UP, float("nan"))])
    assert st.tolist() == [1, 2, 0, 0, 2]
    assert len(c) == 2


def test_insertion_order_sum():
    # Three values at distance 1 with Ri = 16 (w = 16 each), values 1e8, -1e8,
    # 1 in that order: (1.6e9 - 1.6e9) + 16 = 16, over a weight sum of 48 ->
    # 1/3.  Summed in reverse the 16 would be absorbed by -1.6e9 -> 0.
    elems = [((1e8,) * 3, (1, 0, 0), UP, 16.0), ((-1e8,) * 3, (-1, 0, 0), UP, 16.0), ((1,) * 3, (0, 0, 1), UP, 16.0)]
    c, _ = _cache(elems)
    rgb, ws, cnt = c.query(np.zeros((1, 3), np.float32), np.float32([UP]))
    assert cnt[0] == 3 and ws[0] == 48.0
    assert rgb[0].tolist() == [float(np.float32(16) / np.float32(48))] * 3
