"""The Hilbert-curve index used to check the mixed traversal's packet order.

Pinned by the defining property of a Hilbert curve, not by restating it:
every cell of a 2^b grid gets a distinct index, and cells with consecutive
indices are face neighbours.  The curve is nested: the 21-bit indices of the
corners of the cells of a coarse grid order those cells as the coarse
indices do.
"""

import numpy as np
import pytest

import helpers as H


@pytest.mark.parametrize("bits", [1, 2, 3, 4])
def test_hilbert_visits_every_cell_through_face_neighbours(bits):
    side = 1 << bits
    g = np.arange(side ** 3)
    qx, qy, qz = g // (side * side), (g // side) % side, g % side
    k = H.hilbert_keys(qx, qy, qz, bits).astype(np.int64)
    assert np.array_equal(np.sort(k), np.arange(side ** 3))
    o = np.argsort(k)
    p = np.stack([qx[o], qy[o], qz[o]], 1)
    step = np.abs(np.diff(p, axis=0)).sum(1)
    assert np.all(step == 1)
    assert tuple(p[0]) == (0, 0, 0)


def test_hilbert_is_nested():
    bits = 3
    side = 1 << bits
    g = np.arange(side ** 3)
    qx, qy, qz = g // (side * side), (g // side) % side, g % side
    coarse = H.hilbert_keys(qx, qy, qz, bits)
    s = 21 - bits
    fine = H.hilbert_keys(qx << s, qy << s, qz << s, 21)
    assert np.array_equal(np.argsort(coarse, kind="stable"), np.argsort(fine, kind="stable"))
    assert np.array_equal(fine >> np.uint64(3 * s), coarse)


def test_morton_decode_inverts_the_key():
    rng = np.random.default_rng(5)
    q = rng.integers(0, 1 << 21, size=(3, 200))
    keys = np.array([H.morton_key(int(a), int(b), int(c), 21) for a, b, c in q.T],
                    dtype=np.uint64)
    d = H.morton_decode(keys, 21)
    for a in range(3):
        assert np.array_equal(d[a], q[a].astype(np.uint64))
