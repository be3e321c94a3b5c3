"""Pins the oracle's velocity sets (Lattice.h:145-210, 535-612, 614-703).

Checks the isotropy sums the reference's own tests state
(test/TestLattice.cpp:438-531 D2Q9, 1301-1395 D3Q19, 1920-2016 D3Q27), the
opposite-pair block structure (1..faceQ have c_x = -1, faceQ+1..2faceQ are
their mirrors) and the table entries quoted from the headers.
"""
import itertools

import numpy as np
import pytest

LATTICES = [(3, 9, 2, 3), (9, 19, 3, 5), (10, 27, 3, 9)]  # (enum, Q, D, faceQ)


@pytest.mark.parametrize("lat,Q,D,faceQ", LATTICES)
def test_isotropy(oracle, lat, Q, D, faceQ):
    c = oracle.velocities(lat).astype(float)
    w = oracle.weights(lat)
    assert c.shape == (Q, D)
    cs2 = 1.0 / 3.0
    assert abs(w.sum() - 1.0) < 1e-15
    np.testing.assert_allclose(w @ c, 0.0, atol=1e-16)
    M2 = np.einsum("i,ia,ib->ab", w, c, c)
    np.testing.assert_allclose(M2, cs2 * np.eye(D), atol=1e-16)
    M3 = np.einsum("i,ia,ib,ic->abc", w, c, c, c)
    np.testing.assert_allclose(M3, 0.0, atol=1e-16)
    M4 = np.einsum("i,ia,ib,ic,id->abcd", w, c, c, c, c)
    d = np.eye(D)
    iso = cs2 ** 2 * (np.einsum("ab,cd->abcd", d, d) + np.einsum("ac,bd->abcd", d, d)
                      + np.einsum("ad,bc->abcd", d, d))
    np.testing.assert_allclose(M4, iso, atol=1e-16)
    M5 = np.einsum("i,ia,ib,ic,id,ie->abcde", w, c, c, c, c, c)
    np.testing.assert_allclose(M5, 0.0, atol=1e-16)


@pytest.mark.parametrize("lat,Q,D,faceQ", LATTICES)
def test_block_structure(oracle, lat, Q, D, faceQ):
    c = oracle.velocities(lat)
    w = oracle.weights(lat)
    assert (c[0] == 0).all() and w[0] == w.max()
    for i in range(1, faceQ + 1):
        assert c[i, 0] == -1
        np.testing.assert_array_equal(c[i + faceQ], -c[i])   # opposite(i) = i + faceQ
        assert w[i] == w[i + faceQ]
    for i in range(2 * faceQ + 1, Q):
        assert c[i, 0] == 0
    # every velocity appears once and the set is closed under negation
    rows = {tuple(r) for r in c}
    assert len(rows) == Q
    assert all(tuple(-np.array(r)) in rows for r in rows)


def test_quoted_entries(oracle):
    # D2Q9 header order (Lattice.h:188-206); note c4 = (1,-1), which the old
    # test file TestLattice.cpp:595 expects differently: the header wins.
    c = oracle.velocities(3)
    assert [tuple(r) for r in c] == [(0, 0), (-1, 1), (-1, 0), (-1, -1), (1, -1),
                                    (1, 0), (1, 1), (0, -1), (0, 1)]
    np.testing.assert_array_equal(oracle.weights(3) * 36,
                                  [16, 1, 4, 1, 1, 4, 1, 4, 4])
    c19 = oracle.velocities(9)
    assert tuple(c19[2]) == (-1, -1, 0) and tuple(c19[4]) == (-1, 0, -1)
    assert tuple(c19[13]) == (0, -1, -1) and tuple(c19[18]) == (0, 1, -1)
    c27 = oracle.velocities(10)
    assert tuple(c27[6]) == (-1, -1, -1) and tuple(c27[26]) == (0, 1, -1)
    w27 = oracle.weights(10) * 216
    np.testing.assert_allclose(sorted(set(np.round(w27, 9))), [1, 4, 16, 64])


@pytest.mark.parametrize("lat,Q,D,faceQ", LATTICES)
def test_weights_by_speed(oracle, lat, Q, D, faceQ):
    c = oracle.velocities(lat)
    w = oracle.weights(lat)
    by = {}
    for ci, wi in zip(c, w):
        by.setdefault(int((ci ** 2).sum()), set()).add(wi)
    assert all(len(s) == 1 for s in by.values())
    expect = {3: {0: 4 / 9, 1: 1 / 9, 2: 1 / 36},
              9: {0: 1 / 3, 1: 1 / 18, 2: 1 / 36},
              10: {0: 8 / 27, 1: 2 / 27, 2: 1 / 54, 3: 1 / 216}}[lat]
    for k, v in expect.items():
        assert by[k] == {v}
