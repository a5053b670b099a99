"""Lattice templates of the oracle vs textbook values and isotropy identities.

Pins ASL's VectorTemplate tables (src/math/aslTemplates.cxx:30-200): the
direction order, the weights lc_i/6 (+1 on rest) and the inverse table.
"""
import numpy as np
import pytest

from oracle import template


TEXTBOOK = {
    15: {0: 2 / 9, 1: 1 / 9, 7: 1 / 72},
    19: {0: 1 / 3, 1: 1 / 18, 7: 1 / 36},
    27: {0: 8 / 27, 1: 2 / 27, 7: 1 / 54, 19: 1 / 216},
}


@pytest.mark.parametrize("q", [15, 19, 27])
def test_weights_textbook(q):
    c, w, inv = template(q)
    assert c.shape == (q, 3)
    norm2 = (c * c).sum(1)
    for i in range(q):
        # weight class by |c|^2 (D3Q15 corners have |c|^2 = 3, starting at index 7)
        cls = {0: 0, 1: 1, 2: 7, 3: 19 if q == 27 else 7}[int(norm2[i])]
        assert w[i] == pytest.approx(TEXTBOOK[q][cls], rel=1e-15), (q, i)


@pytest.mark.parametrize("q", [15, 19, 27])
def test_isotropy(q):
    c, w, inv = template(q)
    assert abs(w.sum() - 1.0) < 1e-15
    assert np.abs((w[:, None] * c).sum(0)).max() < 1e-16
    second = np.einsum("i,ia,ib->ab", w, c, c)
    assert np.abs(second - np.eye(3) / 3).max() < 1e-15
    third = np.einsum("i,ia,ib,ic->abc", w, c, c, c)
    assert np.abs(third).max() < 1e-16
    # fourth-order isotropy (needed for the Navier-Stokes limit)
    fourth = np.einsum("i,ia,ib,ic,id->abcd", w, c, c, c, c)
    d = np.eye(3)
    iso = (np.einsum("ab,cd->abcd", d, d) + np.einsum("ac,bd->abcd", d, d) + np.einsum("ad,bc->abcd", d, d)) / 9
    assert np.abs(fourth - iso).max() < 1e-15


@pytest.mark.parametrize("q", [15, 19, 27])
def test_inverse_and_order(q):
    c, w, inv = template(q)
    assert (c[inv] == -c).all()
    assert (inv[inv] == np.arange(q)).all()
    # ASL order: rest, +x, +y, +z, -x, -y, -z
    assert (c[:7] == np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [-1, 0, 0], [0, -1, 0], [0, 0, -1]])).all()
    assert len({tuple(v) for v in c}) == q


def test_d3q19_edge_order():
    # d3q19() lists its edges as in src/math/aslTemplates.cxx:170-175
    c, _, _ = template(19)
    edges = [(1, 1, 0), (1, -1, 0), (-1, -1, 0), (-1, 1, 0), (0, 1, 1), (0, 1, -1), (0, -1, -1), (0, -1, 1),
             (1, 0, 1), (1, 0, -1), (-1, 0, -1), (-1, 0, 1)]
    assert [tuple(v) for v in c[7:]] == edges


def test_bad_template():
    with pytest.raises(ValueError):
        template(9)
