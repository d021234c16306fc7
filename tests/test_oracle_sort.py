"""Pins of the oracle's restatement of the reference's spatial sorting
(HilbertSorter, inc/utility/HilbertSort.h:27-104; Simple2DCell.cpp:438-599),
CPU only.

The curve is pinned by what fixes a Hilbert curve: on every grid up to 64 x 64
the index is a bijection onto 0 .. n^2 - 1, consecutive indices are unit steps
(the defining continuity), the order-1 curve visits (0,0), (0,1), (1,1), (1,0)
in that order and every curve starts at (0,0) and ends at (n-1, 0), which fixes
the orientation among the eight symmetric curves. The curve order and the pixel
follow the reference's box arithmetic and are checked at hand-evaluated points.
The vertex-model cell order (first appearance) is checked on a hand-built case.
"""
import numpy as np


def test_curve_order_follows_box_side(oracle):
    # 2^mm >= L with mm >= 1, then mm + 4 (capped at 30)
    assert oracle.hilbert_curve_order(0.5) == 5
    assert oracle.hilbert_curve_order(2.0) == 5
    assert oracle.hilbert_curve_order(2.5) == 6
    assert oracle.hilbert_curve_order(31.6) == 9
    assert oracle.hilbert_curve_order(32.0) == 9
    assert oracle.hilbert_curve_order(512.0) == 13
    assert oracle.hilbert_curve_order(2.0 ** 40) == 30


def test_curve_is_a_continuous_bijection_with_fixed_orientation(oracle):
    assert [int(oracle.hilbert_xy2d(1, x, y)) for x, y in ((0, 0), (0, 1), (1, 1), (1, 0))] == [0, 1, 2, 3]
    for M in range(1, 7):
        n = 1 << M
        X, Y = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
        d = oracle.hilbert_xy2d(M, X.ravel(), Y.ravel())
        assert np.array_equal(np.sort(d), np.arange(n * n))
        o = np.argsort(d)
        px, py = X.ravel()[o], Y.ravel()[o]
        assert np.all(np.abs(np.diff(px)) + np.abs(np.diff(py)) == 1)
        assert (px[0], py[0]) == (0, 0) and (px[-1], py[-1]) == (n - 1, 0)


def test_pixel_and_key_at_hand_evaluated_points(oracle):
    L = 4.0  # order 6 (2^2 >= 4), 64 pixels per side, 16 per unit length
    assert oracle.hilbert_curve_order(L) == 6
    # (0.1, 0.1) -> pixel (1, 1); (3.99, 0.01) -> pixel (63, 0): the curve's end
    k = oracle.hilbert_keys([0.1, 3.99, 0.0], [0.1, 0.01, 0.0], L)
    assert k[1] == 64 * 64 - 1 and k[2] == 0
    assert k[0] == oracle.hilbert_xy2d(6, 1, 1) == 2


def test_sort_order_breaks_ties_by_index(oracle):
    x = np.array([0.5, 0.01, 0.5, 0.02])
    y = np.array([0.5, 0.01, 0.5, 0.01])
    # L = 2: order 5, 16 pixels per unit length. Cells 1 and 3 share pixel
    # (0, 0) (index 0), cells 0 and 2 pixel (8, 8): each pair keeps its order.
    perm = oracle.hilbert_sort_order(x, y, 2.0)
    assert list(perm) == [1, 3, 0, 2]


def test_vertex_cell_order_is_first_appearance(oracle):
    # three vertices already in curve order along the box's first pixel row
    vx = np.array([0.01, 0.2, 0.4])
    vy = np.array([0.01, 0.01, 0.01])
    vcn = np.array([[5, 2, 7], [2, 9, 1], [1, 7, 0]])
    vperm, cperm = oracle.vertex_sort_orders(vx, vy, vcn, 4.0)
    assert list(vperm) == [0, 1, 2]
    assert list(cperm) == [5, 2, 7, 9, 1, 0]
