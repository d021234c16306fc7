"""Rounding bounds of the floating-point gift-wrapping steps of the star kernels
(wrap_select_fast / nearest_fast in cellgpu_amd/csrc/voronoi_kernels.hip).

A step is decided in floating point only when one candidate's interval
[t - e, t + e] lies wholly below every other's, so the decision equals the
exact wraps_before / nearer order exactly when e bounds the rounding error of
t. These tests check that claim against exact rational arithmetic on inputs
chosen to stress it (near-collinear, near-cocircular, wide magnitude ranges),
with the operations in the kernel's order (no fused multiply-add: the library
is built with -ffp-contract=off). The kernels' parity with the oracle on ties
and near-ties is covered by the GPU tests (test_lattice_ties_bit_exact).
"""
from fractions import Fraction

import numpy as np

U4 = 4.0 * 2.0 ** -53


def _t_and_e(cx, cy, ux, uy):
    """The kernel's centre parameter t of the circle through 0, cur = (ux, uy)
    and c = (cx, cy), and its error bound e; None where the kernel defers to
    the exact comparison."""
    with np.errstate(all="ignore"):
        px = cx * (cx - ux)
        py = cy * (cy - uy)
        qx = ux * cy
        qy = uy * cx
        num = px + py
        den = 2.0 * (qx - qy)
        en = U4 * (abs(px) + abs(py))
        ed = 2.0 * U4 * (abs(qx) + abs(qy))
        if not (2.0 * ed <= den):
            return None
        inv = 1.0 / den
        t = num * inv
        e = 4.0 * (abs(t) * ed + en) * inv + U4 * abs(t)
        if not (abs(t) < 1e300 and e < 1e300):
            return None
    return t, e


def _exact_t(cx, cy, ux, uy):
    cx, cy, ux, uy = (Fraction(float(v)) for v in (cx, cy, ux, uy))
    return (cx * (cx - ux) + cy * (cy - uy)) / (2 * (ux * cy - uy * cx))


def _cases(rng, n):
    for k in range(n):
        ux, uy = rng.uniform(-2, 2, 2) * 10.0 ** rng.integers(-3, 3)
        kind = k % 4
        if kind == 0:  # generic
            cx, cy = rng.uniform(-3, 3, 2)
        elif kind == 1:  # nearly on the line through 0 and cur, left of it
            a = rng.uniform(-2, 3)
            off = 10.0 ** rng.uniform(-15, -6)
            cx, cy = a * ux - off * uy, a * uy + off * ux
        elif kind == 2:  # nearly on the circle through 0 and cur centred at m
            ang = rng.uniform(0, 2 * np.pi)
            mx, my = ux / 2 - uy * 0.3, uy / 2 + ux * 0.3
            rr = np.hypot(mx, my)
            cx, cy = mx + rr * np.cos(ang) * (1 + 1e-14 * rng.standard_normal()), my + rr * np.sin(ang)
        else:  # magnitudes far apart
            cx, cy = rng.uniform(-1, 1, 2) * 10.0 ** rng.integers(-4, 4)
        yield float(cx), float(cy), float(ux), float(uy)


def test_centre_parameter_error_bound_holds():
    rng = np.random.default_rng(2024)
    checked = deferred = 0
    for cx, cy, ux, uy in _cases(rng, 20000):
        cross = Fraction(ux) * Fraction(cy) - Fraction(uy) * Fraction(cx)
        if cross <= 0:  # the kernel only ranks candidates strictly left of 0 -> cur
            continue
        r = _t_and_e(cx, cy, ux, uy)
        if r is None:
            deferred += 1
            continue
        t, e = r
        assert abs(Fraction(t) - _exact_t(cx, cy, ux, uy)) <= Fraction(e), (cx, cy, ux, uy)
        checked += 1
    assert checked > 8000 and deferred > 0


def test_decided_order_equals_the_circle_order():
    """Where two candidates' intervals are disjoint, their order is the exact
    order: a lies inside the circle through 0, cur and b exactly when t_a < t_b."""
    rng = np.random.default_rng(7)
    decided = 0
    for _ in range(4000):
        ux, uy = rng.uniform(-2, 2, 2)
        pts = []
        base = rng.uniform(-2, 2, 2)
        for k in range(2):
            p = base + (10.0 ** rng.uniform(-14, -1)) * rng.standard_normal(2) * (k > 0)
            pts.append((float(p[0]), float(p[1])))
        ra = _t_and_e(*pts[0], ux, uy)
        rb = _t_and_e(*pts[1], ux, uy)
        if ra is None or rb is None:
            continue
        (ta, ea), (tb, eb) = ra, rb
        if ta + ea < tb - eb:
            decided += 1
            assert _exact_t(*pts[0], ux, uy) < _exact_t(*pts[1], ux, uy)
        elif tb + eb < ta - ea:
            decided += 1
            assert _exact_t(*pts[1], ux, uy) < _exact_t(*pts[0], ux, uy)
    assert decided > 100


def test_squared_norm_error_bound_holds():
    rng = np.random.default_rng(11)
    for _ in range(20000):
        x, y = rng.uniform(-1, 1, 2) * 10.0 ** rng.integers(-5, 5, 2)
        n2 = x * x + y * y
        exact = Fraction(float(x)) ** 2 + Fraction(float(y)) ** 2
        assert abs(Fraction(n2) - exact) <= Fraction(U4 * n2)
