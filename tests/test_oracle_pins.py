"""CPU pins of the FP64 oracle (no GPU).

Each check is chosen so that a plausible slip in the oracle (dropped term,
wrong sign or index, transposed operand) fails it:
  * exact predicates vs rational arithmetic (random and near-degenerate);
  * periodic Delaunay rings vs scipy Qhull on a 9-sheet copy (golden fixture)
    and a brute-force empty-circumcircle check;
  * Euler invariants (sum of degrees 6N, 2N owned triangles, sum of areas L^2);
  * a centred-rectangular lattice with closed-form area, perimeter and F = 0;
  * forces = -grad E by central differences; sum F = 0; the reference's two
    force paths (GPU force sets, CPU force formula) agree;
  * self-propelled / Brownian closed forms at F = 0 (incl. the GPU-path lag);
  * FIRE returns a perturbed lattice to the energy minimum.
"""
import os
from fractions import Fraction

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "delaunay_rings.npz")


# ------------------------------------------------------------------ predicates
def _exact_orient(a, b):
    ax, ay, bx, by = map(Fraction, (a[0], a[1], b[0], b[1]))
    v = ax * by - ay * bx
    return (v > 0) - (v < 0)


def _exact_incircle(b, c, d):
    bx, by, cx, cy, dx, dy = map(Fraction, (b[0], b[1], c[0], c[1], d[0], d[1]))
    b2, c2, d2 = bx * bx + by * by, cx * cx + cy * cy, dx * dx + dy * dy
    det = bx * (cy * d2 - c2 * dy) - by * (cx * d2 - c2 * dx) + b2 * (cx * dy - cy * dx)
    return -((det > 0) - (det < 0))  # inside <=> det < 0 for CCW (0, b, c)


def test_predicates_random_and_degenerate(oracle):
    rng = np.random.default_rng(1)
    for _ in range(400):
        a, b, c = rng.uniform(-2, 2, (3, 2))
        assert oracle.orient0(a, b) == _exact_orient(a, b)
        if _exact_orient(b, c) > 0:
            d = rng.uniform(-2, 2, 2)
            assert oracle.incircle0(b, c, d) == _exact_incircle(b, c, d)
    # near-degenerate: d on the circle through (0, b, c) up to one ulp, collinear triples
    for _ in range(300):
        b = rng.uniform(-1, 1, 2)
        c = rng.uniform(-1, 1, 2)
        if _exact_orient(b, c) <= 0:
            b, c = c, b
        if _exact_orient(b, c) == 0:
            continue
        cc = oracle.circumcenter0(b, c)
        r = np.hypot(*cc)
        t = rng.uniform(0, 2 * np.pi)
        d = cc + r * np.array([np.cos(t), np.sin(t)])
        d = d + rng.integers(-2, 3, 2) * np.spacing(d)
        assert oracle.incircle0(b, c, d) == _exact_incircle(b, c, d)
        s = rng.uniform(0.1, 3.0)
        e = b * s + rng.integers(-1, 2, 2) * np.spacing(b * s)
        assert oracle.orient0(b, e) == _exact_orient(b, e)


# ------------------------------------------------------------------ Delaunay
def _rings(nn, nbr):
    return [tuple(int(v) for v in nbr[i, : nn[i]]) for i in range(len(nn))]


@pytest.mark.parametrize("key", ["N64_s13377", "N256_s1", "N1000_s13377", "N2000_s42"])
def test_triangulation_matches_qhull_fixture(oracle, key):
    g = np.load(GOLD)
    x, y, L = g[key + "_x"], g[key + "_y"], float(g[key + "_L"][0])
    nn_ref, flat = g[key + "_nn"], g[key + "_rings"]
    ref, o = [], 0
    for z in nn_ref:
        ref.append(tuple(int(v) for v in flat[o: o + z]))
        o += z
    nn, nbr, nm = oracle.triangulate(x, y, L, cap=32)
    assert _rings(nn, nbr) == ref
    assert nn.sum() == 6 * len(x)
    assert nm == (nn.max() + 2 if nn.max() % 2 == 0 else nn.max() + 1)


def test_fixture_positions_are_reference_mt19937(oracle):
    # std::mt19937(13377) + uniform_real_distribution<double>, x then y, then directors
    g = np.load(GOLD)
    x, y, th, L = oracle.random_init(1000, 13377)
    assert np.array_equal(x, g["N1000_s13377_x"]) and np.array_equal(y, g["N1000_s13377_y"])
    assert 0.0 <= th.min() and th.max() < 2 * np.pi


def test_brute_force_empty_circumcircles(oracle):
    N = 48
    x, y, th, L = oracle.random_init(N, 99)
    nn, nbr, _ = oracle.triangulate(x, y, L)
    P = np.stack([x, y], 1)
    for i in range(N):
        for k in range(nn[i]):
            j, l = nbr[i, k], nbr[i, (k + 1) % nn[i]]
            dj = oracle.min_dist(L, P[j], P[i])
            dl = oracle.min_dist(L, P[l], P[i])
            assert _exact_orient(dj, dl) > 0  # counter-clockwise ring
            for m in range(N):
                if m in (i, j, l):
                    continue
                for sx in (-L, 0, L):  # every periodic image near the cell
                    for sy in (-L, 0, L):
                        dm = P[m] - P[i] + np.array([sx, sy])
                        if np.hypot(*dm) > L / 2:
                            continue
                        assert _exact_incircle(dj, dl, dm) <= 0


def test_area_sum_and_symmetry(oracle):
    N = 500
    x, y, th, L = oracle.random_init(N, 5)
    s = oracle.VoronoiSystem(x, y, th, L)
    s.compute_forces()
    g = s.get()
    assert abs(g["A"].sum() - L * L) < 1e-9 * L * L
    nn, nbr, _ = s.topology()
    rings = _rings(nn, nbr)
    for i, r in enumerate(rings):  # neighbour relation is symmetric
        for j in r:
            assert i in rings[j]


# ------------------------------------------------------------------ lattice closed forms
def _lattice(nx, ny, L):
    ax, ay = L / nx, L / ny
    pts = [((i + 0.5 * (j % 2) + 0.25) * ax, (j + 0.5) * ay) for j in range(ny) for i in range(nx)]
    p = np.array(pts)
    return p[:, 0].copy(), p[:, 1].copy(), ax, ay


def test_centred_rectangular_lattice_closed_form(oracle):
    nx, ny = 12, 12
    L = 12.0
    x, y, ax, ay = _lattice(nx, ny, L)
    N = nx * ny
    k = (ay * ay - ax * ax / 4) / (2 * ay)
    h = (ay * ay + ax * ax / 4) / (2 * ay)
    A_cf = ax * ay
    P_cf = 4 * k + 4 * np.hypot(ax / 2, h - k)
    s = oracle.VoronoiSystem(x, y, np.zeros(N), L, A0=A_cf, P0=P_cf)
    nn, _, _ = s.topology()
    assert (nn == 6).all()
    E = s.compute_forces(gpu_style=True)
    g = s.get()
    np.testing.assert_allclose(g["A"], A_cf, rtol=1e-12)
    np.testing.assert_allclose(g["P"], P_cf, rtol=1e-12)
    assert abs(E) < 1e-20
    # with other preferences all cells are still identical: F = 0 by symmetry
    s.set_prefs(1.0, 3.0)
    s.compute_forces(gpu_style=True)
    g = s.get()
    assert np.abs(g["Fx"]).max() < 1e-12 and np.abs(g["Fy"]).max() < 1e-12
    any_, flags = s.test_flags()
    assert not any_


# ------------------------------------------------------------------ forces
def _energy_at(oracle, s, x, y):
    s.set_state(x, y)
    return s.compute_forces(gpu_style=True)


def test_forces_are_minus_energy_gradient(oracle):
    N = 200
    x, y, th, L = oracle.random_init(N, 21)
    s = oracle.VoronoiSystem(x, y, th, L, P0=3.7)
    # relax a little so that no triangle is close to degenerate
    for _ in range(30):
        s.spp_step(np.zeros(N))
    g = s.get()
    x0, y0 = g["x"].copy(), g["y"].copy()
    _energy_at(oracle, s, x0, y0)
    F = s.get()
    h = 1e-6
    rng = np.random.default_rng(3)
    for i in rng.choice(N, 12, replace=False):
        for comp in ("x", "y"):
            xp, yp = x0.copy(), y0.copy()
            (xp if comp == "x" else yp)[i] += h
            Ep = _energy_at(oracle, s, xp, yp)
            xm, ym = x0.copy(), y0.copy()
            (xm if comp == "x" else ym)[i] -= h
            Em = _energy_at(oracle, s, xm, ym)
            fd = -(Ep - Em) / (2 * h)
            an = F["F" + comp][i]
            assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)), (i, comp, fd, an)


def test_force_sum_zero_and_two_reference_paths_agree(oracle):
    N = 800
    x, y, th, L = oracle.random_init(N, 8)
    s = oracle.VoronoiSystem(x, y, th, L, P0=3.85)
    Eg = s.compute_forces(gpu_style=True)
    g = s.get()
    Ec = s.compute_forces(gpu_style=False)
    c = s.get()
    scale = np.abs(np.concatenate([g["Fx"], g["Fy"]])).max()
    assert abs(g["Fx"].sum()) < 1e-10 * scale * N and abs(g["Fy"].sum()) < 1e-10 * scale * N
    assert np.abs(g["Fx"] - c["Fx"]).max() < 1e-12 * scale
    assert np.abs(g["Fy"] - c["Fy"]).max() < 1e-12 * scale
    assert abs(Eg - Ec) <= 1e-13 * abs(Eg)  # area sums run in a different vertex order


# ------------------------------------------------------------------ integrators
def test_spp_gpu_lag_at_zero_force(oracle):
    nx = ny = 10
    L = 10.0
    x, y, ax, ay = _lattice(nx, ny, L)
    N = nx * ny
    k = (ay * ay - ax * ax / 4) / (2 * ay)
    h = (ay * ay + ax * ax / 4) / (2 * ay)
    A_cf, P_cf = ax * ay, 4 * k + 4 * np.hypot(ax / 2, h - k)
    th0 = 0.3
    v0, Dr, dt = 0.2, 1.0, 0.01
    s = oracle.VoronoiSystem(x, y, np.full(N, th0), L, A0=A_cf, P0=P_cf, v0=v0, Dr=Dr, dt=dt)
    etas = [0.7, -1.1, 0.4]
    th = th0
    th_hist = [th0]
    for e in etas:  # uniform noise: the lattice translates rigidly, F stays 0
        s.spp_step(np.full(N, e), gpu_path=True)
        th = np.arctan2(np.sin(th), np.cos(th)) + e * np.sqrt(2 * dt * Dr)
        th_hist.append(th)
    g = s.get()
    # displacement of step n uses the director of step n-2 (theta_0 for steps 1 and 2)
    dx = dt * v0 * (np.cos(th_hist[0]) * 2 + np.cos(th_hist[1]))
    dy = dt * v0 * (np.sin(th_hist[0]) * 2 + np.sin(th_hist[1]))
    ex = (x + dx) % L
    ey = (y + dy) % L
    assert np.abs(g["x"] - ex).max() < 1e-12 and np.abs(g["y"] - ey).max() < 1e-12
    np.testing.assert_allclose(g["theta"], th_hist[-1], rtol=0, atol=1e-12)


def test_spp_cpu_path_has_no_lag(oracle):
    nx = ny = 10
    L = 10.0
    x, y, ax, ay = _lattice(nx, ny, L)
    N = nx * ny
    k = (ay * ay - ax * ax / 4) / (2 * ay)
    h = (ay * ay + ax * ax / 4) / (2 * ay)
    s = oracle.VoronoiSystem(x, y, np.full(N, 0.3), L, A0=ax * ay, P0=4 * k + 4 * np.hypot(ax / 2, h - k),
                             v0=0.2, Dr=1.0, dt=0.01)
    s.spp_step(np.full(N, 0.5), gpu_path=False)
    s.spp_step(np.full(N, 0.5), gpu_path=False)
    g = s.get()
    th1 = 0.3 + 0.5 * np.sqrt(0.02)
    dx = 0.01 * 0.2 * (np.cos(0.3) + np.cos(th1))
    assert np.abs(g["x"] - (x + dx) % L).max() < 1e-12


def test_brownian_zero_force(oracle):
    nx = ny = 8
    L = 8.0
    x, y, ax, ay = _lattice(nx, ny, L)
    N = nx * ny
    k = (ay * ay - ax * ax / 4) / (2 * ay)
    h = (ay * ay + ax * ax / 4) / (2 * ay)
    s = oracle.VoronoiSystem(x, y, np.zeros(N), L, A0=ax * ay, P0=4 * k + 4 * np.hypot(ax / 2, h - k), dt=0.005)
    ex, ey = np.full(N, 0.3), np.full(N, -0.8)
    T = 1e-3
    s.brownian_step(ex, ey, T)
    g = s.get()
    sc = np.sqrt(0.005 * T * 1.0 * 2.0)
    assert np.abs(g["x"] - (x + 0.3 * sc) % L).max() < 1e-13
    assert np.abs(g["y"] - (y - 0.8 * sc) % L).max() < 1e-13


def test_fire_relaxes_perturbed_lattice(oracle):
    nx = ny = 8
    L = 8.0
    x, y, ax, ay = _lattice(nx, ny, L)
    N = nx * ny
    k = (ay * ay - ax * ax / 4) / (2 * ay)
    h = (ay * ay + ax * ax / 4) / (2 * ay)
    rng = np.random.default_rng(4)
    xp = x + rng.normal(0, 0.01, N)
    yp = y + rng.normal(0, 0.01, N)
    s = oracle.VoronoiSystem(xp, yp, np.zeros(N), L, A0=ax * ay, P0=4 * k + 4 * np.hypot(ax / 2, h - k))
    f = oracle.FireMinimizer(s, max_it=20000, force_cutoff=1e-8, dt=0.01, dt_max=0.1, alpha_start=0.1)
    E0 = s.compute_forces()
    it, st = f.minimize()
    E = s.compute_forces()
    assert st["force_max"] ** 0.5 < 1e-8  # GPU-branch test: RMS force (sum |F|^2 / N)
    assert it < 20000
    assert E < 1e-7 * E0
    # a second call continues the iteration count (reference never resets it)
    f2 = oracle.FireMinimizer(s, max_it=it, force_cutoff=1e-14)
    it2, _ = f2.minimize()
    assert it2 == it


@pytest.mark.parametrize("variant", ["simple", "matrix"])
def test_tension_forces_are_minus_energy_gradient(oracle, variant):
    """Line tension between cell types (reference VoronoiQuadraticEnergyWithTension
    force sets): F = -grad(E_AP + sum over cells of gamma times their edges with
    unlike neighbours), by central differences; with one type everywhere the
    tension force sets equal the plain ones."""
    N = 200
    x, y, th, L = oracle.random_init(N, 22)
    s = oracle.VoronoiSystem(x, y, th, L, P0=3.7)
    for _ in range(30):
        s.spp_step(np.zeros(N))
    g = s.get()
    x0, y0 = g["x"].copy(), g["y"].copy()
    types = np.random.default_rng(6).integers(0, 3, N)
    s.set_tension(np.zeros(N, np.int32), gamma=0.7)
    _energy_at(oracle, s, x0, y0)
    F_one_type = s.get()
    plain = oracle.VoronoiSystem(x0, y0, th, L, P0=3.7)
    E_plain = plain.compute_forces(gpu_style=True)
    Fp = plain.get()
    assert np.array_equal(F_one_type["Fx"], Fp["Fx"]) and np.array_equal(F_one_type["Fy"], Fp["Fy"])
    if variant == "simple":
        s.set_tension(types, gamma=0.3)
    else:
        gm = np.array([[0.0, 0.2, 0.5], [0.2, 0.0, 0.35], [0.5, 0.35, 0.0]])
        s.set_tension(types, matrix=gm)
    E0 = _energy_at(oracle, s, x0, y0)
    assert E0 > E_plain  # unlike edges add positive energy
    F = s.get()
    assert np.abs(F["Fx"] - Fp["Fx"]).max() > 1e-3  # the tension changes the forces
    h = 1e-6
    rng = np.random.default_rng(3)
    for i in rng.choice(N, 12, replace=False):
        for comp in ("x", "y"):
            xp, yp = x0.copy(), y0.copy()
            (xp if comp == "x" else yp)[i] += h
            Ep = _energy_at(oracle, s, xp, yp)
            xm, ym = x0.copy(), y0.copy()
            (xm if comp == "x" else ym)[i] -= h
            Em = _energy_at(oracle, s, xm, ym)
            fd = -(Ep - Em) / (2 * h)
            an = F["F" + comp][i]
            assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)), (variant, i, comp, fd, an)
    assert abs(F["Fx"].sum()) < 1e-9 and abs(F["Fy"].sum()) < 1e-9


def test_cell_division_and_death_closed_forms(oracle):
    """Division (reference voronoiModelBase::cellDivision) on the centred
    rectangular lattice, whose Voronoi cell is the hexagon with vertices
    (+-ax/2, +-k), (0, +-h): along theta = 0 the axis leaves the cell at ax/2, so
    parent and daughter sit at x +- frac ax/2; along pi/2 it leaves through the
    vertices (0, +-h). The daughter takes the parent's motility and the given
    director. Death removes the cell. Every result is a valid tiling (sum of
    areas L^2, mean degree 6)."""
    nx = ny = 12
    L = 12.0
    x, y, ax, ay = _lattice(nx, ny, L)
    h = (ay * ay + ax * ax / 4) / (2 * ay)
    N = nx * ny
    s = oracle.VoronoiSystem(x, y, np.zeros(N), L, P0=3.7)
    s.set_motility(0.3, 1.0)
    i, j = 50, 77
    # (fractions chosen so that no four points become cocircular: the lattice
    # itself is full of cocircular quadruples, resolved consistently only by symmetry)
    s.cell_division(i, 0.0, 0.3, 1.2)
    g = s.get()
    assert s.N == N + 1
    assert abs(g["x"][i] - (x[i] + 0.3 * ax / 2)) < 1e-12 and abs(g["y"][i] - y[i]) < 1e-12
    assert abs(g["x"][N] - (x[i] - 0.3 * ax / 2)) < 1e-12 and abs(g["y"][N] - y[i]) < 1e-12
    assert g["theta"][N] == 1.2
    assert abs(g["vx"][N] - 0.3 * np.cos(1.2)) < 1e-15 and abs(g["vy"][N] - 0.3 * np.sin(1.2)) < 1e-15
    s.cell_division(j, np.pi / 2, 0.4, 0.0)
    g = s.get()
    assert abs(g["y"][j] - (y[j] + 0.4 * h)) < 1e-12 and abs(g["y"][N + 1] - (y[j] - 0.4 * h)) < 1e-12
    assert abs(g["x"][j] - x[j]) < 1e-12
    s.compute_forces(gpu_style=True)
    g = s.get()
    nn, _, _ = s.topology()
    assert abs(g["A"].sum() - L * L) < 1e-9 and nn.sum() == 6 * s.N
    before = s.get()
    s.cell_death(10)
    g = s.get()
    assert s.N == N + 1
    keep = np.r_[0:10, 11:N + 2]
    assert np.array_equal(g["x"], before["x"][keep]) and np.array_equal(g["theta"], before["theta"][keep])
    s.compute_forces(gpu_style=True)
    g = s.get()
    nn, _, _ = s.topology()
    assert abs(g["A"].sum() - L * L) < 1e-9 and nn.sum() == 6 * s.N
