"""Pins of the oracle's FIRE update rules and of the reference's circumcircle
search window (CPU only).

FIRE (reference EnergyMinimizerFIRE2D.cpp:156-207, GPU branch): a trajectory on
a quadratic bowl fixed by hand (closed-form velocity-Verlet map of the harmonic
oscillator, and one hand-evaluated mixing step), and the logged decision of
every iteration checked against the stated rules, covering positive power
before and after NMin, negative power with the velocities zeroed, NSince kept
(the reference never resets it), and no dt floor on the GPU branch.

Search window (voronoiModelBase.cu:47-69): hand-built
configurations in which an intruder lies inside a triangle's circumcircle but
outside the bins within wcheck of the triangle's first vertex, so the
reference GPU test does not flag it while the CPU path's window around the
circumcenter (DelaunayLoc.cpp:542-544) does; the same intruder inside the
window is flagged with the triangle; and both across the periodic boundary.
"""
import numpy as np
import pytest


# ---------------------------------------------------------------- FIRE rules
def _vv_map(dt):
    """Velocity Verlet of x'' = -x over one step dt, as a linear map of (x, v)."""
    return np.array([[1 - dt * dt / 2, dt], [-dt * (1 - dt * dt / 4), 1 - dt * dt / 2]])


def test_fire_harmonic_trajectory_fixed_by_hand(oracle):
    """One degree of freedom moving along x in E = x^2 / 2 from rest at x = 1.
    While the power is positive the mixing leaves v (parallel to F) unchanged,
    so the trajectory is the closed-form map above: six iterations at dt = 0.1
    with NSince 0..5, the seventh still at 0.1, after which NSince = 6 > NMin
    raises dt to 0.105 and lowers alpha to 0.09 for the eighth."""
    k = [[1.0, 1.0]]
    for its, x_expect, v_expect in ((6, 0.8251942694005, -0.56414251896502499),
                                    (7, 0.76465404615699495, -0.64363493474289979),
                                    (8, 0.69285722257954996, -0.72015427635156837)):
        it, x, lg, st = oracle.fire_bowl(k, [[1.0, 0.0]], max_it=its, alpha_start=0.1, dt=0.1, dt_max=1.0)
        assert it == its
        assert abs(x[0, 0] - x_expect) < 1e-14 and x[0, 1] == 0.0
        # sum |v|^2 after the last step
        assert abs(lg[-1, 8] - v_expect * v_expect) < 1e-14
    # the decisions of those eight iterations
    assert np.all(lg[:, 0] > 0)
    assert list(lg[:, 5]) == [0, 1, 2, 3, 4, 5, 6, 7]
    assert np.all(lg[:6, 2] == 0.1) and np.all(lg[:6, 4] == 0.1)
    assert lg[6, 2] == 0.1 * 1.05 and lg[6, 4] == 0.1 * 0.9
    assert lg[7, 2] == 0.1 * 1.05 * 1.05 and lg[7, 4] == 0.1 * 0.9 * 0.9


def test_fire_first_mixing_step_by_hand(oracle):
    """E = (x^2 + 4 y^2) / 2 from rest at (1, 1), dt = 0.1, alpha = 0.1: after the
    first velocity-Verlet step x = (0.995, 0.98), v = (-0.09975, -0.396),
    F = (-0.995, -3.92), power 1.65157125; the mixing
    v <- (1 - alpha) v + alpha sqrt(v.v / F.F) F, evaluated by hand, gives
    (-0.099821914926718186, -0.39598181559068879)."""
    it, x, lg, st = oracle.fire_bowl([[1.0, 4.0]], [[1.0, 1.0]], max_it=1, alpha_start=0.1, dt=0.1)
    assert it == 1
    assert np.allclose(x[0], [0.995, 0.98], rtol=0, atol=1e-15)
    assert abs(lg[0, 0] - 1.65157125) < 1e-14
    vv = 0.099821914926718186 ** 2 + 0.39598181559068879 ** 2
    assert abs(lg[0, 8] - vv) < 1e-15
    assert st[0] == pytest.approx((0.995 ** 2 + 3.92 ** 2) / 1.0, rel=1e-15)  # forceMax = sum |F|^2 / N


@pytest.mark.parametrize("dt_dec", [0.95, 0.5])
def test_fire_logged_decisions_follow_the_reference_rules(oracle, dt_dec):
    """Anisotropic bowl (stiffness 1 and 40): the power changes sign many times.
    Every logged iteration obeys the GPU-branch rules."""
    rng = np.random.default_rng(3)
    n = 6
    k = np.column_stack([np.ones(n), np.full(n, 40.0)])
    x0 = rng.uniform(-1, 1, (n, 2))
    dt0, dt_max, inc, dec_a, a0, nmin = 0.1, 0.5, 1.1, 0.9, 0.2, 3
    it, x, lg, st = oracle.fire_bowl(k, x0, max_it=150, force_cutoff=1e-300, alpha_start=a0, dt=dt0,
                                     dt_max=dt_max, dt_inc=inc, dt_dec=dt_dec, alpha_dec=dec_a, n_min=nmin)
    assert it == 150
    pos, neg = lg[:, 0] > 0, lg[:, 0] <= 0
    before, after = pos & (lg[:, 5] <= nmin), pos & (lg[:, 5] > nmin)
    assert before.sum() > 0 and after.sum() > 0 and neg.sum() > 0
    # positive power, NSince not above NMin: dt and alpha unchanged
    assert np.all(lg[before, 2] == lg[before, 1]) and np.all(lg[before, 4] == lg[before, 3])
    # positive power after NMin: dt grows to at most dtMax, alpha decays
    assert np.all(lg[after, 2] == np.minimum(lg[after, 1] * inc, dt_max))
    assert np.all(lg[after, 4] == lg[after, 3] * dec_a)
    assert np.all(lg[pos, 6] == lg[pos, 5] + 1) and np.all(lg[pos, 7] == 0)
    # non-positive power: dt shrinks, alpha restarts, velocities zeroed, NSince kept
    assert np.all(lg[neg, 2] == lg[neg, 1] * dt_dec) and np.all(lg[neg, 4] == a0)
    assert np.all(lg[neg, 7] == 1) and np.all(lg[neg, 8] == 0.0)
    assert np.all(lg[neg, 6] == lg[neg, 5])
    # consecutive iterations chain (the log is the state the next one starts from)
    assert np.all(lg[1:, 1] == lg[:-1, 2]) and np.all(lg[1:, 3] == lg[:-1, 4]) and np.all(lg[1:, 5] == lg[:-1, 6])


def test_fire_gpu_branch_has_no_dt_floor(oracle):
    """E = x^2 / 2 from rest at x = 1 with NMin so large that dt never grows:
    every zero crossing turns the power negative, zeroes the velocity and
    halves dt (dtDec = 0.5). After seven crossings dt is below the CPU branch's
    floor dt0 / 100 (EnergyMinimizerFIRE2D.cpp:255), which the GPU branch does
    not apply (:202)."""
    it, x, lg, st = oracle.fire_bowl([[1.0, 1.0]], [[1.0, 0.0]], max_it=6000, force_cutoff=1e-300,
                                     alpha_start=0.1, dt=0.1, dt_max=1.0, dt_dec=0.5, n_min=10 ** 9)
    neg = np.flatnonzero(lg[:, 0] <= 0)
    assert len(neg) >= 8
    assert np.all(lg[neg, 2] == lg[neg, 1] * 0.5)
    assert lg[neg[7], 2] == 0.1 * 0.5 ** 8 < 0.01 * 0.1
    # between crossings the step is constant
    assert np.all(lg[lg[:, 0] > 0, 2] == lg[lg[:, 0] > 0, 1])


def test_fire_converges_on_the_bowl(oracle):
    """The loop test is the RMS force: sqrt(sum |F|^2 / N) below the cutoff."""
    k = np.array([[1.0, 3.0], [2.0, 5.0], [0.5, 1.0]])
    it, x, lg, st = oracle.fire_bowl(k, [[1, -1], [0.5, 2], [-1, 0.3]], max_it=5000, force_cutoff=1e-10,
                                     alpha_start=0.1, dt=0.05, dt_max=0.5)
    assert it < 5000
    F = -k * x
    assert np.sqrt((F * F).sum() / 3) <= 1e-10
    assert np.sqrt(st[0]) <= 1e-10 and np.sqrt((F * F).sum() / 3) == pytest.approx(np.sqrt(st[0]), rel=1e-12)


# ------------------------------------------------------ N3 search window
L_FIX = 12.0  # 10 x 10 bins of side 1.2


def n3_fixture(intruder, shift=0.0):
    """Triangle (0, 1, 2) inscribed in the circle of radius 1.1 about
    Q = (3.49 + shift, 6.6): cell 0 (the smallest index, the reference's first
    vertex) at the left end of the circle, in bin column 1 near its right edge
    (wcheck = ceil(1.1 / 1.2) = 1, so the window is bin columns 0..2). Cell 3 is
    the candidate intruder at Q + (intruder, 0); the other cells are a jittered
    unit-density grid with a hole of radius 1.6 about Q. x shifted by `shift`
    (periodic). Returns x, y and the rings: cell 0 = (1, 2), every other cell
    (0, 0), so that (0, 1, 2) is the only triangle the reference tests."""
    q = np.array([3.49, 6.6])
    r = 1.1
    pts = [q + r * np.array([np.cos(a), np.sin(a)]) for a in (np.pi, np.pi / 3, -np.pi / 3)]
    pts.append(q + np.array([intruder, 0.0]))
    rng = np.random.default_rng(11)
    for gx in range(12):
        for gy in range(12):
            p = np.array([gx + 0.5, gy + 0.5]) + rng.uniform(-0.2, 0.2, 2)
            if np.hypot(*(p - q)) > 1.6:
                pts.append(p)
    xy = np.array(pts)
    xy[:, 0] = (xy[:, 0] + shift) % L_FIX
    N = len(xy)
    nn = np.full(N, 2, np.int32)
    nbr = np.zeros((N, 4), np.int32)
    nbr[0, :2] = (1, 2)
    return xy[:, 0].copy(), xy[:, 1].copy(), nn, nbr


def _flags(oracle, x, y, nn, nbr):
    s = oracle.VoronoiSystem(x, y, np.zeros(len(x)), L_FIX)
    s.set_topology(nn, nbr, 8)
    return s.test_flags()[1], s.test_flags_cpu()[1]


@pytest.mark.parametrize("shift", [0.0, 9.4])
def test_n3_window_misses_an_intruder_beyond_the_first_vertex_bins(oracle, shift):
    # intruder at Q + (0.81, 0): inside the circle (0.81 < 1.1), bin column 3
    # (shift 9.4: cell 0 in column 9, window 8, 9, 0 across the periodic
    # boundary, intruder in column 1)
    x, y, nn, nbr = n3_fixture(0.81, shift)
    dx = x[3] - (3.49 + shift)
    dx -= L_FIX * np.round(dx / L_FIX)
    assert np.hypot(dx, y[3] - 6.6) < 1.1
    gpu, cpu = _flags(oracle, x, y, nn, nbr)
    assert gpu.sum() == 0
    # the CPU path's window around the circumcenter finds it: cell 0's ring
    # (1, 2) has a point inside the circle of its triangle
    assert cpu[0] == 1


@pytest.mark.parametrize("shift", [0.0, 9.4])
def test_n3_window_flags_an_intruder_in_the_first_vertex_bins(oracle, shift):
    # intruder at Q + (-0.2, 0): inside the circle, bin column 2 (shift 9.4:
    # column 0, inside the wrapped window)
    x, y, nn, nbr = n3_fixture(-0.2, shift)
    gpu, cpu = _flags(oracle, x, y, nn, nbr)
    assert sorted(np.flatnonzero(gpu)) == [0, 1, 2, 3]
    assert cpu[0] == 1


def test_n3_window_no_flag_when_the_circle_is_empty(oracle):
    x, y, nn, nbr = n3_fixture(1.3)  # outside the circle
    gpu, cpu = _flags(oracle, x, y, nn, nbr)
    assert gpu.sum() == 0 and cpu[0] == 0
