"""CPU pins of the vertex-model oracle.

* centred-rectangular lattice: every cell a hexagon of known area/perimeter, F = 0;
* forces = -grad E by central differences on vertex positions;
* Euler / bookkeeping invariants (Nv = 2 Nc, sum of cell vertex counts = 3 Nv);
* single-T1 known answer: a shortened edge flips, the two cells on its sides
  lose a vertex (6 -> 5), the two at its ends gain one (6 -> 7), the new edge is
  perpendicular to the old one and twice as long.
"""
import numpy as np


def _lattice(nx, ny, L):
    ax, ay = L / nx, L / ny
    pts = [((i + 0.5 * (j % 2) + 0.25) * ax, (j + 0.5) * ay) for j in range(ny) for i in range(nx)]
    p = np.array(pts)
    return p[:, 0].copy(), p[:, 1].copy(), ax, ay


def _lattice_vm(oracle, nx=8, ny=8, L=8.0, **kw):
    x, y, ax, ay = _lattice(nx, ny, L)
    nn, nbr, _ = oracle.triangulate(x, y, L, cap=32)
    return oracle.VertexSystem(x, y, L, nn, nbr, **kw), ax, ay


def test_lattice_closed_form(oracle):
    vm, ax, ay = _lattice_vm(oracle)
    k = (ay * ay - ax * ax / 4) / (2 * ay)
    h = (ay * ay + ax * ax / 4) / (2 * ay)
    A_cf, P_cf = ax * ay, 4 * k + 4 * np.hypot(ax / 2, h - k)
    vm.set_prefs(A_cf, P_cf)
    r = vm.compute()
    assert vm.Nv == 2 * vm.Nc
    np.testing.assert_allclose(r["A"], A_cf, rtol=1e-12)
    np.testing.assert_allclose(r["P"], P_cf, rtol=1e-12)
    assert abs(r["E"]) < 1e-20
    assert np.abs(r["Fx"]).max() < 1e-9 and np.abs(r["Fy"]).max() < 1e-9  # at the minimum
    vm.set_prefs(1.0, 3.0)
    r = vm.compute()
    # identical cells under tension: forces balance over the whole tissue
    assert abs(r["Fx"].sum()) < 1e-10 and abs(r["Fy"].sum()) < 1e-10
    vn, vcn, cvn, cv = vm.topology()
    assert (cvn == 6).all() and cvn.sum() == 3 * vm.Nv


def test_forces_are_minus_energy_gradient(oracle):
    vm, (x, y, L, nn, nbr) = oracle.vertex_from_random(120, 4, P0=3.7)
    r0 = vm.compute()
    vx, vy = r0["vx"].copy(), r0["vy"].copy()
    h = 1e-6
    rng = np.random.default_rng(1)
    for v in rng.choice(vm.Nv, 10, replace=False):
        for comp in ("x", "y"):
            a, b = vx.copy(), vy.copy()
            (a if comp == "x" else b)[v] += h
            vm.set_positions(a, b)
            Ep = vm.compute()["E"]
            a, b = vx.copy(), vy.copy()
            (a if comp == "x" else b)[v] -= h
            vm.set_positions(a, b)
            Em = vm.compute()["E"]
            fd = -(Ep - Em) / (2 * h)
            an = r0["F" + comp][v]
            assert abs(fd - an) <= 1e-5 * max(1.0, abs(an)), (v, comp, fd, an)
    vm.set_positions(vx, vy)
    r = vm.compute()
    assert abs(r["Fx"].sum()) < 1e-9 and abs(r["Fy"].sum()) < 1e-9
    assert abs(r["A"].sum() - L * L) < 1e-9 * L * L


def test_single_t1_known_answer(oracle):
    vm, ax, ay = _lattice_vm(oracle, thr=0.04)
    vn, vcn, cvn, cv = vm.topology()
    r = vm.compute()
    vx, vy = r["vx"].copy(), r["vy"].copy()
    L = vm.L
    # pick the vertex pair of the first vertical edge (x equal) and shrink it
    v1 = 0
    v2 = None
    for f in range(3):
        c = vn[3 * v1 + f]
        d = vx[c] - vx[v1]
        d -= L * np.round(d / L)
        if abs(d) < 1e-9:
            v2 = c
    assert v2 is not None
    mid = np.array([vx[v1], (vy[v1] + vy[v2]) / 2])
    dy = vy[v2] - vy[v1]
    dy -= L * np.round(dy / L)
    mid[1] = vy[v1] + dy / 2
    vx[v1], vy[v1] = mid[0], mid[1] - 0.01
    vx[v2], vy[v2] = mid[0], mid[1] + 0.01
    vx, vy = vx % L, vy % L
    vm.set_positions(vx, vy)
    cells_before = set(vcn[3 * v1: 3 * v1 + 3]) | set(vcn[3 * v2: 3 * v2 + 3])
    flips = vm.t1()
    assert flips == 1 and vm.last_flagged() == 1
    vn2, vcn2, cvn2, cv2 = vm.topology()
    assert cvn2.sum() == cvn.sum()
    changed = {c: (cvn[c], cvn2[c]) for c in cells_before}
    assert sorted(changed.values()) == [(6, 5), (6, 5), (6, 7), (6, 7)]
    r2 = vm.compute()
    e_old = np.array([0.0, 0.02])
    e_new = np.array([r2["vx"][v1] - r2["vx"][v2], r2["vy"][v1] - r2["vy"][v2]])
    e_new -= L * np.round(e_new / L)
    assert abs(np.dot(e_new, e_old)) < 1e-12
    assert abs(np.linalg.norm(e_new) - 2 * np.linalg.norm(e_old)) < 1e-12
    # vertex bookkeeping stays consistent: every vertex is in the lists of its 3 cells
    for v in range(vm.Nv):
        for c in vcn2[3 * v: 3 * v + 3]:
            assert v in cv2[c, : cvn2[c]]


def test_brownian_zero_temperature_relaxes(oracle):
    vm, _ = oracle.vertex_from_random(200, 6, P0=3.9, dt=0.005, T=0.0)
    E0 = vm.compute()["E"]
    z = np.zeros(vm.Nv)
    for _ in range(50):
        vm.brownian_step(z, z)
    E1 = vm.compute()["E"]
    assert E1 < E0
    vn, vcn, cvn, cv = vm.topology()
    assert cvn.sum() == 3 * vm.Nv


def _wrapped(d, L):
    return d - L * np.round(d / L)


def test_spp_cell_vertex_closed_forms(oracle):
    """Self-propelled cell-vertex step at the force-free lattice. Uniform
    directors: every vertex moves by dt v0 e(theta). Per-cell directors: each
    cell's self-propulsion is shared by its six vertices, a third each, so the
    summed displacement is 2 dt sum_c v0_c e(theta_c). Directors rotate by
    eta sqrt(2 dt Dr)."""
    vm, ax, ay = _lattice_vm(oracle, dt=0.01)
    k = (ay * ay - ax * ax / 4) / (2 * ay)
    h = (ay * ay + ax * ax / 4) / (2 * ay)
    vm.set_prefs(ax * ay, 4 * k + 4 * np.hypot(ax / 2, h - k))
    L, dt = vm.L, 0.01
    r0 = vm.compute()
    vm.set_motility(0.3, 0.5)
    vm.set_directors(np.full(vm.Nc, 0.7))
    eta = np.random.default_rng(3).standard_normal(vm.Nc)
    assert vm.spp_step(eta) == 0
    r1 = vm.compute()
    dx, dy = _wrapped(r1["vx"] - r0["vx"], L), _wrapped(r1["vy"] - r0["vy"], L)
    np.testing.assert_allclose(dx, dt * 0.3 * np.cos(0.7), atol=1e-12)
    np.testing.assert_allclose(dy, dt * 0.3 * np.sin(0.7), atol=1e-12)
    np.testing.assert_allclose(vm.directors(), 0.7 + eta * np.sqrt(2 * dt * 0.5), rtol=0, atol=1e-15)

    vm2, _, _ = _lattice_vm(oracle, dt=dt)
    vm2.set_prefs(ax * ay, 4 * k + 4 * np.hypot(ax / 2, h - k))
    rng = np.random.default_rng(4)
    v0, th = rng.uniform(0, 0.5, vm2.Nc), rng.uniform(0, 2 * np.pi, vm2.Nc)
    vm2.set_motility(v0, 1.0)
    vm2.set_directors(th)
    r0 = vm2.compute()
    vm2.spp_step(np.zeros(vm2.Nc))
    r1 = vm2.compute()
    dx, dy = _wrapped(r1["vx"] - r0["vx"], L), _wrapped(r1["vy"] - r0["vy"], L)
    assert abs(dx.sum() - 2 * dt * np.sum(v0 * np.cos(th))) < 1e-12
    assert abs(dy.sum() - 2 * dt * np.sum(v0 * np.sin(th))) < 1e-12
    np.testing.assert_array_equal(vm2.directors(), th)  # zero noise: directors unchanged


def test_fire_relaxes_vertex_lattice_to_its_minimum(oracle):
    """FIRE on the vertices from a perturbed force-free lattice with the lattice's
    own (A0, P0): the energy minimum is E = 0 and FIRE reaches it."""
    vm, ax, ay = _lattice_vm(oracle)
    k = (ay * ay - ax * ax / 4) / (2 * ay)
    h = (ay * ay + ax * ax / 4) / (2 * ay)
    vm.set_prefs(ax * ay, 4 * k + 4 * np.hypot(ax / 2, h - k))
    r0 = vm.compute()
    rng = np.random.default_rng(5)
    vm.set_positions(r0["vx"] + 0.02 * rng.standard_normal(vm.Nv), r0["vy"] + 0.02 * rng.standard_normal(vm.Nv))
    E0 = vm.compute()["E"]
    assert E0 > 1e-4
    it, st = vm.fire(max_it=20000, force_cutoff=1e-10, alpha_start=0.1, dt=0.01)
    assert it < 20000 and st["force_max"] ** 0.5 <= 1e-10
    r = vm.compute()
    assert r["E"] < 1e-12 * E0
    np.testing.assert_allclose(r["A"], ax * ay, rtol=1e-7)


def test_cell_division_and_t2_death(oracle):
    """Vertex-model division (reference vertexModelBase::cellDivision) of a lattice
    hexagon through the midpoints of two opposite edges halves it exactly (the
    chord passes through its centre); the cells across those edges gain a vertex;
    bookkeeping stays consistent. Dividing off a triangle (two consecutive edges)
    and killing it by a T2 transition (vertexModelBase::cellDeath) restores the
    original tissue exactly."""
    vm, ax, ay = _lattice_vm(oracle)
    r0 = vm.compute()
    vn0, vcn0, cvn0, cv0 = vm.topology()
    Nc, Nv, L = vm.Nc, vm.Nv, vm.L
    c = 20
    vm.cell_division(c, 0, 3, 0.5)
    assert vm.Nc == Nc + 1 and vm.Nv == Nv + 2
    r = vm.compute()
    vn, vcn, cvn, cv = vm.topology()
    assert cvn[c] == 5 and cvn[Nc] == 5 and cvn.sum() == 3 * vm.Nv
    assert sorted(cvn[:Nc][cvn[:Nc] != 6]) == [5, 7, 7]
    np.testing.assert_allclose([r["A"][c], r["A"][Nc]], r0["A"][c] / 2, rtol=1e-12)
    assert abs(r["A"].sum() - L * L) < 1e-9
    for v in range(vm.Nv):
        for q in vcn[3 * v: 3 * v + 3]:
            assert v in cv[q, : cvn[q]]
    # a triangle divided off and removed by a T2 transition: back to the start
    vm2, _, _ = _lattice_vm(oracle)
    vm2.cell_division(c, 0, 1, 0.0)
    _, _, cvn2, _ = vm2.topology()
    assert cvn2[Nc] == 3
    vm2.cell_death(Nc)
    assert vm2.Nc == Nc and vm2.Nv == Nv
    r2 = vm2.compute()
    vn2, vcn2, cvn2, cv2 = vm2.topology()
    assert np.array_equal(r2["vx"], r0["vx"]) and np.array_equal(r2["vy"], r0["vy"])
    canon = lambda cv_, n_: [tuple(np.roll(cv_[q, :n_[q]], -int(np.argmin(cv_[q, :n_[q]])))) for q in range(len(n_))]
    assert canon(cv2, cvn2) == canon(cv0, cvn0)
    assert [set(vn2[3 * v:3 * v + 3]) for v in range(Nv)] == [set(vn0[3 * v:3 * v + 3]) for v in range(Nv)]
    np.testing.assert_allclose(r2["A"], r0["A"], rtol=1e-13)
