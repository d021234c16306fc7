"""Storage growth (reference neighMax, voronoiModelBase.cpp:309-312, and
vertexMax / growCellVerticesList, vertexModelBase.cpp:475-500, 815-824).

* Ring storage: a Delaunay degree above the starting capacity (24) at creation
  grows the storage, and the rings equal the oracle's; a degree above the
  largest capacity (64, one neighbour per lane of a star wavefront) is refused.
* A system whose repair needs a longer ring in the middle of a run stalls, is
  retriangulated with grown storage and catches up the steps it owes alone: its
  state equals a run that had the storage from the start, bit for bit, and the
  other system of the batch equals its run alone. The same through
  enforceTopology and through FIRE.
* Cell-list storage of the vertex model: a T1 flip that lengthens a cell beyond
  its 32 slots grows the storage and the run equals the oracle's step by step.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cg():
    import cellgpu_amd as cg
    cg.require_device()
    return cg


def ring_config(N, m, r0=1.5, rout=2.6, seed=9, centers=1, split=0.5):
    """A cell (or two, split apart along x) inside a ring of m cells of radius
    r0, the others uniform beyond rout: the centre's Delaunay degree is m."""
    L = float(np.sqrt(N))
    c = np.array([L / 2, L / 2])
    k = np.arange(m)
    r = r0 + 0.01 * np.sin(1.3 * k)
    ring = c + np.stack([r * np.cos(2 * np.pi * k / m + 0.01), r * np.sin(2 * np.pi * k / m + 0.01)], 1)
    rng = np.random.default_rng(seed)
    rest = []
    while len(rest) < N - m - centers:
        p = rng.uniform(0, L, 2)
        if np.hypot(*(p - c)) > rout:
            rest.append(p)
    cs = [c] if centers == 1 else [c + [split, 0.013], c - [split, -0.007]]
    return np.vstack([np.array(cs), ring, np.array(rest)]), L


def _rings(nn, nbr, lo=0, hi=None):
    hi = len(nn) if hi is None else hi
    return [tuple(int(v) - lo for v in nbr[i, : nn[i]]) for i in range(lo, hi)]


def test_ring_storage_grows_at_creation(oracle):
    cg = _cg()
    N = 200
    pts, L = ring_config(N, 30)
    b = cg.VoronoiBatch([N], x=pts[:, 0], y=pts[:, 1], theta=np.zeros(N), precision="fp64")
    assert b.ring_capacity() == 48
    nn, nbr, _ = b.topology()
    onn, onbr, _ = oracle.triangulate(pts[:, 0], pts[:, 1], L, cap=48)
    assert onn[0] == 30 and _rings(nn, nbr) == _rings(onn, onbr)
    o = oracle.VoronoiSystem(pts[:, 0], pts[:, 1], np.zeros(N), L, P0=3.8)
    o.compute_forces()
    b.set_prefs(1.0, 3.8)
    b.compute_forces()
    g, r = b.get(), o.get()
    scale = np.abs(np.concatenate([r["Fx"], r["Fy"]])).max()
    assert np.abs(g["Fx"] - r["Fx"]).max() <= 1e-12 * scale
    assert np.abs(g["Fy"] - r["Fy"]).max() <= 1e-12 * scale


def test_division_of_a_cell_beyond_the_starting_capacity(oracle):
    """The division axis of a degree-30 cell (above the starting 24 slots) is
    found from its 30-slot ring: daughters and rings equal the oracle's."""
    cg = _cg()
    N = 200
    pts, L = ring_config(N, 30)
    th = np.random.default_rng(2).uniform(0, 2 * np.pi, N)
    b = cg.VoronoiBatch([N], x=pts[:, 0], y=pts[:, 1], theta=th, precision="fp64")
    assert b.ring_capacity() == 48
    o = oracle.VoronoiSystem(pts[:, 0], pts[:, 1], th, L, P0=3.8)
    print("division of a degree-30 cell", flush=True)
    b.cell_events([1], [0], [0.7], [0.4], [2.0])
    o.cell_division(0, 0.7, 0.4, 2.0)
    b.check_status()
    g, r = b.get(), o.get()
    assert b.N == o.N == N + 1
    for k in ("x", "y", "theta", "vx", "vy"):
        assert np.array_equal(g[k], r[k]), k
    nn, nbr, _ = b.topology()
    onn, onbr, _ = o.topology(cap=b.ring_capacity())
    assert _rings(nn, nbr) == _rings(onn, onbr)


def test_degree_above_the_largest_ring_capacity_is_refused():
    cg = _cg()
    N = 400
    pts, L = ring_config(N, 80, r0=3.0, rout=4.0)  # centre degree 80 > 64
    with pytest.raises(cg.CellGPUError, match="ring capacity"):
        cg.VoronoiBatch([N], x=pts[:, 0], y=pts[:, 1], theta=np.zeros(N), precision="fp64")


def _stall_pair(cg, precision):
    """System 0 random; system 1 two centres inside a ring of 30 (degrees 17),
    then one centre moved far away without retriangulating: its repair needs a
    degree-30 star. The moduli are zero (the rings kept across the jump would
    give meaningless forces), so the cells move by self-propulsion alone."""
    N0, N1 = 300, 200
    a = np.random.default_rng(4).uniform(0, np.sqrt(N0), (N0, 2))
    x0, L1 = ring_config(N1, 30, centers=2)
    x1 = x0.copy()
    x1[1] = [1.3, 1.1]
    th = np.random.default_rng(5).uniform(0, 2 * np.pi, N0 + N1)
    b = cg.VoronoiBatch([N0, N1], x=np.r_[a[:, 0], x0[:, 0]], y=np.r_[a[:, 1], x0[:, 1]], theta=th,
                        precision=precision)
    assert b.ring_capacity() == 24
    b.set_positions(np.r_[a[:, 0], x1[:, 0]], np.r_[a[:, 1], x1[:, 1]], False)
    # the same system 1 with storage grown beforehand (created from a degree-30 set)
    big, _ = ring_config(N1, 30)
    r = cg.VoronoiBatch([N1], x=big[:, 0], y=big[:, 1], theta=th[N0:], precision=precision)
    assert r.ring_capacity() == 48
    r.set_positions(x0[:, 0], x0[:, 1], True)
    r.set_positions(x1[:, 0], x1[:, 1], False)
    solo = cg.VoronoiBatch([N0], x=a[:, 0], y=a[:, 1], theta=th[:N0], precision=precision)
    for q in (b, r, solo):
        q.set_motility(0.05, 1.0)
        q.set_prefs(1.0, 3.8)
        q.set_params(KA=0.0, KP=0.0)
    return b, r, solo, N0


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_stalled_system_grows_and_catches_up(precision):
    cg = _cg()
    b, r, solo, N0 = _stall_pair(cg, precision)
    print("ring growth mid-run, %s" % precision, flush=True)
    b.spv_steps(6)
    r.spv_steps(6)
    solo.spv_steps(6)
    for q in (b, r, solo):
        q.check_status()
    assert b.ring_capacity() == 48
    assert b.counters(0)["step"] == b.counters(1)["step"] == 6
    g, gr, gs = b.get(), r.get(), solo.get()
    for k in ("x", "y", "theta", "vx", "vy", "Fx", "Fy", "A", "P"):
        assert np.array_equal(g[k][N0:], gr[k]), k
        assert np.array_equal(g[k][:N0], gs[k]), k
    nn, nbr, _ = b.topology()
    rn, rb, _ = r.topology()
    assert _rings(nn, nbr, N0) == _rings(rn, rb)
    assert nn[N0:].max() >= 25


def test_enforce_topology_grows_the_ring_storage(oracle):
    cg = _cg()
    b, r, solo, N0 = _stall_pair(cg, "fp64")
    b.enforce_topology()
    b.check_status()
    assert b.ring_capacity() == 48
    g = b.get()
    onn, onbr, _ = oracle.triangulate(g["x"][N0:], g["y"][N0:], b.L[1], cap=48)
    nn, nbr, _ = b.topology()
    assert _rings(nn, nbr, N0) == _rings(onn, onbr)
    assert b.counters(1)["step"] == 0


def test_fire_stall_grows_and_completes_the_iteration():
    """A FIRE iteration whose repair stalls is finished after the growth (its
    force half and reduction), equal to the run with the storage grown."""
    cg = _cg()
    runs = []
    b, r, solo, N0 = _stall_pair(cg, "fp64")
    for q in (b, r):
        f = cg.EnergyMinimizerFIRE(q)
        f.setMaximumIterations(40)
        f.minimize()
        q.check_status()
        runs.append((q, f))
    assert b.ring_capacity() == 48
    g, gr = b.get(), r.get()
    for k in ("x", "y", "Fx", "Fy"):
        assert np.array_equal(g[k][N0:], gr[k]), k
    # zero moduli: the stalled iteration, once completed, finds zero force
    assert runs[0][1].state(1)["iterations"] == runs[1][1].state(0)["iterations"] == 1
    assert runs[0][1].state(1)["force_max"] == 0.0


def test_vertex_cell_list_grows_for_a_flip(oracle):
    """A cell of 32 vertices (the starting capacity) gains one through a T1 flip
    of a shortened spoke: the storage grows and the steps equal the oracle's."""
    cg = _cg()
    Nc = 300
    pts, L = ring_config(Nc, 32)
    b = cg.VertexBatch([Nc], x=pts[:, 0], y=pts[:, 1], precision="fp64")
    nn, nbr, _ = oracle.triangulate(pts[:, 0], pts[:, 1], L, cap=48)
    o = oracle.VertexSystem(pts[:, 0], pts[:, 1], L, nn, nbr)
    o.CAP = 64
    vn, vcn, cvn, cv = b.topology()
    assert cvn[0] == 32 and b.vertex_capacity() == 32
    # shorten the spoke leaving the big cell at one of its vertices
    v1 = int(cv[0, 0])
    f = [q for q in range(3) if vcn[3 * v1 + q] == 0][0]
    v2 = int(vn[3 * v1 + f])
    g = b.get()
    x, y = g["vx"].copy(), g["vy"].copy()
    d = np.array([x[v2] - x[v1], y[v2] - y[v1]])
    d -= L * np.round(d / L)
    x[v2], y[v2] = np.array([x[v1], y[v1]]) + 0.025 * d / np.hypot(*d)
    x[v2] %= L
    y[v2] %= L
    for q in (b, o):
        q.set_positions(x, y)
        q.set_prefs(1.0, 3.9)
    # a short step keeps the spoke short until the flip and the flipped edge long
    b.set_params(dt=1e-4, T=0.0, t1_threshold=0.04)
    o.set_params(dt=1e-4, T=0.0, thr=0.04)
    z = np.zeros(b.Nv)
    print("vertex cell-list growth: spoke %d-%d" % (v1, v2), flush=True)
    for step in range(4):
        b.brownian_steps(1, use_graph=False)
        b.check_status()
        o.brownian_step(z, z)
        g, r = b.get(), o.compute()
        np.testing.assert_allclose(g["vx"], r["vx"], rtol=0, atol=1e-11)
        np.testing.assert_allclose(g["vy"], r["vy"], rtol=0, atol=1e-11)
        vn, vcn, cvn, cv = b.topology()
        ovn, ovcn, ocvn, ocv = o.topology()
        assert sorted(cvn) == sorted(ocvn), step
        assert [frozenset(vn[3 * v: 3 * v + 3]) for v in range(b.Nv)] == \
            [frozenset(ovn[3 * v: 3 * v + 3]) for v in range(o.Nv)], step
    assert cvn[0] == 33 and b.vertex_capacity() == 40
    assert b.counters(0)["step"] == 4
