"""GPU parity of the vertex-model step (geometry, forces, Brownian move, T1
flips) against the FP64 oracle. Topology is compared as sets (vertex
neighbours, vertex cells) and as cyclic counter-clockwise cell vertex lists."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cg():
    import cellgpu_amd as cg
    cg.require_device()
    return cg


def _canon(cv, cvn):
    out = []
    for c in range(len(cvn)):
        l = list(cv[c, : cvn[c]])
        k = int(np.argmin(l))
        out.append(tuple(l[k:] + l[:k]))
    return out


def _sets(a):
    return [frozenset(a[3 * v: 3 * v + 3]) for v in range(len(a) // 3)]


def _oracle_like(oracle, Nc, seed, **kw):
    o, (x, y, L, nn, nbr) = oracle.vertex_from_random(Nc, seed, **kw)
    return o


@pytest.mark.parametrize("Nc", [64, 500])
def test_vertex_init_and_forces_match_oracle(oracle, Nc):
    cg = _cg()
    b = cg.VertexBatch([Nc], seeds=[5], precision="fp64")
    o = _oracle_like(oracle, Nc, 5)
    assert b.Nv == o.Nv == 2 * Nc
    vn, vcn, cvn, cv = b.topology()
    ovn, ovcn, ocvn, ocv = o.topology()
    assert _sets(vn) == _sets(ovn) and _sets(vcn) == _sets(ovcn)
    assert _canon(cv, cvn) == _canon(ocv, ocvn)
    b.set_prefs(1.0, 3.9)
    b.compute_forces()
    g = b.get()
    r = o.compute()
    assert np.array_equal(g["vx"], r["vx"]) and np.array_equal(g["vy"], r["vy"])
    np.testing.assert_allclose(g["A"], r["A"], rtol=1e-12)
    np.testing.assert_allclose(g["P"], r["P"], rtol=1e-12)
    scale = np.abs(np.concatenate([r["Fx"], r["Fy"]])).max()
    assert np.abs(g["Fx"] - r["Fx"]).max() <= 1e-12 * scale
    assert np.abs(g["Fy"] - r["Fy"]).max() <= 1e-12 * scale


def test_vertex_brownian_t1_stepwise(oracle):
    """Each step: oracle restarted from the GPU state, same noise; T1 flips included."""
    cg = _cg()
    Nc = 300
    b = cg.VertexBatch([Nc], seeds=[9], precision="fp64")
    b.set_params(dt=0.005, T=1e-3, t1_threshold=0.04)
    b.set_prefs(1.0, 3.9)
    b.set_noise_buffer(True)
    o = _oracle_like(oracle, Nc, 9, dt=0.005, T=1e-3, thr=0.04, P0=3.9)
    rng = np.random.default_rng(2)
    t1_total = 0
    for step in range(25):
        g = b.get()
        vn, vcn, cvn, cv = b.topology()
        o.set_positions(g["vx"], g["vy"])
        o.set_topology(vn, vcn, cvn, cv)
        ex, ey = rng.standard_normal(b.Nv), rng.standard_normal(b.Nv)
        b.upload_noise(ex, ey)
        b.brownian_steps(1, use_graph=False)
        b.check_status()
        t1_total += o.brownian_step(ex, ey)
        g, r = b.get(), o.compute()
        np.testing.assert_allclose(g["vx"], r["vx"], rtol=0, atol=1e-11)
        np.testing.assert_allclose(g["vy"], r["vy"], rtol=0, atol=1e-11)
        vn, vcn, cvn, cv = b.topology()
        ovn, ovcn, ocvn, ocv = o.topology()
        assert _sets(vn) == _sets(ovn), step
        assert _sets(vcn) == _sets(ovcn), step
        assert _canon(cv, cvn) == _canon(ocv, ocvn), step
    assert t1_total > 0
    assert b.counters()["t1_events"] == t1_total


def test_vertex_full_size_invariants():
    """Nc = 32 768 (Nv = 65 536, L = 181.02), Brownian T = 1e-3, dt = 0.005, T1
    threshold 0.04, P0 = 3.9: after 200 steps with T1 flips the vertex model keeps
    its bookkeeping (Nv = 2 Nc, sum of cell vertex counts = 3 Nv, every vertex in
    the lists of its three cells, positive signed cell areas), tiles the box
    (sum of areas = L^2) and its forces sum to zero."""
    cg = _cg()
    Nc = 32768
    b = cg.VertexBatch([Nc], seeds=[21], precision="fp64")
    b.set_params(dt=0.005, T=1e-3, t1_threshold=0.04)
    b.set_prefs(1.0, 3.9)
    b.brownian_steps(200)
    b.check_status()
    assert b.counters(0)["t1_events"] > 0
    b.compute_forces()
    g = b.get()
    vn, vcn, cvn, cv = b.topology()
    L = float(np.sqrt(Nc))
    assert b.Nv == 2 * Nc and cvn.sum() == 3 * b.Nv
    for v in range(0, b.Nv, 97):
        for c in vcn[3 * v: 3 * v + 3]:
            assert v in cv[c, : cvn[c]]
    assert abs(g["A"].sum() - L * L) < 1e-9 * L * L
    assert (g["A"] > 0).all()
    F = np.abs(np.concatenate([g["Fx"], g["Fy"]])).max()
    assert abs(g["Fx"].sum()) < 1e-9 * F * b.Nv and abs(g["Fy"].sum()) < 1e-9 * F * b.Nv


def test_vertex_spp_cell_vertex_stepwise(oracle):
    """Self-propelled cell-vertex dynamics (per-cell v0, Dr = 1, T1 threshold 0.1):
    each step the oracle restarts from the GPU state (positions, topology,
    directors) with the same director noise; positions within 1e-11, directors
    within 1e-13, topology identical."""
    cg = _cg()
    Nc = 300
    b = cg.VertexBatch([Nc], seeds=[9], precision="fp64")
    b.set_params(dt=0.01, t1_threshold=0.1)
    b.set_prefs(1.0, 3.9)
    rng = np.random.default_rng(3)
    v0 = rng.uniform(0.1, 0.5, Nc)
    b.set_motility(v0, 1.0)
    b.set_noise_buffer(True)
    o = _oracle_like(oracle, Nc, 9, dt=0.01, thr=0.1, P0=3.9)
    o.set_motility(v0, 1.0)
    t1_total = 0
    for step in range(25):
        g = b.get()
        vn, vcn, cvn, cv = b.topology()
        o.set_positions(g["vx"], g["vy"])
        o.set_topology(vn, vcn, cvn, cv)
        o.set_directors(b.directors())
        eta = rng.standard_normal(Nc)
        b.upload_cell_noise(eta)
        b.spp_steps(1, use_graph=False)
        b.check_status()
        t1_total += o.spp_step(eta)
        g, r = b.get(), o.compute()
        np.testing.assert_allclose(g["vx"], r["vx"], rtol=0, atol=1e-11)
        np.testing.assert_allclose(g["vy"], r["vy"], rtol=0, atol=1e-11)
        np.testing.assert_allclose(b.directors(), o.directors(), rtol=0, atol=1e-13)
        vn, vcn, cvn, cv = b.topology()
        ovn, ovcn, ocvn, ocv = o.topology()
        assert _sets(vn) == _sets(ovn), step
        assert _sets(vcn) == _sets(ovcn), step
        assert _canon(cv, cvn) == _canon(ocv, ocvn), step
    assert t1_total > 0
    assert b.counters()["t1_events"] == t1_total


def test_vertex_spp_graph_equals_single_steps():
    """Graph-captured self-propelled cell-vertex steps (counter-based noise) end
    bit-identical to the same steps launched one at a time."""
    cg = _cg()
    out = []
    for use_graph in (False, True):
        b = cg.VertexBatch([400], seeds=[12], precision="fp32")
        b.set_params(dt=0.01, t1_threshold=0.1)
        b.set_motility(0.3, 1.0)
        b.spp_steps(41, use_graph=use_graph)
        b.check_status()
        g = b.get()
        out.append((g["vx"], g["vy"], b.directors(), b.topology()[3], b.counters()["t1_events"]))
        b.close()
    a, c = out
    assert np.array_equal(a[0], c[0]) and np.array_equal(a[1], c[1]) and np.array_equal(a[2], c[2])
    assert np.array_equal(a[3], c[3]) and a[4] == c[4] and a[4] > 0


def test_vertex_fire_matches_oracle_fp64(oracle):
    """FIRE on the vertices (per-system scalars, FP64 sums) against the oracle's
    GPU-branch FIRE with the T1 pass as enforceTopology."""
    cg = _cg()
    Nc = 200
    b = cg.VertexBatch([Nc], seeds=[11], precision="fp64")
    b.set_params(dt=0.01, t1_threshold=0.04)
    b.set_prefs(1.0, 3.7)
    o = _oracle_like(oracle, Nc, 11, dt=0.01, thr=0.04, P0=3.7)
    f = cg.EnergyMinimizerFIRE(b)
    f.setMaximumIterations(60)
    f.setForceCutoff(1e-12)
    f.setAlphaStart(0.1)
    f.setDeltaT(0.01)
    f.minimize()
    st = f.state()
    it, so = o.fire(max_it=60, force_cutoff=1e-12, alpha_start=0.1, dt=0.01)
    assert st["iterations"] == it == 60
    b.compute_forces()
    g, r = b.get(), o.compute()
    L = float(np.sqrt(Nc))
    # reductions run in a different order (tree vs sequential)
    assert np.abs(g["vx"] - r["vx"]).max() < 1e-10 * L and np.abs(g["vy"] - r["vy"]).max() < 1e-10 * L
    assert abs(st["force_max"] - so["force_max"]) <= 1e-8 * so["force_max"]
    assert abs(st["dt"] - so["dt"]) <= 1e-12 and abs(st["alpha"] - so["alpha"]) <= 1e-12
    vn, vcn, cvn, cv = b.topology()
    ovn, ovcn, ocvn, ocv = o.topology()
    assert _canon(cv, cvn) == _canon(ocv, ocvn)


def test_vertex_fire_batch_converges_independently():
    cg = _cg()
    b = cg.VertexBatch([100, 300], seeds=[1, 2], precision="fp64")
    b.set_params(dt=0.01, t1_threshold=0.04)
    b.set_prefs(1.0, np.concatenate([np.full(100, 3.7), np.full(300, 3.75)]))
    f = cg.EnergyMinimizerFIRE(b)
    f.setMaximumIterations(20000)
    f.setForceCutoff(1e-8)
    f.setAlphaStart(0.1)
    f.minimize()
    for s in range(2):
        st = f.state(s)
        assert st["force_max"] ** 0.5 < 1e-8 and st["iterations"] < 20000
    b.compute_forces()
    g = b.get()
    assert np.sqrt(np.mean(g["Fx"][:200] ** 2 + g["Fy"][:200] ** 2)) < 1e-8
    b.check_status()


def test_vertex_spv_initialize_smooths_the_tissue():
    """runSPV start (100 self-propelled Voronoi steps before the tessellation):
    a valid vertex model whose cells are rounder than the random start's."""
    cg = _cg()
    Nc = 400
    q = []
    for spv in (False, True):
        m = cg.VertexQuadraticEnergy(Nc, 1.0, 3.8, True, runSPV=spv, seed=13377)
        m.computeForces()
        AP = m.returnAreaPeri()
        vn, vcn, cvn, cv = m.batch.topology()
        assert m.Nvertices == 2 * Nc and cvn.sum() == 3 * m.Nvertices
        assert abs(AP[:, 0].sum() - Nc) < 1e-9 * Nc and (AP[:, 0] > 0).all()
        q.append(np.mean(AP[:, 1] / np.sqrt(AP[:, 0])))
    assert q[1] < q[0]


def test_vertex_cell_division_and_death_match_oracle(oracle):
    """Vertex-model division and T2 death on the device against the oracle's
    restatement from the same state: identical new vertex positions, the same
    topology (as sets and canonical cell lists) and forces within 1e-12; then
    Brownian steps with T1 flips still match step by step."""
    cg = _cg()
    Nc = 300
    b = cg.VertexBatch([Nc], seeds=[9], precision="fp64")
    b.set_params(dt=0.005, T=1e-3, t1_threshold=0.04)
    b.set_prefs(1.0, 3.9)
    b.brownian_steps(30, use_graph=False)
    o = _oracle_like(oracle, Nc, 9, dt=0.005, T=1e-3, thr=0.04, P0=3.9)

    def load():
        g = b.get()
        vn, vcn, cvn, cv = b.topology()
        o.set_positions(g["vx"], g["vy"])
        o.set_topology(vn, vcn, cvn, cv)

    def compare():
        g, r = b.get(), o.compute()
        assert np.array_equal(g["vx"], r["vx"]) and np.array_equal(g["vy"], r["vy"])
        vn, vcn, cvn, cv = b.topology()
        ovn, ovcn, ocvn, ocv = o.topology()
        assert _sets(vn) == _sets(ovn) and _sets(vcn) == _sets(ovcn)
        assert _canon(cv, cvn) == _canon(ocv, ocvn)
        b.compute_forces()
        g = b.get()
        scale = np.abs(np.concatenate([r["Fx"], r["Fy"]])).max()
        assert np.abs(g["Fx"] - r["Fx"]).max() <= 1e-12 * scale
        np.testing.assert_allclose(g["A"], r["A"], rtol=1e-12)

    load()
    b.cell_events([1], [17], [0], [3], [0.25])
    o.cell_division(17, 0, 3, 0.25)
    assert b.Nc == o.Nc == Nc + 1 and b.Nv == o.Nv == 2 * Nc + 2
    compare()
    # a triangle divided off cell 40, then removed by a T2 transition
    load()
    b.cell_events([1], [40], [1], [2])
    o.cell_division(40, 1, 2)
    compare()
    tri = Nc + 1
    assert b.topology()[2][tri] == 3
    load()
    b.cell_events([0], [tri])
    o.cell_death(tri)
    assert b.Nc == o.Nc == Nc + 1
    compare()
    b.set_noise_buffer(True)
    rng = np.random.default_rng(4)
    for step in range(10):
        load()
        ex, ey = rng.standard_normal(b.Nv), rng.standard_normal(b.Nv)
        b.upload_noise(ex, ey)
        b.brownian_steps(1, use_graph=False)
        b.check_status()
        o.brownian_step(ex, ey)
        g, r = b.get(), o.compute()
        np.testing.assert_allclose(g["vx"], r["vx"], rtol=0, atol=1e-11)
        vn, vcn, cvn, cv = b.topology()
        ovn, ovcn, ocvn, ocv = o.topology()
        assert _canon(cv, cvn) == _canon(ocv, ocvn), step


def test_vertex_cell_events_in_a_batch():
    """A division in one system of a two-system FP32 batch leaves the other
    system's state bit for bit; graph-captured steps run on the new layout, and a
    FIRE minimizer created before the event adapts to the new size."""
    cg = _cg()
    b = cg.VertexBatch([200, 300], seeds=[1, 2], precision="fp32")
    b.set_params(dt=0.005, T=1e-3, t1_threshold=0.04)
    f = cg.EnergyMinimizerFIRE(b)
    f.setMaximumIterations(5)
    f.minimize()
    b.brownian_steps(10)
    g0 = b.get()
    vn0, vcn0, cvn0, cv0 = b.topology()
    b.cell_events([1], [250], [0], [2])
    assert b.Nc == 501 and b.Nv == 1002
    g1 = b.get()
    assert np.array_equal(g0["vx"][:400], g1["vx"][:400])
    vn1, vcn1, cvn1, cv1 = b.topology()
    assert np.array_equal(vn0[:1200], vn1[:1200]) and np.array_equal(vcn0[:1200], vcn1[:1200])
    assert cvn1.sum() == 3 * b.Nv
    b.brownian_steps(20)
    b.check_status()
    f.setMaximumIterations(20)
    f.minimize()
    b.check_status()
    assert np.isfinite(b.get()["vx"]).all()


def test_vertex_fire_kernel_layer_fused_equals_piecewise():
    """The vertex-model FIRE iteration as one call (cgpu_vm_fire_iteration) and
    as its documented sequence of calls (move, T1 pass, forces and reductions,
    parking) give bit-identical vertex positions, topology and FIRE scalars,
    driven through the C ABI with caller-owned device buffers."""
    import ctypes

    cg = _cg()
    from cellgpu_amd._lib import FireState, VertexBatchView, lib
    hip = ctypes.CDLL("libamdhip64.so")
    L = lib()
    VP, PV = ctypes.c_void_p, ctypes.POINTER(VertexBatchView)
    for name, args in (("cgpu_vm_fire_iteration", [PV, VP, VP, VP, VP]), ("cgpu_vm_fire_vv_move", [PV, VP, VP, VP]),
                       ("cgpu_vm_fire_force_reduce", [PV, VP, VP, VP, VP]),
                       ("cgpu_vm_fire_park_converged", [PV, VP, VP]), ("cgpu_vm_t1", [PV, VP]),
                       ("cgpu_vm_geometry", [PV, VP]), ("cgpu_vm_forces", [PV, VP])):
        getattr(L, name).argtypes = args
        getattr(L, name).restype = ctypes.c_int

    def dev(nbytes):
        p = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(nbytes)) == 0
        return p

    runs = []
    for fused in (True, False):
        b = cg.VertexBatch([200, 300], seeds=[3, 4], precision="fp64")
        b.set_prefs(1.0, 3.9)
        b.get()  # construction work done
        v = VertexBatchView()
        assert L.cgpu_vm_view(b.h, ctypes.byref(v)) == 0
        st = ctypes.c_void_p(L.cgpu_vm_stream(b.h))
        S, Nv = v.num_systems, v.num_vertices
        fs_h = (FireState * S)()
        for f in fs_h:  # EnergyMinimizerFIRE2D defaults, loop armed
            f.alpha = f.alpha_start = 0.99
            f.dt, f.dt_max, f.dt_inc, f.dt_dec, f.alpha_dec = 0.01, 0.1, 1.05, 0.95, 0.9
            f.force_cutoff, f.force_max, f.nmin, f.max_iterations = 1e-12, 110.0, 5, 1000
        fs, vel = dev(ctypes.sizeof(fs_h)), dev(2 * Nv * 8)
        dots = dev((3 * Nv + 4 * v.num_chunks) * 8)
        assert hip.hipMemcpy(fs, ctypes.byref(fs_h), ctypes.c_size_t(ctypes.sizeof(fs_h)), 1) == 0
        assert hip.hipMemset(vel, 0, ctypes.c_size_t(2 * Nv * 8)) == 0
        assert hip.hipDeviceSynchronize() == 0  # the batch stream does not wait for the null stream
        print("vertex FIRE kernel layer, fused=%s, S=%d Nv=%d chunks=%d" % (fused, S, Nv, v.num_chunks), flush=True)
        assert L.cgpu_vm_geometry(ctypes.byref(v), st) == 0
        assert L.cgpu_vm_forces(ctypes.byref(v), st) == 0
        for it in range(40):
            vn = VertexBatchView.from_buffer_copy(v)
            vn.pos, vn.pos_next = v.pos_next, v.pos
            if fused:
                assert L.cgpu_vm_fire_iteration(ctypes.byref(v), fs, vel, dots, st) == 0
            else:
                assert L.cgpu_vm_fire_vv_move(ctypes.byref(v), fs, vel, st) == 0
                assert L.cgpu_vm_t1(ctypes.byref(vn), st) == 0
                assert L.cgpu_vm_fire_force_reduce(ctypes.byref(vn), fs, vel, dots, st) == 0
                assert L.cgpu_vm_fire_park_converged(ctypes.byref(vn), fs, st) == 0
            v = vn
        assert hip.hipStreamSynchronize(st) == 0
        xy = np.zeros(2 * Nv)
        assert hip.hipMemcpy(xy.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(v.pos), ctypes.c_size_t(16 * Nv),
                             2) == 0
        out = (FireState * S)()
        assert hip.hipMemcpy(ctypes.byref(out), fs, ctypes.c_size_t(ctypes.sizeof(out)), 2) == 0
        b.check_status()
        topo = [np.array(a).copy() for a in b.topology()]
        runs.append((xy, [(f.iterations, f.dt, f.alpha, f.force_max) for f in out], topo))
        for p in (fs, vel, dots):
            hip.hipFree(p)
        b.close()
    assert runs[0][1][0][0] == 40
    assert np.array_equal(runs[0][0], runs[1][0])
    assert runs[0][1] == runs[1][1]
    assert all(np.array_equal(a, c) for a, c in zip(runs[0][2], runs[1][2]))
