"""GPU parity of the on-device spatial sorting (reference spatialSorting:
Simple2DCell.cpp:438-599, voronoiModelBase.cpp:669-680, vertexModelBase.cpp:236-242)
against the oracle's restatement of the reference's HilbertSorter.

Bars: the permutation equals the reference order exactly; every per-cell and
per-vertex array follows it bit for bit; the rings after the sort equal the
oracle's triangulation; a run sorted part way equals an unsorted run in label
space (FP64, 1e-10 of the box); the vertex model's forces after the sort are
the unsorted forces re-indexed, bit for bit.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CAP = 24


def _cg():
    import cellgpu_amd as cg
    cg.require_device()
    return cg


def _tags(b):
    from cellgpu_amd._lib import lib
    t = np.zeros(b.N, np.int32)
    assert lib().cgpu_vor_get_tags(b.h, t.ctypes.data_as(ctypes.c_void_p)) == 0
    return t


def test_voronoi_sort_is_the_reference_hilbert_order(oracle):
    cg = _cg()
    n = [500, 1300, 64]
    b = cg.VoronoiBatch(n, seeds=[4, 5, 6], precision="fp64")
    N = sum(n)
    rng = np.random.default_rng(3)
    b.set_prefs(1.0 + 0.1 * rng.random(N), 3.6 + 0.3 * rng.random(N))
    b.set_motility(0.02 + 0.1 * rng.random(N), 0.5 + rng.random(N))
    b.spv_steps(30)
    b.compute_forces()
    g0, t0 = b.get(), _tags(b)
    print("voronoi sort: systems %s" % n, flush=True)
    perm = b.spatial_sort()
    g1 = b.get()
    for s in range(len(n)):
        sl = b.system_slice(s)
        expect = sl.start + oracle.hilbert_sort_order(g0["x"][sl], g0["y"][sl], b.L[s])
        assert np.array_equal(perm[sl], expect)
        nn, nbr, _ = b.topology(CAP)
        onn, onbr, _ = oracle.triangulate(g1["x"][sl], g1["y"][sl], b.L[s], cap=CAP)
        loc = [tuple(int(v) - sl.start for v in nbr[i, : nn[i]]) for i in range(sl.start, sl.stop)]
        assert loc == [tuple(onbr[i, : onn[i]]) for i in range(len(onn))]
    for k in ("x", "y", "theta", "vx", "vy", "Fx", "Fy"):
        assert np.array_equal(g1[k], g0[k][perm]), k
    assert np.array_equal(_tags(b), t0[perm])


def test_voronoi_kernel_layer_order_through_the_c_abi(oracle):
    """cgpu_vor_hilbert_order on a handle's view with a caller-owned workspace."""
    cg = _cg()
    from cellgpu_amd._lib import lib
    hip = ctypes.CDLL("libamdhip64.so")
    b = cg.VoronoiBatch([700, 900], seeds=[8, 9], precision="fp32")
    g = b.get()
    v = b.view()
    N = v.num_cells
    wb = lib().cgpu_vor_sort_workspace(N)
    work, perm = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(work), ctypes.c_size_t(wb)) == 0
    assert hip.hipMalloc(ctypes.byref(perm), ctypes.c_size_t(4 * N)) == 0
    print("kernel-layer order: N=%d workspace %d bytes" % (N, wb), flush=True)
    st = ctypes.c_void_p(b.stream())
    try:
        assert lib().cgpu_vor_hilbert_order(ctypes.byref(v), perm, work, wb, st) == 0
        assert lib().cgpu_vor_hilbert_order(ctypes.byref(v), perm, work, wb - 4096, st) != 0
        b.sync()
        p = np.zeros(N, np.int32)
        assert hip.hipMemcpy(p.ctypes.data_as(ctypes.c_void_p), perm, ctypes.c_size_t(4 * N), 2) == 0
    finally:
        hip.hipFree(work)
        hip.hipFree(perm)
    for s in range(2):
        sl = b.system_slice(s)
        assert np.array_equal(p[sl], sl.start + oracle.hilbert_sort_order(g["x"][sl], g["y"][sl], b.L[s]))


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_sorted_run_equals_unsorted_run_in_label_space(oracle, precision):
    """Two copies step with the same per-label noise; one is sorted after 10
    steps, the other only retriangulated there (the sort's own topology
    rebuild). After 15 more steps they agree label by label."""
    cg = _cg()
    n = [400, 900]
    N = sum(n)
    rng = np.random.default_rng(11)
    P0 = 3.6 + 0.3 * rng.random(N)
    v0, Dr = 0.05 + 0.1 * rng.random(N), 0.5 + rng.random(N)
    runs = []
    for sort in (False, True):
        b = cg.VoronoiBatch(n, seeds=[21, 22], precision=precision)
        b.set_prefs(1.0, P0)
        b.set_motility(v0, Dr)
        b.set_noise_buffer(True)
        runs.append(b)
    A, B = runs
    noise = np.random.default_rng(5).standard_normal((25, N))
    lab = np.arange(N)  # current index -> label of the sorted run
    print("sorted vs unsorted, %s" % precision, flush=True)
    for k in range(25):
        if k == 10:
            perm = B.spatial_sort()
            lab = lab[perm]
            gA = A.get()
            A.set_positions(gA["x"], gA["y"], True)
        A.upload_noise(noise[k])
        B.upload_noise(noise[k][lab])
        A.spv_steps(1, use_graph=False)
        B.spv_steps(1, use_graph=False)
    A.check_status()
    B.check_status()
    gA, gB = A.get(), B.get()
    tol = 1e-10 if precision == "fp64" else 1e-4
    for s in range(2):
        sl = A.system_slice(s)
        L = A.L[s]
        dx = gB["x"][sl] - gA["x"][lab[sl]]
        dy = gB["y"][sl] - gA["y"][lab[sl]]
        dx -= L * np.round(dx / L)
        dy -= L * np.round(dy / L)
        assert np.abs(dx).max() < tol * L and np.abs(dy).max() < tol * L


def test_simulation_sort_period_keeps_the_trajectory():
    """Simulation.setSortPeriod: a run sorted every 7 steps equals the unsorted
    run (v0 = 0, so the index-keyed director noise does not enter the motion)."""
    cg = _cg()
    out = []
    for sp in (-1, 7):
        m = cg.VoronoiQuadraticEnergy(600, P0=3.75, precision="fp64", seed=77)
        m.setv0Dr(0.0, 1.0)
        sim = cg.Simulation()
        sim.setConfiguration(m)
        sim.addUpdater(cg.selfPropelledParticleDynamics(600), m)
        sim.setIntegrationTimestep(0.01)
        sim.setSortPeriod(sp)
        sim.performTimesteps(30)
        x = m.returnPositions()
        if sp > 0:
            xs = np.empty_like(x)
            xs[m.idxToTag] = x
            x = xs
            assert not np.array_equal(m.idxToTag, np.arange(600))
        out.append(x)
    L = np.sqrt(600.0)
    d = out[0] - out[1]
    d -= L * np.round(d / L)
    assert np.abs(d).max() < 1e-10 * L


def test_vertex_sort_is_the_reference_order_and_keeps_the_state(oracle):
    cg = _cg()
    n = [300, 500]
    vb = cg.VertexBatch(n, seeds=[1, 2], precision="fp64")
    rng = np.random.default_rng(7)
    vb.set_prefs(1.0, 3.7 + 0.3 * rng.random(vb.Nc))
    vb.set_motility(0.05 + 0.1 * rng.random(vb.Nc), 0.5 + rng.random(vb.Nc))
    vb.set_directors(2 * np.pi * rng.random(vb.Nc))
    vb.set_params(T=0.0, dt=0.01)
    vb.spp_steps(20, use_graph=False)
    vb.compute_forces()
    g0, d0 = vb.get(), vb.directors()
    vn0, vcn0, cvn0, cv0 = vb.topology()
    voff = np.concatenate([[0], np.cumsum([vb.counters(s)["nv"] for s in range(2)])])
    coff = np.concatenate([[0], np.cumsum(n)])
    print("vertex sort: cells %s vertices %d" % (n, vb.Nv), flush=True)
    vp, cp = vb.spatial_sort()
    for s in range(2):
        vs, cs = slice(voff[s], voff[s + 1]), slice(coff[s], coff[s + 1])
        loc_vcn = vcn0.reshape(-1, 3)[vs] - coff[s]
        ev, ec = oracle.vertex_sort_orders(g0["vx"][vs], g0["vy"][vs], loc_vcn, np.sqrt(n[s]))
        assert np.array_equal(vp[vs], voff[s] + ev)
        assert np.array_equal(cp[cs], coff[s] + ec)
    vinv, cinv = np.empty_like(vp), np.empty_like(cp)
    vinv[vp] = np.arange(len(vp))
    cinv[cp] = np.arange(len(cp))
    vn1, vcn1, cvn1, cv1 = vb.topology()
    assert np.array_equal(vn1.reshape(-1, 3), vinv[vn0.reshape(-1, 3)[vp]])
    assert np.array_equal(vcn1.reshape(-1, 3), cinv[vcn0.reshape(-1, 3)[vp]])
    assert np.array_equal(cvn1, cvn0[cp])
    for c in range(vb.Nc):
        assert np.array_equal(cv1[c, : cvn1[c]], vinv[cv0[cp[c], : cvn0[cp[c]]]])
    g1 = vb.get()
    for k in ("vx", "vy"):
        assert np.array_equal(g1[k], g0[k][vp]), k
    for k in ("A", "P"):
        assert np.array_equal(g1[k], g0[k][cp]), k
    assert np.array_equal(vb.directors(), d0[cp])
    vb.compute_forces()
    g2 = vb.get()
    for k in ("Fx", "Fy"):
        assert np.array_equal(g2[k], g0[k][vp]), k
    vb.check_status()
