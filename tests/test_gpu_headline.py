"""GPU parity at the headline ensemble's shape and of the topology machinery's
intermediate results, against the FP64 oracle.

* The 512-system FP32 ensemble (the benchmark's workload): each checked step's
  new rings equal the oracle's reference test-and-repair (voronoiModelBase.cpp:
  600-663) applied to the GPU's moved positions from the GPU's previous rings,
  on sampled systems (smallest, largest, six between); FP32 forces against FP64
  at the same positions and rings.
* FP32 single system, step by step, the same topology check.
* Circumcircle-test flags and the NeedsFixing set of every step against the
  oracle's, through the kernel layer of the C ABI; the reference search window
  on hand-built fixtures (an intruder outside the first vertex's bins).
* The launch-shape autotuner restores the state it tunes on, and every launch
  shape gives bit-identical results; more work chunks than star work-groups
  (the global queue is filled by a block loop).
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


def _hip():
    return ctypes.CDLL("libamdhip64.so")


def _d2h(hip, dst, src_ptr, nbytes):
    assert hip.hipMemcpy(dst.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(src_ptr), ctypes.c_size_t(nbytes), 2) == 0


def _h2d(hip, dst_ptr, src, nbytes):
    assert hip.hipMemcpy(ctypes.c_void_p(dst_ptr), src.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(nbytes), 1) == 0


def _local(nbr, off):
    return np.where(nbr >= 0, nbr - off, -1).astype(np.int32)


def oracle_topology_step(O, L, x, y, pre_nn, pre_nbr, pre_nm, sl):
    """The oracle's GPU-branch enforceTopology on positions (x, y) of one system
    starting from the rings pre_*: returns its rings (nn, nbr local), neighMax,
    NeedsFixing set and repair mode."""
    off = sl.start
    o = O.VoronoiSystem(x[sl], y[sl], np.zeros(sl.stop - off), L)
    o.set_topology(pre_nn[sl], _local(pre_nbr[sl], off), int(pre_nm))
    o.enforce_topology(gpu_path=True)
    nn, nbr, nm = o.topology(CAP)
    nf, mode = o.last_needs_fixing()
    return nn, nbr, nm, nf, mode


def _same_rings(nn_a, nbr_a, nn_b, nbr_b):
    if not np.array_equal(nn_a, nn_b):
        return False
    return all(np.array_equal(nbr_a[i, :nn_a[i]], nbr_b[i, :nn_b[i]]) for i in range(len(nn_a)))


def _check_steps(O, b, systems, steps):
    """`steps` single steps of batch b; after each, every listed system's rings
    must equal the oracle's repair of the GPU positions from the previous rings.
    Returns the number of (system, step) pairs that repaired."""
    repaired = 0
    for _ in range(steps):
        nn0, nbr0, nm0 = b.topology(CAP)
        b.spv_steps(1, use_graph=False)
        b.check_status()
        g = b.get()
        nn1, nbr1, nm1 = b.topology(CAP)
        for s in systems:
            sl = b.system_slice(s)
            onn, onbr, onm, nf, mode = oracle_topology_step(O, b.L[s], g["x"], g["y"], nn0, nbr0, nm0[s], sl)
            assert _same_rings(nn1[sl], _local(nbr1[sl], sl.start), onn, onbr), "system %d" % s
            assert int(nm1[s]) == onm, "neighMax of system %d" % s
            repaired += mode > 0
    return repaired


def _bench():
    import bench
    return bench


def test_headline_ensemble_topology_and_fp32_forces(oracle):
    """The benchmark's 512-system FP32 ensemble, 20 graph steps from the random
    start, then three checked steps on eight systems; FP32 forces against FP64
    at the same positions and rings: max |dF| / max |F| over the ensemble
    within 1e-5."""
    cg = _cg()
    B = _bench()
    n, P0, v0, Dr, seeds = B.ensemble()
    b = B.make_batch(cg, n, P0, v0, Dr, seeds, "fp32")
    print("ensemble: %d systems, %d cells" % (len(n), int(n.sum())), flush=True)
    b.spv_steps(20, use_graph=True)
    b.check_status()
    order = np.argsort(n)
    systems = sorted(set([int(order[0]), int(order[-1])] + [int(order[k]) for k in np.linspace(0, len(n) - 1, 8)[1:-1].astype(int)]))
    assert len(systems) == 8
    repaired = _check_steps(oracle, b, systems, 3)
    assert repaired > 0
    # FP32 against FP64 at the same state
    g = b.get()
    nn, nbr, _ = b.topology(CAP)
    b64 = B.make_batch(cg, n, P0, v0, Dr, seeds, "fp64")
    b64.set_positions(g["x"], g["y"])
    b64.set_topology(nn, nbr, b.del_other())
    b64.compute_forces()
    b.compute_forces()
    g32, g64 = b.get(), b64.get()
    dF = np.hypot(g32["Fx"] - g64["Fx"], g32["Fy"] - g64["Fy"])
    F = np.hypot(g64["Fx"], g64["Fy"])
    err = dF.max() / F.max()
    off = np.concatenate([[0], np.cumsum(n)[:-1]])
    per = np.maximum.reduceat(dF, off) / np.maximum.reduceat(F, off)
    print("fp32 force error: ensemble %.3g, per system median %.3g, worst %.3g (system %d)" % (
        err, np.median(per), per.max(), per.argmax()), flush=True)
    assert err <= 1e-5, err
    A_err = np.abs(g32["A"] - g64["A"]).max()
    assert A_err <= 1e-5, A_err
    b64.close()
    b.close()


@pytest.mark.parametrize("N", [400, 3000])
def test_fp32_topology_stepwise(oracle, N):
    """FP32 bulk: every step's rings equal the oracle's reference repair of the
    GPU's (FP32-integrated, FP64-stored) positions from the previous rings."""
    cg = _cg()
    b = cg.VoronoiBatch([N], seeds=[17], precision="fp32")
    b.set_prefs(1.0, 3.8)
    b.set_motility(0.1, 1.0)
    assert _check_steps(oracle, b, [0], 40) > 0


def _kernel_layer():
    from cellgpu_amd._lib import VoronoiBatchView, lib
    L = lib()
    VP, PV = ctypes.c_void_p, ctypes.POINTER(VoronoiBatchView)
    for name, args in (("cgpu_step_begin", [PV, VP]), ("cgpu_compute_cell_list", [PV, ctypes.c_int, VP]),
                       ("cgpu_test_and_geometry", [PV, ctypes.c_int, VP]), ("cgpu_repair_topology", [PV, VP]),
                       ("cgpu_spv_force_eom_move", [PV, VP]), ("cgpu_step_end", [PV, VP])):
        getattr(L, name).argtypes = args
        getattr(L, name).restype = ctypes.c_int
    return L, VoronoiBatchView


@pytest.mark.parametrize("relax", [0, 400])
def test_flags_and_needs_fixing_match_oracle(oracle, relax):
    """Kernel layer, step by step (FP64, two systems), from the random start and
    after `relax` steps: the circumcircle-test flags of every cell and each
    system's NeedsFixing set equal the oracle's on the same positions and
    previous rings. From the random start the local repair leaves rings that
    disagree, which turns the test's skip rule off; after relaxation the rule
    is in force and most owned triangles never reach the scan."""
    cg = _cg()
    hip = _hip()
    L, View = _kernel_layer()
    b = cg.VoronoiBatch([500, 1500], seeds=[21, 22], precision="fp64")
    b.set_motility(0.1, 1.0)
    if relax:
        b.spv_steps(relax, use_graph=False)
        b.check_status()
    v = b.view()
    st = ctypes.c_void_p(b.stream())
    N = v.num_cells
    nflags_total = 0
    queued = []
    for step in range(25):
        nn0, nbr0, nm0 = b.topology(CAP)
        vn = View.from_buffer_copy(v)
        vn.pos, vn.pos_next = v.pos_next, v.pos
        assert L.cgpu_step_begin(ctypes.byref(v), st) == 0
        assert L.cgpu_spv_force_eom_move(ctypes.byref(v), st) == 0
        assert L.cgpu_compute_cell_list(ctypes.byref(v), 1, st) == 0
        assert L.cgpu_test_and_geometry(ctypes.byref(vn), 1, st) == 0
        assert hip.hipStreamSynchronize(st) == 0
        flags = np.zeros(N, np.int32)
        _d2h(hip, flags, vn.repair, 4 * N)
        xy = np.zeros(2 * N)
        _d2h(hip, xy, vn.pos, 16 * N)
        x, y = xy[0::2].copy(), xy[1::2].copy()
        assert L.cgpu_repair_topology(ctypes.byref(vn), st) == 0
        assert L.cgpu_step_end(ctypes.byref(vn), st) == 0
        assert hip.hipStreamSynchronize(st) == 0
        cnt = np.zeros(16, np.int32)
        _d2h(hip, cnt, vn.counters, 64)
        queued.append(int(cnt[9]))  # triangles left to the scan in this step
        nfq = np.zeros(max(int(cnt[1]), 1), np.int32)
        if cnt[1] > 0:
            _d2h(hip, nfq, vn.nfq, 4 * int(cnt[1]))
        nfq = np.sort(nfq[:int(cnt[1])])
        nn1, nbr1, nm1 = b.topology(CAP)
        for s in range(2):
            sl = b.system_slice(s)
            o = oracle.VoronoiSystem(x[sl], y[sl], np.zeros(sl.stop - sl.start), b.L[s])
            o.set_topology(nn0[sl], _local(nbr0[sl], sl.start), int(nm0[s]))
            _, of = o.test_flags()
            o.enforce_topology(gpu_path=True)
            onf, mode = o.last_needs_fixing()
            mine = nfq[(nfq >= sl.start) & (nfq < sl.stop)] - sl.start
            n_s = sl.stop - sl.start
            if int(of.sum()) > n_s // 6:
                # more than N/6 flagged cells decide a global retriangulation
                # (|NeedsFixing| >= |flagged|); the GPU test stops scanning such
                # a system, so its flags are a subset with the same decision
                assert np.all(flags[sl] <= of) and int(flags[sl].sum()) > n_s // 6, "step %d system %d" % (step, s)
                assert len(mine) > n_s // 6 and len(onf) > n_s // 6
            else:
                assert np.array_equal(flags[sl], of), "step %d system %d: flags" % (step, s)
                assert np.array_equal(mine, onf), "step %d system %d: NeedsFixing" % (step, s)
            onn, onbr, onm = o.topology(CAP)
            assert _same_rings(nn1[sl], _local(nbr1[sl], sl.start), onn, onbr), "step %d system %d" % (step, s)
            nflags_total += int(of.sum())
        v = vn
    assert nflags_total > 0
    # a triangulation of the torus has 2 N triangles, each owned by one cell
    if relax:
        assert 0 < sum(queued) < 0.5 * 25 * 2 * N, queued
    b.close()


@pytest.mark.parametrize("variant", ["outside", "inside", "outside_wrapped", "inside_wrapped", "empty"])
def test_reference_window_fixture_on_gpu(oracle, variant):
    """The hand-built search-window fixtures of the oracle tests, on the GPU
    test kernel (rings written through the kernel layer's view): an intruder in
    the circle but outside the first vertex's bins is not flagged; inside them
    the triangle and the intruder are; the same across the periodic boundary."""
    import test_oracle_rules as R
    cg = _cg()
    hip = _hip()
    L, View = _kernel_layer()
    intr = {"outside": 0.81, "inside": -0.2, "outside_wrapped": 0.81, "inside_wrapped": -0.2, "empty": 1.3}[variant]
    shift = 9.4 if "wrapped" in variant else 0.0
    x, y, nn, nbr = R.n3_fixture(intr, shift)
    N = len(x)
    b = cg.VoronoiBatch([N], L=[R.L_FIX], x=x, y=y, theta=np.zeros(N), precision="fp64")
    v = b.view()
    st = ctypes.c_void_p(b.stream())
    ring = np.zeros(CAP * N, np.int32)
    for i in range(N):
        for k in range(nn[i]):
            ring[k * N + i] = nbr[i, k]
    _h2d(hip, v.nn, np.ascontiguousarray(nn, np.int32), 4 * N)
    _h2d(hip, v.nbr, ring, 4 * CAP * N)
    assert L.cgpu_step_begin(ctypes.byref(v), st) == 0
    assert L.cgpu_compute_cell_list(ctypes.byref(v), 0, st) == 0
    assert L.cgpu_test_and_geometry(ctypes.byref(v), 1, st) == 0
    assert hip.hipStreamSynchronize(st) == 0
    flags = np.zeros(N, np.int32)
    _d2h(hip, flags, v.repair, 4 * N)
    gpu_ref, _ = R._flags(oracle, x, y, nn, nbr)
    assert np.array_equal(flags, gpu_ref)
    expect = [0, 1, 2, 3] if variant.startswith("inside") else []
    assert list(np.flatnonzero(flags)) == expect
    b.close()


def test_autotune_restores_state_and_shapes_give_identical_results():
    cg = _cg()
    n = [300, 1000, 2500, 64]
    a = cg.VoronoiBatch(n, seeds=[1, 2, 3, 4], precision="fp32")
    c = cg.VoronoiBatch(n, seeds=[1, 2, 3, 4], precision="fp32")
    for b in (a, c):
        b.set_motility(0.1, 1.0)
        b.spv_steps(30)
    g0 = a.get()
    shape, cached = a.autotune(trial_steps=4)
    print("tuned shape", shape, "cached", cached, flush=True)
    g1 = a.get()
    for k in ("x", "y", "theta", "vx", "vy"):
        assert np.array_equal(g0[k], g1[k]), k
    assert a.counters(0)["step"] == c.counters(0)["step"]
    # a second tuning of the same workload comes from the cache
    _, cached2 = a.autotune(trial_steps=4)
    assert cached2
    # an extreme shape on the other copy
    c.set_launch_shape(force_block=128, test_block=512, cell_block=64, star_blocks=64, repair_blocks=64,
                       queue_lanes=16)
    assert c.launch_shape()["star_blocks"] == 64
    for b in (a, c):
        b.spv_steps(40)
        b.check_status()
    ga, gc = a.get(), c.get()
    for k in ("x", "y", "theta", "vx", "vy", "Fx", "Fy", "A", "P"):
        assert np.array_equal(ga[k], gc[k]), k
    na, ra, ma = a.topology(CAP)
    nc, rc, mc = c.topology(CAP)
    assert np.array_equal(na, nc) and np.array_equal(ra, rc) and np.array_equal(ma, mc)
    a.close()
    c.close()


def test_more_chunks_than_star_work_groups(oracle):
    """4200 systems of 64 cells (4200 work chunks) with 64 star work-groups:
    each block of the first star pass queues many systems' chunks for global
    retriangulation in a loop. Sampled systems follow the oracle step by step
    and every step conserves the 6N degree sum of every system."""
    cg = _cg()
    S = 4200
    b = cg.VoronoiBatch([64] * S, seeds=np.arange(1, S + 1), precision="fp32")
    b.set_motility(0.2, 1.0)
    b.set_launch_shape(star_blocks=64)
    print("systems %d, chunks %d" % (S, b.view().num_chunks), flush=True)
    assert b.view().num_chunks > 4096
    _check_steps(oracle, b, [0, 1, 777, 2048, 4199], 6)
    nn, _, _ = b.topology(CAP)
    assert np.all(nn.reshape(S, 64).sum(axis=1) == 6 * 64)
    g = [b.counters(s)["global_fixes"] for s in (0, 1, 777, 2048, 4199)]
    assert sum(g) > 0
    b.close()


def test_profiled_steps_equal_graph_steps():
    """The per-kernel event path the benchmark times (one event after every
    kernel, no graph) runs the same kernels: the state equals graph-captured
    steps bit for bit, and every stage of the step is named and timed."""
    cg = _cg()
    n = [300, 1000, 64]
    a = cg.VoronoiBatch(n, seeds=[5, 6, 7], precision="fp32")
    c = cg.VoronoiBatch(n, seeds=[5, 6, 7], precision="fp32")
    for b in (a, c):
        b.set_motility(0.1, 1.0)
        b.spv_steps(10)
    ms, q = a.profiled_steps(6)
    c.spv_steps(6)
    for b in (a, c):
        b.check_status()
    ga, gc = a.get(), c.get()
    for k in ("x", "y", "theta", "vx", "vy", "Fx", "Fy", "A", "P"):
        assert np.array_equal(ga[k], gc[k]), k
    assert list(ms) == ["k_force_eom_move", "k_bin_scan", "k_bin_scatter", "k_test_geometry", "k_ld_reduce",
                        "k_select_triangles", "k_test_queue", "k_mark_nf", "k_star_local_wave",
                        "k_delsets_local", "k_global_fallback", "k_star_global_wave", "k_global_stats",
                        "k_finalize"]
    assert all(v > 0.0 for v in ms.values())
    assert len(q) == 5 and min(q) >= 0 and q[3] > 0  # triangles queued for the scan
    a.close()
    c.close()
