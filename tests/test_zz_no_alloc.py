"""GPU test of the allocation-free FIRE iteration (cgpu_fire_iteration_noalloc):
through the C ABI, against cgpu_fire_iteration with a caller-allocated
workspace on an identical batch, bit for bit; two batches in a row, the second
smaller (the static scratch keeps the first call's contents); an input beyond
the scratch is refused before anything is launched."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CAP = 24


def _setup():
    import cellgpu_amd as cg
    cg.require_device()
    from cellgpu_amd._lib import FireState, VoronoiBatchView, lib
    hip = ctypes.CDLL("libamdhip64.so")
    L = lib()
    VP, PV = ctypes.c_void_p, ctypes.POINTER(VoronoiBatchView)
    for name, args in (("cgpu_fire_iteration", [PV, VP, VP, VP, VP]), ("cgpu_fire_iteration_noalloc", [PV, VP, VP, VP]),
                       ("cgpu_voronoi_forces", [PV, VP])):
        getattr(L, name).argtypes = args
        getattr(L, name).restype = ctypes.c_int
    L.cgpu_fire_noalloc_capacity.restype = ctypes.c_size_t
    return cg, hip, L, FireState, VoronoiBatchView


def _dev(hip, nbytes):
    p = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(nbytes)) == 0
    return p


def _run(n, seeds, noalloc, iters=40):
    """FIRE iterations on a fresh FP64 batch through the kernel layer; returns
    positions, FIRE scalars per system and rings."""
    cg, hip, L, FireState, View = _setup()
    b = cg.VoronoiBatch(n, seeds=seeds, precision="fp64")
    b.set_prefs(1.0, 3.7)
    b.get()
    v = b.view()
    st = ctypes.c_void_p(b.stream())
    S, N = v.num_systems, v.num_cells
    fs_h = (FireState * S)()
    for f in fs_h:  # EnergyMinimizerFIRE2D defaults, loop armed
        f.alpha = f.alpha_start = 0.99
        f.dt, f.dt_max, f.dt_inc, f.dt_dec, f.alpha_dec = 0.01, 0.1, 1.05, 0.95, 0.9
        f.force_cutoff, f.force_max, f.nmin, f.max_iterations = 1e-12, 110.0, 5, 1000
    fs, vel = _dev(hip, ctypes.sizeof(fs_h)), _dev(hip, 2 * N * 8)
    dots = None if noalloc else _dev(hip, (3 * N + 4 * v.num_chunks) * 8)
    assert hip.hipMemcpy(fs, ctypes.byref(fs_h), ctypes.c_size_t(ctypes.sizeof(fs_h)), 1) == 0
    assert hip.hipMemset(vel, 0, ctypes.c_size_t(2 * N * 8)) == 0
    assert hip.hipDeviceSynchronize() == 0
    print("FIRE %s: S=%d N=%d chunks=%d" % ("no-alloc" if noalloc else "workspace", S, N, v.num_chunks), flush=True)
    assert L.cgpu_voronoi_forces(ctypes.byref(v), st) == 0
    for _ in range(iters):
        vn = View.from_buffer_copy(v)
        vn.pos, vn.pos_next = v.pos_next, v.pos
        if noalloc:
            assert L.cgpu_fire_iteration_noalloc(ctypes.byref(v), fs, vel, st) == 0
        else:
            assert L.cgpu_fire_iteration(ctypes.byref(v), fs, vel, dots, st) == 0
        v = vn
    assert hip.hipStreamSynchronize(st) == 0
    xy = np.zeros(2 * N)
    assert hip.hipMemcpy(xy.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(v.pos), ctypes.c_size_t(16 * N), 2) == 0
    out = (FireState * S)()
    assert hip.hipMemcpy(ctypes.byref(out), fs, ctypes.c_size_t(ctypes.sizeof(out)), 2) == 0
    b.check_status()
    nn, nbr, _ = b.topology(CAP)
    for p in (fs, vel, dots):
        if p is not None:
            hip.hipFree(p)
    b.close()
    return xy, [(f.iterations, f.dt, f.alpha, f.force_max, f.power) for f in out], nn, nbr


def test_noalloc_fire_iteration_equals_workspace_version():
    # first a larger batch, then a smaller one: the second call reuses the
    # scratch that still holds the first call's products
    for n, seeds in (([3000, 1200, 700], [5, 6, 7]), ([400], [9])):
        ref = _run(n, seeds, noalloc=False)
        got = _run(n, seeds, noalloc=True)
        assert got[1][0][0] == 40
        assert np.array_equal(ref[0], got[0])
        assert ref[1] == got[1]
        assert np.array_equal(ref[2], got[2]) and np.array_equal(ref[3], got[3])


def test_noalloc_refuses_inputs_beyond_the_scratch():
    cg, hip, L, FireState, View = _setup()
    cap = int(L.cgpu_fire_noalloc_capacity())
    N = cap // 3 + 1000  # 3 N > capacity
    b = cg.VoronoiBatch([N], seeds=[3], precision="fp64")
    g0 = b.get()
    v = b.view()
    st = ctypes.c_void_p(b.stream())
    fs_h = (FireState * 1)()
    fs, vel = _dev(hip, ctypes.sizeof(fs_h)), _dev(hip, 2 * N * 8)
    r = L.cgpu_fire_iteration_noalloc(ctypes.byref(v), fs, vel, st)
    assert r == -8
    assert b"scratch" in L.cgpu_last_error()
    assert hip.hipStreamSynchronize(st) == 0
    g1 = b.get()
    assert np.array_equal(g0["x"], g1["x"]) and np.array_equal(g0["y"], g1["y"])
    hip.hipFree(fs)
    hip.hipFree(vel)
    b.close()
