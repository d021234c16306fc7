"""Kernel-layer batches in caller-owned memory (cgpu_vor_workspace_query /
cgpu_vor_workspace_bind): one hipMalloc by the caller, the batch carved out of
it, then the same triangulation and self-propelled steps as a handle batch
with the same state. Bar: positions and Delaunay rings bit-identical to the
handle's after 30 steps (FP64 and FP32 bulk), step counters equal, no error."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CAP = 24


def _rings(nn, nbr):
    return [tuple(nbr[i, : nn[i]]) for i in range(len(nn))]


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_workspace_batch_equals_handle_batch(precision):
    import cellgpu_amd as cg
    cg.require_device()
    from cellgpu_amd._lib import SysParams, SysState, VorWorkspace, VoronoiBatchView, lib
    hip = ctypes.CDLL("libamdhip64.so")
    L = lib()
    PV = ctypes.POINTER(VoronoiBatchView)
    for name in ("cgpu_global_triangulation", "cgpu_spv_step"):
        getattr(L, name).argtypes = [PV, ctypes.c_void_p]
        getattr(L, name).restype = ctypes.c_int
    prec = 1 if precision == "fp32" else 0
    real = np.float32 if prec else np.float64

    n = np.array([300, 700], np.int32)
    seeds = np.array([5, 6], np.uint64)
    b = cg.VoronoiBatch(list(n), seeds=list(seeds), precision=precision)
    b.set_prefs(1.0, 3.7)
    b.set_motility(0.1, 1.0)
    g = b.get()
    N = int(n.sum())

    S = len(n)
    params = (SysParams * S)()
    assert L.cgpu_vor_default_params(S, n.ctypes.data, None, seeds.ctypes.data, params) == 0
    assert [p.off for p in params] == [0, 300] and [p.n for p in params] == [300, 700]
    ws = VorWorkspace()
    assert L.cgpu_vor_workspace_query(S, params, prec, ctypes.byref(ws)) == 0
    assert ws.num_cells == N and ws.num_systems == S
    # the handle's arrays plus alignment padding
    L.cgpu_vor_device_bytes.restype = ctypes.c_longlong
    held = L.cgpu_vor_device_bytes(b.h)
    assert held <= ws.bytes <= held + 256 * 40
    print("workspace %s: %d bytes for %d cells" % (precision, ws.bytes, N), flush=True)

    mem = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(mem), ctypes.c_size_t(ws.bytes)) == 0
    try:
        v = VoronoiBatchView()
        # refused before any device work: misaligned, too small
        assert L.cgpu_vor_workspace_bind(ctypes.byref(ws), params, ctypes.c_void_p(mem.value + 8), ws.bytes - 8,
                                         ctypes.byref(v), None) == -1
        assert L.cgpu_vor_workspace_bind(ctypes.byref(ws), params, mem, ws.bytes - 1, ctypes.byref(v), None) == -8
        assert L.cgpu_vor_workspace_bind(ctypes.byref(ws), params, mem, ws.bytes, ctypes.byref(v), None) == 0
        assert v.num_cells == N and v.pos and v.pos_next and v.nbr

        def put(ptr, arr):
            arr = np.ascontiguousarray(arr)
            assert hip.hipMemcpy(ctypes.c_void_p(ptr), arr.ctypes.data_as(ctypes.c_void_p),
                                 ctypes.c_size_t(arr.nbytes), 1) == 0

        def inter(a, c):
            out = np.empty(2 * N, a.dtype)
            out[0::2], out[1::2] = a, c
            return out

        put(v.pos, inter(g["x"], g["y"]))
        put(v.APpref, inter(np.full(N, 1.0, real), np.full(N, 3.7, real)))
        put(v.mot, inter(np.full(N, 0.1, real), np.full(N, 1.0, real)))
        put(v.vel, inter(g["vx"].astype(real), g["vy"].astype(real)))
        put(v.theta, g["theta"].astype(real))
        assert hip.hipDeviceSynchronize() == 0
        assert L.cgpu_global_triangulation(ctypes.byref(v), None) == 0

        def rings_of(view):
            nn = np.zeros(N, np.int32)
            ring = np.zeros(CAP * N, np.int32)
            assert hip.hipDeviceSynchronize() == 0
            assert hip.hipMemcpy(nn.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(view.nn),
                                 ctypes.c_size_t(4 * N), 2) == 0
            assert hip.hipMemcpy(ring.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(view.nbr),
                                 ctypes.c_size_t(4 * CAP * N), 2) == 0
            return _rings(nn, ring.reshape(CAP, N).T)  # slot-major on the device

        hn, hr, _ = b.topology(CAP)
        assert rings_of(v) == _rings(hn, hr)

        steps = 30
        b.spv_steps(steps, use_graph=False)
        for _ in range(steps):
            assert L.cgpu_spv_step(ctypes.byref(v), None) == 0
            v.pos, v.pos_next = v.pos_next, v.pos
        assert hip.hipDeviceSynchronize() == 0
        b.check_status()
        g2 = b.get()
        xy = np.zeros(2 * N)
        assert hip.hipMemcpy(xy.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(v.pos), ctypes.c_size_t(16 * N),
                             2) == 0
        assert np.array_equal(xy[0::2], g2["x"]) and np.array_equal(xy[1::2], g2["y"])
        hn, hr, _ = b.topology(CAP)
        assert rings_of(v) == _rings(hn, hr)
        # per-system state: no error, same step counter as the handle
        st = (SysState * S)()
        assert hip.hipMemcpy(ctypes.byref(st), ctypes.c_void_p(v.st), ctypes.c_size_t(ctypes.sizeof(st)), 2) == 0
        for s in range(S):
            assert st[s].error == 0 and st[s].step == steps == b.counters(s)["step"]
    finally:
        hip.hipFree(mem)
        b.close()
