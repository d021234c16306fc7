"""The secondary workloads at their full sizes.

* Vertex model, Nc = 32 768 (Nv = 65 536, L = 181.02), Brownian T = 1e-3,
  dt = 0.005, T1 threshold 0.04, P0 = 3.9: FP32-bulk forces within 1e-5 of the
  largest FP64 force on the same state; after 10 steps with the same noise the
  two precisions have the same topology and positions within 1e-5; after 200
  FP32 steps the bookkeeping invariants hold.
* Voronoi FIRE, N = 262 144 (L = 512), P0 = 3.75, dtMax 0.1, alpha0 0.1: run
  until the RMS force (the GPU branch's loop test, EnergyMinimizerFIRE2D.cpp:184,
  277) is below 1e-10, then on until the largest force component (the CPU
  branch's test, :227) is below 1e-10; both figures are printed.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cg():
    import cellgpu_amd as cg
    cg.require_device()
    return cg


def test_vertex_full_size_fp32_against_fp64():
    cg = _cg()
    Nc = 32768
    bs = {}
    for prec in ("fp64", "fp32"):
        b = cg.VertexBatch([Nc], seeds=[21], precision=prec)
        b.set_params(dt=0.005, T=1e-3, t1_threshold=0.04)
        b.set_prefs(1.0, 3.9)
        b.compute_forces()
        bs[prec] = b
    g64, g32 = bs["fp64"].get(), bs["fp32"].get()
    F = np.abs(np.concatenate([g64["Fx"], g64["Fy"]])).max()
    err = max(np.abs(g32["Fx"] - g64["Fx"]).max(), np.abs(g32["Fy"] - g64["Fy"]).max())
    print("Nc 32768: FP32 force error %.3e of max |F| %.3e" % (err / F, F), flush=True)
    assert err <= 1e-5 * F
    for b in bs.values():
        b.brownian_steps(10)
        b.check_status()
    g64, g32 = bs["fp64"].get(), bs["fp32"].get()
    for a, c in zip(bs["fp64"].topology(), bs["fp32"].topology()):
        assert np.array_equal(a, c)
    d = max(np.abs(g32["vx"] - g64["vx"]).max(), np.abs(g32["vy"] - g64["vy"]).max())
    print("after 10 steps: max position difference %.3e" % d, flush=True)
    assert d <= 1e-5
    assert bs["fp32"].counters(0)["t1_events"] == bs["fp64"].counters(0)["t1_events"] > 0

    b = bs["fp32"]
    b.brownian_steps(190)
    b.check_status()
    b.compute_forces()
    g = b.get()
    vn, vcn, cvn, cv = b.topology()
    L = float(np.sqrt(Nc))
    assert b.Nv == 2 * Nc and cvn.sum() == 3 * b.Nv
    for v in range(0, b.Nv, 97):
        for c in vcn[3 * v: 3 * v + 3]:
            assert v in cv[c, : cvn[c]]
    assert abs(g["A"].sum() - L * L) < 1e-5 * L * L
    assert (g["A"] > 0).all()
    F = np.abs(np.concatenate([g["Fx"], g["Fy"]])).max()
    assert abs(g["Fx"].sum()) < 1e-5 * F * b.Nv and abs(g["Fy"].sum()) < 1e-5 * F * b.Nv
    for x in bs.values():
        x.close()


def test_voronoi_fire_full_size_to_1e10():
    cg = _cg()
    N = 262144
    b = cg.VoronoiBatch([N], seeds=[13377], precision="fp64")
    b.set_prefs(1.0, 3.75)
    b.set_motility(0.0, 0.0)
    f = cg.EnergyMinimizerFIRE(b)
    f.setDeltaT(0.01)
    f.setDeltaTMax(0.1)
    f.setAlphaStart(0.1)
    f.setForceCutoff(1e-10)
    E0 = b.energy(0)
    it = 0
    while it < 100000:
        it += 10000
        f.setMaximumIterations(it)
        f.minimize()
        st = f.state()
        print("FIRE N=262144: %d iterations, RMS force %.3e" % (st["iterations"], st["force_max"] ** 0.5), flush=True)
        if st["force_max"] ** 0.5 < 1e-10:
            break
    b.check_status()
    b.compute_forces()
    g = b.get()
    rms = np.sqrt(np.mean(g["Fx"] ** 2 + g["Fy"] ** 2))
    mx = np.abs(np.concatenate([g["Fx"], g["Fy"]])).max()
    print("RMS-test stop: %d iterations, RMS %.3e, max |F| %.3e" % (st["iterations"], rms, mx), flush=True)
    assert st["iterations"] < 100000 and rms < 1e-10
    assert b.energy(0) < E0
    # on to the CPU branch's test: the largest component below 1e-10
    f.setForceCutoff(1e-12)
    while mx >= 1e-10 and it < 200000:
        it += 5000
        f.setMaximumIterations(it)
        f.minimize()
        b.compute_forces()
        g = b.get()
        mx = np.abs(np.concatenate([g["Fx"], g["Fy"]])).max()
        print("  %d iterations: max |F| %.3e" % (f.state()["iterations"], mx), flush=True)
    print("max-test stop: %d iterations, RMS %.3e, max |F| %.3e" % (
        f.state()["iterations"], np.sqrt(np.mean(g["Fx"] ** 2 + g["Fy"] ** 2)), mx), flush=True)
    assert mx < 1e-10
    nn, _, _ = b.topology(24)
    assert nn.sum() == 6 * N
    b.check_status()
    f.close()
    b.close()
