"""Failure and special paths, each forced from the host with valid inputs:

* exclusions (reference VoronoiQuadraticEnergy::setExclusions,
  voronoiQuadraticEnergy.cu:59-93): excluded cells get F = 0 and an external
  force equal to minus their summed force sets, other cells keep their forces
  bit for bit; graph steps captured before or after set_exclusions equal
  uncaptured steps, and excluded cells stop from the second step on (their
  stored velocity carries one step, the N1 lag; the box wrap's rounding aside);
* more than N/6 flagged cells (voronoiModelBase.cpp:600-663): enforceTopology
  decides a global retriangulation, the oracle decides the same, and the result
  is the exact Delaunay triangulation;
* a displacement that leaves the representable range: the system reports
  CGPU_ERR_NONFINITE and keeps its last position, the other systems of the batch
  carry on unaffected;
* ring and cell-list capacities: test_gpu_capacity.py (growth, and refusal
  beyond the largest ring capacity).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CAP = 24


def _cg():
    import cellgpu_amd as cg
    cg.require_device()
    return cg


def _rings(nn, nbr):
    return [tuple(nbr[i, : nn[i]]) for i in range(len(nn))]


def test_exclusions_forces_and_graph_steps():
    cg = _cg()
    N = 300
    ex = np.zeros(N, np.int32)
    ex[::7] = 1
    b = cg.VoronoiBatch([N], seeds=[7], precision="fp64")
    b.set_prefs(1.0, 3.75)
    b.compute_forces()
    g0 = b.get()
    fx0, fy0 = b.external_forces()
    assert not fx0.any() and not fy0.any()
    b.set_exclusions(ex)
    b.compute_forces()
    g1 = b.get()
    ext_x, ext_y = b.external_forces()
    e = ex.astype(bool)
    assert not g1["Fx"][e].any() and not g1["Fy"][e].any()
    assert np.array_equal(ext_x[e], -g0["Fx"][e]) and np.array_equal(ext_y[e], -g0["Fy"][e])
    assert not ext_x[~e].any() and not ext_y[~e].any()
    assert np.array_equal(g1["Fx"][~e], g0["Fx"][~e]) and np.array_equal(g1["Fy"][~e], g0["Fy"][~e])
    b.close()

    # exclusions set before the first graph capture, and after one
    for when in ("before", "after"):
        runs = []
        for graph in (True, False):
            x = cg.VoronoiBatch([N], seeds=[7], precision="fp64")
            x.set_prefs(1.0, 3.75)
            x.set_motility(0.1, 1.0)
            if when == "before":
                x.set_exclusions(ex)
            else:
                x.spv_steps(4, use_graph=graph)
                x.set_exclusions(ex)
            x.spv_steps(1, use_graph=False)  # the stored velocity moves excluded cells once more
            p1 = x.get()
            x.spv_steps(10, use_graph=graph)
            x.check_status()
            g = x.get()
            # zero displacement: only the box wrap's rounding (x -> L (x / L), as the
            # reference's putInBoxReal) touches an excluded cell
            L = float(x.L[0])
            assert np.abs(g["x"][e] - p1["x"][e]).max() < 1e-13 * L, when
            assert np.abs(g["y"][e] - p1["y"][e]).max() < 1e-13 * L, when
            assert np.abs(g["x"][~e] - p1["x"][~e]).max() > 1e-4
            runs.append(g)
            x.close()
        for k in ("x", "y", "theta", "vx", "vy"):
            assert np.array_equal(runs[0][k], runs[1][k]), (when, k)


def test_more_than_n_over_6_flags_decide_a_global_retriangulation(oracle):
    cg = _cg()
    N = 400
    b = cg.VoronoiBatch([N], seeds=[5], precision="fp64")
    g = b.get()
    L = float(b.L[0])
    o = oracle.VoronoiSystem(g["x"], g["y"], g["theta"], L)
    rng = np.random.default_rng(3)
    x = np.mod(g["x"] + rng.uniform(-0.35, 0.35, N), L)
    y = np.mod(g["y"] + rng.uniform(-0.35, 0.35, N), L)
    b.set_positions(x, y, retriangulate=False)
    g2 = b.get()
    before = b.counters()["global_fixes"]
    b.enforce_topology()
    b.check_status()
    c = b.counters()
    assert c["global_fixes"] == before + 1 and c["last_mode"] == 2 and c["last_nf"] > N // 6
    o.set_state(g2["x"], g2["y"])
    # NeedsFixing = flagged cells and their ring neighbours, on the old rings
    _, flags = o.test_flags()
    onn, onbr, _ = o.topology(CAP)
    nf = set(np.nonzero(flags)[0].tolist())
    for i in list(nf):
        nf.update(int(v) for v in onbr[i, : onn[i]])
    # the same decision: the GPU stops testing a system once more than N/6 of its
    # cells are flagged (the outcome cannot change), so its count may be smaller
    assert len(nf) > N // 6 and N // 6 < c["last_nf"] <= len(nf)
    o.enforce_topology(gpu_path=True)
    _, mode = o.last_needs_fixing()
    assert mode == 2
    nn, nbr, _ = b.topology(CAP)
    onn, onbr, _ = o.topology(CAP)
    tnn, tnbr, _ = oracle.triangulate(g2["x"], g2["y"], L, cap=CAP)
    assert _rings(nn, nbr) == _rings(onn, onbr) == _rings(tnn, tnbr)
    b.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_runaway_system_reports_nonfinite_and_others_carry_on(precision):
    cg = _cg()
    solo = cg.VoronoiBatch([200], seeds=[1], precision=precision)
    pair = cg.VoronoiBatch([200, 300], seeds=[1, 2], precision=precision)
    for x in (solo, pair):
        x.set_prefs(1.0, 3.8)
        x.set_motility(0.1, 1.0)
    pair.set_motility(np.concatenate([np.full(200, 0.1), np.full(300, 1e300)]), 1.0)
    before = pair.get()
    solo.spv_steps(6, use_graph=False)
    pair.spv_steps(6, use_graph=False)
    with pytest.raises(cg.CellGPUError, match="non-finite"):
        pair.check_status()
    assert pair.counters(1)["error"] == -6 and pair.counters(0)["error"] == 0
    g, gs = pair.get(), solo.get()
    sl = pair.system_slice(1)
    # the kept position (box wrap rounding aside)
    L1 = float(pair.L[1])
    assert np.abs(g["x"][sl] - before["x"][sl]).max() < 1e-13 * L1
    assert np.abs(g["y"][sl] - before["y"][sl]).max() < 1e-13 * L1
    assert np.array_equal(g["x"][:200], gs["x"]) and np.array_equal(g["y"][:200], gs["y"])
    solo.check_status()
    solo.close()
    pair.close()
