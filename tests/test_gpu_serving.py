"""Serving at system granularity: batch merges carry systems exactly, and a
request admitted into a running SPV pool (rebuilt at tick boundaries, which
drop finished systems) ends bit-identical to the same request served alone.
FIRE requests have no iteration cap and run to their cutoff."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CAP = 24


def _cg():
    import cellgpu_amd as cg
    cg.require_device()
    return cg


def _local_rings(b, s):
    nn, nbr, _ = b.topology(CAP)
    sl = b.system_slice(s)
    return [tuple(int(v) - sl.start for v in nbr[i, : nn[i]]) for i in range(sl.start, sl.stop)]


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_merge_continues_trajectories_exactly(precision):
    cg = _cg()
    a = cg.VoronoiBatch([300, 500], seeds=[3, 4], precision=precision)
    b = cg.VoronoiBatch([400], seeds=[8], precision=precision)
    for x, P0, v0 in ((a, 3.7, 0.1), (b, 3.85, 0.05)):
        x.set_prefs(1.0, P0)
        x.set_motility(v0, 1.0)
    a.set_active(np.ones(2, np.int32), np.array([-1, 37], np.int32))
    a.spv_steps(20, use_graph=True)
    b.spv_steps(7, use_graph=False)
    print("merging %s" % precision, flush=True)
    m = cg.VoronoiBatch.merge([(a, 1), (b, 0), (a, 0)])
    assert list(m.n) == [500, 400, 300]
    m.spv_steps(30, use_graph=True)
    a.spv_steps(30, use_graph=True)
    b.spv_steps(30, use_graph=False)
    for x in (a, b, m):
        x.check_status()
    gm, ga, gb = m.get(), a.get(), b.get()
    for ms, (src, g, s) in enumerate(((a, ga, 1), (b, gb, 0), (a, ga, 0))):
        sm, ss = m.system_slice(ms), src.system_slice(s)
        for k in ("x", "y", "theta", "vx", "vy"):
            assert np.array_equal(gm[k][sm], g[k][ss]), (ms, k)
        assert _local_rings(m, ms) == _local_rings(src, s)
        assert m.counters(ms)["step"] == src.counters(s)["step"]
    # the countdown travels with its system: system 1 of a stops after 37 steps in both
    assert m.counters(0)["step"] == 37 and a.counters(1)["step"] == 37
    for x in (a, b, m):
        x.close()


class _TickClock:
    """Deterministic clock: every reading advances by one millisecond."""

    def __init__(self):
        self.t = 0.0

    def __call__(self):
        self.t += 1e-3
        return self.t


def test_request_admitted_into_running_pool_equals_alone():
    _cg()
    from cellgpu_amd.serving import Request, ServingEngine

    def serve(reqs):
        eng = ServingEngine(admit_interval=0.0, tick_steps=16, precision="fp32", keep_results=True,
                            clock=_TickClock())
        done = {r.rid: r for r in eng.run(reqs, t0=0.0)}
        return eng, done

    def reqs():
        return {"B": Request(1, 400, 300, P0=3.7, v0=0.1, seed=5, arrival=0.0),
                "C": Request(2, 150, 40, P0=3.9, v0=0.2, seed=9, arrival=0.0),
                "A": Request(0, 65536, 100, P0=3.8, v0=0.1, Dr=1.0, seed=11, arrival=0.02)}

    r = reqs()
    print("mixed run", flush=True)
    eng, mixed = serve(list(r.values()))
    assert set(mixed) == {0, 1, 2}
    # A joined a pool that was running B, after C had finished
    assert mixed[0].admitted > mixed[1].admitted and mixed[1].done > mixed[0].admitted
    assert mixed[2].done <= mixed[0].admitted
    assert eng.rebuilds >= 2
    for key, rid in (("A", 0), ("B", 1)):
        print("alone run", key, flush=True)
        q = reqs()[key]
        q.arrival = 0.0
        _, alone = serve([q])
        xa, ya = alone[rid].result
        xm, ym = mixed[rid].result
        assert np.array_equal(xa, xm) and np.array_equal(ya, ym), key


def test_fire_request_runs_to_cutoff_without_iteration_cap():
    _cg()
    from cellgpu_amd.serving import Request, ServingEngine
    eng = ServingEngine(admit_interval=0.0, tick_iterations=256, keep_results=True)
    reqs = [Request(0, 1500, P0=3.7, seed=3, kind="vertex_fire", force_cutoff=1e-8),
            Request(1, 300, 500, P0=3.8, v0=0.1, seed=4)]
    done = {r.rid: r for r in eng.run(reqs)}
    f = done[0]
    assert f.converged and f.iterations > 0
    vx, vy = f.result
    assert np.isfinite(vx).all() and np.isfinite(vy).all()
    assert done[1].converged and done[1].iterations == 500
