"""Benchmark: batched self-propelled Voronoi ensemble on one MI355X.

Workload (the ensemble the headline metric is quoted on): 512 independent
periodic monolayers, N log-uniform in [64, 8192] at unit density, P0 ~ U[3.5, 4],
v0 ~ U[0.01, 0.2], Dr ~ U[0.1, 10], dt = 0.01, A0 = 1, KA = KP = 1; FP32 bulk
arithmetic with FP64 Delaunay predicates; one hipGraph-captured batched step
advances every system. A step = one time step of all 512 systems.

Protocol (the reference driver's, voronoi.cpp:95-126): --init-steps untimed
initialization steps from the random start (their rate is reported as
init_phase), the autotuner on the relaxed state (restores it), --warmup steps,
then exactly --steps timed steps.

Prints ONE JSON line. Besides value (aggregate cell-steps/s) it reports the
plain one-system-at-a-time FP64 GPU path on a sample of the same systems over
the same step window, FP32-vs-FP64 force / energy / topology / trajectory errors
at the relaxed state, per-kernel durations and algorithmic bandwidth (HIP
events on the batch stream) with the roofline of the dominant kernel, and the
oracle's restatement of the reference CPU path on one N = 1000 system, timed
on the host.

usage: python bench.py [--gpus 1] [--steps 2000] [--warmup 50] [--init-steps 500]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cell-steps/s batched vs plain FP64 on same GPU at 1e-5 force err; p50/p99 latency"
HBM_PEAK_GBS = 8000.0
# algorithmic (compulsory) bytes per cell for one launch, FP32 bulk, FP64 positions,
# mean Delaunay degree 6 (Euler): neighbour gathers are served on chip.
BYTES_FORCE_EOM_MOVE = 16 + 4 + 24 + 24 + 8 + 8 + 8 + 4 + 4 + 16 + 8 + 8 + 4  # = 136
BYTES_TEST_GEOMETRY = 16 + 4 + 24 + 4 + 8 + 8 + 8 + 4  # = 76 (the last 4: cell_rmax written)
BYTES_CELL_LIST = 16 + 4 + 4 + 8  # = 32


def ensemble(S=512, master_seed=20240607):
    rng = np.random.default_rng(master_seed)
    n = np.rint(np.exp(rng.uniform(np.log(64), np.log(8192), S))).astype(np.int32)
    P0 = rng.uniform(3.5, 4.0, S)
    v0 = rng.uniform(0.01, 0.2, S)
    Dr = rng.uniform(0.1, 10.0, S)
    seeds = rng.integers(1, 2**31 - 1, S).astype(np.uint64)
    return n, P0, v0, Dr, seeds


def per_cell(n, v):
    return np.repeat(np.asarray(v, np.float64), n)


def make_batch(cg, n, P0, v0, Dr, seeds, precision):
    b = cg.VoronoiBatch(n, seeds=seeds, precision=precision)
    b.set_params(KA=1.0, KP=1.0, dt=0.01, mu=1.0)
    # reference spatialSorting: relabel each system's cells along a Hilbert curve
    # (same random positions, neighbour gathers become local); per-cell
    # preferences are uniform within a system, so they are set after the sort
    b.spatial_sort()
    b.set_prefs(1.0, per_cell(n, P0))
    b.set_motility(per_cell(n, v0), per_cell(n, Dr))
    return b


def system_energies(g, n, P0):
    """E_s = sum over the system's cells of (A - 1)^2 + (P - P0)^2 (KA = KP = 1)."""
    e = (g["A"] - 1.0) ** 2 + (g["P"] - per_cell(n, P0)) ** 2
    return np.add.reduceat(e, np.concatenate([[0], np.cumsum(n)[:-1]]))


def _per_system_max(v, n):
    return np.maximum.reduceat(v, np.concatenate([[0], np.cumsum(n)[:-1]]))


def fp32_vs_fp64(cg, n, P0, v0, Dr, seeds, b32, steps=100):
    """FP32 bulk against FP64 at the FP32 batch's current (relaxed) state: an
    FP64 batch gets the same positions, directors, velocities, step counters and
    Delaunay rings, and both compute forces. Reports the force error per system
    (max over cells of |dF| relative to the system's largest |F|, and relative
    to its RMS |F|; worst system), the relative energy error, the position
    difference after `steps` further steps with the same noise, and how many of
    the maintained rings differ from the exact Delaunay triangulation of the
    same positions (the reference's test window, voronoiModelBase.cu:47-69, can
    leave an intruder unflagged: the FP64 path keeps the same rings)."""
    S = len(n)
    g = b32.get()
    st = np.array([b32.counters(s)["step"] for s in range(S)], np.int32)
    nn32, nbr32, _ = b32.topology()
    b64 = make_batch(cg, n, P0, v0, Dr, seeds, "fp64")
    b64.set_positions(g["x"], g["y"])  # exact triangulation of the same positions
    nnx, nbrx, _ = b64.topology()
    # slots past a cell's degree hold -1 in both copies
    not_delaunay = int(np.count_nonzero((nn32 != nnx) | np.any(nbr32 != nbrx, axis=1)))
    b64.set_topology(nn32, nbr32, b32.del_other())
    b64.set_directors(g["theta"])
    b64.set_velocities(g["vx"], g["vy"])
    b64.set_step_counters(st)
    b64.compute_forces()
    b32.compute_forces()
    g64, g32 = b64.get(), b32.get()
    dF = np.hypot(g32["Fx"] - g64["Fx"], g32["Fy"] - g64["Fy"])
    F = np.hypot(g64["Fx"], g64["Fy"])
    off = np.concatenate([[0], np.cumsum(n)[:-1]])
    err_max = _per_system_max(dF, n) / _per_system_max(F, n)
    rms = np.sqrt(np.add.reduceat(F * F, off) / n)
    err_rms = _per_system_max(dF, n) / rms
    E64, E32 = system_energies(g64, n, P0), system_energies(g32, n, P0)
    eerr = float(np.max(np.abs(E32 - E64) / np.abs(E64)))
    # trajectories from the same state with the same (counter-based) noise
    b32.spv_steps(steps, use_graph=True)
    b64.spv_steps(steps, use_graph=True)
    b32.check_status()
    b64.check_status()
    a, c = b32.get(), b64.get()
    L = np.repeat(np.sqrt(n.astype(np.float64)), n)
    dx = a["x"] - c["x"]
    dy = a["y"] - c["y"]
    dx -= L * np.round(dx / L)
    dy -= L * np.round(dy / L)
    d = np.hypot(dx, dy)
    b64.close()
    return {
        "force_rel_err_ensemble": float(dF.max() / F.max()),
        "force_rel_err_max_norm": float(err_max.max()),
        "force_rel_err_max_norm_worst_system": int(err_max.argmax()),
        "force_rel_err_vs_rms": float(err_rms.max()),
        "force_rel_err_vs_rms_worst_system": int(err_rms.argmax()),
        "force_rel_err_median_system": float(np.median(err_max)),
        "energy_rel_err": eerr,
        "cells_with_ring_not_exact_delaunay": not_delaunay,
        "trajectory": {"steps": steps, "max": float(d.max()), "rms": float(np.sqrt(np.mean(d * d))),
                       "unit": "mean cell spacing"},
    }


def plain_fp64(cg, n, P0, v0, Dr, seeds, untimed_steps, steps, sample_every=8):
    """Reference-structured plain path: one FP64 system at a time, host sync per
    step (the reference reads its "any circumcircle failed" flag back every
    step). Each sampled system first runs the same untimed steps as the batched
    run from the same initial condition, so both time the same window."""
    cells = 0
    t = 0.0
    for s in range(0, len(n), sample_every):
        b = make_batch(cg, n[s:s + 1], P0[s:s + 1], v0[s:s + 1], Dr[s:s + 1], seeds[s:s + 1], "fp64")
        b.spv_steps(untimed_steps, use_graph=True)
        b.sync()
        b.check_status()
        t0 = time.perf_counter()
        for _ in range(steps):
            b.spv_steps(1, use_graph=False)
            b.status()
        t += time.perf_counter() - t0
        cells += int(n[s]) * steps
        b.close()
    return cells / t, "every %dth system (%d systems), %d steps each after %d untimed steps, per-step host sync" % (
        sample_every, len(range(0, len(n), sample_every)), steps, untimed_steps)


def _cpu_info():
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count()
    return model, avail


def cpu_baseline(runs=5, steps=1000, N=1000):
    """The reference's CPU-runnable case on the oracle's restatement of its CPU path
    (local test-and-repair, CPU forces and SPP update): N = 1000, L = sqrt(N),
    reference initial condition mt19937(13377), A0 = 1, P0 = 3.8, KA = KP = 1,
    v0 = 0.05, Dr = 1, dt = 0.01, noise from a pre-generated seeded N(0,1)
    buffer; 1 warm-up step, then `steps` steps timed by std::chrono inside one C
    loop; median of `runs` runs; g++ -O3, one thread (the reference has no host
    threading)."""
    from oracle import oracle as O
    eta = np.random.default_rng(20240607).standard_normal((steps + 1, N))
    times = []
    for _ in range(runs):
        x, y, th, L = O.random_init(N, 13377)
        o = O.VoronoiSystem(x, y, th, L, A0=1.0, P0=3.8, v0=0.05, Dr=1.0, dt=0.01)
        o.spp_step(eta[0], gpu_path=False)
        times.append(o.spp_run(eta[1:], gpu_path=False))
        del o
    med = float(np.median(times))
    model, avail = _cpu_info()
    return dict(value=N * steps / med, unit="cell-steps/s", cores=1, kind="port",
                sample="N=1000, %d steps after 1 warm-up step from the mt19937(13377) start, P0=3.8, "
                       "v0=0.05, Dr=1, dt=0.01, oracle restatement of the reference CPU path (local repair), "
                       "g++ -O3, 1 thread, median of %d runs (%s s)" % (
                           steps, runs, ", ".join("%.2f" % x for x in times)),
                cpu_model=model, host_cpus_available=avail)


def serving(count=48, rate=8.0, steps_range=(1000, 100000), fire_fraction=0.25, fire_budget=100000):
    """Serving workload: Poisson arrivals; SPV requests with N log-uniform
    [100, 65536] and steps log-uniform over steps_range, and a fraction of
    vertex-model FIRE requests (N log-uniform [100, 65536], run until the RMS
    vertex force is below 1e-8 within a budget of fire_budget iterations),
    served by the continuous batching engine (SPV pool rebuilt at tick
    boundaries, FIRE batches in worker threads). Latency = completion - arrival
    over the requests that reached their target; FIRE requests that spent their
    budget are counted apart and add no cell-steps."""
    from cellgpu_amd.serving import ServingEngine, poisson_requests
    reqs = poisson_requests(count, rate, seed=20240607, steps_range=steps_range, fire_fraction=fire_fraction,
                            fire_max_iterations=fire_budget)
    t0 = time.perf_counter()

    def progress(r):
        print("[bench] serving: request %d (%s, n %d) done after %.2f s%s" % (
            r.rid, r.kind, r.n, r.latency, "" if r.converged else ", budget spent"), flush=True)

    eng = ServingEngine(admit_interval=0.05, tick_steps=64, precision="fp32", on_complete=progress)
    done = eng.run(reqs, t0=t0)
    wall = time.perf_counter() - t0
    ok = [r for r in done if r.converged]
    lat = np.array([r.latency for r in ok])
    fire = [r for r in done if r.kind == "vertex_fire"]
    spv = [r for r in done if r.kind == "spv"]
    return {"serving_p50_ms": round(1e3 * float(np.percentile(lat, 50)), 1),
            "serving_p99_ms": round(1e3 * float(np.percentile(lat, 99)), 1),
            "serving_spv_p50_ms": round(1e3 * float(np.percentile([r.latency for r in spv], 50)), 1) if spv else None,
            "serving_fire_p50_ms": round(1e3 * float(np.percentile([r.latency for r in fire], 50)), 1) if fire else None,
            "serving_cell_steps_per_s": round(eng.cell_steps / wall, 1),
            "serving_wall_s": round(wall, 2),
            "serving_pool_rebuilds": eng.rebuilds,
            "serving_fire_requests": len(fire),
            "serving_fire_converged": sum(bool(r.converged) for r in fire),
            "serving_fire_budget_spent": sum(not r.converged for r in fire),
            "serving_fire_max_iterations": max((r.iterations for r in fire), default=0),
            "serving_workload": "%d requests, Poisson %.1f/s; SPV: N log-uniform [100, 65536], steps log-uniform "
                                "[%d, %d], P0~U[3.5,4], v0~U[0,0.2], FP32 bulk; %.0f%% vertex-model FIRE: N "
                                "log-uniform [100, 65536], P0~U[3.5,4], until RMS force < 1e-8 within %d "
                                "iterations, FP64; one SPV pool rebuilt at tick boundaries, FIRE in worker threads; "
                                "latency and cell-steps (incl. FIRE cell-iterations) over requests that reached "
                                "their target" % (count, rate, steps_range[0], steps_range[1], 100 * fire_fraction,
                                                  fire_budget)}


def other_workload(cg, name, steps, warmup, use_graph=True, init_steps=0):
    """Secondary workloads (a single large system and its FIRE minimization, the
    vertex model, a larger ensemble), one JSON line each; the
    headline stays the ensemble."""
    if name == "single":
        # one N = 262 144 system (L = 512), P0 = 3.75, v0 = 0.1, Dr = 1, dt = 0.01
        N = 262144
        b = cg.VoronoiBatch([N], seeds=[13377], precision="fp32")
        b.set_params(KA=1.0, KP=1.0, dt=0.01, mu=1.0)
        b.spatial_sort()
        b.set_prefs(1.0, 3.75)
        b.set_motility(0.1, 1.0)
        b.spv_steps(warmup)
        b.sync()
        b.check_status()
        t0 = time.perf_counter()
        b.spv_steps(steps)
        b.sync()
        el = time.perf_counter() - t0
        b.check_status()
        return {"metric": "cell-steps/s, one N=262144 SPV system", "value": round(N * steps / el, 1),
                "unit": "cell-steps/s", "steps": steps, "warmup": warmup, "ms_per_step": round(1e3 * el / steps, 5),
                "dtype": "f32", "config": {"workload": "single SPV system N=262144 L=512 P0=3.75 v0=0.1 Dr=1 dt=0.01"}}
    if name == "fire":
        # FIRE on the same system shape (dtMax 0.1, alpha0 0.1), a bounded number of iterations
        N = 262144
        b = cg.VoronoiBatch([N], seeds=[13377], precision="fp64")
        b.spatial_sort()
        b.set_prefs(1.0, 3.75)
        f = cg.EnergyMinimizerFIRE(b)
        f.setDeltaT(0.01)
        f.setDeltaTMax(0.1)
        f.setAlphaStart(0.1)
        f.setForceCutoff(1e-10)
        f.setMaximumIterations(warmup)
        f.minimize()
        b.sync()
        it0 = f.state()["iterations"]
        f.setMaximumIterations(warmup + steps)
        t0 = time.perf_counter()
        f.minimize()
        b.sync()
        el = time.perf_counter() - t0
        st = f.state()
        its = st["iterations"] - it0
        return {"metric": "cell-iterations/s, FIRE on one N=262144 system (FP64)", "value": round(N * its / el, 1),
                "unit": "cell-iterations/s", "iterations": its, "ms_per_iteration": round(1e3 * el / max(its, 1), 5),
                "dtype": "f64", "rms_force": st["force_max"] ** 0.5,
                "config": {"workload": "FIRE N=262144 L=512 P0=3.75 dt0=0.01 dtMax=0.1 alpha0=0.1 cutoff 1e-10"}}
    if name == "vertex":
        # vertex model: Nc = 32 768 (Nv = 65 536), Brownian T = 1e-3, dt = 0.005, T1 threshold 0.04, P0 = 3.9
        Nc = 32768
        b = cg.VertexBatch([Nc], seeds=[21], precision="fp32")
        b.set_params(dt=0.005, T=1e-3, t1_threshold=0.04)
        b.set_prefs(1.0, 3.9)
        b.brownian_steps(warmup)
        b.check_status()
        t0 = time.perf_counter()
        b.brownian_steps(steps)
        b.check_status()
        el = time.perf_counter() - t0
        return {"metric": "cell-steps/s, vertex model Nc=32768 Brownian with T1", "value": round(Nc * steps / el, 1),
                "unit": "cell-steps/s", "steps": steps, "warmup": warmup, "ms_per_step": round(1e3 * el / steps, 5),
                "dtype": "f32", "t1_events": b.counters(0)["t1_events"],
                "config": {"workload": "vertex Nc=32768 Nv=65536 T=1e-3 dt=0.005 thr=0.04 P0=3.9"}}
    if name == "ensemble_large":
        # the same ensemble distribution with 8x the systems (about 6.7 million
        # cells, 2.35 GB of device state): beyond the 256 MB Infinity Cache;
        # the headline's protocol (untimed initialization steps, warm-up, timed
        # steps, per-kernel event pass)
        n, P0, v0, Dr, seeds = ensemble(S=4096, master_seed=20240608)
        N = int(n.sum())
        bins = int(sum(max(2, int(np.floor(np.sqrt(k) / 1.25)) + (int(np.floor(np.sqrt(k) / 1.25)) % 2)) ** 2
                       for k in n))
        b = make_batch(cg, n, P0, v0, Dr, seeds, "fp32")
        print("[bench] ensemble_large: %d systems, %d cells" % (len(n), N), flush=True)
        if use_graph:
            b.spv_steps(init_steps, use_graph=True)
        else:
            # eager launches (profiler runs, where each dispatch is slow): progress
            # per 100 steps
            for k in range(0, init_steps, 100):
                b.spv_steps(min(100, init_steps - k), use_graph=False)
                b.sync()
                print("[bench] initialization step %d" % min(k + 100, init_steps), flush=True)
        b.spv_steps(warmup, use_graph=use_graph)
        b.check_status()
        b.sync()
        t0 = time.perf_counter()
        b.spv_steps(steps, use_graph=use_graph)
        b.sync()
        el = time.perf_counter() - t0
        b.check_status()
        kt, queues = kernel_roofline(cg, b, N, bins, min(steps, 50))
        dev_bytes = int(cg.lib().cgpu_vor_device_bytes(b.h))
        return {"metric": "cell-steps/s, batched SPV ensemble of 4096 systems (device state beyond the Infinity Cache)",
                "value": round(N * steps / el, 1), "unit": "cell-steps/s", "steps": steps, "warmup": warmup,
                "init_steps": init_steps, "ms_per_step": round(1e3 * el / steps, 5), "dtype": "f32",
                "device_bytes": dev_bytes, "kernels": kt, "queues_per_step": queues,
                "config": {"workload": "4096 systems, N log-uniform [64,8192] (%d cells), unit density, "
                                       "P0~U[3.5,4], v0~U[0.01,0.2], Dr~U[0.1,10], dt=0.01" % N}}
    if name == "vertex_spp":
        # self-propelled cell-vertex dynamics (the reference Vertex.cpp active path):
        # Nc = 32 768, v0 = 0.05, Dr = 1, dt = 0.01, T1 threshold 0.04, P0 = 3.9
        Nc = 32768
        b = cg.VertexBatch([Nc], seeds=[21], precision="fp32")
        b.set_params(dt=0.01, t1_threshold=0.04)
        b.set_prefs(1.0, 3.9)
        b.set_motility(0.05, 1.0)
        b.spp_steps(warmup)
        b.check_status()
        t0 = time.perf_counter()
        b.spp_steps(steps)
        b.check_status()
        el = time.perf_counter() - t0
        return {"metric": "cell-steps/s, vertex model Nc=32768 self-propelled cell-vertex dynamics with T1",
                "value": round(Nc * steps / el, 1), "unit": "cell-steps/s", "steps": steps, "warmup": warmup,
                "ms_per_step": round(1e3 * el / steps, 5), "dtype": "f32", "t1_events": b.counters(0)["t1_events"],
                "config": {"workload": "vertex SPP Nc=32768 Nv=65536 v0=0.05 Dr=1 dt=0.01 thr=0.04 P0=3.9"}}
    if name == "vertex_fire":
        # FIRE on the vertices of the same tissue (FP64), a bounded number of iterations
        Nc = 32768
        b = cg.VertexBatch([Nc], seeds=[21], precision="fp64")
        b.set_params(dt=0.01, t1_threshold=0.04)
        b.set_prefs(1.0, 3.9)
        f = cg.EnergyMinimizerFIRE(b)
        f.setDeltaT(0.01)
        f.setDeltaTMax(0.1)
        f.setAlphaStart(0.1)
        f.setForceCutoff(1e-8)
        f.setMaximumIterations(warmup)
        f.minimize()
        it0 = f.state()["iterations"]
        f.setMaximumIterations(warmup + steps)
        t0 = time.perf_counter()
        f.minimize()
        b.check_status()
        el = time.perf_counter() - t0
        st = f.state()
        its = st["iterations"] - it0
        return {"metric": "cell-iterations/s, FIRE on a vertex model Nc=32768 (FP64)", "value": round(Nc * its / el, 1),
                "unit": "cell-iterations/s", "iterations": its, "ms_per_iteration": round(1e3 * el / max(its, 1), 5),
                "dtype": "f64", "rms_force": st["force_max"] ** 0.5,
                "config": {"workload": "vertex FIRE Nc=32768 Nv=65536 P0=3.9 dt0=0.01 dtMax=0.1 alpha0=0.1 cutoff 1e-8"}}
    raise SystemExit("unknown workload " + name)


def copy_probe(mib, reps=5):
    """Plain copy kernel of known size: the reference point that scales the
    FETCH_SIZE / WRITE_SIZE counters to HBM bytes."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_mib": mib, "reps": reps}), flush=True)


# Algorithmic (compulsory) bytes per unit of each kernel of one step (FP32
# bulk, FP64 positions, mean Delaunay degree 6, neighbour gathers on chip), and
# the unit: cells of the batch, bins, or the entries of one of the step's
# queues (flagged cells, NeedsFixing cells, cells of retriangulated systems).
KERNEL_BYTES = {
    # pos 16 + nn 4 + ring 24 + delOther 24 + dAP 8 + vel 8 + (v0, Dr) 8 + theta 4 + system 4;
    # writes pos 16, F 8, vel 8, theta 4; binning pass: bin and rank 8
    "k_force_eom_move": ("cells", BYTES_FORCE_EOM_MOVE + 8),
    "k_bin_scan": ("bins", 12),           # count read + clear, start write
    "k_bin_scatter": ("cells", 52),       # bin 4, rank r/w 8, start 4, pos 16; id 4 + pos 16 written
    "k_test_geometry": ("cells", BYTES_TEST_GEOMETRY),
    # one record (3 x 8) and cell_sys (4) per 64 cells; the stamps of the regions
    "k_ld_reduce": ("cells", 28 / 64),
    # cell_sys 4 + cell_bin 4 + stamp 4 + cell_rmax 4 (system state and rule stay in
    # cache); cells near a stamp read their rings and queue their triangles
    "k_select_triangles": ("cells", 16),
    # the triangle 12 and its corners' slots 12 and positions 48; the window's bin runs
    # and candidate positions are served on chip
    "k_test_queue": ("queued", 72),
    "k_mark_nf": ("flagged", 88),         # queue 4, nn 4, ring 24, 7 marks r/w + NeedsFixing entries
    "k_star_local_wave": ("needs_fixing", 72),  # pos 16, nn 4, ring 24 read; ring 24 + nn 4 written
    "k_delsets_local": ("needs_fixing", 56),       # queue 4, nn 4, ring 24, delOther 24 written
    "k_global_fallback": ("global", 4),
    "k_star_global_wave": ("global", 72),
    "k_global_stats": ("global", 4),
    "k_finalize": ("repaired", 48),       # queue 4, delOther 24, AP + dAP 16, mark 4
}


def kernel_roofline(cg, b, N, bins, steps):
    """Per-kernel durations of `steps` profiled steps (a HIP event after every
    launch on the batch stream) and each kernel's algorithmic bytes / duration."""
    ms, q = b.profiled_steps(steps)
    units = {"cells": N, "bins": bins, "systems": len(b.L), "flagged": q[0] / steps, "needs_fixing": q[1] / steps,
             "global": q[2] / steps, "repaired": (q[1] + q[2]) / steps, "queued": q[3] / steps}
    table = {}
    for k, tot in ms.items():
        unit, per = KERNEL_BYTES[k]
        avg = tot / steps
        nbytes = per * units[unit]
        table[k] = {"avg_ms": round(avg, 5), "unit": unit, "units_per_launch": round(units[unit], 1),
                    "bytes_per_unit": per, "achieved_gbs": round(nbytes / (avg * 1e-3) / 1e9, 1) if avg > 0 else None}
    return table, {"flagged": q[0] / steps, "needs_fixing": q[1] / steps, "global": q[2] / steps,
                   "test_queue": q[3] / steps, "stamp_regions": q[4] / steps}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--init-steps", type=int, default=500,
                    help="untimed initialization steps from the random start before the warm-up (the reference "
                         "driver's initSteps, voronoi.cpp:95-101); their rate is reported separately")
    ap.add_argument("--timed-stage-steps", type=int, default=200)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--launch-shape", default="",
                    help="fixed launch shape instead of autotuning, as field=value,... (fields of cgpu_launch_shape)")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch every step's kernels on the stream instead of replaying a hipGraph (profiler "
                         "runs: rocprofv3 faults inside hipGraphLaunch on this image); implies --no-autotune")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-errors", action="store_true",
                    help="skip the FP32-vs-FP64 force/energy/trajectory comparison (profiling runs)")
    ap.add_argument("--workload", default="ensemble", choices=["ensemble", "ensemble_large", "single", "fire", "vertex", "vertex_spp",
                                                        "vertex_fire"],
                    help="ensemble = the headline; the others print their own line")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (counter calibration) and exit")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: every system of the ensemble runs on a single MI355X "
              "(an ensemble would shard across GPUs by whole systems, one process per GPU)", file=sys.stderr)
        sys.exit(2)

    import cellgpu_amd as cg
    cg.require_device()
    if args.workload != "ensemble":
        print(json.dumps(other_workload(cg, args.workload, args.steps, args.warmup, not args.no_graph,
                                        args.init_steps)), flush=True)
        return
    n, P0, v0, Dr, seeds = ensemble()
    N = int(n.sum())
    bins = int(sum(max(2, int(np.floor(np.sqrt(k) / 1.25)) + (int(np.floor(np.sqrt(k) / 1.25)) % 2)) ** 2 for k in n))
    b = make_batch(cg, n, P0, v0, Dr, seeds, "fp32")
    print("[bench] ensemble: %d systems, %d cells" % (len(n), N), flush=True)

    # untimed initialization from the random start (its rate is reported)
    use_graph = not args.no_graph
    if args.no_graph:
        args.no_autotune = True
    b.sync()
    t0 = time.perf_counter()
    if use_graph:
        b.spv_steps(args.init_steps, use_graph=True)
    else:
        # eager launches (profiler runs, where each dispatch is slow): progress
        # per 100 steps
        for k in range(0, args.init_steps, 100):
            b.spv_steps(min(100, args.init_steps - k), use_graph=False)
            b.sync()
            print("[bench] initialization step %d" % min(k + 100, args.init_steps), flush=True)
    b.sync()
    t_init = time.perf_counter() - t0
    b.check_status()
    print("[bench] %d initialization steps in %.3f s" % (args.init_steps, t_init), flush=True)

    if args.launch_shape:
        b.set_launch_shape(**{k: int(v) for k, v in (kv.split("=") for kv in args.launch_shape.split(","))})
        args.no_autotune = True
    shape, cached, t_tune = b.launch_shape(), False, 0.0
    if not args.no_autotune:
        t0 = time.perf_counter()
        shape, cached = b.autotune(trial_steps=20)
        t_tune = time.perf_counter() - t0
        b.check_status()
        print("[bench] autotuned launch shape %s (%.2f s%s)" % (shape, t_tune, ", cached" if cached else ""),
              flush=True)

    b.spv_steps(args.warmup, use_graph=use_graph)
    b.sync()
    b.check_status()
    cg.lib().cgpu_device_synchronize()
    t0 = time.perf_counter()
    b.spv_steps(args.steps, use_graph=use_graph)
    b.sync()
    cg.lib().cgpu_device_synchronize()
    t1 = time.perf_counter()
    b.check_status()
    elapsed = t1 - t0
    value = N * args.steps / elapsed
    print("[bench] timed %d steps in %.3f s" % (args.steps, elapsed), flush=True)

    # per-kernel durations (HIP events on the batch stream, no graph), continuing
    # the same trajectory after the timed steps
    ks = max(1, min(args.timed_stage_steps, args.steps))
    kt, queues = kernel_roofline(cg, b, N, bins, ks)
    b.check_status()
    dom = max(kt, key=lambda k: kt[k]["avg_ms"])
    d = kt[dom]
    roofline = {"bound": "hbm", "kernel": dom, "achieved": d["achieved_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(d["achieved_gbs"] / HBM_PEAK_GBS, 4) if d["achieved_gbs"] else None,
                "traffic": None, "bytes_per_unit": d["bytes_per_unit"], "unit_of_work": d["unit"],
                "units_per_launch": d["units_per_launch"], "avg_launch_ms": d["avg_ms"],
                "timing": "HIP events on the batch stream over %d steps after the timed steps" % ks}
    # HBM bytes per launch of that kernel from the committed counter summary
    # (rocprofv3 FETCH_SIZE / WRITE_SIZE, scaled by a plain copy probe)
    tpath = os.path.join(ROOT, "profiles", "voronoi_spv_ensemble512_hbm_traffic.json")
    if os.path.exists(tpath):
        tk = json.load(open(tpath)).get("kernels", {})
        hit = [v for k, v in tk.items() if k.split("<")[0] == dom.split("<")[0] and
               ("<" not in dom or k.startswith(dom.split(">")[0]))]
        if hit:
            roofline["traffic"] = round(hit[0]["traffic_bytes"])
            roofline["traffic_source"] = "profiles/voronoi_spv_ensemble512_hbm_traffic.json"
            roofline["algorithmic_bytes_per_launch"] = round(d["bytes_per_unit"] * d["units_per_launch"])
    # the bound that limits the kernel: its share of the VALU issue (rocprofv3
    # SQ_INSTS_VALU x 4 cycles per wavefront instruction / SIMD-cycles of the
    # dispatch), from the committed counter summary
    ipath = os.path.join(ROOT, "profiles", "voronoi_spv_ensemble512_sq_issue.json")
    if os.path.exists(ipath):
        ik = json.load(open(ipath))
        hit = [v for k, v in ik.items() if k.split("<")[0] == dom.split("<")[0] and
               ("<" not in dom or k.startswith(dom.split(">")[0]))]
        if hit:
            roofline["valu_issue_frac"] = round(hit[0]["issue"], 3)
            roofline["valu_per_wavefront"] = round(hit[0]["valu_per_wave"])
            roofline["valu_issue_source"] = "profiles/voronoi_spv_ensemble512_sq_issue.json"

    errs = None
    if not args.no_errors:
        errs = fp32_vs_fp64(cg, n, P0, v0, Dr, seeds, b)
        print("[bench] fp32 vs fp64 at the relaxed state: %s" % json.dumps(errs), flush=True)
    b.close()

    plain_val, plain_desc = (None, None)
    if not args.no_plain:
        print("[bench] plain FP64 one-system-at-a-time leg", flush=True)
        plain_val, plain_desc = plain_fp64(cg, n, P0, v0, Dr, seeds, args.init_steps + args.warmup,
                                           min(args.steps, 200))
    lat = None
    if not args.no_serving:
        lat = serving()
        print("[bench] serving: p50 %.1f ms, p99 %.1f ms" % (lat["serving_p50_ms"], lat["serving_p99_ms"]), flush=True)
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline()

    per_step_ms = 1e3 * elapsed / args.steps
    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(per_step_ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "dtype_detail": "f32 geometry/forces/integrator, f64 positions, f64 Delaunay predicates and cell list",
        "data": "synthetic",
        "config": {
            "workload": "batched Voronoi SPV ensemble: 512 systems, N log-uniform [64,8192], unit density, "
                        "P0~U[3.5,4], v0~U[0.01,0.2], Dr~U[0.1,10], dt=0.01",
            "systems": int(len(n)), "cells": N, "precision": "fp32-bulk/fp64-predicates",
            "graph": use_graph,
            "init_steps": args.init_steps,
            "timed_window": "steps %d-%d after the random start" % (
                args.init_steps + args.warmup, args.init_steps + args.warmup + args.steps),
            "launch_shape": shape, "autotuned": not args.no_autotune,
        },
        "init_phase": {"steps": args.init_steps, "seconds": round(t_init, 4),
                       "cell_steps_per_s": round(N * args.init_steps / t_init, 1) if t_init > 0 else None,
                       "note": "from the random start (repair-heavy transient), untimed in value"},
        "autotune_s": round(t_tune, 3),
        "plain_fp64_value": None if plain_val is None else round(plain_val, 1),
        "plain_fp64_sample": plain_desc,
        "ratio_vs_plain_fp64": None if plain_val is None else round(value / plain_val, 2),
        "fp32_vs_fp64": errs,
        "kernels": kt,
        "queues_per_step": queues,
        "latency_ms": lat,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
