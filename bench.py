"""Benchmark: batched self-propelled Voronoi ensemble on one MI355X.

Workload (the ensemble the headline metric is quoted on): 512 independent
periodic monolayers, N log-uniform in [64, 8192] at unit density, P0 ~ U[3.5, 4],
v0 ~ U[0.01, 0.2], Dr ~ U[0.1, 10], dt = 0.01, A0 = 1, KA = KP = 1; FP32 bulk
arithmetic with FP64 Delaunay predicates; one hipGraph-captured batched step
advances every system. A step = one time step of all 512 systems.

Prints ONE JSON line. Besides value (aggregate cell-steps/s) it reports the
plain one-system-at-a-time FP64 GPU path on a sample of the same systems, the
FP32-vs-FP64 force error, the roofline of the dominant kernel (HIP events on the
batch stream) and the oracle's CPU path timed on the host (bounded sample).

usage: python bench.py [--gpus 1] [--steps 2000] [--warmup 50]
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
BYTES_TEST_GEOMETRY = 16 + 4 + 24 + 4 + 8 + 8 + 8  # = 72
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


def fp32_errors(cg, n, P0, v0, Dr, seeds, b32, steps=100):
    """FP32-bulk vs FP64 at the same state (max-norm force error relative to the
    largest force; largest relative system energy error), and after `steps`
    steps from the same start with the same noise: the largest and RMS
    minimum-image position difference (unit density: in mean cell spacings)."""
    b64 = make_batch(cg, n, P0, v0, Dr, seeds, "fp64")
    b64.compute_forces()
    b32.compute_forces()
    g64, g32 = b64.get(), b32.get()
    F64 = np.concatenate([g64["Fx"], g64["Fy"]])
    F32 = np.concatenate([g32["Fx"], g32["Fy"]])
    ferr = float(np.abs(F32 - F64).max() / np.abs(F64).max())
    E64, E32 = system_energies(g64, n, P0), system_energies(g32, n, P0)
    eerr = float(np.max(np.abs(E32 - E64) / np.abs(E64)))
    t32 = make_batch(cg, n, P0, v0, Dr, seeds, "fp32")
    t32.spv_steps(steps, use_graph=True)
    b64.spv_steps(steps, use_graph=True)
    t32.check_status()
    b64.check_status()
    a, c = t32.get(), b64.get()
    L = np.repeat(np.sqrt(n.astype(np.float64)), n)
    dx = a["x"] - c["x"]
    dy = a["y"] - c["y"]
    dx -= L * np.round(dx / L)
    dy -= L * np.round(dy / L)
    d = np.hypot(dx, dy)
    t32.close()
    b64.close()
    return ferr, eerr, {"steps": steps, "max": float(d.max()), "rms": float(np.sqrt(np.mean(d * d))),
                        "unit": "mean cell spacing"}


def plain_fp64(cg, n, P0, v0, Dr, seeds, sample_every=8, steps=20):
    """Reference-structured plain path: one FP64 system at a time, host sync per step."""
    cells = 0
    t = 0.0
    for s in range(0, len(n), sample_every):
        b = make_batch(cg, n[s:s + 1], P0[s:s + 1], v0[s:s + 1], Dr[s:s + 1], seeds[s:s + 1], "fp64")
        b.spv_steps(2, use_graph=False)
        b.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            b.spv_steps(1, use_graph=False)
            b.status()  # the reference reads the "any circumcircle failed" flag back every step
        t += time.perf_counter() - t0
        cells += int(n[s]) * steps
        b.close()
    return cells / t, "every %dth system (%d systems), %d steps each, per-step host sync" % (
        sample_every, len(range(0, len(n), sample_every)), steps)


def cpu_baseline(n, P0, v0, Dr, seeds, budget_s=12.0, steps=10):
    """Oracle restatement of the reference CPU path (local test-and-repair), 1 thread."""
    from oracle import oracle as O
    cells = 0
    t = 0.0
    used = 0
    rng = np.random.default_rng(7)
    for s in range(len(n)):
        if t > budget_s:
            break
        x, y, th, L = O.random_init(int(n[s]), int(seeds[s]))
        o = O.VoronoiSystem(x, y, th, L, P0=P0[s], v0=v0[s], Dr=Dr[s], dt=0.01)
        etas = rng.standard_normal((steps, int(n[s])))
        t0 = time.perf_counter()
        for k in range(steps):
            o.spp_step(etas[k], gpu_path=False)
        t += time.perf_counter() - t0
        cells += int(n[s]) * steps
        used += 1
    return dict(value=cells / t, unit="cell-steps/s", cores=1, kind="port",
                sample="first %d systems of the ensemble, %d steps each, oracle CPU path (local repair), 1 thread"
                       % (used, steps))


def serving(count=48, rate=8.0, steps_range=(1000, 10000), fire_fraction=0.25):
    """Serving workload (bounded): Poisson arrivals; SPV requests with N
    log-uniform [100, 65536] and steps log-uniform over steps_range, and a
    fraction of vertex-model FIRE requests (N log-uniform [100, 4096], RMS force
    below 1e-8 within 2000 iterations), admitted into batches by the continuous
    batching engine; latency = completion - arrival."""
    from cellgpu_amd.serving import ServingEngine, poisson_requests
    reqs = poisson_requests(count, rate, seed=20240607, steps_range=steps_range, fire_fraction=fire_fraction)
    eng = ServingEngine(admit_interval=0.05, tick_steps=64, precision="fp32")
    t0 = time.perf_counter()
    done = eng.run(reqs, t0=t0)
    wall = time.perf_counter() - t0
    lat = np.array([r.latency for r in done])
    fire = [r for r in done if r.kind == "vertex_fire"]
    return {"serving_p50_ms": round(1e3 * float(np.percentile(lat, 50)), 1),
            "serving_p99_ms": round(1e3 * float(np.percentile(lat, 99)), 1),
            "serving_cell_steps_per_s": round(eng.cell_steps / wall, 1),
            "serving_fire_requests": len(fire),
            "serving_fire_converged": sum(r.iterations < r.steps for r in fire),
            "serving_workload": "%d requests, Poisson %.1f/s; SPV: N log-uniform [100, 65536], steps log-uniform "
                                "[%d, %d], P0~U[3.5,4], v0~U[0,0.2], FP32 bulk; %.0f%% vertex-model FIRE: N "
                                "log-uniform [100, 4096], P0~U[3.5,4], RMS force < 1e-8 within 2000 iterations, "
                                "FP64; continuous batching (cell-steps include FIRE cell-iterations)" % (
                                    count, rate, steps_range[0], steps_range[1], 100 * fire_fraction)}


def other_workload(cg, name, steps, warmup):
    """Secondary workloads (survey cfg2 / cfg3), one JSON line each; the
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
        # cells, several GB of device state): beyond the 256 MB Infinity Cache
        n, P0, v0, Dr, seeds = ensemble(S=4096, master_seed=20240608)
        N = int(n.sum())
        b = make_batch(cg, n, P0, v0, Dr, seeds, "fp32")
        b.spv_steps(warmup)
        b.check_status()
        t0 = time.perf_counter()
        b.spv_steps(steps)
        b.sync()
        el = time.perf_counter() - t0
        b.check_status()
        import ctypes as C
        ks = 50
        ms = np.zeros(4)
        cg.check(cg.lib().cgpu_vor_spv_steps_timed(b.h, ks, ms.ctypes.data_as(C.c_void_p)), "timed steps")
        avg = ms / ks
        dev_bytes = int(cg.lib().cgpu_vor_device_bytes(b.h))
        return {"metric": "cell-steps/s, batched SPV ensemble of 4096 systems (device state beyond the Infinity Cache)",
                "value": round(N * steps / el, 1), "unit": "cell-steps/s", "steps": steps, "warmup": warmup,
                "ms_per_step": round(1e3 * el / steps, 5), "dtype": "f32", "device_bytes": dev_bytes,
                "stage_ms": {"force_eom_move": round(avg[0], 5), "cell_list": round(avg[1], 5),
                             "test_geometry": round(avg[2], 5), "repair_chain": round(avg[3], 5)},
                "test_geometry_achieved_gbs": round(BYTES_TEST_GEOMETRY * N / (avg[2] * 1e-3) / 1e9, 1),
                "force_eom_move_achieved_gbs": round(BYTES_FORCE_EOM_MOVE * N / (avg[0] * 1e-3) / 1e9, 1),
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--timed-stage-steps", type=int, default=200)
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
              "(multi-GPU sharding of an ensemble is by system; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import cellgpu_amd as cg
    cg.require_device()
    if args.workload != "ensemble":
        print(json.dumps(other_workload(cg, args.workload, args.steps, args.warmup)), flush=True)
        return
    n, P0, v0, Dr, seeds = ensemble()
    N = int(n.sum())
    b = make_batch(cg, n, P0, v0, Dr, seeds, "fp32")
    ferr = eerr = traj = None
    if not args.no_errors:
        ferr, eerr, traj = fp32_errors(cg, n, P0, v0, Dr, seeds, b)
    print("[bench] ensemble: %d systems, %d cells; fp32-vs-fp64 force error %s" % (
        len(n), N, "skipped" if ferr is None else "%.2e" % ferr), flush=True)

    b.spv_steps(args.warmup, use_graph=True)
    b.sync()
    b.check_status()
    cg.lib().cgpu_device_synchronize()
    t0 = time.perf_counter()
    b.spv_steps(args.steps, use_graph=True)
    b.sync()
    cg.lib().cgpu_device_synchronize()
    t1 = time.perf_counter()
    b.check_status()
    elapsed = t1 - t0
    value = N * args.steps / elapsed
    print("[bench] timed %d steps in %.3f s" % (args.steps, elapsed), flush=True)

    # per-stage durations with HIP events on the batch stream (no graph)
    ks = max(1, min(args.timed_stage_steps, args.steps))
    ms = np.zeros(4)
    import ctypes as C
    r = cg.lib().cgpu_vor_spv_steps_timed(b.h, ks, ms.ctypes.data_as(C.c_void_p))
    cg.check(r, "timed steps")
    b.check_status()
    avg_ms = ms / ks
    stage = {"force_eom_move": avg_ms[0], "cell_list": avg_ms[1], "test_geometry": avg_ms[2], "repair_chain": avg_ms[3]}
    bytes_map = {"force_eom_move": BYTES_FORCE_EOM_MOVE, "test_geometry": BYTES_TEST_GEOMETRY}
    dom = max(bytes_map, key=lambda k: stage[k])
    achieved = bytes_map[dom] * N / (stage[dom] * 1e-3) / 1e9
    roofline = {"bound": "hbm", "kernel": {"force_eom_move": "k_force_eom_move<float, 0, false>",
                                           "test_geometry": "k_test_geometry<float>"}[dom],
                "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "bytes_per_cell": bytes_map[dom], "cells_per_launch": N,
                "avg_launch_ms": round(stage[dom], 5)}
    # HBM bytes per launch of the dominant kernel from the committed counter
    # summary (rocprofv3 FETCH_SIZE / WRITE_SIZE, scaled by a plain copy probe)
    tpath = os.path.join(ROOT, "profiles", "voronoi_spv_ensemble512_hbm_traffic.json")
    if os.path.exists(tpath):
        tk = json.load(open(tpath)).get("kernels", {}).get(roofline["kernel"])
        if tk:
            roofline["traffic"] = round(tk["traffic_bytes"])
            roofline["traffic_source"] = "profiles/voronoi_spv_ensemble512_hbm_traffic.json"
            roofline["algorithmic_bytes_per_launch"] = bytes_map[dom] * N
    others = {k: round(bytes_map[k] * N / (stage[k] * 1e-3) / 1e9, 1) for k in bytes_map}

    plain_val, plain_desc = (None, None)
    if not args.no_plain:
        plain_val, plain_desc = plain_fp64(cg, n, P0, v0, Dr, seeds)
    b.close()
    lat = None
    if not args.no_serving:
        lat = serving()
        print("[bench] serving: p50 %.1f ms, p99 %.1f ms" % (lat["serving_p50_ms"], lat["serving_p99_ms"]), flush=True)
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(n, P0, v0, Dr, seeds)

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
            "graph": True,
        },
        "plain_fp64_value": None if plain_val is None else round(plain_val, 1),
        "plain_fp64_sample": plain_desc,
        "ratio_vs_plain_fp64": None if plain_val is None else round(value / plain_val, 2),
        "force_rel_err_vs_fp64": ferr,
        "energy_rel_err_vs_fp64": eerr,
        "trajectory_err_vs_fp64": traj,
        "stage_ms": {k: round(v, 5) for k, v in stage.items()},
        "stage_achieved_gbs": others,
        "latency_ms": lat,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
