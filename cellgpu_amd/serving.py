"""Continuous batching of independent simulations on one GPU.

The reference runs one simulation per process. Here requests arrive over time:
self-propelled Voronoi runs (a system size, a number of steps, the model
parameters) and FIRE minimizations of vertex-model tissues (a size, a force
cutoff). The engine groups the requests of one kind that have arrived into a new
batch, admitted against an HBM budget. Every live batch advances on its own HIP
stream. SPV systems stop on the device when their step countdown
(`cgpu_vor_set_active`) reaches zero; FIRE batches run ticks of iterations and a
system is parked once its RMS force is below its cutoff (or its iteration budget
is spent). A request completes at the first poll after its system has stopped.

A request's result does not depend on the batch it lands in: systems are
independent, the noise is keyed by (seed, cell, step) and every FIRE reduction is
per system.
"""
import time

import numpy as np

from . import EnergyMinimizerFIRE, VertexBatch, VoronoiBatch
from ._lib import check, lib


class Request:
    """One simulation: n cells in a periodic box of side sqrt(n). kind "spv":
    `steps` self-propelled Voronoi steps at (P0, v0, Dr), noise seed `seed`. kind
    "vertex_fire": FIRE on the vertex model built from the seed's random cells,
    until the RMS vertex force is below force_cutoff or `steps` iterations."""

    __slots__ = ("rid", "kind", "n", "steps", "P0", "v0", "Dr", "seed", "force_cutoff", "arrival", "admitted",
                 "done", "result", "iterations")

    def __init__(self, rid, n, steps, P0=3.8, v0=0.05, Dr=1.0, seed=13377, arrival=0.0, kind="spv",
                 force_cutoff=1e-8):
        if kind not in ("spv", "vertex_fire"):
            raise ValueError("request kind is spv or vertex_fire")
        self.rid, self.kind, self.n, self.steps = int(rid), kind, int(n), int(steps)
        self.P0, self.v0, self.Dr, self.seed = float(P0), float(v0), float(Dr), int(seed)
        self.force_cutoff = float(force_cutoff)
        self.arrival = float(arrival)
        self.admitted = None
        self.done = None
        self.result = None
        self.iterations = None

    @property
    def latency(self):
        return None if self.done is None else self.done - self.arrival


class _Live:
    def __init__(self, batch, reqs, fire=None):
        self.batch, self.reqs, self.fire = batch, reqs, fire
        self.open = np.ones(len(reqs), bool)
        self.iters = 0
        if fire is None:
            self.bytes = int(lib().cgpu_vor_device_bytes(batch.h))
        else:  # FP64 vertex state, topology and FIRE buffers
            self.bytes = int(batch.Nv) * 400 + int(batch.Nc) * 400


class ServingEngine:
    """Admit, advance and retire batches of requests.

    hbm_budget: bytes of device memory the live batches may hold;
    admit_interval: seconds the oldest waiting request may wait for company;
    tick_steps: steps every live batch advances between polls;
    keep_results: copy each finished system's positions into request.result.
    """

    BYTES_PER_CELL_ESTIMATE = 600  # FP32 bulk, ring capacity 24, before the exact figure is known

    def __init__(self, hbm_budget=64 << 30, admit_interval=0.05, tick_steps=64, precision="fp32", dt=0.01,
                 max_batch_cells=1 << 22, keep_results=False, clock=time.perf_counter, tick_iterations=64):
        self.budget, self.admit_interval, self.tick_steps = int(hbm_budget), float(admit_interval), int(tick_steps)
        self.tick_iterations = int(tick_iterations)
        self.precision, self.dt, self.max_batch_cells = precision, float(dt), int(max_batch_cells)
        self.keep_results, self.clock = keep_results, clock
        self.pending, self.live, self.finished = [], [], []
        self.used = 0
        self.cell_steps = 0

    # ------------------------------------------------------------------ admission
    def submit(self, req):
        self.pending.append(req)

    def _admit(self, now):
        ready = [r for r in self.pending if r.arrival <= now]
        if not ready:
            return
        if now - min(r.arrival for r in ready) < self.admit_interval and self.live:
            return
        kind = min(ready, key=lambda q: q.arrival).kind  # the oldest request's kind goes first
        ready = [r for r in ready if r.kind == kind]
        group, cells = [], 0
        for r in sorted(ready, key=lambda q: q.arrival):
            need = (cells + r.n) * self.BYTES_PER_CELL_ESTIMATE
            if group and (cells + r.n > self.max_batch_cells or self.used + need > self.budget):
                break
            if not group and self.used + r.n * self.BYTES_PER_CELL_ESTIMATE > self.budget and self.live:
                return  # wait for memory to free up
            group.append(r)
            cells += r.n
        for r in group:
            self.pending.remove(r)
        ns = np.array([r.n for r in group], np.int32)
        if kind == "vertex_fire":
            self._admit_fire(group, ns)
            return
        b = VoronoiBatch(ns, seeds=np.array([r.seed for r in group], np.uint64), precision=self.precision)
        b.set_params(KA=1.0, KP=1.0, dt=self.dt, mu=1.0)
        b.spatial_sort()
        b.set_prefs(1.0, np.repeat([r.P0 for r in group], ns))
        b.set_motility(np.repeat([r.v0 for r in group], ns), np.repeat([r.Dr for r in group], ns))
        b.set_active(np.ones(len(group), np.int32), np.array([r.steps for r in group], np.int32))
        t = self.clock()
        for r in group:
            r.admitted = t
        live = _Live(b, group)
        self.used += live.bytes
        self.live.append(live)

    def _admit_fire(self, group, ns):
        """A vertex-model batch (FP64: the cutoffs are far below FP32 resolution)
        and its FIRE minimizer (dt0 0.01, dtMax 0.1, alpha0 0.1)."""
        b = VertexBatch(ns, seeds=np.array([r.seed for r in group], np.uint64), precision="fp64")
        b.set_params(dt=0.01, t1_threshold=0.04)
        b.set_prefs(1.0, np.repeat([r.P0 for r in group], ns))
        f = EnergyMinimizerFIRE(b)
        f.setDeltaT(0.01)
        f.setDeltaTMax(0.1)
        f.setAlphaStart(0.1)
        f.setForceCutoff(min(r.force_cutoff for r in group))
        t = self.clock()
        for r in group:
            r.admitted = t
        live = _Live(b, group, fire=f)
        self.used += live.bytes
        self.live.append(live)

    def _advance_fire(self, lv):
        """One tick of FIRE iterations; systems that reached their cutoff or
        iteration budget are parked (inactive) so later ticks skip them."""
        lv.iters += self.tick_iterations
        lv.fire.setMaximumIterations(lv.iters)
        lv.fire.minimize()
        t = self.clock()
        active = np.zeros(len(lv.reqs), np.int32)
        check(lib().cgpu_vm_get_active(lv.batch.h, active.ctypes.data), "vm_get_active")
        newly = []
        for k, r in enumerate(lv.reqs):
            if not lv.open[k]:
                continue
            st = lv.fire.state(k)
            if st["force_max"] ** 0.5 <= r.force_cutoff or st["iterations"] >= r.steps:
                newly.append(k)
                r.done = t
                r.iterations = st["iterations"]
                self.cell_steps += r.n * st["iterations"]
                if self.keep_results:
                    g = lv.batch.get()
                    lo = int(np.sum([q.n for q in lv.reqs[:k]]) * 2)
                    r.result = (g["vx"][lo:lo + 2 * r.n].copy(), g["vy"][lo:lo + 2 * r.n].copy())
                self.finished.append(r)
                lv.open[k] = False
                active[k] = 0
        if newly:
            check(lib().cgpu_vm_set_active(lv.batch.h, active.ctypes.data), "vm_set_active")

    # -------------------------------------------------------------------- advance
    def _advance(self):
        for lv in self.live:
            if lv.fire is None:
                lv.batch.spv_steps(self.tick_steps, use_graph=True)
        done_batches = []
        for lv in self.live:
            if lv.fire is not None:
                self._advance_fire(lv)
                if not lv.open.any():
                    done_batches.append(lv)
                continue
            active = np.zeros(len(lv.reqs), np.int32)
            check(lib().cgpu_vor_get_active(lv.batch.h, active.ctypes.data, None), "get_active")
            t = self.clock()
            newly = lv.open & (active == 0)
            if newly.any():
                g = lv.batch.get() if self.keep_results else None
                for k in np.nonzero(newly)[0]:
                    r = lv.reqs[k]
                    r.done = t
                    self.cell_steps += r.n * r.steps
                    if g is not None:
                        sl = lv.batch.system_slice(k)
                        r.result = (g["x"][sl].copy(), g["y"][sl].copy())
                    self.finished.append(r)
                lv.open &= ~newly
            if not lv.open.any():
                done_batches.append(lv)
        for lv in done_batches:
            lv.batch.check_status()
            if lv.fire is not None:
                lv.fire.close()
            lv.batch.close()
            self.used -= lv.bytes
            self.live.remove(lv)

    def run(self, requests, t0=None):
        """Serve `requests` (arrival times relative to the start) until all are
        done. Returns the finished requests with absolute done times."""
        t0 = self.clock() if t0 is None else t0
        for r in requests:
            r.arrival += t0
            self.submit(r)
        while self.pending or self.live:
            now = self.clock()
            self._admit(now)
            if self.live:
                self._advance()
            elif self.pending:
                nxt = min(r.arrival for r in self.pending)
                time.sleep(max(0.0, min(nxt - self.clock(), 0.01)))
        return self.finished


def poisson_requests(count, rate, seed=1, n_range=(100, 65536), steps_range=(1000, 100000), fire_fraction=0.0,
                     fire_n_range=(100, 4096), fire_max_iterations=2000, force_cutoff=1e-8):
    """Requests of the serving workload: Poisson arrivals at `rate` per second.
    SPV: N and steps log-uniform, P0 ~ U[3.5, 4], v0 ~ U[0, 0.2], Dr = 1. A
    fraction fire_fraction are vertex-model FIRE minimizations instead (N
    log-uniform over fire_n_range, P0 ~ U[3.5, 4], RMS force cutoff, at most
    fire_max_iterations iterations)."""
    rng = np.random.default_rng(seed)
    arr = np.cumsum(rng.exponential(1.0 / rate, count))
    n = np.rint(np.exp(rng.uniform(np.log(n_range[0]), np.log(n_range[1]), count))).astype(int)
    st = np.rint(np.exp(rng.uniform(np.log(steps_range[0]), np.log(steps_range[1]), count))).astype(int)
    P0 = rng.uniform(3.5, 4.0, count)
    v0 = rng.uniform(0.0, 0.2, count)
    seeds = rng.integers(1, 2**31 - 1, count)
    fire = rng.uniform(0, 1, count) < fire_fraction
    nf = np.rint(np.exp(rng.uniform(np.log(fire_n_range[0]), np.log(fire_n_range[1]), count))).astype(int)
    out = []
    for k in range(count):
        if fire[k]:
            out.append(Request(k, nf[k], fire_max_iterations, P0[k], 0.0, 1.0, seeds[k], arr[k], kind="vertex_fire",
                               force_cutoff=force_cutoff))
        else:
            out.append(Request(k, n[k], st[k], P0[k], v0[k], 1.0, seeds[k], arr[k]))
    return out
