"""Continuous batching of independent simulations on one GPU.

The reference runs one simulation per process. Here requests arrive over time:
self-propelled Voronoi runs (a system size, a number of steps, the model
parameters) and FIRE minimizations of vertex-model tissues (a size, a force
cutoff).

SPV requests share one pool batch that advances in ticks of `tick_steps` steps
(two-step hipGraph). A system stops on the device when its step countdown
(`cgpu_vor_set_active`) reaches zero and its request completes at the next
poll. At a tick boundary the pool is rebuilt with `VoronoiBatch.merge`: the
running systems are carried over exactly (positions, stored velocities,
directors, rings, step counters, remaining steps), finished systems are dropped
and their memory released, and the requests that arrived are admitted into the
freed room (against an HBM budget), after the same initialisation (seeded
random start, spatial sort) they get when served alone.

Vertex-model FIRE requests are grouped into batches that run in worker threads
(ctypes releases the GIL, and each batch has its own non-blocking HIP stream),
so a minimization never holds up the SPV ticks. A FIRE system is parked once its
RMS force is below its cutoff; a request that exhausts its iteration budget
instead completes with `converged = False`.

A request's result does not depend on the batch it lands in or on when the pool
is rebuilt: systems are independent, the noise is keyed by (seed, cell, step),
every FIRE reduction is per system and the merge copies state bit for bit.
"""
import threading
import time

import numpy as np

from . import EnergyMinimizerFIRE, VertexBatch, VoronoiBatch
from ._lib import check, lib

UNLIMITED_ITERATIONS = 10 ** 8


class Request:
    """One simulation: n cells in a periodic box of side sqrt(n). kind "spv":
    `steps` self-propelled Voronoi steps at (P0, v0, Dr), noise seed `seed`. kind
    "vertex_fire": FIRE on the vertex model built from the seed's random cells,
    until the RMS vertex force is below force_cutoff; `steps` (default: no limit)
    bounds the iterations."""

    __slots__ = ("rid", "kind", "n", "steps", "P0", "v0", "Dr", "seed", "force_cutoff", "arrival", "admitted",
                 "done", "result", "iterations", "converged")

    def __init__(self, rid, n, steps=None, P0=3.8, v0=0.05, Dr=1.0, seed=13377, arrival=0.0, kind="spv",
                 force_cutoff=1e-8):
        if kind not in ("spv", "vertex_fire"):
            raise ValueError("request kind is spv or vertex_fire")
        if steps is None:
            if kind == "spv":
                raise ValueError("an SPV request needs a number of steps")
            steps = UNLIMITED_ITERATIONS
        self.rid, self.kind, self.n, self.steps = int(rid), kind, int(n), int(steps)
        self.P0, self.v0, self.Dr, self.seed = float(P0), float(v0), float(Dr), int(seed)
        self.force_cutoff = float(force_cutoff)
        self.arrival = float(arrival)
        self.admitted = None
        self.done = None
        self.result = None
        self.iterations = None
        self.converged = None

    @property
    def latency(self):
        return None if self.done is None else self.done - self.arrival


class _Pool:
    """The SPV pool: a batch and the request of each of its systems (None once
    the request completed)."""

    def __init__(self, batch, reqs):
        self.batch, self.reqs = batch, reqs
        self.bytes = int(lib().cgpu_vor_device_bytes(batch.h))
        self.captured = False  # the first multi-step tick captures the step graph

    def live_cells(self):
        return sum(r.n for r in self.reqs if r is not None)

    def dead_cells(self):
        return sum(int(self.batch.n[k]) for k, r in enumerate(self.reqs) if r is None)


class _FireJob(threading.Thread):
    """One FIRE batch in a worker thread: ticks of `tick_iterations` iterations,
    then every open system is checked against its own cutoff and budget."""

    def __init__(self, engine, group):
        super().__init__(daemon=True)
        self.engine, self.group = engine, group
        self.error = None
        ns = np.array([r.n for r in group], np.int32)
        self.bytes = int(ns.sum()) * 2 * 400 + int(ns.sum()) * 400  # FP64 vertices, topology, FIRE buffers
        self.ns = ns

    def run(self):
        try:
            self._run()
        except Exception as e:  # reported by the engine's main loop
            self.error = e

    def _run(self):
        # every HIP call of this thread runs under the engine's lock: a hipGraph
        # capture on the SPV stream fails when another thread makes synchronous
        # HIP calls meanwhile, so the pool captures only while holding it
        eng, group, ns = self.engine, self.group, self.ns
        with eng._gpu_lock:
            b = VertexBatch(ns, seeds=np.array([r.seed for r in group], np.uint64), precision="fp64")
        f = None
        try:
            with eng._gpu_lock:
                b.set_params(dt=0.01, t1_threshold=0.04)
                b.set_prefs(1.0, np.repeat([r.P0 for r in group], ns))
                f = EnergyMinimizerFIRE(b)
                f.setDeltaT(0.01)
                f.setDeltaTMax(0.1)
                f.setAlphaStart(0.1)
                f.setForceCutoff(min(r.force_cutoff for r in group))
            open_ = np.ones(len(group), bool)
            iters = 0
            while open_.any():
                while eng._capture_pending.is_set():  # let the pool's graph capture in first
                    time.sleep(0.001)
                with eng._gpu_lock:
                    self._tick(group, ns, open_, iters + eng.tick_iterations, b, f)
                iters += eng.tick_iterations
            with eng._gpu_lock:
                b.check_status()
        finally:
            with eng._gpu_lock:
                if f is not None:
                    f.close()
                b.close()

    def _tick(self, group, ns, open_, iters, b, f):
        """FIRE up to `iters` iterations, then retire the systems that reached
        their cutoff or budget (parked: later ticks skip them)."""
        eng = self.engine
        f.setMaximumIterations(iters)
        f.minimize()
        t = eng.clock()
        active = np.zeros(len(group), np.int32)
        check(lib().cgpu_vm_get_active(b.h, active.ctypes.data), "vm_get_active")
        newly = []
        for k, r in enumerate(group):
            if not open_[k]:
                continue
            st = f.state(k)
            conv = st["force_max"] ** 0.5 <= r.force_cutoff
            if conv or st["iterations"] >= r.steps:
                newly.append(k)
                r.iterations, r.converged = int(st["iterations"]), bool(conv)
                if eng.keep_results:
                    g = b.get()
                    lo = 2 * int(ns[:k].sum())
                    r.result = (g["vx"][lo:lo + 2 * r.n].copy(), g["vy"][lo:lo + 2 * r.n].copy())
                open_[k] = False
                active[k] = 0
                # work that never reached its target is not counted as served
                eng._complete(r, t, r.n * r.iterations if conv else 0)
        if newly:
            check(lib().cgpu_vm_set_active(b.h, active.ctypes.data), "vm_set_active")


class ServingEngine:
    """Admit, advance and retire requests.

    hbm_budget: bytes of device memory the pool and the FIRE batches may hold;
    admit_interval: seconds the oldest waiting request may wait for company once
    a pool is running; tick_steps: SPV steps between polls; tick_iterations: FIRE
    iterations between checks; reclaim_fraction: rebuild the pool when finished
    systems hold more than this fraction of its cells; keep_results: copy each
    finished system's positions into request.result; on_complete: called with
    each request as it completes (from the thread that completed it).
    """

    BYTES_PER_CELL_ESTIMATE = 600  # FP32 bulk, ring capacity 24, before the exact figure is known

    def __init__(self, hbm_budget=64 << 30, admit_interval=0.05, tick_steps=64, precision="fp32", dt=0.01,
                 max_batch_cells=1 << 22, keep_results=False, clock=time.perf_counter, tick_iterations=64,
                 reclaim_fraction=0.25, on_complete=None):
        self.budget, self.admit_interval, self.tick_steps = int(hbm_budget), float(admit_interval), int(tick_steps)
        self.tick_iterations = int(tick_iterations)
        self.precision, self.dt, self.max_batch_cells = precision, float(dt), int(max_batch_cells)
        self.keep_results, self.clock = keep_results, clock
        self.reclaim_fraction = float(reclaim_fraction)
        self.pending, self.finished = [], []
        self.pool = None
        self.jobs = []
        self.cell_steps = 0
        self.rebuilds = 0
        self._lock = threading.Lock()  # completions from worker threads
        self._gpu_lock = threading.Lock()  # a graph capture vs the workers' HIP calls
        self._capture_pending = threading.Event()
        self.on_complete = on_complete

    @property
    def used(self):
        return (self.pool.bytes if self.pool else 0) + sum(j.bytes for j in self.jobs)

    def submit(self, req):
        self.pending.append(req)

    def _complete(self, r, t, cell_steps):
        with self._lock:
            r.done = t
            self.cell_steps += cell_steps
            self.finished.append(r)
            if self.on_complete is not None:
                self.on_complete(r)

    # ------------------------------------------------------------------ SPV pool
    def _new_systems(self, group):
        """A batch of the admitted requests, initialised as when served alone."""
        ns = np.array([r.n for r in group], np.int32)
        b = VoronoiBatch(ns, seeds=np.array([r.seed for r in group], np.uint64), precision=self.precision)
        b.set_params(KA=1.0, KP=1.0, dt=self.dt, mu=1.0)
        b.spatial_sort()
        b.set_prefs(1.0, np.repeat([r.P0 for r in group], ns))
        b.set_motility(np.repeat([r.v0 for r in group], ns), np.repeat([r.Dr for r in group], ns))
        b.set_active(np.ones(len(group), np.int32), np.array([r.steps for r in group], np.int32))
        return b

    def _rebuild(self, group):
        """New pool: the running systems of the old one, then `group`."""
        fresh = self._new_systems(group) if group else None
        keep = [(k, r) for k, r in enumerate(self.pool.reqs) if r is not None] if self.pool else []
        t = self.clock()
        for r in group:
            r.admitted = t
        if not keep and fresh is not None:
            batch = fresh
        else:
            parts = [(self.pool.batch, k) for k, _ in keep] + [(fresh, k) for k in range(len(group))]
            batch = VoronoiBatch.merge(parts) if parts else None
            if fresh is not None:
                fresh.close()
        if self.pool:
            self.pool.batch.check_status()
            self.pool.batch.close()
        self.pool = _Pool(batch, [r for _, r in keep] + list(group)) if batch is not None else None
        self.rebuilds += 1

    def _admissible(self, ready, extra_bytes_free):
        group, cells = [], 0
        for r in sorted(ready, key=lambda q: q.arrival):
            if cells + r.n > self.max_batch_cells or (cells + r.n) * self.BYTES_PER_CELL_ESTIMATE > extra_bytes_free:
                break
            group.append(r)
            cells += r.n
        return group

    def _admit(self, now):
        ready = [r for r in self.pending if r.arrival <= now]
        spv = [r for r in ready if r.kind == "spv"]
        fire = [r for r in ready if r.kind == "vertex_fire"]
        busy = bool(self.jobs or self.pool)
        if fire and (not busy or now - min(r.arrival for r in fire) >= self.admit_interval):
            free = self.budget - self.used
            group = self._admissible(fire, free)
            if group or not busy:
                group = group or [min(fire, key=lambda q: q.arrival)]
                for r in group:
                    self.pending.remove(r)
                    r.admitted = now
                job = _FireJob(self, group)
                self.jobs.append(job)
                job.start()
        waited = spv and now - min(r.arrival for r in spv) >= self.admit_interval
        reclaim = self.pool is not None and self.pool.dead_cells() > self.reclaim_fraction * self.pool.batch.N
        if not spv and not reclaim:
            return
        if spv and self.pool is not None and not waited and not reclaim:
            return  # let arrivals gather before the next rebuild
        live = self.pool.live_cells() if self.pool else 0
        free = self.budget - self.used + (self.pool.bytes if self.pool else 0) - live * self.BYTES_PER_CELL_ESTIMATE
        group = self._admissible(spv, free) if spv else []
        if spv and not group and self.pool is None and not self.jobs:
            group = [min(spv, key=lambda q: q.arrival)]  # larger than the budget: run it alone
        for r in group:
            self.pending.remove(r)
        if group or reclaim:
            self._rebuild(group)

    def _advance_pool(self):
        p = self.pool
        if p.captured:
            p.batch.spv_steps(self.tick_steps, use_graph=True)
        else:
            self._capture_pending.set()
            try:
                with self._gpu_lock:
                    p.batch.spv_steps(self.tick_steps, use_graph=True)
            finally:
                self._capture_pending.clear()
            p.captured = self.tick_steps >= 2
        active = np.zeros(p.batch.S, np.int32)
        check(lib().cgpu_vor_get_active(p.batch.h, active.ctypes.data, None), "get_active")
        t = self.clock()
        newly = [k for k, r in enumerate(p.reqs) if r is not None and active[k] == 0]
        if newly:
            g = p.batch.get() if self.keep_results else None
            for k in newly:
                r = p.reqs[k]
                if g is not None:
                    sl = p.batch.system_slice(k)
                    r.result = (g["x"][sl].copy(), g["y"][sl].copy())
                r.iterations, r.converged = r.steps, True
                self._complete(r, t, r.n * r.steps)
                p.reqs[k] = None
        if all(r is None for r in p.reqs):
            p.batch.check_status()
            p.batch.close()
            self.pool = None

    def _reap_jobs(self):
        for j in list(self.jobs):
            if not j.is_alive():
                j.join()
                self.jobs.remove(j)
                if j.error is not None:
                    raise j.error

    def run(self, requests, t0=None):
        """Serve `requests` (arrival times relative to the start) until all are
        done. Returns the finished requests with absolute done times."""
        t0 = self.clock() if t0 is None else t0
        for r in requests:
            r.arrival += t0
            self.submit(r)
        while self.pending or self.pool or self.jobs:
            now = self.clock()
            self._admit(now)
            if self.pool:
                self._advance_pool()
            else:
                nxt = min((r.arrival for r in self.pending), default=now + 0.01)
                time.sleep(max(0.0, min(nxt - self.clock(), 0.01)))
            self._reap_jobs()
        return self.finished


def poisson_requests(count, rate, seed=1, n_range=(100, 65536), steps_range=(1000, 100000), fire_fraction=0.0,
                     fire_n_range=(100, 65536), fire_max_iterations=None, force_cutoff=1e-8):
    """Requests of the serving workload: Poisson arrivals at `rate` per second.
    SPV: N and steps log-uniform, P0 ~ U[3.5, 4], v0 ~ U[0, 0.2], Dr = 1. A
    fraction fire_fraction are vertex-model FIRE minimizations instead (N
    log-uniform over fire_n_range, P0 ~ U[3.5, 4], RMS force cutoff; no
    iteration limit unless fire_max_iterations is given)."""
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
