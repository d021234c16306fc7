"""cellgpu_amd — MI355X-native Voronoi / vertex-model time step.

Python front end of libcellgpu_amd.so. The classes keep the reference's
(abaso/cellGPU) plugin API names — Simulation, VoronoiQuadraticEnergy,
selfPropelledParticleDynamics, brownianParticleDynamics, EnergyMinimizerFIRE,
VertexQuadraticEnergy — on top of device-resident batches (VoronoiBatch,
VertexBatch). Everything runs on the GPU; there is no CPU path.
"""
import ctypes as C

import numpy as np

from ._lib import CellGPUError, check, lib, require_device, LaunchShape, VoronoiBatchView  # noqa: F401

__all__ = [
    "CellGPUError", "VoronoiBatch", "VertexBatch", "VertexQuadraticEnergy", "VoronoiQuadraticEnergy", "VoronoiQuadraticEnergyWithTension", "Simulation", "selfPropelledParticleDynamics",
    "selfPropelledCellVertexDynamics", "brownianParticleDynamics", "EnergyMinimizerFIRE", "RING_CAP",
]

RING_CAP = 64  # largest ring capacity (CGPU_RING_CAP_MAX): getters copy up to this many slots


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f64(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if n is not None and a.shape != (n,):
        a = np.ascontiguousarray(np.broadcast_to(a, (n,)), dtype=np.float64)
    return a


class VoronoiBatch:
    """A device-resident batch of independent periodic Voronoi monolayers.

    n: cells per system; L: box sides (default sqrt(n), unit density);
    seeds: reference mt19937 seeds of the initial condition (used when x, y are
    None); precision: "fp64" (plain) or "fp32" (FP32 bulk, FP64 predicates).
    """

    def __init__(self, n, L=None, seeds=None, x=None, y=None, theta=None, precision="fp32", device=0):
        require_device()
        n = np.atleast_1d(np.asarray(n, dtype=np.int32))
        self.S = len(n)
        self.n = n
        self.offsets = np.concatenate([[0], np.cumsum(n)]).astype(np.int64)
        self.N = int(self.offsets[-1])
        Ls = None if L is None else _f64(np.broadcast_to(np.asarray(L, np.float64), (self.S,)))
        self.L = np.sqrt(n.astype(np.float64)) if Ls is None else Ls
        sd = None if seeds is None else np.ascontiguousarray(np.broadcast_to(np.asarray(seeds, np.uint64), (self.S,)))
        xs = None if x is None else _f64(x)
        ys = None if y is None else _f64(y)
        ts = None if theta is None else _f64(theta)
        self.precision = precision
        prec = 1 if precision == "fp32" else 0
        self._keep = (n, Ls, sd, xs, ys, ts)
        self.h = lib().cgpu_vor_create(self.S, _p(n), _p(Ls), _p(sd), _p(xs), _p(ys), _p(ts), prec, device)
        if not self.h:
            raise CellGPUError("cgpu_vor_create failed: " + lib().cgpu_last_error().decode())

    @classmethod
    def merge(cls, parts):
        """A new batch of the systems parts = [(batch, s), ...] (system s of each
        batch, in that order), each continuing its trajectory exactly; the sources
        stay valid (cgpu_vor_merge)."""
        require_device()
        if not parts:
            raise ValueError("merge needs at least one system")
        hs = (C.c_void_p * len(parts))(*[b.h for b, _ in parts])
        ss = np.ascontiguousarray([int(s) for _, s in parts], np.int32)
        h = lib().cgpu_vor_merge(len(parts), hs, _p(ss))
        if not h:
            raise CellGPUError("cgpu_vor_merge failed: " + lib().cgpu_last_error().decode())
        self = cls.__new__(cls)
        self.h = h
        self.n = np.array([b.n[s] for b, s in parts], np.int32)
        self.S = len(parts)
        self.offsets = np.concatenate([[0], np.cumsum(self.n)]).astype(np.int64)
        self.N = int(self.offsets[-1])
        self.L = np.array([b.L[s] for b, s in parts], np.float64)
        self.precision = parts[0][0].precision
        self._keep = ()
        return self

    def close(self):
        if getattr(self, "h", None):
            lib().cgpu_vor_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    # ---------------------------------------------------------------- setters
    def set_params(self, KA=1.0, KP=1.0, dt=0.01, mu=1.0, T=0.0, system=-1):
        check(lib().cgpu_vor_set_params(self.h, system, KA, KP, dt, mu, T), "set_params")

    def set_prefs(self, A0, P0):
        a, p = _f64(A0, self.N), _f64(P0, self.N)
        check(lib().cgpu_vor_set_prefs(self.h, _p(a), _p(p)), "set_prefs")

    def set_motility(self, v0, Dr):
        a, b = _f64(v0, self.N), _f64(Dr, self.N)
        check(lib().cgpu_vor_set_motility(self.h, _p(a), _p(b)), "set_motility")

    def set_exclusions(self, ex):
        e = np.ascontiguousarray(ex, dtype=np.int32)
        check(lib().cgpu_vor_set_exclusions(self.h, _p(e)), "set_exclusions")

    def external_forces(self):
        """(fx, fy): minus the summed force sets of excluded cells, zero elsewhere."""
        fx, fy = np.zeros(self.N), np.zeros(self.N)
        check(lib().cgpu_vor_get_external_forces(self.h, _p(fx), _p(fy)), "get_external_forces")
        return fx, fy

    def enforce_topology(self):
        """Test and repair of the rings at the current positions (reference
        voronoiModelBase::enforceTopology); step counters do not advance."""
        check(lib().cgpu_vor_enforce_topology(self.h), "enforce_topology")

    def set_directors(self, theta):
        t = _f64(theta, self.N)
        check(lib().cgpu_vor_set_directors(self.h, _p(t)), "set_directors")

    def set_velocities(self, vx, vy):
        check(lib().cgpu_vor_set_velocities(self.h, _p(_f64(vx, self.N)), _p(_f64(vy, self.N))), "set_velocities")

    def set_topology(self, nn, nbr, del_other=None):
        """Replace every cell's Delaunay ring (nbr[i, :nn[i]], counter-clockwise,
        batch indices); geometry and neighMax are derived, delOther too unless
        given (same layout as nbr)."""
        nn = np.ascontiguousarray(nn, np.int32)
        nbr = np.ascontiguousarray(nbr, np.int32)
        oth = None
        if del_other is not None:
            oth = np.ascontiguousarray(del_other, np.int32)
            if oth.ndim != 2 or oth.shape[0] != nbr.shape[0]:
                raise ValueError("del_other must have one row per cell, as nbr")
            if oth.shape[1] != nbr.shape[1]:  # the slots past a degree are unused
                w = min(oth.shape[1], nbr.shape[1])
                o2 = np.full(nbr.shape, -1, np.int32)
                o2[:, :w] = oth[:, :w]
                oth = o2
        check(lib().cgpu_vor_set_topology(self.h, _p(nn), _p(nbr), int(nbr.shape[1]), _p(oth)), "set_topology")

    def del_other(self, cap=RING_CAP):
        """delOther of every slot, [N, cap] (slots past a cell's degree: -1)."""
        out = np.full(self.N * cap, -1, np.int32)
        check(lib().cgpu_vor_get_del_other(self.h, _p(out), cap), "get_del_other")
        return out.reshape(self.N, cap)

    def set_step_counters(self, step):
        st = np.ascontiguousarray(np.broadcast_to(np.asarray(step, np.int32), (self.S,)))
        check(lib().cgpu_vor_set_step_counters(self.h, _p(st)), "set_step_counters")

    def tags(self):
        """Label of every cell within its system at creation (spatial sorting permutes cells)."""
        out = np.zeros(self.N, np.int32)
        check(lib().cgpu_vor_get_tags(self.h, _p(out)), "get_tags")
        return out

    def set_cell_types(self, types):
        """Cell types (reference Simple2DCell::setCellType); None clears them."""
        self._types = None if types is None else np.ascontiguousarray(types, dtype=np.int32)
        if self._types is not None and self._types.shape != (self.N,):
            raise CellGPUError("one type per cell")
        check(lib().cgpu_vor_set_cell_types(self.h, _p(self._types)), "set_cell_types")

    def set_tension(self, mode, gamma=0.0, system=-1):
        """Line tension of system `system` (-1: all): mode 0 off, 1 one value
        gamma between unlike types, 2 the tension matrix."""
        check(lib().cgpu_vor_set_tension(self.h, int(system), int(mode), float(gamma)), "set_tension")

    def set_tension_matrix(self, g):
        """g[a, b]: tension between types a and b (reference setSurfaceTension(vector))."""
        g = np.ascontiguousarray(g, dtype=np.float64)
        if g.ndim != 2 or g.shape[0] != g.shape[1]:
            raise CellGPUError("the tension matrix is square")
        check(lib().cgpu_vor_set_tension_matrix(self.h, g.shape[0], _p(g)), "set_tension_matrix")

    def set_positions(self, x, y, retriangulate=True):
        a, b = _f64(x, self.N), _f64(y, self.N)
        check(lib().cgpu_vor_set_positions(self.h, _p(a), _p(b), 1 if retriangulate else 0), "set_positions")

    def set_noise_buffer(self, on=True):
        check(lib().cgpu_vor_set_noise_mode(self.h, 1 if on else 0), "set_noise_mode")

    def upload_noise(self, eta, eta2=None):
        buf = np.zeros(2 * self.N)
        buf[: self.N] = eta
        if eta2 is not None:
            buf[self.N:] = eta2
        check(lib().cgpu_vor_upload_noise(self.h, _p(buf)), "upload_noise")

    def set_active(self, active=None, steps_left=None):
        a = None if active is None else np.ascontiguousarray(active, np.int32)
        s = None if steps_left is None else np.ascontiguousarray(steps_left, np.int32)
        check(lib().cgpu_vor_set_active(self.h, _p(a), _p(s)), "set_active")

    # ---------------------------------------------------------------- actions
    def compute_forces(self):
        check(lib().cgpu_vor_compute_forces(self.h), "compute_forces")

    def spv_steps(self, steps, use_graph=True):
        check(lib().cgpu_vor_spv_steps(self.h, int(steps), 1 if use_graph else 0), "spv_steps")

    def brownian_steps(self, steps):
        check(lib().cgpu_vor_brownian_steps(self.h, int(steps)), "brownian_steps")

    def sync(self):
        check(lib().cgpu_vor_sync(self.h), "sync")

    def status(self):
        return lib().cgpu_vor_status(self.h)

    def check_status(self):
        check(self.status(), "batch status")

    # ---------------------------------------------------------------- getters
    def get(self):
        N = self.N
        a = {k: np.zeros(N) for k in ("x", "y", "theta", "vx", "vy", "Fx", "Fy", "A", "P")}
        check(lib().cgpu_vor_get(self.h, *[_p(a[k]) for k in ("x", "y", "theta", "vx", "vy", "Fx", "Fy", "A", "P")]),
              "get")
        return a

    def topology(self, cap=RING_CAP):
        nn = np.zeros(self.N, np.int32)
        nbr = np.full(self.N * cap, -1, np.int32)
        nm = np.zeros(self.S, np.int32)
        check(lib().cgpu_vor_get_topology(self.h, _p(nn), _p(nbr), cap, _p(nm)), "get_topology")
        return nn, nbr.reshape(self.N, cap), nm

    def counters(self, s=0):
        out = np.zeros(8, np.int64)
        check(lib().cgpu_vor_get_counters(self.h, s, _p(out)), "get_counters")
        keys = ["global_fixes", "local_repairs", "repair_events", "step", "last_mode", "last_nf", "neigh_max", "error"]
        return dict(zip(keys, out.tolist()))

    def energy(self, s=0):
        return lib().cgpu_vor_energy(self.h, s)

    def view(self):
        v = VoronoiBatchView()
        lib().cgpu_vor_view(self.h, C.byref(v))
        return v

    def stream(self):
        return lib().cgpu_vor_stream(self.h)

    SHAPE_FIELDS = ("force_block", "test_block", "cell_block", "star_blocks", "repair_blocks", "queue_lanes",
                    "force_waves")

    def launch_shape(self):
        sh = LaunchShape()
        check(lib().cgpu_vor_get_launch_shape(self.h, C.byref(sh)), "get_launch_shape")
        return {k: getattr(sh, k) for k in self.SHAPE_FIELDS}

    def set_launch_shape(self, **kw):
        """Launch shapes of the kernels (fields of cgpu_launch_shape; 0 = default)."""
        sh = LaunchShape()
        for k in self.SHAPE_FIELDS:
            setattr(sh, k, int(kw.pop(k, 0)))
        if kw:
            raise TypeError("unknown launch shape fields: %s" % sorted(kw))
        check(lib().cgpu_vor_set_launch_shape(self.h, C.byref(sh)), "set_launch_shape")

    def autotune(self, trial_steps=20):
        """Search the launch shapes on the current state (restored afterwards),
        apply the best; returns (shape dict, from_cache)."""
        sh = LaunchShape()
        cached = C.c_int(0)
        check(lib().cgpu_vor_autotune(self.h, int(trial_steps), C.byref(sh), C.byref(cached)), "autotune")
        return {k: getattr(sh, k) for k in self.SHAPE_FIELDS}, bool(cached.value)

    def profiled_steps(self, steps):
        """Steps with a HIP event after every kernel. Returns ({kernel: summed
        ms}, summed queue sizes (flagged, NeedsFixing, global, triangles left to
        the circumcircle test, regions stamped))."""
        from ._lib import PROFILE_SLOTS
        ms = np.zeros(PROFILE_SLOTS)
        q = np.zeros(5, np.int64)
        check(lib().cgpu_vor_spv_steps_profiled(self.h, int(steps), _p(ms), _p(q)), "profiled steps")
        names = [lib().cgpu_profile_slot_name(k).decode() for k in range(PROFILE_SLOTS)]
        return dict(zip(names, ms.tolist())), q.tolist()

    def ring_capacity(self):
        """ring slots per cell now allocated (grows when a star needs more)."""
        return int(lib().cgpu_vor_ring_capacity(self.h))

    def spatial_sort(self):
        """Relabel cells along a Hilbert curve (per system) and retriangulate.
        Returns perm with new index -> old index."""
        perm = np.zeros(self.N, np.int32)
        check(lib().cgpu_vor_spatial_sort(self.h, _p(perm)), "spatial_sort")
        return perm

    def system_slice(self, s):
        return slice(int(self.offsets[s]), int(self.offsets[s + 1]))

    def cell_events(self, kinds, cells, theta=None, frac=None, new_director=None):
        """Deaths (kind 0) and divisions (kind 1) of batch cells, at most one per
        system (reference voronoiModelBase::cellDeath / cellDivision); the batch is
        laid out again and retriangulated."""
        k = np.ascontiguousarray(np.atleast_1d(kinds), np.int32)
        c = np.ascontiguousarray(np.atleast_1d(cells), np.int32)
        m = len(k)
        t = _f64(np.zeros(m) if theta is None else np.atleast_1d(theta), m)
        f = _f64(np.zeros(m) if frac is None else np.atleast_1d(frac), m)
        d = None if new_director is None else _f64(np.atleast_1d(new_director), m)
        check(lib().cgpu_vor_cell_events(self.h, m, _p(k), _p(c), _p(t), _p(f), _p(d)), "cell_events")
        n = np.zeros(self.S, np.int32)
        lib().cgpu_vor_get_sizes(self.h, _p(n))
        self.n = n
        self.offsets = np.concatenate([[0], np.cumsum(n)]).astype(np.int64)
        self.N = int(self.offsets[-1])


# ----------------------------------------------------------------------------
# Reference-named single-system API (abaso/cellGPU inc/models, inc/updaters,
# inc/Simulation.h). A model owns a one-system VoronoiBatch.
# ----------------------------------------------------------------------------
class VoronoiQuadraticEnergy:
    """reference inc/models/voronoiQuadraticEnergy.h:19-64 (GPU branch only)."""

    def __init__(self, n, A0=1.0, P0=3.8, reprod=True, precision="fp64", x=None, y=None, theta=None, seed=13377,
                 device=0):
        self.Ncells = int(n)
        self.batch = VoronoiBatch([n], seeds=[seed if reprod else np.random.randint(1, 2**31)], x=x, y=y,
                                  theta=theta, precision=precision, device=device)
        self.KA = self.KP = 1.0
        self.deltaT = 0.01
        self.setCellPreferencesUniform(A0, P0)
        self.setv0Dr(0.05, 1.0)

    def setCellPreferencesUniform(self, A0, P0):
        self.batch.set_prefs(np.full(self.Ncells, A0), np.full(self.Ncells, P0))

    def setCellPreferences(self, AP):
        AP = np.asarray(AP, np.float64)
        self.batch.set_prefs(AP[:, 0], AP[:, 1])

    def setv0Dr(self, v0, Dr):
        self.batch.set_motility(np.full(self.Ncells, v0), np.full(self.Ncells, Dr))

    def setModuliUniform(self, KA, KP):
        self.KA, self.KP = KA, KP
        self.batch.set_params(KA=KA, KP=KP, dt=self.deltaT)

    def setDeltaT(self, dt):
        self.deltaT = dt
        self.batch.set_params(KA=self.KA, KP=self.KP, dt=dt)

    def setExclusions(self, ex):
        self.batch.set_exclusions(ex)

    def setCPU(self, globalOnly=True):
        raise CellGPUError("cellgpu_amd has no CPU path: the time step runs on the GPU only")

    def computeForces(self):
        self.batch.compute_forces()

    def computeEnergy(self):
        return self.batch.energy(0)

    def returnPositions(self):
        g = self.batch.get()
        return np.stack([g["x"], g["y"]], 1)

    def returnForces(self):
        g = self.batch.get()
        return np.stack([g["Fx"], g["Fy"]], 1)

    def returnAreaPeri(self):
        g = self.batch.get()
        return np.stack([g["A"], g["P"]], 1)

    def reportq(self):
        g = self.batch.get()
        return float(np.mean(g["P"] / np.sqrt(g["A"])))

    def reportMeanCellForce(self, verbose=False):
        g = self.batch.get()
        return float(np.mean(g["Fx"])), float(np.mean(g["Fy"]))

    def setCellType(self, types):
        """reference Simple2DCell::setCellType."""
        self.batch.set_cell_types(types)

    def spatialSorting(self):
        """Hilbert relabelling of the cells on the device, then retriangulation
        (reference voronoiModelBase::spatialSorting, voronoiModelBase.cpp:669-680).
        itt[new] = old index of the last sort; idxToTag / tagToIdx relate current
        indices to the labels at creation (Simple2DCell.h:241-253)."""
        self.itt = self.batch.spatial_sort()
        self.tti = np.empty_like(self.itt)
        self.tti[self.itt] = np.arange(len(self.itt), dtype=self.itt.dtype)
        t = np.zeros(self.Ncells, np.int32)
        check(lib().cgpu_vor_get_tags(self.batch.h, _p(t)), "tags")
        self.idxToTag = t
        self.tagToIdx = np.empty_like(t)
        self.tagToIdx[t] = np.arange(self.Ncells, dtype=np.int32)

    def cellDeath(self, cellIndex):
        """reference voronoiModelBase::cellDeath (voronoiModelBase.cpp:1309-1316)."""
        self.batch.cell_events([0], [cellIndex])
        self.Ncells = self.batch.N

    def cellDivision(self, parameters, dParams):
        """reference voronoiModelBase::cellDivision: parameters = [cell], dParams =
        [theta, separation fraction]; the new cell is the last one."""
        self.batch.cell_events([1], [parameters[0]], [dParams[0]], [dParams[1]])
        self.Ncells = self.batch.N


class VoronoiQuadraticEnergyWithTension(VoronoiQuadraticEnergy):
    """reference inc/models/voronoiQuadraticEnergyWithTension.h:16-60 (GPU branch):
    line tension between cells of different types, one value for every unlike pair
    or a matrix over types. computeEnergy adds gamma times the length of every
    edge between unlike cells, counted from both cells (the energy whose gradient
    the reference's force sets are)."""

    def __init__(self, n, A0=1.0, P0=3.8, reprod=True, **kw):
        super().__init__(n, A0, P0, reprod, **kw)
        self.gamma = 0.0
        self.Tension = False
        self.simpleTension = True

    def _push(self):
        mode = 0 if not self.Tension else (1 if self.simpleTension else 2)
        self.batch.set_tension(mode, self.gamma)

    def setUseSurfaceTension(self, use_tension):
        self.Tension = bool(use_tension)
        self._push()

    def setSurfaceTension(self, g):
        """A number: one tension for every unlike pair. A sequence of n*n values
        (or an n x n matrix): g[a*n + b] between types a and b."""
        if np.ndim(g) == 0:
            self.gamma = float(g)
            self.simpleTension = True
        else:
            m = np.asarray(g, np.float64)
            n = int(round(np.sqrt(m.size)))
            if n * n != m.size:
                raise CellGPUError("the tension matrix has n*n entries")
            self.batch.set_tension_matrix(m.reshape(n, n))
            self.simpleTension = False
        self._push()

    def getSurfaceTension(self):
        return self.gamma


class _Updater:
    def __init__(self, n=None):
        self.Period = 1
        self.Phase = 0
        self.model = None
        self.deltaT = 0.01

    def set2DModel(self, m):
        self.model = m

    def setDeltaT(self, dt):
        self.deltaT = dt

    def setGPU(self):
        pass

    def setCPU(self):
        raise CellGPUError("cellgpu_amd has no CPU path: the time step runs on the GPU only")

    def setReproducible(self, r):
        pass


class selfPropelledParticleDynamics(_Updater):
    """reference inc/updaters/selfPropelledParticleDynamics.h:13-44 (GPU semantics)."""

    def performUpdate(self):
        self.model.batch.spv_steps(1, use_graph=False)

    def advance(self, steps):
        self.model.batch.spv_steps(steps, use_graph=True)


class brownianParticleDynamics(_Updater):
    """reference inc/updaters/brownianParticleDynamics.h:14-51."""

    def __init__(self, n=None, T=0.0, mu=1.0):
        super().__init__(n)
        self.T = T
        self.mu = mu

    def setT(self, T):
        self.T = T


    def performUpdate(self):
        m = self.model
        if isinstance(m, VertexQuadraticEnergy):
            m.push_params(mu=self.mu, T=self.T)
        else:
            m.batch.set_params(KA=m.KA, KP=m.KP, dt=m.deltaT, mu=self.mu, T=self.T)
        m.batch.brownian_steps(1)


class selfPropelledCellVertexDynamics(_Updater):
    """reference inc/updaters/selfPropelledCellVertexDynamics.h (GPU branch,
    selfPropelledCellVertexDynamics.cpp:51-76): vertices move by dt (mean of the
    three cells' v0 e(theta) + mu F), directors diffuse with Dr, then T1 flips."""

    def __init__(self, Ncells=None, Nvertices=None, mu=1.0):
        super().__init__(Ncells)
        self.mu = mu

    def set2DModel(self, m):
        if not isinstance(m, VertexQuadraticEnergy):
            raise CellGPUError("selfPropelledCellVertexDynamics drives a vertex model")
        self.model = m

    def performUpdate(self):
        self.model.push_params(mu=self.mu)
        self.model.batch.spp_steps(1, use_graph=False)

    def advance(self, steps):
        self.model.push_params(mu=self.mu)
        self.model.batch.spp_steps(steps, use_graph=True)


class Simulation:
    """reference inc/Simulation.h:18-87."""

    def __init__(self):
        self.updaters = []
        self.config = None
        self.integerTimestep = 0
        self.Time = 0.0
        self.integrationTimestep = 0.01

    def setConfiguration(self, c):
        self.config = c

    def addUpdater(self, upd, config):
        upd.set2DModel(config)
        self.updaters.append(upd)

    def setIntegrationTimestep(self, dt):
        self.integrationTimestep = dt
        self.config.setDeltaT(dt)
        for u in self.updaters:
            u.setDeltaT(dt)

    def setCPUOperation(self, setcpu):
        if setcpu:
            raise CellGPUError("cellgpu_amd has no CPU path: the time step runs on the GPU only")

    def setReproducible(self, r):
        for u in self.updaters:
            u.setReproducible(r)

    def setSortPeriod(self, sp):
        """Spatial sorting every sp steps (reference Simulation::setSortPeriod;
        sp <= 0: never)."""
        self.sortPeriod = int(sp)

    def _sort_if_due(self):
        # reference Simulation.cpp:155-164: after the updaters of a step whose
        # number is a multiple of the period
        sp = getattr(self, "sortPeriod", -1)
        if sp > 0 and self.integerTimestep % sp == 0:
            self.config.spatialSorting()
            for u in self.updaters:
                if hasattr(u, "spatialSorting"):
                    u.spatialSorting()

    def performTimestep(self):
        self.integerTimestep += 1
        self.Time += self.integrationTimestep
        for u in self.updaters:
            u.performUpdate()
        self._sort_if_due()

    def performTimesteps(self, k):
        """k steps at once (graph-captured between sorts); same result as k
        performTimestep calls."""
        if len(self.updaters) == 1 and isinstance(self.updaters[0],
                                                  (selfPropelledParticleDynamics, selfPropelledCellVertexDynamics)):
            sp = getattr(self, "sortPeriod", -1)
            while k > 0:
                run = k if sp <= 0 else min(k, sp - self.integerTimestep % sp)
                self.updaters[0].advance(run)
                self.integerTimestep += run
                self.Time += run * self.integrationTimestep
                k -= run
                self._sort_if_due()
        else:
            for _ in range(k):
                self.performTimestep()


class EnergyMinimizerFIRE:
    """reference inc/updaters/EnergyMinimizerFIRE2D.h:20-132 (GPU branch)."""

    def __init__(self, model):
        self.model = model
        batch = model.batch if hasattr(model, "batch") else model
        self.batch = batch
        create = lib().cgpu_vm_fire_create if isinstance(batch, VertexBatch) else lib().cgpu_fire_create
        if isinstance(model, VertexQuadraticEnergy):
            model.push_params()
        self.h = create(batch.h)
        if not self.h:
            raise CellGPUError("cgpu_fire_create failed: " + lib().cgpu_last_error().decode())
        self.params = dict(max_it=1000, force_cutoff=1e-7, alpha_start=0.99, dt=0.01, dt_max=0.1, dt_inc=1.05,
                           dt_dec=0.95, alpha_dec=0.9, nmin=5)
        self._push()

    def close(self):
        if getattr(self, "h", None):
            lib().cgpu_fire_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _push(self):
        p = self.params
        check(lib().cgpu_fire_set(self.h, p["max_it"], p["force_cutoff"], p["alpha_start"], p["dt"], p["dt_max"],
                                  p["dt_inc"], p["dt_dec"], p["alpha_dec"], p["nmin"]), "fire_set")

    # one parameter at a time, as the reference setters (EnergyMinimizerFIRE2D.h:37-54):
    # setAlphaStart also sets alpha, setDeltaT the current dt
    _WHICH = dict(max_it=0, force_cutoff=1, alpha_start=2, dt=3, dt_max=4, dt_inc=5, dt_dec=6, alpha_dec=7, nmin=8)

    def _set(self, name, v):
        self.params[name] = v
        check(lib().cgpu_fire_set_param(self.h, self._WHICH[name], float(v)), "fire_set_param")

    def setMaximumIterations(self, m):
        self._set("max_it", int(m))

    def setForceCutoff(self, f):
        self._set("force_cutoff", f)

    def setAlphaStart(self, a):
        self._set("alpha_start", a)

    def setDeltaT(self, dt):
        self._set("dt", dt)

    def setDeltaTMax(self, t):
        self._set("dt_max", t)

    def setDeltaTInc(self, t):
        self._set("dt_inc", t)

    def setDeltaTDec(self, t):
        self._set("dt_dec", t)

    def setAlphaDec(self, a):
        self._set("alpha_dec", a)

    def setNMin(self, n):
        self._set("nmin", int(n))

    def minimize(self):
        check(lib().cgpu_fire_minimize(self.h), "fire_minimize")

    def state(self, s=0):
        out = np.zeros(6)
        check(lib().cgpu_fire_get(self.h, s, _p(out)), "fire_get")
        return dict(force_max=out[0], power=out[1], alpha=out[2], dt=out[3], max_f2=out[4], iterations=int(out[5]))

    def getMaxForce(self, s=0):
        return self.state(s)["force_max"]


class VertexBatch:
    """A device-resident batch of vertex-model monolayers, each built from the
    Voronoi tessellation of its cell positions (reference
    vertexModelBase::setCellsVoronoiTesselation). Cells: ncells per system;
    positions from x, y (concatenated) or the reference random start (seeds)."""

    VCAP = 32

    def __init__(self, ncells, L=None, seeds=None, x=None, y=None, precision="fp32", device=0, spv_initialize=False):
        require_device()
        nc = np.atleast_1d(np.asarray(ncells, dtype=np.int32))
        self.S = len(nc)
        self.Nc = int(nc.sum())
        Ls = None if L is None else _f64(np.broadcast_to(np.asarray(L, np.float64), (self.S,)))
        self.L = np.sqrt(nc.astype(np.float64)) if Ls is None else Ls
        sd = None if seeds is None else np.ascontiguousarray(np.broadcast_to(np.asarray(seeds, np.uint64), (self.S,)))
        xs = None if x is None else _f64(x)
        ys = None if y is None else _f64(y)
        self._keep = (nc, Ls, sd, xs, ys)
        self.h = lib().cgpu_vm_create(self.S, _p(nc), _p(Ls), _p(sd), _p(xs), _p(ys), 1 if precision == "fp32" else 0,
                                      device, 1 if spv_initialize else 0)
        if not self.h:
            raise CellGPUError("cgpu_vm_create failed")
        self.Nv = lib().cgpu_vm_num_vertices(self.h)

    def close(self):
        if getattr(self, "h", None):
            lib().cgpu_vm_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_params(self, KA=1.0, KP=1.0, dt=0.005, mu=1.0, T=0.0, t1_threshold=0.04, system=-1):
        check(lib().cgpu_vm_set_params(self.h, system, KA, KP, dt, mu, T, t1_threshold), "vm_set_params")

    def set_prefs(self, A0, P0):
        a, p = _f64(A0, self.Nc), _f64(P0, self.Nc)
        check(lib().cgpu_vm_set_prefs(self.h, _p(a), _p(p)), "vm_set_prefs")

    def set_positions(self, vx, vy):
        a, b = _f64(vx, self.Nv), _f64(vy, self.Nv)
        check(lib().cgpu_vm_set_positions(self.h, _p(a), _p(b)), "vm_set_positions")

    def set_noise_buffer(self, on=True):
        check(lib().cgpu_vm_set_noise_mode(self.h, 1 if on else 0), "vm_set_noise_mode")

    def upload_noise(self, etax, etay):
        buf = np.concatenate([_f64(etax, self.Nv), _f64(etay, self.Nv)])
        check(lib().cgpu_vm_upload_noise(self.h, _p(buf)), "vm_upload_noise")

    def upload_cell_noise(self, eta):
        """N(0,1) draws of the director rotation, one per cell (noise-buffer mode)."""
        check(lib().cgpu_vm_upload_cell_noise(self.h, _p(_f64(eta, self.Nc))), "vm_upload_cell_noise")

    def set_motility(self, v0, Dr):
        a, b = _f64(v0, self.Nc), _f64(Dr, self.Nc)
        check(lib().cgpu_vm_set_motility(self.h, _p(a), _p(b)), "vm_set_motility")

    def set_directors(self, theta):
        t = _f64(theta, self.Nc)
        check(lib().cgpu_vm_set_directors(self.h, _p(t)), "vm_set_directors")

    def directors(self):
        out = np.zeros(self.Nc)
        check(lib().cgpu_vm_get_directors(self.h, _p(out)), "vm_get_directors")
        return out

    def compute_forces(self):
        check(lib().cgpu_vm_compute_forces(self.h), "vm_compute_forces")

    def brownian_steps(self, steps, use_graph=True):
        check(lib().cgpu_vm_brownian_steps(self.h, int(steps), 1 if use_graph else 0), "vm_brownian_steps")

    def cell_events(self, kinds, cells, p1=None, p2=None, new_director=None):
        """T2 deaths (kind 0) and divisions (kind 1, through the edges after
        vertex-list positions p1, p2) of batch cells, at most one per system
        (reference vertexModelBase::cellDeath / cellDivision)."""
        k = np.ascontiguousarray(np.atleast_1d(kinds), np.int32)
        c = np.ascontiguousarray(np.atleast_1d(cells), np.int32)
        m = len(k)
        a = np.ascontiguousarray(np.zeros(m) if p1 is None else np.atleast_1d(p1), np.int32)
        b = np.ascontiguousarray(np.zeros(m) if p2 is None else np.atleast_1d(p2), np.int32)
        d = None if new_director is None else _f64(np.atleast_1d(new_director), m)
        check(lib().cgpu_vm_cell_events(self.h, m, _p(k), _p(c), _p(a), _p(b), _p(d)), "vm_cell_events")
        nc = np.zeros(self.S, np.int32)
        self.Nv = int(lib().cgpu_vm_get_sizes(self.h, _p(nc), None))
        self.Nc = int(nc.sum())

    def spatial_sort(self):
        """Relabel vertices along a Hilbert curve (per system) and cells in order of
        first appearance around them (reference vertexModelBase::spatialSorting).
        Returns (vperm, cperm), new index -> old index."""
        vp = np.zeros(self.Nv, np.int32)
        cp = np.zeros(self.Nc, np.int32)
        check(lib().cgpu_vm_spatial_sort(self.h, _p(vp), _p(cp)), "vm_spatial_sort")
        return vp, cp

    def set_topology(self, vn, vcn, cvn, cv):
        """Replace the topology (checked for consistency first); cv is [Nc, cap]."""
        cv = np.ascontiguousarray(cv, np.int32)
        a = [np.ascontiguousarray(v, np.int32) for v in (vn, vcn, cvn)]
        check(lib().cgpu_vm_set_topology(self.h, _p(a[0]), _p(a[1]), _p(a[2]), _p(cv), cv.shape[1]), "vm_set_topology")

    def spp_steps(self, steps, use_graph=True):
        """Self-propelled cell-vertex steps (reference selfPropelledCellVertexDynamics)."""
        check(lib().cgpu_vm_spp_steps(self.h, int(steps), 1 if use_graph else 0), "vm_spp_steps")

    def check_status(self):
        check(lib().cgpu_vm_status(self.h), "vertex batch status")

    def get(self):
        a = {k: np.zeros(self.Nv) for k in ("vx", "vy", "Fx", "Fy")}
        a.update({k: np.zeros(self.Nc) for k in ("A", "P")})
        check(lib().cgpu_vm_get(self.h, *[_p(a[k]) for k in ("vx", "vy", "Fx", "Fy", "A", "P")]), "vm_get")
        return a

    def topology(self):
        vn = np.zeros(3 * self.Nv, np.int32)
        vcn = np.zeros(3 * self.Nv, np.int32)
        cvn = np.zeros(self.Nc, np.int32)
        cap = max(self.VCAP, self.vertex_capacity())
        cv = np.full(self.Nc * cap, -1, np.int32)
        check(lib().cgpu_vm_get_topology(self.h, _p(vn), _p(vcn), _p(cvn), _p(cv), cap), "vm_get_topology")
        return vn, vcn, cvn, cv.reshape(self.Nc, cap)

    def vertex_capacity(self):
        """cell-list slots per cell now allocated (grows when a flip needs more)."""
        return int(lib().cgpu_vm_vertex_capacity(self.h))

    def counters(self, s=0):
        out = np.zeros(4, np.int64)
        check(lib().cgpu_vm_get_counters(self.h, s, _p(out)), "vm_get_counters")
        return dict(zip(["t1_events", "step", "error", "nv"], out.tolist()))


class VertexQuadraticEnergy:
    """reference inc/models/vertexQuadraticEnergy.h:16-36 with vertexModelBase
    (vertexModelBase.h:26-109) and Simple2DActiveCell's motility (GPU branch),
    one system."""

    def __init__(self, n, A0=1.0, P0=3.9, reprod=True, runSPV=False, precision="fp64", seed=13377, device=0):
        self.Ncells = int(n)
        self.batch = VertexBatch([n], seeds=[seed], precision=precision, device=device, spv_initialize=runSPV)
        self.Nvertices = self.batch.Nv
        self.KA = self.KP = 1.0
        self.deltaT = 0.005
        self.T1Threshold = 0.01  # reference initializeVertexModelBase default
        self.setCellPreferencesUniform(A0, P0)

    def getNumberOfDegreesOfFreedom(self):
        return self.Nvertices

    def setCellPreferencesUniform(self, A0, P0):
        self.A0, self.P0 = float(A0), float(P0)
        self.batch.set_prefs(np.full(self.Ncells, A0), np.full(self.Ncells, P0))

    def setModuliUniform(self, KA, KP):
        self.KA, self.KP = float(KA), float(KP)

    def setT1Threshold(self, t):
        self.T1Threshold = t

    def spatialSorting(self):
        """Vertices along a Hilbert curve, cells by first appearance around them,
        on the device (reference vertexModelBase::spatialSorting,
        vertexModelBase.cpp:236-242). ittVertex / itt: new -> old of the last
        sort; idxToTag(Vertex): current index -> label at creation."""
        vp, cp = self.batch.spatial_sort()
        self.ittVertex, self.itt = vp, cp
        if not hasattr(self, "idxToTagVertex") or len(self.idxToTagVertex) != len(vp):
            self.idxToTagVertex = np.arange(len(vp), dtype=np.int32)
        if not hasattr(self, "idxToTag") or len(self.idxToTag) != len(cp):
            self.idxToTag = np.arange(len(cp), dtype=np.int32)
        self.idxToTagVertex = self.idxToTagVertex[vp]
        self.idxToTag = self.idxToTag[cp]

    def setDeltaT(self, dt):
        self.deltaT = dt

    def setv0Dr(self, v0, Dr):
        """reference Simple2DActiveCell::setv0Dr (Simple2DActiveCell.cpp:76-92)."""
        self.batch.set_motility(np.full(self.Ncells, v0), np.full(self.Ncells, Dr))

    def setCellDirectors(self, theta):
        self.batch.set_directors(theta)

    def cellDeath(self, cellIndex):
        """reference vertexModelBase::cellDeath (T2 of a triangular cell)."""
        self.batch.cell_events([0], [cellIndex])
        self.Ncells, self.Nvertices = self.batch.Nc, self.batch.Nv

    def cellDivision(self, parameters, dParams=()):
        """reference vertexModelBase::cellDivision: parameters = [cell, list
        position 1, list position 2]."""
        self.batch.cell_events([1], [parameters[0]], [parameters[1]], [parameters[2]])
        self.Ncells, self.Nvertices = self.batch.Nc, self.batch.Nv

    def returnCellDirectors(self):
        return self.batch.directors()

    def push_params(self, mu=1.0, T=0.0):
        """Hands the model's constants and an updater's (mu, T) to the device."""
        self.batch.set_params(KA=self.KA, KP=self.KP, dt=self.deltaT, mu=mu, T=T, t1_threshold=self.T1Threshold)

    def computeForces(self):
        self.push_params()
        self.batch.compute_forces()

    def computeEnergy(self):
        """reference VertexQuadraticEnergy::computeEnergy (vertexQuadraticEnergy.cpp:30-44)."""
        self.computeForces()
        g = self.batch.get()
        return float(np.sum(self.KA * (g["A"] - self.A0) ** 2 + self.KP * (g["P"] - self.P0) ** 2))

    def returnForces(self):
        g = self.batch.get()
        return np.stack([g["Fx"], g["Fy"]], 1)

    def returnPositions(self):
        g = self.batch.get()
        return np.stack([g["vx"], g["vy"]], 1)

    def returnAreaPeri(self):
        g = self.batch.get()
        return np.stack([g["A"], g["P"]], 1)
