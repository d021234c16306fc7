"""ORACLE — test infrastructure only.

ctypes front end of ``liboracle.so``, the CPU FP64 restatement of the reference
(abaso/cellGPU) Voronoi/vertex time step. Only ``tests/``, the smoke check and
``bench.py``'s CPU-baseline leg import this module; the product never does.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle` or __graft_entry__.build()")
        L = C.CDLL(path)
        vp = C.c_void_p
        sig = {
            "orc_orient0": (C.c_int, [C.c_double] * 4),
            "orc_incircle0": (C.c_int, [C.c_double] * 6),
            "orc_circumcenter0": (None, [C.c_double] * 4 + [_dp]),
            "orc_min_dist": (None, [C.c_double] * 5 + [_dp]),
            "orc_put_in_box": (None, [C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
            "orc_random_init": (None, [C.c_int, C.c_uint, C.c_double, _dp, _dp, _dp]),
            "orc_triangulate": (C.c_int, [C.c_int, C.c_double, _dp, _dp, C.c_int, _ip, _ip]),
            "orc_vsys_create": (vp, [C.c_int, C.c_double, _dp, _dp, _dp]),
            "orc_vsys_destroy": (None, [vp]),
            "orc_vsys_set_params": (None, [vp, C.c_double, C.c_double, C.c_double, C.c_double]),
            "orc_vsys_set_prefs": (None, [vp, _dp, _dp]),
            "orc_vsys_set_motility": (None, [vp, _dp, _dp]),
            "orc_vsys_set_exclusions": (None, [vp, _ip]),
            "orc_vsys_set_state": (None, [vp, _dp, _dp, vp, vp, vp]),
            "orc_vsys_set_topology": (None, [vp, _ip, _ip, C.c_int, C.c_int]),
            "orc_vsys_get": (C.c_int, [vp] + [vp] * 9),
            "orc_vsys_get_topology": (C.c_int, [vp, _ip, _ip, C.c_int]),
            "orc_vsys_get_counters": (None, [vp, np.ctypeslib.ndpointer(dtype=np.int64)]),
            "orc_vsys_compute_forces": (C.c_double, [vp, C.c_int]),
            "orc_vsys_set_tension": (None, [vp, C.c_int, C.c_double, C.c_int, vp, _ip]),
            "orc_vsys_cell_death": (C.c_int, [vp, C.c_int]),
            "orc_vsys_cell_division": (C.c_int, [vp, C.c_int, C.c_double, C.c_double, C.c_double]),
            "orc_vsys_num_cells": (C.c_int, [vp]),
            "orc_vsys_test_flags": (C.c_int, [vp, _ip]),
            "orc_vsys_enforce_topology": (C.c_int, [vp, C.c_int]),
            "orc_vsys_last_needs_fixing": (C.c_int, [vp, vp]),
            "orc_vsys_last_repair_mode": (C.c_int, [vp]),
            "orc_vsys_test_flags_cpu": (C.c_int, [vp, _ip]),
            "orc_vsys_spp_step": (C.c_int, [vp, _dp, C.c_int]),
            "orc_vsys_spp_run": (C.c_int, [vp, _dp, C.c_int, C.c_int, C.POINTER(C.c_double)]),
            "orc_vsys_brownian_step": (C.c_int, [vp, _dp, _dp, C.c_double]),
            "orc_fire_create": (vp, [vp]),
            "orc_fire_destroy": (None, [vp]),
            "orc_fire_set": (None, [vp, C.c_int] + [C.c_double] * 7 + [C.c_int]),
            "orc_fire_minimize": (C.c_int, [vp, vp, _dp]),
            "orc_fire_bowl": (C.c_int, [C.c_int, _dp, _dp, C.c_int] + [C.c_double] * 7 + [C.c_int, _dp, _dp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def orient0(a, b):
    return lib().orc_orient0(a[0], a[1], b[0], b[1])


def incircle0(b, c, d):
    return lib().orc_incircle0(b[0], b[1], c[0], c[1], d[0], d[1])


def circumcenter0(x1, x2):
    out = np.zeros(2)
    lib().orc_circumcenter0(x1[0], x1[1], x2[0], x2[1], out)
    return out


def min_dist(L, p1, p2):
    out = np.zeros(2)
    lib().orc_min_dist(L, p1[0], p1[1], p2[0], p2[1], out)
    return out


def put_in_box(L, x, y):
    cx, cy = C.c_double(x), C.c_double(y)
    lib().orc_put_in_box(L, C.byref(cx), C.byref(cy))
    return cx.value, cy.value


def random_init(N, seed=13377, L=None):
    """Reference initial condition: mt19937(seed), x,y per cell then directors."""
    if L is None:
        L = float(np.sqrt(N))
    x, y, th = np.zeros(N), np.zeros(N), np.zeros(N)
    lib().orc_random_init(N, seed, L, x, y, th)
    return x, y, th, L


def triangulate(x, y, L, cap=32):
    N = len(x)
    nn = np.zeros(N, np.int32)
    nbr = np.zeros(N * cap, np.int32)
    nm = lib().orc_triangulate(N, L, np.ascontiguousarray(x, np.float64), np.ascontiguousarray(y, np.float64), cap, nn, nbr)
    if nm < 0:
        raise RuntimeError("oracle triangulation failed (%d)" % nm)
    return nn, nbr.reshape(N, cap), nm


class VoronoiSystem:
    """FP64 oracle of one periodic Voronoi monolayer (reference voronoiModelBase +
    VoronoiQuadraticEnergy + selfPropelledParticleDynamics semantics)."""

    def __init__(self, x, y, theta, L, A0=1.0, P0=3.8, KA=1.0, KP=1.0, dt=0.01, mu=1.0, v0=0.05, Dr=1.0):
        self.N = len(x)
        self.L = float(L)
        x = np.ascontiguousarray(x, np.float64)
        y = np.ascontiguousarray(y, np.float64)
        th = np.ascontiguousarray(theta, np.float64)
        self.h = lib().orc_vsys_create(self.N, self.L, x, y, th)
        if not self.h:
            raise RuntimeError("oracle: initial triangulation failed")
        lib().orc_vsys_set_params(self.h, KA, KP, dt, mu)
        self.set_prefs(A0, P0)
        self.set_motility(v0, Dr)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_vsys_destroy(self.h)
            self.h = None

    def _full(self, v):
        return np.ascontiguousarray(np.broadcast_to(np.asarray(v, np.float64), (self.N,)))

    def set_prefs(self, A0, P0):
        lib().orc_vsys_set_prefs(self.h, self._full(A0), self._full(P0))

    def set_motility(self, v0, Dr):
        lib().orc_vsys_set_motility(self.h, self._full(v0), self._full(Dr))

    def set_exclusions(self, ex):
        lib().orc_vsys_set_exclusions(self.h, np.ascontiguousarray(ex, np.int32))

    def set_state(self, x, y, theta=None, vx=None, vy=None):
        x = np.ascontiguousarray(x, np.float64)
        y = np.ascontiguousarray(y, np.float64)
        keep = [np.ascontiguousarray(a, np.float64) if a is not None else None for a in (theta, vx, vy)]
        lib().orc_vsys_set_state(self.h, x, y, _ptr(keep[0]), _ptr(keep[1]), _ptr(keep[2]))

    def set_topology(self, nn, nbr, neigh_max):
        nbr = np.ascontiguousarray(nbr, np.int32)
        cap = nbr.shape[1]
        lib().orc_vsys_set_topology(self.h, np.ascontiguousarray(nn, np.int32), nbr.reshape(-1), cap, int(neigh_max))

    def get(self):
        N = self.N
        a = {k: np.zeros(N) for k in ("x", "y", "theta", "vx", "vy", "Fx", "Fy", "A", "P")}
        lib().orc_vsys_get(self.h, *[_ptr(a[k]) for k in ("x", "y", "theta", "vx", "vy", "Fx", "Fy", "A", "P")])
        return a

    def topology(self, cap=32):
        nn = np.zeros(self.N, np.int32)
        nbr = np.zeros(self.N * cap, np.int32)
        nm = lib().orc_vsys_get_topology(self.h, nn, nbr, cap)
        return nn, nbr.reshape(self.N, cap), nm

    def counters(self):
        out = np.zeros(6, np.int64)
        lib().orc_vsys_get_counters(self.h, out)
        return dict(zip(["global_fixes", "local_repairs", "repair_events", "skipped", "last_mode", "last_nf"], out.tolist()))

    def cell_death(self, i):
        """reference voronoiModelBase::cellDeath: remove cell i, retriangulate."""
        if lib().orc_vsys_cell_death(self.h, int(i)) != 0:
            raise RuntimeError("oracle: cell death failed")
        self.N = lib().orc_vsys_num_cells(self.h)

    def cell_division(self, i, theta, frac, daughter_director):
        """reference voronoiModelBase::cellDivision (parameters {i}, dParams {theta, frac});
        the daughter is appended with the given director."""
        if lib().orc_vsys_cell_division(self.h, int(i), float(theta), float(frac), float(daughter_director)) != 0:
            raise RuntimeError("oracle: cell division failed")
        self.N = lib().orc_vsys_num_cells(self.h)

    def set_tension(self, types, gamma=0.0, matrix=None):
        """Line tension between cell types (VoronoiQuadraticEnergyWithTension):
        one value gamma for every unlike pair, or a symmetric matrix g[a, b]."""
        types = np.ascontiguousarray(types, np.int32)
        if matrix is None:
            lib().orc_vsys_set_tension(self.h, 1, float(gamma), 0, None, types)
        else:
            g = np.ascontiguousarray(matrix, np.float64)
            lib().orc_vsys_set_tension(self.h, 2, 0.0, g.shape[0], g.ctypes.data, types)

    def compute_forces(self, gpu_style=True):
        return lib().orc_vsys_compute_forces(self.h, 0 if gpu_style else 1)

    def test_flags(self):
        r = np.zeros(self.N, np.int32)
        any_ = lib().orc_vsys_test_flags(self.h, r)
        return bool(any_), r

    def test_flags_cpu(self):
        """Flags of the reference CPU path's test (window around the circumcenter)."""
        r = np.zeros(self.N, np.int32)
        any_ = lib().orc_vsys_test_flags_cpu(self.h, r)
        return bool(any_), r

    def last_needs_fixing(self):
        """NeedsFixing set (sorted) and repair mode (0 none, 1 local, 2 global) of
        the last enforceTopology."""
        n = lib().orc_vsys_last_needs_fixing(self.h, None)
        out = np.zeros(max(n, 1), np.int32)
        lib().orc_vsys_last_needs_fixing(self.h, _ptr(out))
        return out[:n], lib().orc_vsys_last_repair_mode(self.h)

    def enforce_topology(self, gpu_path=True):
        if lib().orc_vsys_enforce_topology(self.h, 1 if gpu_path else 0) != 0:
            raise RuntimeError("oracle: topology repair failed")

    def spp_step(self, eta, gpu_path=True):
        if lib().orc_vsys_spp_step(self.h, np.ascontiguousarray(eta, np.float64), 1 if gpu_path else 0) != 0:
            raise RuntimeError("oracle: topology repair failed")

    def spp_run(self, eta, gpu_path=True):
        """len(eta) / N self-propelled steps in one C loop (eta[k, i]: noise of
        cell i at step k); returns the loop's steady_clock seconds."""
        eta = np.ascontiguousarray(eta, np.float64).reshape(-1)
        steps = len(eta) // self.N
        sec = C.c_double(0.0)
        if lib().orc_vsys_spp_run(self.h, eta, steps, 1 if gpu_path else 0, C.byref(sec)) != 0:
            raise RuntimeError("oracle: topology repair failed")
        return sec.value

    def brownian_step(self, etax, etay, T):
        r = lib().orc_vsys_brownian_step(self.h, np.ascontiguousarray(etax, np.float64), np.ascontiguousarray(etay, np.float64), T)
        if r != 0:
            raise RuntimeError("oracle: topology repair failed")


def fire_bowl(k, x0, max_it=100, force_cutoff=1e-12, alpha_start=0.1, dt=0.1, dt_max=1.0, dt_inc=1.05,
              dt_dec=0.95, alpha_dec=0.9, n_min=5):
    """FIRE (GPU branch) on the quadratic bowl E = sum (kx x^2 + ky y^2) / 2.
    k, x0: (n, 2). Returns (iterations, final x (n, 2), per-iteration log with
    columns power, dt_before, dt_after, alpha_before, alpha_after,
    nsince_before, nsince_after, zeroed, vv_after, state [forceMax, Power,
    alpha, dt, max |F|^2])."""
    k = np.ascontiguousarray(k, np.float64).reshape(-1)
    x = np.ascontiguousarray(x0, np.float64).reshape(-1).copy()
    lg = np.zeros(9 * max_it)
    out = np.zeros(5)
    it = lib().orc_fire_bowl(len(x) // 2, k, x, max_it, force_cutoff, alpha_start, dt, dt_max, dt_inc, dt_dec,
                             alpha_dec, n_min, lg, out)
    return it, x.reshape(-1, 2), lg[:9 * it].reshape(it, 9), out


class FireMinimizer:
    """Oracle of EnergyMinimizerFIRE (GPU branch)."""

    def __init__(self, sys, max_it=1000, force_cutoff=1e-7, alpha_start=0.99, dt=0.01, dt_max=0.1,
                 dt_inc=1.05, dt_dec=0.95, alpha_dec=0.9, n_min=5):
        self.sys = sys
        self.h = lib().orc_fire_create(sys.h)
        lib().orc_fire_set(self.h, max_it, force_cutoff, alpha_start, dt, dt_max, dt_inc, dt_dec, alpha_dec, n_min)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_fire_destroy(self.h)
            self.h = None

    def minimize(self):
        out = np.zeros(5)
        it = lib().orc_fire_minimize(self.h, self.sys.h, out)
        if it < 0:
            raise RuntimeError("oracle: topology repair failed during FIRE")
        return it, dict(force_max=out[0], power=out[1], alpha=out[2], dt=out[3], max_f2=out[4])


# ----------------------------------------------------------------------------- vertex model
def _vlib():
    L = lib()
    if not getattr(L, "_vm_ready", False):
        vp = C.c_void_p
        sig = {
            "vmo_create": (vp, [C.c_int, C.c_double, _dp, _dp, _ip, _ip, C.c_int]),
            "vmo_destroy": (None, [vp]),
            "vmo_nv": (C.c_int, [vp]),
            "vmo_set_params": (None, [vp] + [C.c_double] * 6),
            "vmo_set_prefs": (None, [vp, _dp, _dp]),
            "vmo_set_positions": (None, [vp, _dp, _dp]),
            "vmo_get_topology": (None, [vp, _ip, _ip, _ip, _ip, C.c_int]),
            "vmo_set_topology": (None, [vp, _ip, _ip, _ip, _ip, C.c_int]),
            "vmo_compute": (C.c_double, [vp] + [vp] * 6),
            "vmo_brownian_step": (C.c_int, [vp, _dp, _dp]),
            "vmo_t1": (C.c_int, [vp]),
            "vmo_last_flagged": (C.c_int, [vp]),
            "vmo_set_motility": (None, [vp, _dp, _dp]),
            "vmo_set_directors": (None, [vp, _dp]),
            "vmo_get_directors": (None, [vp, _dp]),
            "vmo_spp_step": (C.c_int, [vp, _dp]),
            "vmo_fire_set": (None, [vp, C.c_int] + [C.c_double] * 7 + [C.c_int]),
            "vmo_fire_minimize": (C.c_int, [vp, _dp]),
            "vmo_cell_death": (C.c_int, [vp, C.c_int]),
            "vmo_cell_division": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_double]),
            "vmo_nc": (C.c_int, [vp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        L._vm_ready = True
    return L


class VertexSystem:
    """FP64 oracle of the reference vertex model (GPU semantics, deterministic T1 rounds)."""

    CAP = 32

    def __init__(self, x, y, L, nn, nbr, KA=1.0, KP=1.0, dt=0.005, mu=1.0, T=0.0, thr=0.04, A0=1.0, P0=3.9):
        self.Nc = len(x)
        self.L = float(L)
        nbr = np.ascontiguousarray(nbr, np.int32)
        self.h = _vlib().vmo_create(self.Nc, self.L, np.ascontiguousarray(x, np.float64),
                                    np.ascontiguousarray(y, np.float64), np.ascontiguousarray(nn, np.int32),
                                    nbr.reshape(-1), nbr.shape[1])
        self.Nv = _vlib().vmo_nv(self.h)
        _vlib().vmo_set_params(self.h, KA, KP, dt, mu, T, thr)
        self.set_prefs(A0, P0)

    def __del__(self):
        if getattr(self, "h", None):
            _vlib().vmo_destroy(self.h)
            self.h = None

    def set_params(self, KA=1.0, KP=1.0, dt=0.005, mu=1.0, T=0.0, thr=0.04):
        _vlib().vmo_set_params(self.h, KA, KP, dt, mu, T, thr)

    def set_prefs(self, A0, P0):
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(A0, np.float64), (self.Nc,)))
        p = np.ascontiguousarray(np.broadcast_to(np.asarray(P0, np.float64), (self.Nc,)))
        _vlib().vmo_set_prefs(self.h, a, p)

    def set_positions(self, vx, vy):
        _vlib().vmo_set_positions(self.h, np.ascontiguousarray(vx, np.float64), np.ascontiguousarray(vy, np.float64))

    def topology(self):
        vn = np.zeros(3 * self.Nv, np.int32)
        vcn = np.zeros(3 * self.Nv, np.int32)
        cvn = np.zeros(self.Nc, np.int32)
        cv = np.full(self.Nc * self.CAP, -1, np.int32)
        _vlib().vmo_get_topology(self.h, vn, vcn, cvn, cv, self.CAP)
        return vn, vcn, cvn, cv.reshape(self.Nc, self.CAP)

    def set_topology(self, vn, vcn, cvn, cv):
        cv = np.ascontiguousarray(cv, np.int32)
        _vlib().vmo_set_topology(self.h, np.ascontiguousarray(vn, np.int32), np.ascontiguousarray(vcn, np.int32),
                                 np.ascontiguousarray(cvn, np.int32), cv.reshape(-1), cv.shape[1])

    def compute(self):
        a = {k: np.zeros(self.Nv) for k in ("vx", "vy", "Fx", "Fy")}
        a.update({k: np.zeros(self.Nc) for k in ("A", "P")})
        E = _vlib().vmo_compute(self.h, *[_ptr(a[k]) for k in ("vx", "vy", "Fx", "Fy", "A", "P")])
        a["E"] = E
        return a

    def brownian_step(self, etax, etay):
        return _vlib().vmo_brownian_step(self.h, np.ascontiguousarray(etax, np.float64),
                                         np.ascontiguousarray(etay, np.float64))

    def t1(self):
        return _vlib().vmo_t1(self.h)

    def last_flagged(self):
        return _vlib().vmo_last_flagged(self.h)

    def cell_death(self, c):
        """reference vertexModelBase::cellDeath (T2 of a triangular cell)."""
        if _vlib().vmo_cell_death(self.h, int(c)) != 0:
            raise RuntimeError("oracle: cell death needs a simple triangular cell")
        self.Nv, self.Nc = _vlib().vmo_nv(self.h), _vlib().vmo_nc(self.h)

    def cell_division(self, c, p1, p2, daughter_director=0.0):
        """reference vertexModelBase::cellDivision (parameters c, p1, p2)."""
        if _vlib().vmo_cell_division(self.h, int(c), int(p1), int(p2), float(daughter_director)) != 0:
            raise RuntimeError("oracle: cell division failed")
        self.Nv, self.Nc = _vlib().vmo_nv(self.h), _vlib().vmo_nc(self.h)

    def set_motility(self, v0, Dr):
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(v0, np.float64), (self.Nc,)))
        b = np.ascontiguousarray(np.broadcast_to(np.asarray(Dr, np.float64), (self.Nc,)))
        _vlib().vmo_set_motility(self.h, a, b)

    def set_directors(self, theta):
        _vlib().vmo_set_directors(self.h, np.ascontiguousarray(theta, np.float64))

    def directors(self):
        out = np.zeros(self.Nc)
        _vlib().vmo_get_directors(self.h, out)
        return out

    def spp_step(self, eta_cell):
        """One self-propelled cell-vertex step (reference selfPropelledCellVertexDynamics,
        GPU branch); eta_cell: the N(0,1) draw of every cell's director."""
        return _vlib().vmo_spp_step(self.h, np.ascontiguousarray(eta_cell, np.float64))

    def fire(self, max_it=1000, force_cutoff=1e-7, alpha_start=0.99, dt=0.01, dt_max=0.1, dt_inc=1.05,
             dt_dec=0.95, alpha_dec=0.9, n_min=5):
        """FIRE (GPU branch) on the vertices; returns (iterations, state). The
        iteration count accumulates over calls, as in the reference."""
        _vlib().vmo_fire_set(self.h, max_it, force_cutoff, alpha_start, dt, dt_max, dt_inc, dt_dec, alpha_dec,
                             n_min)
        out = np.zeros(5)
        it = _vlib().vmo_fire_minimize(self.h, out)
        return it, dict(force_max=out[0], power=out[1], alpha=out[2], dt=out[3], max_f2=out[4])


def vertex_from_random(Nc, seed=13377, **kw):
    x, y, th, L = random_init(Nc, seed)
    nn, nbr, _ = triangulate(x, y, L, cap=32)
    return VertexSystem(x, y, L, nn, nbr, **kw), (x, y, L, nn, nbr)


# ------------------------------------------------------------ spatial sorting
def hilbert_curve_order(L):
    """Curve order of the reference HilbertSorter for a square box of side L
    (inc/utility/HilbertSort.h:33-44): the smallest mm >= 1 with 2^mm >= L,
    order min(30, mm + 4)."""
    mm, t = 1, 2
    while float(t) < L:
        t *= 2
        mm += 1
    return min(30, mm + 4)


def hilbert_xy2d(M, x, y):
    """Index along the order-M Hilbert curve of pixel (x, y), the iterative
    rotate-and-reflect construction the reference calls (xy2d in
    HilbertSort.h:95-101); vectorised over integer arrays."""
    x = np.asarray(x, np.int64).copy()
    y = np.asarray(y, np.int64).copy()
    d = np.zeros(np.broadcast(x, y).shape, np.int64)
    s = 1 << (M - 1)
    while s > 0:
        rx = ((x & s) > 0).astype(np.int64)
        ry = ((y & s) > 0).astype(np.int64)
        d += s * s * ((3 * rx) ^ ry)
        flip = (ry == 0) & (rx == 1)
        x = np.where(flip, s - 1 - x, x)
        y = np.where(flip, s - 1 - y, y)
        swap = ry == 0
        x, y = np.where(swap, y, x), np.where(swap, x, y)
        s >>= 1
    return d


def hilbert_keys(x, y, L):
    """HilbertSorter::getIdx (HilbertSort.h:88-101) of points in a square box:
    the general-box inverse transform (1 / (L L)) L times the coordinate
    (gpubox.h:86-101, 117-129), then floor(2^M v)."""
    M = hilbert_curve_order(L)
    inv = (1.0 / (L * L)) * L
    n = float(1 << M)
    px = np.floor(n * (inv * np.asarray(x, np.float64))).astype(np.int64)
    py = np.floor(n * (inv * np.asarray(y, np.float64))).astype(np.int64)
    return hilbert_xy2d(M, px, py)


def hilbert_sort_order(x, y, L):
    """new -> old order of Simple2DCell::spatiallySortCells
    (Simple2DCell.cpp:438-477): std::sort of (index, i) pairs."""
    k = hilbert_keys(x, y, L)
    return np.lexsort((np.arange(len(k)), k)).astype(np.int32)


def vertex_sort_orders(vx, vy, vcn, L):
    """Simple2DCell::spatiallySortVertices (Simple2DCell.cpp:484-599): vertices
    by curve index, then cells numbered by first appearance walking the sorted
    vertices and each vertex's three cells (vcn, 3 per vertex, old indices).
    Returns (vertex new -> old, cell new -> old)."""
    vperm = hilbert_sort_order(vx, vy, L)
    vcn = np.asarray(vcn).reshape(-1, 3)
    order, seen = [], set()
    for v in vperm:
        for c in vcn[v]:
            if int(c) not in seen:
                seen.add(int(c))
                order.append(int(c))
    return vperm, np.asarray(order, np.int32)
