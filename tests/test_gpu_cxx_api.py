"""The C++ drop-in layer (include/cellgpu_amd.hpp) against the oracle.

cellgpu_amd/bin/api_check drives the models through the reference's class API
(abaso/cellGPU Simulation, VoronoiQuadraticEnergy, VertexQuadraticEnergy,
ArrayHandle, EnergyMinimizerFIRE) and dumps what it read through host array
handles. Bars: Delaunay rings and vertex topology bit-exact, FP64 forces within
1e-12 of the largest component, sorted and unsorted runs equal cell by cell
(tag order) within 1e-12 of the box, FIRE converged below its cutoff. The
reference drivers' command lines (voronoi.cpp, Vertex.cpp) are run too.
"""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "..", "cellgpu_amd", "bin")


def _run(args, timeout=120):
    exe = os.path.join(BIN, args[0])
    if not os.path.exists(exe):
        pytest.fail("%s is not built (run __graft_entry__.build())" % exe)
    print("running", " ".join(args), flush=True)
    r = subprocess.run([exe] + args[1:], capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, (r.stdout, r.stderr)
    return r.stdout


@pytest.fixture(scope="module")
def dumps(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("api") / "api_check.txt")
    log = _run(["api_check", out])
    assert "api_check done" in log, log
    d = {}
    for line in open(out):
        parts = line.split()
        name, n = parts[0], int(parts[1])
        vals = parts[2:]
        assert len(vals) == n, name
        d[name] = np.array([float(v) for v in vals])
    return d


def _xy(a):
    return a[0::2].copy(), a[1::2].copy()


def _rings(nn, nbr):
    return [tuple(int(v) for v in nbr[i, : int(nn[i])]) for i in range(len(nn))]


def _canon(cv, cvn):
    out = []
    for c in range(len(cvn)):
        l = [int(v) for v in cv[c, : int(cvn[c])]]
        k = int(np.argmin(l))
        out.append(tuple(l[k:] + l[:k]))
    return out


def _sets(a):
    a = a.astype(int)
    return [frozenset(a[3 * v: 3 * v + 3]) for v in range(len(a) // 3)]


def _close_forces(gx, gy, r):
    scale = np.abs(np.concatenate([r["Fx"], r["Fy"]])).max()
    assert np.abs(gx - r["Fx"]).max() <= 1e-12 * scale
    assert np.abs(gy - r["Fy"]).max() <= 1e-12 * scale


def test_voronoi_model_through_the_class_api(dumps, oracle):
    N = 300
    x0, y0 = _xy(dumps["v_pos0"])
    ox, oy, th, L = oracle.random_init(N, 13377)
    assert np.array_equal(x0, ox) and np.array_equal(y0, oy)
    o = oracle.VoronoiSystem(ox, oy, th, L, P0=3.75)
    E = o.compute_forces(gpu_style=True)
    r = o.get()
    fx, fy = _xy(dumps["v_force0"])
    _close_forces(fx, fy, r)
    A, P = _xy(dumps["v_ap0"])
    np.testing.assert_allclose(A, r["A"], rtol=1e-12)
    np.testing.assert_allclose(P, r["P"], rtol=1e-12)
    assert abs(dumps["v_energy0"][0] - E) <= 1e-12 * abs(E)

    # moveDegreesOfFreedom: x + d wrapped into the box, on the host, bit-exact
    dx, dy = _xy(dumps["v_disp"])
    x1, y1 = _xy(dumps["v_pos1"])
    for i in range(N):
        wx, wy = oracle.put_in_box(L, x0[i] + dx[i], y0[i] + dy[i])
        assert (wx, wy) == (x1[i], y1[i]), i
    # enforceTopology: the GPU-path test and repair of the old rings at the new positions
    o.set_state(x1, y1)
    o.enforce_topology(gpu_path=True)
    onn, onbr, onm = o.topology(32)
    nm = int(dumps["v_nm1"][0])
    assert nm == onm
    nn1 = dumps["v_nn1"].astype(int)
    nbr1 = dumps["v_nbr1"].astype(int).reshape(N, nm)
    assert _rings(nn1, nbr1) == _rings(onn, onbr)
    o.compute_forces(gpu_style=True)
    f1x, f1y = _xy(dumps["v_force1"])
    _close_forces(f1x, f1y, o.get())

    # positions written through an ArrayHandle: stored back and retriangulated
    x2, y2 = _xy(dumps["v_pos2"])
    assert np.array_equal(x2[10:], x1[10:]) and not np.array_equal(x2[:10], x1[:10])
    tnn, tnbr, tnm = oracle.triangulate(x2, y2, L, cap=32)
    nm2 = int(dumps["v_nm2"][0])
    assert nm2 == tnm
    assert _rings(dumps["v_nn2"].astype(int), dumps["v_nbr2"].astype(int).reshape(N, nm2)) == _rings(tnn, tnbr)


def test_simulation_sort_period_relabels_consistently(dumps):
    N = 200
    tags = dumps["s_tags"].astype(int)
    itt = dumps["s_itt"].astype(int)
    assert sorted(tags) == list(range(N)) and not np.array_equal(tags, np.arange(N))
    assert sorted(itt) == list(range(N))
    xs, ys = _xy(dumps["s_pos_sorted"])
    xp, yp = _xy(dumps["s_pos_plain"])
    L = np.sqrt(N)
    # cell i of the sorted system is cell tags[i] of the unsorted one
    ddx = xs - xp[tags]
    ddy = ys - yp[tags]
    ddx -= L * np.round(ddx / L)
    ddy -= L * np.round(ddy / L)
    assert np.abs(ddx).max() < 1e-12 * L and np.abs(ddy).max() < 1e-12 * L
    assert abs(dumps["s_time"][0] - 0.07) < 1e-12 and dumps["s_model_time"][0] == dumps["s_time"][0]


def test_fire_through_the_simulation(dumps):
    fm = dumps["f_forcemax"][0]
    assert np.sqrt(fm) < 1e-8 and dumps["f_iterations"][0] < 20000
    assert dumps["f_e1"][0] < dumps["f_e0"][0]
    fx, fy = _xy(dumps["f_force"])
    assert np.sqrt(np.mean(fx ** 2 + fy ** 2)) < 1e-8


def test_vertex_model_through_the_class_api(dumps, oracle):
    Nc = 64
    o, _ = oracle.vertex_from_random(Nc, 5, thr=0.04)
    Nv = int(dumps["m_nv"][0])
    vmax = int(dumps["m_vmax"][0])
    assert Nv == o.Nv == 2 * Nc
    ovn, ovcn, ocvn, ocv = o.topology()
    assert _sets(dumps["m_vn0"]) == _sets(ovn) and _sets(dumps["m_vcn0"]) == _sets(ovcn)
    cv0 = dumps["m_cv0"].astype(int).reshape(Nc, vmax)
    assert _canon(cv0, dumps["m_cvn0"]) == _canon(ocv, ocvn)
    o.set_prefs(1.0, 3.9)
    r = o.compute()
    vx, vy = _xy(dumps["m_pos0"])
    assert np.array_equal(vx, r["vx"]) and np.array_equal(vy, r["vy"])
    fx, fy = _xy(dumps["m_force0"])
    _close_forces(fx, fy, r)
    A, P = _xy(dumps["m_ap0"])
    np.testing.assert_allclose(A, r["A"], rtol=1e-12)
    np.testing.assert_allclose(P, r["P"], rtol=1e-12)
    # cell positions: inside the box
    cx, cy = _xy(dumps["m_cellpos0"])
    L = np.sqrt(Nc)
    assert (cx >= 0).all() and (cx < L).all() and (cy >= 0).all() and (cy < L).all()

    # move + T1 enforcement against the oracle's T1 rounds at the same positions
    # (moveDegreesOfFreedom: x + d wrapped into the box, bit-exact)
    dx, dy = _xy(dumps["m_disp"])
    xm, ym = _xy(dumps["m_pos_moved"])
    for i in range(Nv):
        assert oracle.put_in_box(L, vx[i] + dx[i], vy[i] + dy[i]) == (xm[i], ym[i]), i
    o.set_topology(dumps["m_vn0"].astype(np.int32), dumps["m_vcn0"].astype(np.int32),
                   dumps["m_cvn0"].astype(np.int32), cv0.astype(np.int32))
    o.set_positions(xm, ym)
    assert o.t1() > 0
    r = o.compute()
    # the flips rotate their edges' vertices
    x1, y1 = _xy(dumps["m_pos1"])
    assert np.abs(x1 - r["vx"]).max() < 1e-12 * L and np.abs(y1 - r["vy"]).max() < 1e-12 * L
    ovn, ovcn, ocvn, ocv = o.topology()
    assert _sets(dumps["m_vn1"]) == _sets(ovn) and _sets(dumps["m_vcn1"]) == _sets(ovcn)
    cv1 = dumps["m_cv1"].astype(int).reshape(Nc, vmax)
    assert _canon(cv1, dumps["m_cvn1"]) == _canon(ocv, ocvn)
    r = o.compute()
    f1x, f1y = _xy(dumps["m_force1"])
    _close_forces(f1x, f1y, r)
    assert abs(dumps["m_energy1"][0] - r["E"]) <= 1e-12 * abs(r["E"])

    # one Brownian step at T = 0 through the Simulation: the oracle's step with zero noise
    o.set_params(dt=0.005, thr=0.04)
    o.set_topology(dumps["m_vn1"].astype(np.int32), dumps["m_vcn1"].astype(np.int32),
                   dumps["m_cvn1"].astype(np.int32), cv1.astype(np.int32))
    o.brownian_step(np.zeros(Nv), np.zeros(Nv))
    r = o.compute()
    x2, y2 = _xy(dumps["m_pos2"])
    assert np.abs(x2 - r["vx"]).max() < 1e-11 and np.abs(y2 - r["vy"]).max() < 1e-11
    ovn, ovcn, ocvn, ocv = o.topology()
    assert _canon(dumps["m_cv2"].astype(int).reshape(Nc, vmax), dumps["m_cvn2"]) == _canon(ocv, ocvn)


def test_reference_driver_command_lines():
    """voronoi.cpp and Vertex.cpp flags: a few steps each, GPU 0."""
    out = _run(["voronoi_driver", "-n", "500", "-t", "20", "-i", "20", "-g", "0"])
    assert "current q" in out and "timestep" in out
    out = _run(["vertex_driver", "-n", "200", "-t", "20", "-i", "20", "-g", "0"])
    assert "Mean q" in out
    out = _run(["vertex_driver", "-n", "200", "-t", "20", "-i", "20", "-g", "0", "-z", "1"])
    assert "Mean q" in out
