"""NetCDF trajectory databases (reference SPVDatabaseNetCDF / AVMDatabaseNetCDF
layouts) through the C ABI, on the CPU: files written by the library are read by
scipy's independent NetCDF classic reader (dimensions, variables, types, values),
files written by scipy in the same layout are read by the library, records
append in order, and misuse is reported."""
import ctypes as C
import os

import numpy as np
import pytest
from scipy.io import netcdf_file

SO = os.path.join(os.path.dirname(__file__), "..", "cellgpu_amd", "libcellgpu_amd.so")


@pytest.fixture(scope="module")
def L():
    if not os.path.exists(SO):
        pytest.skip("libcellgpu_amd.so not built")
    from cellgpu_amd import _lib
    return _lib.lib()


def _p(a):
    return None if a is None else a.ctypes.data


def _spv_state(rng, n):
    return dict(x=rng.uniform(0, 10, n), y=rng.uniform(0, 10, n), d=rng.uniform(0, 6.28, n),
                t=rng.integers(-5, 3, n).astype(np.int32), ex=rng.standard_normal(n), ey=rng.standard_normal(n))


@pytest.mark.parametrize("excl", [0, 1])
def test_spv_layout_roundtrip(L, tmp_path, excl):
    n = 37
    path = str(tmp_path / "spv.nc").encode()
    db = L.cgpu_db_open(path, 0, n, 0, excl)
    assert db
    rng = np.random.default_rng(1)
    states = []
    for r in range(3):
        s = _spv_state(rng, n)
        states.append(s)
        assert L.cgpu_db_spv_write(db, -1, 0.5 * r, 10.0, 3.8 + r, _p(s["x"]), _p(s["y"]), _p(s["d"]), _p(s["t"]),
                                   _p(s["ex"]), _p(s["ey"])) == 0
    assert L.cgpu_db_records(db) == 3
    L.cgpu_db_close(db)
    with netcdf_file(path.decode(), "r", mmap=False) as f:
        assert f.dimensions == {"rec": None, "Nv": n, "dof": 2 * n, "boxdim": 4, "unit": 1}
        names = ["time", "means0", "pos", "type", "director", "BoxMatrix"] + (["externalForce"] if excl else [])
        assert list(f.variables) == names
        assert f.variables["pos"].shape == (3, 2 * n) and f.variables["pos"].typecode() == "d"
        assert f.variables["type"].typecode() == "i"
        for r, s in enumerate(states):
            assert f.variables["time"][r, 0] == 0.5 * r and f.variables["means0"][r, 0] == 3.8 + r
            pos = f.variables["pos"][r]
            assert np.array_equal(pos[0::2], s["x"]) and np.array_equal(pos[1::2], s["y"])
            assert np.array_equal(f.variables["type"][r], s["t"])
            assert np.array_equal(f.variables["director"][r], s["d"])
            assert np.array_equal(f.variables["BoxMatrix"][r], [10.0, 0.0, 0.0, 10.0])
            if excl:
                ef = f.variables["externalForce"][r]
                assert np.array_equal(ef[0::2], s["ex"]) and np.array_equal(ef[1::2], s["ey"])
    # append mode, then read back through the library
    db = L.cgpu_db_open(path, 0, n, 1, 0)
    s = _spv_state(rng, n)
    assert L.cgpu_db_spv_write(db, -1, 9.0, 10.0, 4.0, _p(s["x"]), _p(s["y"]), _p(s["d"]), _p(s["t"]), _p(s["ex"]),
                               _p(s["ey"])) == 0
    L.cgpu_db_close(db)
    db = L.cgpu_db_open(path, 0, n, 2, 0)
    assert L.cgpu_db_records(db) == 4
    out = {k: np.zeros(n) for k in ("x", "y", "d", "ex", "ey")}
    t = np.zeros(n, np.int32)
    tm, box = np.zeros(1), np.zeros(4)
    for r, s in enumerate(states + [s]):
        assert L.cgpu_db_spv_read(db, r, _p(tm), _p(box), _p(out["x"]), _p(out["y"]), _p(out["d"]), _p(t),
                                  _p(out["ex"]), _p(out["ey"])) == 0
        assert np.array_equal(out["x"], s["x"]) and np.array_equal(out["d"], s["d"]) and np.array_equal(t, s["t"])
        if excl:
            assert np.array_equal(out["ex"], s["ex"])
    # read-only files refuse writes; a different size is refused at open
    assert L.cgpu_db_spv_write(db, -1, 0.0, 10.0, 0.0, _p(s["x"]), _p(s["y"]), None, None, None, None) != 0
    L.cgpu_db_close(db)
    assert not L.cgpu_db_open(path, 0, n + 1, 2, 0)
    assert b"different number" in L.cgpu_db_last_error()


def test_spv_reads_scipy_written_file(L, tmp_path):
    """A file in the reference layout written by another NetCDF encoder (scipy,
    CDF-1 with attributes) is read by the library."""
    n = 20
    path = str(tmp_path / "other.nc")
    rng = np.random.default_rng(2)
    x, y, d = rng.uniform(0, 5, n), rng.uniform(0, 5, n), rng.uniform(0, 6, n)
    t = rng.integers(0, 4, n).astype(np.int32)
    with netcdf_file(path, "w", version=1) as f:
        f.history = "written by another encoder"
        f.createDimension("rec", None)
        f.createDimension("Nv", n)
        f.createDimension("dof", 2 * n)
        f.createDimension("boxdim", 4)
        f.createDimension("unit", 1)
        v = {}
        for name, typ, dims in [("time", "d", ("rec", "unit")), ("means0", "d", ("rec", "unit")),
                                ("pos", "d", ("rec", "dof")), ("type", "i", ("rec", "Nv")),
                                ("director", "d", ("rec", "Nv")), ("BoxMatrix", "d", ("rec", "boxdim"))]:
            v[name] = f.createVariable(name, typ, dims)
        v["pos"].units = "box lengths"
        for r in range(2):
            v["time"][r] = [r + 0.25]
            v["means0"][r] = [3.9]
            pos = np.empty(2 * n)
            pos[0::2], pos[1::2] = x + r, y
            v["pos"][r] = pos
            v["type"][r] = t
            v["director"][r] = d
            v["BoxMatrix"][r] = [5.0, 0.0, 0.0, 5.0]
    db = L.cgpu_db_open(path.encode(), 0, n, 2, 0)
    assert db, L.cgpu_db_last_error()
    assert L.cgpu_db_records(db) == 2
    ox, oy, od = np.zeros(n), np.zeros(n), np.zeros(n)
    ot = np.zeros(n, np.int32)
    tm, box = np.zeros(1), np.zeros(4)
    assert L.cgpu_db_spv_read(db, 1, _p(tm), _p(box), _p(ox), _p(oy), _p(od), _p(ot), None, None) == 0
    assert tm[0] == 1.25 and np.array_equal(box, [5.0, 0.0, 0.0, 5.0])
    assert np.array_equal(ox, x + 1) and np.array_equal(oy, y) and np.array_equal(od, d) and np.array_equal(ot, t)
    L.cgpu_db_close(db)


def test_avm_layout_roundtrip(L, tmp_path):
    nv = 24
    nc = nv // 2
    path = str(tmp_path / "avm.nc").encode()
    db = L.cgpu_db_open(path, 1, nv, 0, 0)
    rng = np.random.default_rng(3)
    vx, vy, fx, fy = (rng.standard_normal(nv) for _ in range(4))
    vn = rng.integers(0, nv, 3 * nv).astype(np.int32)
    vcn = rng.integers(0, nc, 3 * nv).astype(np.int32)
    ct = rng.integers(0, 3, nc).astype(np.int32)
    d, cx, cy = rng.uniform(0, 6, nc), rng.uniform(0, 4, nc), rng.uniform(0, 4, nc)
    for r in range(2):
        assert L.cgpu_db_avm_write(db, -1, 0.1 * r, 4.0, 3.9, _p(vx), _p(vy), _p(fx), _p(fy), _p(vn), _p(vcn), _p(ct),
                                   _p(d), _p(cx), _p(cy)) == 0
    L.cgpu_db_close(db)
    with netcdf_file(path.decode(), "r", mmap=False) as f:
        assert f.dimensions == {"rec": None, "Nv": nv, "Nc": nc, "Nc2": 2 * nc, "dof": 2 * nv, "Nvn": 3 * nv,
                                "boxdim": 4, "unit": 1}
        assert list(f.variables) == ["pos", "force", "VertexCellNeighbors", "Vneighs", "cellType", "director",
                                     "cellPositions", "BoxMatrix", "meanQ", "time"]
        assert f.variables["Vneighs"].typecode() == "i" and f.variables["cellType"].typecode() == "d"
        assert np.array_equal(f.variables["Vneighs"][1], vn) and np.array_equal(f.variables["VertexCellNeighbors"][1], vcn)
        assert np.array_equal(f.variables["cellType"][1], ct.astype(float))
        assert np.array_equal(f.variables["pos"][1][0::2], vx) and np.array_equal(f.variables["force"][1][1::2], fy)
        assert np.array_equal(f.variables["cellPositions"][0][0::2], cx)
        assert f.variables["time"][1, 0] == 0.1 and f.variables["meanQ"][0, 0] == 3.9
    db = L.cgpu_db_open(path, 1, nv, 2, 0)
    ovx, ovy, ofx, ofy = (np.zeros(nv) for _ in range(4))
    ovn, ovcn = np.zeros(3 * nv, np.int32), np.zeros(3 * nv, np.int32)
    od = np.zeros(nc)
    assert L.cgpu_db_avm_read(db, 1, None, None, _p(ovx), _p(ovy), _p(ofx), _p(ofy), _p(ovn), _p(ovcn), _p(od)) == 0
    assert np.array_equal(ovx, vx) and np.array_equal(ofy, fy) and np.array_equal(ovn, vn)
    assert np.array_equal(ovcn, vcn) and np.array_equal(od, d)
    assert L.cgpu_db_avm_read(db, 2, None, None, _p(ovx), None, None, None, None, None, None) != 0  # no record 2
    L.cgpu_db_close(db)
