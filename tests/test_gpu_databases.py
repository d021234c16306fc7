"""NetCDF databases written from and read into device batches (reference
SPVDatabaseNetCDF / AVMDatabaseNetCDF WriteState / ReadState)."""
import numpy as np
import pytest
from scipy.io import netcdf_file

pytestmark = pytest.mark.gpu

CAP = 24


def _cg():
    import cellgpu_amd as cg
    cg.require_device()
    return cg


def test_spv_write_and_read_state(tmp_path):
    """A sorted two-system FP64 batch: system 1 written in tag order (types,
    excluded cells as type - 5, means0 = mean q), read into a fresh batch: same
    positions, directors and types, and the same forces after the read."""
    cg = _cg()
    from cellgpu_amd.databases import SPVDatabaseNetCDF, NcFile
    n = [200, 300]
    b = cg.VoronoiBatch(n, seeds=[3, 4], precision="fp64")
    b.set_prefs(1.0, 3.8)
    b.set_motility(0.1, 1.0)
    b.spv_steps(20, use_graph=False)
    perm = b.spatial_sort()
    types = (np.arange(500) % 3).astype(np.int32)[perm]
    b.set_cell_types(types)
    b.spv_steps(5, use_graph=False)
    b.compute_forces()
    g = b.get()
    tags = b.tags()
    sl = b.system_slice(1)
    path = str(tmp_path / "spv.nc")
    db = SPVDatabaseNetCDF(300, path, NcFile.Replace)
    db.WriteState(b, time=1.5, system=1)
    db.WriteState(b, system=1)
    assert db.Records == 2
    rec = db.read_record(0)
    order = np.argsort(tags[sl])  # index of each tag
    assert np.array_equal(rec["x"], g["x"][sl][order]) and np.array_equal(rec["y"], g["y"][sl][order])
    assert np.array_equal(rec["director"], g["theta"][sl][order])
    assert np.array_equal(rec["type"], types[sl][order])
    assert rec["time"] == 1.5 and np.array_equal(rec["box"], [b.L[1], 0, 0, b.L[1]])
    assert db.read_record(1)["time"] == pytest.approx(25 * 0.01)
    db.close()
    with netcdf_file(path, "r", mmap=False) as f:
        q = np.mean(g["P"][sl] / np.sqrt(g["A"][sl]))
        assert abs(f.variables["means0"][0, 0] - q) < 1e-12 * q
    fresh = cg.VoronoiBatch(n, seeds=[3, 4], precision="fp64")
    fresh.set_prefs(1.0, 3.8)
    db = SPVDatabaseNetCDF(300, path, NcFile.ReadOnly)
    db.ReadState(fresh, 0, system=1)
    fresh.compute_forces()
    h = fresh.get()
    sl2 = fresh.system_slice(1)
    assert np.array_equal(h["x"][sl2], rec["x"]) and np.array_equal(h["theta"][sl2], rec["director"])
    F = np.stack([g["Fx"][sl][order], g["Fy"][sl][order]], 1)
    F2 = np.stack([h["Fx"][sl2], h["Fy"][sl2]], 1)
    assert np.abs(F - F2).max() <= 1e-12 * np.abs(F).max()
    fresh.check_status()


def test_avm_write_and_read_state(tmp_path):
    """A vertex model after Brownian steps with T1 flips, written and read into
    a fresh batch: positions, directors and the topology (up to the rotation of
    each cell's list) come back, and so do the forces."""
    cg = _cg()
    from cellgpu_amd.databases import AVMDatabaseNetCDF, NcFile
    Nc = 300
    b = cg.VertexBatch([Nc], seeds=[9], precision="fp64")
    b.set_params(dt=0.005, T=2e-3, t1_threshold=0.04)
    b.set_prefs(1.0, 3.9)
    b.brownian_steps(60, use_graph=False)
    assert b.counters()["t1_events"] > 0
    b.compute_forces()
    g = b.get()
    vn, vcn, cvn, cv = b.topology()
    path = str(tmp_path / "avm.nc")
    with AVMDatabaseNetCDF(b.Nv, path, NcFile.Replace) as db:
        db.WriteState(b)
    with netcdf_file(path, "r", mmap=False) as f:
        assert np.array_equal(f.variables["Vneighs"][0], vn) and np.array_equal(f.variables["VertexCellNeighbors"][0], vcn)
        assert np.array_equal(f.variables["pos"][0][0::2], g["vx"])
        cp = f.variables["cellPositions"][0]
        assert np.all((cp >= 0) & (cp < b.L[0]))
    fresh = cg.VertexBatch([Nc], seeds=[9], precision="fp64")
    fresh.set_prefs(1.0, 3.9)
    with AVMDatabaseNetCDF(b.Nv, path, NcFile.ReadOnly) as db:
        db.ReadState(fresh, 0)
    h = fresh.get()
    fvn, fvcn, fcvn, fcv = fresh.topology()
    assert np.array_equal(h["vx"], g["vx"]) and np.array_equal(fresh.directors(), b.directors())
    assert np.array_equal(fvcn, vcn) and np.array_equal(fvn, vn)
    canon = lambda cv_, n_: [tuple(np.roll(cv_[c, :n_[c]], -int(np.argmin(cv_[c, :n_[c]])))) for c in range(Nc)]
    assert canon(fcv, fcvn) == canon(cv, cvn)
    F = np.concatenate([g["Fx"], g["Fy"]])
    F2 = np.concatenate([h["Fx"], h["Fy"]])
    assert np.abs(F - F2).max() <= 1e-12 * np.abs(F).max()
    fresh.brownian_steps(5, use_graph=False)
    fresh.check_status()
