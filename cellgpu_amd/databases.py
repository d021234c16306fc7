"""Trajectory databases in the reference's NetCDF layouts (reference
inc/databases/DatabaseNetCDFSPV.h, DatabaseNetCDFAVM.h), written and read by the
library's NetCDF classic encoder: files that the reference's analysis scripts
(and any NetCDF reader) open.

    db = SPVDatabaseNetCDF(N, "traj.nc", NcFile.Replace)
    db.WriteState(model)            # one record: time, means0, pos, type, director, BoxMatrix
    db.ReadState(model, rec)        # positions, directors, types; retriangulates
"""
import numpy as np

from ._lib import CellGPUError, check, lib


class NcFile:
    """The reference's NcFile::FileMode values."""

    Replace, Write, ReadOnly, New = 0, 1, 2, 3


def _target(model, system):
    batch = getattr(model, "batch", model)
    return batch, int(system)


class _Database:
    kind = None

    def __init__(self, n, fn, mode=NcFile.ReadOnly, exclude=False):
        self.filename = str(fn)
        self.Mode = int(mode)
        self.n = int(n)
        self.h = lib().cgpu_db_open(self.filename.encode(), self.kind, int(n), self.Mode, 1 if exclude else 0)
        if not self.h:
            raise CellGPUError("cannot open database %s: %s" % (self.filename, lib().cgpu_db_last_error().decode()))

    def close(self):
        if getattr(self, "h", None):
            lib().cgpu_db_close(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def Records(self):
        return lib().cgpu_db_records(self.h)

    def _check(self, r, what):
        if r != 0:
            msg = lib().cgpu_db_last_error().decode()
            check(r, "%s (%s)" % (what, msg) if msg else what)


class SPVDatabaseNetCDF(_Database):
    """reference SPVDatabaseNetCDF (DatabaseNetCDFSPV.cpp): n cells."""

    kind = 0

    def WriteState(self, model, time=-1.0, rec=-1, system=0):
        batch, s = _target(model, system)
        self._check(lib().cgpu_db_spv_write_batch(self.h, batch.h, s, float(time), int(rec)), "WriteState")

    def ReadState(self, model, rec, geometry=True, system=0):
        batch, s = _target(model, system)
        self._check(lib().cgpu_db_spv_read_batch(self.h, batch.h, s, int(rec)), "ReadState")

    def read_record(self, rec):
        """The raw record (time, box, positions, directors, types) as arrays in tag order."""
        n = self.n
        t, box = np.zeros(1), np.zeros(4)
        x, y, d = np.zeros(n), np.zeros(n), np.zeros(n)
        ty = np.zeros(n, np.int32)
        p = [a.ctypes.data for a in (t, box, x, y, d, ty)]
        self._check(lib().cgpu_db_spv_read(self.h, int(rec), *p, None, None), "read")
        return dict(time=t[0], box=box, x=x, y=y, director=d, type=ty)


class AVMDatabaseNetCDF(_Database):
    """reference AVMDatabaseNetCDF (DatabaseNetCDFAVM.cpp): nv vertices."""

    kind = 1

    def WriteState(self, model, time=-1.0, rec=-1, system=0):
        batch, s = _target(model, system)
        self._check(lib().cgpu_db_avm_write_batch(self.h, batch.h, s, float(time), int(rec)), "WriteState")

    def ReadState(self, model, rec, geometry=True, system=0):
        batch, s = _target(model, system)
        self._check(lib().cgpu_db_avm_read_batch(self.h, batch.h, s, int(rec)), "ReadState")
