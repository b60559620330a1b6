"""SHY output files (the format SHYFEM's post-processing reads) through the
library's native writer (shyhip_shy_*): a header from a basin, then data
records from host arrays or from a model's device state.

    w = ShyWriter("run.hydro.shy", basin, ftype=1, nvar=4, title="lagoon")
    w.hydro(model, dtime)              # znv, zenv, utlnv, vtlnv
    w.close()
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import ShyhipError, load_library


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class ShyWriter:
    """ftype 1: hydro file (npr 3), 2: scalars on nodes (npr 1)."""

    def __init__(self, path, basin, ftype=1, nvar=4, date=0, time=0, title="", femver="shyfemcm-ismar_amd",
                 simpar=(0.01, 0.05, 0.0, 0.0), ext_nodes=None, ext_elems=None):
        self.lib = lib = load_library()
        lib.shyhip_shy_error.restype = C.c_char_p
        b = self.basin = basin
        i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)      # noqa: E731
        f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)    # noqa: E731
        meta = getattr(basin, "meta", {}) or {}
        ipv = i32(ext_nodes if ext_nodes is not None else meta.get("node_ext", np.arange(1, b.nkn + 1)))
        ipev = i32(ext_elems if ext_elems is not None else meta.get("elem_ext", np.arange(1, b.nel + 1)))
        keep = dict(nen=i32(b.nen3v + 1), ipev=ipev, ipv=ipv, x=f32(b.xgv), y=f32(b.ygv), hm3v=f32(b.hm3v),
                    hlv=f32(b.hlv), ilhv=i32(b.ilhv), ilhkv=i32(b.ilhkv), simpar=f32(simpar))
        self._keep = keep
        f = C.c_void_p()
        npr = 3 if ftype == 1 else 1
        rc = lib.shyhip_shy_open(str(path).encode(), ftype, b.nkn, b.nel, npr, b.nlv, nvar, date, time,
                                 title.encode(), femver.encode(), _p(keep["simpar"]), _p(keep["nen"]),
                                 _p(keep["ipev"]), _p(keep["ipv"]), None, None, _p(keep["x"]), _p(keep["y"]),
                                 _p(keep["hm3v"]), _p(keep["hlv"]), _p(keep["ilhv"]), _p(keep["ilhkv"]), C.byref(f))
        if rc:
            raise ShyhipError(f"shyhip_shy_open({path}) failed")
        self.f = f

    def _chk(self, rc, what):
        if rc:
            raise ShyhipError(f"{what}: {self.lib.shyhip_shy_error(self.f).decode()}")

    def record(self, dtime, ivar, values, belem=False, m=1):
        """values: (n,) or (n, m) for 2D, (n, nlv) for 3D, in the basin's
        numbering (the reference's c(l, i) transposed)."""
        a = np.ascontiguousarray(values, dtype=np.float32)
        n = self.basin.nel if belem else self.basin.nkn
        if a.ndim == 2 and m == 1 and a.shape[1] > 1:
            lmax = nlvddi = a.shape[1]
        else:
            lmax = nlvddi = 1
        self._chk(self.lib.shyhip_shy_write_record(self.f, C.c_double(dtime), ivar, 1 if belem else 0, n, m, lmax,
                                                   nlvddi, _p(a)), "shy_write_record")

    def hydro(self, model, dtime):
        self._chk(self.lib.shyhip_shy_write_hydro(model.h, self.f, C.c_double(dtime)), "shy_write_hydro")

    def node3d(self, model, dtime, ivar, field):
        fid = model.field_id(field)
        self._chk(self.lib.shyhip_shy_write_node3d(model.h, self.f, C.c_double(dtime), ivar, fid),
                  "shy_write_node3d")

    def close(self):
        if self.f:
            rc = self.lib.shyhip_shy_close(self.f)      # frees the handle, error text included
            self.f = None
            if rc:
                raise ShyhipError("shy_close: could not close the file")


class DbgWriter:
    """Debug dumps in the reference's format (shympi_debug), the files
    check_shympi_debug lists and compares."""

    TYPES = {np.dtype(np.int32): 1, np.dtype(np.float32): 2, np.dtype(np.float64): 3}

    def __init__(self, path):
        self.lib = lib = load_library()
        lib.shyhip_dbg_error.restype = C.c_char_p
        d = C.c_void_p()
        if lib.shyhip_dbg_open(str(path).encode(), C.byref(d)):
            raise ShyhipError(f"shyhip_dbg_open({path}) failed")
        self.d = d

    def _chk(self, rc, what):
        if rc:
            raise ShyhipError(f"{what}: {self.lib.shyhip_dbg_error(self.d).decode()}")

    def time(self, dtime):
        self._chk(self.lib.shyhip_dbg_time(self.d, C.c_double(dtime)), "dbg_time")

    def record(self, name, values):
        """values: (nn,) or (nn, lmax) of int32, float32 or float64."""
        a = np.ascontiguousarray(values)
        if a.dtype not in self.TYPES:
            a = np.ascontiguousarray(a, dtype=np.float32)
        nn = a.shape[0]
        lmax = 1 if a.ndim == 1 else a.shape[1]
        self._chk(self.lib.shyhip_dbg_record(self.d, name.encode(), nn, lmax, self.TYPES[a.dtype], _p(a)),
                  "dbg_record")

    def end_time(self):
        self._chk(self.lib.shyhip_dbg_end_time(self.d), "dbg_end_time")

    def state(self, model):
        """The model's state records (inside a time block)."""
        self._chk(self.lib.shyhip_dbg_write_state(model.h, self.d), "dbg_write_state")

    def header(self, basin):
        """The basin records the reference's first debug block holds
        (shyfem_subs.f90:1198-1204): external numbers, coordinates, depths."""
        meta = getattr(basin, "meta", {}) or {}
        self.record("ipv", np.asarray(meta.get("node_ext", np.arange(1, basin.nkn + 1)), np.int32))
        self.record("ipev", np.asarray(meta.get("elem_ext", np.arange(1, basin.nel + 1)), np.int32))
        self.record("xgv", np.asarray(basin.xgv, np.float32))
        self.record("ygv", np.asarray(basin.ygv, np.float32))
        self.record("hm3v", np.asarray(basin.hm3v, np.float32))

    def close(self):
        if self.d:
            rc = self.lib.shyhip_dbg_close(self.d)
            self.d = None
            if rc:
                raise ShyhipError("dbg_close: could not close the file")
