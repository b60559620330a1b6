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
