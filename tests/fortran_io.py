"""Case and result files of the Fortran host program
shyfemcm-ismar_amd/fortran/shyhip_run, for tests/test_gpu_fortran_host.py
(stream records in native byte order, layouts as in the reference: Fortran
column-major, 1-based connectivity)."""
from __future__ import annotations

import os

import numpy as np

FORTRAN_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "shyfemcm-ismar_amd", "fortran")
RUN_EXE = os.path.join(FORTRAN_DIR, "shyhip_run")
DROPIN_EXE = os.path.join(FORTRAN_DIR, "shyhip_dropin_check")


def write_case(path, basin, precision: str, dt: float, levels, ireib=5, itvd=0, itvdv=0, aapar=0.0,
               czv=2.5e-3, visv=0.0, difv=0.0, wind=(0.0, 0.0), tracers=None):
    b = basin
    tr = list(tracers.values()) if tracers else []
    levels = np.asarray(levels, dtype=np.float32)
    nbnd = len(b.ibtyp) - 1
    hdr = np.array([b.nkn, b.nel, b.nlv, nbnd, len(levels), 0 if precision == "fp64" else 1, len(tr),
                    ireib, itvd, itvdv], dtype=np.int32)
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32).reshape(-1)  # noqa: E731
    i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32).reshape(-1)  # noqa: E731
    with open(path, "wb") as f:
        hdr.tofile(f)
        np.array([dt], dtype=np.float64).tofile(f)
        np.array([aapar, czv, visv, difv, wind[0], wind[1]], dtype=np.float32).tofile(f)
        i32(b.nen3v + 1).tofile(f)                      # nen3v(3,nel): vertex index fastest
        for a in (b.xgv, b.ygv, b.hm3v, b.hev, b.hlv, b.hldv):
            f32(a).tofile(f)
        for a in (b.ilhv, b.ilhkv, b.ibnd, b.ibtyp, b.iextb):
            i32(a).tofile(f)
        levels.tofile(f)
        for c in tr:
            f32(c).tofile(f)                            # conc(nlv,nkn): layer fastest


def read_result(path, basin, ntracer):
    b = basin
    with open(path, "rb") as f:
        znv = np.fromfile(f, dtype=np.float32, count=b.nkn)
        utlnv = np.fromfile(f, dtype=np.float32, count=b.nel * b.nlv).reshape(b.nel, b.nlv)
        conc = [np.fromfile(f, dtype=np.float32, count=b.nkn * b.nlv).reshape(b.nkn, b.nlv) for _ in range(ntracer)]
        iwegv = np.fromfile(f, dtype=np.int32, count=b.nel)
        vol = float(np.fromfile(f, dtype=np.float64, count=1)[0])
    return dict(znv=znv, utlnv=utlnv, conc=conc, iwegv=iwegv, volume=vol)
