"""The reference's debug-file comparison (check_shympi_debug, built into
oracle/_ref/check_dbg) applied to the device path: the Venice lagoon after
30 tidal steps on the device and in the oracle, each dumped as SHYFEM's debug
output does (one time block: ipv, ipev, then the state this path updates),
differ by no more than 1e-4 in any value of any record (levels, transports,
vertical velocities, depths, the wet/dry mask)."""
import os
import subprocess
import sys

import numpy as np
import pytest

from shyhip import grd, shyout
from test_shy_cpu import INLETS, LEVELS, VENICE

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import pyoracle as po  # noqa: E402

CHECK = os.path.join(ROOT, "oracle", "_ref", "check_dbg")
pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not os.path.exists(CHECK), reason="reference debug checker not built")]
PARAMS = dict(dt=60.0, ireib=5, itvd=2, itvdv=1, aapar=0.0, block_rows=256, tri_sweeps=3, rtol=1e-12, atol=0.0)


def test_device_and_oracle_dumps_agree(gpu, tmp_path):
    b = grd.basin_from_grd(grd.read_grd(VENICE), LEVELS, bounds=INLETS, name="venlag62")
    tr = dict(salt=np.full((b.nkn, b.nlv), 30.0, np.float32))
    z0 = float(grd.oscillation(0.0, 0.4, 43200.0, 32400.0))
    m = gpu.Model(b, "fp64", params=PARAMS, visv=1e-3, difv=1e-4, tracers=tr, zconst=z0)
    om = po.OracleModel(b, "f64", params=PARAMS, visv=1e-3, difv=1e-4, tracers=tr, zconst=z0)
    t = 0.0
    for k in range(1, 31):
        t = PARAMS["dt"] * k
        lev = float(grd.oscillation(t, 0.4, 43200.0, 32400.0))
        m.set_boundary_levels([lev] * 3)
        m.step(None)
        om.step(lev)
    fd, fo = str(tmp_path / "device.dbg"), str(tmp_path / "oracle.dbg")
    w = shyout.DbgWriter(fd)
    w.time(t)
    w.header(b)
    w.state(m)
    w.end_time()
    w.close()
    m.close()
    a = om.a
    w = shyout.DbgWriter(fo)
    w.time(t)
    w.header(b)
    for name in ("znv", "zenv", "utlnv", "vtlnv", "wlnv", "hdknv", "hdenv"):
        w.record(name, np.asarray(a[name], np.float32))
    w.record("iwegv", np.asarray(a["iwegv"], np.int32))
    w.end_time()
    w.close()
    # the tool keeps file names in 60 characters: run it beside the files
    p = subprocess.run([CHECK, "-quiet", "-maxdiff", "1e-4", "device.dbg", "oracle.dbg"], capture_output=True,
                       text=True, cwd=str(tmp_path))
    print(p.stdout[-2000:])
    assert "differences found:  0" in p.stdout
