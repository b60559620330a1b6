"""SHY output of a model's device state (shyhip_shy_write_hydro,
shyhip_shy_write_node3d): the Venice lagoon stepped on the device, its hydro
time records and a tracer record written and read back equal what
shyhip_pull returns, in the reference's record layout."""
import os

import numpy as np
import pytest

from shyhip import grd, shyout
from test_shy_cpu import DUMP, INLETS, LEVELS, VENICE, _dump, _packed, read_shy

pytestmark = pytest.mark.gpu

PARAMS = dict(dt=60.0, ireib=5, itvd=2, itvdv=1, aapar=0.0, block_rows=256, tri_sweeps=3)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_hydro_output_of_device_state(gpu, prec, tmp_path):
    b = grd.basin_from_grd(grd.read_grd(VENICE), LEVELS, bounds=INLETS, name="venlag62")
    tr = dict(salt=np.full((b.nkn, b.nlv), 30.0, np.float32))
    z0 = float(grd.oscillation(0.0, 0.4, 43200.0, 32400.0))
    m = gpu.Model(b, prec, params=PARAMS, visv=1e-3, difv=1e-4, tracers=tr, zconst=z0)
    hyd = shyout.ShyWriter(str(tmp_path / "v.hydro.shy"), b, ftype=1, nvar=4, title="venice")
    ts = shyout.ShyWriter(str(tmp_path / "v.ts.shy"), b, ftype=2, nvar=1, title="venice")
    want = []
    for k in range(1, 11):
        t = PARAMS["dt"] * k
        lev = float(grd.oscillation(t, 0.4, 43200.0, 32400.0))
        m.set_boundary_levels([lev] * 3)
        m.step(None)
        if k % 5 == 0:
            hyd.hydro(m, t)
            ts.node3d(m, t, 12, "salt")
            want.append((t, m.pull("znv").astype(np.float32), m.pull("zenv").astype(np.float32),
                         m.pull("utlnv").astype(np.float32), m.pull("vtlnv").astype(np.float32),
                         m.pull("salt").astype(np.float32)))
    hyd.close()
    ts.close()
    m.close()
    _, recs = read_shy(str(tmp_path / "v.hydro.shy"))
    _, srec = read_shy(str(tmp_path / "v.ts.shy"))
    assert len(recs) == 8 and len(srec) == 2
    for j, (t, z, ze, u, v, s) in enumerate(want):
        r = recs[4 * j:4 * j + 4]
        assert all(x["dtime"] == t for x in r)
        assert np.array_equal(r[0]["values"], z) and np.array_equal(r[1]["values"], ze.ravel())
        assert np.array_equal(r[2]["values"], _packed(u, b.ilhv))
        assert np.array_equal(r[3]["values"], _packed(v, b.ilhv))
        assert srec[j]["ivar"] == 12 and np.array_equal(srec[j]["values"], _packed(s, b.ilhkv))
    if os.path.exists(DUMP):                     # built from the reference tree where it is present
        _, rr = _dump(str(tmp_path / "v.hydro.shy"))
        assert len(rr) == 8 and np.array_equal(rr[6]["values"], _packed(want[1][3], b.ilhv))
