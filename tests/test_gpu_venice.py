"""A real grid through the device path: the reference's Venice lagoon test
case (tests/golden/venlag62.grd: 4359 nodes, 7842 elements, element depths
-0.4..38.8 m, 17 z-layers as its STR's $levels), read by shyhip.grd, with
the STR's tide at its three inlets (ampli 0.4 m, period 43200 s, phase
32400 s: get_oscil) and the start level set from the boundary as sp111(1)
sets zconst.  Wetting and drying is active on the salt marshes.  The device
run is compared with the oracle run on the same basin, to the bars of the
other parity tests."""
import os
import sys

import numpy as np
import pytest

from shyhip import grd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import pyoracle as po  # noqa: E402

pytestmark = pytest.mark.gpu

VENICE = os.path.join(ROOT, "tests", "golden", "venlag62.grd")
LEVELS = [1, 2, 3, 4, 5, 6, 8, 10, 12, 14, 16, 20, 24, 28, 32, 36, 40]
INLETS = [[150, 173, 172, 151], [4311, 4315, 4314, 4309], [2750, 2757, 2758, 2751]]
PARAMS = dict(dt=60.0, ireib=5, itvd=2, itvdv=1, aapar=0.0, block_rows=256, tri_sweeps=3, rtol=1e-12, atol=0.0)


def _tide(t):
    return float(grd.oscillation(t, 0.4, 43200.0, 32400.0, 0.0))


@pytest.mark.parametrize("prec,oprec,ztol", [("fp64", "f64", 1e-6), ("fp32", "f32", 2e-5)])
def test_venice_tide_matches_oracle(gpu, prec, oprec, ztol):
    b = grd.basin_from_grd(grd.read_grd(VENICE), LEVELS, bounds=INLETS, name="venlag62")
    tr = dict(salt=np.full((b.nkn, b.nlv), 30.0, np.float32))
    z0 = _tide(0.0)
    m = gpu.Model(b, prec, params=PARAMS, visv=1e-3, difv=1e-4, tracers=tr, zconst=z0)
    om = po.OracleModel(b, oprec, params=PARAMS, visv=1e-3, difv=1e-4, tracers=tr, zconst=z0)
    nsteps = 30
    for k in range(1, nsteps + 1):
        lev = _tide(PARAMS["dt"] * k)
        m.set_boundary_levels([lev] * 3)
        m.step(None)
        om.step(lev)
    z, zo = m.pull("znv"), om.a["znv"].astype(np.float64)
    w, wo = m.pull("iwegv"), om.a["iwegv"]
    print(prec, "dry elements", int((wo != 0).sum()), "max |dz|", float(np.abs(z - zo).max()))
    assert int((wo != 0).sum()) > 100                   # the marshes are dry
    assert np.array_equal(w, wo)
    assert np.abs(z - zo).max() < ztol
    valid_e = np.arange(b.nlv)[None, :] < b.ilhv[:, None]
    u, uo = m.pull("utlnv"), om.a["utlnv"].astype(np.float64)
    assert np.abs(u - uo)[valid_e].max() < 1e-4 * np.abs(uo[valid_e]).max()
    valid_n = np.arange(b.nlv)[None, :] < b.ilhkv[:, None]
    s = m.pull("salt")[valid_n]
    assert np.abs(s - 30.0).max() < (1e-9 if prec == "fp64" else 2e-4 * 30.0)   # a constant stays constant
    m.close()
