"""Forcing series on the device (shyhip_set_boundary_series,
shyhip_set_wind_series, shyhip_forcing and the forcing inside shyhip_step)
against the oracle's restatement of the reference's record handling
(iff_populate_records / iff_read_and_interpolate / rd_intp_neville in
ring-slot order) and of meteo_convert_wind_data.  The bar is bit-identical
f32 values: boundary levels in rzv, stresses in tauxnv/tauynv, and whole
steps driven by device series equal to steps driven by levels and stresses
the oracle computed on the host."""
import dataclasses
import math
import os
import sys

import numpy as np
import pytest

from shyhip import mesh

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import pyoracle as po  # noqa: E402

pytestmark = pytest.mark.gpu

PARAMS = dict(dt=60.0, ireib=5, itvd=2, itvdv=1, aapar=0.0, block_rows=256, tri_sweeps=3)


def _basin_two_boundaries():
    """The lagoon's inlet split into two level boundaries."""
    b = mesh.lagoon(2500, seed=61, lx=7000.0, ly=3500.0, nlv=5, tile=16)
    nodes = np.sort(b.bnd_nodes)
    ibnd = b.ibnd.copy()
    half = len(nodes) // 2
    ibnd[nodes[half:]] = 2
    b2 = dataclasses.replace(b, ibnd=ibnd, ibtyp=np.array([0, 1, 1], np.int32))
    return b2, nodes[:half], nodes[half:]


def _f32(a):
    return np.asarray(a).astype(np.float32)


def _level_records():
    rng = np.random.default_rng(20261202)
    t1 = np.cumsum(np.r_[0.0, rng.uniform(500, 1300, 40)]) - 1200.0      # irregular, starts before 0
    v1 = (0.4 * np.sin(2 * math.pi * t1 / mesh.M2_PERIOD)).astype(np.float32)
    return t1, v1


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_boundary_series_levels(gpu, prec):
    b, n1, n2 = _basin_two_boundaries()
    m = gpu.Model(b, prec, params=PARAMS, visv=1e-3)
    t1, v1 = _level_records()
    # boundary 2: node values, linear, zfact, one record without data at one node
    t2 = np.arange(0.0, 40000.0, 1800.0)
    x = np.linspace(0, 1, len(n2))
    v2 = (0.3 * np.cos(2 * math.pi * t2[:, None] / mesh.K1_PERIOD + x[None, :])).astype(np.float32)
    v2[5, 1] = -999.0
    m.set_boundary_series(1, t1, v1, intpol=4, dtime0=0.0)
    m.set_boundary_series(2, t2, v2, intpol=2, zfact=0.9, dtime0=0.0)
    ts = np.arange(60.0, 30000.0, 60.0 * 37)
    o1, _ = po.iff_series(t1, v1[:, None], 4, 0.0, ts)
    o2, _ = po.iff_series(t2, v2, 2, 0.0, ts)
    for k, t in enumerate(ts):
        m.forcing(t)
        rzv = _f32(m.pull("rzv"))
        assert np.array_equal(rzv[n1], np.full(len(n1), o1[k, 0], np.float32)), (t, "boundary 1")
        assert np.array_equal(rzv[n2], o2[k] * np.float32(0.9)), (t, "boundary 2")
    with pytest.raises(gpu.ShyhipError, match="outside the records"):
        m.forcing(t2[-1] + 60.0)
    m.close()


@pytest.mark.parametrize("prec,kw,npt", [("fp64", dict(itdrag=1, wslim=25.0), "nodes"),
                                         ("fp32", dict(itdrag=2), "nodes"),
                                         ("fp32", dict(iwtype=2), "one")])
def test_wind_series_stress(gpu, prec, kw, npt):
    b = mesh.lagoon(2000, seed=62, lx=6000.0, ly=3000.0, nlv=4, tile=16)
    m = gpu.Model(b, prec, params=PARAMS, visv=1e-3)
    rng = np.random.default_rng(5)
    tw = np.arange(-3600.0, 90000.0, 3 * 3600.0)
    nv = b.nkn if npt == "nodes" else 1
    sp = rng.uniform(0.0, 30.0, (len(tw), nv))
    d = rng.uniform(0, 2 * math.pi, (len(tw), nv))
    scale = 1.0 if kw.get("iwtype", 1) == 1 else 0.02          # stresses in N/m2 for iwtype 2
    wx = (scale * sp * np.cos(d)).astype(np.float32)
    wy = (scale * sp * np.sin(d)).astype(np.float32)
    w = dict(iwtype=1, itdrag=0, dragco=2.5e-3, roluft=1.225, wslim=-1.0, wsmax=50.0)
    w.update(kw)
    m.set_wind_series(tw, wx if nv > 1 else wx[:, 0], wy if nv > 1 else wy[:, 0], intpol=2, dtime0=0.0, **kw)
    vals = np.stack([wx, wy], axis=1)                          # (nrec, 2, nval)
    ts = np.arange(60.0, 80000.0, 60.0 * 211)
    ox, _ = po.iff_series(tw, vals, 2, 0.0, ts, comp=0)
    oy, _ = po.iff_series(tw, vals, 2, 0.0, ts, comp=1)
    for k, t in enumerate(ts):
        m.forcing(t)
        tx, ty, _, _, _, rc = po.convert_wind(ox[k], oy[k], rowass=1025.0, **w)
        assert rc == 0
        gx, gy = _f32(m.pull("tauxnv")), _f32(m.pull("tauynv"))
        if nv == 1:
            tx, ty = np.full(b.nkn, tx[0], np.float32), np.full(b.nkn, ty[0], np.float32)
        assert np.array_equal(gx, tx) and np.array_equal(gy, ty), t
    m.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_steps_with_device_series_equal_host_forcing(gpu, prec):
    """A run whose levels and wind come from series on the device equals, bit
    for bit, a run fed with the oracle's interpolated levels and stresses."""
    b = mesh.lagoon(2500, seed=41, lx=7000.0, ly=3500.0, nlv=5, tile=16)
    tr = dict(salt=mesh.smooth_tracer(b, 42, 25.0, 38.0))
    t1, v1 = _level_records()
    tw = np.arange(0.0, 20000.0, 1800.0)
    rng = np.random.default_rng(9)
    wx = rng.uniform(-12, 12, len(tw)).astype(np.float32)
    wy = rng.uniform(-12, 12, len(tw)).astype(np.float32)
    nsteps = 40
    ts = PARAMS["dt"] * np.arange(1, nsteps + 1)
    lev, _ = po.iff_series(t1, v1[:, None], 4, 0.0, ts)
    ox, _ = po.iff_series(tw, wx[:, None], 2, 0.0, ts)
    oy, _ = po.iff_series(tw, wy[:, None], 2, 0.0, ts)
    out = []
    for device in (True, False):
        m = gpu.Model(b, prec, params=PARAMS, visv=1e-3, difv=1e-4, tracers=tr)
        if device:
            m.set_boundary_series(1, t1, v1, intpol=4, dtime0=0.0)
            m.set_wind_series(tw, wx, wy, intpol=2, dtime0=0.0)
            m.set_deferred_checks(True)
        for k in range(nsteps):
            if device:
                m.step(None)
            else:
                tx, ty, _, _, _, _ = po.convert_wind(ox[k], oy[k])
                m.set_wind(tx[0], ty[0])
                m.step(float(lev[k, 0]))
        m.sync()
        out.append({f: m.pull(f) for f in ("znv", "utlnv", "rzv", "tauxnv", "salt")})
        m.close()
    for f in out[0]:
        assert np.array_equal(out[0][f], out[1][f]), f


def test_wind_above_wsmax_is_an_error(gpu):
    b = mesh.lagoon(1500, seed=63, lx=5000.0, ly=2500.0, nlv=3, tile=16)
    m = gpu.Model(b, "fp64", params=PARAMS, visv=1e-3)
    tw = np.array([0.0, 120.0])
    m.set_wind_series(tw, np.array([10.0, 70.0], np.float32), np.array([0.0, 0.0], np.float32), wsmax=50.0)
    m.step(None)                                        # 60 s: 40 m/s
    with pytest.raises(gpu.ShyhipError, match="wsmax"):
        m.step(None)                                    # 120 s: 70 m/s
    m.close()


def test_ensemble_level_series_equal_host_levels(gpu):
    """Ensemble members' boundary levels as device series sampled at every
    step: interpolation at a record's time returns the record, so the run
    equals the one that uploads the same levels each step."""
    from shyhip import ensemble
    mems = []
    for n, s, phase in ((1200, 71, 0.0), (1800, 72, 2.0)):
        b = mesh.lagoon(n, seed=s, lx=6000.0, ly=3000.0, nlv=4, tile=16)
        mems.append(dict(basin=b, wind=mesh.wind_stress(6.0, phase),
                         level=(lambda t, ph=phase: mesh.tide_level(t, amps=(0.4, 0.1),
                                                                   periods=(mesh.M2_PERIOD, mesh.K1_PERIOD),
                                                                   phases=(ph, 0.7)) - 0.2),
                         tracers=dict(salt=mesh.smooth_tracer(b, s + 1, 25.0, 38.0).astype(np.float32))))
    res = []
    for series in (False, True):
        ens = ensemble.Ensemble(mems, "fp32", params=PARAMS, visv=1e-3, difv=1e-4, deferred=True)
        if series:
            ens.set_level_series(PARAMS["dt"] * 20)
        for _ in range(15):
            ens.step()
        ens.sync()
        res.append([(ens.pull("znv", i), ens.pull("salt", i)) for i in range(len(mems))])
        ens.close()
    for (za, sa), (zb, sb) in zip(*res):
        assert np.array_equal(za, zb) and np.array_equal(sa, sb)


def test_step_outside_records_leaves_state(gpu):
    """A step whose time the series do not cover fails before it starts
    (status 2, as the reference stops in iff_time_interpolate): the state
    is the one of the last good step."""
    b = mesh.lagoon(1500, seed=64, lx=5000.0, ly=2500.0, nlv=3, tile=16)
    m = gpu.Model(b, "fp64", params=PARAMS, visv=1e-3)
    t1, v1 = _level_records()
    m.set_boundary_series(1, t1, v1, intpol=4, dtime0=0.0)
    for _ in range(3):
        m.step(None)
    before = {f: m.pull(f) for f in ("znv", "utlnv", "zenv")}
    m.t = float(t1[-1])                  # the next step's time lies past the last record
    with pytest.raises(gpu.ShyhipError, match="outside the records"):
        m.step(None)
    for f, v in before.items():
        assert np.array_equal(m.pull(f), v), f
    m.close()
