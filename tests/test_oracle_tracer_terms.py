"""The oracle's own terms of a scalar (conz3d_omp's robs/cobs/rtauv, wsink/
wsinkv, rload/load) pinned by closed forms in a closed flat basin at rest.

  * nudging (concentration_omp.f90:697-701,727-735, an = 0): a uniform
    field relaxes explicitly, c(n+1) = c(n) + dt robs rtau (cobs - c(n)),
    in layers of any thickness; scalar_stability's index is robs rtau;
  * sinking (:479, the vertical fluxes of vertical_flux_ie): without
    vertical diffusion each substep moves the fraction e = dt ws / h of a
    layer's content into the layer below (upwind), nothing through the
    surface or the bottom;
  * loads (conz3d_nodes, :816-834): erosion adds dt L / V to the node's
    concentration, V = h * (sum of the incident elements' areas) / 3;
    deposition takes c (1 - exp(-dt |L| / (V c))) and records the load it
    took.
"""
import dataclasses

import numpy as np

import pyoracle
from shyhip import mesh


def _closed_flat(nlv):
    b = mesh.tidal_basin(noise=0.0, nx=9, ny=5, depth=5.0 if nlv == 1 else float(nlv), nlv=nlv)
    return dataclasses.replace(b, ibnd=np.zeros(b.nkn, np.int32), ibtyp=np.array([0], np.int32),
                               bnd_nodes=np.zeros(0, np.int32))


def _model(b, c0, **params):
    p = dict(dt=300.0, ireib=5, itpre=1, rtol=1e-15, atol=1e-15, difmol=0.0, itvdv=0, aapar=0.0)
    p.update(params)
    return pyoracle.OracleModel(b, "f64", params=p, tracers=dict(salt=c0))


def test_nudging_relaxation():
    for nlv in (1, 3):
        b = _closed_flat(nlv)
        c0 = np.full((b.nkn, b.nlv), 10.0)
        om = _model(b, c0)
        dt, tau, cob, robs = 300.0, 3600.0, 20.0, 0.8
        om.set_tracer_terms("salt", robs=robs, cobs=np.full((b.nkn, b.nlv), cob),
                            rtauv=np.full((b.nkn, b.nlv), 1.0 / tau))
        istot, rindex = om.scalar_stability(tracer="salt")
        # the index passes through a default real (the reference's shympi_max)
        assert abs(rindex - dt * robs / tau) < 1e-7 * dt * robs / tau and istot == 1 + int(dt * robs / tau)
        for n in range(1, 6):
            om.step(None, istot=1)
            exp = cob + (10.0 - cob) * (1.0 - dt * robs / tau) ** n
            assert np.abs(om.tracers["salt"] - exp).max() < 1e-11, (nlv, n)


def test_sinking_upwind_layers():
    b = _closed_flat(4)
    L = b.nlv
    prof = np.array([8.0, 4.0, 2.0, 1.0])
    c0 = np.tile(prof, (b.nkn, 1))
    om = _model(b, c0)
    dt, ws, h = 300.0, 5e-4, 1.0
    wsv = np.ones((b.nkn, L + 1))
    om.set_tracer_terms("salt", wsink=ws, wsinkv=wsv)
    e = dt * ws / h
    c = prof.copy()
    for n in range(4):
        om.step(None, istot=1)
        new = c.copy()
        new[0] = c[0] * (1 - e)
        for l in range(1, L - 1):
            new[l] = c[l] + e * (c[l - 1] - c[l])
        new[L - 1] = c[L - 1] + e * c[L - 2]
        c = new
        got = om.tracers["salt"]
        assert np.abs(got - c[None, :]).max() < 1e-12, (n, got[0], c)
    assert abs(c.sum() - prof.sum()) < 1e-12                 # the column keeps its content


def _node_volume(b, h):
    x, y = b.xgv.astype(np.float64), b.ygv.astype(np.float64)
    vol = np.zeros(b.nkn)
    for kn in b.nen3v:
        area = 0.5 * abs((x[kn[1]] - x[kn[0]]) * (y[kn[2]] - y[kn[0]]) - (x[kn[2]] - x[kn[0]]) * (y[kn[1]] - y[kn[0]]))
        vol[kn] += area / 3.0 * h
    return vol


def test_loads_erosion_and_deposition():
    b = _closed_flat(1)
    h, dt = 5.0, 300.0
    vol = _node_volume(b, h)
    c0 = np.full((b.nkn, 1), 10.0)
    # erosion
    om = _model(b, c0)
    load = np.full((b.nkn, 1), 2.0e3)
    om.set_tracer_terms("salt", rload=0.5, load=load)
    om.step(None, istot=1)
    exp = 10.0 + dt * 0.5 * 2.0e3 / vol
    assert np.abs(om.tracers["salt"][:, 0] - exp).max() < 1e-11
    # deposition: the exponential bound, and the load actually taken
    om = _model(b, c0)
    om.set_tracer_terms("salt", rload=0.5, load=np.full((b.nkn, 1), -4.0e3))
    om.step(None, istot=1)
    cn = 10.0 * vol
    cload = cn * (1.0 - np.exp(-dt * 0.5 * 4.0e3 / cn))
    assert np.abs(om.tracers["salt"][:, 0] - (cn - cload) / vol).max() < 1e-11
    taken = om.tracer_terms["salt"]["load"][:, 0]
    assert np.abs(taken - (-cload / dt) / 0.5).max() < 1e-9 * np.abs(taken).max()
