"""The oracle's optional hydro forcing terms pinned by closed forms.

hydro_intern (hydrodynamic.f90:1025-1033,1235-1251) adds the atmospheric
pressure gradient (h/rho0) grad(p) and measures the level against the
equilibrium tide (zeta - zeqv); hydro_zeta (:612,679) adds the level nudging
term 4 aj dt andgzv to each vertex's right-hand side; elements with
iuvfix = 1 keep the transports they hold (:1383-1392), and their momentum
leaves the zeta matrix (delta = 0, :627-629) and the final transports
(rfix = 0, :1525).  Checks (FP64 oracle, tight solves):
  * inverse barometer: a closed basin under a uniform pressure gradient G
    settles at the slope d(zeta)/dx = -G / (rho0 g), d(zeta)/dy = 0;
  * equilibrium tide: under zeqv = E x the level settles at slope E;
  * nudging: a uniform andgzv = r raises a basin at rest by r dt per step,
    at every node, with no flow;
  * fixed transports: every element fixed at a uniform transport U keeps it
    bit for bit, and each step moves each node's level by the lumped P1
    continuity dt U . (integral of grad phi_n) / (integral of phi_n), the
    element gradients computed independently (plane through the vertices)
    -- zero inside the basin, non-zero only along the walls.
"""
import dataclasses

import numpy as np

import pyoracle
from shyhip import mesh

RHO0, GRAV = 1025.0, 9.81


def _closed(b):
    return dataclasses.replace(b, ibnd=np.zeros(b.nkn, np.int32), ibtyp=np.array([0], np.int32),
                               bnd_nodes=np.zeros(0, np.int32))


def _slopes(om, b):
    z = om.a["znv"]
    gx = (om.ev[:, 3:6] * z[b.nen3v]).sum(1)
    gy = (om.ev[:, 6:9] * z[b.nen3v]).sum(1)
    return gx, gy


def _steady(forcing, value):
    b = _closed(mesh.tidal_basin(noise=0.0))
    om = pyoracle.OracleModel(b, "f64", params=dict(dt=300.0, ireib=1, itpre=1, rtol=1e-15, atol=1e-15), czv=1e-3)
    om.set_forcing(forcing, value(b.xgv.astype(np.float64), b.ygv.astype(np.float64)))
    v0 = om.volume()
    for _ in range(1000):
        om.step(None)
    assert abs(om.volume() - v0) / v0 < 1e-13
    return b, om


def test_inverse_barometer():
    G = 0.02                                            # Pa/m
    b, om = _steady("ppv", lambda x, y: 101325.0 + G * x)
    gx, gy = _slopes(om, b)
    ratio = -RHO0 * GRAV * gx / G
    print("-rho0 g dz/dx / G: mean", ratio.mean(), "range", ratio.min(), ratio.max())
    assert abs(ratio.mean() - 1.0) < 2e-3
    assert abs(gy.mean()) < 1e-3 * abs(gx.mean())
    # residual seiche << the transport the unbalanced gradient would drive
    # against linear friction r = 1e-3 m/s: h G / (rho0 r)
    assert np.abs(om.a["utlnv"]).max() < 2e-3 * 10.0 * G / (RHO0 * 1e-3)


def test_equilibrium_tide():
    E = 2e-6                                            # m/m
    b, om = _steady("zeqv", lambda x, y: E * x - 0.01)
    gx, gy = _slopes(om, b)
    ratio = gx / E
    print("dz/dx / E: mean", ratio.mean(), "range", ratio.min(), ratio.max())
    assert abs(ratio.mean() - 1.0) < 2e-3
    assert abs(gy.mean()) < 1e-3 * abs(gx.mean())
    assert np.abs(om.a["utlnv"]).max() < 2e-3 * 10.0 * GRAV * E / 1e-3


def test_nudging_uniform_rise():
    b = _closed(mesh.tidal_basin(noise=1.0, nlv=3))
    dt, r = 300.0, 1.5e-5
    om = pyoracle.OracleModel(b, "f64", params=dict(dt=dt, ireib=5, itpre=1, rtol=1e-15, atol=1e-15))
    om.set_forcing("andgzv", np.full(b.nkn, r))
    for n in range(1, 6):
        om.step(None)
        dz = np.abs(om.a["znv"] - n * r * dt)
        print("step", n, "max |zeta - n r dt| =", dz.max())
        assert dz.max() < 1e-12
        assert np.abs(om.a["utlnv"]).max() < 1e-12 and np.abs(om.a["vtlnv"]).max() < 1e-12


def _lumped_continuity(b, ut, vt):
    """dt-free level rate of each node under a uniform transport (ut, vt):
    sum_e A_e (ut dphi/dx + vt dphi/dy) / sum_e A_e / 3."""
    x, y = b.xgv.astype(np.float64), b.ygv.astype(np.float64)
    num = np.zeros(b.nkn)
    mass = np.zeros(b.nkn)
    for ie in range(b.nel):
        kn = b.nen3v[ie]
        mtx = np.stack([np.ones(3), x[kn], y[kn]], axis=1)
        area = 0.5 * abs((x[kn[1]] - x[kn[0]]) * (y[kn[2]] - y[kn[0]]) -
                         (x[kn[2]] - x[kn[0]]) * (y[kn[1]] - y[kn[0]]))
        inv = np.linalg.inv(mtx)                         # column j: coefficients of phi_j
        for j in range(3):
            num[kn[j]] += area * (ut * inv[1, j] + vt * inv[2, j])
            mass[kn[j]] += area / 3.0
    return num / mass


def test_fixed_transports_lumped_continuity():
    b = _closed(mesh.tidal_basin(noise=0.0, nx=21, ny=11, depth=10.0))
    dt, U, V = 300.0, 0.02, -0.005                    # m2/s
    om = pyoracle.OracleModel(b, "f64", params=dict(dt=dt, ireib=5, itpre=1, rtol=1e-15, atol=1e-15))
    ut = np.full((b.nel, b.nlv), U)
    vt = np.full((b.nel, b.nlv), V)
    om.set_uvfix(np.ones(b.nel, np.int32), ut, vt)
    rate = _lumped_continuity(b, U, V)
    interior = np.abs(rate) < 1e-12 * max(abs(U), abs(V))
    assert interior.sum() > 0.5 * b.nkn and (~interior).sum() > 0
    z_prev = None
    for n in range(4):
        om.step(None)
        assert np.all(om.a["utlnv"] == U) and np.all(om.a["vtlnv"] == V)
        z = om.a["znv"].copy()
        if z_prev is not None:                         # both old and new transports fixed
            dz = z - z_prev
            print("step", n + 1, "max |dz - dt rate| =", np.abs(dz - dt * rate).max(), "max |dz|", np.abs(dz).max())
            assert np.abs(dz - dt * rate).max() < 1e-10 * np.abs(dt * rate).max()
            assert np.abs(dz[interior]).max() < 1e-13
        z_prev = z
