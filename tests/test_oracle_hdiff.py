"""The oracle's horizontal diffusion of scalars pinned by what the geometry
and the mathematics fix.

diffweight (diffusion_admin.f90:182-340) gives each element a 3x3 weight
matrix: idtype 0 the P1 stiffness grad(phi_a) . grad(phi_b); idtype 1 the
same with each positive off-diagonal weight moved onto the third vertex of
its column; idtype 2 (the reference's default) -(2/3)/s^2 between the ends
of each edge of length s, with zero row sums.  conz3d_element
(concentration_omp.f90:560-570,727-735) subtracts dt 3 rkpar difhv h
wdiff from each vertex's explicit term.  Checks:
  * idtype 0 against the gradients of the plane through the vertices
    (numpy inverse of [1 x y]), idtype 2 against the edge lengths from the
    coordinates, idtype 1 against the move rule applied to that
    independent idtype 0; on right triangles idtype 1 equals idtype 0;
  * a closed flat basin at rest with c = c0 + cos(k x): one substep scales
    the cosine at every node inside the walls by
    1 - dt rkpar difhv (4/dx^2) sin^2(k dx/2), for idtype 0 and 2 alike on
    this mesh of right triangles (idtype 2 with the reference's 2/3 rounded
    to single precision).
"""
import ctypes as C
import dataclasses

import numpy as np
import pytest

import pyoracle
from shyhip import mesh


def _closed(b):
    return dataclasses.replace(b, ibnd=np.zeros(b.nkn, np.int32), ibtyp=np.array([0], np.int32),
                               bnd_nodes=np.zeros(0, np.int32))


def _weights(lib, ev, idtype):
    nel = ev.shape[0]
    w = np.zeros((nel, 3, 3))                       # [ie][b][a] = wdifhv(a, b, ie)
    rc = lib.oc_diffweight(nel, ev.ctypes.data_as(C.c_void_p), idtype, w.ctypes.data_as(C.c_void_p))
    return rc, w


def _grad_products(b):
    x, y = b.xgv.astype(np.float64), b.ygv.astype(np.float64)
    out = np.zeros((b.nel, 3, 3))
    for ie in range(b.nel):
        kn = b.nen3v[ie]
        inv = np.linalg.inv(np.stack([np.ones(3), x[kn], y[kn]], axis=1))   # column j: phi_j
        g = inv[1:, :]                                                       # [xy][j]
        out[ie] = g.T @ g
    return out


@pytest.mark.parametrize("basin", ["lagoon", "right"])
def test_diffweight_types(basin):
    b = (mesh.lagoon(700, seed=21, lx=3000.0, ly=1500.0, nlv=2, tile=8) if basin == "lagoon"
         else mesh.tidal_basin(noise=0.0, nx=9, ny=5))
    om = pyoracle.OracleModel(b, "f64")
    lib = om.lib
    w0_ref = _grad_products(b)
    rc0, w0 = _weights(lib, om.ev, 0)
    assert rc0 == 0
    assert np.allclose(w0, w0_ref, rtol=1e-9, atol=1e-12 * np.abs(w0_ref).max())
    # idtype 2 from the coordinates
    x, y = b.xgv.astype(np.float64), b.ygv.astype(np.float64)
    w2_ref = np.zeros((b.nel, 3, 3))
    for ie in range(b.nel):
        kn = b.nen3v[ie]
        for a in range(3):
            for c in range(3):
                if a != c:
                    s2 = (x[kn[a]] - x[kn[c]]) ** 2 + (y[kn[a]] - y[kn[c]]) ** 2
                    w2_ref[ie, a, c] = -float(np.float32(2.0) / np.float32(3.0)) / s2
        for a in range(3):
            w2_ref[ie, a, a] = -w2_ref[ie, a].sum()
    rc2, w2 = _weights(lib, om.ev, 2)
    assert rc2 == 0
    assert np.allclose(w2, w2_ref, rtol=1e-9, atol=0)
    assert np.allclose(w2.sum(axis=2), 0.0, atol=1e-12 * np.abs(w2).max())
    # idtype 1: each positive off-diagonal weight of a column moved onto the
    # column's third vertex
    w1_ref = w0_ref.copy()
    moved = 0
    for ie in range(b.nel):
        for col in range(3):
            for row in range(3):
                v = w0_ref[ie, col, row]
                if row != col and v > 0:
                    third = 3 - row - col
                    w1_ref[ie, col, row] = 0.0
                    w1_ref[ie, col, third] -= v
                    moved += 1
    rc1, w1 = _weights(lib, om.ev, 1)
    assert rc1 == 0
    assert np.allclose(w1, w1_ref, rtol=1e-9, atol=1e-12 * np.abs(w0_ref).max())
    if basin == "right":
        assert moved == 0 and np.array_equal(w1, w0)
    else:
        assert moved > 0                            # the lagoon has obtuse triangles


@pytest.mark.parametrize("idtype", [0, 2])
def test_cosine_mode_decay(idtype):
    nx, ny, dx = 33, 9, 100.0
    b = _closed(mesh.tidal_basin(noise=0.0, nx=nx, ny=ny, dx=dx, depth=5.0))
    dt, rkpar, D = 60.0, 0.5, 20.0
    x = b.xgv.astype(np.float64)
    L = (nx - 1) * dx
    kw = 8 * np.pi / L
    c0 = 20.0 + np.cos(kw * x)
    om = pyoracle.OracleModel(b, "f64", params=dict(dt=dt, ireib=5, itpre=1, rkpar=rkpar, rtol=1e-15, atol=1e-15),
                              tracers=dict(salt=c0[:, None].copy()), idtype=idtype)
    om.a["difhv"][:] = D
    om.step(None, istot=1)
    assert np.all(om.a["utlnv"] == 0.0) and np.all(om.a["znv"] == 0.0)     # lake at rest
    c1 = om.tracers["salt"][:, 0]
    # idtype 2 weighs with the reference's single-precision 2/3
    scale = 1.5 * float(np.float32(2.0) / np.float32(3.0)) if idtype == 2 else 1.0
    factor = 1.0 - dt * rkpar * D * (4.0 / dx ** 2) * np.sin(kw * dx / 2) ** 2 * scale
    ix = np.rint(x / dx).astype(int)
    iy = np.rint(b.ygv.astype(np.float64) / dx).astype(int)
    inner = (ix > 0) & (ix < nx - 1) & (iy > 0) & (iy < ny - 1)
    got = (c1 - 20.0)[inner]
    exp = factor * (c0 - 20.0)[inner]
    print("idtype", idtype, "decay factor", factor, "max error", np.abs(got - exp).max())
    assert np.abs(got - exp).max() < 1e-12
    assert factor < 0.99                              # the diffusion is visible
