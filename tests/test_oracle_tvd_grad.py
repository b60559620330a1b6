"""The oracle's gradient TVD (itvd = 1) pinned by what the mathematics fixes.

tvd_grad_3d (tvd_admin.f90:478-543) averages the P1 element gradients of a
scalar over each node's incident elements that have the layer, weighted by
the element area.  tvd_fluxes (tvd_admin.f90:673-823) then takes, for a
pair (ic -> id) of an element, grad = 2 (g(kc) . (x_d - x_c)) with g(kc) the
gradient at the upstream node.  Checks:
  * a field linear in x and y is differentiated exactly at every node-layer;
  * a general field: the per-element gradient from the plane through the
    three vertices (numpy solve, not the element's shape coefficients) and
    the shoelace area, averaged over the deep-enough incident elements;
  * a field linear on the elements around one node and random elsewhere:
    each TVD pair leaving that node sees the ratio 1 of a linear field, so
    the flux is the Lax-Wendroff face value c_c + (c_d - c_c)(1 - alfa)/2
    with alfa = dt |u . (x_d - x_c)| / |x_d - x_c|^2 (taking the gradient at
    the downstream node, a sign error in the edge vector, a dropped factor
    of 2 or transposed gradient components all fail it).
"""
import numpy as np
import pytest

import pyoracle
from shyhip import mesh


def _basin():
    return mesh.lagoon(900, seed=11, lx=3000.0, ly=1500.0, nlv=4, tile=8)


def _layer_valid(b):
    return np.arange(b.nlv)[None, :] < b.ilhkv[:, None]


@pytest.mark.parametrize("prec,tol", [("f64", 1e-11), ("f32", 2e-4)])
def test_grad_linear_field_exact(prec, tol):
    b = _basin()
    om = pyoracle.OracleModel(b, prec, params=dict(itvd=1))
    x, y = b.xgv.astype(np.float64), b.ygv.astype(np.float64)
    ax = np.array([2e-3, -1e-3, 5e-4, 3e-3])
    ay = np.array([-7e-4, 1.5e-3, 2e-3, -2.5e-3])
    c = 10.0 + x[:, None] * ax[None, :] + y[:, None] * ay[None, :]
    gx, gy = om.tvd_grad(c)
    valid = _layer_valid(b)
    exp_x = np.broadcast_to(ax, c.shape)
    exp_y = np.broadcast_to(ay, c.shape)
    assert np.max(np.abs(gx - exp_x)[valid]) <= tol * np.max(np.abs(ax))
    assert np.max(np.abs(gy - exp_y)[valid]) <= tol * np.max(np.abs(ay))
    assert np.all(gx[~valid] == 0) and np.all(gy[~valid] == 0)


def test_grad_general_field_area_weighted():
    b = _basin()
    om = pyoracle.OracleModel(b, "f64", params=dict(itvd=1))
    rng = np.random.default_rng(5)
    c = rng.uniform(5.0, 30.0, size=(b.nkn, b.nlv))
    gx, gy = om.tvd_grad(c)
    x, y = b.xgv.astype(np.float64), b.ygv.astype(np.float64)
    sx = np.zeros((b.nkn, b.nlv))
    sy = np.zeros((b.nkn, b.nlv))
    sa = np.zeros((b.nkn, b.nlv))
    for ie in range(b.nel):
        kn = b.nen3v[ie]
        mtx = np.stack([np.ones(3), x[kn], y[kn]], axis=1)
        area = 0.5 * abs((x[kn[1]] - x[kn[0]]) * (y[kn[2]] - y[kn[0]]) -
                         (x[kn[2]] - x[kn[0]]) * (y[kn[1]] - y[kn[0]]))
        for l in range(b.ilhv[ie]):
            coef = np.linalg.solve(mtx, c[kn, l])
            sx[kn, l] += area * coef[1]
            sy[kn, l] += area * coef[2]
            sa[kn, l] += area
    valid = _layer_valid(b)
    assert np.all(sa[valid] > 0)
    ex, ey = sx[valid] / sa[valid], sy[valid] / sa[valid]
    scale = max(np.max(np.abs(ex)), np.max(np.abs(ey)))
    assert np.max(np.abs(gx[valid] - ex)) < 1e-9 * scale
    assert np.max(np.abs(gy[valid] - ey)) < 1e-9 * scale
    # the layer mask matters on this basin: some node-layers have incident
    # elements that stop above them
    mixed = 0
    for ie in range(b.nel):
        mixed += int(np.any(b.ilhkv[b.nen3v[ie]] > b.ilhv[ie]))
    assert mixed > 0


def _stars(b):
    star = [[] for _ in range(b.nkn)]
    for ie in range(b.nel):
        for k in b.nen3v[ie]:
            star[k].append(ie)
    return star


def test_flux_lax_wendroff_on_linear_patch():
    b = _basin()
    dt = 30.0
    u, v, ut, vt = 0.25, 0.1, 1.2, 0.5
    models = {}
    for itvd in (0, 1):
        om = pyoracle.OracleModel(b, "f64", params=dict(itvd=itvd, azpar=0.5))
        for n, val in (("utlnv", ut), ("utlov", ut), ("vtlnv", vt), ("vtlov", vt), ("ulnv", u), ("vlnv", v)):
            om.a[n][:] = val
        models[itvd] = om
    star = _stars(b)
    x, y = b.xgv.astype(np.float64), b.ygv.astype(np.float64)
    rng = np.random.default_rng(9)
    ext = b.iextb.astype(bool)
    checked = 0
    for kc in rng.permutation(b.nkn)[:400]:
        elems = star[kc]
        nodes = np.unique(b.nen3v[elems].ravel())
        if ext[nodes].any() or b.ilhkv[kc] < 1:
            continue
        # linear on the elements around kc, random elsewhere
        g1 = np.array([3e-3, -2e-3])
        c = rng.uniform(5.0, 30.0, size=(b.nkn, b.nlv))
        c[nodes, :] = 12.0 + g1[0] * x[nodes, None] + g1[1] * y[nodes, None]
        for ie in elems:
            kn = list(b.nen3v[ie])
            ic = kn.index(kc)
            for l in range(1, b.ilhv[ie] + 1):
                # upwind with one upstream vertex: fl[ii] = f[ii] c(upstream),
                # c > 0, so the signs of fl are those of f
                fl0 = models[0].hflux(c, ie, l, dt)
                if int(np.sum(fl0 < 0)) != 1 or not fl0[ic] < 0:
                    continue                                 # kc must be the single upstream vertex
                f = fl0 / c[kc, l - 1]
                fl1 = models[1].hflux(c, ie, l, dt)
                exp = np.zeros(3)
                cc = c[kc, l - 1]
                for ii in range(3):
                    if ii == ic:
                        continue
                    d = np.array([x[kn[ii]] - x[kc], y[kn[ii]] - y[kc]])
                    alfa = dt * abs(u * d[0] + v * d[1]) / float(d @ d)
                    conf = cc + 0.5 * (c[kn[ii], l - 1] - cc) * (1.0 - alfa)
                    exp[ic] -= conf * f[ii]
                    exp[ii] += conf * f[ii]
                assert np.allclose(fl1, exp, rtol=1e-9, atol=1e-12 * np.max(np.abs(exp))), (kc, ie, l, fl1, exp)
                checked += 1
    assert checked >= 50
