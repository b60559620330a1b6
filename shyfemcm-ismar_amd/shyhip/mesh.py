"""Synthetic basins, layer structures and forcing for the hydro step.

Produces the inputs the SHYFEM time step consumes, in the reference's own
array layouts (Fortran column-major, layer index fastest; node and element
numbers are 0-based here and +1 at the Fortran boundary):

* ``nen3v(3,nel)``, ``xgv/ygv(nkn)``          basin (src/utils/shyutil/basin.f90:122-124)
* ``hm3v(3,nel)``, ``hev(nel)``               depths (depth on element vertices)
* ``hlv/hldv(nlv)``, ``ilhv(nel)``, ``ilhkv(nkn)``  z-layer structure as set by
  set_ilhv / set_ilhkv (src/shyfem/initialize.f90:787-881)
* open boundaries of type 1 (prescribed level) and the static
  external-boundary flag used by the TVD scheme.

Every generator is seeded with a uint64 seed; the same seed gives the same
basin.  Two families are provided:

* ``tidal_basin``   -- structured rectangle split into right triangles
  (41x21 nodes, 1600 elements for the 2D tidal case),
* ``lagoon``        -- jittered points + Delaunay triangulation with a smooth
  log-normal depth field and intertidal flats (the 3D lagoon family).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

FLAG = np.float32(-9988765.0)          # mkonst flag, src/shyfem/const.f90:119
GRAV = 9.81
ROWASS = 1025.0
ROLUFT = 1.225


@dataclass
class Basin:
    nkn: int
    nel: int
    nlv: int
    xgv: np.ndarray          # (nkn,) f32
    ygv: np.ndarray
    nen3v: np.ndarray        # (nel,3) int32, 0-based, counter-clockwise
    hm3v: np.ndarray         # (nel,3) f32
    hev: np.ndarray          # (nel,) f32
    hlv: np.ndarray          # (nlv,) f32 bottom depth of layer l
    hldv: np.ndarray         # (nlv,) f32 layer thickness
    ilhv: np.ndarray         # (nel,) int32 number of layers in element
    ilhkv: np.ndarray        # (nkn,) int32
    ibnd: np.ndarray         # (nkn,) int32 open boundary number (0 none)
    ibtyp: np.ndarray        # (nbnd+1,) int32 boundary type (index 0 unused)
    iextb: np.ndarray        # (nkn,) int32 node on the external boundary
    bnd_nodes: np.ndarray    # (nb,) int32 nodes of boundary 1
    seed: int = 0
    name: str = ""
    meta: dict = field(default_factory=dict)


def _layers(nlv: int, dz: float = 1.0):
    if nlv == 1:
        hlv = np.array([10000.0], dtype=np.float32)
    else:
        hlv = (dz * np.arange(1, nlv + 1)).astype(np.float32)   # z-layers of dz metres
    hldv = np.diff(np.concatenate([[0.0], hlv])).astype(np.float32)
    return hlv, hldv


def _set_levels(hm3v, hlv, nlv):
    """set_ilhv (initialize.f90:787-840) with a last-layer adjustment that
    keeps every vertex of a multi-layer element below the top of its last
    layer (the role of set_last_layer, initialize.f90:1021)."""
    hm3v = hm3v.astype(np.float32).copy()
    for _ in range(4):
        hev = (hm3v.astype(np.float32).sum(axis=1, dtype=np.float32) / np.float32(3.0)).astype(np.float32)
        if nlv == 1:
            ilhv = np.ones(len(hev), dtype=np.int32)
            break
        ilhv = np.searchsorted(hlv, hev, side="left").astype(np.int32) + 1
        ilhv = np.minimum(ilhv, nlv).astype(np.int32)
        multi = ilhv > 1
        floor = np.zeros_like(hev)
        floor[multi] = hlv[ilhv[multi] - 2] + np.float32(0.1) * (hlv[ilhv[multi] - 1] - hlv[ilhv[multi] - 2])
        new = np.maximum(hm3v, floor[:, None]).astype(np.float32)
        if np.array_equal(new, hm3v):
            break
        hm3v = new
    hev = (hm3v.sum(axis=1, dtype=np.float32) / np.float32(3.0)).astype(np.float32)
    return hm3v, hev, ilhv


def _ilhkv(nkn, nen3v, ilhv):
    out = np.zeros(nkn, dtype=np.int32)
    np.maximum.at(out, nen3v.reshape(-1), np.repeat(ilhv, 3))
    return out


def _external_boundary(nkn, nen3v):
    """Nodes on edges that belong to exactly one element."""
    e = np.concatenate([nen3v[:, [0, 1]], nen3v[:, [1, 2]], nen3v[:, [2, 0]]])
    e = np.sort(e, axis=1)
    key = e[:, 0].astype(np.int64) * nkn + e[:, 1]
    uniq, cnt = np.unique(key, return_counts=True)
    b = uniq[cnt == 1]
    flag = np.zeros(nkn, dtype=np.int32)
    flag[(b // nkn).astype(np.int64)] = 1
    flag[(b % nkn).astype(np.int64)] = 1
    return flag


def tidal_basin(seed: int = 20261202, nx: int = 41, ny: int = 21, dx: float = 250.0,
                depth: float = 10.0, noise: float = 1.0, nlv: int = 1) -> Basin:
    """Rectangular basin, nx*ny structured nodes, each cell split into two
    counter-clockwise right triangles along the (i,j)-(i+1,j+1) diagonal.
    Depth = depth + U(-noise, noise) per node; open boundary (type 1) on the
    nodes at x = 0."""
    rng = np.random.default_rng(np.uint64(seed))
    ii, jj = np.meshgrid(np.arange(nx), np.arange(ny), indexing="xy")
    x = (ii * dx).astype(np.float32).reshape(-1)
    y = (jj * dx).astype(np.float32).reshape(-1)
    nkn = nx * ny
    k = lambda i, j: i + nx * j  # noqa: E731
    tris = []
    for j in range(ny - 1):
        for i in range(nx - 1):
            k00, k10, k01, k11 = k(i, j), k(i + 1, j), k(i, j + 1), k(i + 1, j + 1)
            tris.append((k00, k10, k11))
            tris.append((k00, k11, k01))
    nen3v = np.array(tris, dtype=np.int32)
    hnode = (depth + rng.uniform(-noise, noise, nkn)).astype(np.float32)
    hm3v = hnode[nen3v]
    hlv, hldv = _layers(nlv)
    hm3v, hev, ilhv = _set_levels(hm3v, hlv, nlv)
    ibnd = np.zeros(nkn, dtype=np.int32)
    bnd = np.array([k(0, j) for j in range(ny)], dtype=np.int32)
    ibnd[bnd] = 1
    return Basin(nkn=nkn, nel=len(nen3v), nlv=nlv, xgv=x, ygv=y, nen3v=nen3v, hm3v=hm3v,
                 hev=hev, hlv=hlv, hldv=hldv, ilhv=ilhv, ilhkv=_ilhkv(nkn, nen3v, ilhv),
                 ibnd=ibnd, ibtyp=np.array([0, 1], dtype=np.int32),
                 iextb=_external_boundary(nkn, nen3v), bnd_nodes=bnd, seed=seed,
                 name=f"tidal_basin_{nx}x{ny}")


def _smooth_field(rng, nx, ny, sigma):
    from scipy.ndimage import gaussian_filter
    g = gaussian_filter(rng.standard_normal((ny, nx)), sigma=sigma, mode="reflect")
    g = (g - g.mean()) / (g.std() + 1e-12)
    return g


def lagoon(nkn_target: int = 50000, seed: int = 20261202, lx: float = 40000.0, ly: float = 20000.0,
           nlv: int = 12, median: float = 1.5, sigma_ln: float = 1.06,
           hmin: float = 0.05, hmax: float = 20.0, jitter: float = 0.3,
           tile: int = 32, inlet_depth: float = 2.0, dz: float = 1.0) -> Basin:
    """Jittered-point Delaunay lagoon with a smooth log-normal depth.

    Nodes are numbered tile by tile (tile x tile grid points), so contiguous
    node ranges are compact patches of the basin.  The western side (x = 0)
    is an open boundary of type 1 in water at least inlet_depth deep: the
    depth floor is inlet_depth up to 3 grid steps from the boundary and
    falls linearly to zero at 8 steps, so boundary nodes stay wet through
    the tide (a dry open-boundary node stops the model, topo_update.f90
    setnod)."""
    from scipy.spatial import Delaunay
    rng = np.random.default_rng(np.uint64(seed))
    ny = max(3, int(round(math.sqrt(nkn_target * ly / lx))))
    nx = max(3, int(round(nkn_target / ny)))
    dx, dy = lx / (nx - 1), ly / (ny - 1)
    ii, jj = np.meshgrid(np.arange(nx), np.arange(ny), indexing="xy")
    x = ii * dx + rng.uniform(-jitter, jitter, ii.shape) * dx
    y = jj * dy + rng.uniform(-jitter, jitter, jj.shape) * dy
    x[:, 0] = 0.0
    x[:, -1] = lx
    y[0, :] = 0.0
    y[-1, :] = ly
    # tile-major numbering
    ti, tj = ii // tile, jj // tile
    ntx = (nx + tile - 1) // tile
    order_key = ((tj * ntx + ti) * tile * tile + (jj % tile) * tile + (ii % tile)).reshape(-1)
    perm = np.argsort(order_key, kind="stable")
    xs = x.reshape(-1)[perm].astype(np.float32)
    ys = y.reshape(-1)[perm].astype(np.float32)
    g = _smooth_field(rng, nx, ny, sigma=6.0).reshape(-1)[perm]
    hnode = np.clip(median * np.exp(sigma_ln * g), hmin, hmax)
    floor = inlet_depth * np.clip((8.0 * dx - xs.astype(np.float64)) / (5.0 * dx), 0.0, 1.0)
    hnode = np.maximum(hnode, floor).astype(np.float32)
    tri = Delaunay(np.stack([xs.astype(np.float64), ys.astype(np.float64)], axis=1))
    nen3v = tri.simplices.astype(np.int32)
    x1, y1 = xs[nen3v[:, 0]].astype(np.float64), ys[nen3v[:, 0]].astype(np.float64)
    x2, y2 = xs[nen3v[:, 1]].astype(np.float64), ys[nen3v[:, 1]].astype(np.float64)
    x3, y3 = xs[nen3v[:, 2]].astype(np.float64), ys[nen3v[:, 2]].astype(np.float64)
    area2 = (x2 - x1) * (y3 - y1) - (x3 - x1) * (y2 - y1)
    cw = area2 < 0
    nen3v[cw] = nen3v[cw][:, [0, 2, 1]]
    area2 = np.abs(area2)
    keep = area2 > 1e-6 * dx * dy
    nen3v = nen3v[keep]
    # order elements by their first (smallest) node for locality
    eorder = np.lexsort((nen3v.max(axis=1), nen3v.min(axis=1)))
    nen3v = np.ascontiguousarray(nen3v[eorder])
    hm3v = hnode[nen3v]
    hlv, hldv = _layers(nlv, dz)
    hm3v, hev, ilhv = _set_levels(hm3v, hlv, nlv)
    nkn = len(xs)
    ibnd = np.zeros(nkn, dtype=np.int32)
    bnd = np.nonzero(xs == 0.0)[0].astype(np.int32)
    bnd = bnd[np.argsort(ys[bnd])]
    ibnd[bnd] = 1
    return Basin(nkn=nkn, nel=len(nen3v), nlv=nlv, xgv=xs, ygv=ys, nen3v=nen3v, hm3v=hm3v,
                 hev=hev, hlv=hlv, hldv=hldv, ilhv=ilhv, ilhkv=_ilhkv(nkn, nen3v, ilhv),
                 ibnd=ibnd, ibtyp=np.array([0, 1], dtype=np.int32),
                 iextb=_external_boundary(nkn, nen3v), bnd_nodes=bnd, seed=seed,
                 name=f"lagoon_{nkn}", meta=dict(nx=nx, ny=ny, dx=dx, dy=dy))


# ---------------------------------------------------------------- forcing

M2_PERIOD = 44712.0        # 12.42 h
K1_PERIOD = 86148.0        # 23.93 h


def tide_level(t: float, amps=(0.5,), periods=(M2_PERIOD,), phases=(0.0,)) -> float:
    return float(sum(a * math.sin(2 * math.pi * t / p + ph) for a, p, ph in zip(amps, periods, phases)))


def wind_stress(speed: float, direction: float, dragco: float = 2.5e-3):
    """Uniform wind stress per unit density, meteo_forcing.f90:801-802,973:
    tau = (roluft/rowass) * cd * |w| * w."""
    wfact = np.float32(ROLUFT) / np.float32(ROWASS)
    wx = np.float32(speed * math.cos(direction))
    wy = np.float32(speed * math.sin(direction))
    sp = np.float32(math.sqrt(float(wx) ** 2 + float(wy) ** 2))
    return (np.float32(wfact * np.float32(dragco) * sp * wx), np.float32(wfact * np.float32(dragco) * sp * wy))


def smooth_tracer(basin: Basin, seed: int, lo: float, hi: float) -> np.ndarray:
    """Seeded smooth field on the nodes, lo..hi, layer-independent apart from
    a small linear stratification; returned as (nkn, nlv) f32."""
    rng = np.random.default_rng(np.uint64(seed))
    kx, ky = rng.uniform(0.5, 2.0, 2) * 2 * math.pi / 20000.0
    ph = rng.uniform(0, 2 * math.pi, 2)
    base = 0.5 + 0.25 * np.sin(kx * basin.xgv + ph[0]) + 0.25 * np.cos(ky * basin.ygv + ph[1])
    out = np.empty((basin.nkn, basin.nlv), dtype=np.float32)
    for l in range(basin.nlv):
        out[:, l] = (lo + (hi - lo) * np.clip(base - 0.02 * l, 0, 1)).astype(np.float32)
    return out
