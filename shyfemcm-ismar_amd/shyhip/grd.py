"""Mesh ingestion: the FEM grid format (.grd) into a basin for the device.

read_grd follows rdgrd (src/utils/generic/grd.f90:594-1000): every line is
a list of numbers whose first one says what it holds -- 0 comment, 1 node
(number, type, x, y[, depth]), 2 element (number, type, vertex count,
vertex numbers[, depth]; the vertex list may continue on the next lines),
3 line (same layout; kept but not used by the model).  Elements must be
triangles.  A line that does not fit its kind is an error, as in rdnode /
rdelem.

basin_from_grd builds the model basin: nodes renumbered for a narrow
bandwidth (reverse Cuthill-McKee, the role of shypre's optimisation) with
the file's numbers kept as external numbers, elements made counter-
clockwise and sorted by their first node, depths from the elements (or the
nodes), z-layers from the level list with the same layer rules as the
generated basins, open boundaries from lists of external node numbers
(the STR file's kbound).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .mesh import Basin, _external_boundary, _ilhkv, _set_levels

NO_DEPTH = -999.0


class GrdError(ValueError):
    pass


@dataclass
class Grd:
    node_id: np.ndarray            # (nkn,) external node numbers, file order
    node_type: np.ndarray
    x: np.ndarray                  # (nkn,) f32
    y: np.ndarray
    node_depth: np.ndarray         # (nkn,) f32, -999 where not given
    elem_id: np.ndarray            # (nel,)
    elem_type: np.ndarray
    elem_nodes: np.ndarray         # (nel, 3) external node numbers
    elem_depth: np.ndarray         # (nel,) f32, -999 where not given
    lines: list = field(default_factory=list)   # (number, type, external nodes, depth)
    comments: int = 0
    vertex_depth: np.ndarray | None = None      # (nel, 3) f32 depth at each vertex (.bas files)


def _numbers(line, lineno):
    try:
        return [float(t) for t in line.split()]
    except ValueError:
        raise GrdError(f"line {lineno}: not a list of numbers: {line.strip()!r}") from None


def read_grd(path) -> Grd:
    nodes, elems, lines = [], [], []
    ncom = 0
    with open(path) as f:
        it = enumerate(f, start=1)
        for lineno, line in it:
            tok = line.split()
            if not tok:
                continue
            if tok[0] == "0":                   # comment (may carry text and FEM parameters)
                ncom += 1
                continue
            v = _numbers(line, lineno)
            kind = int(round(v[0]))
            if kind == 1:
                if not 5 <= len(v) <= 6:
                    raise GrdError(f"line {lineno}: a node needs 5 or 6 values, got {len(v)}")
                nodes.append((int(round(v[1])), int(round(v[2])), v[3], v[4], v[5] if len(v) == 6 else NO_DEPTH))
            elif kind in (2, 3):
                if len(v) < 4:
                    raise GrdError(f"line {lineno}: not enough data")
                num, typ, nvert = int(round(v[1])), int(round(v[2])), int(round(v[3]))
                if kind == 2 and nvert != 3:
                    raise GrdError(f"line {lineno}: element {num} has {nvert} vertices (triangles only)")
                rest = v[4:]
                while len(rest) < nvert:            # the vertex list continues on the next lines
                    try:
                        lineno, more = next(it)
                    except StopIteration:
                        raise GrdError(f"element/line {num}: could not read all vertices") from None
                    rest += _numbers(more, lineno)
                # a depth follows the last vertex on its line; more values are ignored
                verts = [int(round(a)) for a in rest[:nvert]]
                depth = rest[nvert] if len(rest) > nvert else NO_DEPTH
                (elems if kind == 2 else lines).append((num, typ, verts, depth))
            else:
                raise GrdError(f"line {lineno}: unknown item {kind}")
    if not nodes:
        raise GrdError(f"{path}: no nodes")
    nd = np.array([n[0] for n in nodes], np.int64)
    if len(np.unique(nd)) != len(nd):
        raise GrdError("duplicate node numbers")
    return Grd(node_id=nd, node_type=np.array([n[1] for n in nodes], np.int32),
               x=np.array([n[2] for n in nodes], np.float32), y=np.array([n[3] for n in nodes], np.float32),
               node_depth=np.array([n[4] for n in nodes], np.float32),
               elem_id=np.array([e[0] for e in elems], np.int64),
               elem_type=np.array([e[1] for e in elems], np.int32),
               elem_nodes=np.array([e[2] for e in elems], np.int64).reshape(-1, 3),
               elem_depth=np.array([e[3] for e in elems], np.float32),
               lines=[(l[0], l[1], l[2], l[3]) for l in lines], comments=ncom)


def _rcm(nkn, nen3v):
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import reverse_cuthill_mckee
    e = np.concatenate([nen3v[:, [0, 1]], nen3v[:, [1, 2]], nen3v[:, [2, 0]]])
    a = coo_matrix((np.ones(2 * len(e)), (np.r_[e[:, 0], e[:, 1]], np.r_[e[:, 1], e[:, 0]])), shape=(nkn, nkn))
    return np.asarray(reverse_cuthill_mckee(a.tocsr(), symmetric_mode=True), dtype=np.int64)


def basin_from_grd(g: Grd, levels, bounds=(), name: str = "") -> Basin:
    """levels: bottom depths of the z-layers (the STR's $levels; one value
    or none: a 2D basin); bounds: for each open level boundary (ibtyp 1),
    its external node numbers."""
    pos = {int(k): i for i, k in enumerate(g.node_id)}
    try:
        nen = np.array([[pos[int(k)] for k in row] for row in g.elem_nodes], np.int64)
    except KeyError as e:
        raise GrdError(f"element refers to node {e.args[0]}, which the grid does not have") from None
    used = np.zeros(len(g.node_id), bool)
    used[nen.ravel()] = True
    if not used.all():
        raise GrdError(f"{int((~used).sum())} nodes belong to no element")
    nkn = len(g.node_id)
    order = _rcm(nkn, nen)                          # new -> old
    new_of = np.empty(nkn, np.int64)
    new_of[order] = np.arange(nkn)
    nen = new_of[nen]
    x, y = g.x[order], g.y[order]
    # counter-clockwise elements (a clockwise one has its last two vertices swapped)
    ax = x[nen[:, 1]].astype(np.float64) - x[nen[:, 0]]
    ay = y[nen[:, 1]].astype(np.float64) - y[nen[:, 0]]
    bx = x[nen[:, 2]].astype(np.float64) - x[nen[:, 0]]
    by = y[nen[:, 2]].astype(np.float64) - y[nen[:, 0]]
    area2 = ax * by - ay * bx
    if np.any(area2 == 0):
        raise GrdError("degenerate element (zero area)")
    cw = area2 < 0
    nen[cw] = nen[cw][:, [0, 2, 1]]
    eorder = np.lexsort((nen.max(axis=1), nen.min(axis=1)))
    nen = nen[eorder].astype(np.int32)
    if g.vertex_depth is not None:
        vd = np.array(g.vertex_depth, np.float32)
        vd[cw] = vd[cw][:, [0, 2, 1]]
        hm3v = vd[eorder]
    elif np.all(g.elem_depth != NO_DEPTH):
        hm3v = np.repeat(g.elem_depth[eorder][:, None], 3, axis=1).astype(np.float32)
    elif np.all(g.node_depth != NO_DEPTH):
        hm3v = g.node_depth[order][nen].astype(np.float32)
    else:
        raise GrdError("depths are needed on every element or on every node")
    hlv = np.atleast_1d(np.asarray(levels, np.float32))
    if hlv.size <= 1:
        hlv = np.array([10000.0], np.float32)
    nlv = len(hlv)
    hldv = np.diff(np.concatenate([[0.0], hlv])).astype(np.float32)
    hm3v, hev, ilhv = _set_levels(hm3v, hlv, nlv)
    ibnd = np.zeros(nkn, np.int32)
    for b, ks in enumerate(bounds, start=1):
        for k in ks:
            if int(k) not in pos:
                raise GrdError(f"boundary {b}: node {k} is not in the grid")
            ibnd[new_of[pos[int(k)]]] = b
    ibtyp = np.array([0] + [1] * len(bounds), np.int32)
    bnd = np.nonzero(ibnd)[0].astype(np.int32)
    return Basin(nkn=nkn, nel=len(nen), nlv=nlv, xgv=x.astype(np.float32), ygv=y.astype(np.float32), nen3v=nen,
                 hm3v=hm3v, hev=hev, hlv=hlv, hldv=hldv, ilhv=ilhv, ilhkv=_ilhkv(nkn, nen, ilhv), ibnd=ibnd,
                 ibtyp=ibtyp, iextb=_external_boundary(nkn, nen), bnd_nodes=bnd, name=name,
                 meta=dict(node_ext=g.node_id[order], elem_ext=g.elem_id[eorder]))


def oscillation(t, ampli, period, phase, zref=0.0):
    """A level boundary given by a sine in the STR (get_oscil,
    utils/shyutilmpi/bnd_admin.f90:1303-1316): the parameters are single
    precision, dtime is double, so the argument and the sine are double and
    the level is rounded to single at the end."""
    f = np.float32
    twopi = float(f(2.0) * f(3.141592653))
    arg = twopi * (float(t) + float(f(phase))) / float(f(period))
    return f(float(f(zref)) + float(f(ampli)) * math.sin(arg))
