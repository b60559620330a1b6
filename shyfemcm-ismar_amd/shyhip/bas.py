"""Basin files (.bas), the reference's binary grid format (module basin,
utils/shyutil/basin.f90:403-483): Fortran unformatted sequential records,
version 6 as basin_write_internal writes it --

    ftype (789233567), nvers
    nkn, nel, ngr, mbw
    dcorbas, dirnbas
    sphebas                          (version > 5)
    descrr                           (80 characters)
    nen3v(3, nel), ipv(nkn), ipev(nel), iarv(nel)
    xgv(nkn), ygv(nkn), hm3v(3, nel)
    npart_node, npart_elem           (version >= 5), then the partitions

read_bas returns a grd.Grd (external numbers, vertex depths), so
grd.basin_from_grd builds the device basin from either format.

    g = read_bas("venlag62.bas")
    b = grd.basin_from_grd(g, levels, bounds=[...])
"""
from __future__ import annotations

import struct

import numpy as np

from .grd import NO_DEPTH, Grd, GrdError

BAS_FTYPE = 789233567
BAS_VERSION = 6


def _rec(f, payload: bytes):
    n = struct.pack("<i", len(payload))
    f.write(n + payload + n)


def _records(path):
    out = []
    with open(path, "rb") as f:
        data = f.read()
    pos = 0
    while pos < len(data):
        if pos + 4 > len(data):
            raise GrdError(f"{path}: truncated record marker")
        (n,) = struct.unpack_from("<i", data, pos)
        end = pos + 4 + n
        if n < 0 or end + 4 > len(data) or struct.unpack_from("<i", data, end)[0] != n:
            raise GrdError(f"{path}: bad record at byte {pos}")
        out.append(data[pos + 4:end])
        pos = end + 4
    return out


def grade_and_bandwidth(nkn, nen3v):
    """ngr (most neighbours of a node) and mbw (largest node-number
    difference within an element), for 0-based nen3v."""
    e = np.asarray(nen3v, np.int64)
    pairs = np.concatenate([e[:, [0, 1]], e[:, [1, 2]], e[:, [2, 0]]])
    pairs = np.unique(np.sort(pairs, axis=1), axis=0)
    deg = np.bincount(pairs.ravel(), minlength=nkn)
    mbw = int((e.max(axis=1) - e.min(axis=1)).max()) if len(e) else 0
    return int(deg.max()) if nkn else 0, mbw


def write_bas(path, basin, descr="", dcor=0.0, dirn=0.0, sphe=0, iarv=None):
    """Writes a Basin (its node and element order, external numbers from
    basin.meta, vertex depths hm3v) as a version-6 basin file."""
    b = basin
    meta = getattr(b, "meta", {}) or {}
    nen = np.asarray(b.nen3v, np.int32)
    ngr, mbw = grade_and_bandwidth(b.nkn, nen)
    ipv = np.asarray(meta.get("node_ext", np.arange(1, b.nkn + 1)), "<i4")
    ipev = np.asarray(meta.get("elem_ext", np.arange(1, b.nel + 1)), "<i4")
    ia = np.zeros(b.nel, "<i4") if iarv is None else np.asarray(iarv, "<i4")
    text = descr.encode()[:80].ljust(80)
    with open(path, "wb") as f:
        _rec(f, struct.pack("<2i", BAS_FTYPE, BAS_VERSION))
        _rec(f, struct.pack("<4i", b.nkn, b.nel, ngr, mbw))
        _rec(f, struct.pack("<2f", dcor, dirn))
        _rec(f, struct.pack("<i", sphe))
        _rec(f, text)
        _rec(f, (nen + 1).astype("<i4").tobytes())
        for a in (ipv, ipev, ia):
            _rec(f, a.tobytes())
        for a in (b.xgv, b.ygv, b.hm3v):
            _rec(f, np.asarray(a, "<f4").tobytes())
        _rec(f, struct.pack("<2i", 0, 0))


def read_bas(path) -> Grd:
    """Reads versions 4-6 (the ones with a file-type record); returns a Grd with vertex depths and the header
    values in .header (nkn, nel, ngr, mbw, dcor, dirn, sphe, descr)."""
    r = _records(path)
    if len(r) < 2 or len(r[0]) != 8:
        raise GrdError(f"{path}: not a basin file")
    ftype, nvers = struct.unpack("<2i", r[0])
    if ftype != BAS_FTYPE:
        raise GrdError(f"{path}: not a basin file (file type {ftype})")
    if not 4 <= nvers <= BAS_VERSION:          # basin_test (basin.f90:597)
        raise GrdError(f"{path}: basin version {nvers} is not supported")
    nkn, nel, ngr, mbw = struct.unpack("<4i", r[1])
    dcor, dirn = struct.unpack("<2f", r[2])
    i = 3
    sphe = -1
    if nvers > 5:
        (sphe,) = struct.unpack("<i", r[i])
        i += 1
    descr = r[i].decode(errors="replace").rstrip()
    i += 1
    ints = lambda b, n: np.frombuffer(b, "<i4", count=n).astype(np.int64)      # noqa: E731
    flts = lambda b, n: np.frombuffer(b, "<f4", count=n).copy()                # noqa: E731
    try:
        nen = ints(r[i], 3 * nel).reshape(nel, 3)
        ipv, ipev, iarv = ints(r[i + 1], nkn), ints(r[i + 2], nel), ints(r[i + 3], nel)
        x, y = flts(r[i + 4], nkn), flts(r[i + 5], nkn)
        hm3v = flts(r[i + 6], 3 * nel).reshape(nel, 3)
    except (IndexError, ValueError):
        raise GrdError(f"{path}: truncated basin file") from None
    if nen.min() < 1 or nen.max() > nkn:
        raise GrdError(f"{path}: element refers to a node outside 1..{nkn}")
    g = Grd(node_id=ipv, node_type=np.zeros(nkn, np.int32), x=x, y=y, node_depth=np.full(nkn, NO_DEPTH, np.float32),
            elem_id=ipev, elem_type=iarv.astype(np.int32), elem_nodes=ipv[nen - 1],
            elem_depth=np.full(nel, NO_DEPTH, np.float32), vertex_depth=hm3v)
    g.header = dict(nvers=nvers, nkn=nkn, nel=nel, ngr=ngr, mbw=mbw, dcor=dcor, dirn=dirn, sphe=sphe, descr=descr)
    return g
