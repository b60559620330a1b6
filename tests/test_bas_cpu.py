"""Basin files (.bas, shyhip.bas): the Venice grid written by write_bas is
read back (a) by read_bas and (b) by the reference's own reader, module
basin of utils/shyutil/basin.f90 compiled from the reference tree into
oracle/_ref/bas_dump (skipped where that was not built).  Both must return
what was written, bit for bit; a basin rebuilt from the file has the same
elements, depths and layers, keyed by external numbers."""
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "shyfemcm-ismar_amd"))
from shyhip import bas, grd  # noqa: E402

DUMP = os.path.join(ROOT, "oracle", "_ref", "bas_dump")
VENICE = os.path.join(ROOT, "tests", "golden", "venlag62.grd")
LEVELS = [1, 2, 3, 4, 5, 6, 8, 10, 12, 14, 16, 20, 24, 28, 32, 36, 40]
INLETS = [[150, 173, 172, 151], [4311, 4315, 4314, 4309], [2750, 2757, 2758, 2751]]


@pytest.fixture(scope="module")
def venice():
    return grd.basin_from_grd(grd.read_grd(VENICE), LEVELS, bounds=INLETS, name="venlag62")


def _by_ext(b):
    ext_n, ext_e = b.meta["node_ext"], b.meta["elem_ext"]
    nodes = {int(ext_n[k]): (float(b.xgv[k]), float(b.ygv[k])) for k in range(b.nkn)}
    elems = {int(ext_e[i]): (tuple(sorted((int(ext_n[b.nen3v[i, j]]), float(b.hm3v[i, j])) for j in range(3))),
                             int(b.ilhv[i])) for i in range(b.nel)}
    return nodes, elems


def test_round_trip(tmp_path, venice):
    b = venice
    path = str(tmp_path / "venlag62.bas")
    bas.write_bas(path, b, descr="Venice lagoon 62", dcor=45.0, dirn=10.0, sphe=0)
    g = bas.read_bas(path)
    h = g.header
    ngr, mbw = bas.grade_and_bandwidth(b.nkn, b.nen3v)
    assert (h["nvers"], h["nkn"], h["nel"], h["ngr"], h["mbw"]) == (6, b.nkn, b.nel, ngr, mbw)
    assert (h["dcor"], h["dirn"], h["sphe"], h["descr"]) == (45.0, 10.0, 0, "Venice lagoon 62")
    ext = b.meta["node_ext"]
    assert np.array_equal(g.node_id, ext) and np.array_equal(g.elem_id, b.meta["elem_ext"])
    assert np.array_equal(g.elem_nodes, ext[b.nen3v])
    assert np.array_equal(g.x, b.xgv) and np.array_equal(g.y, b.ygv)
    assert np.array_equal(g.vertex_depth, b.hm3v)
    # the basin rebuilt from the file: same nodes, elements, depths, layers
    b2 = grd.basin_from_grd(g, LEVELS, bounds=INLETS)
    assert _by_ext(b2) == _by_ext(b)
    assert sorted(b2.meta["node_ext"][b2.bnd_nodes].tolist()) == sorted(ext[b.bnd_nodes].tolist())


def test_reference_reader(tmp_path, venice):
    if not os.path.exists(DUMP):
        pytest.skip("the reference's basin reader was not built (oracle/_ref/bas_dump)")
    b = venice
    path = str(tmp_path / "v.bas")
    bas.write_bas(path, b, descr="Venice lagoon 62", dcor=45.0, dirn=10.0, sphe=0)
    out = subprocess.run([DUMP, path], capture_output=True, text=True, check=True, cwd=tmp_path).stdout
    d = {}
    for line in out.splitlines():
        if line.startswith("descr|"):
            d["descr"] = line[6:-1].rstrip()
            continue
        k, *v = line.split()
        d[k] = np.array(v, np.float64)
    ngr, mbw = bas.grade_and_bandwidth(b.nkn, b.nen3v)
    assert d["params"].astype(int).tolist() == [b.nkn, b.nel, ngr, mbw]
    assert d["geom"].tolist() == [45.0, 10.0, 0.0] and d["descr"] == "Venice lagoon 62"
    assert np.array_equal(d["nen3v"].astype(int), (b.nen3v + 1).ravel())
    assert np.array_equal(d["ipv"].astype(int), b.meta["node_ext"])
    assert np.array_equal(d["ipev"].astype(int), b.meta["elem_ext"])
    assert np.array_equal(d["iarv"], np.zeros(b.nel))
    assert np.array_equal(d["xgv"].astype(np.float32), b.xgv)
    assert np.array_equal(d["ygv"].astype(np.float32), b.ygv)
    assert np.array_equal(d["hm3v"].astype(np.float32), b.hm3v.ravel())


def test_bad_files(tmp_path, venice):
    p = tmp_path / "x.bas"
    p.write_bytes(b"\x08\x00\x00\x00" + b"\x01\x00\x00\x00\x06\x00\x00\x00" + b"\x08\x00\x00\x00")
    with pytest.raises(grd.GrdError, match="not a basin file"):
        bas.read_bas(str(p))
    good = tmp_path / "g.bas"
    bas.write_bas(str(good), venice)
    p.write_bytes(good.read_bytes()[:5000])
    with pytest.raises(grd.GrdError):
        bas.read_bas(str(p))
