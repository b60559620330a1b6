"""The oracle's init-time geometry against the reference's own routines.

oracle/_ref/geom_ref runs, compiled from the reference's sources (ref.mk):
set_ev (evgeom.f90), domain_clusterization_ggu (subset.f90), quad_tree_search
at the mirrored points of tvd_upwind_init_elem (quad_tree.f90,
tvd_admin.f90:361-474) and femintp_xi (evgeom.f90:1413-1445).  The oracle's
restatements must return the same bits on the structured tidal basin (whose
mirrored points fall exactly on nodes and edges, so the search's tie rule
matters), jittered lagoons up to the 50k-node workload, and the reference
test case's Venice grid.  Skipped where the reference tree (and so
oracle/_ref) is absent."""
import ctypes as C
import os
import subprocess
import tempfile

import numpy as np
import pytest

import pyoracle
from shyhip import grd, mesh

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GEOM_REF = os.path.join(ROOT, "oracle", "_ref", "geom_ref")
VENICE = os.path.join(ROOT, "tests", "golden", "venlag62.grd")

pytestmark = pytest.mark.skipif(not os.path.exists(GEOM_REF), reason="oracle/_ref/geom_ref not built")


def _run_ref(b, cv):
    """Runs the reference's routines on basin b; returns its arrays in the
    oracle's 0-based layouts."""
    nkn, nel = b.nkn, b.nel
    nen = np.ascontiguousarray(b.nen3v, np.int32)
    e = np.sort(np.concatenate([nen[:, [0, 1]], nen[:, [1, 2]], nen[:, [2, 0]]]), axis=1)
    e = np.unique(e, axis=0)
    ngr = int(np.bincount(e.reshape(-1), minlength=nkn).max())
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(fin, "wb") as f:
            np.array([nkn, nel, ngr], "<i4").tofile(f)
            np.asarray(b.xgv, "<f4").tofile(f)
            np.asarray(b.ygv, "<f4").tofile(f)
            (nen + 1).astype("<i4").tofile(f)
            np.asarray(cv, "<f4").tofile(f)
        subprocess.run([GEOM_REF, fin, fout], check=True, stdout=subprocess.DEVNULL)
        raw = open(fout, "rb").read()
    o = 0

    def take(dt, n):
        nonlocal o
        a = np.frombuffer(raw, dtype=dt, count=n, offset=o)
        o += a.nbytes
        return a
    ev = take("<f8", 19 * nel).reshape(nel, 19)
    ncolor = int(take("<i4", 1)[0])
    color = take("<i4", nel)
    ie = take("<i4", 9 * nel).reshape(nel, 3, 3) - 1        # [ie][ic][id], -1 = none
    x = take("<f4", 9 * nel).reshape(nel, 3, 3)
    y = take("<f4", 9 * nel).reshape(nel, 3, 3)
    cu = take("<f4", 9 * nel).reshape(nel, 3, 3)
    assert o == len(raw)
    return dict(ev=ev, ncolor=ncolor, color=color, tvd_ie=ie, tvd_x=x, tvd_y=y, cu=cu)


BASINS = {
    "tidal_basin": lambda: mesh.tidal_basin(),
    "lagoon3k": lambda: mesh.lagoon(3000, seed=11, lx=8000.0, ly=4000.0, nlv=1, tile=16),
    "venice": lambda: grd.basin_from_grd(grd.read_grd(VENICE), [1.0], name="venlag62"),
    "lagoon50k": lambda: mesh.lagoon(50000, nlv=1),
}


@pytest.fixture(scope="module", params=list(BASINS))
def case(request):
    name = request.param
    b = BASINS[name]()
    rng = np.random.default_rng(7)
    cv = rng.uniform(10.0, 30.0, b.nkn).astype(np.float32)
    ref = _run_ref(b, cv)
    om = pyoracle.OracleModel(b, "f32", params=dict(itvd=2))
    return name, b, cv, ref, om


def test_set_ev_bitexact(case):
    """set_ev (evgeom.f90:164-345): all 19 coefficients of every element."""
    _, b, _, ref, om = case
    assert np.array_equal(om.ev.view(np.uint64), ref["ev"].view(np.uint64))


def test_color_bitexact(case):
    """domain_clusterization_ggu (subset.f90:371-711): the number of colours
    and every element's colour, after the balancing passes."""
    _, b, _, ref, om = case
    assert om.ncolor == ref["ncolor"]
    assert np.array_equal(om.color, ref["color"])


def test_tvd_upwind_points_bitexact(case):
    """tvd_upwind_init_elem (tvd_admin.f90:361-474): the mirrored points and
    the element quad_tree_search returns for each, including its rule for
    points on shared edges (the highest element number)."""
    name, b, _, ref, om = case
    off = ~np.eye(3, dtype=bool)[None].repeat(b.nel, 0)
    assert np.array_equal(om.tvd_ie[off], ref["tvd_ie"][off])
    assert np.array_equal(om.tvd_x[off].view(np.uint32), ref["tvd_x"][off].view(np.uint32))
    assert np.array_equal(om.tvd_y[off].view(np.uint32), ref["tvd_y"][off].view(np.uint32))
    if name == "tidal_basin":
        # the structured grid puts mirrored points on nodes and edges: the tie
        # rule decides many of them
        assert np.count_nonzero(ref["tvd_ie"][off] >= 0) > 0.8 * off.sum()


def test_tvd_upwind_values_bitexact(case):
    """tvd_get_upwind_c's value (femintp_xi at the upwind point, xy2xi with
    adjust_xi) for a random nodal field."""
    _, b, cv, ref, om = case
    out = np.zeros((b.nel, 3, 3), np.float32)
    om.lib.oc_tvd_upwind_table(C.byref(om.mesh), 1, pyoracle.ptr(np.ascontiguousarray(cv[:, None])),
                               pyoracle.ptr(out))
    assert np.array_equal(out.view(np.uint32), ref["cu"].view(np.uint32))
