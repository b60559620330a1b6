"""set_mass_flux on the device (shyhip_set_sources, shyhip_set_mass_flux)
against the oracle's restatement of bnd_routines.f90:1035-1183 on the same
layer volumes (bit-identical mfluxv and rqv in both precisions), its error
stop, and steps driven by it equal to steps driven by the same arrays
pushed from the host."""
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


def _sources(b, rng):
    """Seven sources: six nodes (one of them twice) with layer limits."""
    inner = np.setdiff1d(np.arange(b.nkn), b.bnd_nodes)
    nodes = rng.choice(inner, 6, replace=False)
    nodes = np.r_[nodes, nodes[2]]
    lo = np.array([0, 2, 0, -1, 0, 1, 0])
    hi = np.array([0, 0, -2, 0, 1, 3, 0])
    q = rng.uniform(5.0, 50.0, len(nodes))
    return nodes, lo, hi, q


def _lagoon():
    return mesh.lagoon(2000, seed=71, lx=6000.0, ly=3000.0, nlv=5, tile=16)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_mass_flux_matches_oracle(gpu, prec):
    b = _lagoon()
    m = gpu.Model(b, prec, params=PARAMS, visv=1e-3)
    for _ in range(3):
        m.step(bnd_level=0.2)
    rng = np.random.default_rng(11)
    nodes, lo, hi, q = _sources(b, rng)
    rq = rng.normal(0.0, 0.05, b.nkn)
    m.set_sources(nodes, lo, hi)
    m.set_mass_flux(q, rq)
    dt = np.float64 if prec == "fp64" else np.float32
    hd, ar = m.pull("hdknv").astype(dt), m.pull("areakv").astype(dt)
    il = b.ilhkv
    want = po.set_mass_flux(il, hd, ar, nodes, lo, hi, q, rq, nlv=b.nlv, precision="f64" if prec == "fp64" else "f32")
    assert want is not None
    mf, rqv = m.pull("mfluxv").astype(dt), m.pull("rqv").astype(dt)
    assert np.array_equal(mf, want[0])
    assert np.array_equal(rqv, want[1])
    assert abs(rqv.sum() - (q.astype(np.float32).sum() + rq.astype(np.float32).sum())) < 1e-3
    # distributed sources only
    m.set_sources([])
    m.set_mass_flux(None, rq)
    want = po.set_mass_flux(il, hd, ar, [], [], [], [], rq, nlv=b.nlv, precision="f64" if prec == "fp64" else "f32")
    assert np.array_equal(m.pull("mfluxv").astype(dt), want[0])
    m.close()


def test_mass_flux_error_stop(gpu):
    b = _lagoon()
    m = gpu.Model(b, "fp64", params=PARAMS, visv=1e-3)
    k = int(np.setdiff1d(np.arange(b.nkn), b.bnd_nodes)[0])
    m.set_sources([k], [3], [2])
    with pytest.raises(gpu.ShyhipError, match="set_mass_flux"):
        m.set_mass_flux([1.0])
    with pytest.raises(gpu.ShyhipError, match="out of range"):
        m.set_sources([b.nkn], [0], [0])
    m.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_steps_with_sources(gpu, prec):
    """Steps with discharges computed on the device equal steps with the same
    mfluxv and rqv pushed from the host; the discharges raise the volume."""
    b = _lagoon()
    rng = np.random.default_rng(12)
    nodes, lo, hi, q = _sources(b, rng)
    a = gpu.Model(b, prec, params=PARAMS, visv=1e-3)
    c = gpu.Model(b, prec, params=PARAMS, visv=1e-3)
    r = gpu.Model(b, prec, params=PARAMS, visv=1e-3)      # no sources
    a.set_sources(nodes, lo, hi)
    for step in range(6):
        lvl = 0.1 * np.sin(step / 3.0)
        a.set_mass_flux(q * (1 + 0.1 * step))
        mf, rqv = a.pull("mfluxv"), a.pull("rqv")
        a.push("rqv", rqv)               # the same host history (a push restarts the level history)
        c.push("mfluxv", mf)
        c.push("rqv", rqv)
        r.push("rqv", np.zeros(b.nkn))
        for mm in (a, c, r):
            mm.step(bnd_level=lvl)
    for name in ("znv", "utlnv", "vtlnv", "wlnv", "hdknv"):
        assert np.array_equal(a.pull(name), c.pull(name)), name
    assert a.volume() > r.volume()
    for mm in (a, c, r):
        mm.close()
