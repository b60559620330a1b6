"""set_mass_flux (bnd_routines.f90:1035-1183) in the oracle, pinned by
hand-computed cases and by what the routine must conserve: a source's
discharge is split over its layers in proportion to the layer volumes
hdknv*areakv, levmin/levmax select the layers (negative values count from
the bottom), sources at one node add, the distributed source goes to the
top layer, rqv is the column sum, and a bad layer range or a zero volume
is the reference's error stop."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import pyoracle as po  # noqa: E402

# one column of three layers with volumes 10, 20, 30 and a shallower
# column of two layers
HD = np.array([[1.0, 2.0, 3.0], [1.0, 1.0, 0.0]])
AR = np.full((2, 3), 10.0)
IL = np.array([3, 2])


def _run(nodes, lo, hi, q, rq=None, precision="f32", hd=HD, il=IL):
    return po.set_mass_flux(il, hd, AR[: len(il)], nodes, lo, hi, q, rq, precision=precision)


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_volume_weights(precision):
    mf, rqv = _run([0], [0], [0], [6.0], precision=precision)
    assert np.allclose(mf[0], [1.0, 2.0, 3.0], rtol=1e-6) and np.all(mf[1] == 0)
    assert rqv[0] == pytest.approx(6.0, rel=1e-6) and rqv[1] == 0


@pytest.mark.parametrize("lo,hi,want", [
    (2, 0, [0.0, 2.4, 3.6]),        # layers 2..3: volumes 20, 30
    (0, 1, [6.0, 0.0, 0.0]),        # top layer only
    (-1, 0, [0.0, 0.0, 6.0]),       # levmin -1: the bottom layer
    (0, -2, [2.0, 4.0, 0.0]),       # levmax -2: all but the bottom layer
    (0, 9, [1.0, 2.0, 3.0]),        # levmax beyond the column: clipped
])
def test_layer_selection(lo, hi, want):
    mf, rqv = _run([0], [lo], [hi], [6.0])
    assert np.allclose(mf[0], want, rtol=1e-6, atol=1e-7)
    assert rqv[0] == pytest.approx(6.0, rel=1e-6)


def test_shallow_column_and_two_sources():
    # node 1 has two layers of volume 10 each; two sources at node 0 add
    mf, rqv = _run([1, 0, 0], [0, 0, -1], [3, 0, 0], [4.0, 6.0, 3.0])
    assert np.allclose(mf[1], [2.0, 2.0, 0.0], rtol=1e-6)
    assert np.allclose(mf[0], [1.0, 2.0, 6.0], rtol=1e-6)
    assert np.allclose(rqv, [9.0, 4.0], rtol=1e-6)


def test_distributed_source_top_layer():
    mf, rqv = _run([], [], [], [], rq=[0.5, -0.25])
    assert mf[0].tolist() == [0.5, 0.0, 0.0] and mf[1].tolist() == [-0.25, 0.0, 0.0]
    assert rqv.tolist() == [0.5, -0.25]
    mf, rqv = _run([0], [0], [0], [6.0], rq=[0.5, 0.0])
    assert np.allclose(mf[0], [1.5, 2.0, 3.0], rtol=1e-6) and rqv[0] == pytest.approx(6.5, rel=1e-6)


@pytest.mark.parametrize("lo,hi", [(3, 2), (2, -3)])
def test_bad_layer_range_stops(lo, hi):
    assert _run([0], [lo], [hi], [1.0]) is None


def test_zero_volume_stops():
    hd = HD.copy()
    hd[0] = 0.0
    assert _run([0], [0], [0], [1.0], hd=hd) is None


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_random_conservation(precision):
    rng = np.random.default_rng(7)
    nkn, nl = 400, 6
    il = rng.integers(1, nl + 1, nkn)
    hd = rng.uniform(0.2, 2.0, (nkn, nl)) * (np.arange(nl)[None, :] < il[:, None])
    ar = rng.uniform(1e3, 1e4, (nkn, nl))
    nodes = rng.integers(0, nkn, 60)
    lo = rng.choice([0, 1, -1], 60)
    hi = rng.choice([0, 2, -1], 60)
    q = rng.normal(0, 5, 60)
    rq = rng.normal(0, 0.1, nkn)
    r = po.set_mass_flux(il, hd, ar, nodes, np.zeros(60), np.zeros(60), q, rq, precision=precision)
    mf, rqv = r
    dt = np.float32 if precision == "f32" else np.float64
    tol = 1e-5 if precision == "f32" else 1e-12
    # every column's parts sum to its sources, nothing below the column
    want = np.zeros(nkn)
    np.add.at(want, nodes, q.astype(np.float32).astype(np.float64))
    want += rq.astype(np.float32)
    assert np.allclose(mf.sum(axis=1), want, rtol=tol, atol=tol * 10)
    assert np.all(mf[np.arange(nl)[None, :] >= il[:, None]] == 0)
    assert np.array_equal(rqv, np.array([mf[k, :il[k]].astype(np.float64).sum() for k in range(nkn)]).astype(dt))
    # with layer limits the parts stay inside the limits
    r2 = po.set_mass_flux(il, hd, ar, nodes, lo, hi, q, None, precision=precision)
    if r2 is not None:
        mf2, _ = r2
        assert np.allclose(mf2.sum(axis=1), want - rq.astype(np.float32), rtol=tol, atol=tol * 10)
