"""The oracle's forcing restatement (oracle/shyoracle_forcing.c) pinned by
what the mathematics fixes: Lagrange interpolation through the window's
records (polynomial reproduction, an independent barycentric evaluation,
data returned at the records), the window the reference's ring selects
(iff_populate_records + iff_read_and_interpolate: for the cubic window two
records before t and two at or after it, for the linear one and one), and
the wind-to-stress closed forms of meteo_convert_wind_data (stress =
(roluft/rowass) cd |w| w, the Smith-Banke and Large-Pond drag laws, the
speed limit, the wsmax stop, stress input inverted to the same wind).
The reference holds no vectors for these routines."""
import math
import os
import sys

import numpy as np
import pytest
from scipy.interpolate import BarycentricInterpolator

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "shyfemcm-ismar_amd"))
import pyoracle as po  # noqa: E402


@pytest.mark.parametrize("n", [2, 4])
def test_neville_reproduces_polynomials(n):
    rng = np.random.default_rng(20261202 + n)
    for _ in range(50):
        xa = np.sort(rng.uniform(0, 1e5, n))
        coef = rng.normal(size=n)
        ya = np.polyval(coef, xa / 1e5).astype(np.float32)       # data as the files hold it (f32)
        for x in rng.uniform(xa[0], xa[-1], 5):
            ref = BarycentricInterpolator(xa, ya.astype(np.float64))(x)
            got = po.intp_neville(xa, ya, x)
            assert abs(got - ref) <= 1e-12 * max(1.0, abs(ref))
        # a polynomial of the interpolation's degree is reproduced exactly
        # (up to round-off) from exactly representable data
        ya_e = np.polyval(coef, xa / 1e5).astype(np.float32)
        fit = np.polyfit(xa / 1e5, ya_e.astype(np.float64), n - 1)
        x = float(rng.uniform(xa[0], xa[-1]))
        assert abs(po.intp_neville(xa, ya_e, x) - np.polyval(fit, x / 1e5)) < 1e-9


def test_neville_point_order_and_nodes():
    """The value does not depend on the order of the points beyond
    round-off, and at a record's time it is that record's value."""
    xa = np.array([0.0, 600.0, 1200.0, 1800.0])
    ya = np.array([0.3, -0.1, 0.25, 0.4], np.float32)
    for perm in ([0, 1, 2, 3], [2, 3, 0, 1], [3, 1, 0, 2]):
        v = po.intp_neville(xa[perm], ya[perm], 700.0)
        assert abs(v - po.intp_neville(xa, ya, 700.0)) < 1e-15
    for j in range(4):
        assert abs(po.intp_neville(xa, ya, xa[j]) - float(ya[j])) < 1e-15


def test_first_record_rule():
    t = np.arange(0.0, 36000.0, 600.0)
    for m in range(3, 40):
        # start at a record's time: the cubic ring starts two records earlier
        # (two before, two at or after), the linear one record earlier
        assert po.iff_first_record(t, 4, t[m]) == m - 2
        assert po.iff_first_record(t, 2, t[m]) == m - 1
        assert po.iff_first_record(t, 4, t[m] + 100.0) == m - 1
        assert po.iff_first_record(t, 2, t[m] + 100.0) == m
    assert po.iff_first_record(t, 2, 0.0) == 0
    assert po.iff_first_record(t, 4, 0.0) == 0
    assert po.iff_first_record(t, 2, -1.0) == -1          # before the first record
    assert po.iff_first_record(t, 4, t[-1]) == -1         # no full ring left
    assert po.iff_first_record(t[:1], 4, 123.0) == 0      # one record: constant field


@pytest.mark.parametrize("nintp", [2, 4])
def test_window_centres_on_time(nintp):
    rng = np.random.default_rng(7 + nintp)
    time = np.cumsum(rng.uniform(300, 900, 80))
    vals = rng.normal(size=(80, 3)).astype(np.float32)
    t0 = time[5]
    ts = np.sort(rng.uniform(t0, time[-3], 200))
    out, newest = po.iff_series(time, vals, nintp, t0, ts)
    for t, j in zip(ts, newest):
        if nintp == 4:
            # two records before t and two at or after it, unless the ring has
            # not moved yet (then t is before its third record)
            assert time[j - 1] >= t and time[j - 3] <= t
            assert time[j - 2] < t or j == po.iff_first_record(time, 4, t0) + 3
        else:
            assert time[j] >= t and (time[j - 1] < t or j == po.iff_first_record(time, 2, t0) + 1)
    # the value equals the Lagrange interpolant through the window's records
    for k in range(0, len(ts), 17):
        j = newest[k]
        w = slice(j - nintp + 1, j + 1)
        ref = BarycentricInterpolator(time[w], vals[w].astype(np.float64))(ts[k])
        assert np.allclose(out[k], ref.astype(np.float32), rtol=1e-6, atol=1e-6)


def test_series_times_outside_and_flags():
    time = np.arange(0.0, 6000.0, 600.0)
    vals = np.linspace(0, 1, 10, dtype=np.float32)[:, None].repeat(2, axis=1)
    assert po.iff_series(time, vals, 2, 0.0, [0.0, 6000.0]) is None          # after the last record
    assert po.iff_series(time, vals, 4, 3000.0, [1700.0]) is None            # before the ring's first (1800)
    assert po.iff_series(time, vals, 4, 3000.0, [1800.0]) is not None
    vals[4, 1] = -999.0
    out, _ = po.iff_series(time, vals, 2, 0.0, [2500.0, 3000.0, 3100.0], vdefault=0.0)
    assert out[0, 1] == 0.0 and out[1, 1] == 0.0                              # window holds the flag
    assert out[2, 1] != 0.0 and out[0, 0] != 0.0


def test_constant_series():
    out, _ = po.iff_series([100.0], np.array([[0.7, -999.0]], np.float32), 4, 0.0, [0.0, 1e6])
    assert np.array_equal(out, np.array([[0.7, 0.0], [0.7, 0.0]], np.float32))


def test_wind_stress_closed_form():
    rng = np.random.default_rng(3)
    sp = rng.uniform(0, 30, 500)
    d = rng.uniform(0, 2 * math.pi, 500)
    wx, wy = (sp * np.cos(d)).astype(np.float32), (sp * np.sin(d)).astype(np.float32)
    w = np.hypot(wx.astype(np.float64), wy.astype(np.float64))
    f = 1.225 / 1025.0
    tx, ty, ws, _, _, rc = po.convert_wind(wx, wy)
    assert rc == 0
    assert np.allclose(ws, w, rtol=1e-7)
    assert np.allclose(tx, f * 2.5e-3 * w * wx, rtol=2e-6, atol=1e-12)
    assert np.allclose(ty, f * 2.5e-3 * w * wy, rtol=2e-6, atol=1e-12)
    # Smith and Banke (1975), Large and Pond (1981)
    tx1, _, _, _, _, _ = po.convert_wind(wx, wy, itdrag=1)
    assert np.allclose(tx1, f * 1e-3 * (0.63 + 0.066 * w) * w * wx, rtol=2e-6, atol=1e-12)
    tx2, _, _, _, _, _ = po.convert_wind(wx, wy, itdrag=2)
    cd2 = np.where(w > 11.0, 1e-3 * (0.49 + 0.066 * w), 1.2e-3)
    assert np.allclose(tx2, f * cd2 * w * wx, rtol=2e-6, atol=1e-12)


def test_wind_limit_and_stop():
    wx = np.array([5.0, 25.0, -40.0], np.float32)
    wy = np.array([0.0, 10.0, 30.0], np.float32)
    tx, ty, ws, lx, ly, rc = po.convert_wind(wx, wy, wslim=20.0, wsmax=50.0)
    assert rc == 0
    assert ws[0] == 5.0 and ws[1] == 20.0 and ws[2] == 20.0
    for k in (1, 2):       # direction kept, speed capped, stress from the capped wind
        assert abs(math.hypot(lx[k], ly[k]) - 20.0) < 1e-4
        assert abs(math.atan2(ly[k], lx[k]) - math.atan2(wy[k], wx[k])) < 1e-6
        assert abs(tx[k] - 1.225 / 1025.0 * 2.5e-3 * 20.0 * lx[k]) < 1e-9
    assert po.convert_wind(wx, wy, wsmax=45.0)[5] == 1       # 50 m/s > wsmax
    assert po.convert_wind(wx, wy, wslim=20.0, wsmax=45.0)[5] == 0


def test_stress_input_round_trip():
    """iwtype 2 (stress in N/m2): the normalised stress is tau/rowass and
    the derived speed is the wind that gives that stress -- a wind of that
    speed in the stress direction converted with iwtype 1 returns it.  (The
    derived components wx, wy = tx/(cd ws) are diagnostics hydro never
    reads; only the speed is a physical wind.)"""
    rng = np.random.default_rng(11)
    tx_in = rng.uniform(-0.5, 0.5, 200).astype(np.float32)
    ty_in = rng.uniform(-0.5, 0.5, 200).astype(np.float32)
    tx, ty, ws, wx, wy, rc = po.convert_wind(tx_in, ty_in, iwtype=2)
    assert rc == 0
    assert np.allclose(tx, tx_in / 1025.0, rtol=1e-6)
    mag = np.hypot(tx_in.astype(np.float64), ty_in.astype(np.float64))
    assert np.allclose(ws, np.sqrt(mag / (2.5e-3 * 1.225)), rtol=1e-6)
    ux = (ws * tx_in / mag).astype(np.float32)
    uy = (ws * ty_in / mag).astype(np.float32)
    bx, by, _, _, _, _ = po.convert_wind(ux, uy, iwtype=1)
    assert np.allclose(bx, tx, rtol=1e-5, atol=1e-12) and np.allclose(by, ty, rtol=1e-5, atol=1e-12)


def test_uniform_wind_helper_matches_oracle():
    """mesh.wind_stress (the uniform stress the workloads push) is the
    oracle's conversion of a uniform wind, bit for bit."""
    from shyhip import mesh
    for speed, d in ((8.0, 0.9), (10.0, 0.7), (0.0, 1.0), (27.5, -2.0)):
        tx, ty = mesh.wind_stress(speed, d)
        wx = np.float32(speed * math.cos(d))
        wy = np.float32(speed * math.sin(d))
        ox, oy, _, _, _, _ = po.convert_wind([wx], [wy])
        assert ox[0] == tx and oy[0] == ty
