"""Serving helpers on the host: request streams, forcing, footprints."""
import numpy as np

from shyhip import serve


def test_stream_is_seeded_and_poisson():
    a = serve.poisson_stream(6, rate=4.0, seed=3, sizes=(2000, 6000), steps=(10, 20), nlv=4)
    b = serve.poisson_stream(6, rate=4.0, seed=3, sizes=(2000, 6000), steps=(10, 20), nlv=4)
    assert [r.arrival for r in a] == [r.arrival for r in b]
    assert all(np.array_equal(x.basin.nen3v, y.basin.nen3v) for x, y in zip(a, b))
    arr = np.array([r.arrival for r in a])
    assert np.all(np.diff(arr) > 0)
    assert all(1800 <= r.basin.nkn <= 6600 and 10 <= r.nsteps <= 20 for r in a)
    assert all(0.3 <= r.surge <= 1.2 for r in a)


def test_forcing_ramps_wind_and_surge():
    r = serve.poisson_stream(1, rate=1.0, seed=5, sizes=(2000, 2001), steps=(10, 10), nlv=4)[0]
    tx0, ty0, l0 = serve.forcing(r, 1, 60.0)
    tx1, ty1, l1 = serve.forcing(r, r.ramp_steps, 60.0)
    tx2, ty2, _ = serve.forcing(r, 10 * r.ramp_steps, 60.0)
    assert abs(tx1) + abs(ty1) > abs(tx0) + abs(ty0)
    assert (tx1, ty1) == (tx2, ty2)                 # ramp complete
    # full wind stress of 30 m/s: (rho_air/rho_water) cd |w|^2
    assert np.isclose(np.hypot(tx1, ty1), 1.225 / 1025.0 * 2.5e-3 * 900.0, rtol=1e-5)


def test_footprint_grows_with_size_and_precision():
    a, b = serve.poisson_stream(2, rate=1.0, seed=9, sizes=(2000, 8000), steps=(10, 10), nlv=4)
    small, big = sorted((a.basin, b.basin), key=lambda x: x.nkn)
    assert serve.footprint_bytes(big, "fp32") > serve.footprint_bytes(small, "fp32")
    assert serve.footprint_bytes(small, "fp64") > serve.footprint_bytes(small, "fp32")
