"""GPU parity for the paths the default workloads leave unexercised:

* scalar_stability (stability.f90:220-260, conzstab concentration.f90:
  1573-2097) computed on the device, and scal3sh with several substeps
  (istot >= 2), chosen by the device as the reference chooses them or given
  by the caller;
* the Coriolis-coupled element column (the pentadiagonal elimination of
  penta_fact/penta_solve, penta.f90:104-321, on the 2L x 2L band of
  hydrodynamic.f90:1274-1277), with its coupled final transports.

Bars as in test_gpu_parity: masks bit-exact; zeta 1e-6 m (FP64 vs FP64
oracle), 2e-5 m (FP32 vs reference-precision oracle); tracers 1e-6 of the
field scale (FP64); the stability index equal to the oracle's to round-off
and istot equal.
"""
import numpy as np
import pytest

from shyhip import mesh

pytestmark = pytest.mark.gpu


def _pair(shyhip, basin, precision, oracle_prec, params, **kw):
    import pyoracle
    om = pyoracle.OracleModel(basin, oracle_prec, params=params, **kw)
    gm = shyhip.Model(basin, precision, params=params, **kw)
    return om, gm


def _case(dt, nlv=6, n=3000, seed=7):
    b = mesh.lagoon(n, seed=seed, lx=8000.0, ly=4000.0, nlv=nlv, tile=16)
    tr = dict(temp=mesh.smooth_tracer(b, seed + 1, 15.0, 25.0), salt=mesh.smooth_tracer(b, seed + 2, 25.0, 38.0))
    params = dict(dt=dt, ireib=5, itpre=1, itvd=2, itvdv=1, aapar=0.0, block_rows=0)
    return b, tr, params


def _level(t):
    return mesh.tide_level(t, amps=(0.4, 0.1), periods=(mesh.M2_PERIOD, mesh.K1_PERIOD)) - 0.3


SCALE = dict(temp=25.0, salt=38.0)            # the fields' initial maxima


def _check_tracers(b, gm, om, rel=1e-6):
    valid = np.arange(b.nlv)[None, :] < b.ilhkv[:, None]
    for name in ("temp", "salt"):
        g, o = gm.pull(name), om.tracers[name]
        d = float(np.max(np.abs(g - o)[valid]))
        print(name, "max |dc| =", d)
        assert d < rel * SCALE[name]


# the state scalar_stability reads
STAB_FIELDS = ("utlnv", "utlov", "vtlnv", "vtlov", "wlnv", "hdknv", "hdkov", "hdeov", "difv")


@pytest.mark.parametrize("prec,oprec", [("fp64", "f64"), ("fp32", "f32")])
def test_stability_index_bitexact_on_oracle_state(gpu, prec, oprec):
    """The device's scalar_stability on the oracle's own state (pushed into
    the model after each oracle step) returns the oracle's istot and the
    same bits of rindex, for two values of rstol."""
    b, tr, params = _case(dt=300.0)
    om, gm = _pair(gpu, b, prec, oprec, params, visv=1e-3, difv=1e-4, tracers=tr)
    om.set_wind(*mesh.wind_stress(12.0, 0.7))
    seen = set()
    for n in range(5):
        om.step(_level((n + 1) * 300.0), istot=0)
        for name in STAB_FIELDS:
            gm.push(name, om.a[name])
        for rstol in (1.0, 0.5):
            io, ro = om.scalar_stability(rstol=rstol)
            ig, rg = gm.scalar_stability(rstol=rstol)
            print(f"step {n + 1} rstol {rstol}: oracle istot {io} rindex {ro!r}, device {ig} {rg!r}")
            assert ig == io and rg == ro
            seen.add(io)
    assert max(seen) >= 2                     # the case reaches several substeps


@pytest.mark.parametrize("prec,oprec", [("fp64", "f64"), ("fp32", "f32")])
def test_stability_index_along_runs(gpu, prec, oprec):
    """Device and oracle runs side by side: the same istot after every step
    and rindex within the float rounding the index passes through (the two
    states differ at round-off)."""
    b, tr, params = _case(dt=60.0)
    om, gm = _pair(gpu, b, prec, oprec, params, visv=1e-3, difv=1e-4, tracers=tr)
    tx, ty = mesh.wind_stress(12.0, 0.7)
    om.set_wind(tx, ty)
    gm.set_wind(tx, ty)
    for n in range(6):
        t = (n + 1) * 60.0
        om.step(_level(t), istot=0)
        gm.step(_level(t), istot=0)
        io, ro = om.scalar_stability()
        ig, rg = gm.scalar_stability()
        print(f"step {n + 1}: oracle istot {io} rindex {ro!r}, device {ig} {rg!r}, last step ran {gm.last_istot()}")
        assert ig == io
        assert abs(rg - ro) <= 1e-6 * max(1.0, abs(ro))


@pytest.mark.parametrize("dt,istot", [(300.0, 0), (60.0, 0), (60.0, 12), (60.0, 16)])
def test_substeps_match_oracle(gpu, dt, istot):
    """Steps whose transport runs in istot substeps (0: as scalar_stability
    chooses; given values above what it chooses, so the run stays stable)
    match the oracle: masks, zeta, tracers."""
    b, tr, params = _case(dt=dt)
    om, gm = _pair(gpu, b, "fp64", "f64", params, visv=1e-3, difv=1e-4, tracers=tr)
    tx, ty = mesh.wind_stress(12.0, 0.7)
    om.set_wind(tx, ty)
    gm.set_wind(tx, ty)
    ran = []
    for n in range(6):
        t = (n + 1) * dt
        om.step(_level(t), istot=istot)
        gm.step(_level(t), istot=istot)
        ran.append(gm.last_istot())
    print("substeps run by the device:", ran)
    assert np.array_equal(gm.pull("iwegv"), om.a["iwegv"])
    assert float(np.max(np.abs(gm.pull("znv") - om.a["znv"]))) < 1e-6
    _check_tracers(b, gm, om)
    if istot == 0:
        assert max(ran) >= 2
    else:
        assert ran == [istot] * 6


@pytest.mark.parametrize("prec,oprec,ztol", [("fp64", "f64", 1e-6), ("fp32", "f32", 2e-5)])
def test_coriolis_coupled_columns(gpu, prec, oprec, ztol):
    """fcorv != 0 switches every element column to the coupled u/v band
    (pentadiagonal elimination) and the final transports to their coupled
    form; 12 steps of the 6-layer lagoon under wind match the oracle."""
    b, tr, params = _case(dt=60.0)
    f = 1.0e-4
    om, gm = _pair(gpu, b, prec, oprec, params, visv=1e-3, difv=1e-4, tracers=tr, fcor=f)
    tx, ty = mesh.wind_stress(10.0, 0.3)
    om.set_wind(tx, ty)
    gm.set_wind(tx, ty)
    for n in range(12):
        t = (n + 1) * 60.0
        om.step(_level(t))
        gm.step(_level(t))
    assert np.array_equal(gm.pull("iwegv"), om.a["iwegv"])
    dz = float(np.max(np.abs(gm.pull("znv") - om.a["znv"])))
    print(prec, "coriolis 12 steps |dz| =", dz, gm.stats())
    assert dz < ztol
    for name in ("utlnv", "vtlnv"):
        g, o = gm.pull(name), om.a[name]
        assert float(np.max(np.abs(g - o))) < 1e-4 * float(np.abs(o).max()), name
    # the rotation matters: the same run without Coriolis differs
    om0, gm0 = _pair(gpu, b, prec, oprec, params, visv=1e-3, difv=1e-4, tracers=tr, fcor=0.0)
    gm0.set_wind(tx, ty)
    for n in range(12):
        gm0.step(_level((n + 1) * 60.0))
    assert float(np.max(np.abs(gm0.pull("vtlnv") - gm.pull("vtlnv")))) > 1e-3 * float(np.abs(om.a["vtlnv"]).max())
    if prec == "fp64":
        # the TVD limiter turns differences of round-off size in the
        # transports into somewhat larger local tracer differences
        _check_tracers(b, gm, om, rel=1e-5)
