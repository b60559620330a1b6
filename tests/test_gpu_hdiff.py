"""GPU parity of the horizontal diffusion of scalars (rkpar > 0): the element
term 3 rkpar difhv h wdiff of conz3d_element (concentration_omp.f90:560-570,
727-735) with the weights of diffweight (diffusion_admin.f90:182-340), and
its share in scalar_stability (conzstab, concentration.f90:1890-1985),
against the oracle (pinned in test_oracle_hdiff.py).

Bars as in test_gpu_parity: masks bit-exact, zeta 1e-6 m, tracers 1e-6 of
the field scale (FP64); the stability index's istot equal and rindex within
1e-6 relative; FP32 tracers within a tenth of what the diffusion changes.
The diffusion changes the tracers by far more than the bar.
"""
import numpy as np
import pytest

from shyhip import mesh

pytestmark = pytest.mark.gpu

SCALE = dict(temp=25.0, salt=38.0)


def _case(n=3000, nlv=6, seed=7):
    b = mesh.lagoon(n, seed=seed, lx=8000.0, ly=4000.0, nlv=nlv, tile=16)
    tr = dict(temp=mesh.smooth_tracer(b, seed + 1, 15.0, 25.0), salt=mesh.smooth_tracer(b, seed + 2, 25.0, 38.0))
    rng = np.random.default_rng(seed + 5)
    difhv = rng.uniform(2.0, 20.0, (b.nel, b.nlv))       # m2/s, varying between elements
    return b, tr, difhv


def _level(t):
    return mesh.tide_level(t, amps=(0.4, 0.1), periods=(mesh.M2_PERIOD, mesh.K1_PERIOD)) - 0.3


def _pair(gpu, b, tr, difhv, prec, oprec, params, idtype):
    import pyoracle
    om = pyoracle.OracleModel(b, oprec, params=params, visv=1e-3, difv=1e-4, tracers=tr, idtype=idtype)
    gm = gpu.Model(b, prec, params=dict(params, idtype=idtype), visv=1e-3, difv=1e-4, tracers=tr)
    om.a["difhv"][:] = difhv
    gm.push("difhv", difhv)
    tx, ty = mesh.wind_stress(12.0, 0.7)
    om.set_wind(tx, ty)
    gm.set_wind(tx, ty)
    return om, gm


def _dc(b, g, o):
    valid = np.arange(b.nlv)[None, :] < b.ilhkv[:, None]
    return float(np.max(np.abs(g - o)[valid]))


# substeps as scalar_stability chooses them (the diffusion needs several:
# one substep would be unstable), or a given count above that
@pytest.mark.parametrize("idtype,istot", [(2, 0), (2, 12), (0, 0), (1, 0)])
def test_horizontal_diffusion_fp64(gpu, idtype, istot):
    b, tr, difhv = _case()
    params = dict(dt=60.0, ireib=5, itpre=1, itvd=2, itvdv=1, aapar=0.0, block_rows=0, rkpar=0.5)
    om, gm = _pair(gpu, b, tr, difhv, "fp64", "f64", params, idtype)
    for n in range(8):
        om.step(_level((n + 1) * 60.0), istot=istot)
        gm.step(_level((n + 1) * 60.0), istot=istot)
        io, ro = om.scalar_stability()
        ig, rg = gm.scalar_stability()
        assert ig == io and abs(rg - ro) <= 1e-6 * max(1.0, abs(ro)), (n, io, ro, ig, rg)
    assert np.array_equal(gm.pull("iwegv"), om.a["iwegv"])
    assert float(np.max(np.abs(gm.pull("znv") - om.a["znv"]))) < 1e-6
    g0 = gpu.Model(b, "fp64", params=dict(params, rkpar=0.0), visv=1e-3, difv=1e-4, tracers=tr)
    g0.set_wind(*mesh.wind_stress(12.0, 0.7))
    for n in range(8):
        g0.step(_level((n + 1) * 60.0), istot=istot)
    for name in ("temp", "salt"):
        d = _dc(b, gm.pull(name), om.tracers[name])
        effect = _dc(b, gm.pull(name), g0.pull(name))
        print(f"idtype {idtype} istot {istot} {name}: device vs oracle {d}, effect of the diffusion {effect}, "
              f"last istot {gm.last_istot()}")
        assert d < 1e-6 * SCALE[name]
        assert effect > 100 * d


def test_horizontal_diffusion_fp32(gpu):
    b, tr, difhv = _case()
    params = dict(dt=60.0, ireib=5, itpre=1, itvd=2, itvdv=1, aapar=0.0, block_rows=0, rkpar=0.5)
    om, gm = _pair(gpu, b, tr, difhv, "fp32", "f32", params, 2)
    import pyoracle
    o0 = pyoracle.OracleModel(b, "f32", params=dict(params, rkpar=0.0), visv=1e-3, difv=1e-4, tracers=tr)
    o0.set_wind(*mesh.wind_stress(12.0, 0.7))
    for n in range(8):
        for m in (om, gm, o0):
            m.step(_level((n + 1) * 60.0), istot=12)
    assert np.array_equal(gm.pull("iwegv"), om.a["iwegv"])
    assert float(np.max(np.abs(gm.pull("znv") - om.a["znv"]))) < 2e-5
    for name in ("temp", "salt"):
        d = _dc(b, gm.pull(name), om.tracers[name])
        effect = _dc(b, om.tracers[name], o0.tracers[name])
        print(f"fp32 {name}: device vs oracle {d}, effect of the diffusion {effect}")
        assert d < 0.1 * effect
