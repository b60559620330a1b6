"""GPU parity of the gradient TVD scheme (itvd = 1): the node gradients of
tvd_grad_3d (tvd_admin.f90:478-543), computed on the device before every
substep, and the gradient branch of tvd_fluxes (tvd_admin.f90:787-790),
against the oracle run with itvd = 1 (pinned in test_oracle_tvd_grad.py).

Bars as in test_gpu_parity: masks bit-exact; zeta 1e-6 m (FP64 vs the FP64
oracle), 2e-5 m (FP32 vs the reference-precision oracle); tracers 1e-6 of
the field scale (FP64).  FP32 tracers: within a tenth of what separates the
itvd = 1 run from the itvd = 2 run of the same oracle, so the bar tells the
two schemes apart.  Ensemble members (several substep counts) match their
runs alone to the zeta-solver tolerance, as for itvd = 2.
"""
import numpy as np
import pytest

from shyhip import ensemble, mesh

pytestmark = pytest.mark.gpu

SCALE = dict(temp=25.0, salt=38.0)            # the fields' initial maxima


def _case(itvd=1, dt=60.0, n=3000, nlv=6, seed=7):
    b = mesh.lagoon(n, seed=seed, lx=8000.0, ly=4000.0, nlv=nlv, tile=16)
    tr = dict(temp=mesh.smooth_tracer(b, seed + 1, 15.0, 25.0), salt=mesh.smooth_tracer(b, seed + 2, 25.0, 38.0))
    params = dict(dt=dt, ireib=5, itpre=1, itvd=itvd, itvdv=1, aapar=0.0, block_rows=0)
    return b, tr, params


def _level(t):
    return mesh.tide_level(t, amps=(0.4, 0.1), periods=(mesh.M2_PERIOD, mesh.K1_PERIOD)) - 0.3


def _oracle_run(b, tr, params, prec, nsteps, istot):
    import pyoracle
    om = pyoracle.OracleModel(b, prec, params=params, visv=1e-3, difv=1e-4, tracers=tr)
    om.set_wind(*mesh.wind_stress(12.0, 0.7))
    for n in range(nsteps):
        om.step(_level((n + 1) * params["dt"]), istot=istot)
    return om


def _device_run(gpu, b, tr, params, prec, nsteps, istot):
    gm = gpu.Model(b, prec, params=params, visv=1e-3, difv=1e-4, tracers=tr)
    gm.set_wind(*mesh.wind_stress(12.0, 0.7))
    ran = []
    for n in range(nsteps):
        gm.step(_level((n + 1) * params["dt"]), istot=istot)
        ran.append(gm.last_istot())
    return gm, ran


def _tracer_diff(b, g, o):
    valid = np.arange(b.nlv)[None, :] < b.ilhkv[:, None]
    return float(np.max(np.abs(g - o)[valid]))


@pytest.mark.parametrize("istot", [1, 0])
def test_gradient_tvd_fp64(gpu, istot):
    """10 steps, one substep each or as scalar_stability chooses."""
    b, tr, params = _case()
    om = _oracle_run(b, tr, params, "f64", 10, istot)
    gm, ran = _device_run(gpu, b, tr, params, "fp64", 10, istot)
    print("substeps run by the device:", ran)
    assert np.array_equal(gm.pull("iwegv"), om.a["iwegv"])
    assert float(np.max(np.abs(gm.pull("znv") - om.a["znv"]))) < 1e-6
    o2 = _oracle_run(b, tr, dict(params, itvd=2), "f64", 10, istot)
    for name in ("temp", "salt"):
        d = _tracer_diff(b, gm.pull(name), om.tracers[name])
        scheme = _tracer_diff(b, o2.tracers[name], om.tracers[name])
        print(name, "device vs oracle", d, "itvd 1 vs 2 (oracle)", scheme)
        assert d < 1e-6 * SCALE[name]
        assert scheme > 100 * d                 # the bar separates the two schemes
    if istot == 0:
        assert max(ran) >= 2
    gm.close()


def test_gradient_tvd_fp32(gpu):
    b, tr, params = _case()
    om = _oracle_run(b, tr, params, "f32", 10, 1)
    gm, _ = _device_run(gpu, b, tr, params, "fp32", 10, 1)
    assert np.array_equal(gm.pull("iwegv"), om.a["iwegv"])
    assert float(np.max(np.abs(gm.pull("znv") - om.a["znv"]))) < 2e-5
    o2 = _oracle_run(b, tr, dict(params, itvd=2), "f32", 10, 1)
    for name in ("temp", "salt"):
        d = _tracer_diff(b, gm.pull(name), om.tracers[name])
        scheme = _tracer_diff(b, o2.tracers[name], om.tracers[name])
        print(name, "device vs oracle", d, "itvd 1 vs 2 (oracle)", scheme)
        assert d < 0.1 * scheme
    gm.close()


def test_gradient_tvd_ensemble_members_match_alone(gpu):
    """Members of different sizes and winds, each choosing its own substep
    count, batched in one model: each matches its run alone."""
    params = dict(dt=300.0, ireib=5, itvd=1, itvdv=1, aapar=0.0, block_rows=256, tri_sweeps=3)
    mems = []
    for n, s, lx, speed, phase in ((1500, 11, 6000.0, 5.0, 0.0), (2600, 12, 8000.0, 16.0, 1.3)):
        mb = mesh.lagoon(n, seed=s, lx=lx, ly=lx / 2, nlv=4, tile=16)
        salt = mesh.smooth_tracer(mb, s + 100, 25.0, 38.0).astype(np.float32)
        mems.append(dict(basin=mb, wind=mesh.wind_stress(speed, 0.5 + phase),
                         level=(lambda t, ph=phase: mesh.tide_level(t, amps=(0.4, 0.1),
                                                                   periods=(mesh.M2_PERIOD, mesh.K1_PERIOD),
                                                                   phases=(ph, 0.7)) - 0.2),
                         tracers=dict(salt=salt)))
    nsteps = 8
    ens = ensemble.Ensemble(mems, "fp64", params=params, visv=1e-3, difv=1e-4)
    counts = []
    for _ in range(nsteps):
        ens.step(istot=0)
        counts.append(ens.model.member_istot(len(mems)))
    print("member substeps per step:", counts)
    for i, mem in enumerate(mems):
        m = gpu.Model(mem["basin"], "fp64", params=params, visv=1e-3, difv=1e-4, tracers=mem["tracers"])
        m.set_wind(*mem["wind"])
        for n in range(nsteps):
            m.step(mem["level"]((n + 1) * params["dt"]), istot=0)
        b = mem["basin"]
        dz = float(np.max(np.abs(ens.pull("znv", i) - m.pull("znv"))))
        ds = _tracer_diff(b, ens.pull("salt", i), m.pull("salt"))
        print("member", i, "|dz| =", dz, "|dsalt| =", ds)
        assert np.array_equal(ens.pull("iwegv", i), m.pull("iwegv"))
        assert dz < 1e-6
        assert ds < 1e-5 * 38.0
        m.close()
    ens.close()
