"""GPU parity of a tracer's own terms (conz3d_omp's robs/cobs/rtauv nudging,
wsink/wsinkv sinking, rload/load loads; concentration_omp.f90:479,697-735,
816-834) against the oracle (pinned in test_oracle_tracer_terms.py).  The
tracer with terms is transported alone with the substeps its own
scalar_stability chooses (nudging and sinking count there), the other
tracer as before.

Bars as in test_gpu_parity: masks bit-exact, zeta 1e-6 m, tracers 1e-6 of
the field scale (its initial or final maximum, FP64); the loads deposition leaves within 1e-6 of their
scale; each term changes the tracer by far more than the bar.
"""
import numpy as np
import pytest

from shyhip import mesh

pytestmark = pytest.mark.gpu

SCALE = dict(temp=25.0, salt=38.0)


def _case(n=3000, nlv=6, seed=7):
    b = mesh.lagoon(n, seed=seed, lx=8000.0, ly=4000.0, nlv=nlv, tile=16)
    tr = dict(temp=mesh.smooth_tracer(b, seed + 1, 15.0, 25.0), salt=mesh.smooth_tracer(b, seed + 2, 25.0, 38.0))
    rng = np.random.default_rng(seed + 9)
    terms = dict(
        nudge=dict(robs=1.0, cobs=mesh.smooth_tracer(b, seed + 3, 30.0, 36.0),
                   rtauv=rng.uniform(1 / 7200.0, 1 / 1800.0, (b.nkn, b.nlv))),
        sink=dict(wsink=2e-4, wsinkv=rng.uniform(0.5, 1.5, (b.nkn, b.nlv + 1))),
        load=dict(rload=1.0, load=np.where(rng.uniform(size=(b.nkn, b.nlv)) < 0.5, -1.0, 1.0)
                  * rng.uniform(1.0, 10.0, (b.nkn, b.nlv))))
    terms["all"] = dict(**terms["nudge"], **terms["sink"], **terms["load"])
    return b, tr, terms


def _level(t):
    return mesh.tide_level(t, amps=(0.4, 0.1), periods=(mesh.M2_PERIOD, mesh.K1_PERIOD)) - 0.3


def _dc(b, g, o):
    valid = np.arange(b.nlv)[None, :] < b.ilhkv[:, None]
    return float(np.max(np.abs(g - o)[valid]))


@pytest.mark.parametrize("which", ["nudge", "sink", "load", "all"])
def test_tracer_terms_fp64(gpu, which):
    import pyoracle
    b, tr, terms = _case()
    params = dict(dt=60.0, ireib=5, itpre=1, itvd=2, itvdv=1, aapar=0.0, block_rows=0)
    om = pyoracle.OracleModel(b, "f64", params=params, visv=1e-3, difv=1e-4, tracers=tr)
    gm = gpu.Model(b, "fp64", params=params, visv=1e-3, difv=1e-4, tracers=tr)
    g0 = gpu.Model(b, "fp64", params=params, visv=1e-3, difv=1e-4, tracers=tr)
    om.set_tracer_terms("salt", **terms[which])
    gm.set_tracer_terms("salt", **terms[which])
    tx, ty = mesh.wind_stress(12.0, 0.7)
    for m in (om, gm, g0):
        m.set_wind(tx, ty)
    for n in range(8):
        for m in (om, gm, g0):
            m.step(_level((n + 1) * 60.0), istot=0)
    assert np.array_equal(gm.pull("iwegv"), om.a["iwegv"])
    assert float(np.max(np.abs(gm.pull("znv") - om.a["znv"]))) < 1e-6
    for name in ("temp", "salt"):
        d = _dc(b, gm.pull(name), om.tracers[name])
        effect = _dc(b, gm.pull(name), g0.pull(name))
        scale = max(SCALE[name], float(np.abs(om.tracers[name]).max()))
        print(f"{which} {name}: device vs oracle {d}, effect of the terms {effect}, field scale {scale}")
        assert d < 1e-6 * scale
        if name == "salt":
            assert effect > 100 * max(d, 1e-12)
        else:
            assert effect < 1e-6 * SCALE[name]          # the other tracer is untouched
    if "load" in terms[which]:
        lg, lo = gm.tracer_load("salt"), om.tracer_terms["salt"]["load"]
        dl = _dc(b, lg, lo)
        print("load after deposition: max |d|", dl)
        assert dl < 1e-6 * float(np.abs(lo).max())


def test_tracer_terms_fp32(gpu):
    """FP32 path (all three terms) against the reference-precision oracle:
    within a tenth of what the terms change."""
    import pyoracle
    b, tr, terms = _case()
    params = dict(dt=60.0, ireib=5, itpre=1, itvd=2, itvdv=1, aapar=0.0, block_rows=0)
    om = pyoracle.OracleModel(b, "f32", params=params, visv=1e-3, difv=1e-4, tracers=tr)
    o0 = pyoracle.OracleModel(b, "f32", params=params, visv=1e-3, difv=1e-4, tracers=tr)
    gm = gpu.Model(b, "fp32", params=params, visv=1e-3, difv=1e-4, tracers=tr)
    om.set_tracer_terms("salt", **terms["all"])
    gm.set_tracer_terms("salt", **terms["all"])
    tx, ty = mesh.wind_stress(12.0, 0.7)
    for m in (om, o0, gm):
        m.set_wind(tx, ty)
    for n in range(8):
        for m in (om, o0, gm):
            m.step(_level((n + 1) * 60.0), istot=0)
    assert np.array_equal(gm.pull("iwegv"), om.a["iwegv"])
    assert float(np.max(np.abs(gm.pull("znv") - om.a["znv"]))) < 2e-5
    d = _dc(b, gm.pull("salt"), om.tracers["salt"])
    effect = _dc(b, om.tracers["salt"], o0.tracers["salt"])
    print("fp32 salt: device vs oracle", d, "effect of the terms", effect)
    assert d < 0.1 * effect


def test_ensemble_with_diffusion_terms_and_gradient_tvd(gpu):
    """Members batched with horizontal diffusion, gradient TVD and one
    tracer's terms match their runs alone (each member and tracer with its
    own substeps)."""
    from shyhip import ensemble
    params = dict(dt=60.0, ireib=5, itvd=1, itvdv=1, aapar=0.0, block_rows=256, tri_sweeps=3, rkpar=0.3)
    mems = []
    for n, s, lx in ((1500, 31, 6000.0), (2600, 32, 8000.0)):
        mb = mesh.lagoon(n, seed=s, lx=lx, ly=lx / 2, nlv=4, tile=16)
        mems.append(dict(basin=mb, wind=mesh.wind_stress(10.0, 0.5 + s),
                         level=(lambda t, ph=float(s): mesh.tide_level(t, amps=(0.4,), phases=(ph,)) - 0.2),
                         tracers=dict(temp=mesh.smooth_tracer(mb, s + 1, 15.0, 25.0),
                                      salt=mesh.smooth_tracer(mb, s + 2, 25.0, 38.0))))
    ens = ensemble.Ensemble(mems, "fp64", params=params, visv=1e-3, difv=1e-4)
    nodes = sum(m["basin"].nkn for m in mems)
    L = mems[0]["basin"].nlv
    rng = np.random.default_rng(5)
    ens.model.push("difhv", np.full((sum(m["basin"].nel for m in mems), L), 5.0))
    cobs = rng.uniform(30.0, 34.0, (nodes, L))
    rtau = np.full((nodes, L), 1 / 3600.0)
    ens.model.set_tracer_terms("salt", robs=1.0, cobs=cobs, rtauv=rtau)
    for _ in range(6):
        ens.step(istot=0)
    k0 = 0
    for i, mem in enumerate(mems):
        b = mem["basin"]
        m = gpu.Model(b, "fp64", params=params, visv=1e-3, difv=1e-4, tracers=mem["tracers"])
        m.push("difhv", np.full((b.nel, L), 5.0))
        m.set_tracer_terms("salt", robs=1.0, cobs=cobs[k0:k0 + b.nkn], rtauv=rtau[k0:k0 + b.nkn])
        m.set_wind(*mem["wind"])
        for n in range(6):
            m.step(mem["level"]((n + 1) * 60.0), istot=0)
        dz = float(np.max(np.abs(ens.pull("znv", i) - m.pull("znv"))))
        for name in ("temp", "salt"):
            ds = _dc(b, ens.pull(name, i), m.pull(name))
            print("member", i, name, "|dz|", dz, "|dc|", ds)
            assert ds < 1e-5 * SCALE[name]
        assert dz < 1e-6
        k0 += b.nkn
        m.close()
    ens.close()
