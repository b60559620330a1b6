"""Deferred checks through the C ABI: steps that do not wait for the device
at their end (shyhip_set_deferred_checks) run the same kernels in the same
order, so every field is bit-identical to steps that wait.  In this mode a
step's tracer transport runs on streams of its own while the next step's
hydro runs (the arrays it reads rotate over three buffers): the cases
below cover one and two tracers, switching the mode during a run, pulls
between steps, and the ensemble path (pinned boundary-level staging)."""
import numpy as np
import pytest

from shyhip import ensemble, mesh

pytestmark = pytest.mark.gpu

PARAMS = dict(dt=60.0, ireib=5, itvd=2, itvdv=1, aapar=0.0, block_rows=256, tri_sweeps=3)


def _run(gpu, b, tr, prec, deferred, nsteps=20):
    m = gpu.Model(b, prec, params=PARAMS, visv=1e-3, difv=1e-4, tracers=tr)
    m.set_deferred_checks(deferred)
    m.set_wind(*mesh.wind_stress(10.0, 0.7))
    t = 0.0
    for _ in range(nsteps):
        t += PARAMS["dt"]
        m.step(mesh.tide_level(t, amps=(0.4, 0.1), periods=(mesh.M2_PERIOD, mesh.K1_PERIOD), phases=(0.0, 0.7)))
    m.sync()
    out = {k: m.pull(k) for k in ("znv", "utlnv", "wlnv", "iwegv", "salt", "temp")}
    out["stats"] = m.stats()
    m.close()
    return out


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_deferred_steps_bit_identical(gpu, prec):
    b = mesh.lagoon(3000, seed=21, lx=8000.0, ly=4000.0, nlv=6, tile=16)
    tr = dict(temp=mesh.smooth_tracer(b, 22, 15.0, 25.0), salt=mesh.smooth_tracer(b, 23, 25.0, 38.0))
    ref = _run(gpu, b, tr, prec, False)
    got = _run(gpu, b, tr, prec, True)
    print(prec, "stats", ref["stats"], got["stats"])
    for k in ("znv", "utlnv", "wlnv", "iwegv", "salt", "temp"):
        assert np.array_equal(ref[k], got[k]), k


def test_deferred_ensemble_levels(gpu):
    mems = []
    for n, s, phase in ((1200, 31, 0.0), (2000, 32, 2.0)):
        b = mesh.lagoon(n, seed=s, lx=6000.0, ly=3000.0, nlv=4, tile=16)
        mems.append(dict(basin=b, wind=mesh.wind_stress(6.0, phase),
                         level=(lambda t, ph=phase: mesh.tide_level(t, amps=(0.4, 0.1),
                                                                   periods=(mesh.M2_PERIOD, mesh.K1_PERIOD),
                                                                   phases=(ph, 0.7)) - 0.2),
                         tracers=dict(salt=mesh.smooth_tracer(b, s + 1, 25.0, 38.0).astype(np.float32))))
    res = []
    for deferred in (False, True):
        ens = ensemble.Ensemble(mems, "fp32", params=PARAMS, visv=1e-3, difv=1e-4)
        ens.model.set_deferred_checks(deferred)
        for _ in range(15):
            ens.step()
        ens.model.sync()
        res.append([ens.pull("znv", i) for i in range(len(mems))])
        ens.close()
    for a, c in zip(*res):
        assert np.array_equal(a, c)


def _lag():
    b = mesh.lagoon(2500, seed=41, lx=7000.0, ly=3500.0, nlv=5, tile=16)
    return b, dict(salt=mesh.smooth_tracer(b, 42, 25.0, 38.0))


def _lev(t):
    return mesh.tide_level(t, amps=(0.4, 0.1), periods=(mesh.M2_PERIOD, mesh.K1_PERIOD), phases=(0.3, 0.7))


def _steps(m, t0, n, pulls=False):
    t, seen = t0, []
    for _ in range(n):
        t += PARAMS["dt"]
        m.step(_lev(t))
        if pulls:
            seen.append(m.pull("salt"))       # waits for the transport still running
    return t, seen


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_pipelined_one_tracer_modes_and_pulls(gpu, prec):
    b, tr = _lag()
    ref = gpu.Model(b, prec, params=PARAMS, visv=1e-3, difv=1e-4, tracers=tr)
    ref.set_wind(*mesh.wind_stress(12.0, 1.1))
    _, ref_seen = _steps(ref, 0.0, 18, pulls=True)
    ref_end = {k: ref.pull(k) for k in ("znv", "utlnv", "salt")}
    ref.close()
    m = gpu.Model(b, prec, params=PARAMS, visv=1e-3, difv=1e-4, tracers=tr)
    m.set_wind(*mesh.wind_stress(12.0, 1.1))
    m.set_deferred_checks(True)
    t, seen = _steps(m, 0.0, 6, pulls=True)          # pipelined, drained by every pull
    t, _ = _steps(m, t, 5)                            # pipelined, nothing waits
    m.set_deferred_checks(False)                      # drains; buffers back to two
    t, _ = _steps(m, t, 3)
    m.set_deferred_checks(True)
    t, _ = _steps(m, t, 4)
    m.sync()
    for a, c in zip(seen, ref_seen[:6]):
        assert np.array_equal(a, c)
    for k, v in ref_end.items():
        assert np.array_equal(m.pull(k), v), k
    m.close()
