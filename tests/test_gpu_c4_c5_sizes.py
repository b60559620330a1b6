"""The batched and serving workloads at their own sizes.

* Ensemble (64 members of 5k-120k nodes, log-uniform, 10 z-layers, each
  with its own wind 0-25 m/s, direction and tidal phase; FP32 + FP64
  refinement; all members in one model): every step's launches cover the
  whole union.  Checks: a constant tracer stays constant in every member
  (the transport's consistency with continuity), a varying tracer stays
  inside its initial bounds (the TVD property with boundary values inside
  them), and three members match the reference-precision oracle run alone
  (zeta 2e-5 m, wet/dry masks equal).
* Serving: a Poisson stream whose requests include a 300k-node basin; every
  result is the same alone and in the mix.
"""
import math

import numpy as np
import pytest

from shyhip import ensemble, mesh, serve

pytestmark = pytest.mark.gpu

PARAMS = dict(dt=60.0, ireib=5, itvd=2, itvdv=1, aapar=0.0, block_rows=256, tri_sweeps=3)


def _members(nmem=64, seed=20261202):
    rng = np.random.default_rng(np.uint64(seed))
    out = []
    for i in range(nmem):
        n = int(math.exp(rng.uniform(math.log(5000), math.log(120000))))
        lx = 40000.0 * math.sqrt(n / 50000.0)
        b = mesh.lagoon(n, seed=seed + 1000 + i, lx=lx, ly=lx / 2, nlv=10, tile=32)
        ph = float(rng.uniform(0, 2 * math.pi))
        wind = mesh.wind_stress(float(rng.uniform(0, 25)), float(rng.uniform(0, 2 * math.pi)))
        level = (lambda t, ph=ph: mesh.tide_level(t, amps=(0.4, 0.1), periods=(mesh.M2_PERIOD, mesh.K1_PERIOD),
                                                  phases=(ph, 0.7)))
        tr = dict(salt=mesh.smooth_tracer(b, seed + 3000 + i, 25.0, 38.0).astype(np.float32),
                  const=np.full((b.nkn, b.nlv), 20.0, np.float32))
        out.append(dict(basin=b, wind=wind, level=level, tracers=tr))
    return out


def test_ensemble_64_members(gpu):
    import pyoracle
    mems = _members()
    nkn = sum(m["basin"].nkn for m in mems)
    print("members", len(mems), "nodes", nkn, "sizes", min(m["basin"].nkn for m in mems),
          max(m["basin"].nkn for m in mems), flush=True)
    assert len(mems) == 64 and nkn > 1_000_000
    ens = ensemble.Ensemble(mems, "fp32", params=PARAMS, visv=1e-3, difv=1e-4)
    nsteps = 10
    for n in range(nsteps):
        ens.step(istot=0)           # every member its own substeps, as it runs alone
    print("solver", ens.model.stats(), flush=True)
    for i, mem in enumerate(mems):
        b = mem["basin"]
        valid = np.arange(b.nlv)[None, :] < b.ilhkv[:, None]
        c = ens.pull("const", i)[valid]
        assert np.abs(c - 20.0).max() < 2e-4 * 20.0, f"member {i}: constant tracer drifted"
        s = ens.pull("salt", i)[valid]
        s0 = mem["tracers"]["salt"][valid]
        assert s.min() >= s0.min() - 1e-4 and s.max() <= s0.max() + 1e-4, f"member {i}: TVD bound"
    # three members against the reference-precision oracle, each alone
    small = sorted(range(len(mems)), key=lambda i: mems[i]["basin"].nkn)[:3]
    for i in small:
        mem = mems[i]
        om = pyoracle.OracleModel(mem["basin"], "f32", params=dict(PARAMS, itpre=1), visv=1e-3, difv=1e-4,
                                  tracers=mem["tracers"])
        om.set_wind(*mem["wind"])
        t = 0.0
        for n in range(nsteps):
            t += PARAMS["dt"]
            om.step(mem["level"](t), istot=0)
        dz = float(np.max(np.abs(ens.pull("znv", i) - om.a["znv"])))
        print("member", i, "nkn", mem["basin"].nkn, "|dz| vs oracle =", dz, flush=True)
        assert np.array_equal(ens.pull("iwegv", i), om.a["iwegv"])
        assert dz < 2e-5
    ens.close()


def test_serving_stream_with_300k_request(gpu):
    reqs = serve.poisson_stream(4, rate=4.0, seed=20261203, sizes=(5000, 20000), steps=(20, 40), nlv=8)
    big = mesh.lagoon(301000, seed=99, lx=40000.0 * math.sqrt(6.0), ly=20000.0 * math.sqrt(6.0), nlv=8, tile=32,
                      median=10.0, dz=3.0)
    reqs.append(serve.Request(rid=len(reqs), basin=big, nsteps=20, arrival=reqs[1].arrival, wind_max=30.0,
                              wind_dir=0.7, surge=0.8))
    params = dict(dt=60.0, ireib=5, block_rows=256, tri_sweeps=3)
    alone = [serve.run_request(r, "fp32", params, {})["znv"] for r in reqs]
    srv = serve.Server("fp32", params=params, slots=3, hbm_budget=32 << 30)
    done, wall = srv.serve(reqs)
    s = serve.summary(done, wall)
    print(s, "largest request nodes", big.nkn)
    assert big.nkn >= 300000
    assert s["failed"] == 0, [r.error for r in done if r.error]
    for r, z in zip(done, alone):
        assert np.array_equal(r.result["znv"], z), f"request {r.rid} differs in the mix"
