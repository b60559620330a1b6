"""Continuous batching (serve.BatchServer) through the C ABI: the resident
runs are the members of one union model, stepped together; runs join and
leave between steps, and each change moves the continuing runs' state into
the next union (shyhip_copy_member_state).  The copy itself is exact; a
run's fields agree with its run alone to the solver tolerance (the union's
convergence test is on the union residual)."""
import numpy as np
import pytest

from shyhip import ensemble, mesh, serve

pytestmark = pytest.mark.gpu

PARAMS = dict(dt=60.0, ireib=5, block_rows=256, tri_sweeps=3)
STATE = ("znv", "zenv", "utlnv", "vtlnv", "utlov", "vtlov", "wlnv", "hdknv", "hdenv", "iwegv", "tauxnv", "salt")


def test_copy_member_state_is_exact(gpu):
    bs = [mesh.lagoon(n, seed=s, lx=6000.0, ly=3000.0, nlv=4, tile=16) for n, s in ((1500, 41), (2200, 42), (900, 43))]
    tr = lambda b, s: dict(salt=mesh.smooth_tracer(b, s, 25.0, 38.0))
    ua = ensemble.union_basin([bs[0], bs[1]])
    trs = [tr(b, 50 + i) for i, b in enumerate(bs)]
    a = gpu.Model(ua.basin, "fp32", params=PARAMS, tracers={"salt": np.concatenate([trs[0]["salt"], trs[1]["salt"]])},
                  members=(ua.node_start, ua.elem_start))
    for i, w in enumerate((5.0, 12.0)):
        a.set_member_wind(i, *mesh.wind_stress(w, 0.3 + i))
    for n in range(6):
        a.step(mesh.tide_level((n + 1) * 60.0) + 0.2, istot=0)
    a.sync()
    ub = ensemble.union_basin([bs[2], bs[1], bs[0]])
    b = gpu.Model(ub.basin, "fp32", params=PARAMS,
                  tracers={"salt": np.concatenate([trs[2]["salt"], trs[1]["salt"], trs[0]["salt"]])},
                  members=(ub.node_start, ub.elem_start))
    b.copy_member_state(2, a, 0)
    b.copy_member_state(1, a, 1)
    for name in STATE:
        fa, fb = a.pull(name), b.pull(name)
        for ma, mb in ((0, 2), (1, 1)):
            if fa.shape[0] == ua.basin.nkn:
                sa = fa[ua.node_start[ma]:ua.node_start[ma + 1]]
                sb = fb[ub.node_start[mb]:ub.node_start[mb + 1]]
            else:
                sa = fa[ua.elem_start[ma]:ua.elem_start[ma + 1]]
                sb = fb[ub.elem_start[mb]:ub.elem_start[mb + 1]]
            assert np.array_equal(sa, sb), (name, ma)
    # the member that was not copied keeps its own initial state
    c = gpu.Model(bs[2], "fp32", params=PARAMS, tracers=trs[2])
    assert np.array_equal(b.pull("salt")[:bs[2].nkn], c.pull("salt"))
    for m in (a, b, c):
        m.close()


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_batched_serving_matches_runs_alone(gpu, prec):
    reqs = serve.poisson_stream(6, rate=12.0, seed=20261203, sizes=(2500, 9000), steps=(40, 90), nlv=8)
    alone = [serve.run_request(r, prec, PARAMS, {})["znv"] for r in reqs]
    srv = serve.BatchServer(prec, params=PARAMS, slots=3, hbm_budget=8 << 30)
    done, wall = srv.serve(reqs)
    s = serve.summary(done, wall)
    print(prec, s, "rebuilds", srv.rebuilds, "batch steps", srv.batch_steps, "run steps", srv.run_steps)
    assert s["failed"] == 0, [r.error for r in done if r.error]
    assert srv.run_steps == sum(r.nsteps for r in reqs)
    assert srv.batch_steps < srv.run_steps          # steps of different runs shared launches
    assert srv.rebuilds >= 3                        # runs joined and left while others ran
    tol = 5e-5 if prec == "fp32" else 1e-6
    for r, z in zip(done, alone):
        err = float(np.max(np.abs(r.result["znv"] - z)))
        print(f"request {r.rid}: {r.basin.nkn} nodes, {r.nsteps} steps, |dz| max {err:.3e} m")
        assert err < tol, f"request {r.rid} differs from its run alone by {err} m"
        assert r.t_done >= r.t_admit >= r.arrival - 1e-6


def test_batched_admission_respects_budget(gpu):
    reqs = serve.poisson_stream(3, rate=100.0, seed=7, sizes=(2000, 3000), steps=(5, 10), nlv=4)
    need = max(serve.footprint_bytes(r.basin, "fp32") for r in reqs)
    srv = serve.BatchServer("fp32", params=PARAMS, slots=3, hbm_budget=int(1.5 * need))
    done, wall = srv.serve(reqs)
    assert all(r.error is None for r in done)
    spans = sorted((r.t_admit, r.t_done) for r in done)
    for (a0, d0), (a1, d1) in zip(spans, spans[1:]):
        assert a1 >= d0 - 1e-3, "two runs resident over the budget"
