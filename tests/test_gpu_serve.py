"""Serving: requests of different mesh sizes admitted while others run,
stepped concurrently on the GPU (one model and stream each); every result is
the same alone and in the mix -- bit for bit, since runs share no device
memory and every kernel is deterministic."""
import numpy as np
import pytest

from shyhip import serve

pytestmark = pytest.mark.gpu

PARAMS = dict(dt=60.0, ireib=5, block_rows=256, tri_sweeps=3)


@pytest.mark.parametrize("prec", ["fp32"])
def test_serving_results_identical_alone_and_mixed(gpu, prec):
    reqs = serve.poisson_stream(5, rate=20.0, seed=20261202, sizes=(2500, 8000), steps=(30, 60), nlv=8)
    alone = [serve.run_request(r, prec, PARAMS, {})["znv"] for r in reqs]
    srv = serve.Server(prec, params=PARAMS, slots=3, hbm_budget=8 << 30)
    done, wall = srv.serve(reqs)
    s = serve.summary(done, wall)
    print(s)
    assert s["failed"] == 0, [r.error for r in done if r.error]
    for r, z in zip(done, alone):
        assert np.array_equal(r.result["znv"], z), f"request {r.rid} differs in the mix"
        assert r.t_done >= r.t_admit >= r.arrival - 1e-6
    assert s["p99_s"] >= s["p50_s"] > 0


def test_admission_respects_footprint_budget(gpu):
    """With a budget below two runs' footprint, runs are admitted one at a time."""
    reqs = serve.poisson_stream(3, rate=100.0, seed=7, sizes=(2000, 3000), steps=(5, 10), nlv=4)
    need = max(serve.footprint_bytes(r.basin, "fp32") for r in reqs)
    srv = serve.Server("fp32", params=PARAMS, slots=3, hbm_budget=int(1.5 * need))
    done, wall = srv.serve(reqs)
    assert all(r.error is None for r in done)
    spans = sorted((r.t_admit, r.t_done) for r in done)
    for (a0, d0), (a1, d1) in zip(spans, spans[1:]):
        assert a1 >= d0 - 1e-3, "two runs resident over the budget"
