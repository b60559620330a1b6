"""HIP graphs through the C ABI: a tracer group's transport substeps and the
zeta solver's iteration chunks are captured once per argument set and
replayed (shyhip_set_graphs).  The replayed launches are the same kernels
with the same arguments, so every field is bit-identical to a run that
launches each kernel directly.  The cases cover both precisions, substeps
chosen every step by scalar_stability (so the graphs differ in length from
step to step), the pipelined transport whose input arrays rotate over three
buffers (the graphs differ in their pointers), pulls between steps (eager
work between replays), and an ensemble whose members run different numbers
of substeps."""
import numpy as np
import pytest

from shyhip import ensemble, mesh

pytestmark = pytest.mark.gpu

PARAMS = dict(dt=60.0, ireib=5, itvd=2, itvdv=1, aapar=0.0, block_rows=256, tri_sweeps=3)
FIELDS = ("znv", "utlnv", "wlnv", "iwegv", "salt", "temp")


def _level(t):
    return mesh.tide_level(t, amps=(0.4, 0.1), periods=(mesh.M2_PERIOD, mesh.K1_PERIOD), phases=(0.0, 0.7))


def _run(gpu, b, tr, prec, graphs, deferred, nsteps=24):
    m = gpu.Model(b, prec, params=PARAMS, visv=1e-3, difv=1e-4, tracers=tr)
    m.set_graphs(graphs)
    m.set_deferred_checks(deferred)
    m.set_wind(*mesh.wind_stress(12.0, 0.7))
    t, snaps, istots = 0.0, [], []
    for n in range(nsteps):
        t += PARAMS["dt"]
        m.step(_level(t), istot=0)
        if n % 6 == 5:                     # eager work between replays
            m.sync()
            snaps.append({k: m.pull(k) for k in FIELDS})
            istots.append(m.last_istot())
    m.sync()
    out = dict(snaps=snaps, istots=istots, graphs=m.graph_stats(), stats=m.stats())
    m.close()
    return out


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("deferred", [False, True])
def test_graph_replays_bit_identical(gpu, prec, deferred):
    b = mesh.lagoon(3000, seed=21, lx=8000.0, ly=4000.0, nlv=6, tile=16)
    tr = dict(temp=mesh.smooth_tracer(b, 22, 15.0, 25.0), salt=mesh.smooth_tracer(b, 23, 25.0, 38.0))
    ref = _run(gpu, b, tr, prec, False, deferred)
    got = _run(gpu, b, tr, prec, True, deferred)
    print(prec, deferred, "istot", got["istots"], "graphs (built, replayed)", got["graphs"], "stats", got["stats"])
    assert ref["graphs"] == (0, 0)
    # (a short run: the pipelined transport's rotating buffers and the
    # varying substep and iteration counts make most argument sets new)
    builds, replays = got["graphs"]
    assert builds > 0 and replays > 0, got["graphs"]
    assert ref["istots"] == got["istots"]
    for s_ref, s_got in zip(ref["snaps"], got["snaps"]):
        for k in FIELDS:
            assert np.array_equal(s_ref[k], s_got[k]), k


def test_graph_ensemble_members_own_substeps(gpu):
    mems = []
    for n, s, ws in ((1200, 31, 2.0), (2400, 32, 14.0)):
        b = mesh.lagoon(n, seed=s, lx=6000.0, ly=3000.0, nlv=4, tile=16)
        mems.append(dict(basin=b, wind=mesh.wind_stress(ws, 0.4),
                         level=(lambda t: mesh.tide_level(t, amps=(0.4, 0.1),
                                                          periods=(mesh.M2_PERIOD, mesh.K1_PERIOD),
                                                          phases=(0.0, 0.7)) - 0.2),
                         tracers=dict(salt=mesh.smooth_tracer(b, s + 1, 25.0, 38.0).astype(np.float32))))
    res = []
    for graphs in (False, True):
        ens = ensemble.Ensemble(mems, "fp32", params=PARAMS, visv=1e-3, difv=1e-4)
        ens.model.set_graphs(graphs)
        for _ in range(15):
            ens.step()
        ens.model.sync()
        res.append(([ens.pull("znv", i) for i in range(len(mems))], [ens.pull("salt", i) for i in range(len(mems))],
                    ens.model.graph_stats()))
    print("graphs (built, replayed)", res[1][2])
    assert res[1][2][1] > 0
    for a, b_ in zip(res[0][0] + res[0][1], res[1][0] + res[1][1]):
        assert np.array_equal(a, b_)


def test_graphs_with_models_in_threads(gpu):
    """Models stepped from several threads at once, each capturing and
    replaying its own graphs while the others allocate, free and copy: the
    results equal the same runs one after another, bit for bit.  (The
    models' streams are non-blocking and a capture holds the library's
    device lock, so another thread's allocation or copy never lands inside a
    capture.)"""
    import threading
    basins = [mesh.lagoon(n, seed=60 + i, lx=6000.0, ly=3000.0, nlv=5, tile=16) for i, n in enumerate((1500, 2500, 3500))]

    def run(b, out, i):
        tr = dict(temp=mesh.smooth_tracer(b, 70 + i, 15.0, 25.0), salt=mesh.smooth_tracer(b, 80 + i, 25.0, 38.0))
        for rep in range(2):               # a model created and closed while others capture
            m = gpu.Model(b, "fp32", params=PARAMS, visv=1e-3, difv=1e-4, tracers=tr)
            m.set_wind(*mesh.wind_stress(8.0 + 3 * i, 0.5))
            t = 0.0
            for _ in range(12):
                t += PARAMS["dt"]
                m.step(_level(t), istot=0)
            m.sync()
            out[i] = {k: m.pull(k) for k in FIELDS}
            m.close()

    seq = [None] * 3
    for i, b in enumerate(basins):
        run(b, seq, i)
    par = [None] * 3
    ths = [threading.Thread(target=run, args=(b, par, i)) for i, b in enumerate(basins)]
    for th in ths:
        th.start()
    for th in ths:
        th.join()
    for i in range(3):
        for k in FIELDS:
            assert np.array_equal(seq[i][k], par[i][k]), (i, k)
