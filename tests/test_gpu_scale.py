"""The third workload's scale through the C ABI: a 500k-node, 20-layer
lagoon, closed (no open boundary) under a 10 m/s wind with wetting and
drying, on the FP32 + FP64-refined path and the plain FP64 path.  The oracle
does not run at this size; the checks are the size-independent properties
of the scheme:

- total volume (FP64 sum, masscont) and tracer mass are conserved: the
  transports are computed from the solved levels, so the discrete
  continuity holds up to the residual of the zeta solve -- with the solve
  run to 1e-13 here, to round-off of the stored state (1e-12 relative with
  FP64 state; FP32 state rounds every stored level and transport);
- the TVD scheme keeps the tracer within its initial bounds;
- the two paths solve the same systems: their levels agree to the FP32
  storage and solver tolerances on all but the nodes whose wet/dry state
  differs.
"""
import dataclasses

import numpy as np
import pytest

from shyhip import mesh

pytestmark = pytest.mark.gpu

PARAMS = dict(dt=60.0, ireib=5, itvd=2, itvdv=1, aapar=0.0, block_rows=256, tri_sweeps=3, rtol=1e-13, atol=0.0,
              refine_rtol=1e-13)
NSTEPS = 8


@pytest.fixture(scope="module")
def basin():
    b = mesh.lagoon(500000, seed=20261202, lx=40000.0 * 10 ** 0.5, ly=20000.0 * 10 ** 0.5, nlv=20, tile=32,
                    inlet_depth=0.0)
    b = dataclasses.replace(b, ibnd=np.zeros(b.nkn, np.int32), ibtyp=np.array([0], np.int32),
                            bnd_nodes=np.zeros(0, np.int32))
    salt = mesh.smooth_tracer(b, 20261204, 25.0, 38.0)
    return b, salt


def _run(gpu, b, salt, prec):
    m = gpu.Model(b, prec, params=PARAMS, visv=1e-3, difv=1e-4, tracers=dict(salt=salt))
    m.set_wind(*mesh.wind_stress(10.0, 0.6))
    v0, m0 = m.volume(), m.tracer_mass("salt")
    for _ in range(NSTEPS):
        m.step(None)
    m.sync()
    out = dict(v0=v0, m0=m0, v=m.volume(), m=m.tracer_mass("salt"), znv=m.pull("znv"), iwegv=m.pull("iwegv"),
               salt=m.pull("salt"), stats=m.stats())
    m.close()
    return out


@pytest.fixture(scope="module")
def runs(gpu, basin):
    b, salt = basin
    return {p: _run(gpu, b, salt, p) for p in ("fp64", "fp32")}


@pytest.mark.parametrize("prec,vtol,mtol", [("fp64", 1e-12, 1e-12), ("fp32", 1e-9, 1e-8)])
def test_scale_conservation(runs, basin, prec, vtol, mtol):
    b, salt = basin
    r = runs[prec]
    dv = abs(r["v"] - r["v0"]) / r["v0"]
    dm = abs(r["m"] - r["m0"]) / r["m0"]
    print(prec, "nkn", b.nkn, "EL", int(b.ilhv.sum()), "stats", r["stats"], "dV/V", dv, "dM/M", dm)
    assert (r["iwegv"] != 0).sum() > 0          # some elements are dry
    assert dv < vtol
    assert dm < mtol
    valid = np.arange(b.nlv)[None, :] < b.ilhkv[:, None]
    lo, hi = float(salt[valid].min()), float(salt[valid].max())
    c = r["salt"][valid]
    slack = 1e-9 if prec == "fp64" else 1e-5 * hi
    assert c.min() >= lo - slack and c.max() <= hi + slack


def test_scale_paths_agree(runs):
    z64, z32 = runs["fp64"]["znv"], runs["fp32"]["znv"]
    dz = np.abs(z32 - z64)
    same = np.mean(runs["fp64"]["iwegv"] == runs["fp32"]["iwegv"])
    print("fp32 vs fp64 |dz| max", dz.max(), "p99", np.percentile(dz, 99), "median", np.median(dz),
          "wet/dry masks equal on", same)
    assert same > 0.999
    assert np.percentile(dz, 99) < 1e-6        # m
    assert np.median(dz) < 1e-7                # m
