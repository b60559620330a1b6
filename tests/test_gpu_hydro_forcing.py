"""GPU parity of the optional hydro forcing terms against the oracle (pinned
in test_oracle_hydro_forcing.py): atmospheric pressure ppv and equilibrium
tide zeqv in hydro_intern (hydrodynamic.f90:1025-1033,1235-1251), the level
nudging term andgzv in hydro_zeta (:612,679), and elements with fixed
transports (iuvfix, :627-629,1383-1392,1525) taking prescribed transports
each step as bclfix writes them (bcvel.f90:332-338) or keeping the previous
step's.

Bars as in test_gpu_parity: masks bit-exact; zeta within 1e-6 m (FP64 vs the
FP64 oracle) and 2e-5 m (FP32 vs the reference-precision oracle);
transports within 1e-5 (FP64) / 1e-3 (FP32) of their scale; fixed
elements' transports bit-exact.  Each term moves zeta by far more than
the bar (checked against the same run without it).
"""
import numpy as np
import pytest

from shyhip import mesh

pytestmark = pytest.mark.gpu


def _case(n=3000, nlv=6, seed=7):
    b = mesh.lagoon(n, seed=seed, lx=8000.0, ly=4000.0, nlv=nlv, tile=16)
    tr = dict(salt=mesh.smooth_tracer(b, seed + 2, 25.0, 38.0))
    params = dict(dt=60.0, ireib=5, itpre=1, itvd=2, itvdv=1, aapar=0.0, block_rows=0)
    return b, tr, params


def _forcing(b, which):
    x, y = b.xgv.astype(np.float64), b.ygv.astype(np.float64)
    rng = np.random.default_rng(3)
    f = {}
    if which in ("ppv", "all"):
        f["ppv"] = 101325.0 - 1500.0 * np.exp(-((x - 5000.0) ** 2 + (y - 2500.0) ** 2) / 2000.0 ** 2)
    if which in ("zeqv", "all"):
        f["zeqv"] = 0.05 * np.sin(2 * np.pi * x / 16000.0) * np.cos(2 * np.pi * y / 12000.0)
    if which in ("andgzv", "all"):
        f["andgzv"] = 2e-5 * np.cos(2 * np.pi * x / 8000.0)
    if which in ("uvfix", "uvkeep", "all"):
        fix = np.zeros(b.nel, np.int32)
        fix[rng.choice(b.nel, b.nel // 20, replace=False)] = 1
        f["iuvfix"] = fix
        if which != "uvkeep":
            f["utfix"] = rng.uniform(-0.05, 0.05, (b.nel, b.nlv))
            f["vtfix"] = rng.uniform(-0.05, 0.05, (b.nel, b.nlv))
    return f


def _run(gpu, b, tr, params, prec, oprec, f, nsteps=10):
    import pyoracle
    om = pyoracle.OracleModel(b, oprec, params=params, visv=1e-3, difv=1e-4, tracers=tr)
    gm = gpu.Model(b, prec, params=params, visv=1e-3, difv=1e-4, tracers=tr)
    tx, ty = mesh.wind_stress(8.0, 0.7)
    om.set_wind(tx, ty)
    gm.set_wind(tx, ty)
    for name in ("ppv", "zeqv", "andgzv"):
        if name in f:
            om.set_forcing(name, f[name])
            gm.push(name, f[name])
    if "iuvfix" in f:
        om.set_uvfix(f["iuvfix"], f.get("utfix"), f.get("vtfix"))
        gm.push("iuvfix", f["iuvfix"])
        if "utfix" in f:
            gm.push("utfix", f["utfix"])
            gm.push("vtfix", f["vtfix"])
    for n in range(nsteps):
        lev = mesh.tide_level((n + 1) * 60.0, amps=(0.4, 0.1), periods=(mesh.M2_PERIOD, mesh.K1_PERIOD)) - 0.3
        om.step(lev)
        gm.step(lev)
    return om, gm


@pytest.mark.parametrize("which", ["ppv", "zeqv", "andgzv", "uvfix", "uvkeep", "all"])
def test_forcing_term_fp64(gpu, which):
    b, tr, params = _case()
    f = _forcing(b, which)
    om, gm = _run(gpu, b, tr, params, "fp64", "f64", f)
    z = gm.pull("znv")
    dz = float(np.max(np.abs(z - om.a["znv"])))
    _, g0 = _run(gpu, b, tr, params, "fp64", "f64", {}, nsteps=10)
    effect = float(np.max(np.abs(z - g0.pull("znv"))))
    print(which, "|dz| vs oracle", dz, "effect of the term", effect, gm.stats())
    assert np.array_equal(gm.pull("iwegv"), om.a["iwegv"])
    assert dz < 1e-6
    assert effect > 1e3 * max(dz, 1e-9)
    ut, vt = gm.pull("utlnv"), gm.pull("vtlnv")
    scale = max(float(np.abs(om.a["utlnv"]).max()), float(np.abs(om.a["vtlnv"]).max()))
    assert float(np.abs(ut - om.a["utlnv"]).max()) < 1e-5 * scale
    assert float(np.abs(vt - om.a["vtlnv"]).max()) < 1e-5 * scale
    if "iuvfix" in f:
        fixed = f["iuvfix"] != 0
        valid = np.arange(b.nlv)[None, :] < b.ilhv[:, None]
        sel = fixed[:, None] & valid
        assert np.array_equal(ut[sel], om.a["utlnv"][sel]) and np.array_equal(vt[sel], om.a["vtlnv"][sel])
        if "utfix" in f:
            assert np.array_equal(ut[sel], f["utfix"][sel])


def test_forcing_all_fp32(gpu):
    b, tr, params = _case()
    f = _forcing(b, "all")
    om, gm = _run(gpu, b, tr, params, "fp32", "f32", f)
    dz = float(np.max(np.abs(gm.pull("znv") - om.a["znv"])))
    print("fp32 all terms |dz| vs oracle", dz, gm.stats())
    assert np.array_equal(gm.pull("iwegv"), om.a["iwegv"])
    assert dz < 2e-5
    ut = gm.pull("utlnv", dtype=np.float32)
    fixed = f["iuvfix"] != 0
    valid = np.arange(b.nlv)[None, :] < b.ilhv[:, None]
    sel = fixed[:, None] & valid
    assert np.array_equal(ut[sel], om.a["utlnv"][sel])
    scale = float(np.abs(om.a["utlnv"]).max())
    assert float(np.abs(gm.pull("utlnv") - om.a["utlnv"]).max()) < 1e-3 * scale


def test_optional_field_pull_before_push(gpu):
    b, tr, params = _case(n=800, nlv=2)
    gm = gpu.Model(b, "fp64", params=params)
    with pytest.raises(gpu.ShyhipError, match="never pushed"):
        gm.pull("ppv")
    gm.push("ppv", np.full(b.nkn, 101325.0))
    assert np.all(gm.pull("ppv") == 101325.0)
