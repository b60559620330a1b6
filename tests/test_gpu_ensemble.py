"""Ensembles through the C ABI: members advanced together in one model give
the results they give alone (isolation), within the zeta-solver tolerance;
integer structures (colouring, wet/dry masks) are identical."""
import numpy as np
import pytest

from shyhip import ensemble, mesh

pytestmark = pytest.mark.gpu

PARAMS = dict(dt=60.0, ireib=5, itvd=2, itvdv=1, aapar=0.0, block_rows=256, tri_sweeps=3)


def _members():
    out = []
    for n, s, lx, speed, phase in ((1500, 11, 6000.0, 5.0, 0.0), (2600, 12, 8000.0, 12.0, 1.3),
                                   (1100, 13, 5000.0, 20.0, 2.9)):
        b = mesh.lagoon(n, seed=s, lx=lx, ly=lx / 2, nlv=4, tile=16)
        salt = mesh.smooth_tracer(b, s + 100, 25.0, 38.0).astype(np.float32)
        out.append(dict(basin=b, wind=mesh.wind_stress(speed, 0.5 + phase),
                        level=(lambda t, ph=phase: mesh.tide_level(t, amps=(0.4, 0.1),
                                                                  periods=(mesh.M2_PERIOD, mesh.K1_PERIOD),
                                                                  phases=(ph, 0.7)) - 0.2),
                        tracers=dict(salt=salt)))
    return out


def _alone(gpu, mem, prec, nsteps):
    m = gpu.Model(mem["basin"], prec, params=PARAMS, visv=1e-3, difv=1e-4, tracers=mem["tracers"])
    m.set_wind(*mem["wind"])
    t = 0.0
    for _ in range(nsteps):
        t += PARAMS["dt"]
        m.step(mem["level"](t))
    out = dict(znv=m.pull("znv"), iwegv=m.pull("iwegv"), salt=m.pull("salt"), color=m.pattern()[3])
    m.close()
    return out


@pytest.mark.parametrize("prec,ztol", [("fp64", 1e-6), ("fp32", 2e-5)])
def test_ensemble_members_match_alone(gpu, prec, ztol):
    mems = _members()
    nsteps = 12
    ens = ensemble.Ensemble(mems, prec, params=PARAMS, visv=1e-3, difv=1e-4)
    for _ in range(nsteps):
        ens.step()
    color = ens.model.pattern()[3]
    for i, mem in enumerate(mems):
        ref = _alone(gpu, mem, prec, nsteps)
        e0, e1 = ens.u.elem_start[i], ens.u.elem_start[i + 1]
        assert np.array_equal(color[e0:e1], ref["color"]), "member colouring differs"
        assert np.array_equal(ens.pull("iwegv", i), ref["iwegv"])
        dz = float(np.max(np.abs(ens.pull("znv", i) - ref["znv"])))
        b = mem["basin"]
        valid = np.arange(b.nlv)[None, :] < b.ilhkv[:, None]
        ds = float(np.max(np.abs(ens.pull("salt", i) - ref["salt"])[valid]))
        print(prec, "member", i, "nkn", b.nkn, "|dz| =", dz, "|dsalt| =", ds)
        assert dz < ztol                                   # m: solver tolerance of the union solve
        assert ds < 1e-5 * 38.0                            # relative to the field scale
    ens.close()


def test_ensemble_members_choose_own_substeps(gpu):
    """With istot from scalar_stability, each member runs the substeps it
    chooses alone (the kernels leave a member alone past its last substep),
    so the members still match their stand-alone runs; the members here
    choose different counts."""
    mems = _members()
    nsteps = 8
    ens = ensemble.Ensemble(mems, "fp64", params=PARAMS, visv=1e-3, difv=1e-4)
    seen = []
    for _ in range(nsteps):
        ens.step(istot=0)
        seen.append(tuple(ens.member_istot()))
    print("member istot per step:", seen)
    assert any(len(set(s)) > 1 for s in seen)
    for i, mem in enumerate(mems):
        m = gpu.Model(mem["basin"], "fp64", params=PARAMS, visv=1e-3, difv=1e-4, tracers=mem["tracers"])
        m.set_wind(*mem["wind"])
        t = 0.0
        alone = []
        for _ in range(nsteps):
            t += PARAMS["dt"]
            m.step(mem["level"](t), istot=0)
            alone.append(m.last_istot())
        assert alone == [s[i] for s in seen]
        assert np.array_equal(ens.pull("iwegv", i), m.pull("iwegv"))
        b = mem["basin"]
        valid = np.arange(b.nlv)[None, :] < b.ilhkv[:, None]
        dz = float(np.max(np.abs(ens.pull("znv", i) - m.pull("znv"))))
        ds = float(np.max(np.abs(ens.pull("salt", i) - m.pull("salt"))[valid]))
        print("member", i, "|dz| =", dz, "|dsalt| =", ds)
        assert dz < 1e-6 and ds < 1e-5 * 38.0
        m.close()
    ens.close()
