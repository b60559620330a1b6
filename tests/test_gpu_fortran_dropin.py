"""The Fortran drop-ins called by the reference's own names, from a program
built with flang (fortran/shyhip_dropin_check):

* the system_* solver plugin (simsys_shyhip.f90, replacing simsys_spkaux.f90
  at the SOLVER seam), driven as hydro_zeta drives it through
  mod_zeta_system's kn/hia/hik: the same bits as the plugin called through
  the C ABI, for the implicit system and for the lumped one
  (system_set_explicit);
* conz3d_omp (conz3d_shyhip.f90, the argument list of
  concentration_omp.f90:65-151) inside scal3sh's substep loop (clip,
  conz3d_omp, open-boundary treatment, for each isact of istot): the same
  bits as the library's own scal3sh with the same istot on the FP32 model,
  whose tracers are default reals as conz3d_omp's arguments are; on the
  FP64 model the tracer passes through a default real at every substep,
  so there the tracers agree to that rounding and the levels bit for bit."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from shyhip import mesh
import fortran_io
from test_gpu_system_plugin import _element_systems

pytestmark = pytest.mark.gpu


def _need_exe():
    if not os.path.exists(fortran_io.DROPIN_EXE):
        pytest.skip("fortran/shyhip_dropin_check not built (flang missing)")


def _plugin_capi(lib, h, b, hia, hik, rq, dt, explicit):
    if explicit:
        assert lib.shyhip_system_set_explicit(h) == 0
    assert lib.shyhip_system_init(h) == 0
    for ie in range(b.nel):
        kn = np.ascontiguousarray(b.nen3v[ie] + 1, dtype=np.int32)
        a = np.asfortranarray(hia[ie])
        assert lib.shyhip_system_assemble(h, ie + 1, kn.ctypes.data_as(C.c_void_p), a.ctypes.data_as(C.c_void_p),
                                          hik[ie].ctypes.data_as(C.c_void_p)) == 0
    assert lib.shyhip_system_add_rhs(h, C.c_float(dt), b.nkn, rq.ctypes.data_as(C.c_void_p)) == 0
    z0 = np.zeros(b.nkn, np.float32)
    assert lib.shyhip_system_solve(h, b.nkn, z0.ctypes.data_as(C.c_void_p)) == 0
    z = np.zeros(b.nkn, np.float32)
    assert lib.shyhip_system_get(h, b.nkn, z.ctypes.data_as(C.c_void_p)) == 0
    return z


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("explicit", [0, 1])
def test_system_plugin_by_reference_names(gpu, tmp_path, prec, explicit):
    _need_exe()
    b = mesh.tidal_basin()
    hia, hik = _element_systems(b, 3)
    if explicit:
        hia = (np.eye(3, dtype=np.float32)[None] * hia).astype(np.float32)     # lumped: diagonal only
    rq = np.random.default_rng(4).uniform(-1e-3, 1e-3, b.nkn).astype(np.float32)
    dt = np.float32(300.0)
    case, systems, res = (str(tmp_path / n) for n in ("case.bin", "sys.bin", "z.bin"))
    fortran_io.write_case(case, b, prec, 300.0, np.zeros(1, np.float32), ireib=5)
    with open(systems, "wb") as f:
        np.array([explicit], np.int32).tofile(f)
        np.array([dt], np.float32).tofile(f)
        np.ascontiguousarray(hia.transpose(0, 2, 1)).tofile(f)       # hia(i,j,ie): i fastest
        hik.tofile(f)
        rq.tofile(f)
        np.zeros(b.nkn, np.float32).tofile(f)
    print("running", fortran_io.DROPIN_EXE, "plugin", prec, explicit, flush=True)
    out = subprocess.run([fortran_io.DROPIN_EXE, "plugin", case, systems, res], capture_output=True, text=True,
                         timeout=120)
    print(out.stdout[-2000:], out.stderr[-2000:])
    assert out.returncode == 0
    zf = np.fromfile(res, np.float32)
    m = gpu.Model(b, prec, params=dict(dt=300.0, ireib=5))
    zc = _plugin_capi(m.lib, m.h, b, hia, hik, rq, dt, explicit)
    m.close()
    assert np.array_equal(zf, zc)
    if explicit:
        diag = np.zeros(b.nkn)
        rhs = np.zeros(b.nkn)
        np.add.at(diag, b.nen3v.reshape(-1), np.einsum("eii->ei", hia).astype(np.float64).reshape(-1))
        np.add.at(rhs, b.nen3v.reshape(-1), hik.astype(np.float64).reshape(-1))
        rhs += (dt * rq).astype(np.float64)
        assert np.allclose(zf, rhs / diag, rtol=1e-6, atol=0)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_conz3d_omp_by_reference_name(gpu, tmp_path, prec):
    _need_exe()
    b = mesh.lagoon(3000, seed=7, lx=8000.0, ly=4000.0, nlv=6, tile=16)
    tr = dict(temp=mesh.smooth_tracer(b, 8, 15.0, 25.0).astype(np.float32),
              salt=mesh.smooth_tracer(b, 9, 25.0, 38.0).astype(np.float32))
    nsteps, istot = 6, 2
    levels = np.asarray([mesh.tide_level((n + 1) * 60.0, amps=(0.4, 0.1), periods=(mesh.M2_PERIOD, mesh.K1_PERIOD))
                         - 0.3 for n in range(nsteps)], dtype=np.float32)
    wind = mesh.wind_stress(8.0, 0.7)
    case, res = str(tmp_path / "case.bin"), str(tmp_path / "result.bin")
    fortran_io.write_case(case, b, prec, 60.0, levels, ireib=5, itvd=2, itvdv=1, aapar=0.0,
                          czv=2.5e-3, visv=1e-3, difv=1e-4, wind=wind, tracers=tr)
    print("running", fortran_io.DROPIN_EXE, "transport", prec, flush=True)
    out = subprocess.run([fortran_io.DROPIN_EXE, "transport", case, str(istot), res], capture_output=True, text=True,
                         timeout=120)
    print(out.stdout[-2000:], out.stderr[-2000:])
    assert out.returncode == 0
    f = fortran_io.read_result(res, b, 2)
    params = dict(dt=60.0, ireib=5, itvd=2, itvdv=1, aapar=0.0)
    m = gpu.Model(b, prec, params=params, visv=np.float32(1e-3), difv=np.float32(1e-4), tracers=tr)
    m.set_wind(np.float32(wind[0]), np.float32(wind[1]))
    for n in range(nsteps):
        m.step(float(levels[n]), istot=istot)
    assert np.array_equal(m.pull("znv", np.float32), f["znv"])
    assert np.array_equal(m.pull("iwegv"), f["iwegv"])
    valid = np.arange(b.nlv)[None, :] < b.ilhkv[:, None]
    for i, (name, scale) in enumerate((("temp", 25.0), ("salt", 38.0))):
        g = m.pull(name, np.float32)
        d = float(np.max(np.abs(g.astype(np.float64) - f["conc"][i])[valid]))
        print(prec, name, "max |dc| Fortran drop-in vs library scal3sh:", d)
        if prec == "fp32":
            assert np.array_equal(g, f["conc"][i])
        else:
            assert d < 1e-6 * scale
    m.close()
