"""The Fortran host program (fortran/shyhip_run, ISO_C_BINDING over the C
ABI) drives the same device path as the Python host and produces the same
state: the two hosts make the same calls in the same order, so results are
bit-identical."""
import os
import subprocess

import numpy as np
import pytest

from shyhip import mesh
import fortran_io

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_fortran_host_matches_python_host(gpu, tmp_path, prec):
    if not os.path.exists(fortran_io.RUN_EXE):
        pytest.skip("fortran/shyhip_run not built (flang missing)")
    b = mesh.lagoon(3000, seed=7, lx=8000.0, ly=4000.0, nlv=6, tile=16)
    tr = dict(temp=mesh.smooth_tracer(b, 8, 15.0, 25.0).astype(np.float32),
              salt=mesh.smooth_tracer(b, 9, 25.0, 38.0).astype(np.float32))
    nsteps = 8
    levels = [mesh.tide_level((n + 1) * 60.0, amps=(0.4, 0.1), periods=(mesh.M2_PERIOD, mesh.K1_PERIOD)) - 0.3
              for n in range(nsteps)]
    levels = np.asarray(levels, dtype=np.float32)
    wind = mesh.wind_stress(8.0, 0.7)
    params = dict(dt=60.0, ireib=5, itvd=2, itvdv=1, aapar=0.0)
    case, res = str(tmp_path / "case.bin"), str(tmp_path / "result.bin")
    fortran_io.write_case(case, b, prec, 60.0, levels, ireib=5, itvd=2, itvdv=1, aapar=0.0,
                          czv=2.5e-3, visv=1e-3, difv=1e-4, wind=wind, tracers=tr)
    print("running", fortran_io.RUN_EXE, prec, flush=True)
    out = subprocess.run([fortran_io.RUN_EXE, case, res], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr)
    assert out.returncode == 0
    f = fortran_io.read_result(res, b, 2)

    m = gpu.Model(b, prec, params=params, visv=np.float32(1e-3), difv=np.float32(1e-4), tracers=tr)
    m.set_wind(np.float32(wind[0]), np.float32(wind[1]))
    for n in range(nsteps):
        m.step(float(levels[n]))
    assert np.array_equal(m.pull("znv", np.float32), f["znv"])
    assert np.array_equal(m.pull("utlnv", np.float32), f["utlnv"])
    assert np.array_equal(m.pull("temp", np.float32), f["conc"][0])
    assert np.array_equal(m.pull("salt", np.float32), f["conc"][1])
    assert np.array_equal(m.pull("iwegv"), f["iwegv"])
    assert m.volume() == f["volume"]
    m.close()
