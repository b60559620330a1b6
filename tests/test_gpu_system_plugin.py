"""The system_* plugin (simsys_spkaux.f90:85-721) through the C ABI: element
matrices assembled by system_assemble(ie) in element order, the right-hand
side completed by system_add_rhs, solved on the device, read back by
system_get -- against scipy's direct solve of the same assembled system."""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from shyhip import mesh

pytestmark = pytest.mark.gpu


def _element_systems(b, seed):
    """Lumped mass + stiffness element matrices of a zeta-like system, in
    float32 as hydro_zeta hands them over (hia(3,3), hik(3))."""
    rng = np.random.default_rng(seed)
    x, y = b.xgv.astype(np.float64), b.ygv.astype(np.float64)
    xe, ye = x[b.nen3v], y[b.nen3v]
    area = 0.5 * ((xe[:, 1] - xe[:, 0]) * (ye[:, 2] - ye[:, 0]) - (xe[:, 2] - xe[:, 0]) * (ye[:, 1] - ye[:, 0]))
    bb = np.stack([ye[:, 1] - ye[:, 2], ye[:, 2] - ye[:, 0], ye[:, 0] - ye[:, 1]], 1) / (2 * area[:, None])
    cc = np.stack([xe[:, 2] - xe[:, 1], xe[:, 0] - xe[:, 2], xe[:, 1] - xe[:, 0]], 1) / (2 * area[:, None])
    k = 300.0 ** 2 * 9.81 * 0.25 * b.hev.astype(np.float64)
    hia = (area[:, None, None] / 3 * np.eye(3)[None]
           + (area * k)[:, None, None] * (bb[:, :, None] * bb[:, None, :] + cc[:, :, None] * cc[:, None, :]))
    hia = hia.astype(np.float32)
    hik = (area[:, None] / 3 * rng.uniform(-0.5, 0.5, (b.nel, 3))).astype(np.float32)
    return hia, hik


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_system_plugin_solves_assembled_system(gpu, prec):
    b = mesh.tidal_basin()
    m = gpu.Model(b, prec, params=dict(dt=300.0, ireib=5))
    lib, h = m.lib, m.h
    hia, hik = _element_systems(b, 3)
    rq = np.random.default_rng(4).uniform(-1e-3, 1e-3, b.nkn).astype(np.float32)
    dt = np.float32(300.0)
    assert lib.shyhip_system_init(h) == 0
    for ie in range(b.nel):
        kn = np.ascontiguousarray(b.nen3v[ie] + 1, dtype=np.int32)
        a = np.asfortranarray(hia[ie])                     # hia(i,j), column-major
        rc = lib.shyhip_system_assemble(h, ie + 1, kn.ctypes.data_as(C.c_void_p),
                                        a.ctypes.data_as(C.c_void_p), hik[ie].ctypes.data_as(C.c_void_p))
        assert rc == 0
    assert lib.shyhip_system_add_rhs(h, C.c_float(dt), b.nkn, rq.ctypes.data_as(C.c_void_p)) == 0
    z0 = np.zeros(b.nkn, np.float32)
    assert lib.shyhip_system_solve(h, b.nkn, z0.ctypes.data_as(C.c_void_p)) == 0
    z = np.zeros(b.nkn, np.float32)
    assert lib.shyhip_system_get(h, b.nkn, z.ctypes.data_as(C.c_void_p)) == 0

    rows = np.repeat(b.nen3v, 3, axis=1).reshape(-1)
    cols = np.tile(b.nen3v, (1, 3)).reshape(-1)
    A = sp.coo_matrix((hia.astype(np.float64).reshape(-1), (rows, cols)), shape=(b.nkn, b.nkn)).tocsc()
    rhs = np.zeros(b.nkn)
    np.add.at(rhs, b.nen3v.reshape(-1), hik.astype(np.float64).reshape(-1))
    rhs += (dt * rq).astype(np.float64)
    ref = spla.spsolve(A, rhs)
    err = np.abs(z - ref).max() / np.abs(ref).max()
    print(prec, "plugin solve rel err", err)
    assert err < 1e-6                   # solver tolerance 1e-8 + the float32 result of system_get
    # a size mismatch is an error, not a crash
    assert lib.shyhip_system_get(h, b.nkn - 1, z.ctypes.data_as(C.c_void_p)) != 0
    m.close()
