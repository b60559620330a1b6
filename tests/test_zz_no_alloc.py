"""The allocation-free entry point shyhip_masscont_noalloc against the
entry points it is a variant of (shyhip_volume, shyhip_tracer_mass), through
the C ABI: the same node arrays, copied into device buffers the test owns
(allocated here through the HIP runtime the library itself uses), must give
the same bits.  Two calls in a row, the second on a smaller basin (fewer
partial sums than the first left in the shared array), on both precisions;
an oversized nkn is refused before any launch."""
import ctypes as C

import numpy as np
import pytest

from shyhip import mesh

pytestmark = pytest.mark.gpu


class _Hip:
    """hipMalloc/hipMemcpy/hipFree of the runtime libshyhip.so is linked to."""

    def __init__(self):
        self.rt = C.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.rt.hipDeviceSynchronize.argtypes = []
        self.bufs = []

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.rt.hipMalloc(C.byref(p), max(8, a.nbytes)) == 0
        self.bufs.append(p)
        assert self.rt.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0      # host to device
        return p

    def download(self, p, n, dtype):
        out = np.zeros(n, dtype)
        assert self.rt.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, out.nbytes, 2) == 0   # device to host
        return out

    def free(self):
        for p in self.bufs:
            self.rt.hipFree(p)
        self.bufs = []


def _noalloc(lib, hip, gm, conc_name=None):
    """shyhip_masscont_noalloc on device copies of the model's arrays (in
    the library's device layout a[l*nkn + k]); returns the FP64 result."""
    b = gm.basin
    dt = np.float64 if gm.prec == 0 else np.float32
    dtype = 1 if gm.prec == 0 else 0
    ilhkv = hip.upload(np.asarray(b.ilhkv, np.int32))
    hdk = hip.upload(gm.pull("hdknv", dtype=dt).T.copy())
    area = hip.upload(gm.pull("areakv", dtype=dt).T.copy())
    conc = hip.upload(gm.pull(conc_name, dtype=dt).T.copy()) if conc_name else None
    out = hip.upload(np.zeros(1, np.float64))
    rc = lib.shyhip_masscont_noalloc(dtype, b.nkn, ilhkv, hdk, area, conc, out, None)
    assert rc == 0, lib.shyhip_last_error(None)
    assert hip.rt.hipDeviceSynchronize() == 0
    v = float(hip.download(out, 1, np.float64)[0])
    hip.free()
    return v


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_noalloc_matches_volume_and_mass(gpu, prec):
    lib = gpu.load_library()
    lib.shyhip_masscont_noalloc.argtypes = [C.c_int, C.c_int] + [C.c_void_p] * 6
    hip = _Hip()
    big = mesh.lagoon(6000, seed=3, lx=10000.0, ly=5000.0, nlv=5, tile=16)
    small = mesh.tidal_basin()
    for b in (big, small):
        tr = dict(salt=mesh.smooth_tracer(b, 5, 25.0, 38.0))
        gm = gpu.Model(b, prec, params=dict(dt=60.0, ireib=5), tracers=tr)
        for n in range(3):
            gm.step(mesh.tide_level((n + 1) * 60.0) - 0.2)
        v_ref = gm.volume(mode=1)
        m_ref = gm.tracer_mass("salt")
        v = _noalloc(lib, hip, gm)
        m = _noalloc(lib, hip, gm, "salt")
        print(prec, b.name, "volume", v, v_ref, "mass", m, m_ref)
        assert v == v_ref
        assert m == m_ref
        gm.close()


def test_noalloc_refuses_oversized(gpu):
    lib = gpu.load_library()
    lib.shyhip_masscont_noalloc.argtypes = [C.c_int, C.c_int] + [C.c_void_p] * 6
    one = C.c_void_p(1)
    rc = lib.shyhip_masscont_noalloc(1, 65536 * 256 + 1, one, one, one, None, one, None)
    assert rc != 0
    assert b"allocation-free limit" in lib.shyhip_last_error(None)
