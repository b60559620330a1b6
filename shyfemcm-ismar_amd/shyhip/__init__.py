"""Python host-side driver of libshyhip.so (the gfx950 SHYFEM hydro step).

Mirrors the reference's call surface for the time step -- ``hydro``
(src/shyfem/hydrodynamic.f90:284), ``scal3sh``/``conz3d``
(src/shyfem/concentration.f90:565, concentration_omp.f90:65) and the
``system_*`` solver plugin (src/shyfem/simsys_spkaux.f90:85-721) -- through the
C ABI declared in include/shyhip.h.  Arrays cross the boundary in the
reference's own layout: ``a(l, ie)`` with the layer index fastest, which in
numpy is a C-ordered array of shape ``(nel, nlvdi)``.

The native library is required: importing works without it, but creating a
:class:`Model` raises if ``libshyhip.so`` is missing or no GPU is visible.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
LIBPATH = os.path.join(PKG, "libshyhip.so")

FP64, FP32 = 0, 1
F32, F64, I32 = 0, 1, 2

FIELDS = dict(znv=1, zov=2, zenv=3, zeov=4, utlnv=5, vtlnv=6, utlov=7, vtlov=8, unv=9, vnv=10, uov=11, vov=12,
              ulnv=13, vlnv=14, wlnv=15, wlov=16, hdknv=17, hdkov=18, hdenv=19, hdeov=20, iwegv=21, iwetv=22,
              inodv=23, rfricv=24, ddxv=25, ddyv=26, visv=27, difv=28, difhv=29, tauxnv=30, tauynv=31, rzv=32,
              rqv=33, mfluxv=34, areakv=35, csr_val=36, rhs=37, rhov=38,
              # optional hydro forcing (absent until pushed)
              ppv=39, zeqv=40, andgzv=41, iuvfix=42, utfix=43, vtfix=44)
TRACER0 = 100
TRACER_BC0 = 200
INT_FIELDS = {"iwegv", "iwetv", "inodv", "iuvfix"}


class ShyParams(C.Structure):
    _fields_ = [("dt", C.c_double), ("azpar", C.c_float), ("ampar", C.c_float), ("afpar", C.c_float),
                ("atpar", C.c_float), ("adpar", C.c_float), ("aapar", C.c_float), ("grav", C.c_float),
                ("rowass", C.c_float), ("vismol", C.c_float), ("ireib", C.c_int), ("uvmin", C.c_float),
                ("hzmin", C.c_float), ("hzoff", C.c_float), ("hzon", C.c_float), ("volmin", C.c_float),
                ("iclin", C.c_int), ("rtol", C.c_double), ("atol", C.c_double), ("itermax", C.c_int),
                ("refine_rtol", C.c_double), ("block_rows", C.c_int), ("tri_sweeps", C.c_int),
                ("inner_rtol", C.c_double), ("rkpar", C.c_float),
                ("difmol", C.c_float), ("itvd", C.c_int), ("itvdv", C.c_int), ("node_tile", C.c_int),
                ("idtype", C.c_int)]


DEFAULTS = dict(dt=300.0, azpar=0.5, ampar=0.5, afpar=0.5, atpar=1.0, adpar=1.0, aapar=0.0, grav=9.81,
                rowass=1025.0, vismol=1.0e-6, ireib=5, uvmin=0.2, hzmin=0.01, hzoff=0.05, hzon=0.10, volmin=1.0,
                iclin=0, rtol=float(np.float32(1.0e-8)), atol=float(np.float32(1.0e-8)), itermax=1000,
                refine_rtol=1.0e-10, block_rows=0, tri_sweeps=0, inner_rtol=1.0e-5, rkpar=0.0, difmol=1.0e-6, itvd=0, itvdv=0,
                node_tile=0, idtype=2)

# wind conversion settings (shyhip_wind_params; STR defaults, def_para.f90:782-799)
class WindParams(C.Structure):
    _fields_ = [("iwtype", C.c_int), ("itdrag", C.c_int), ("dragco", C.c_float), ("roluft", C.c_float),
                ("wslim", C.c_float), ("wsmax", C.c_float)]


WIND_DEFAULTS = dict(iwtype=1, itdrag=0, dragco=2.5e-3, roluft=1.225, wslim=-1.0, wsmax=50.0)

_lib = None


class ShyhipError(RuntimeError):
    pass


def load_library(path: str = LIBPATH):
    """Load libshyhip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ShyhipError(f"{path} is missing: build it with `make -C {PKG}` (or __graft_entry__.build())")
    lib = C.CDLL(path)
    lib.shyhip_last_error.restype = C.c_char_p
    lib.shyhip_last_error.argtypes = [C.c_void_p]
    _lib = lib
    return lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def check_params(params: dict | None = None) -> None:
    """Raises ShyhipError if the device path refuses these parameters (the
    checks of shyhip_set_params, run without a GPU)."""
    lib = load_library()
    p = dict(DEFAULTS)
    if params:
        p.update(params)
    if lib.shyhip_check_params(C.byref(ShyParams(**p))):
        raise ShyhipError(f"check_params: {lib.shyhip_last_error(None).decode()}")


class Model:
    """One basin resident on the GPU, stepped through libshyhip."""

    def __init__(self, basin, precision: str = "fp64", params: dict | None = None, czv: float = 2.5e-3,
                 fcor: float = 0.0, visv: float = 0.0, difv: float = 0.0, zconst: float = 0.0,
                 tracers: dict | None = None, members=None):
        """members = (node_start, elem_start): the basin is a union of
        independent member basins (shyhip.ensemble)."""
        self.lib = lib = load_library()
        self.basin = b = basin
        self.prec = FP64 if precision == "fp64" else FP32
        self.np_real = np.float64 if self.prec == FP64 else np.float32
        p = dict(DEFAULTS)
        if params:
            p.update(params)
        self.params_dict = p
        self.tracer_names = list(tracers.keys()) if tracers else []
        h = C.c_void_p()
        self.nlvdi = b.nlv
        rc = lib.shyhip_create(C.byref(h), self.prec, b.nkn, b.nel, b.nlv, b.nlv, len(self.tracer_names))
        if rc:
            raise ShyhipError(f"shyhip_create: {lib.shyhip_last_error(None).decode()}")
        self.h = h
        self._chk(lib.shyhip_set_params(h, C.byref(ShyParams(**p))), "set_params")
        if members is not None:
            ns = np.ascontiguousarray(members[0], dtype=np.int32)
            es = np.ascontiguousarray(members[1], dtype=np.int32)
            self._chk(lib.shyhip_set_members(h, len(ns) - 1, _p(ns), _p(es)), "set_members")
        nen = np.ascontiguousarray(b.nen3v + 1, dtype=np.int32)
        f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)  # noqa: E731
        self._keep = [nen]
        self._chk(lib.shyhip_set_basin(h, _p(nen), _p(f32(b.xgv)), _p(f32(b.ygv)), _p(f32(b.hm3v)), _p(f32(b.hev)),
                                       _p(f32(b.hlv)), _p(f32(b.hldv)), _p(np.ascontiguousarray(b.ilhv, np.int32)),
                                       _p(np.ascontiguousarray(b.ilhkv, np.int32))), "set_basin")
        ibtyp = np.ascontiguousarray(b.ibtyp, dtype=np.int32)
        self._chk(lib.shyhip_set_boundaries(h, len(ibtyp) - 1, _p(np.ascontiguousarray(b.ibnd, np.int32)),
                                            _p(ibtyp), _p(np.ascontiguousarray(b.iextb, np.int32))), "set_boundaries")
        self._chk(lib.shyhip_set_coefficients(h, _p(np.full(b.nel, fcor, np.float32)),
                                              _p(np.full(b.nel, czv, np.float32))), "set_coefficients")
        self.push("visv", np.full((b.nkn, b.nlv + 1), visv, np.float32))
        self.push("difv", np.full((b.nkn, b.nlv + 1), difv, np.float32))
        if p["itvd"] == 2:
            self._chk(lib.shyhip_tvd_init(h), "tvd_init")
        self._chk(lib.shyhip_init_state(h, C.c_float(zconst)), "init_state")
        if tracers:
            for i, (name, arr) in enumerate(tracers.items()):
                a = np.asarray(arr, dtype=np.float64)
                self.push(TRACER0 + i, a)
                # open-boundary values: the initial values on the boundary nodes
                bc = np.full((b.nkn, b.nlv), -999.0)
                bc[b.bnd_nodes] = a[b.bnd_nodes]
                self.push(TRACER_BC0 + i, bc)
        self.t = 0.0

    # ------------------------------------------------------------------
    def _chk(self, rc, what):
        if rc:
            msg = self.lib.shyhip_last_error(self.h).decode()
            raise ShyhipError(f"shyhip_{what} failed ({rc}): {msg}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.shyhip_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _shape(self, name):
        b = self.basin
        L = self.nlvdi
        shapes = dict(znv=(b.nkn,), zov=(b.nkn,), zenv=(b.nel, 3), zeov=(b.nel, 3), unv=(b.nel,), vnv=(b.nel,),
                      uov=(b.nel,), vov=(b.nel,), wlnv=(b.nkn, L + 1), wlov=(b.nkn, L + 1), visv=(b.nkn, L + 1),
                      difv=(b.nkn, L + 1), iwegv=(b.nel,), iwetv=(b.nel,), inodv=(b.nkn,), rfricv=(b.nel,),
                      ddxv=(b.nel, 2 * L), ddyv=(b.nel, 2 * L), tauxnv=(b.nkn,), tauynv=(b.nkn,), rzv=(b.nkn,),
                      rqv=(b.nkn,), hdknv=(b.nkn, L), hdkov=(b.nkn, L), mfluxv=(b.nkn, L), areakv=(b.nkn, L),
                      rhs=(b.nkn,), rhov=(b.nkn, L), ppv=(b.nkn,), zeqv=(b.nkn,), andgzv=(b.nkn,),
                      iuvfix=(b.nel,))
        if name in shapes:
            return shapes[name]
        if name in ("utlnv", "vtlnv", "utlov", "vtlov", "ulnv", "vlnv", "hdenv", "hdeov", "difhv", "utfix", "vtfix"):
            return (b.nel, L)
        if isinstance(name, int) and name >= TRACER0:
            return (b.nkn, L)
        raise KeyError(name)

    def field_id(self, name):
        """Field id of a state array or a tracer (by name)."""
        if isinstance(name, int):
            return name
        if name in self.tracer_names:
            return TRACER0 + self.tracer_names.index(name)
        return FIELDS[name]

    def push(self, name, arr):
        fid = name if isinstance(name, int) else FIELDS[name]
        if name in INT_FIELDS:
            a = np.ascontiguousarray(arr, dtype=np.int32)
            dt = I32
        else:
            a = np.ascontiguousarray(arr, dtype=np.float64)
            dt = F64
        self._chk(self.lib.shyhip_push(self.h, fid, _p(a), dt, C.c_size_t(a.size)), f"push({name})")

    def pull(self, name, dtype=np.float64):
        if isinstance(name, str) and name in self.tracer_names:
            fid = TRACER0 + self.tracer_names.index(name)
            shape = (self.basin.nkn, self.nlvdi)
        elif name == "csr_val":
            fid, shape = FIELDS[name], (self.nnz(),)
        else:
            fid = name if isinstance(name, int) else FIELDS[name]
            shape = self._shape(name)
        if name in INT_FIELDS:
            out = np.zeros(shape, dtype=np.int32)
            self._chk(self.lib.shyhip_pull(self.h, fid, _p(out), I32, C.c_size_t(out.size)), f"pull({name})")
            return out
        out = np.zeros(shape, dtype=np.float64 if dtype == np.float64 else np.float32)
        self._chk(self.lib.shyhip_pull(self.h, fid, _p(out), F64 if out.dtype == np.float64 else F32,
                                       C.c_size_t(out.size)), f"pull({name})")
        return out

    def nnz(self):
        n = C.c_int()
        nc = C.c_int()
        self._chk(self.lib.shyhip_get_pattern(self.h, C.byref(n), None, None, C.byref(nc), None), "get_pattern")
        return n.value

    def pattern(self):
        n = self.nnz()
        ntg = np.zeros(self.basin.nkn + 1, np.int32)
        cols = np.zeros(n, np.int32)
        color = np.zeros(self.basin.nel, np.int32)
        nc = C.c_int()
        self._chk(self.lib.shyhip_get_pattern(self.h, None, _p(ntg), _p(cols), C.byref(nc), _p(color)), "get_pattern")
        return ntg, cols, nc.value, color

    # ------------------------------------------------------------------ forcing / step
    def set_boundary(self, level, ibc=1):
        self._chk(self.lib.shyhip_set_boundary_level(self.h, ibc, C.c_float(level)), "set_boundary_level")

    def set_boundary_levels(self, levels):
        """Levels of all open boundaries (boundary b = 1..nbnd gets levels[b-1])."""
        lv = np.ascontiguousarray(levels, dtype=np.float32)
        self._chk(self.lib.shyhip_set_boundary_levels(self.h, len(lv), _p(lv)), "set_boundary_levels")

    def set_wind(self, taux, tauy):
        self._chk(self.lib.shyhip_set_wind_uniform(self.h, C.c_float(taux), C.c_float(tauy)), "set_wind_uniform")

    def set_member_wind(self, m, taux, tauy):
        """Uniform wind stress on the nodes of member m of a union basin."""
        self._chk(self.lib.shyhip_set_member_wind(self.h, m, C.c_float(taux), C.c_float(tauy)), "set_member_wind")

    def copy_member_state(self, m, src, sm):
        """Member m of this model takes the state of member sm of model src
        (the same basin), device to device."""
        self._chk(self.lib.shyhip_copy_member_state(self.h, m, src.h, sm), "copy_member_state")

    def set_boundary_series(self, ibc, time, values, intpol=0, zfact=1.0, dtime0=None):
        """Boundary-level file of level boundary ibc held on the device:
        time (nrec,) in seconds, values (nrec,) for one level on the whole
        boundary or (nrec, nk) per node (the boundary's nodes in ascending
        order); intpol 2 or 4 records (0: 4); the first ring is chosen for
        dtime0 (default: the model's current time).  Every step then sets
        the level from it (sp111(2))."""
        t = np.ascontiguousarray(time, dtype=np.float64)
        v = np.ascontiguousarray(values, dtype=np.float32)
        nval = 1 if v.ndim == 1 else v.shape[1]
        d0 = self.t if dtime0 is None else dtime0
        self._chk(self.lib.shyhip_set_boundary_series(self.h, ibc, len(t), _p(t), _p(v), nval, intpol,
                                                      C.c_float(zfact), C.c_double(d0)), "set_boundary_series")

    def set_wind_series(self, time, wx, wy, intpol=0, dtime0=None, **wind):
        """Wind file held on the device: time (nrec,), wx/wy (nrec,) for one
        point or (nrec, nkn) node values; wind settings as WIND_DEFAULTS
        (iwtype 1: wind velocity, 2: stress).  Every step converts it to the
        stress (meteo_convert_wind_data).  time=None removes the series."""
        if time is None:
            self._chk(self.lib.shyhip_set_wind_series(self.h, 0, None, None, None, 1, 0, C.c_double(0.0), None),
                      "set_wind_series")
            return
        w = dict(WIND_DEFAULTS)
        w.update(wind)
        t = np.ascontiguousarray(time, dtype=np.float64)
        x = np.ascontiguousarray(wx, dtype=np.float32)
        y = np.ascontiguousarray(wy, dtype=np.float32)
        nval = 1 if x.ndim == 1 else x.shape[1]
        d0 = self.t if dtime0 is None else dtime0
        self._chk(self.lib.shyhip_set_wind_series(self.h, len(t), _p(t), _p(x), _p(y), nval, intpol, C.c_double(d0),
                                                  C.byref(WindParams(**w))), "set_wind_series")

    def set_baroclinic(self, temp="temp", salt="salt", ibarcl=1):
        """Baroclinic coupling with the named tracers as T and S (ibarcl 1;
        0 switches it off): starts rhov as barocl(0) does, then every step
        adds the baroclinic term before hydro and updates rhov after the
        transport."""
        it = self.tracer_names.index(temp) if ibarcl else 0
        isl = self.tracer_names.index(salt) if ibarcl else 0
        self._chk(self.lib.shyhip_set_baroclinic(self.h, ibarcl, it, isl), "set_baroclinic")

    def rhoset(self, niter=1):
        self._chk(self.lib.shyhip_rhoset(self.h, niter), "rhoset")

    def set_tracer_terms(self, name, robs=0.0, cobs=None, rtauv=None, wsink=0.0, wsinkv=None, rload=0.0,
                         load=None):
        """conz3d_omp's own terms of one tracer: nudging robs * rtauv towards
        cobs ([nkn, nlv]), sinking wsink * wsinkv ([nkn, nlv+1]), loads
        rload * load ([nkn, nlv]; tracer_load() reads what deposition left)."""
        i = self.tracer_names.index(name)
        for base, arr in ((300, cobs), (400, rtauv), (500, wsinkv), (600, load)):
            if arr is not None:
                a = np.ascontiguousarray(arr, dtype=np.float64)
                self._chk(self.lib.shyhip_push(self.h, base + i, _p(a), F64, C.c_size_t(a.size)), "push(terms)")
        self._chk(self.lib.shyhip_set_tracer_terms(self.h, i, C.c_float(robs), C.c_float(wsink), C.c_float(rload)),
                  "set_tracer_terms")

    def tracer_load(self, name):
        i = self.tracer_names.index(name)
        out = np.zeros((self.basin.nkn, self.nlvdi))
        self._chk(self.lib.shyhip_pull(self.h, 600 + i, _p(out), F64, C.c_size_t(out.size)), "pull(load)")
        return out

    def set_sources(self, nodes, levmin=None, levmax=None):
        """Point sources of set_mass_flux: 0-based nodes in the reference's
        boundary loop order, with their boundary's levmin/levmax (0: the
        whole column).  An empty list keeps the distributed sources only."""
        n = np.ascontiguousarray(np.asarray(nodes, np.int64) + 1, dtype=np.int32)
        lo = np.ascontiguousarray(np.zeros(len(n)) if levmin is None else levmin, dtype=np.int32)
        hi = np.ascontiguousarray(np.zeros(len(n)) if levmax is None else levmax, dtype=np.int32)
        if not (len(lo) == len(hi) == len(n)):
            raise ValueError("set_sources: nodes, levmin and levmax differ in length")
        self._src = (n, lo, hi)
        self._chk(self.lib.shyhip_set_sources(self.h, len(n), _p(n), _p(lo), _p(hi)), "set_sources")

    def set_mass_flux(self, rflux=None, rqdsv=None):
        """set_mass_flux on the device: discharges per source (m3/s) and
        distributed sources per node (None: none) into mfluxv and rqv."""
        q = None if rflux is None else np.ascontiguousarray(rflux, dtype=np.float32)
        d = None if rqdsv is None else np.ascontiguousarray(rqdsv, dtype=np.float32)
        if d is not None and d.shape != (self.basin.nkn,):
            raise ValueError("set_mass_flux: rqdsv must hold one value per node")
        self._chk(self.lib.shyhip_set_mass_flux(self.h, None if q is None else _p(q), None if d is None else _p(d)),
                  "set_mass_flux")

    def forcing(self, dtime):
        """Evaluate the installed series at dtime (sp111(2) outside a step)."""
        self._chk(self.lib.shyhip_forcing(self.h, C.c_double(dtime)), "forcing")

    def step(self, bnd_level=None, istot=1):
        """set_dry, set_timestep, copy_uvz/copy_depth, sp111 level, hydro,
        scal3sh (istot <= 0: substeps from scalar_stability, as the
        reference chooses them)."""
        if bnd_level is not None:
            self.set_boundary(bnd_level)
        self.t += self.params_dict["dt"]
        self._chk(self.lib.shyhip_step(self.h, C.c_double(self.t), istot), "step")

    def sync(self):
        self._chk(self.lib.shyhip_sync(self.h), "sync")

    def set_deferred_checks(self, on=True):
        """Steps return without waiting for the device; a device error of a
        step's final phase is raised by the next call that waits."""
        self._chk(self.lib.shyhip_set_deferred_checks(self.h, 1 if on else 0), "set_deferred_checks")

    def last_istot(self):
        """Substeps the last step's scal3sh ran (a host-side value)."""
        d = C.c_int()
        self._chk(self.lib.shyhip_get_istot(self.h, C.byref(d)), "get_istot")
        return d.value

    def member_istot(self, nmem):
        """Substeps each ensemble member ran in the last step."""
        out = np.zeros(nmem, np.int32)
        self._chk(self.lib.shyhip_member_istot(self.h, nmem, _p(out)), "member_istot")
        return out

    def scalar_stability(self, rstol=1.0):
        """scalar_stability on the current state: (istot, rindex)."""
        n = C.c_int()
        r = C.c_double()
        self._chk(self.lib.shyhip_scalar_stability(self.h, C.c_float(rstol), C.byref(n), C.byref(r)),
                  "scalar_stability")
        return n.value, r.value

    def stats(self):
        a, b_, c, d = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        self._chk(self.lib.shyhip_stats(self.h, C.byref(a), C.byref(b_), C.byref(c)), "stats")
        self._chk(self.lib.shyhip_get_istot(self.h, C.byref(d)), "get_istot")
        r = C.c_long()
        self._chk(self.lib.shyhip_get_retries(self.h, C.byref(r)), "get_retries")
        return dict(nloop=a.value, solver_iters=b_.value, refinements=c.value, istot=d.value,
                    solver_retries=r.value)

    def set_timing(self, on=True):
        self._chk(self.lib.shyhip_set_timing(self.h, 1 if on else 0), "set_timing")

    def set_graphs(self, on=True):
        """Replay the transport substeps and solver iteration chunks from HIP
        graphs (on by default); off launches every kernel directly."""
        self._chk(self.lib.shyhip_set_graphs(self.h, 1 if on else 0), "set_graphs")

    def graph_stats(self):
        """(graphs instantiated, graph replays) since create."""
        b, r = C.c_long(), C.c_long()
        self._chk(self.lib.shyhip_graph_stats(self.h, C.byref(b), C.byref(r)), "graph_stats")
        return b.value, r.value

    def timing(self):
        """{kernel group: (total_ms, launches)} since set_timing(True)."""
        out = {}
        name = C.c_char_p()
        ms = C.c_double()
        cnt = C.c_long()
        for i in range(20):
            self._chk(self.lib.shyhip_get_timing(self.h, i, C.byref(name), C.byref(ms), C.byref(cnt)), "get_timing")
            out[name.value.decode()] = (ms.value, cnt.value)
        return out

    def volume(self, mode=1):
        v = C.c_double()
        self._chk(self.lib.shyhip_volume(self.h, mode, C.byref(v)), "volume")
        return v.value

    def tracer_mass(self, name):
        v = C.c_double()
        self._chk(self.lib.shyhip_tracer_mass(self.h, self.tracer_names.index(name), C.byref(v)), "tracer_mass")
        return v.value


def gpu_available() -> bool:
    try:
        lib = load_library()
    except (OSError, ShyhipError):
        return False
    n = C.c_int(0)
    lib.shyhip_device_count(C.byref(n))
    return n.value > 0
