"""ctypes driver for the CPU oracle (TEST INFRASTRUCTURE ONLY).

Loads liboracle_f32.so (reference precision policy: f32 state, f64 element
locals and solve) or liboracle_f64.so (all-FP64 oracle) built from the C
restatement in this directory, and runs the reference's initialisation and
time-step sequence on a synthetic basin:

  init   (shyfem_subs.f90:392-453, bnd_routines.f90:414-426):
         setweg(-1), setnod, init_z, setweg(0), setznv, set_area,
         make_new_depth + copy_depth, init_uvt (zero transports),
         init_uv -> hydro_vertical, copy_uvz, copy_depth
  step   (shyfem_subs.f90:577-607):
         set_dry, set_timestep, copy_uvz, copy_depth, sp111(2) boundary
         levels and wind, hydro, run_scalar (scal3sh per tracer)

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline use this.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
FLAG = np.float32(-9988765.0)

_P = C.c_void_p


class OcMesh(C.Structure):
    _fields_ = [("nkn", C.c_int), ("nel", C.c_int), ("nlvdi", C.c_int), ("nlv", C.c_int),
                ("nen3v", _P), ("ev", _P), ("ilhv", _P), ("ilhkv", _P), ("hm3v", _P), ("hev", _P),
                ("hlv", _P), ("hldv", _P), ("areakv", _P), ("ibnd", _P), ("ibtyp", _P), ("iextb", _P),
                ("fcorv", _P), ("czv", _P), ("wdifhv", _P), ("ntg", _P), ("cols", _P), ("ijp", _P),
                ("nnz", C.c_int), ("ncolor", C.c_int), ("color_ptr", _P), ("color_elems", _P),
                ("tvd_ie", _P), ("tvd_x", _P), ("tvd_y", _P), ("nbndo", C.c_int), ("bndo_k", _P),
                ("bndo_nx", _P), ("bndo_ny", _P), ("bndo_ptr", _P), ("bndo_nb", _P), ("bndo_w", _P)]


_STATE_FIELDS = ["znv", "zov", "zenv", "zeov", "utlnv", "vtlnv", "utlov", "vtlov", "unv", "vnv", "uov",
                 "vov", "ulnv", "vlnv", "wlnv", "wlov", "hdknv", "hdkov", "hdenv", "hdeov", "iwegv",
                 "iwetv", "inodv", "rfricv", "ddxv", "ddyv", "visv", "difv", "difhv", "fxv", "fyv",
                 "tauxnv", "tauynv", "ppv", "zeqv", "andgzv", "iuvfix", "rzv", "rqv", "mfluxv"]


def _state_class(real_t):
    class S(C.Structure):
        _fields_ = [(n, _P) for n in _STATE_FIELDS] + [("dtime", C.c_double), ("dtime_old", C.c_double),
                                                      ("robs", real_t), ("wsink", real_t), ("rload", real_t),
                                                      ("cobs", _P), ("rtauv", _P), ("wsinkv", _P), ("load", _P)]
    return S


_PARAM_FIELDS = [("dt", C.c_double), ("azpar", None), ("ampar", None), ("afpar", None), ("atpar", None),
                ("adpar", None), ("aapar", None), ("grav", None), ("rowass", None), ("vismol", None),
                ("ireib", C.c_int), ("uvmin", None), ("hzmin", None), ("hzoff", None), ("hzon", None),
                ("volmin", None), ("iclin", C.c_int), ("itpre", C.c_int), ("lfil", C.c_int),
                ("droptol", C.c_double), ("rtol", C.c_double), ("atol", C.c_double), ("itermax", C.c_int),
                ("rkpar", None), ("difmol", None), ("itvd", C.c_int), ("itvdv", C.c_int)]


class OcStats(C.Structure):
    _fields_ = [("nloop", C.c_int), ("solver_iters", C.c_int), ("solver_ierr", C.c_int),
                ("total_solver_iters", C.c_int), ("err", C.c_int), ("msg", C.c_char * 160)]


def _params_class(real_t):
    fields = [(n, (real_t if t is None else t)) for n, t in _PARAM_FIELDS]

    class P(C.Structure):
        _fields_ = fields
    return P


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load(precision: str = "f32"):
    path = os.path.join(HERE, f"liboracle_{precision}.so")
    if not os.path.exists(path):
        build()
    lib = C.CDLL(path)
    return lib


def ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


DEFAULT_PARAMS = dict(dt=300.0, azpar=0.5, ampar=0.5, afpar=0.5, atpar=1.0, adpar=1.0, aapar=0.0,
                      grav=9.81, rowass=1025.0, vismol=1.0e-6, ireib=5, uvmin=0.2, hzmin=0.01,
                      hzoff=0.05, hzon=0.10, volmin=1.0, iclin=0, itpre=2, lfil=5, droptol=0.01,
                      rtol=float(np.float32(1.0e-8)), atol=float(np.float32(1.0e-8)), itermax=1000,
                      rkpar=0.0, difmol=1.0e-6, itvd=0, itvdv=0)


class OracleModel:
    """The reference time step on one basin, in the oracle."""

    def __init__(self, basin, precision="f32", params=None, czv=2.5e-3, fcor=0.0, visv=0.0, difv=0.0,
                 zconst=0.0, tracers=None, idtype=2):
        self.basin = b = basin
        self.prec = precision
        self.real = np.float32 if precision == "f32" else np.float64
        R = self.real
        self.lib = lib = load(precision)
        self.real_c = C.c_float if precision == "f32" else C.c_double
        self.P = _params_class(self.real_c)
        p = dict(DEFAULT_PARAMS)
        if params:
            p.update(params)
        self.params_dict = p
        self.params = self.P(**p)
        nkn, nel, nlv = b.nkn, b.nel, b.nlv
        self.nlvdi = nlvdi = nlv
        # --- topology / geometry (init)
        self.nen3v = np.ascontiguousarray(b.nen3v, dtype=np.int32)
        self.xgv = np.ascontiguousarray(b.xgv, dtype=R)
        self.ygv = np.ascontiguousarray(b.ygv, dtype=R)
        self.ev = np.zeros((nel, 19), dtype=np.float64)
        lib.oc_set_ev(nkn, nel, ptr(self.nen3v), ptr(self.xgv), ptr(self.ygv), ptr(self.ev))
        maxnnz = 8 * nkn + 64
        self.ntg = np.zeros(nkn + 1, dtype=np.int32)
        self.cols = np.zeros(maxnnz, dtype=np.int32)
        self.ijp = np.zeros((nel, 3, 3), dtype=np.int32)  # [ie][j][i]
        nnz = lib.oc_node_list(nkn, nel, ptr(self.nen3v), ptr(self.ntg), ptr(self.cols), maxnnz, ptr(self.ijp))
        assert nnz > 0, "node list overflow"
        self.nnz = nnz
        self.cols = self.cols[:nnz].copy()
        self.color = np.zeros(nel, dtype=np.int32)
        self.color_ptr = np.zeros(33, dtype=np.int32)
        self.color_elems = np.zeros(nel, dtype=np.int32)
        nc = lib.oc_color(nkn, nel, ptr(self.nen3v), ptr(self.color), ptr(self.color_ptr), ptr(self.color_elems))
        assert nc > 0, f"colouring failed ({nc})"
        self.ncolor = nc
        self.color_ptr = self.color_ptr[:nc + 1].copy()
        self.hm3v = np.ascontiguousarray(b.hm3v, dtype=R)
        self.hev = np.ascontiguousarray(b.hev, dtype=R)
        self.hlv = np.ascontiguousarray(b.hlv, dtype=R)
        self.hldv = np.ascontiguousarray(b.hldv, dtype=R)
        self.ilhv = np.ascontiguousarray(b.ilhv, dtype=np.int32)
        self.ilhkv = np.ascontiguousarray(b.ilhkv, dtype=np.int32)
        self.ibnd = np.ascontiguousarray(b.ibnd, dtype=np.int32)
        self.ibtyp = np.ascontiguousarray(b.ibtyp, dtype=np.int32)
        self.iextb = np.ascontiguousarray(b.iextb, dtype=np.int32)
        self.fcorv = np.full(nel, fcor, dtype=R)
        self.czv = np.full(nel, czv, dtype=R)
        # horizontal diffusion weights (diffweight, the reference's default
        # idtype 2); they act only where rkpar * difhv != 0
        self.wdifhv = np.zeros((nel, 3, 3), dtype=R)
        if lib.oc_diffweight(nel, ptr(self.ev), idtype, ptr(self.wdifhv)) != 0:
            raise RuntimeError("diffweight: error in matrix (idtype %d)" % idtype)
        self.areakv = np.zeros((nkn, nlvdi), dtype=R)
        self.tvd_ie = np.full((nel, 3, 3), -1, dtype=np.int32)
        self.tvd_x = np.zeros((nel, 3, 3), dtype=R)
        self.tvd_y = np.zeros((nel, 3, 3), dtype=R)
        self.mesh = OcMesh(nkn=nkn, nel=nel, nlvdi=nlvdi, nlv=nlv, nen3v=ptr(self.nen3v), ev=ptr(self.ev),
                           ilhv=ptr(self.ilhv), ilhkv=ptr(self.ilhkv), hm3v=ptr(self.hm3v), hev=ptr(self.hev),
                           hlv=ptr(self.hlv), hldv=ptr(self.hldv), areakv=ptr(self.areakv),
                           ibnd=ptr(self.ibnd), ibtyp=ptr(self.ibtyp), iextb=ptr(self.iextb),
                           fcorv=ptr(self.fcorv), czv=ptr(self.czv), wdifhv=ptr(self.wdifhv),
                           ntg=ptr(self.ntg), cols=ptr(self.cols), ijp=ptr(self.ijp), nnz=nnz,
                           ncolor=nc, color_ptr=ptr(self.color_ptr), color_elems=ptr(self.color_elems),
                           tvd_ie=ptr(self.tvd_ie), tvd_x=ptr(self.tvd_x), tvd_y=ptr(self.tvd_y))
        lib.oc_set_area(C.byref(self.mesh), ptr(self.areakv))
        # scalar open-boundary tables (bndo_init)
        nb = len(b.bnd_nodes)
        maxnb = 64
        self.bndo_k = np.zeros(max(nb, 1), np.int32)
        self.bndo_nx = np.zeros(max(nb, 1), R)
        self.bndo_ny = np.zeros(max(nb, 1), R)
        self.bndo_ptr = np.zeros(nb + 1, np.int32)
        self.bndo_nb = np.zeros(max(nb, 1) * maxnb, np.int32)
        self.bndo_w = np.zeros(max(nb, 1) * maxnb, R)
        bnodes = np.ascontiguousarray(b.bnd_nodes, np.int32)
        if nb:
            rc = lib.oc_bndo_init(C.byref(self.mesh), ptr(self.xgv), ptr(self.ygv), nb, ptr(bnodes),
                                  ptr(self.bndo_k), ptr(self.bndo_nx), ptr(self.bndo_ny), ptr(self.bndo_ptr),
                                  ptr(self.bndo_nb), ptr(self.bndo_w), maxnb)
            assert rc == nb, "bndo_init failed"
        self.mesh.nbndo = nb
        for n in ("bndo_k", "bndo_nx", "bndo_ny", "bndo_ptr", "bndo_nb", "bndo_w"):
            setattr(self.mesh, n, ptr(getattr(self, n)))
        if p["itvd"] == 2:
            lib.oc_tvd_init(C.byref(self.mesh), ptr(self.xgv), ptr(self.ygv), ptr(self.tvd_ie), ptr(self.tvd_x),
                            ptr(self.tvd_y))
        # --- state
        z = lambda *shape, dt=R: np.zeros(shape, dtype=dt)  # noqa: E731
        self.a = a = {}
        a["znv"] = z(nkn); a["zov"] = z(nkn); a["zenv"] = z(nel, 3); a["zeov"] = z(nel, 3)
        for n in ["utlnv", "vtlnv", "utlov", "vtlov", "ulnv", "vlnv", "hdenv", "hdeov", "difhv"]:
            a[n] = z(nel, nlvdi)
        for n in ["unv", "vnv", "uov", "vov", "rfricv"]:
            a[n] = z(nel)
        for n in ["wlnv", "wlov"]:
            a[n] = z(nkn, nlvdi + 1)
        a["visv"] = np.full((nkn, nlvdi + 1), visv, dtype=R)
        a["difv"] = np.full((nkn, nlvdi + 1), difv, dtype=R)
        for n in ["hdknv", "hdkov", "mfluxv"]:
            a[n] = z(nkn, nlvdi)
        for n in ["iwegv", "iwetv"]:
            a[n] = z(nel, dt=np.int32)
        a["inodv"] = z(nkn, dt=np.int32)
        a["ddxv"] = z(nel, 2 * nlvdi, dt=np.float64)
        a["ddyv"] = z(nel, 2 * nlvdi, dt=np.float64)
        a["tauxnv"] = z(nkn); a["tauynv"] = z(nkn)
        a["rzv"] = np.full(nkn, FLAG, dtype=R)
        a["rqv"] = z(nkn)
        self.state = _state_class(self.real_c)(**{n: ptr(a[n]) if n in a else None for n in _STATE_FIELDS})
        self.tracer_terms = {}
        self.stats = OcStats()
        self.tracers = {}
        self.t = 0.0
        self._init(zconst)
        self.tracer_bc = {}
        if tracers:
            for name, arr in tracers.items():
                self.tracers[name] = np.ascontiguousarray(arr, dtype=R).copy()
                # boundary values: the initial values on the open-boundary nodes,
                # -999 (no value) elsewhere
                bc = np.full((b.nkn, nlvdi), -999.0, dtype=R)
                bc[b.bnd_nodes] = self.tracers[name][b.bnd_nodes]
                self.tracer_bc[name] = bc

    # ---------------------------------------------------------------- init
    def _init(self, zconst):
        lib, m, s, p = self.lib, C.byref(self.mesh), C.byref(self.state), C.byref(self.params)
        a, b = self.a, self.basin
        iw = C.c_int(0)
        self.state.dtime = 0.0
        self.state.dtime_old = -1.0          # setweg's saved initial value (wetdry.f90:186)
        lib.oc_setweg(m, s, p, -1, C.byref(iw))
        lib.oc_setnod(m, s)
        a["znv"][:] = zconst
        a["zenv"][:] = zconst
        lib.oc_setweg(m, s, p, 0, C.byref(iw))
        self._setznv()
        rc = lib.oc_setdepth(m, ptr(a["hdknv"]), ptr(a["hdenv"]), ptr(a["zenv"]))
        assert rc == 0, "setdepth failed at init"
        lib.oc_copy_depth(m, s)
        dz = np.zeros(b.nkn, dtype=self.real)
        lib.oc_hydro_vertical(m, s, p, ptr(dz))
        lib.oc_ttov(m, s)
        lib.oc_copy_uvz(m, s)
        lib.oc_copy_depth(m, s)
        self.state.dtime = 0.0
        self.state.dtime_old = 0.0

    def _setznv(self):
        """setznv: wetdry.f90:619-700"""
        a, b = self.a, self.basin
        R = self.real
        znv = np.full(b.nkn, FLAG, dtype=R)
        v1 = np.zeros(b.nkn, dtype=R)
        v2 = np.zeros(b.nkn, dtype=R)
        for ie in range(b.nel):
            if a["iwegv"][ie] == 0:
                for ii in range(3):
                    k = b.nen3v[ie, ii]
                    zz = a["zenv"][ie, ii]
                    if znv[k] != FLAG and znv[k] != zz:
                        zz = max(zz, znv[k])
                    znv[k] = zz
            else:
                area = R(4.0 * self.ev[ie, 9])
                for ii in range(3):
                    k = b.nen3v[ie, ii]
                    v1[k] += a["zenv"][ie, ii] * area
                    v2[k] += area
        dry = znv == FLAG
        znv[dry] = v1[dry] / v2[dry]
        a["znv"][:] = znv

    def set_initial_level(self, znv):
        """A non-uniform start level on an all-wet basin: znv and zenv, the
        new depths from it (setdepth), then copy_uvz/copy_depth so the old
        time level equals the new."""
        lib, m, s = self.lib, C.byref(self.mesh), C.byref(self.state)
        a, b = self.a, self.basin
        a["znv"][:] = np.asarray(znv, dtype=self.real)
        a["zenv"][:] = a["znv"][b.nen3v]
        rc = lib.oc_setdepth(m, ptr(a["hdknv"]), ptr(a["hdenv"]), ptr(a["zenv"]))
        assert rc == 0, "setdepth failed"
        lib.oc_copy_uvz(m, s)
        lib.oc_copy_depth(m, s)

    def transports(self):
        """hydro_transports alone (bottom friction + the element columns) on
        the current state: utlnv, vtlnv, ddxv, ddyv, rfricv."""
        self.lib.oc_hydro_transports(C.byref(self.mesh), C.byref(self.state), C.byref(self.params))

    def scalar_stability(self, dt=None, rstol=1.0, tracer=None):
        """scalar_stability on the current state: (istot, rindex), with the
        nudging and sinking of `tracer` (set_tracer_terms) if given."""
        if tracer is not None:
            self._apply_terms(tracer)
        ri = self.real_c(0)
        dt = self.params_dict["dt"] if dt is None else dt
        istot = self.lib.oc_scalar_stability(C.byref(self.mesh), C.byref(self.state), C.byref(self.params),
                                             self.real_c(dt), self.real_c(rstol), C.byref(ri))
        return istot, ri.value

    def tvd_grad(self, cv):
        """tvd_grad_3d of cv ([nkn, nlv]): the node gradients (gx, gy)."""
        cv = np.ascontiguousarray(cv, dtype=self.real)
        gx, gy = np.zeros_like(cv), np.zeros_like(cv)
        self.lib.oc_tvd_grad_3d(C.byref(self.mesh), ptr(cv), ptr(gx), ptr(gy))
        return gx, gy

    def hflux(self, cv, ie, l, dt, grads=None):
        """The horizontal advective flux of element ie (0-based) at layer l
        (1-based) on cv, as conz3d computes it for a substep of length dt:
        fl[3].  grads: (gx, gy) of cv for itvd = 1 (computed if None)."""
        cv = np.ascontiguousarray(cv, dtype=self.real)
        gx = gy = None
        if self.params_dict["itvd"] == 1:
            gx, gy = grads if grads is not None else self.tvd_grad(cv)
        fl = np.zeros(3)
        self.lib.oc_conz_hflux(C.byref(self.mesh), C.byref(self.state), C.byref(self.params), ptr(cv), ptr(gx),
                               ptr(gy), int(ie), int(l), C.c_double(dt), ptr(fl))
        return fl

    # ---------------------------------------------------------------- step
    def set_boundary(self, level):
        self.a["rzv"][:] = FLAG
        self.a["rzv"][self.basin.bnd_nodes] = self.real(level)

    def set_wind(self, taux, tauy):
        self.a["tauxnv"][:] = taux
        self.a["tauynv"][:] = tauy

    def set_forcing(self, name, arr):
        """Optional hydro forcing the default runs leave out: atmospheric
        pressure ppv (Pa), equilibrium tide zeqv (m), level nudging andgzv
        (m/s) at the nodes, or the fixed-transport flags iuvfix (elements)."""
        dt = np.int32 if name == "iuvfix" else self.real
        self.a[name] = np.ascontiguousarray(arr, dtype=dt).copy()
        setattr(self.state, name, ptr(self.a[name]))

    def set_uvfix(self, iuvfix, utfix=None, vtfix=None):
        """Elements with fixed transports; utfix/vtfix ([nel, nlv]): the
        transports they take each step, written into utlnv/vtlnv after
        copy_uvz as bclfix does (bcvel.f90:332-338) -- None: they keep the
        previous step's."""
        self.set_forcing("iuvfix", iuvfix)
        self.uvfix = None if utfix is None else (np.asarray(utfix), np.asarray(vtfix))

    def set_tracer_terms(self, name, robs=0.0, cobs=None, rtauv=None, wsink=0.0, wsinkv=None, rload=0.0,
                         load=None):
        """conz3d_omp's terms of one scalar: nudging robs * rtauv towards
        cobs ([nkn, nlv]), sinking wsink * wsinkv ([nkn, nlv+1]), loads
        rload * load ([nkn, nlv], updated by deposition)."""
        R = self.real
        arr = lambda x: None if x is None else np.ascontiguousarray(x, dtype=R).copy()  # noqa: E731
        self.tracer_terms[name] = dict(robs=robs, wsink=wsink, rload=rload, cobs=arr(cobs), rtauv=arr(rtauv),
                                       wsinkv=arr(wsinkv), load=arr(load))

    def _apply_terms(self, name):
        t = self.tracer_terms.get(name, {})
        st = self.state
        st.robs, st.wsink, st.rload = t.get("robs", 0.0), t.get("wsink", 0.0), t.get("rload", 0.0)
        for n in ("cobs", "rtauv", "wsinkv", "load"):
            setattr(st, n, ptr(t.get(n)))

    def set_baroclinic(self, temp="temp", salt="salt"):
        """ibarcl = 1: barocl(0)'s start (rhov = 0, rhoset three times,
        baroclinic.f90:398-404); then every step adds the baroclinic term to
        the explicit momentum terms before hydro (set_explicit,
        explicit.f90:193) and updates rhov after the transport
        (rhoset_shell(1), baroclinic.f90:542-548)."""
        R = self.real
        self.bcl = (temp, salt)
        self.rhov = np.zeros((self.basin.nkn, self.nlvdi), dtype=R)
        self.a["fxv"] = np.zeros((self.basin.nel, self.nlvdi), dtype=R)
        self.a["fyv"] = np.zeros((self.basin.nel, self.nlvdi), dtype=R)
        self.state.fxv = ptr(self.a["fxv"])
        self.state.fyv = ptr(self.a["fyv"])
        for _ in range(3):
            self.rhoset()

    def rhoset(self):
        t, sname = self.bcl
        self.lib.oc_rhoset(C.byref(self.mesh), C.byref(self.params), ptr(self.tracers[t]),
                           ptr(self.tracers[sname]), ptr(self.rhov))

    def step(self, bnd_level, istot=1):
        lib, m, s, p = self.lib, C.byref(self.mesh), C.byref(self.state), C.byref(self.params)
        lib.oc_set_dry(m, s, p)
        self.t += self.params_dict["dt"]
        self.state.dtime = self.t
        lib.oc_copy_uvz(m, s)
        lib.oc_copy_depth(m, s)
        if getattr(self, "uvfix", None) is not None:
            # the element's layers only, as bclfix writes them
            sel = (self.a["iuvfix"] != 0)[:, None] & (np.arange(self.nlvdi)[None, :] < self.ilhv[:, None])
            self.a["utlnv"][sel] = self.uvfix[0][sel]
            self.a["vtlnv"][sel] = self.uvfix[1][sel]
        self.set_boundary(bnd_level)
        if getattr(self, "bcl", None):
            lib.oc_barocl(m, p, ptr(self.rhov), ptr(self.a["fxv"]), ptr(self.a["fyv"]))
        rc = lib.oc_hydro(m, s, p, C.byref(self.stats))
        if rc != 0:
            raise RuntimeError(f"oracle hydro failed with code {rc}")
        for name, arr in self.tracers.items():
            self._apply_terms(name)
            rc = lib.oc_scal3sh(m, s, p, ptr(arr), ptr(self.tracer_bc[name]), istot, 0 if name == "temp" else 1)
            if rc != 0:
                raise RuntimeError(f"oracle scal3sh({name}) failed with code {rc}")
        if getattr(self, "bcl", None):
            self.rhoset()
        return self.stats

    def volume(self, mode=1):
        self.lib.oc_masscont.restype = C.c_double
        return self.lib.oc_masscont(C.byref(self.mesh), C.byref(self.state), mode)

    def tracer_mass(self, name, mode=1):
        self.lib.oc_massconc.restype = C.c_double
        return self.lib.oc_massconc(C.byref(self.mesh), C.byref(self.state), ptr(self.tracers[name]), mode)


# ------------------------------------------------------------------ forcing
def intp_neville(xa, ya, x):
    """rd_intp_neville over the points in the given order (f32 data)."""
    lib = load("f32")
    lib.oc_intp_neville.restype = C.c_double
    xa = np.ascontiguousarray(xa, np.float64)
    ya = np.ascontiguousarray(ya, np.float32)
    return lib.oc_intp_neville(len(xa), ptr(xa), ptr(ya), C.c_double(x))


def iff_first_record(time, nintp, dtime0):
    lib = load("f32")
    t = np.ascontiguousarray(time, np.float64)
    return lib.oc_iff_first_record(len(t), ptr(t), nintp, C.c_double(dtime0))


def iff_series(time, vals, nintp, dtime0, times, comp=0, vdefault=0.0):
    """Values of component comp of a record series at each of `times`, as a
    run with one interpolation per step gets them.  vals: (nrec, nval) or
    (nrec, ncomp, nval).  Returns (out (nt, nval) f32, newest record per
    time), or None when a time is outside the records."""
    lib = load("f32")
    t = np.ascontiguousarray(time, np.float64)
    v = np.asarray(vals, np.float32)
    if v.ndim == 2:
        v = v[:, None, :]
    v = np.ascontiguousarray(v)
    nrec, ncomp, nval = v.shape
    ts = np.ascontiguousarray(times, np.float64)
    out = np.zeros((len(ts), nval), np.float32)
    newest = np.zeros(len(ts), np.int32)
    rc = lib.oc_iff_series(nrec, ptr(t), ptr(v), ncomp, comp, nval, nintp, C.c_double(dtime0),
                           C.c_float(vdefault), len(ts), ptr(ts), ptr(out), ptr(newest))
    return None if rc else (out, newest)


def convert_wind(wx, wy, iwtype=1, itdrag=0, dragco=2.5e-3, roluft=1.225, rowass=1025.0, wslim=-1.0, wsmax=50.0):
    """meteo_convert_wind_data: returns (tx, ty, ws, wx, wy, rc) (f32 arrays;
    wx/wy after the limit; rc = 1 when a speed exceeds wsmax)."""
    lib = load("f32")
    x = np.array(wx, np.float32, copy=True).ravel()
    y = np.array(wy, np.float32, copy=True).ravel()
    n = len(x)
    tx, ty, ws = (np.zeros(n, np.float32) for _ in range(3))
    f = C.c_float
    rc = lib.oc_convert_wind(n, ptr(x), ptr(y), iwtype, itdrag, f(dragco), f(roluft), f(rowass), f(wslim), f(wsmax),
                             ptr(tx), ptr(ty), ptr(ws))
    return tx, ty, ws, x, y, rc


def set_mass_flux(ilhkv, hdknv, areakv, nodes, levmin, levmax, rflux, rqdsv=None, nlv=None, precision="f32"):
    """set_mass_flux: hdknv/areakv (nkn, nlvdi) node-major; sources with
    0-based nodes.  Returns (mfluxv (nkn, nlvdi), rqv) or None on the
    error stop."""
    lib = load(precision)
    dt = np.float32 if precision == "f32" else np.float64
    hd = np.ascontiguousarray(hdknv, dtype=dt)
    ar = np.ascontiguousarray(areakv, dtype=dt)
    nkn, nlvdi = hd.shape
    il = np.ascontiguousarray(ilhkv, dtype=np.int32)
    nd, lo, hi = (np.ascontiguousarray(a, dtype=np.int32) for a in (nodes, levmin, levmax))
    q = np.ascontiguousarray(rflux, dtype=np.float32)
    rq = None if rqdsv is None else np.ascontiguousarray(rqdsv, dtype=np.float32)
    mf = np.zeros((nkn, nlvdi), dt)
    rqv = np.zeros(nkn, dt)
    rc = lib.oc_set_mass_flux(nkn, nlvdi, nlvdi if nlv is None else nlv, ptr(il), ptr(hd), ptr(ar), len(nd), ptr(nd),
                              ptr(lo), ptr(hi), ptr(q), ptr(rq), ptr(mf), ptr(rqv))
    return None if rc else (mf, rqv)


# ------------------------------------------------------------------ baroclinic
def sigma_stp(s, t, p, precision="f32"):
    """sigma = rho - 1000 of the UNESCO 1980 equation of state (p in bar)."""
    lib = load(precision)
    T = C.c_float if precision == "f32" else C.c_double
    f = lib.oc_sigma_stp
    f.restype = T
    f.argtypes = [T, T, T]
    return f(s, t, p)
