// shyhip_hydro.hip -- gfx950 kernels of the semi-implicit hydro step.
//
// Every kernel is one work-item per element column or per node column, with
// layer-major arrays so that a wavefront reads 64 consecutive elements of one
// layer.  Element -> node accumulations are gathers over the node's incident
// elements in ascending element order: atomic-free, deterministic, and the
// same summation order as the reference's sequential element loops.
// Arithmetic order and operand types follow the reference (`real` locals are
// the storage type S, double precision locals are double), and the build
// disables FP contraction, so per-kernel results reproduce the CPU path.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include <cmath>

#include "shyhip_internal.h"

namespace shy {

template <typename S>
struct HydroArgs {
    int nkn, nel, nlvdi, nlv;
    double dt;
    float az, am, af, at, grav, rowass, vismol, hzmin, hzoff, hzon, volmin, uvmin;
    int ireib, iclin;
    const int *nen3v;
    const double *geo, *ang;
    const int *ilhv, *ilhkv;
    const S *hm3v, *hev, *hlv, *hldv, *areakv, *fcorv, *czv;
    const int *ibnd, *ibtyp, *zbound;
    const int *nk_ptr, *nk_el;
    const int *inc_cnt, *inc_el;     // incidence slots (element order)
    const int *rowptr, *ijp;
    S *zn, *zo, *zen, *zeo;
    S *utn, *vtn, *uto, *vto, *uln, *vln;
    S *un, *vn, *uo, *vo;
    S *wln;
    S *hdkn, *hdko, *hden, *hdeo;
    int *iweg, *iwet, *inod;
    S *rfric, *ddx, *ddy, *visv;
    S *taux, *tauy, *rzv, *rqv, *mflux;
    const S *rhov;      // density anomaly (baroclinic runs), null otherwise
    // optional forcing (null when never pushed): atmospheric pressure ppv
    // (Pa) and equilibrium tide zeqv (m) at the nodes, the level nudging
    // term andgzv (m/s), elements with fixed transports (iuvfix) and the
    // transports they take (utfix/vtfix; null: the previous step's)
    const S *ppv, *zeqv, *andgz;
    const int *iuvfix;
    const S *utfix, *vtfix;
    double *zel, *aval, *rhs, *x;
    S *zrow;
    // inputs of the zeta solve written by the row assembly (null: not fed)
    int *iweg_snap;     // copy of the final wet/dry mask for the transport (may be null)
    S *sv_ell;          // ELL slot values
    float *sv_a32;      // FP32 copy of the values (FP32 path)
    float *sv_zprev;    // FP32 path: level of the previous step (extrapolated guess)
    int sv_have_prev;
    int *sv_ctrl;
    int *err;
};

template <typename S>
static HydroArgs<S> make_args(Model &M) {
    HydroArgs<S> a{};
    a.nkn = M.nkn; a.nel = M.nel; a.nlvdi = M.nlvdi; a.nlv = M.nlv;
    a.dt = M.p.dt;
    a.az = M.p.azpar; a.am = M.p.ampar; a.af = M.p.afpar; a.at = M.p.atpar;
    a.grav = M.p.grav; a.rowass = M.p.rowass; a.vismol = M.p.vismol;
    a.hzmin = M.p.hzmin; a.hzoff = M.p.hzoff; a.hzon = M.p.hzon; a.volmin = M.p.volmin; a.uvmin = M.p.uvmin;
    a.ireib = M.p.ireib; a.iclin = M.p.iclin;
    a.nen3v = M.m.nen3v; a.geo = M.m.geo; a.ang = M.m.ang;
    a.ilhv = M.m.ilhv; a.ilhkv = M.m.ilhkv;
    a.hm3v = (const S *)M.m.hm3v; a.hev = (const S *)M.m.hev; a.hlv = (const S *)M.m.hlv;
    a.hldv = (const S *)M.m.hldv; a.areakv = (const S *)M.m.areakv;
    a.fcorv = (const S *)M.m.fcorv; a.czv = (const S *)M.m.czv;
    a.ibnd = M.m.ibnd; a.ibtyp = M.m.ibtyp; a.zbound = M.m.zbound;
    a.nk_ptr = M.m.nk_ptr; a.nk_el = M.m.nk_el; a.rowptr = M.m.rowptr; a.ijp = M.m.ijp;
    a.inc_cnt = M.m.inc_cnt; a.inc_el = M.m.inc_el; a.zrow = (S *)M.s.zrow;
    a.zn = (S *)M.s.zn; a.zo = (S *)M.s.zo; a.zen = (S *)M.s.zen; a.zeo = (S *)M.s.zeo;
    a.utn = (S *)M.s.utn; a.vtn = (S *)M.s.vtn; a.uto = (S *)M.s.uto; a.vto = (S *)M.s.vto;
    a.uln = (S *)M.s.uln; a.vln = (S *)M.s.vln;
    a.un = (S *)M.s.un; a.vn = (S *)M.s.vn; a.uo = (S *)M.s.uo; a.vo = (S *)M.s.vo;
    a.wln = (S *)M.s.wln;
    a.hdkn = (S *)M.s.hdkn; a.hdko = (S *)M.s.hdko; a.hden = (S *)M.s.hden; a.hdeo = (S *)M.s.hdeo;
    a.iweg = M.s.iweg; a.iwet = M.s.iwet; a.inod = M.s.inod;
    a.rfric = (S *)M.s.rfric; a.ddx = (S *)M.s.ddx; a.ddy = (S *)M.s.ddy; a.visv = (S *)M.s.visv;
    a.taux = (S *)M.s.taux; a.tauy = (S *)M.s.tauy; a.rzv = (S *)M.s.rzv; a.rqv = (S *)M.s.rqv;
    a.mflux = M.has_mflux ? (S *)M.s.mflux : nullptr;
    a.rhov = M.bcl_on ? (const S *)M.s.rhov : nullptr;
    a.ppv = (const S *)M.s.ppv; a.zeqv = (const S *)M.s.zeqv; a.andgz = (const S *)M.s.andgz;
    a.iuvfix = M.s.iuvfix;
    a.utfix = (const S *)M.s.utfix; a.vtfix = (const S *)M.s.vtfix;
    a.zel = M.s.zel; a.aval = M.s.aval; a.rhs = M.s.rhs; a.x = M.s.x;
    a.err = M.d_err;
    return a;
}

static inline int nblocks(int n) { return (n + kThreads - 1) / kThreads; }

// ------------------------------------------------------------------ friction
// bottom_friction, chezy.f90:239-398: ireib 0-7 and 10 (8 and 9 take z0 from
// the sediment and fluid-mud modules and are rejected by shyhip_set_params)
template <typename S>
__device__ inline S friction(const HydroArgs<S> &a, int ie, int lmax) {
    using R = S;
    const int nel = a.nel;
    const R uso = a.uto[(size_t)(lmax - 1) * nel + ie], vso = a.vto[(size_t)(lmax - 1) * nel + ie];
    const R uv = sqrt(uso * uso + vso * vso);
    R hzg = a.hdeo[(size_t)(lmax - 1) * nel + ie];
    if (hzg < (R)a.hzoff) hzg = (R)a.hzoff;
    const R rfric = a.czv[ie];
    const R grav = (R)a.grav;
    R rr = 0, rcd = 0;
    switch (a.ireib) {
    case 1: rr = rfric; break;
    case 2: rcd = grav / ((rfric * rfric) * (R)pow((double)hzg, (double)((R)1 / (R)3))); rr = rcd * uv / (hzg * hzg); break;
    case 3: rcd = grav / (rfric * rfric); rr = rcd * uv / (hzg * hzg); break;
    case 4: rr = rfric / hzg; break;
    case 5: rcd = rfric; rr = rcd * uv / (hzg * hzg); break;
    case 6: {
        // rfric is the roughness length z0: cd = (kappa / log((z0 + h/2) / z0))^2,
        // kappa = 0.4 (cdf, chezy.f90:402-419); log in double, rounded
        const R cds = (R)0.4 / (R)log((double)((rfric + (R)0.5 * hzg) / rfric));
        rcd = cds * cds;
        rr = rcd * uv / (hzg * hzg);
        break;
    }
    case 7:
        rcd = (rfric >= (R)1) ? grav / ((rfric * rfric) * (R)pow((double)hzg, (double)((R)1 / (R)3))) : rfric;
        rr = rcd * uv / (hzg * hzg);
        break;
    case 10:
        if (uv > (R)a.uvmin * hzg) { rcd = rfric; rr = rcd * uv / (hzg * hzg); }
        else rr = (rfric * (R)a.uvmin) / hzg;
        break;
    default: rr = 0; break;
    }
    return rr;
}

// ------------------------------------------------------------------ hydro_intern
// hydro_transports / hydro_intern, hydrodynamic.f90:758-1416.  One work-item
// per element column.  Without Coriolis (gamma = 0) the 2L x 2L pentadiagonal
// system decouples into two identical tridiagonal systems in u and v; the
// tridiagonal elimination below performs exactly the operations that
// penta_fact/penta_solve (penta.f90:104-321) perform on the decoupled bands.
// The coupled case runs the pentadiagonal elimination itself.
template <typename S, int NL, bool CORI>
__global__ __launch_bounds__(kThreads) void k_hydro_intern(HydroArgs<S> a) {
    using R = S;
    const int ie = blockIdx.x * blockDim.x + threadIdx.x;
    if (ie >= a.nel) return;
    const int nel = a.nel, nkn = a.nkn;
    const R dt = (R)a.dt, af = (R)a.af, at = (R)a.at, grav = (R)a.grav, rowass = (R)a.rowass;
    const int ilevel = a.ilhv[ie];
    const int k1 = a.nen3v[3 * ie], k2 = a.nen3v[3 * ie + 1], k3 = a.nen3v[3 * ie + 2];
    const int kn[3] = {k1, k2, k3};
    double b[3], c[3];
    double bz = 0., cz = 0., zmm = 0., taux = 0., tauy = 0., bpres = 0., cpres = 0.;
    for (int ii = 0; ii < 3; ii++) {
        b[ii] = a.geo[(size_t)ii * nel + ie];
        c[ii] = a.geo[(size_t)(3 + ii) * nel + ie];
        const R ze = a.zeo[(size_t)ii * nel + ie];
        // the level relative to the equilibrium tide, and the atmospheric
        // pressure gradient (hydrodynamic.f90:1025-1033)
        const double zz = (double)(a.zeqv ? (R)(ze - a.zeqv[kn[ii]]) : ze);
        zmm += (double)ze;
        bz += zz * b[ii];
        cz += zz * c[ii];
        if (a.ppv) {
            bpres += (double)a.ppv[kn[ii]] * b[ii];
            cpres += (double)a.ppv[kn[ii]] * c[ii];
        }
        taux += (double)a.taux[kn[ii]];
        tauy += (double)a.tauy[kn[ii]];
    }
    zmm = zmm * kDrittl;
    taux = taux * kDrittl;
    tauy = tauy * kDrittl;
    const R gammat = a.fcorv[ie];
    const R gamma = af * dt * gammat;
    const R rfric = friction(a, ie, ilevel);
    a.rfric[ie] = rfric;

    // layer thickness, inverse thickness, element viscosity
    // The column's loads are issued at once, unconditionally: rows below the
    // bottom layer read the bottom layer again (a valid address) and are
    // replaced by zeros or not used.
    R hact[NL + 2], rhact[NL + 2], alev[NL + 1];
    R vsum[NL + 1];
    R uoc[NL], voc[NL];
#pragma unroll
    for (int l = 1; l <= NL; l++) {
        const size_t o = (size_t)(min(l, ilevel) - 1) * nel + ie;
        hact[l] = a.hdeo[o];
        uoc[l - 1] = a.uto[o];
        voc[l - 1] = a.vto[o];
    }
#pragma unroll
    for (int l = 0; l <= NL; l++) {
        const size_t o = (size_t)min(l, ilevel) * nkn;
        vsum[l] = a.visv[o + k1] + a.visv[o + k2] + a.visv[o + k3];
    }
    hact[0] = 0;
    hact[NL + 1] = 0;
#pragma unroll
    for (int l = 1; l <= NL; l++)
        if (l > ilevel) hact[l] = 0;
    if (a.iclin) hact[1] = (R)(hact[1] - zmm);
#pragma unroll
    for (int l = 0; l <= NL + 1; l++) rhact[l] = (hact[l] <= (R)0) ? (R)0 : (R)1 / hact[l];
#pragma unroll
    for (int l = 0; l <= NL; l++) {
        if (l <= ilevel) {
            double vis = (double)a.vismol;
            vis = vis + (double)(vsum[l] / (R)3);
            alev[l] = (R)vis;
        } else {
            alev[l] = 0;
        }
    }

    // explicit terms: the baroclinic pressure gradient (set_explicit ->
    // set_barocl_new_interface, explicit.f90:193,1226-1530, z-layers) when
    // the density anomaly is carried, else zero.  Per layer the anomaly
    // gradient from the element's P1 gradients, averaged onto the interface
    // above (the first layer uses its own), integrated down over the
    // interface thicknesses, times g/rho0 and the nominal layer thickness;
    // rounded to the reference's real as fxv/fyv are.
    R xex[NL], yex[NL];
    if (a.rhov) {
        const double raux = (double)(grav / rowass);
        double px = 0., py = 0., brup = 0., crup = 0., hhup = 0.;
#pragma unroll
        for (int l = 1; l <= NL; l++) {
            const size_t o = (size_t)(min(l, ilevel) - 1) * nkn;
            const R r[3] = {a.rhov[o + k1], a.rhov[o + k2], a.rhov[o + k3]};
            const double hlayer = (double)a.hldv[l - 1];
            const double hh = 0.5 * hlayer;
            const double hint = hh + hhup;
            double br = 0., cr = 0.;
#pragma unroll
            for (int ii = 0; ii < 3; ii++) {
                br = br + b[ii] * (double)r[ii];
                cr = cr + c[ii] * (double)r[ii];
            }
            const double brint = l == 1 ? br : 0.5 * (br + brup);
            const double crint = l == 1 ? cr : 0.5 * (cr + crup);
            brup = br;
            crup = cr;
            hhup = hh;
            px = px + hint * brint;
            py = py + hint * crint;
            xex[l - 1] = (R)(0.0 + raux * hlayer * px);
            yex[l - 1] = (R)(0.0 + raux * hlayer * py);
        }
    } else {
#pragma unroll
        for (int l = 0; l < NL; l++) xex[l] = yex[l] = 0;
    }

    double sxu[NL], sxv[NL], sh[NL];      // solutions: -F^x, -F^y and H rows
    if constexpr (!CORI) {
        double xd[NL], up[NL], gx[NL], gy[NL], gh[NL];
#pragma unroll
        for (int l = 1; l <= NL; l++) {
            if (l > ilevel) continue;
            const bool bfirst = (l == 1), blast = (l == ilevel);
            // layers min(l+1, ilevel) and max(l-1, 1), from registers
            const bool below = l < NL && l + 1 <= ilevel;
            const R uui = uoc[l - 1], uuip = below ? uoc[l < NL ? l : l - 1] : uoc[l - 1];
            const R uuim = l > 1 ? uoc[l > 1 ? l - 2 : 0] : uoc[0];
            const R vvi = voc[l - 1], vvip = below ? voc[l < NL ? l : l - 1] : voc[l - 1];
            const R vvim = l > 1 ? voc[l > 1 ? l - 2 : 0] : voc[0];
            const R hhi = hact[l], hhip = hact[l + 1], hhim = hact[l - 1];
            R rhp = 0, rhm = 0;
            if (hhi > (R)0) {
                if (hhip > (R)0) rhp = (R)2.0 * alev[l] / (hhi + hhip);
                if (hhim > (R)0) rhm = (R)2.0 * alev[l - 1] / (hhi + hhim);
            }
            const R aux = dt * at;
            R aa = aux * rhact[l] * (rhm + rhp);
            R aat = (R)1 * rhact[l] * (rhm + rhp);
            const R bb = aux * rhact[l + 1] * rhp;
            const R bbt = (R)1 * rhact[l + 1] * rhp;
            const R cc = aux * rhact[l - 1] * rhm;
            const R cct = (R)1 * rhact[l - 1] * rhm;
            double ppx = 0., ppy = 0.;
            if (bfirst) { ppx = ppx - taux; ppy = ppy - tauy; }
            if (blast) { aa = aa + dt * rfric; aat = aat + rfric; }
            const R gravx = (R)(1.0 * grav * hhi * bz);
            const R gravy = (R)(1.0 * grav * hhi * cz);
            const R presx = (R)bpres, presy = (R)cpres, xexpl = xex[l - 1], yexpl = yex[l - 1];
            const R ppx_aux = aat * uui - bbt * uuip - cct * uuim - gammat * vvi + gravx + (hhi / rowass) * presx + xexpl;
            const R ppy_aux = aat * vvi - bbt * vvip - cct * vvim + gammat * uui + gravy + (hhi / rowass) * presy + yexpl;
            ppx = ppx + (double)ppx_aux;
            ppy = ppy + (double)ppy_aux;
            const double d = (double)((R)1 + aa);
            up[l - 1] = blast ? 0.0 : (double)(-bb);
            const double lo = bfirst ? 0.0 : (double)(-cc);
            if (l == 1) {
                xd[0] = d;
                gx[0] = ppx;
                gy[0] = ppy;
                gh[0] = (double)hhi;
            } else {
                const double w = lo / xd[l - 2];
                xd[l - 1] = d - w * up[l - 2];
                gx[l - 1] = ppx - w * gx[l - 2];
                gy[l - 1] = ppy - w * gy[l - 2];
                gh[l - 1] = (double)hhi - w * gh[l - 2];
            }
        }
        double su = 0., sv = 0., shh = 0.;
#pragma unroll
        for (int l = NL; l >= 1; l--) {
            if (l > ilevel) continue;
            if (l == ilevel) {
                su = gx[l - 1] / xd[l - 1];
                sv = gy[l - 1] / xd[l - 1];
                shh = gh[l - 1] / xd[l - 1];
            } else {
                su = (gx[l - 1] - up[l - 1] * su) / xd[l - 1];
                sv = (gy[l - 1] - up[l - 1] * sv) / xd[l - 1];
                shh = (gh[l - 1] - up[l - 1] * shh) / xd[l - 1];
            }
            sxu[l - 1] = su;
            sxv[l - 1] = sv;
            sh[l - 1] = shh;
        }
    } else {
        // coupled u/v columns: pentadiagonal elimination of penta.f90
        constexpr int N2 = 2 * NL;
        double vm2[N2], vm1[N2], v0[N2], vp1[N2], vp2[N2], r0[N2], r1[N2], r2[N2];
#pragma unroll
        for (int i = 0; i < N2; i++) { vm2[i] = vm1[i] = v0[i] = vp1[i] = vp2[i] = 0.; r0[i] = r1[i] = r2[i] = 0.; }
        const int n = 2 * ilevel;
        for (int l = 1; l <= ilevel; l++) {
            const bool bfirst = (l == 1), blast = (l == ilevel);
            const int lp = (l + 1 < ilevel) ? l + 1 : ilevel;
            const int lm = (l - 1 > 1) ? l - 1 : 1;
            const R uui = a.uto[(size_t)(l - 1) * nel + ie], uuip = a.uto[(size_t)(lp - 1) * nel + ie];
            const R uuim = a.uto[(size_t)(lm - 1) * nel + ie];
            const R vvi = a.vto[(size_t)(l - 1) * nel + ie], vvip = a.vto[(size_t)(lp - 1) * nel + ie];
            const R vvim = a.vto[(size_t)(lm - 1) * nel + ie];
            const R hhi = hact[l], hhip = hact[l + 1], hhim = hact[l - 1];
            R rhp = 0, rhm = 0;
            if (hhi > (R)0) {
                if (hhip > (R)0) rhp = (R)2.0 * alev[l] / (hhi + hhip);
                if (hhim > (R)0) rhm = (R)2.0 * alev[l - 1] / (hhi + hhim);
            }
            const R aux = dt * at;
            R aa = aux * rhact[l] * (rhm + rhp);
            R aat = (R)1 * rhact[l] * (rhm + rhp);
            const R bb = aux * rhact[l + 1] * rhp;
            const R bbt = (R)1 * rhact[l + 1] * rhp;
            const R cc = aux * rhact[l - 1] * rhm;
            const R cct = (R)1 * rhact[l - 1] * rhm;
            double ppx = 0., ppy = 0.;
            if (bfirst) { ppx = ppx - taux; ppy = ppy - tauy; }
            if (blast) { aa = aa + dt * rfric; aat = aat + rfric; }
            const R gravx = (R)(1.0 * grav * hhi * bz);
            const R gravy = (R)(1.0 * grav * hhi * cz);
            const R presx = (R)bpres, presy = (R)cpres, xexpl = xex[l - 1], yexpl = yex[l - 1];
            const R ppx_aux = aat * uui - bbt * uuip - cct * uuim - gammat * vvi + gravx + (hhi / rowass) * presx + xexpl;
            const R ppy_aux = aat * vvi - bbt * vvip - cct * vvim + gammat * uui + gravy + (hhi / rowass) * presy + yexpl;
            ppx = ppx + (double)ppx_aux;
            ppy = ppy + (double)ppy_aux;
            const int ju = 2 * l - 2, jv = 2 * l - 1;
            v0[ju] = (double)((R)1 + aa);
            v0[jv] = (double)((R)1 + aa);
            vm1[jv] = (double)gamma;
            vp1[ju] = (double)(-gamma);
            if (!blast) { vp2[ju] = -bb; vp2[jv] = -bb; }
            if (!bfirst) { vm2[ju] = -cc; vm2[jv] = -cc; }
            r0[ju] = ppx; r0[jv] = ppy;
            r1[ju] = (double)hhi; r2[jv] = (double)hhi;
        }
        double xx[N2], yy[N2], zz[N2 + 1];
        int bad = 0;
        zz[0] = 0.;
        for (int i = 0; i < n; i++) {
            double t = v0[i];
            if (i >= 1) t = t - yy[i - 1] * zz[i];
            if (i >= 2) t = t - vm2[i] * vp2[i - 2];
            if (t == 0.) bad = 1;
            xx[i] = t;
            yy[i] = (i < n - 1) ? vp1[i] - ((i >= 1) ? zz[i] * vp2[i - 1] : 0.) : 0.;
            if (i == 0) yy[0] = vp1[0];
            if (i + 1 < n) zz[i + 1] = (vm1[i + 1] - ((i >= 1) ? vm2[i + 1] * yy[i - 1] : 0.)) / t;
            if (i + 2 < n) vm2[i + 2] = vm2[i + 2] / t;
        }
        if (bad) atomicCAS(a.err, 0, ERR_PENTA);
        double g0[N2], g1[N2], g2[N2];
        for (int i = 0; i < n; i++) {
            g0[i] = r0[i]; g1[i] = r1[i]; g2[i] = r2[i];
            if (i >= 1) { g0[i] -= zz[i] * g0[i - 1]; g1[i] -= zz[i] * g1[i - 1]; g2[i] -= zz[i] * g2[i - 1]; }
            if (i >= 2) { g0[i] -= vm2[i] * g0[i - 2]; g1[i] -= vm2[i] * g1[i - 2]; g2[i] -= vm2[i] * g2[i - 2]; }
        }
        double s0[N2], s1[N2], s2[N2];
        for (int i = n - 1; i >= 0; i--) {
            double t0 = g0[i], t1 = g1[i], t2 = g2[i];
            if (i + 1 < n) { t0 -= yy[i] * s0[i + 1]; t1 -= yy[i] * s1[i + 1]; t2 -= yy[i] * s2[i + 1]; }
            if (i + 2 < n) { t0 -= vp2[i] * s0[i + 2]; t1 -= vp2[i] * s1[i + 2]; t2 -= vp2[i] * s2[i + 2]; }
            s0[i] = t0 / xx[i]; s1[i] = t1 / xx[i]; s2[i] = t2 / xx[i];
        }
        // store the coupled result into the same per-layer slots
        double dbb = 0., dbc = 0., dcb = 0., dcc = 0.;
        for (int l = 1; l <= ilevel; l++) {
            const int ju = 2 * l - 2, jv = 2 * l - 1;
            sxu[l - 1] = s0[ju];
            sxv[l - 1] = s0[jv];
            a.ddx[(size_t)ju * nel + ie] = (S)s1[ju];
            a.ddx[(size_t)jv * nel + ie] = (S)s1[jv];
            a.ddy[(size_t)ju * nel + ie] = (S)s2[ju];
            a.ddy[(size_t)jv * nel + ie] = (S)s2[jv];
            dbb += s1[ju]; dbc += s2[ju]; dcb += s1[jv]; dcc += s2[jv];
        }
        a.zel[(size_t)0 * nel + ie] = dbb;
        a.zel[(size_t)1 * nel + ie] = dbc;
        a.zel[(size_t)2 * nel + ie] = dcb;
        a.zel[(size_t)3 * nel + ie] = dcc;
    }
    // u^hat = u_old - dt * A^-1 F ; store A^-1 H.  Elements with fixed
    // transports (afix = 0, hydrodynamic.f90:1383-1392) keep the ones they
    // hold: the prescribed transports, or those of the previous step.
    const int afix = a.iuvfix ? 1 - a.iuvfix[ie] : 1;
    const R dtafix = dt * (R)afix;
    double uold = 0., vold = 0., uhat = 0., vhat = 0.;
    double dsum = 0.;
#pragma unroll
    for (int l = 1; l <= NL; l++) {
        if (l > ilevel) continue;
        const size_t ol = (size_t)(l - 1) * nel + ie;
        const R uo = a.uto[ol], vo = a.vto[ol];
        S unew, vnew;
        if (afix != 0) {
            unew = (S)((double)uo - (double)dtafix * sxu[l - 1]);
            vnew = (S)((double)vo - (double)dtafix * sxv[l - 1]);
        } else {
            unew = a.utfix ? a.utfix[ol] : uo;
            vnew = a.vtfix ? a.vtfix[ol] : vo;
        }
        a.utn[(size_t)(l - 1) * nel + ie] = unew;
        a.vtn[(size_t)(l - 1) * nel + ie] = vnew;
        uold += (double)uo; vold += (double)vo;
        uhat += (double)unew; vhat += (double)vnew;
        if constexpr (!CORI) {
            // decoupled columns: the cross terms ddx(jv), ddy(ju) are zero and
            // stay zero in the arrays (never written on this path)
            const int ju = 2 * l - 2, jv = 2 * l - 1;
            a.ddx[(size_t)ju * nel + ie] = (S)sh[l - 1];
            a.ddy[(size_t)jv * nel + ie] = (S)sh[l - 1];
            dsum += sh[l - 1];
        }
    }
    if constexpr (!CORI) {
        a.zel[(size_t)0 * nel + ie] = dsum;
        a.zel[(size_t)1 * nel + ie] = 0.;
        a.zel[(size_t)2 * nel + ie] = 0.;
        a.zel[(size_t)3 * nel + ie] = dsum;
    }
    const double az = (double)a.az;
    a.zel[(size_t)4 * nel + ie] = az * uhat + (1. - az) * uold;
    a.zel[(size_t)5 * nel + ie] = az * vhat + (1. - az) * vold;
}

// ------------------------------------------------------------------ hydro_zeta
// element matrix of hydro_zeta (hydrodynamic.f90:600-728) for element ie,
// returned row i (after level boundary conditions and the excluded-area rule)
template <typename S>
__device__ inline void zeta_elem(const HydroArgs<S> &a, int ie, S hia[3][3], S hik[3]) {
    using R = S;
    const int nel = a.nel;
    const double dt = (double)(R)a.dt;
    const double az = a.az, am = a.am;
    const int afix = a.iuvfix ? 1 - a.iuvfix[ie] : 1;
    const double delta = dt * dt * az * am * (double)a.grav * afix;
    const double aj = a.geo[(size_t)6 * nel + ie];
    double b[3], c[3], z[3];
    int kn[3];
    for (int q = 0; q < 3; q++) {
        kn[q] = a.nen3v[3 * ie + q];
        b[q] = a.geo[(size_t)q * nel + ie];
        c[q] = a.geo[(size_t)(3 + q) * nel + ie];
        z[q] = (double)a.zeo[(size_t)q * nel + ie];
    }
    const double dbb = a.zel[ie], dbc = a.zel[(size_t)nel + ie], dcb = a.zel[(size_t)2 * nel + ie];
    const double dcc = a.zel[(size_t)3 * nel + ie], ut = a.zel[(size_t)4 * nel + ie], vt = a.zel[(size_t)5 * nel + ie];
    for (int n = 0; n < 3; n++) {
        for (int m = 0; m < 3; m++) {
            const double abn = b[n] * (b[m] * dbb + c[m] * dbc);
            const double acn = c[n] * (b[m] * dcb + c[m] * dcc);
            const double h11 = delta * (abn + acn);
            hia[n][m] = (R)(aj * ((n == m ? 4. : 0.) + 12. * h11));
        }
        const double acu = (double)hia[n][0] * z[0] + (double)hia[n][1] * z[1] + (double)hia[n][2] * z[2];
        const double andg = 4. * aj * dt * (a.andgz ? (double)a.andgz[kn[n]] : 0.);
        hik[n] = (R)(acu + andg + 12. * aj * dt * (ut * b[n] + vt * c[n]));
    }
    for (int q = 0; q < 3; q++) {
        const R rz = a.rzv[kn[q]];
        if (rz != (R)kFlag) {
            const double rw = (double)rz;
            const int j1 = (q + 1) % 3, j2 = (q + 2) % 3;
            hik[j1] = (R)((double)hik[j1] - rw * (double)hia[j1][q]);
            hik[j2] = (R)((double)hik[j2] - rw * (double)hia[j2][q]);
            hia[q][j1] = 0; hia[q][j2] = 0;
            hia[j1][q] = 0; hia[j2][q] = 0;
            hik[q] = (R)(rw * (double)hia[q][q]);
        }
    }
    if (a.iweg[ie] != 0) {
        const double hh999 = aj * 12.;
        for (int n = 0; n < 3; n++) {
            for (int m = 0; m < 3; m++) hia[n][m] = (R)(hh999 * (b[n] * b[m] + c[n] * c[m]));
            hik[n] = 0;
        }
        for (int n = 0; n < 3; n++) {
            const int iv = a.inod[kn[n]];
            if (iv != 0 && iv != -2) {
                for (int m = 0; m < 3; m++) hia[n][m] = 0;
                hik[n] = 0;
            }
        }
    }
}

// hydro_zeta's element matrix and vector (all three rows) of every element
template <typename S>
__global__ __launch_bounds__(kThreads) void k_zeta_elem(HydroArgs<S> a) {
    const int ie = blockIdx.x * blockDim.x + threadIdx.x;
    if (ie >= a.nel) return;
    S hia[3][3], hik[3];
    zeta_elem(a, ie, hia, hik);
    const size_t nel = a.nel;
    for (int n = 0; n < 3; n++) {
        for (int m = 0; m < 3; m++) a.zrow[(3 * n + m) * nel + ie] = hia[n][m];
        a.zrow[(9 + n) * nel + ie] = hik[n];
    }
}

// system_init + system_assemble + system_add_rhs as a row gather: row k
// sums its incident elements' rows in element order, the order of the
// reference's scatter into c2coo / rvec2d.  The row is accumulated in
// registers by position in the row (all element loads issued first); rows
// longer than kEllWidth or nodes with more than kIncWidth elements
// accumulate in place.
template <typename S>
__global__ __launch_bounds__(kThreads) void k_zeta_rows(HydroArgs<S> a) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.nkn) return;
    const size_t nel = a.nel;
    const int r0 = a.rowptr[k], len = a.rowptr[k + 1] - r0;
    IncSlots in;
    in.load(a.inc_cnt, a.inc_el, a.nkn, k);
    const S dt = (S)a.dt;
    const S rq = a.rqv[k];
    double rhs = 0.;
    if (len <= kEllWidth && in.cnt <= kIncWidth) {
        S hv[kIncWidth][3], hk[kIncWidth];
        int pos[kIncWidth][3];
#pragma unroll
        for (int s = 0; s < kIncWidth; s++) {
            const int ie = in.pk[s] >> 2, i = in.pk[s] & 3;
#pragma unroll
            for (int j = 0; j < 3; j++) {
                hv[s][j] = a.zrow[(3 * i + j) * nel + ie];
                pos[s][j] = a.ijp[9 * (size_t)ie + 3 * i + j] - r0;
            }
            hk[s] = a.zrow[(9 + i) * nel + ie];
        }
        double acc[kEllWidth];
#pragma unroll
        for (int t = 0; t < kEllWidth; t++) acc[t] = 0.;
#pragma unroll
        for (int s = 0; s < kIncWidth; s++) {
            if (s < in.cnt) {
#pragma unroll
                for (int j = 0; j < 3; j++)
#pragma unroll
                    for (int t = 0; t < kEllWidth; t++)
                        if (t == pos[s][j]) acc[t] = acc[t] + (double)hv[s][j];
                rhs = rhs + (double)hk[s];
            }
        }
#pragma unroll
        for (int t = 0; t < kEllWidth; t++)
            if (t < len) a.aval[r0 + t] = acc[t];
    } else {
        for (int q = r0; q < r0 + len; q++) a.aval[q] = 0.;
        for (int q = a.nk_ptr[k]; q < a.nk_ptr[k + 1]; q++) {
            const int pk = a.nk_el[q];
            const int ie = pk >> 2, i = pk & 3;
            for (int j = 0; j < 3; j++) {
                const int slot = a.ijp[9 * (size_t)ie + 3 * i + j];
                a.aval[slot] = a.aval[slot] + (double)a.zrow[(3 * i + j) * nel + ie];
            }
            rhs = rhs + (double)a.zrow[(9 + i) * nel + ie];
        }
    }
    rhs = rhs + (double)(dt * rq);
    a.rhs[k] = rhs;
    if (a.sv_ell) {
        // the solve's inputs: the row in the ELL slots (and in FP32), the
        // initial guess (FP32 path: extrapolated in time), cleared words
        for (int q = r0; q < r0 + len; q++) {
            const double v = a.aval[q];
            if (q - r0 < kEllWidth) a.sv_ell[(size_t)(q - r0) * a.nkn + k] = (S)v;
            if (a.sv_a32) a.sv_a32[q] = (float)v;
        }
        const double z = (double)a.zn[k];
        if (a.sv_zprev) {
            a.x[k] = a.sv_have_prev ? 2. * z - (double)a.sv_zprev[k] : z;
            a.sv_zprev[k] = (float)a.zn[k];
        } else {
            a.x[k] = z;
        }
        if (k == 0)
            for (int w = CT_ITS; w <= CT_NREF; w++) a.sv_ctrl[w] = 0;
    }
}

// ------------------------------------------------------------------ wet/dry
// setweg(iweich), wetdry.f90:128-347, one work-item per element; the number
// of elements that changed state is accumulated in *cnt.
// one setweg(iweich) update of element ie's state (weg, wet); returns the
// number of state changes it counts
template <typename S>
__device__ inline int setweg_el(const HydroArgs<S> &a, int ie, int iweich, int bnewtime, int &weg, int &wet) {
    using R = S;
    const int nel = a.nel;
    int iwh = 0;
    const int kn[3] = {a.nen3v[3 * ie], a.nen3v[3 * ie + 1], a.nen3v[3 * ie + 2]};
    if (iweich == -1) {
        weg = 0;
        wet = 0;
    } else if (iweich == 0) {
        const R aomega = (R)a.geo[(size_t)6 * nel + ie];
        R zmin = (R)a.hzmin + (R)a.volmin / ((R)4 * aomega);
        if (zmin > (R)a.hzoff) zmin = (R)a.hzmin + (R)0.5 * ((R)a.hzmin + (R)a.hzoff);
        int iweg = 0;
        for (int ii = 0; ii < 3; ii++) {
            const R hm = a.hm3v[(size_t)ii * nel + ie];
            const R hzg = a.zen[(size_t)ii * nel + ie] + hm;
            if (hzg < (R)a.hzoff) iweg++;
            if (hzg < zmin) a.zen[(size_t)ii * nel + ie] = zmin - hm;
        }
        weg = iweg;
        if (iweg == 0) wet = 1;
        if (iweg > 0) iwh++;
    } else if (iweich == 1 || iweich == 2) {
        const R hzlim = (iweich == 1) ? (R)a.hzmin : (R)a.hzoff;
        if (weg == 0) {
            int iweg = 0;
            for (int ii = 0; ii < 3; ii++) {
                const R hzg = a.zn[kn[ii]] + a.hm3v[(size_t)ii * nel + ie];
                if (hzg < hzlim) iweg++;
            }
            if (iweg > 0) {
                if (weg == 0) { iwh++; wet = 0; }
                weg = iweg;
            }
        }
    } else if (iweich == 3) {
        if (weg > 0) {
            int iweg = 0;
            R hztot = 0;
            for (int ii = 0; ii < 3; ii++) {
                const R hzg = a.zn[kn[ii]] + a.hm3v[(size_t)ii * nel + ie];
                hztot = hztot + hzg;
                if (hzg < (R)a.hzon) iweg++;
            }
            hztot = hztot / (R)3;
            if (hztot >= (R)a.hzon && -wet > 5) iweg = 0;
            if (iweg == 0) {
                if (weg > 0) { iwh++; wet = 0; }
                weg = 0;
            }
        }
    }
    if (bnewtime || wet == 0) {
        if (weg == 0) wet++;
        else wet--;
    }
    return iwh;
}

template <typename S>
__global__ __launch_bounds__(kThreads) void k_setweg(HydroArgs<S> a, int iweich, int bnewtime, int *cnt,
                                                     const int *gate) {
    if (gate && !*gate) return;
    const int ie = blockIdx.x * blockDim.x + threadIdx.x;
    if (ie >= a.nel) return;
    int weg = a.iweg[ie], wet = a.iwet[ie];
    const int iwh = setweg_el<S>(a, ie, iweich, bnewtime, weg, wet);
    a.iweg[ie] = weg;
    a.iwet[ie] = wet;
    if (iwh && cnt) atomicAdd(cnt, iwh);
}

// The start of a step in one launch: set_dry (setweg(2) then setweg(3),
// per element), then the copies of copy_uvz/copy_depth that follow it
// (zov <- znv per node, zeov <- zenv per element); work-item i handles
// element i and node i.
template <typename S>
__global__ __launch_bounds__(kThreads) void k_step_start(HydroArgs<S> a, int bnewtime) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const size_t nel = a.nel;
    if (i < a.nel) {
        int weg = a.iweg[i], wet = a.iwet[i];
        setweg_el<S>(a, i, 2, bnewtime, weg, wet);
        setweg_el<S>(a, i, 3, 0, weg, wet);
        a.iweg[i] = weg;
        a.iwet[i] = wet;
        for (int ii = 0; ii < 3; ii++) a.zeo[ii * nel + i] = a.zen[ii * nel + i];
    }
    if (i < a.nkn) a.zo[i] = a.zn[i];
}

// setnod (topo_update.f90:166-270) + adjust_mass_flux (bnd_routines.f90:1185)
template <typename S>
__global__ __launch_bounds__(kThreads) void k_setnod(HydroArgs<S> a) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.nkn) return;
    IncSlots in;
    in.load(a.inc_cnt, a.inc_el, a.nkn, k);
    int wg[kIncWidth];
    double an[kIncWidth];
#pragma unroll
    for (int s = 0; s < kIncWidth; s++) {
        const int ie = in.pk[s] >> 2, ii = in.pk[s] & 3;
        wg[s] = a.iweg[ie];
        an[s] = a.ang[(size_t)ii * a.nel + ie];
    }
    double win = 0.;
#pragma unroll
    for (int s = 0; s < kIncWidth; s++)
        if (s < in.cnt && wg[s] == 0) win = win + an[s];
    if (in.cnt > kIncWidth)
        for (int q = a.nk_ptr[k] + kIncWidth; q < a.nk_ptr[k + 1]; q++) {
            const int pk = a.nk_el[q];
            const int ie = pk >> 2, ii = pk & 3;
            if (a.iweg[ie] == 0) win = win + a.ang[(size_t)ii * a.nel + ie];
        }
    int inod;
    if (win > 359.99) inod = 0;
    else if (win == 0.) inod = -2;
    else inod = -1;
    const int ibc = a.ibnd[k];
    if (ibc > 0) {
        const int typ = a.ibtyp[ibc];
        if (typ >= 3) {
            if (inod == -1) inod = ibc;
        } else if (typ > 0) {
            if (inod != -1) atomicCAS(a.err, 0, ERR_SETNOD);
            inod = ibc;
        }
    }
    a.inod[k] = inod;
    if (inod == -2) {
        a.rqv[k] = 0;
        if (a.mflux)
            for (int l = 0; l < a.nlv; l++) a.mflux[(size_t)l * a.nkn + k] = 0;
    }
}

// system_get: z = real(x).  With a gate (a solver flag), nothing happens
// unless the flag is set: the kernel can be enqueued before the host knows
// whether the solve has finished.
// With dx (FP64 path) the Krylov correction is added first, x = x + dx
// (tidycg), and kept in x.
template <typename S>
__global__ __launch_bounds__(kThreads) void k_system_get(HydroArgs<S> a, int *cnt, const int *gate,
                                                         const double *dx) {
    if (gate && !*gate) return;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k == 0) *cnt = 0;     // counter of the setweg that follows
    if (k >= a.nkn) return;
    double v = a.x[k];
    if (dx) {
        v = v + dx[k];
        a.x[k] = v;
    }
    a.zn[k] = (S)v;
}

// ------------------------------------------------------------------ final transports
// hydro_transports_final, hydrodynamic.f90:1468-1578; without Coriolis the
// cross terms ddx(jv), ddy(ju) are zero and are not read
template <typename S, bool CORI>
__device__ inline void transports_final_el(const HydroArgs<S> &a, int ie) {
    using R = S;
    const int nel = a.nel;
    const R dt = (R)a.dt;
    const double beta = (double)(dt * (R)a.grav) * (double)a.am;
    double bz = 0., cz = 0.;
    for (int ii = 0; ii < 3; ii++) {
        const int kk = a.nen3v[3 * ie + ii];
        const double dz = (double)(a.zn[kk] - a.zeo[(size_t)ii * nel + ie]);
        bz = bz + dz * a.geo[(size_t)ii * nel + ie];
        cz = cz + dz * a.geo[(size_t)(3 + ii) * nel + ie];
    }
    const int ilevel = a.ilhv[ie];
    const double rfix = a.iuvfix ? 1 - a.iuvfix[ie] : 1;      // hydrodynamic.f90:1525-1527
    double um = 0., vm = 0.;
    for (int l = 1; l <= ilevel; l++) {
        const int ju = 2 * l - 2, jv = 2 * l - 1;
        double du, dv;
        if constexpr (CORI) {
            du = beta * ((double)a.ddx[(size_t)ju * nel + ie] * bz + (double)a.ddy[(size_t)ju * nel + ie] * cz);
            dv = beta * ((double)a.ddx[(size_t)jv * nel + ie] * bz + (double)a.ddy[(size_t)jv * nel + ie] * cz);
        } else {
            du = beta * ((double)a.ddx[(size_t)ju * nel + ie] * bz);
            dv = beta * ((double)a.ddy[(size_t)jv * nel + ie] * cz);
        }
        const size_t o = (size_t)(l - 1) * nel + ie;
        const S u = (S)((double)a.utn[o] - du * rfix);
        const S v = (S)((double)a.vtn[o] - dv * rfix);
        a.utn[o] = u;
        a.vtn[o] = v;
        um = um + (double)u;
        vm = vm + (double)v;
    }
    a.un[ie] = (S)um;
    a.vn[ie] = (S)vm;
}
template <typename S, bool CORI>
__global__ __launch_bounds__(kThreads) void k_transports_final(HydroArgs<S> a) {
    const int ie = blockIdx.x * blockDim.x + threadIdx.x;
    if (ie < a.nel) transports_final_el<S, CORI>(a, ie);
}

// ------------------------------------------------------------------ setzev/setuvd/baro2l
// setzev (wetdry.f90:594), setuvd (wetdry.f90:350-590), baro2l
// (transforms.f90:437-481).  The layer thickness baro2l reads is the one
// before make_new_depth, i.e. the old level's.
template <typename S>
__device__ inline void setzev_setuvd_el(const HydroArgs<S> &a, int ie) {
    using R = S;
    const int nel = a.nel;
    if (a.iweg[ie] == 0) {
        for (int ii = 0; ii < 3; ii++) a.zen[(size_t)ii * nel + ie] = a.zn[a.nen3v[3 * ie + ii]];
        return;
    }
    if (!(a.iweg[ie] > 0)) return;
    const R dt = (R)a.dt, hzmin = (R)a.hzmin, az = (R)a.az, azt = (R)1 - az;
    const R eps = (R)1.e-5;
    const R uo = a.uo[ie], vo = a.vo[ie];
    const R axdt = (R)3 * dt, adt = (R)1 / axdt;
    double zm = 0., zmed, d1, d2, det;
    R z[3], b[3], c[3], zn[3], hm[3], u, v, zz;
    for (int ii = 0; ii < 3; ii++) {
        zn[ii] = a.zen[(size_t)ii * nel + ie];
        zm = zm + (double)zn[ii];
        b[ii] = (R)a.geo[(size_t)ii * nel + ie];
        c[ii] = (R)a.geo[(size_t)(3 + ii) * nel + ie];
        hm[ii] = a.hm3v[(size_t)ii * nel + ie];
    }
    zmed = zm / 3.;
    int itot = 0, isum = 0;
    for (int ii = 0; ii < 3; ii++) {
        if (zmed + (double)hm[ii] < (double)hzmin) {
            z[ii] = hzmin - hm[ii];
            itot++;
            isum += ii + 1;
        } else {
            z[ii] = (R)zmed;
        }
        zm = zm - (double)z[ii];
    }
    bool fail = false;
    if (itot == 1) {
        const int i1 = (isum % 3) + 1, i2 = (i1 % 3) + 1;
        if ((double)z[i1 - 1] + zm * 0.5 + (double)hm[i1 - 1] < (double)hzmin) {
            z[i1 - 1] = hzmin - hm[i1 - 1];
            z[i2 - 1] = (R)(3. * zmed - (double)z[i1 - 1] - (double)z[isum - 1]);
        } else if ((double)z[i2 - 1] + zm * 0.5 + (double)hm[i2 - 1] < (double)hzmin) {
            z[i2 - 1] = hzmin - hm[i2 - 1];
            z[i1 - 1] = (R)(3. * zmed - (double)z[i2 - 1] - (double)z[isum - 1]);
        } else {
            z[i1 - 1] = (R)((double)z[i1 - 1] + 0.5 * zm);
            z[i2 - 1] = (R)((double)z[i2 - 1] + 0.5 * zm);
        }
    } else if (itot == 2) {
        const int i1 = 6 - isum;
        z[i1 - 1] = (R)((double)z[i1 - 1] + zm);
    } else if (itot == 3) {
        fail = true;
    }
    zm = 0.;
    for (int ii = 0; ii < 3; ii++) {
        zm = zm + (double)z[ii];
        if (z[ii] + hm[ii] < hzmin - eps) fail = true;
    }
    if (fabs(zm - 3. * zmed) > (double)eps) fail = true;
    if (fail) { atomicCAS(a.err, 0, ERR_SETUVD); return; }
    d1 = (double)((z[0] - zn[0]) * adt - azt * (b[0] * uo + c[0] * vo));
    d2 = (double)((z[1] - zn[1]) * adt - azt * (b[1] * uo + c[1] * vo));
    det = (double)((R)1 / (b[0] * c[1] - b[1] * c[0]));
    u = (R)(det * ((double)c[1] * d1 - (double)c[0] * d2));
    v = (R)(det * (-(double)b[1] * d1 + (double)b[0] * d2));
    itot = 0;
    for (int ii = 0; ii < 3; ii++) {
        zz = axdt * (b[ii] * (azt * uo + az * u) + c[ii] * (azt * vo + az * v)) + zn[ii];
        if (zz + hm[ii] < hzmin - eps) itot++;
    }
    if (itot > 0) {
        if (az == (R)0) { atomicCAS(a.err, 0, ERR_SETUVD); return; }
        det = (double)((R)1 / (az * (b[0] * c[1] - b[1] * c[0])));
        u = (R)(det * ((double)c[1] * d1 - (double)c[0] * d2));
        v = (R)(det * (-(double)b[1] * d1 + (double)b[0] * d2));
    }
    itot = 0;
    zm = 0.;
    for (int ii = 0; ii < 3; ii++) {
        zz = axdt * (b[ii] * (azt * uo + az * u) + c[ii] * (azt * vo + az * v)) + zn[ii];
        if (zz + hm[ii] < hzmin - eps) itot++;
        a.zen[(size_t)ii * nel + ie] = zz;
        zm = zm + (double)zz;
    }
    if (fabs(zm - 3. * zmed) > (double)eps || itot > 0) { atomicCAS(a.err, 0, ERR_SETUVD); return; }
    a.un[ie] = u;
    a.vn[ie] = v;
    // baro2l
    const int ilevel = a.ilhv[ie];
    R htot = 0;
    for (int l = 1; l <= ilevel; l++) htot = htot + a.hdeo[(size_t)(l - 1) * nel + ie];
    for (int l = 1; l <= ilevel; l++) {
        const R weight = a.hdeo[(size_t)(l - 1) * nel + ie] / htot;
        a.utn[(size_t)(l - 1) * nel + ie] = u * weight;
        a.vtn[(size_t)(l - 1) * nel + ie] = v * weight;
    }
}
template <typename S>
__global__ __launch_bounds__(kThreads) void k_setzev_setuvd(HydroArgs<S> a) {
    const int ie = blockIdx.x * blockDim.x + threadIdx.x;
    if (ie < a.nel) setzev_setuvd_el<S>(a, ie);
}

// ------------------------------------------------------------------ setdepth
// setdepth for z-layers (elems_dealing.f90:930-1156): element thickness
template <typename S>
__device__ inline void setdepth_elem_el(const HydroArgs<S> &a, int ie, const S *zen, S *hden) {
    using R = S;
    const int nel = a.nel, lmax = a.ilhv[ie];
    R zmed = 0;
    for (int ii = 0; ii < 3; ii++) zmed = zmed + zen[(size_t)ii * nel + ie];
    zmed = zmed / (R)3;
    const R htot = a.hev[ie];
    bool bad = false;
    // layers below the element's bottom keep their zero thickness
    for (int l = 1; l <= lmax; l++) {
        R h = 0;
        if (lmax == 1) {
            h = htot + zmed;
        } else if (l < lmax) {
            h = a.hldv[l - 1];
            if (l == 1) h = h + zmed;
        } else {
            h = htot - a.hlv[lmax - 2];
            if (h < (R)0) bad = true;
        }
        if (h <= (R)0) bad = true;
        hden[(size_t)(l - 1) * nel + ie] = h;
    }
    if (bad) atomicCAS(a.err, 0, ERR_SETDEPTH);
}
template <typename S>
__global__ __launch_bounds__(kThreads) void k_setdepth_elem(HydroArgs<S> a, const S *zen, S *hden) {
    const int ie = blockIdx.x * blockDim.x + threadIdx.x;
    if (ie < a.nel) setdepth_elem_el<S>(a, ie, zen, hden);
}

// node thickness: volume contributions of the incident elements (element
// order, per layer) divided by the nodal area
template <typename S, int NL>
__global__ __launch_bounds__(kThreads) void k_setdepth_node(HydroArgs<S> a, const S *zen, S *hdkn) {
    using R = S;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.nkn) return;
    const int nel = a.nel, nkn = a.nkn;
    R acc[NL];
#pragma unroll
    for (int l = 0; l < NL; l++) acc[l] = 0;
    bool bad = false;
    // the element terms, added in list order; the slots' loads are issued
    // at once
    auto add = [&](R areael, int lmax, R htot, R z) {
        const R areafv = areael / (R)3;
        if (lmax == 1) {
            acc[0] = acc[0] + areafv * (htot + z);
        } else {
            const R hlast = htot - a.hlv[lmax - 2];
            if (hlast < (R)0) bad = true;
#pragma unroll
            for (int l = 1; l <= NL; l++) {
                if (l < lmax) acc[l - 1] = acc[l - 1] + areafv * a.hldv[l - 1];
                if (l == 1) acc[0] = acc[0] + areafv * z;
                if (l == lmax) acc[l - 1] = acc[l - 1] + areafv * hlast;
            }
        }
    };
    IncSlots in;
    in.load(a.inc_cnt, a.inc_el, nkn, k);
    // the node's layer areas, loaded with the slots (used after the sums)
    R sar[NL];
    {
        const int lk = a.ilhkv[k];
#pragma unroll
        for (int l = 0; l < NL; l++) sar[l] = l < lk ? a.areakv[(size_t)l * nkn + k] : (R)0;
    }
    R sarea[kIncWidth], shtot[kIncWidth], sz[kIncWidth];
    int slmax[kIncWidth];
#pragma unroll
    for (int s = 0; s < kIncWidth; s++) {
        const int ie = in.pk[s] >> 2, ii = in.pk[s] & 3;
        sarea[s] = (R)(12. * a.geo[(size_t)6 * nel + ie]);
        slmax[s] = a.ilhv[ie];
        shtot[s] = a.hm3v[(size_t)ii * nel + ie];
        sz[s] = zen[(size_t)ii * nel + ie];
    }
#pragma unroll
    for (int s = 0; s < kIncWidth; s++)
        if (s < in.cnt) add(sarea[s], slmax[s], shtot[s], sz[s]);
    if (in.cnt > kIncWidth)
        for (int q = a.nk_ptr[k] + kIncWidth; q < a.nk_ptr[k + 1]; q++) {
            const int pk = a.nk_el[q];
            const int ie = pk >> 2, ii = pk & 3;
            add((R)(12. * a.geo[(size_t)6 * nel + ie]), a.ilhv[ie], a.hm3v[(size_t)ii * nel + ie],
                zen[(size_t)ii * nel + ie]);
        }
    const int lmaxk = a.ilhkv[k];
    bool stop = false;
#pragma unroll
    for (int l = 1; l <= NL; l++) {
        if (l <= lmaxk && !stop) {
            const R ar = sar[l - 1];
            if (ar > (R)0) acc[l - 1] = acc[l - 1] / ar;
            else stop = true;
        }
        if (l <= lmaxk && acc[l - 1] <= (R)0) bad = true;
        if (l <= lmaxk) hdkn[(size_t)(l - 1) * nkn + k] = acc[l - 1];
    }
    if (bad) atomicCAS(a.err, 0, ERR_SETDEPTH);
}

// ------------------------------------------------------------------ hydro_vertical
// hydrodynamic.f90:1582-1781: divergence gathered per node and layer, then
// the column is integrated from the bottom; w = 0 on zeta boundaries.  One
// workgroup per tile of consecutive nodes: one work-item per node-layer
// gathers the incident elements (element order, the reference's scatter
// order) into LDS, then one work-item per node integrates its column.
template <int NL>
struct VertTile {
    static constexpr int TN = NL <= 16 ? 64 : 32;
};

template <typename S, int NL>
__global__ __launch_bounds__(kThreads) void k_hydro_vertical(HydroArgs<S> a) {
    using R = S;
    constexpr int TN = VertTile<NL>::TN;
    __shared__ R vfL[NL][TN], vaL[NL][TN], dvL[NL][TN];
    const int nel = a.nel, nkn = a.nkn, nlvdi = a.nlvdi;
    const int k0 = xcd_block() * TN;
    const R az = (R)a.az, azt = (R)1 - az, dt = (R)a.dt;
    const int nlv = min(a.nlv, NL);
    // phase 1: a work-item takes one node of the tile (TN divides kThreads)
    // and every (kThreads/TN)-th layer of it; the node's incidence slots and
    // their elements' geometry are loaded once for all its layers
    static_assert(kThreads % TN == 0, "tile must divide the workgroup");
    {
        const int kk = threadIdx.x % TN, k = k0 + kk;
        const int lstep = kThreads / TN;
        const int lim = k < nkn ? min(nlv, a.ilhkv[k]) : 0;
        int slv[kIncWidth], sie[kIncWidth];
        R saj[kIncWidth], sb[kIncWidth], sc[kIncWidth];
        IncSlots in;
        if (lim > 0) {
            in.load(a.inc_cnt, a.inc_el, nkn, k);
#pragma unroll
            for (int s = 0; s < kIncWidth; s++) {
                const int ie = in.pk[s] >> 2, ii = in.pk[s] & 3;
                sie[s] = ie;
                slv[s] = 0;
                saj[s] = sb[s] = sc[s] = (R)0;
                if (s < in.cnt) {       // (padding slots load nothing)
                    slv[s] = a.ilhv[ie];
                    saj[s] = (R)(4. * a.geo[(size_t)6 * nel + ie]);
                    sb[s] = (R)a.geo[(size_t)ii * nel + ie];
                    sc[s] = (R)a.geo[(size_t)(3 + ii) * nel + ie];
                }
            }
        }
        for (int l = threadIdx.x / TN + 1; l <= lim; l += lstep) {
            R vf = 0, va = 0;
            // layer l of every slot's element loaded at once (only the
            // slots that hold an element reaching layer l), then added in
            // list order
            R sun[kIncWidth], svn[kIncWidth], suo[kIncWidth], svo[kIncWidth];
#pragma unroll
            for (int s = 0; s < kIncWidth; s++) {
                sun[s] = svn[s] = suo[s] = svo[s] = (R)0;
                if (s < in.cnt && slv[s] >= l) {
                    const size_t o = (size_t)(l - 1) * nel + sie[s];
                    sun[s] = a.utn[o]; svn[s] = a.vtn[o]; suo[s] = a.uto[o]; svo[s] = a.vto[o];
                }
            }
            const size_t on = (size_t)(l - 1) * nkn + k;
            const R ar = a.areakv[on];
            const R hn = a.hdkn[on], ho = a.hdko[on];
            const R qq = a.mflux ? a.mflux[on] : (R)0;
#pragma unroll
            for (int s = 0; s < kIncWidth; s++) {
                if (s < in.cnt && slv[s] >= l) {
                    const R ffn = sun[s] * sb[s] + svn[s] * sc[s];
                    const R ffo = suo[s] * sb[s] + svo[s] * sc[s];
                    const R ff = ffn * az + ffo * azt;
                    vf = vf + (R)3 * saj[s] * ff;
                    va = va + saj[s];
                }
            }
            if (in.cnt > kIncWidth)
                for (int q = a.nk_ptr[k] + kIncWidth; q < a.nk_ptr[k + 1]; q++) {
                    const int ie = a.nk_el[q] >> 2, ii = a.nk_el[q] & 3;
                    if (a.ilhv[ie] < l) continue;
                    const R aj = (R)(4. * a.geo[(size_t)6 * nel + ie]);
                    const R b = (R)a.geo[(size_t)ii * nel + ie], c = (R)a.geo[(size_t)(3 + ii) * nel + ie];
                    const size_t o = (size_t)(l - 1) * nel + ie;
                    const R ffn = a.utn[o] * b + a.vtn[o] * c;
                    const R ffo = a.uto[o] * b + a.vto[o] * c;
                    const R ff = ffn * az + ffo * azt;
                    vf = vf + (R)3 * aj * ff;
                    va = va + aj;
                }
            // the column terms of this layer (read here, so that the column
            // scan below waits on LDS only)
            const R voln = hn * ar, volo = ho * ar;
            vfL[l - 1][kk] = vf + qq;
            vaL[l - 1][kk] = va;
            dvL[l - 1][kk] = (voln - volo) / dt;
        }
    }
    __syncthreads();
    const int kk = threadIdx.x, k = k0 + kk;
    if (kk >= TN || k >= nkn) return;
    const int lmax = a.ilhkv[k];
    const bool zb = a.zbound[k] != 0;
    // bottom-up: w(l-1) = w(l) + div - dV/dt, then divide by the top area
    R wb = 0;
    for (int l = lmax; l >= 1; l--) {
        const R wdiv = vfL[l - 1][kk];
        const R dvdt = dvL[l - 1][kk];
        const R wl = wb + wdiv - dvdt;
        R wout = wl;
        if (l == 1) wout = 0;
        else if (vaL[l - 1][kk] > (R)0) wout = wl / vaL[l - 1][kk];
        a.wln[(size_t)(l - 1) * nkn + k] = zb ? (S)0 : wout;
        wb = wl;
    }
    for (int l = lmax; l <= nlvdi; l++) a.wln[(size_t)l * nkn + k] = (S)0;
}

// ttov (transforms.f90:98-122) + uvint (transforms.f90:336-362)
template <typename S>
__device__ inline void ttov_el(const HydroArgs<S> &a, int ie) {
    const int nel = a.nel;
    double u = 0., v = 0.;
    // layers below the element's bottom keep their zero velocities
    const int lmax = a.ilhv[ie];
    for (int l = 1; l <= lmax; l++) {
        const size_t o = (size_t)(l - 1) * nel + ie;
        const S h = a.hden[o];
        const S ut = a.utn[o], vt = a.vtn[o];
        a.uln[o] = (h > (S)0) ? ut / h : (S)0;
        a.vln[o] = (h > (S)0) ? vt / h : (S)0;
        u = u + (double)ut;
        v = v + (double)vt;
    }
    a.un[ie] = (S)u;
    a.vn[ie] = (S)v;
}
template <typename S>
__global__ __launch_bounds__(kThreads) void k_ttov_uvint(HydroArgs<S> a) {
    const int ie = blockIdx.x * blockDim.x + threadIdx.x;
    if (ie < a.nel) ttov_el<S>(a, ie);
}

// The element work of hydro after the solve, one work-item per element:
// hydro_transports_final, setzev + setuvd + baro2l, the element depths of
// make_new_depth, compute_velocities (ttov + uvint).  Each reads the final
// node levels and its own element's data only, in the reference's order.
template <typename S, bool CORI>
__global__ __launch_bounds__(kThreads) void k_elem_after_solve(HydroArgs<S> a) {
    const int ie = blockIdx.x * blockDim.x + threadIdx.x;
    if (ie >= a.nel) return;
    transports_final_el<S, CORI>(a, ie);
    setzev_setuvd_el<S>(a, ie);
    setdepth_elem_el<S>(a, ie, a.zen, a.hden);
    ttov_el<S>(a, ie);
    if (a.iweg_snap) a.iweg_snap[ie] = a.iweg[ie];
}

// setznv (wetdry.f90:619-700) at start-up
template <typename S>
__global__ __launch_bounds__(kThreads) void k_setznv(HydroArgs<S> a) {
    using R = S;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.nkn) return;
    const int nel = a.nel;
    R z = (R)kFlag, v1 = 0, v2 = 0;
    for (int q = a.nk_ptr[k]; q < a.nk_ptr[k + 1]; q++) {
        const int pk = a.nk_el[q];
        const int ie = pk >> 2, ii = pk & 3;
        const R ze = a.zen[(size_t)ii * nel + ie];
        if (a.iweg[ie] == 0) {
            R zz = ze;
            if (z != (R)kFlag && z != zz) zz = (zz > z) ? zz : z;
            z = zz;
        } else {
            const R area = (R)(4. * a.geo[(size_t)6 * nel + ie]);
            v1 = v1 + ze * area;
            v2 = v2 + area;
        }
    }
    if (z == (R)kFlag) z = v1 / v2;
    a.zn[k] = z;
}

template <typename S>
__global__ void k_fill(S *p, S v, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

template <typename S>
__global__ void k_copy(S *dst, const S *src, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

template <typename S>
__global__ void k_set_rzv(S *rzv, const int *nodes, const int *bid, int n, int ibc, S level) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && bid[i] == ibc) rzv[nodes[i]] = level;
}

// all level boundaries (type 1) at once: level[b-1] on the nodes of boundary b
template <typename S>
__global__ void k_set_rzv_all(S *rzv, const int *nodes, const int *bid, const int *ibtyp, int n, const float *level) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && ibtyp[bid[i]] == 1) rzv[nodes[i]] = (S)level[bid[i] - 1];
}

// sigma_stp, eq_state.f90:154-235: the UNESCO 1980 equation of state (its
// published coefficients) in double precision; sigma = rho - 1000 (kg/m3)
// for salinity s, temperature t (deg C) and pressure p (bar)
template <typename R>
__device__ inline R sigma_stp(R s, R t, R p) {
    const double t1 = t, t2 = t1 * t1, t3 = t2 * t1, t4 = t3 * t1, t5 = t4 * t1;
    const double s1 = s, s2 = s1 * s1, s32 = sqrt(s2 * s1);
    const double p1 = p, p2 = p1 * p1;
    const double rho0 = 999.842594 + 6.793952e-2 * t1 + -9.095290e-3 * t2 + 1.001685e-4 * t3 + -1.120083e-6 * t4 +
                        6.536332e-9 * t5 +
                        (8.24493e-1 + -4.0899e-3 * t1 + 7.6438e-5 * t2 + -8.2467e-7 * t3 + 5.3875e-9 * t4) * s1 +
                        (-5.72466e-3 + 1.0227e-4 * t1 + -1.6546e-6 * t2) * s32 + 4.8314e-4 * s2;
    const double sbmk = 19652.21 + 148.4206 * t1 + -2.327105 * t2 + 1.360477e-2 * t3 + -5.155288e-5 * t4 +
                        (3.239908 + 1.43713e-3 * t1 + 1.16092e-4 * t2 + -5.77905e-7 * t3) * p1 +
                        (8.50935e-5 + -6.12293e-6 * t1 + 5.2787e-8 * t2) * p2 +
                        (54.6746 + -0.603459 * t1 + 1.09987e-2 * t2 + -6.1670e-5 * t3) * s1 +
                        (7.944e-2 + 1.6483e-2 * t1 + -5.3009e-4 * t2 + 1.91075e-4 * p1) * s32 +
                        (2.2838e-3 + -1.0981e-5 * t1 + -1.6078e-6 * t2) * p1 * s1 +
                        (-9.9348e-7 + 2.0816e-8 * t1 + 9.1697e-10 * t2) * p2 * s1;
    return (R)(rho0 / (1.0 - (double)p / sbmk) - 1000.0);
}

// rhoset, baroclinic.f90:654-720, z-layers: one work-item per node column,
// top down; the mid-layer pressure (bar) is the barotropic rho0 g depth plus
// the baroclinic part integrated from the anomaly of the previous pass, and
// the new anomaly is sigma(max(S, 0), T, p) - (rho0 - 1000)
template <typename S>
__global__ __launch_bounds__(kThreads) void k_rhoset(HydroArgs<S> a, const S *temp, const S *salt, S *rhov) {
    using R = S;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.nkn) return;
    const R grav = (R)a.grav, rho0 = (R)a.rowass;
    const R sigma0 = rho0 - (R)1000.;
    R depth = 0, presbc = 0;
    const int lmax = a.ilhkv[k];
    for (int l = 1; l <= lmax; l++) {
        const size_t o = (size_t)(l - 1) * a.nkn + k;
        const R hh = (R)0.5 * a.hldv[l - 1];
        depth = depth + hh;
        const R dpresc = rhov[o] * grav * hh;
        presbc = presbc + dpresc;
        const R presbt = rho0 * grav * depth;
        const R pres = (R)1.e-5 * (presbt + presbc);
        const R sl = salt[o] > (R)0 ? salt[o] : (R)0;
        rhov[o] = sigma_stp<R>(sl, temp[o], pres) - sigma0;
        depth = depth + hh;
        presbc = presbc + dpresc;
    }
}

// FP64 total volume (masscont, elems_dealing.f90:1191-1220): per-block sums
// in a fixed tree order, then a single-block pass over the partials.
template <typename S>
__global__ __launch_bounds__(kThreads) void k_volume_part(int nkn, const int *ilhkv, const S *areakv, const S *hdk,
                                                          const S *conc, double *part) {
    __shared__ double red[kThreads];
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    double t = 0.;
    if (k < nkn) {
        for (int l = 1; l <= ilhkv[k]; l++) {
            const size_t o = (size_t)(l - 1) * nkn + k;
            const S vol = hdk[o] * areakv[o];
            t += conc ? (double)conc[o] * (double)vol : (double)vol;
        }
    }
    red[threadIdx.x] = t;
    __syncthreads();
    for (int s = kThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(kThreads) void k_sum_parts(const double *part, int n, double *out) {
    __shared__ double red[kThreads];
    double t = 0.;
    for (int i = threadIdx.x; i < n; i += kThreads) t += part[i];
    red[threadIdx.x] = t;
    __syncthreads();
    for (int s = kThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

// ------------------------------------------------------------------ launchers
#define LAUNCH(M, kern, grid, ...)                                                         \
    do {                                                                                   \
        hipLaunchKernelGGL(kern, dim3(grid), dim3(kThreads), 0, (M).stream, __VA_ARGS__);  \
        HIPCHK(M, hipGetLastError());                                                      \
    } while (0)

// setweg(iweich); with `changed`, the number of elements that changed state
// is read back together with the device error word (d_err[0], kept in
// M.h_pinned[0] for the caller).  cnt_zeroed: the previous kernel in the
// stream already cleared the counter.
template <typename S>
static int set_dry_t(Model &M, int iweich, int bnewtime, int *changed, bool cnt_zeroed = false) {
    auto a = make_args<S>(M);
    int *cnt = M.d_err + 1;
    if (changed && !cnt_zeroed) HIPCHK(M, hipMemsetAsync(cnt, 0, sizeof(int), M.stream));
    LAUNCH(M, k_setweg<S>, nblocks(M.nel), a, iweich, bnewtime, cnt, (const int *)nullptr);
    if (changed) {
        HIPCHK(M, hipMemcpyAsync(M.h_pinned, M.d_err, 2 * sizeof(int), hipMemcpyDeviceToHost, M.stream));
        HIPCHK(M, hipStreamSynchronize(M.stream));
        *changed = M.h_pinned[1];
    }
    return 0;
}
int launch_set_dry(Model &M, int iweich, int bnewtime, int *changed) {
    return M.prec == SHYHIP_FP64 ? set_dry_t<double>(M, iweich, bnewtime, changed)
                                 : set_dry_t<float>(M, iweich, bnewtime, changed);
}

template <typename S>
static int setnod_t(Model &M) {
    auto a = make_args<S>(M);
    LAUNCH(M, k_setnod<S>, nblocks(M.nkn), a);
    return 0;
}
int launch_setnod(Model &M) { return M.prec == SHYHIP_FP64 ? setnod_t<double>(M) : setnod_t<float>(M); }

template <typename S, int NL>
static int hydro_intern_nl(Model &M) {
    auto a = make_args<S>(M);
    if (M.has_coriolis) LAUNCH(M, (k_hydro_intern<S, NL, true>), nblocks(M.nel), a);
    else LAUNCH(M, (k_hydro_intern<S, NL, false>), nblocks(M.nel), a);
    return 0;
}
template <typename S>
static int hydro_intern_t(Model &M) {
    switch (M.nl_tmpl) {
    case 1: return hydro_intern_nl<S, 1>(M);
    case 2: return hydro_intern_nl<S, 2>(M);
    case 4: return hydro_intern_nl<S, 4>(M);
    case 8: return hydro_intern_nl<S, 8>(M);
    case 12: return hydro_intern_nl<S, 12>(M);
    case 16: return hydro_intern_nl<S, 16>(M);
    case 20: return hydro_intern_nl<S, 20>(M);
    case 24: return hydro_intern_nl<S, 24>(M);
    default: return hydro_intern_nl<S, 32>(M);
    }
}
int launch_hydro_intern(Model &M) { return M.prec == SHYHIP_FP64 ? hydro_intern_t<double>(M) : hydro_intern_t<float>(M); }

template <typename S>
static int zeta_t(Model &M) {
    auto a = make_args<S>(M);
    a.sv_ell = (S *)M.sv.ell_a;
    a.sv_a32 = (float *)M.sv.a32;
    a.sv_zprev = (float *)M.sv.zprev;
    a.sv_have_prev = M.sv.have_prev ? 1 : 0;
    a.sv_ctrl = M.sv.ctrl;
    LAUNCH(M, k_zeta_elem<S>, nblocks(M.nel), a);
    LAUNCH(M, k_zeta_rows<S>, nblocks(M.nkn), a);
    M.sv.fed = true;
    if (M.sv.zprev) M.sv.have_prev = true;
    return 0;
}
int launch_zeta_assemble(Model &M) { return M.prec == SHYHIP_FP64 ? zeta_t<double>(M) : zeta_t<float>(M); }

template <typename S>
static int sysget_t(Model &M, int bnewtime, const int *gate, const double *dx) {
    auto a = make_args<S>(M);
    LAUNCH(M, k_system_get<S>, nblocks(M.nkn), a, M.d_err + 1, gate, dx);
    LAUNCH(M, k_setweg<S>, nblocks(M.nel), a, 1, bnewtime, M.d_err + 1, gate);
    return 0;
}
int launch_system_get_setweg1(Model &M, int bnewtime, const int *gate, const double *dx) {
    return M.prec == SHYHIP_FP64 ? sysget_t<double>(M, bnewtime, gate, dx) : sysget_t<float>(M, bnewtime, gate, dx);
}

template <typename S>
static int final_t(Model &M) {
    auto a = make_args<S>(M);
    if (M.has_coriolis) LAUNCH(M, (k_transports_final<S, true>), nblocks(M.nel), a);
    else LAUNCH(M, (k_transports_final<S, false>), nblocks(M.nel), a);
    return 0;
}
int launch_transports_final(Model &M) { return M.prec == SHYHIP_FP64 ? final_t<double>(M) : final_t<float>(M); }

template <typename S>
static int setuvd_t(Model &M) {
    auto a = make_args<S>(M);
    LAUNCH(M, k_setzev_setuvd<S>, nblocks(M.nel), a);
    return 0;
}
int launch_setzev_setuvd(Model &M) { return M.prec == SHYHIP_FP64 ? setuvd_t<double>(M) : setuvd_t<float>(M); }

template <typename S, int NL>
static int setdepth_nl(Model &M, bool old_level, bool node_only) {
    auto a = make_args<S>(M);
    const S *zen = old_level ? a.zeo : a.zen;
    S *hden = old_level ? a.hdeo : a.hden;
    S *hdkn = old_level ? a.hdko : a.hdkn;
    if (!node_only) LAUNCH(M, k_setdepth_elem<S>, nblocks(M.nel), a, zen, hden);
    LAUNCH(M, (k_setdepth_node<S, NL>), nblocks(M.nkn), a, zen, hdkn);
    return 0;
}
template <typename S>
static int setdepth_t(Model &M, bool o, bool n) {
    switch (M.nl_tmpl) {
    case 1: return setdepth_nl<S, 1>(M, o, n);
    case 2: return setdepth_nl<S, 2>(M, o, n);
    case 4: return setdepth_nl<S, 4>(M, o, n);
    case 8: return setdepth_nl<S, 8>(M, o, n);
    case 12: return setdepth_nl<S, 12>(M, o, n);
    case 16: return setdepth_nl<S, 16>(M, o, n);
    case 20: return setdepth_nl<S, 20>(M, o, n);
    case 24: return setdepth_nl<S, 24>(M, o, n);
    default: return setdepth_nl<S, 32>(M, o, n);
    }
}
int launch_setdepth(Model &M, bool old_level) {
    return M.prec == SHYHIP_FP64 ? setdepth_t<double>(M, old_level, false) : setdepth_t<float>(M, old_level, false);
}
int launch_setdepth_node(Model &M) {
    return M.prec == SHYHIP_FP64 ? setdepth_t<double>(M, false, true) : setdepth_t<float>(M, false, true);
}

template <typename S>
static int after_solve_t(Model &M) {
    auto a = make_args<S>(M);
    a.iweg_snap = M.iweg_snap;
    if (M.has_coriolis) LAUNCH(M, (k_elem_after_solve<S, true>), nblocks(M.nel), a);
    else LAUNCH(M, (k_elem_after_solve<S, false>), nblocks(M.nel), a);
    return 0;
}
int launch_elem_after_solve(Model &M) {
    return M.prec == SHYHIP_FP64 ? after_solve_t<double>(M) : after_solve_t<float>(M);
}

template <typename S, int NL>
static int vertical_nl(Model &M) {
    auto a = make_args<S>(M);
    LAUNCH(M, (k_hydro_vertical<S, NL>), (M.nkn + VertTile<NL>::TN - 1) / VertTile<NL>::TN, a);
    return 0;
}
template <typename S>
static int vertical_t(Model &M) {
    if (M.nlv == 1) {
        auto a = make_args<S>(M);
        const size_t n = (size_t)(M.nlvdi + 1) * M.nkn;
        LAUNCH(M, k_fill<S>, (int)((n + kThreads - 1) / kThreads), a.wln, (S)0, n);
        return 0;
    }
    switch (M.nl_tmpl) {
    case 2: return vertical_nl<S, 2>(M);
    case 4: return vertical_nl<S, 4>(M);
    case 8: return vertical_nl<S, 8>(M);
    case 12: return vertical_nl<S, 12>(M);
    case 16: return vertical_nl<S, 16>(M);
    case 20: return vertical_nl<S, 20>(M);
    case 24: return vertical_nl<S, 24>(M);
    default: return vertical_nl<S, 32>(M);
    }
}
int launch_hydro_vertical(Model &M) { return M.prec == SHYHIP_FP64 ? vertical_t<double>(M) : vertical_t<float>(M); }

template <typename S>
static int ttov_t(Model &M) {
    auto a = make_args<S>(M);
    LAUNCH(M, k_ttov_uvint<S>, nblocks(M.nel), a);
    return 0;
}
int launch_ttov_uvint(Model &M) { return M.prec == SHYHIP_FP64 ? ttov_t<double>(M) : ttov_t<float>(M); }

template <typename S>
static int copy_old_t(Model &M) {
    // copy_uvz + copy_depth: arrays that are rewritten before being read are
    // exchanged; z and the vertex levels are copied (they are read first).
    std::swap(M.s.utn, M.s.uto);
    std::swap(M.s.vtn, M.s.vto);
    std::swap(M.s.un, M.s.uo);
    std::swap(M.s.vn, M.s.vo);
    std::swap(M.s.wln, M.s.wlo);
    std::swap(M.s.hdkn, M.s.hdko);
    std::swap(M.s.hden, M.s.hdeo);
    auto a = make_args<S>(M);
    LAUNCH(M, k_copy<S>, nblocks(M.nkn), a.zo, (const S *)a.zn, (size_t)M.nkn);
    LAUNCH(M, k_copy<S>, nblocks(3 * M.nel), a.zeo, (const S *)a.zen, (size_t)3 * M.nel);
    return 0;
}
int launch_copy_old(Model &M) { return M.prec == SHYHIP_FP64 ? copy_old_t<double>(M) : copy_old_t<float>(M); }

// new <- spare, old <- new, spare <- old
static void rotate3(void *&nw, void *&old, void *&spare) {
    void *t = old;
    old = nw;
    nw = spare;
    spare = t;
}

template <typename S>
static int step_start_t(Model &M, int bnewtime, bool rot) {
    if (rot) {
        rotate3(M.s.utn, M.s.uto, M.spare_ut);
        rotate3(M.s.vtn, M.s.vto, M.spare_vt);
        rotate3(M.s.wln, M.s.wlo, M.spare_wl);
        rotate3(M.s.hdkn, M.s.hdko, M.spare_hdk);
        rotate3(M.s.hden, M.s.hdeo, M.spare_hde);
        std::swap(M.s.uln, M.uln_b);
        std::swap(M.s.vln, M.vln_b);
    } else {
        std::swap(M.s.utn, M.s.uto);
        std::swap(M.s.vtn, M.s.vto);
        std::swap(M.s.wln, M.s.wlo);
        std::swap(M.s.hdkn, M.s.hdko);
        std::swap(M.s.hden, M.s.hdeo);
    }
    std::swap(M.s.un, M.s.uo);
    std::swap(M.s.vn, M.s.vo);
    auto a = make_args<S>(M);
    LAUNCH(M, k_step_start<S>, nblocks(std::max(M.nel, M.nkn)), a, bnewtime);
    return 0;
}
int launch_step_start(Model &M, int bnewtime, bool rot) {
    return M.prec == SHYHIP_FP64 ? step_start_t<double>(M, bnewtime, rot) : step_start_t<float>(M, bnewtime, rot);
}

template <typename S>
static int set_rzv_t(Model &M, int ibc, float level) {
    if (M.m.nbnd_nodes == 0) return 0;
    LAUNCH(M, k_set_rzv<S>, nblocks(M.m.nbnd_nodes), (S *)M.s.rzv, M.m.bnd_nodes, M.m.bnd_id, M.m.nbnd_nodes, ibc,
           (S)level);
    return 0;
}
int launch_set_rzv(Model &M, int ibc, float level) {
    return M.prec == SHYHIP_FP64 ? set_rzv_t<double>(M, ibc, level) : set_rzv_t<float>(M, ibc, level);
}

int launch_set_rzv_all(Model &M) {
    const int n = M.m.nbnd_nodes;
    if (n == 0) return 0;
    if (M.prec == SHYHIP_FP64)
        LAUNCH(M, k_set_rzv_all<double>, nblocks(n), (double *)M.s.rzv, M.m.bnd_nodes, M.m.bnd_id, M.m.ibtyp, n,
               (const float *)M.d_levels);
    else
        LAUNCH(M, k_set_rzv_all<float>, nblocks(n), (float *)M.s.rzv, M.m.bnd_nodes, M.m.bnd_id, M.m.ibtyp, n,
               (const float *)M.d_levels);
    return 0;
}

int launch_fill(Model &M, void *ptr, double value, size_t n) {
    const int g = (int)((n + kThreads - 1) / kThreads);
    if (M.prec == SHYHIP_FP64) LAUNCH(M, k_fill<double>, g, (double *)ptr, value, n);
    else LAUNCH(M, k_fill<float>, g, (float *)ptr, (float)value, n);
    return 0;
}

template <typename S>
static int init_levels_t(Model &M, float zconst) {
    // setweg(-1), setnod, init_z, setweg(0), setznv, make_new_depth, copy_depth
    if (set_dry_t<S>(M, -1, 1, nullptr)) return 1;
    if (setnod_t<S>(M)) return 1;
    auto a = make_args<S>(M);
    LAUNCH(M, k_fill<S>, nblocks(M.nkn), a.zn, (S)zconst, (size_t)M.nkn);
    LAUNCH(M, k_fill<S>, nblocks(3 * M.nel), a.zen, (S)zconst, (size_t)3 * M.nel);
    if (set_dry_t<S>(M, 0, 0, nullptr)) return 1;
    LAUNCH(M, k_setznv<S>, nblocks(M.nkn), a);
    if (setdepth_t<S>(M, false, false)) return 1;
    return 0;
}
int launch_init_levels(Model &M, float zconst) {
    return M.prec == SHYHIP_FP64 ? init_levels_t<double>(M, zconst) : init_levels_t<float>(M, zconst);
}

template <typename S>
static int volume_t(Model &M, const void *hdk, const void *conc, double *out) {
    auto a = make_args<S>(M);
    const int nb = nblocks(M.nkn);
    double *part = M.sv.part;   // large enough: >= nblocks(nkn)
    LAUNCH(M, k_volume_part<S>, nb, a.nkn, a.ilhkv, a.areakv, (const S *)hdk, (const S *)conc, part);
    hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(kThreads), 0, M.stream, part, nb, M.sv.scal + 15);
    HIPCHK(M, hipGetLastError());
    HIPCHK(M, hipMemcpyAsync(out, M.sv.scal + 15, sizeof(double), hipMemcpyDeviceToHost, M.stream));
    HIPCHK(M, hipStreamSynchronize(M.stream));
    return 0;
}
template <typename S>
static int rhoset_t(Model &M, int niter, hipStream_t st) {
    auto a = make_args<S>(M);
    for (int i = 0; i < niter; i++)
        hipLaunchKernelGGL(k_rhoset<S>, dim3(nblocks(M.nkn)), dim3(kThreads), 0, st, a,
                           (const S *)M.s.tracer[M.bcl_temp], (const S *)M.s.tracer[M.bcl_salt], (S *)M.s.rhov);
    HIPCHK(M, hipGetLastError());
    return 0;
}
int launch_rhoset(Model &M, int niter, hipStream_t st) {
    if (!M.bcl_on) return 0;
    return M.prec == SHYHIP_FP64 ? rhoset_t<double>(M, niter, st) : rhoset_t<float>(M, niter, st);
}

// ------------------------------------------------------------------ allocation-free volume / mass
// The partial sums of the allocation-free variant of the volume and tracer
// mass sums live in this array, one copy per GPU, held until the process
// ends.  Host code never uses its name as an address: the address comes from
// hipGetSymbolAddress, once per device (the first call on that device), and
// is kept per device.
constexpr int kNoallocParts = 65536;                 // 512 KiB: nkn <= 65536 * kThreads
static __device__ __align__(256) double scratch[kNoallocParts];

static std::mutex g_scratch_mu;
static void *g_scratch_ptr[64] = {};

int noalloc_masscont(int dtype, int nkn, const int *ilhkv, const void *hdk, const void *areakv, const void *conc,
                     double *out, hipStream_t st, std::string &err) {
    if (dtype != SHYHIP_F32 && dtype != SHYHIP_F64) { err = "dtype must be SHYHIP_F32 or SHYHIP_F64"; return 1; }
    if (nkn <= 0 || !ilhkv || !hdk || !areakv || !out) { err = "invalid arguments"; return 1; }
    const int nb = nblocks(nkn);
    if (nb > kNoallocParts) {
        err = "nkn = " + std::to_string(nkn) + " exceeds the allocation-free limit of " +
              std::to_string((long long)kNoallocParts * kThreads) + " nodes";
        return 1;
    }
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) { err = "no current HIP device"; return 1; }
    double *part = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if (!g_scratch_ptr[dev]) {
            void *ptr = nullptr;
            const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !ptr) {
                err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
                return 1;
            }
            g_scratch_ptr[dev] = ptr;
        }
        part = (double *)g_scratch_ptr[dev];
    }
    if (dtype == SHYHIP_F64)
        hipLaunchKernelGGL(k_volume_part<double>, dim3(nb), dim3(kThreads), 0, st, nkn, ilhkv, (const double *)areakv,
                           (const double *)hdk, (const double *)conc, part);
    else
        hipLaunchKernelGGL(k_volume_part<float>, dim3(nb), dim3(kThreads), 0, st, nkn, ilhkv, (const float *)areakv,
                           (const float *)hdk, (const float *)conc, part);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(kThreads), 0, st, part, nb, out);
        e = hipGetLastError();
    }
    if (e != hipSuccess) { err = std::string("launch: ") + hipGetErrorString(e); return 1; }
    return 0;
}

int launch_volume(Model &M, int mode, double *vol) {
    const void *hdk = mode > 0 ? M.s.hdkn : M.s.hdko;
    return M.prec == SHYHIP_FP64 ? volume_t<double>(M, hdk, nullptr, vol) : volume_t<float>(M, hdk, nullptr, vol);
}
int launch_tracer_mass(Model &M, int itr, double *mass) {
    return M.prec == SHYHIP_FP64 ? volume_t<double>(M, M.s.hdkn, M.s.tracer[itr], mass)
                                 : volume_t<float>(M, M.s.hdkn, M.s.tracer[itr], mass);
}

}  // namespace shy
