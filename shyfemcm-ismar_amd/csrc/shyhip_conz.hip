// shyhip_conz.hip -- 3D scalar advection-diffusion (conz3d) on gfx950.
//
// Two kernels per substep:
//   k_conz_elem  one work-item per element column: horizontal advective
//                fluxes (upwind rule or the itvd=2 TVD scheme with
//                precomputed upwind interpolation weights) -> 3*fl per vertex
//                and layer, written layer-major to a scratch array.
//   k_conz_node  one work-item per node column: gathers the incident
//                elements in colour order -- the order in which conz3d_omp
//                accumulates them (concentration_omp.f90:243-290) -- adds the
//                vertical diffusion/advection terms of each element
//                (conz3d_element :355-753), then the boundary mass flux and
//                the implicit vertical solve of conz3d_nodes (:757-886).
// Nudging, settling velocity and loads are not part of the supported
// configurations; horizontal diffusion (rkpar) must be zero.
#include <hip/hip_runtime.h>

#include <cmath>

#include "shyhip_internal.h"

namespace shy {

template <typename S>
struct ConzArgs {
    int nkn, nel, nlvdi, nlv, itvd, itvdv, clip;
    double dt, az, ad, aa, rso, rsn;
    float difmol;
    const int *nen3v, *ilhv, *ilhkv, *iextb, *ibnd;
    const double *geo, *dis;
    const int *nkc_ptr, *nkc_el;
    const int *tvd_el, *tvd_kn;
    const double *tvd_w;
    const S *utn, *vtn, *uto, *vto, *uln, *vln;
    const S *hdkn, *hdko, *hden, *hdeo, *wln, *difv, *mflux;
    S *c;            // tracer (in/out)
    const S *cbound; // boundary concentration (may be null)
    S *hbuf;
    int *err;
};

template <typename S>
__device__ inline double cval(const ConzArgs<S> &a, int l, int k) {
    const S v = a.c[(size_t)(l - 1) * a.nkn + k];
    if (a.clip && v < (S)1.e-15) return 0.;
    return (double)v;
}

template <typename S>
__device__ inline S cval_s(const ConzArgs<S> &a, int l, int k) {
    const S v = a.c[(size_t)(l - 1) * a.nkn + k];
    if (a.clip && v < (S)1.e-15) return (S)0;
    return v;
}

// ------------------------------------------------------------------ element phase
template <typename S>
__global__ __launch_bounds__(kThreads) void k_conz_elem(ConzArgs<S> a) {
    using R = S;
    const int ie = blockIdx.x * blockDim.x + threadIdx.x;
    if (ie >= a.nel) return;
    const int nel = a.nel;
    int kn[3];
    double b[3], c[3];
    int iext = 0;
    for (int ii = 0; ii < 3; ii++) {
        kn[ii] = a.nen3v[3 * ie + ii];
        b[ii] = a.geo[(size_t)ii * nel + ie];
        c[ii] = a.geo[(size_t)(3 + ii) * nel + ie];
        if (a.iextb[kn[ii]]) iext++;
    }
    const double azt = 1. - a.az;
    const int ilevel = a.ilhv[ie];
    const bool tvd = a.itvd > 0 && iext == 0;
    const R ajr = (R)(24. * a.geo[(size_t)6 * nel + ie]);
    for (int l = 1; l <= ilevel; l++) {
        const size_t o = (size_t)(l - 1) * nel + ie;
        const double us = a.az * (double)a.utn[o] + azt * (double)a.uto[o];
        const double vs = a.az * (double)a.vtn[o] + azt * (double)a.vto[o];
        double f[3], cl[3], fl[3];
        int itot = 0, isum = 0;
        for (int ii = 0; ii < 3; ii++) {
            f[ii] = us * b[ii] + vs * c[ii];
            if (f[ii] < 0.) { itot++; isum += ii + 1; }
            cl[ii] = cval(a, l, kn[ii]);
        }
        if (itot == 1) {
            for (int ii = 0; ii < 3; ii++) fl[ii] = f[ii] * cl[isum - 1];
        } else if (itot == 2) {
            const int is2 = 6 - isum;
            for (int ii = 0; ii < 3; ii++) fl[ii] = f[ii] * cl[ii];
            fl[is2 - 1] = 0.;
            fl[is2 - 1] = -(fl[0] + fl[1] + fl[2]);
        } else {
            fl[0] = fl[1] = fl[2] = 0.;
        }
        if (tvd) {
            // tvd_fluxes, tvd_admin.f90:673-823
            fl[0] = fl[1] = fl[2] = 0.;
            if (itot >= 1 && itot <= 2) {
                const int itot2 = itot - 1, itot1 = 2 - itot;
                const R u = a.uln[o], v = a.vln[o];
                const int ip = itot2 * (6 - isum) + itot1 * isum;
                for (int ii = 1; ii <= 3; ii++) {
                    if (ii == ip) continue;
                    const int ic = itot2 * ii + itot1 * ip;
                    const int id = itot2 * ip + itot1 * ii;
                    const R fact = (R)(-itot2 + itot1);
                    const R conc = (R)cl[ic - 1], cond = (R)cl[id - 1];
                    const int iop = 6 - (id + ic);
                    const int tet1 = 1 + iop % 3;
                    R dx = (R)((double)ajr * c[iop - 1]);
                    dx = (R)(-2 * (tet1 == id ? 1 : 0)) * dx + dx;
                    R dy = (R)((double)ajr * b[iop - 1]);
                    dy = (R)(-2 * (tet1 == ic ? 1 : 0)) * dy + dy;
                    const R dis = (R)a.dis[(size_t)(iop - 1) * nel + ie];
                    const R vel = (R)fabs((double)(u * dx + v * dy)) / dis;
                    const R alfa = (R)((a.dt * (double)vel) / (double)dis);
                    // upwind value: pair index of (ic,id)
                    const int pidx = 2 * (ic - 1) + ((id == (ic % 3) + 1) ? 0 : 1);
                    R conu = cond;
                    const int up = a.tvd_el[(size_t)pidx * nel + ie];
                    if (up >= 0 && a.ilhv[up] >= l) {
                        double zh = 0.;
                        for (int q = 0; q < 3; q++) {
                            const int kq = a.tvd_kn[((size_t)q * 6 + pidx) * nel + ie];
                            zh = zh + (double)cval_s(a, l, kq) * a.tvd_w[((size_t)q * 6 + pidx) * nel + ie];
                        }
                        conu = (R)zh;
                    }
                    const R grad = cond - conu;
                    R rf;
                    if ((R)fabs((double)(conc - cond)) < (R)1.e-8) rf = -1;
                    else rf = grad / (cond - conc) - (R)1;
                    const double rfd = (double)rf;
                    const R psi = (R)fmax(0., fmax(fmin(1., 2. * rfd), fmin(2., rfd)));
                    const R conf = conc + (R)0.5 * psi * (cond - conc) * ((R)1 - alfa);
                    const R term = (R)((double)(fact * conf) * f[ii - 1]);
                    fl[ic - 1] = fl[ic - 1] - (double)term;
                    fl[id - 1] = fl[id - 1] + (double)term;
                }
            }
        }
        for (int ii = 0; ii < 3; ii++) a.hbuf[((size_t)(l - 1) * 3 + ii) * nel + ie] = (S)(3. * fl[ii]);
    }
}

// ------------------------------------------------------------------ node phase
template <typename S>
__device__ inline double vflux_at(const ConzArgs<S> &a, int k, int lp, int lmax, const double *cl,
                                  const double *hold, double dt, bool btvdv) {
    // vertical_flux_ie (tvd_admin.f90:930-1016) at interface lp of a column
    // truncated at lmax; cl/hold are 1-based column arrays, zero beyond lmax
    if (lp <= 0 || lp >= lmax) return 0.;
    const double w = (double)a.wln[(size_t)lp * a.nkn + k];
    const double conc = (w > 0.) ? cl[lp + 1] : cl[lp];
    double conf = conc;
    if (btvdv) {
        double cond, conu;
        if (w > 0.) {
            cond = cl[lp];
            conu = cond;
            if (lp + 2 <= lmax) conu = cl[lp + 2];
        } else {
            cond = cl[lp + 1];
            conu = cond;
            if (lp - 1 >= 1) conu = cl[lp - 1];
        }
        const double hdis = 0.5 * (hold[lp] + hold[lp + 1]);
        const double alfa = dt * fabs(w) / hdis;
        double rf;
        if (fabs(conc - cond) < (double)1.e-8f) rf = -1.;
        else rf = (cond - conu) / (cond - conc) - 1.;
        const double psi = fmax(0., fmax(fmin(1., 2. * rf), fmin(2., rf)));
        conf = conc + 0.5 * psi * (cond - conc) * (1. - alfa);
    }
    return w * conf;
}

template <typename S, int NL>
__global__ __launch_bounds__(kThreads) void k_conz_node(ConzArgs<S> a) {
    using R = S;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.nkn) return;
    const int nkn = a.nkn, nel = a.nel;
    const double dt = a.dt, ad = a.ad, aa = a.aa, adt = 1. - a.ad, aat = 1. - a.aa;
    const double rso = a.rso, rsn = a.rsn, rsot = 1. - rso, rsnt = 1. - rsn;
    const int lmaxk = a.ilhkv[k];
    const bool btvdv = a.itvdv > 0;
    // node column (1-based, zero padded)
    double ck[NL + 3], hok[NL + 3], hnk[NL + 3];
#pragma unroll
    for (int l = 0; l < NL + 3; l++) { ck[l] = 0.; hok[l] = 0.; hnk[l] = 0.; }
#pragma unroll
    for (int l = 1; l <= NL; l++) {
        if (l <= lmaxk) {
            const size_t o = (size_t)(l - 1) * nkn + k;
            const double hn = (double)a.hdkn[o], ho = (double)a.hdko[o];
            hok[l] = rso * hn + rsot * ho;
            hnk[l] = rsn * hn + rsnt * ho;
            ck[l] = cval(a, l, k);
        }
    }
    double cn[NL], cd[NL], clo[NL], chi[NL];
#pragma unroll
    for (int l = 0; l < NL; l++) { cn[l] = 0.; cd[l] = 0.; clo[l] = 0.; chi[l] = 0.; }
    for (int q = a.nkc_ptr[k]; q < a.nkc_ptr[k + 1]; q++) {
        const int pk = a.nkc_el[q];
        const int ie = pk >> 2, ii = pk & 3;
        const double aj4 = 4. * a.geo[(size_t)6 * nel + ie];
        const int ilev = a.ilhv[ie];
        // element-local column: zero beyond the element's bottom layer
        double cl[NL + 3], hold[NL + 3];
#pragma unroll
        for (int l = 0; l < NL + 3; l++) {
            cl[l] = (l >= 1 && l <= ilev) ? ck[l] : 0.;
            hold[l] = (l >= 1 && l <= ilev) ? hok[l] : 0.;
        }
        double vf_prev = 0.;  // vflux at interface l-1
#pragma unroll
        for (int l = 1; l <= NL; l++) {
            if (l > ilev) break;
            const double pm = (l - 1 >= 1) ? 1. : 0.;
            const double pp = (l + 1 <= ilev) ? 1. : 0.;
            const double hom = hold[l - 1], hoc = hold[l], hop = hold[l + 1];
            const double hnm = (l - 1 >= 1) ? hnk[l - 1] : 0., hnc = hnk[l], hnp = (l + 1 <= ilev) ? hnk[l + 1] : 0.;
            const double rvptop = (double)(R)(a.difv[(size_t)(l - 1) * nkn + k] + (R)a.difmol);
            const double rvpbot = (double)(R)(a.difv[(size_t)l * nkn + k] + (R)a.difmol);
            const double hmotop = 2. * rvptop * pm / (hom + hoc);
            const double hmobot = 2. * rvpbot * pp / (hoc + hop);
            const double hmntop = 2. * rvptop * pm / (hnm + hnc);
            const double hmnbot = 2. * rvpbot * pp / (hnc + hnp);
            const double fd = adt * ((cl[l] - cl[l + 1]) * hmobot - (cl[l - 1] - cl[l]) * hmotop);
            double clc = 0. + ad * (hmntop + hmnbot);
            double clm = 0. - ad * hmntop;
            double clp = 0. - ad * hmnbot;
            double w = (l == 1) ? 0. : (double)a.wln[(size_t)(l - 1) * nkn + k];
            if (w >= 0.) clc += aa * w;
            else clm += aa * w;
            w = (l == ilev) ? 0. : (double)a.wln[(size_t)l * nkn + k];
            if (w > 0.) clp -= aa * w;
            else clc -= aa * w;
            const double vf_here = vflux_at(a, k, l, ilev, cl, hold, dt, btvdv);
            const double fw = aat * (vf_prev - vf_here);
            vf_prev = vf_here;
            const double hf = (double)a.hbuf[((size_t)(l - 1) * 3 + ii) * nel + ie];
            const double cexpl = aj4 * (hoc * cl[l] + dt * (0. + hf - fw - fd));
            cn[l - 1] += cexpl;
            clo[l - 1] += aj4 * dt * clm;
            chi[l - 1] += aj4 * dt * clp;
            cd[l - 1] += aj4 * dt * clc + aj4 * (1. + dt * 0.) * hnc;
        }
    }
    // conz3d_nodes: boundary mass flux, then the column solve
    if (a.mflux) {
#pragma unroll
        for (int l = 1; l <= NL; l++) {
            if (l > lmaxk) break;
            const size_t o = (size_t)(l - 1) * nkn + k;
            double cconz = a.cbound ? (double)a.cbound[o] : 0.;
            const double qflux = (double)a.mflux[o];
            if (qflux < 0. && a.ibnd[k] > 0) cconz = ck[l];
            cn[l - 1] += dt * (qflux * cconz);
        }
    }
    bool bad = false;
    if ((aa == 0. && ad == 0.) || a.nlv == 1) {
#pragma unroll
        for (int l = 1; l <= NL; l++)
            if (l <= lmaxk && cd[l - 1] != 0.) cn[l - 1] = cn[l - 1] / cd[l - 1];
    } else {
        double aux = 1. / cd[0];
        chi[0] *= aux;
        cn[0] *= aux;
#pragma unroll
        for (int l = 2; l <= NL; l++) {
            if (l > lmaxk) break;
            if (cd[l - 1] == 0.) bad = true;
            aux = 1. / (cd[l - 1] - clo[l - 1] * chi[l - 2]);
            chi[l - 1] *= aux;
            cn[l - 1] = (cn[l - 1] - clo[l - 1] * cn[l - 2]) * aux;
        }
#pragma unroll
        for (int l = NL - 1; l >= 1; l--) {
            if (l > lmaxk - 1) continue;
            cn[l - 1] = cn[l - 1] - cn[l] * chi[l - 1];
        }
    }
    if (bad) { atomicCAS(a.err, 0, ERR_CONZ); return; }
    for (int l = 1; l <= a.nlvdi; l++) {
        double v = 0.;
#pragma unroll
        for (int m = 1; m <= NL; m++)
            if (m == l && l <= lmaxk) v = cn[m - 1];
        a.c[(size_t)(l - 1) * nkn + k] = (S)v;
    }
}

template <typename S, int NL>
static int conz_nl(Model &M, int itr, int istot, int isact, int clip) {
    ConzArgs<S> a{};
    a.nkn = M.nkn; a.nel = M.nel; a.nlvdi = M.nlvdi; a.nlv = M.nlv;
    a.itvd = M.p.itvd; a.itvdv = M.p.itvdv; a.clip = clip;
    const double rstot = istot;
    a.rso = (isact - 1) / rstot;
    a.rsn = isact / rstot;
    a.dt = (double)(S)M.p.dt / rstot;
    a.az = M.p.azpar; a.ad = M.p.adpar; a.aa = M.p.aapar; a.difmol = M.p.difmol;
    a.nen3v = M.m.nen3v; a.ilhv = M.m.ilhv; a.ilhkv = M.m.ilhkv; a.iextb = M.m.iextb; a.ibnd = M.m.ibnd;
    a.geo = M.m.geo; a.dis = M.m.dis;
    a.nkc_ptr = M.m.nk_ptr; a.nkc_el = M.m.nkc_el;
    a.tvd_el = M.m.tvd_el; a.tvd_kn = M.m.tvd_kn; a.tvd_w = (const double *)M.m.tvd_w;
    a.utn = (const S *)M.s.utn; a.vtn = (const S *)M.s.vtn; a.uto = (const S *)M.s.uto; a.vto = (const S *)M.s.vto;
    a.uln = (const S *)M.s.uln; a.vln = (const S *)M.s.vln;
    a.hdkn = (const S *)M.s.hdkn; a.hdko = (const S *)M.s.hdko; a.hden = (const S *)M.s.hden;
    a.hdeo = (const S *)M.s.hdeo; a.wln = (const S *)M.s.wln; a.difv = (const S *)M.s.difv;
    a.mflux = M.has_mflux ? (const S *)M.s.mflux : nullptr;
    a.c = (S *)M.s.tracer[itr];
    a.cbound = nullptr;
    a.hbuf = (S *)M.s.hbuf;
    a.err = M.d_err;
    const int ge = (M.nel + kThreads - 1) / kThreads, gn = (M.nkn + kThreads - 1) / kThreads;
    timer_begin(M, T_CONZ_ELEM);
    hipLaunchKernelGGL(k_conz_elem<S>, dim3(ge), dim3(kThreads), 0, M.stream, a);
    timer_end(M);
    HIPCHK(M, hipGetLastError());
    timer_begin(M, T_CONZ_NODE);
    hipLaunchKernelGGL((k_conz_node<S, NL>), dim3(gn), dim3(kThreads), 0, M.stream, a);
    timer_end(M);
    HIPCHK(M, hipGetLastError());
    return 0;
}

template <typename S>
static int conz_t(Model &M, int itr, int istot, int isact, int clip) {
    switch (M.nl_tmpl) {
    case 1: return conz_nl<S, 1>(M, itr, istot, isact, clip);
    case 2: return conz_nl<S, 2>(M, itr, istot, isact, clip);
    case 4: return conz_nl<S, 4>(M, itr, istot, isact, clip);
    case 8: return conz_nl<S, 8>(M, itr, istot, isact, clip);
    case 12: return conz_nl<S, 12>(M, itr, istot, isact, clip);
    case 16: return conz_nl<S, 16>(M, itr, istot, isact, clip);
    case 20: return conz_nl<S, 20>(M, itr, istot, isact, clip);
    case 24: return conz_nl<S, 24>(M, itr, istot, isact, clip);
    default: return conz_nl<S, 32>(M, itr, istot, isact, clip);
    }
}

// uvtopr (transforms.f90:215-270) at the open-boundary nodes followed by
// bndo_setbc = bndo_impbc + bndo_adjbc (bndo.f90:620-890): impose the
// boundary value, and on outflow take the weighted mean of the inner
// neighbours.  One work-item per open-boundary node; it writes only its own
// column and reads only inner nodes.
template <typename S, int NL>
__global__ __launch_bounds__(kThreads) void k_bndo(int nbndo, const int *bk, const S *bnx, const S *bny, const int *bptr,
                                                   const int *bnb, const S *bw, const int *ibnd, const int *ibtyp,
                                                   const int *ilhkv, const int *ilhv, const int *iweg,
                                                   const int *nk_ptr, const int *nk_el, const double *geo, const S *uln,
                                                   const S *vln, S *cv, const S *rbc, int nkn, int nel, int *err) {
    using R = S;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nbndo) return;
    const int k = bk[i];
    if (ibtyp[ibnd[k]] != 1) return;
    R vv[NL], up[NL], vp[NL];
#pragma unroll
    for (int l = 0; l < NL; l++) { vv[l] = 0; up[l] = 0; vp[l] = 0; }
    for (int q = nk_ptr[k]; q < nk_ptr[k + 1]; q++) {
        const int ie = nk_el[q] >> 2;
        if (iweg[ie] != 0) continue;
        const R aj = (R)geo[(size_t)6 * nel + ie];
        const int lmax = ilhv[ie];
#pragma unroll
        for (int l = 1; l <= NL; l++) {
            if (l > lmax) break;
            vv[l - 1] = vv[l - 1] + aj;
            up[l - 1] = up[l - 1] + aj * uln[(size_t)(l - 1) * nel + ie];
            vp[l - 1] = vp[l - 1] + aj * vln[(size_t)(l - 1) * nel + ie];
        }
    }
    const int lmax = ilhkv[k];
    const R flag = (R)-999.;
    if (rbc) {
        for (int l = 1; l <= lmax; l++) {
            const R rb = rbc[(size_t)(l - 1) * nkn + k];
            if (rb != flag) cv[(size_t)(l - 1) * nkn + k] = rb;
        }
    }
    const R dx = bnx[i], dy = bny[i];
#pragma unroll
    for (int l = 1; l <= NL; l++) {
        if (l > lmax) break;
        R u = up[l - 1], v = vp[l - 1];
        if (vv[l - 1] > (R)0) { u = u / vv[l - 1]; v = v / vv[l - 1]; }
        const R scal = dx * u + dy * v;
        const size_t o = (size_t)(l - 1) * nkn + k;
        const bool binside = scal <= (R)0 || cv[o] == flag;
        R bc = 0, tw = 0;
        int ntbc = 0;
        if (binside) {
            for (int j = bptr[i]; j < bptr[i + 1]; j++) {
                const int ko = bnb[j];
                if (l <= ilhkv[ko]) {
                    ntbc++;
                    tw = tw + bw[j];
                    bc = bc + bw[j] * cv[(size_t)(l - 1) * nkn + ko];
                }
            }
        } else {
            ntbc = bptr[i + 1] - bptr[i];
            tw = 1;
            bc = cv[o];
        }
        if (ntbc > 0) cv[o] = bc / tw;
        else if (l > 1) cv[o] = cv[o - nkn];
        else atomicCAS(err, 0, ERR_CONZ);
    }
}

template <typename S, int NL>
static int bndo_nl(Model &M, int itr) {
    if (M.m.nbndo == 0) return 0;
    const int g = (M.m.nbndo + kThreads - 1) / kThreads;
    hipLaunchKernelGGL((k_bndo<S, NL>), dim3(g), dim3(kThreads), 0, M.stream, M.m.nbndo, M.m.bndo_k,
                       (const S *)M.m.bndo_nx, (const S *)M.m.bndo_ny, M.m.bndo_ptr, M.m.bndo_nb,
                       (const S *)M.m.bndo_w, M.m.ibnd, M.m.ibtyp, M.m.ilhkv, M.m.ilhv, M.s.iweg, M.m.nk_ptr,
                       M.m.nk_el, M.m.geo, (const S *)M.s.uln, (const S *)M.s.vln, (S *)M.s.tracer[itr],
                       (const S *)M.s.tracer_bc[itr], M.nkn, M.nel, M.d_err);
    HIPCHK(M, hipGetLastError());
    return 0;
}

template <typename S>
static int bndo_t(Model &M, int itr) {
    switch (M.nl_tmpl) {
    case 1: return bndo_nl<S, 1>(M, itr);
    case 2: return bndo_nl<S, 2>(M, itr);
    case 4: return bndo_nl<S, 4>(M, itr);
    case 8: return bndo_nl<S, 8>(M, itr);
    case 12: return bndo_nl<S, 12>(M, itr);
    case 16: return bndo_nl<S, 16>(M, itr);
    case 20: return bndo_nl<S, 20>(M, itr);
    case 24: return bndo_nl<S, 24>(M, itr);
    default: return bndo_nl<S, 32>(M, itr);
    }
}

int launch_bndo(Model &M, int itr) { return M.prec == SHYHIP_FP64 ? bndo_t<double>(M, itr) : bndo_t<float>(M, itr); }

int launch_conz3d(Model &M, int itr, int istot, int isact, int clip_small) {
    return M.prec == SHYHIP_FP64 ? conz_t<double>(M, itr, istot, isact, clip_small)
                                 : conz_t<float>(M, itr, istot, isact, clip_small);
}

}  // namespace shy
