// shyhip_conz.hip -- 3D scalar advection-diffusion (conz3d) on gfx950.
//
// Two kernels per substep:
//   k_conz_elem  one workgroup per tile of elements, one work-item per
//                element-layer: horizontal advective fluxes (upwind rule or
//                the itvd=2 TVD scheme with precomputed upwind interpolation
//                weights) -> 3*fl per vertex, written layer-major to a
//                scratch array.
//   k_conz_node  one workgroup per tile of consecutive nodes.  Phase 1, one
//                work-item per node-layer: the vertical diffusion/advection
//                terms of the node column (they depend on an element only
//                through its bottom layer, so three variants cover every
//                element), then the incident elements gathered in colour
//                order -- the order in which conz3d_omp accumulates them
//                (concentration_omp.f90:243-290) -- with the element terms of
//                conz3d_element (:355-753), and the boundary mass flux of
//                conz3d_nodes; the column coefficients go to LDS.  Phase 2,
//                one work-item per node: the implicit vertical solve
//                (:757-886) and the write-back.
// Nudging, settling velocity and loads are not part of the supported
// configurations; horizontal diffusion (rkpar) must be zero.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>

#include "shyhip_internal.h"

namespace shy {

// Up to kMaxTracers tracers are transported by one launch of each phase:
// they share every hydrodynamic load, the element flux coefficients and the
// node column matrix; each keeps its own arithmetic, identical to its
// transport alone.
constexpr int kMaxTracers = 2;

template <typename S>
struct ConzArgs {
    int nkn, nel, nlvdi, nlv, itvd, itvdv;
    int clip[kMaxTracers];
    double dt, az, ad, aa, rso, rsn;
    float difmol;
    const int *nen3v, *ilhv, *ilhkv, *iextb, *ibnd;
    const double *geo, *dis;
    const int *nkc_ptr, *nkc_el;
    const int *slot_of, *slot_lev;   // colour-order list positions (see Mesh)
    const double *slot_aj4;
    int nslot;                       // 3 * nel
    const int *tvd_el, *tvd_kn, *tvd_lev, *elem_iext;
    const double *tvd_w;
    const S *utn, *vtn, *uto, *vto, *uln, *vln;
    const S *hdkn, *hdko, *hden, *hdeo, *wln, *difv, *mflux;
    S *c[kMaxTracers];               // tracers (in/out)
    const S *cbound[kMaxTracers];    // source concentrations (may be null)
    S *hbuf;                         // element terms [(l-1)*NT + t][slot_of[ii][ie]]
    int *err;
    // ensembles whose members run different numbers of substeps: member
    // istot (null: all run istot), node/element -> member, this substep and
    // the step as (double)(S)dt; members already past their last substep
    // are left alone
    const int *mem_istot, *node_mem, *elem_mem;
    int isact;
    double dtp;
    int node_tile;                   // nodes per workgroup of the node phase
};

template <typename S>
__device__ inline double cval(const ConzArgs<S> &a, int t, int l, int k) {
    const S v = a.c[t][(size_t)(l - 1) * a.nkn + k];
    if (a.clip[t] && v < (S)1.e-15) return 0.;
    return (double)v;
}

// ------------------------------------------------------------------ element phase
// One workgroup per tile of kElemTile consecutive elements.  The tile's
// geometry, levels and TVD upwind tables are staged in LDS once, four
// work-items per element, every load independent of the others (the upwind
// elements' levels and the elements' external-node counts are precomputed
// tables).  Then one work-item per element-layer of the tile: the layer's
// transports and every tracer's node concentrations, the flux coefficients
// once, and per tracer the upwind rule or the TVD limiter.
constexpr int kElemTile = 64;

// v[i] for a run-time i in 0..2 without indexing a register array
template <typename T>
__device__ inline T sel3(const T (&v)[3], int i) {
    return i == 0 ? v[0] : (i == 1 ? v[1] : v[2]);
}

struct ElemTileLds {
    double b[3][kElemTile], c[3][kElemTile], dis[3][kElemTile];
    double tw[3][6][kElemTile];
    int kn[3][kElemTile], so[3][kElemTile], tkn[3][6][kElemTile], tel[6][kElemTile], tlev[6][kElemTile];
    int ilev[kElemTile], iext[kElemTile];
    double ajr[kElemTile];
};

template <typename S, int NT>
__global__ __launch_bounds__(kThreads) void k_conz_elem(ConzArgs<S> a) {
    using R = S;
    __shared__ ElemTileLds T;
    const int nel = a.nel, nkn = a.nkn, e0 = xcd_block() * kElemTile;
    const int ne = min(kElemTile, nel - e0);
    const bool use_tvd = a.itvd > 0;
    __shared__ int lmax_tile;
    if (threadIdx.x == 0) lmax_tile = 0;
    __syncthreads();
    {
        const int e = threadIdx.x % kElemTile, part = threadIdx.x / kElemTile;   // kThreads = 4 * kElemTile
        if (e < ne) {
            const int ie = e0 + e;
            if (part == 0) {
                for (int ii = 0; ii < 3; ii++) {
                    T.kn[ii][e] = a.nen3v[3 * ie + ii];
                    T.so[ii][e] = a.slot_of[(size_t)ii * nel + ie];
                    T.b[ii][e] = a.geo[(size_t)ii * nel + ie];
                    T.c[ii][e] = a.geo[(size_t)(3 + ii) * nel + ie];
                    T.dis[ii][e] = a.dis[(size_t)ii * nel + ie];
                }
                T.iext[e] = a.elem_iext[ie];
                T.ajr[e] = a.geo[(size_t)6 * nel + ie];
                const int lv = a.ilhv[ie];
                T.ilev[e] = lv;
                atomicMax(&lmax_tile, lv);
            } else if (use_tvd) {
                for (int p = 2 * (part - 1); p < 2 * part; p++) {
                    T.tel[p][e] = a.tvd_el[(size_t)p * nel + ie];
                    T.tlev[p][e] = a.tvd_lev[(size_t)p * nel + ie];
                    for (int q = 0; q < 3; q++) {
                        T.tkn[q][p][e] = a.tvd_kn[((size_t)q * 6 + p) * nel + ie];
                        T.tw[q][p][e] = a.tvd_w[((size_t)q * 6 + p) * nel + ie];
                    }
                }
            }
        }
    }
    __syncthreads();
    // the tile's element-layers that exist (members still substepping),
    // layer by layer with the elements ascending in each, listed by the
    // first wavefront (work-item e owns element e of the tile)
    static_assert(kElemTile == 64, "one wavefront lists the tile");
    __shared__ unsigned short items[kElemTile * 32];      // e | l << 8
    __shared__ int nitems_s;
    if (threadIdx.x < kElemTile) {
        const int e = threadIdx.x;
        int lev = e < ne ? T.ilev[e] : 0;
        if (lev > 0 && a.mem_istot && a.isact > a.mem_istot[a.elem_mem[e0 + e]]) lev = 0;
        const unsigned long long below = e == 0 ? 0ull : (~0ull >> (64 - e));
        int off = 0;
        for (int l = 1; l <= lmax_tile; l++) {
            const unsigned long long m = __ballot(lev >= l);
            if (lev >= l) items[off + __popcll(m & below)] = (unsigned short)(e | (l << 8));
            off += __popcll(m);
        }
        if (e == 0) nitems_s = off;
    }
    __syncthreads();
    const int nitems = nitems_s;
    const double azt = 1. - a.az;
    for (int it = threadIdx.x; it < nitems; it += kThreads) {
        const int e = items[it] & 0xff, l = items[it] >> 8;
        const int ie = e0 + e;
        // the layer's transports, velocities and node concentrations
        const size_t o = (size_t)(l - 1) * nel + ie;
        const S un = a.utn[o], uo = a.uto[o], vn = a.vtn[o], vo = a.vto[o];
        const bool tvd = use_tvd && T.iext[e] == 0;
        S ul = 0, vl = 0;
        if (tvd) { ul = a.uln[o]; vl = a.vln[o]; }
        S cr[NT][3];
#pragma unroll
        for (int t = 0; t < NT; t++)
            for (int ii = 0; ii < 3; ii++) cr[t][ii] = a.c[t][(size_t)(l - 1) * nkn + T.kn[ii][e]];
        auto clip = [&](int t, S v) -> S { return (a.clip[t] && v < (S)1.e-15) ? (S)0 : v; };
        // upwind rule (the flux coefficients are the tracers' common part)
        const double us = a.az * (double)un + azt * (double)uo;
        const double vs = a.az * (double)vn + azt * (double)vo;
        int itot = 0, isum = 0;
        double f[3];
        for (int ii = 0; ii < 3; ii++) {
            f[ii] = us * T.b[ii][e] + vs * T.c[ii][e];
            if (f[ii] < 0.) { itot++; isum += ii + 1; }
        }
        double fl[NT][3];
#pragma unroll
        for (int t = 0; t < NT; t++) {
            double cl[3];
            for (int ii = 0; ii < 3; ii++) cl[ii] = (double)clip(t, cr[t][ii]);
            if (itot == 1) {
                const double cu1 = sel3(cl, isum - 1);
                for (int ii = 0; ii < 3; ii++) fl[t][ii] = f[ii] * cu1;
            } else if (itot == 2) {
                const int is2 = 6 - isum;
                for (int ii = 0; ii < 3; ii++) fl[t][ii] = (ii == is2 - 1) ? 0. : f[ii] * cl[ii];
                const double sum = fl[t][0] + fl[t][1] + fl[t][2];
                for (int ii = 0; ii < 3; ii++)
                    if (ii == is2 - 1) fl[t][ii] = -sum;
            } else {
                fl[t][0] = fl[t][1] = fl[t][2] = 0.;
            }
        }
        // TVD fluxes (tvd_fluxes, tvd_admin.f90:673-823): the pairs of the
        // flow direction, then each tracer's upwind values and limiter
        if (tvd) {
#pragma unroll
            for (int t = 0; t < NT; t++) fl[t][0] = fl[t][1] = fl[t][2] = 0.;
            if (itot >= 1 && itot <= 2) {
                const int itot2 = itot - 1, itot1 = 2 - itot;
                const int ip = itot2 * (6 - isum) + itot1 * isum;
                const R ajr = (R)(24. * T.ajr[e]);
                const double dte = a.mem_istot ? a.dtp / (double)a.mem_istot[a.elem_mem[ie]] : a.dt;
                for (int k = 0; k < 2; k++) {
                    // the vertices other than ip, ascending
                    const int ii = k == 0 ? (ip == 1 ? 2 : 1) : (ip == 3 ? 2 : 3);
                    const int ic = itot2 * ii + itot1 * ip;
                    const int id = itot2 * ip + itot1 * ii;
                    const int pidx = 2 * (ic - 1) + ((id == (ic % 3) + 1) ? 0 : 1);
                    const bool pup = T.tel[pidx][e] >= 0 && T.tlev[pidx][e] >= l;
                    S cu[NT][3];
                    if (pup)
#pragma unroll
                        for (int t = 0; t < NT; t++)
                            for (int q = 0; q < 3; q++) cu[t][q] = a.c[t][(size_t)(l - 1) * nkn + T.tkn[q][pidx][e]];
                    const R fact = (R)(-itot2 + itot1);
                    const int iop = 6 - (id + ic);
                    const int tet1 = 1 + iop % 3;
                    R dx = (R)((double)ajr * T.c[iop - 1][e]);
                    dx = (R)(-2 * (tet1 == id ? 1 : 0)) * dx + dx;
                    R dy = (R)((double)ajr * T.b[iop - 1][e]);
                    dy = (R)(-2 * (tet1 == ic ? 1 : 0)) * dy + dy;
                    const R dis = (R)T.dis[iop - 1][e];
                    const R vel = (R)fabs((double)(ul * dx + vl * dy)) / dis;
                    const R alfa = (R)((dte * (double)vel) / (double)dis);
#pragma unroll
                    for (int t = 0; t < NT; t++) {
                        const R conc = clip(t, sel3(cr[t], ic - 1)), cond = clip(t, sel3(cr[t], id - 1));
                        R conu = cond;
                        if (pup) {
                            double zh = 0.;
                            for (int q = 0; q < 3; q++) zh = zh + (double)clip(t, cu[t][q]) * T.tw[q][pidx][e];
                            conu = (R)zh;
                        }
                        const R grad = cond - conu;
                        R rf;
                        if ((R)fabs((double)(conc - cond)) < (R)1.e-8) rf = -1;
                        else rf = grad / (cond - conc) - (R)1;
                        const double rfd = (double)rf;
                        const R psi = (R)fmax(0., fmax(fmin(1., 2. * rfd), fmin(2., rfd)));
                        const R conf = conc + (R)0.5 * psi * (cond - conc) * ((R)1 - alfa);
                        const R term = (R)((double)(fact * conf) * sel3(f, ii - 1));
                        for (int q = 0; q < 3; q++) {
                            if (q == ic - 1) fl[t][q] = fl[t][q] - (double)term;
                            if (q == id - 1) fl[t][q] = fl[t][q] + (double)term;
                        }
                    }
                }
            }
        }
#pragma unroll
        for (int t = 0; t < NT; t++)
            for (int ii = 0; ii < 3; ii++)
                a.hbuf[((size_t)(l - 1) * NT + t) * a.nslot + T.so[ii][e]] = (S)(3. * fl[t][ii]);
    }
}

// ------------------------------------------------------------------ node phase
// Vertical flux at interface lp of a column truncated at lmax
// (vertical_flux_ie, tvd_admin.f90:930-1016).  cl and hold are windows of
// the column around layer l (entry x - l + 2 holds layer x, x = l-2..l+2),
// zero beyond lmax; w is the vertical velocity at the interface.
__device__ inline double vflux_win(int lp, int l, int lmax, double w, const double *cl, const double *hold, double dt,
                                   bool btvdv) {
    if (lp <= 0 || lp >= lmax) return 0.;
    const int o = 2 - l;       // window offset: layer x at cl[x + o]
    const double conc = (w > 0.) ? cl[lp + 1 + o] : cl[lp + o];
    double conf = conc;
    if (btvdv) {
        double cond, conu;
        if (w > 0.) {
            cond = cl[lp + o];
            conu = cond;
            if (lp + 2 <= lmax) conu = cl[lp + 2 + o];
        } else {
            cond = cl[lp + 1 + o];
            conu = cond;
            if (lp - 1 >= 1) conu = cl[lp - 1 + o];
        }
        const double hdis = 0.5 * (hold[lp + o] + hold[lp + 1 + o]);
        const double alfa = dt * fabs(w) / hdis;
        double rf;
        if (fabs(conc - cond) < (double)1.e-8f) rf = -1.;
        else rf = (cond - conu) / (cond - conc) - 1.;
        const double psi = fmax(0., fmax(fmin(1., 2. * rf), fmin(2., rf)));
        conf = conc + 0.5 * psi * (cond - conc) * (1. - alfa);
    }
    return w * conf;
}

// Nodes per workgroup of the node phase (a.node_tile, a power of two).  The
// default was the fastest measured on the 12-layer lagoon (tiles of 8 to
// 64 nodes): phase 2's serial column solves are then shared by the most
// nodes for the LDS a workgroup holds.
static int conz_node_tile(int nlv, int requested) {
    if (requested > 0) return requested;
    return nlv <= 16 ? 64 : 32;
}
// LDS of one node-phase workgroup: right-hand sides of NT tracers and the
// three diagonals of the column matrix, [NL][TN] each, and the list of the
// tile's node-layers
static size_t conz_node_lds(int NL, int NT, int TN) {
    return (size_t)(NT + 3) * NL * TN * sizeof(double) + (size_t)NL * TN * sizeof(int);
}

template <typename S, int NL, int NT>
__global__ __launch_bounds__(kThreads) void k_conz_node(ConzArgs<S> a) {
    using R = S;
    const int TN = a.node_tile;
    extern __shared__ double node_lds[];
    double *const cnP = node_lds, *const cdP = cnP + NT * NL * TN, *const cloP = cdP + NL * TN,
                  *const chiP = cloP + NL * TN;
    // entry (layer index l0, node kk of the tile) of a [NL][TN] array
    auto at = [TN](double *p, int l0, int kk) -> double & { return p[l0 * TN + kk]; };
    const int nkn = a.nkn;
    const int k0 = xcd_block() * TN;
    const double ad = a.ad, aa = a.aa, adt = 1. - a.ad, aat = 1. - a.aa;
    const bool btvdv = a.itvdv > 0;
    const int nlv = min(a.nlv, NL);

    // The tile's node-layers that exist (l <= ilhkv of the node, members
    // still substepping), layer by layer with the nodes ascending in each:
    // phase 1 takes exactly these, so no work-item idles below a node's
    // bottom and a layer's loads stay in node order.  Work-item kk < TN owns
    // node kk; each wavefront counts its nodes per layer with a ballot.
    constexpr int kWaves = kThreads / 64;
    __shared__ int wpos[kWaves][NL + 1];
    __shared__ int nitems_s;
    int *const items = (int *)(chiP + NL * TN);     // kk | l << 16
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nwv = (TN + 63) / 64;
    int lmk = 0;
    if ((int)threadIdx.x < TN && k0 + (int)threadIdx.x < nkn) {
        const int k = k0 + threadIdx.x;
        lmk = min(a.ilhkv[k], nlv);
        if (a.mem_istot && a.isact > a.mem_istot[a.node_mem[k]]) lmk = 0;
    }
    if (wv < nwv)
        for (int l = 1; l <= nlv; l++) {
            const unsigned long long m = __ballot(lmk >= l);
            if (lane == 0) wpos[wv][l] = __popcll(m);
        }
    __syncthreads();
    if (threadIdx.x == 0) {
        int off = 0;
        for (int l = 1; l <= nlv; l++)
            for (int w = 0; w < nwv; w++) {
                const int c = wpos[w][l];
                wpos[w][l] = off;
                off += c;
            }
        nitems_s = off;
    }
    __syncthreads();
    if (wv < nwv) {
        const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
        for (int l = 1; l <= nlv; l++) {
            const unsigned long long m = __ballot(lmk >= l);
            if (lmk >= l) items[wpos[wv][l] + __popcll(m & below)] = (int)threadIdx.x | (l << 16);
        }
    }
    __syncthreads();
    const int nitems = nitems_s;

    // phase 1: one work-item per node-layer
    for (int it = threadIdx.x; it < nitems; it += blockDim.x) {
        const int kk = items[it] & 0xffff, l = items[it] >> 16, k = k0 + kk;
        const int lmaxk = a.ilhkv[k];
        double dt = a.dt, rso = a.rso, rsn = a.rsn;
        if (a.mem_istot) {
            const int ist = a.mem_istot[a.node_mem[k]];
            const double rstot = ist;
            rso = (a.isact - 1) / rstot;
            rsn = a.isact / rstot;
            dt = a.dtp / rstot;
        }
        const double rsot = 1. - rso, rsnt = 1. - rsn;
        // incident elements in colour order: list positions q0 .. q0+cnt-1
        const int q0 = a.nkc_ptr[k], cnt = a.nkc_ptr[k + 1] - q0;
        // node column around l (zero outside 1..lmaxk)
        double ckw[NT][5], hokw[5], hnkw[5];
        for (int j = 0; j < 5; j++) {
            const int x = l - 2 + j;
            hokw[j] = 0.; hnkw[j] = 0.;
#pragma unroll
            for (int t = 0; t < NT; t++) ckw[t][j] = 0.;
            if (x >= 1 && x <= lmaxk) {
#pragma unroll
                for (int t = 0; t < NT; t++) ckw[t][j] = cval(a, t, x, k);
                if (j >= 1 && j <= 3) {
                    const size_t o = (size_t)(x - 1) * nkn + k;
                    const double hn = (double)a.hdkn[o], ho = (double)a.hdko[o];
                    hokw[j] = rso * hn + rsot * ho;
                    hnkw[j] = rsn * hn + rsnt * ho;
                }
            }
        }
        const double rvptop = (double)(R)(a.difv[(size_t)(l - 1) * nkn + k] + (R)a.difmol);
        const double rvpbot = (double)(R)(a.difv[(size_t)l * nkn + k] + (R)a.difmol);
        const double wtop = (l == 1) ? 0. : (double)a.wln[(size_t)(l - 1) * nkn + k];
        const double wbot = (double)a.wln[(size_t)l * nkn + k];
        // element terms of the three column truncations: bottom layer l,
        // l+1, or deeper (nothing below l+2 is read)
        double fdv[NT][3], fwv[NT][3], clcv[3], clmv[3], clpv[3];
        for (int v = 0; v < 3; v++) {
            const int lmax = l + v;
            double holdt[5];
            for (int j = 0; j < 5; j++) {
                const int x = l - 2 + j;
                holdt[j] = (x >= 1 && x <= lmax) ? hokw[j] : 0.;
            }
            const double pm = (l - 1 >= 1) ? 1. : 0.;
            const double pp = (l + 1 <= lmax) ? 1. : 0.;
            const double hom = holdt[1], hoc = holdt[2], hop = holdt[3];
            const double hnm = (l - 1 >= 1) ? hnkw[1] : 0., hnc = hnkw[2], hnp = (l + 1 <= lmax) ? hnkw[3] : 0.;
            const double hmotop = 2. * rvptop * pm / (hom + hoc);
            const double hmobot = 2. * rvpbot * pp / (hoc + hop);
            const double hmntop = 2. * rvptop * pm / (hnm + hnc);
            const double hmnbot = 2. * rvpbot * pp / (hnc + hnp);
            double clc = 0. + ad * (hmntop + hmnbot);
            double clm = 0. - ad * hmntop;
            double clp = 0. - ad * hmnbot;
            double w = wtop;
            if (w >= 0.) clc += aa * w;
            else clm += aa * w;
            w = (l == lmax) ? 0. : wbot;
            if (w > 0.) clp -= aa * w;
            else clc -= aa * w;
            clcv[v] = clc; clmv[v] = clm; clpv[v] = clp;
#pragma unroll
            for (int t = 0; t < NT; t++) {
                double clt[5];
                for (int j = 0; j < 5; j++) {
                    const int x = l - 2 + j;
                    clt[j] = (x >= 1 && x <= lmax) ? ckw[t][j] : 0.;
                }
                fdv[t][v] = adt * ((clt[2] - clt[3]) * hmobot - (clt[1] - clt[2]) * hmotop);
                const double vf_prev = vflux_win(l - 1, l, lmax, wtop, clt, holdt, dt, btvdv);
                const double vf_here = vflux_win(l, l, lmax, wbot, clt, holdt, dt, btvdv);
                fwv[t][v] = aat * (vf_prev - vf_here);
            }
        }
        // incident elements in colour order (conz3d_element's scatter)
        const double hoc = hokw[2], hnc = hnkw[2];
        double cn[NT], cd = 0., clo = 0., chi = 0.;
#pragma unroll
        for (int t = 0; t < NT; t++) cn[t] = 0.;
        auto add = [&](int ilev, double aj4, const double *hf) {
            if (ilev < l) return;
            const int v = min(ilev - l, 2);
#pragma unroll
            for (int t = 0; t < NT; t++) {
                const double cexpl = aj4 * (hoc * ckw[t][2] + dt * (0. + hf[t] - fwv[t][v] - fdv[t][v]));
                cn[t] += cexpl;
            }
            clo += aj4 * dt * clmv[v];
            chi += aj4 * dt * clpv[v];
            cd += aj4 * dt * clcv[v] + aj4 * (1. + dt * 0.) * hnc;
        };
        // the first kIncWidth positions are read unconditionally (clamped
        // to the list), contiguous for the node; the rest one by one
        int slv[kIncWidth];
        double saj[kIncWidth], shf[kIncWidth][NT];
        const size_t hl = (size_t)(l - 1) * NT * a.nslot;
#pragma unroll
        for (int s = 0; s < kIncWidth; s++) {
            const int q = min(q0 + s, a.nslot - 1);
            slv[s] = a.slot_lev[q];
            saj[s] = a.slot_aj4[q];
#pragma unroll
            for (int t = 0; t < NT; t++) shf[s][t] = (double)a.hbuf[hl + (size_t)t * a.nslot + q];
        }
#pragma unroll
        for (int s = 0; s < kIncWidth; s++)
            if (s < cnt) add(slv[s], saj[s], shf[s]);
        for (int q = q0 + kIncWidth; q < q0 + cnt; q++) {
            double hf[NT];
#pragma unroll
            for (int t = 0; t < NT; t++) hf[t] = (double)a.hbuf[hl + (size_t)t * a.nslot + q];
            add(a.slot_lev[q], a.slot_aj4[q], hf);
        }
        // conz3d_nodes: boundary mass flux
        if (a.mflux) {
            const size_t o = (size_t)(l - 1) * nkn + k;
            const double qflux = (double)a.mflux[o];
#pragma unroll
            for (int t = 0; t < NT; t++) {
                double cconz = a.cbound[t] ? (double)a.cbound[t][o] : 0.;
                if (qflux < 0. && a.ibnd[k] > 0) cconz = ckw[t][2];
                cn[t] += dt * (qflux * cconz);
            }
        }
#pragma unroll
        for (int t = 0; t < NT; t++) at(cnP + t * NL * TN, l - 1, kk) = cn[t];
        at(cdP, l - 1, kk) = cd; at(cloP, l - 1, kk) = clo; at(chiP, l - 1, kk) = chi;
    }
    __syncthreads();

    // phase 2: the column solve of conz3d_nodes, one work-item per node; the
    // tracers share the column matrix and its elimination
    const int kk = threadIdx.x, k = k0 + kk;
    if (kk >= TN || k >= nkn) return;
    if (a.mem_istot && a.isact > a.mem_istot[a.node_mem[k]]) return;
    const int lmaxk = a.ilhkv[k];
    auto cn = [&](int t, int l0) -> double & { return at(cnP + t * NL * TN, l0, kk); };
    auto cd = [&](int l0) -> double & { return at(cdP, l0, kk); };
    auto clo = [&](int l0) -> double & { return at(cloP, l0, kk); };
    auto chi = [&](int l0) -> double & { return at(chiP, l0, kk); };
    bool bad = false;
    if ((aa == 0. && ad == 0.) || a.nlv == 1) {
        for (int l = 1; l <= lmaxk; l++)
            if (cd(l - 1) != 0.)
#pragma unroll
                for (int t = 0; t < NT; t++) cn(t, l - 1) = cn(t, l - 1) / cd(l - 1);
    } else {
        // the elimination's chain runs through registers (chip, cnp); LDS
        // supplies the coefficients and keeps the forward results for the
        // back substitution
        const double aux0 = 1. / cd(0);
        double chip = chi(0) * aux0, cnp[NT];
        chi(0) = chip;
#pragma unroll
        for (int t = 0; t < NT; t++) {
            cnp[t] = cn(t, 0) * aux0;
            cn(t, 0) = cnp[t];
        }
        for (int l = 2; l <= lmaxk; l++) {
            const double cdl = cd(l - 1), clol = clo(l - 1), chil = chi(l - 1);
            double cnl[NT];
#pragma unroll
            for (int t = 0; t < NT; t++) cnl[t] = cn(t, l - 1);
            if (cdl == 0.) bad = true;
            const double aux = 1. / (cdl - clol * chip);
            chip = chil * aux;
            chi(l - 1) = chip;
#pragma unroll
            for (int t = 0; t < NT; t++) {
                cnp[t] = (cnl[t] - clol * cnp[t]) * aux;
                cn(t, l - 1) = cnp[t];
            }
        }
#pragma unroll
        for (int t = 0; t < NT; t++) {
            double nx = cnp[t];
            for (int l = lmaxk - 1; l >= 1; l--) {
                nx = cn(t, l - 1) - nx * chi(l - 1);
                cn(t, l - 1) = nx;
            }
        }
    }
    if (bad) { atomicCAS(a.err, 0, ERR_CONZ); return; }
#pragma unroll
    for (int t = 0; t < NT; t++)
        for (int l = 1; l <= a.nlvdi; l++) a.c[t][(size_t)(l - 1) * nkn + k] = (S)(l <= lmaxk ? cn(t, l - 1) : 0.);
}

template <typename S, int NL, int NT>
static int conz_nl(Model &M, const TracerSet &ts, int istot, int isact, hipStream_t st, void *hbuf) {
    ConzArgs<S> a{};
    a.nkn = M.nkn; a.nel = M.nel; a.nlvdi = M.nlvdi; a.nlv = M.nlv;
    a.itvd = M.p.itvd; a.itvdv = M.p.itvdv;
    const double rstot = istot;
    a.rso = (isact - 1) / rstot;
    a.rsn = isact / rstot;
    a.dt = (double)(S)M.p.dt / rstot;
    a.az = M.p.azpar; a.ad = M.p.adpar; a.aa = M.p.aapar; a.difmol = M.p.difmol;
    a.nen3v = M.m.nen3v; a.ilhv = M.m.ilhv; a.ilhkv = M.m.ilhkv; a.iextb = M.m.iextb; a.ibnd = M.m.ibnd;
    a.geo = M.m.geo; a.dis = M.m.dis;
    a.nkc_ptr = M.m.nk_ptr; a.nkc_el = M.m.nkc_el;
    a.slot_of = M.m.slot_of; a.slot_lev = M.m.slot_lev; a.slot_aj4 = M.m.slot_aj4; a.nslot = 3 * M.nel;
    a.tvd_el = M.m.tvd_el; a.tvd_kn = M.m.tvd_kn; a.tvd_w = (const double *)M.m.tvd_w;
    a.tvd_lev = M.m.tvd_lev; a.elem_iext = M.m.elem_iext;
    a.utn = (const S *)M.s.utn; a.vtn = (const S *)M.s.vtn; a.uto = (const S *)M.s.uto; a.vto = (const S *)M.s.vto;
    a.uln = (const S *)M.s.uln; a.vln = (const S *)M.s.vln;
    a.hdkn = (const S *)M.s.hdkn; a.hdko = (const S *)M.s.hdko; a.hden = (const S *)M.s.hden;
    a.hdeo = (const S *)M.s.hdeo; a.wln = (const S *)M.s.wln; a.difv = (const S *)M.s.difv;
    a.mflux = M.has_mflux ? (const S *)M.s.mflux : nullptr;
    for (int t = 0; t < NT; t++) {
        a.c[t] = (S *)M.s.tracer[ts.itr[t]];
        a.cbound[t] = (const S *)ts.cbound[t];
        a.clip[t] = ts.clip[t];
    }
    a.isact = isact;
    a.dtp = (double)(S)M.p.dt;
    if (M.mem_istot_on) {
        a.mem_istot = M.cur_mem_istot; a.node_mem = M.d_node_mem; a.elem_mem = M.d_elem_mem;
    }
    a.hbuf = (S *)hbuf;
    a.err = M.d_err;
    a.node_tile = conz_node_tile(M.nlv, M.p.node_tile);
    const int ge = (M.nel + kElemTile - 1) / kElemTile;
    launch_timed(M, T_CONZ_ELEM, k_conz_elem<S, NT>, dim3(ge), dim3(kThreads), 0, st, a);
    HIPCHK(M, hipGetLastError());
    const int tn = a.node_tile;
    if (conz_node_lds(NL, NT, tn) > 65536)
        return hip_check(M, hipErrorInvalidValue, "conz3d: node_tile too large for the LDS of one workgroup");
    launch_timed(M, T_CONZ_NODE, k_conz_node<S, NL, NT>, dim3((M.nkn + tn - 1) / tn),
                 dim3(std::min(kThreads, (tn * M.nlv + 63) / 64 * 64)), conz_node_lds(NL, NT, tn), st, a);
    HIPCHK(M, hipGetLastError());
    return 0;
}

template <typename S, int NT>
static int conz_t(Model &M, const TracerSet &ts, int istot, int isact, hipStream_t st, void *hbuf) {
    switch (M.nl_tmpl) {
    case 1: return conz_nl<S, 1, NT>(M, ts, istot, isact, st, hbuf);
    case 2: return conz_nl<S, 2, NT>(M, ts, istot, isact, st, hbuf);
    case 4: return conz_nl<S, 4, NT>(M, ts, istot, isact, st, hbuf);
    case 8: return conz_nl<S, 8, NT>(M, ts, istot, isact, st, hbuf);
    case 12: return conz_nl<S, 12, NT>(M, ts, istot, isact, st, hbuf);
    case 16: return conz_nl<S, 16, NT>(M, ts, istot, isact, st, hbuf);
    case 20: return conz_nl<S, 20, NT>(M, ts, istot, isact, st, hbuf);
    case 24: return conz_nl<S, 24, NT>(M, ts, istot, isact, st, hbuf);
    default: return conz_nl<S, 32, NT>(M, ts, istot, isact, st, hbuf);
    }
}

// uvtopr (transforms.f90:215-270) at the open-boundary nodes followed by
// bndo_setbc = bndo_impbc + bndo_adjbc (bndo.f90:620-890): impose the
// boundary value, and on outflow take the weighted mean of the inner
// neighbours.  One wavefront per open-boundary node, one lane per layer; a
// layer with no inner neighbour deep enough takes the final value of the
// layer above, resolved top-down after the lanes finish.  Each wavefront
// writes only its own column and reads only inner nodes.  The node's static
// data (open flag, bottom layer, incident elements in element order with
// their bottom layers and areas, the inner neighbours' bottom layers) comes
// from tables built at set-up, so a lane's loads depend on one another at
// most two deep.
constexpr int kBndoLanes = 64;

template <typename S>
struct BndoArgs {
    int nbndo, nkn, nel;
    const int *bk, *open, *lmax, *eptr, *el, *elev, *bptr, *bnb, *nblev;
    const double *eaj;
    const S *bnx, *bny, *bw, *uln, *vln;
    const int *iweg;
    S *cv[2];               // the tracers of the launch (grid row 0 and 1)
    const S *rbc[2];        // their boundary values (may be null)
    int *err;
    const int *mem_istot, *node_mem;
    int isact;
};

template <typename S>
__global__ __launch_bounds__(kBndoLanes) void k_bndo(BndoArgs<S> a) {
    using R = S;
    __shared__ R val[kBndoLanes];
    __shared__ int status[kBndoLanes];
    const int i = blockIdx.x;
    if (i >= a.nbndo || !a.open[i]) return;
    const int k = a.bk[i];
    if (a.mem_istot && a.isact > a.mem_istot[a.node_mem[k]]) return;
    S *const cv = a.cv[blockIdx.y];
    const S *const rbc = a.rbc[blockIdx.y];
    const int lane = threadIdx.x, l = lane + 1;
    const int lmax = a.lmax[i];
    const int nkn = a.nkn, nel = a.nel;
    const R flag = (R)-999.;
    status[lane] = 0;
    if (l <= lmax) {
        R vv = 0, up = 0, vp = 0;
        for (int q = a.eptr[i]; q < a.eptr[i + 1]; q++) {
            const int ie = a.el[q];
            if (a.iweg[ie] != 0 || l > a.elev[q]) continue;
            const R aj = (R)a.eaj[q];
            vv = vv + aj;
            up = up + aj * a.uln[(size_t)(l - 1) * nel + ie];
            vp = vp + aj * a.vln[(size_t)(l - 1) * nel + ie];
        }
        const size_t o = (size_t)(l - 1) * nkn + k;
        R cur = cv[o];
        if (rbc) {
            const R rb = rbc[o];
            if (rb != flag) cur = rb;
        }
        R u = up, v = vp;
        if (vv > (R)0) { u = u / vv; v = v / vv; }
        const R scal = a.bnx[i] * u + a.bny[i] * v;
        const bool binside = scal <= (R)0 || cur == flag;
        R bc = 0, tw = 0;
        int ntbc = 0;
        if (binside) {
            for (int j = a.bptr[i]; j < a.bptr[i + 1]; j++) {
                if (l <= a.nblev[j]) {
                    ntbc++;
                    tw = tw + a.bw[j];
                    bc = bc + a.bw[j] * cv[(size_t)(l - 1) * nkn + a.bnb[j]];
                }
            }
        } else {
            ntbc = a.bptr[i + 1] - a.bptr[i];
            tw = 1;
            bc = cur;
        }
        if (ntbc > 0) { val[lane] = bc / tw; status[lane] = 1; }
        else status[lane] = l > 1 ? 2 : 3;
    }
    __syncthreads();
    if (lane == 0) {
        for (int m = 1; m < lmax && m < kBndoLanes; m++)
            if (status[m] == 2) val[m] = val[m - 1];
        if (status[0] == 3) atomicCAS(a.err, 0, ERR_CONZ);
    }
    __syncthreads();
    if (l <= lmax && status[0] != 3) cv[(size_t)(l - 1) * nkn + k] = val[lane];
}

template <typename S>
static int bndo_t(Model &M, const TracerSet &ts, hipStream_t st, int isact) {
    if (M.m.nbndo == 0) return 0;
    if (M.nlv > kBndoLanes) return hip_check(M, hipErrorInvalidValue, "bndo: more layers than lanes");
    BndoArgs<S> a{};
    a.nbndo = M.m.nbndo; a.nkn = M.nkn; a.nel = M.nel;
    a.bk = M.m.bndo_k; a.open = M.m.bndo_open; a.lmax = M.m.bndo_lmax; a.eptr = M.m.bndo_eptr; a.el = M.m.bndo_el;
    a.elev = M.m.bndo_elev; a.bptr = M.m.bndo_ptr; a.bnb = M.m.bndo_nb; a.nblev = M.m.bndo_nblev;
    a.eaj = M.m.bndo_eaj;
    a.bnx = (const S *)M.m.bndo_nx; a.bny = (const S *)M.m.bndo_ny; a.bw = (const S *)M.m.bndo_w;
    a.uln = (const S *)M.s.uln; a.vln = (const S *)M.s.vln;
    a.iweg = M.iweg_conz ? M.iweg_conz : M.s.iweg;
    for (int t = 0; t < 2; t++) {
        const int itr = ts.itr[t < ts.n ? t : 0];
        a.cv[t] = (S *)M.s.tracer[itr];
        a.rbc[t] = (const S *)M.s.tracer_bc[itr];
    }
    a.err = M.d_err;
    if (M.mem_istot_on) { a.mem_istot = M.cur_mem_istot; a.node_mem = M.d_node_mem; }
    a.isact = isact;
    hipLaunchKernelGGL(k_bndo<S>, dim3(M.m.nbndo, ts.n), dim3(kBndoLanes), 0, st, a);
    HIPCHK(M, hipGetLastError());
    return 0;
}

int launch_bndo(Model &M, const TracerSet &ts, hipStream_t st, int isact) {
    return M.prec == SHYHIP_FP64 ? bndo_t<double>(M, ts, st, isact) : bndo_t<float>(M, ts, st, isact);
}

int launch_bndo(Model &M, int itr, hipStream_t st, int isact) {
    TracerSet ts;
    ts.n = 1;
    ts.itr[0] = itr;
    return launch_bndo(M, ts, st, isact);
}

int launch_conz3d(Model &M, const TracerSet &ts, int istot, int isact, hipStream_t st, void *hbuf) {
    if (ts.n == 2)
        return M.prec == SHYHIP_FP64 ? conz_t<double, 2>(M, ts, istot, isact, st, hbuf)
                                     : conz_t<float, 2>(M, ts, istot, isact, st, hbuf);
    return M.prec == SHYHIP_FP64 ? conz_t<double, 1>(M, ts, istot, isact, st, hbuf)
                                 : conz_t<float, 1>(M, ts, istot, isact, st, hbuf);
}

int launch_conz3d(Model &M, int itr, int istot, int isact, int clip_small, hipStream_t st, void *hbuf,
                  const void *cbound) {
    TracerSet ts;
    ts.n = 1;
    ts.itr[0] = itr;
    ts.clip[0] = clip_small;
    ts.cbound[0] = cbound;
    return launch_conz3d(M, ts, istot, isact, st, hbuf);
}

// ------------------------------------------------------------------ stability
// scalar_stability (stability.f90:220-260): conzstab (concentration.f90:
// 1573-2097) for one second (one substep) with difmol = 0, no nudging and no
// sinking, then rindex = dt * (index / rstol) and istot = 1 + rindex.  One
// work-item per node-layer gathers its incident elements in element order --
// the order conzstab accumulates them -- with each element's advective term
// for this vertex (the larger of the last out- and in-flux seen over the
// element's vertices up to this one, as conzstab carries them), its explicit
// vertical diffusion and its share of the node volume; point sources add
// their inflow.  Horizontal diffusion (rkpar) is zero on the device path and
// adds nothing.  The node-layer ratios are non-negative, so their IEEE bit
// patterns order as unsigned integers; their maximum is taken across the
// wavefront and the workgroup first (exact in any order), and each workgroup
// then issues one atomic max per member it touches, none when the value
// already held is not smaller.

// maximum of v over the wavefront, in every lane
__device__ inline unsigned long long wave_max_u64(unsigned long long v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned lo = (unsigned)__shfl_xor((int)(unsigned)v, off);
        const unsigned hi = (unsigned)__shfl_xor((int)(unsigned)(v >> 32), off);
        const unsigned long long o = ((unsigned long long)hi << 32) | lo;
        v = o > v ? o : v;
    }
    return v;
}

// atomic max that skips the atomic when the value held is already at least v
// (a stale read is never larger than the value held, so skipping is exact)
__device__ inline void max_into(unsigned long long *p, unsigned long long v) {
    if (v == 0ull) return;
    const unsigned long long cur = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (v > cur) atomicMax(p, v);
}

// the ratio of one node-layer (l, k) as the bits of a non-negative double,
// 0 when it adds nothing
template <typename S>
__device__ inline unsigned long long stab_ratio(int l, int k, int nkn, int nel, const int *ilhv, const int *nk_ptr,
                                                const int *nk_el, const int *nen3v, const double *geo, double az,
                                                double ad, double aa, const S *utn, const S *uto, const S *vtn,
                                                const S *vto, const S *wln, const S *difv, const S *hdeo,
                                                const S *hdkn, const S *hdko, const S *mflux) {
    const double azt = 1. - az, adt = 1. - ad, aat = 1. - aa;
    double chigh = 0., cn = 0., cdiag = 0.;
    double co = 0.;
    if (mflux) {
        const S q = mflux[(size_t)(l - 1) * nkn + k];
        if (q > (S)0) co = (double)q;
    }
    for (int p = nk_ptr[k]; p < nk_ptr[k + 1]; p++) {
        const int ie = nk_el[p] >> 2, iv = nk_el[p] & 3;
        const int ilevel = ilhv[ie];
        if (l > ilevel) continue;
        const double aj4 = 4. * geo[(size_t)6 * nel + ie];
        const size_t le = (size_t)(l - 1) * nel + ie;
        const double us = az * (double)utn[le] + azt * (double)uto[le];
        const double vs = az * (double)vtn[le] + azt * (double)vto[le];
        double coadv = 0., ciadv = 0., cadv = 0.;
        for (int jj = 0; jj <= iv; jj++) {
            const int kj = nen3v[3 * ie + jj];
            const double f = us * geo[(size_t)jj * nel + ie] + vs * geo[(size_t)(3 + jj) * nel + ie];
            double fwin = 0., fwout = 0.;
            const double wt = l > 1 ? (double)wln[(size_t)(l - 1) * nkn + kj] : 0.;
            const double wb = l < ilevel ? (double)wln[(size_t)l * nkn + kj] : 0.;
            if (wt > 0.) fwout = fwout + aat * wt;
            else fwin = fwin - aat * wt;
            if (wb > 0.) fwin = fwin + aat * wb;
            else fwout = fwout - aat * wb;
            const double chadv = aj4 * 3. * f;
            const double cviadv = aj4 * fwin, cvoadv = aj4 * fwout;
            if (chadv < 0.) coadv = cvoadv - chadv;
            else ciadv = cviadv + chadv;
            cadv = coadv > ciadv ? coadv : ciadv;
        }
        chigh = chigh + cadv;
        const double hdv = (double)hdeo[le];
        double hmtop = 0., hmbot = 0.;
        if (l > 1)
            hmtop = 2. * (double)difv[(size_t)(l - 1) * nkn + k] / ((double)hdeo[le - nel] + hdv);
        if (l < ilevel)
            hmbot = 2. * (double)difv[(size_t)l * nkn + k] / (hdv + (double)hdeo[le + nel]);
        cn = cn + aj4 * (adt * (hmtop + hmbot));
        const double hn = (double)hdkn[(size_t)(l - 1) * nkn + k], ho = (double)hdko[(size_t)(l - 1) * nkn + k];
        cdiag = cdiag + aj4 * (ho < hn ? ho : hn);
    }
    if (cdiag > 0.) {
        const double r = ((chigh + 0.) + cn + co) / cdiag;
        if (r > 0.) return (unsigned long long)__double_as_longlong(r);
    }
    return 0ull;
}

template <typename S>
__global__ void __launch_bounds__(kThreads) k_conz_stab(int nkn, int nel, int nlv, const int *ilhkv, const int *ilhv,
                                                        const int *zbound, const int *nk_ptr, const int *nk_el,
                                                        const int *nen3v, const double *geo, double az, double ad,
                                                        double aa, const S *utn, const S *uto, const S *vtn,
                                                        const S *vto, const S *wln, const S *difv, const S *hdeo,
                                                        const S *hdkn, const S *hdko, const S *mflux,
                                                        const int *node_mem, unsigned long long *dmax) {
    // (blocks in launch order: the layer-major index puts the work of the
    // upper layers in the first blocks, which round-robin spreads over XCDs)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long val = 0ull;    // this node-layer's ratio as bits (0: none)
    int key = 0;                      // its member
    if (i < nlv * nkn) {
        const int l = i / nkn + 1, k = i - (l - 1) * nkn;
        if (node_mem) key = node_mem[k];
        if (l <= ilhkv[k] && !zbound[k]) val = stab_ratio(l, k, nkn, nel, ilhv, nk_ptr, nk_el, nen3v, geo, az, ad,
                                                         aa, utn, uto, vtn, vto, wln, difv, hdeo, hdkn, hdko, mflux);
    }
    // the wavefront's maximum when its node-layers belong to one member
    // (always, for a single model); lanes of mixed waves go one by one
    __shared__ unsigned long long wval[kThreads / 64];
    __shared__ int wkey[kThreads / 64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int key0 = __shfl(key, 0);
    if (__all(key == key0)) {
        val = wave_max_u64(val);
        if (lane == 0) { wval[wv] = val; wkey[wv] = key0; }
    } else {
        max_into(dmax + key, val);
        if (lane == 0) { wval[wv] = 0ull; wkey[wv] = key0; }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        // one atomic per distinct member of the workgroup's waves
        for (int w = 0; w < kThreads / 64; w++) {
            if (wval[w] == 0ull) continue;
            unsigned long long m = wval[w];
            for (int q = w + 1; q < kThreads / 64; q++)
                if (wkey[q] == wkey[w] && wval[q] != 0ull) {
                    m = wval[q] > m ? wval[q] : m;
                    wval[q] = 0ull;
                }
            max_into(dmax + wkey[w], m);
        }
    }
}

// the index of each member rounded through a default real (the reference's
// shympi_max of a real), rindex = dt * index / rstol in the model's real,
// istot = 1 + rindex truncated.  w holds nmem maxima, then nmem istot and
// nmem rindex bits (as double); istot also goes to mem_istot (may be null);
// the maxima are cleared for the next call.  One work-item per member.
template <typename S>
__global__ void k_conz_stab_final(unsigned long long *w, int nmem, S dt, float rstol, int *mem_istot) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nmem) return;
    double stabind = __longlong_as_double((long long)w[m]);
    stabind = (double)(float)stabind;
    const S sindex = (S)(stabind / (double)rstol);
    const S ri = dt * sindex;
    const int ist = (int)((S)1 + ri);
    w[nmem + m] = (unsigned long long)(long long)ist;
    w[2 * nmem + m] = (unsigned long long)__double_as_longlong((double)ri);
    if (mem_istot) mem_istot[m] = ist;
    w[m] = 0ull;
}

// per member (nmem = 1: the model): istot[m] and rindex[m] on the host
template <typename S>
static int stab_t(Model &M, float rstol, int *istot, double *rindex) {
    const int nmem = M.nmem_dev > 1 ? M.nmem_dev : 1;
    unsigned long long *w = nmem > 1 ? M.d_mem_stab : (unsigned long long *)(M.d_err + 8);
    const int n = M.nlv * M.nkn;
    hipLaunchKernelGGL(k_conz_stab<S>, dim3((n + kThreads - 1) / kThreads), dim3(kThreads), 0, M.stream, M.nkn,
                       M.nel, M.nlv, M.m.ilhkv, M.m.ilhv, M.m.zbound, M.m.nk_ptr, M.m.nk_el, M.m.nen3v, M.m.geo,
                       (double)M.p.azpar, (double)M.p.adpar, (double)M.p.aapar, (const S *)M.s.utn,
                       (const S *)M.s.uto, (const S *)M.s.vtn, (const S *)M.s.vto, (const S *)M.s.wln,
                       (const S *)M.s.difv, (const S *)M.s.hdeo, (const S *)M.s.hdkn, (const S *)M.s.hdko,
                       M.has_mflux ? (const S *)M.s.mflux : (const S *)nullptr,
                       nmem > 1 ? M.d_node_mem : nullptr, w);
    HIPCHK(M, hipGetLastError());
    hipLaunchKernelGGL(k_conz_stab_final<S>, dim3((nmem + 63) / 64), dim3(64), 0, M.stream, w, nmem, (S)M.p.dt,
                       rstol, nmem > 1 ? M.cur_mem_istot : nullptr);
    HIPCHK(M, hipGetLastError());
    unsigned long long *hbuf = nmem > 1 ? M.h_mem_stab : (unsigned long long *)(M.h_pinned + 8);
    HIPCHK(M, hipMemcpyAsync(hbuf, w + nmem, 2 * nmem * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                             M.stream));
    HIPCHK(M, hipStreamSynchronize(M.stream));
    for (int m = 0; m < nmem; m++) {
        istot[m] = (int)(long long)hbuf[m];
        std::memcpy(&rindex[m], &hbuf[nmem + m], sizeof(double));
    }
    return 0;
}

int launch_scalar_stability(Model &M, float rstol, int *istot, double *rindex) {
    return M.prec == SHYHIP_FP64 ? stab_t<double>(M, rstol, istot, rindex) : stab_t<float>(M, rstol, istot, rindex);
}

}  // namespace shy
