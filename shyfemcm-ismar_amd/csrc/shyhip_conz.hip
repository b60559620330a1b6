// shyhip_conz.hip -- 3D scalar advection-diffusion (conz3d) on gfx950.
//
// Two kernels per substep:
//   k_conz_elem  one work-item per element-layer of a precomputed work list
//                (workgroups of equal work whatever the depth): horizontal
//                advective fluxes (upwind rule, the itvd=2 TVD scheme with
//                precomputed upwind interpolation weights, or the itvd=1
//                scheme with node gradients) -> 3*fl per vertex, stored by
//                the vertex's position in its node's colour-order incidence
//                list.
//   k_conz_node  one workgroup per node tile (at most 64 nodes and 256
//                node-layers).  Phase 1, one work-item per node-layer: the
//                vertical diffusion/advection terms of the node column (they
//                depend on an element only through its bottom layer, so
//                three variants cover every element), then the incident
//                elements gathered in colour order -- the order in which
//                conz3d_omp accumulates them (concentration_omp.f90:243-290)
//                -- with the element terms of conz3d_element (:355-753), and
//                the boundary mass flux of conz3d_nodes; the column
//                coefficients go to LDS.  Phase 2, one work-item per node:
//                the implicit vertical solve (:757-886) and the write-back.
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
    // element-layer work list and node tiles (see Mesh)
    const int *eitem, *ntile;
    int n_eitem;
    // the step's element-layer records (k_conz_prep): int and double columns
    int *rec_i;
    double *rec_d;
    // the launch's single tracer's own terms (null: none): nudging factor
    // robs with target cobs and rate rtauv, sinking wsink * wsinkv, loads
    // rload * load (node arrays like the tracer; wsinkv on interfaces)
    float robs, wsink, rload;
    const S *cobs, *rtauv, *wsinkv;
    S *load;
    // horizontal diffusion (rkpar > 0): factor, element diffusivities
    // [l*nel + ie] and the weights [(3*(ii-1) + (iii-1))*nel + ie]
    float rkpar;
    const S *difhv, *wdif;
    // itvd = 1: the node-layer list, the element-order incidence lists and
    // the node gradients [((l-1)*NT + t)*2 + xy][nkn] (after hbuf's terms)
    const int *nitem, *inc_cnt, *inc_el, *nk_el;
    int n_nitem;
    S *grad;
};

template <typename S>
__device__ inline double cval(const ConzArgs<S> &a, int t, int l, int k) {
    const S v = a.c[t][(size_t)(l - 1) * a.nkn + k];
    if (a.clip[t] && v < (S)1.e-15) return 0.;
    return (double)v;
}

// ------------------------------------------------------------------ element phase
// One work-item per entry of the element-layer work list (Mesh::conz_eitem):
// chunks of kConzChunk consecutive elements, each chunk layer by layer with
// its elements ascending, so that a layer's loads are contiguous.  A
// workgroup takes kThreads consecutive entries, whatever the depth of the
// elements they come from, so every workgroup does the same work.
//
// Everything of an element-layer that depends on the flow only -- the flux
// coefficients f, the upwind case, and for TVD the two pairs of the flow
// direction with their Courant numbers, upwind nodes and interpolation
// weights -- holds for all the substeps of a step (the transports,
// velocities and the substep length do not change within it).  k_conz_prep
// computes it once per step into a record per list entry (ConzRec: columns
// indexed by the entry, so every load is contiguous); k_conz_elem then
// reads the record, gathers the tracers at the element's nodes and upwind
// points, and applies the upwind rule or the limiter.  The arithmetic is
// that of the element terms computed in one pass (the record holds the
// values, not approximations of them).

// v[i] for a run-time i in 0..2 without indexing a register array
template <typename T>
__device__ inline T sel3(const T (&v)[3], int i) {
    return i == 0 ? v[0] : (i == 1 ? v[1] : v[2]);
}

// record columns (int: RI_*, double: RD_*), each n_eitem long
enum { RI_IE = 0, RI_L, RI_KN, RI_SO = RI_KN + 3, RI_CODE = RI_SO + 3, RI_TKN, RI_N = RI_TKN + 6 };
enum { RD_F = 0, RD_TW = 3, RD_ALFA = 9, RD_N = 11 };
// the case word: itot (bits 0-1), isum (2-4), TVD pairs present (5), and per
// pair k (bits 8 + 8k ..): ic (2 bits), id (2 bits), upwind point present (1)
__host__ __device__ inline size_t conz_rec_bytes(int n) {
    return ((size_t)RI_N * n * sizeof(int) + 7) / 8 * 8 + (size_t)RD_N * n * sizeof(double);
}

template <typename S>
__global__ __launch_bounds__(kThreads) void k_conz_prep(ConzArgs<S> a) {
    using R = S;
    const int i = xcd_block() * kThreads + threadIdx.x;
    if (i >= a.n_eitem) return;
    const int n = a.n_eitem;
    const unsigned w = (unsigned)a.eitem[i];
    const int ie = (int)(w >> 5), l = (int)(w & 31u) + 1;
    const int nel = a.nel;
    const bool use_tvd = a.itvd > 0;
    int kn[3], so[3];
    double gb[3], gc[3];
#pragma unroll
    for (int ii = 0; ii < 3; ii++) {
        kn[ii] = a.nen3v[3 * ie + ii];
        so[ii] = a.slot_of[(size_t)ii * nel + ie];
        gb[ii] = a.geo[(size_t)ii * nel + ie];
        gc[ii] = a.geo[(size_t)(3 + ii) * nel + ie];
    }
    const size_t o = (size_t)(l - 1) * nel + ie;
    const S un = a.utn[o], uo = a.uto[o], vn = a.vtn[o], vo = a.vto[o];
    int iext = 1;
    S ul = 0, vl = 0;
    double dis3[3] = {0., 0., 0.}, ajr0 = 0.;
    if (use_tvd) {
        iext = a.elem_iext[ie];
        ul = a.uln[o];
        vl = a.vln[o];
#pragma unroll
        for (int ii = 0; ii < 3; ii++) dis3[ii] = a.dis[(size_t)ii * nel + ie];
        ajr0 = a.geo[(size_t)6 * nel + ie];
    }
    const double azt = 1. - a.az;
    const double us = a.az * (double)un + azt * (double)uo;
    const double vs = a.az * (double)vn + azt * (double)vo;
    int itot = 0, isum = 0;
    double f[3];
    for (int ii = 0; ii < 3; ii++) {
        f[ii] = us * gb[ii] + vs * gc[ii];
        if (f[ii] < 0.) { itot++; isum += ii + 1; }
    }
    int code = itot | (isum << 2);
    int *ri = a.rec_i;
    double *rd = a.rec_d;
    if (use_tvd && iext == 0 && itot >= 1 && itot <= 2) {
        // TVD fluxes (tvd_fluxes, tvd_admin.f90:673-823): the two pairs
        // (ic, id) of the flow direction -- the vertices other than ip,
        // ascending -- and their Courant numbers and upwind points
        code |= 1 << 5;
        const int itot2 = itot - 1, itot1 = 2 - itot;
        const int ip = itot2 * (6 - isum) + itot1 * isum;
        const R ajr = (R)(24. * ajr0);
        const double dte = a.mem_istot ? a.dtp / (double)a.mem_istot[a.elem_mem[ie]] : a.dt;
#pragma unroll
        for (int k = 0; k < 2; k++) {
            const int ii = k == 0 ? (ip == 1 ? 2 : 1) : (ip == 3 ? 2 : 3);
            const int ic = itot2 * ii + itot1 * ip;
            const int id = itot2 * ip + itot1 * ii;
            const int pidx = 2 * (ic - 1) + ((id == (ic % 3) + 1) ? 0 : 1);
            const int iop = 6 - (id + ic);
            const int tet1 = 1 + iop % 3;
            R dx = (R)((double)ajr * sel3(gc, iop - 1));
            dx = (R)(-2 * (tet1 == id ? 1 : 0)) * dx + dx;
            R dy = (R)((double)ajr * sel3(gb, iop - 1));
            dy = (R)(-2 * (tet1 == ic ? 1 : 0)) * dy + dy;
            const R dis = (R)sel3(dis3, iop - 1);
            const R vel = (R)fabs((double)(ul * dx + vl * dy)) / dis;
            const R alfa = (R)((dte * (double)vel) / (double)dis);
            if (a.itvd == 2) {
                const size_t po = (size_t)pidx * nel + ie;
                const bool pup = a.tvd_el[po] >= 0 && a.tvd_lev[po] >= l;
                code |= (ic | (id << 2) | ((pup ? 1 : 0) << 4)) << (8 + 8 * k);
#pragma unroll
                for (int q = 0; q < 3; q++) {
                    ri[(size_t)(RI_TKN + 3 * k + q) * n + i] = a.tvd_kn[((size_t)q * 6 + pidx) * nel + ie];
                    rd[(size_t)(RD_TW + 3 * k + q) * n + i] = a.tvd_w[((size_t)q * 6 + pidx) * nel + ie];
                }
            } else {
                // itvd = 1: the edge vector, for the gradient's projection
                code |= (ic | (id << 2)) << (8 + 8 * k);
                rd[(size_t)(RD_TW + 3 * k) * n + i] = (double)dx;
                rd[(size_t)(RD_TW + 3 * k + 1) * n + i] = (double)dy;
            }
            rd[(size_t)(RD_ALFA + k) * n + i] = (double)alfa;
        }
    }
    ri[(size_t)RI_IE * n + i] = ie;
    ri[(size_t)RI_L * n + i] = l;
#pragma unroll
    for (int ii = 0; ii < 3; ii++) {
        ri[(size_t)(RI_KN + ii) * n + i] = kn[ii];
        ri[(size_t)(RI_SO + ii) * n + i] = so[ii];
        rd[(size_t)(RD_F + ii) * n + i] = f[ii];
    }
    ri[(size_t)RI_CODE * n + i] = code;
}

// ------------------------------------------------------------------ gradient TVD
// itvd = 1 (tvd_grad_3d, tvd_admin.f90:478-543, called before each substep's
// conz3d_omp, concentration.f90:744): each tracer's gradient at every
// node-layer, the mean of the gradients of the incident elements that have
// the layer, weighted by the element area ev(10), in the precision of the
// tracers.  k_tvd_grad_elem, one work-item per element-layer entry, writes
// the element's gradient times its area (the reference's ggx*area,
// ggy*area) into the element-term scratch, which k_conz_elem overwrites
// only afterwards; k_tvd_grad_node, one work-item per node-layer entry,
// adds the incident elements' products and areas in element order -- the
// order of the reference's element loop -- and divides.

// element-layer (ie, l) gradient product of tracer t, component xy, in hbuf
__device__ inline size_t egrad_at(int nel, int NT, int t, int xy, int l, int ie) {
    return ((size_t)((l - 1) * NT + t) * 2 + xy) * nel + ie;
}
__device__ inline size_t ngrad_at(int nkn, int NT, int t, int xy, int l, int k) {
    return ((size_t)((l - 1) * NT + t) * 2 + xy) * nkn + k;
}

template <typename S, int NT>
__global__ __launch_bounds__(kThreads) void k_tvd_grad_elem(ConzArgs<S> a) {
    using R = S;
    const int i = xcd_block() * kThreads + threadIdx.x;
    if (i >= a.n_eitem) return;
    const unsigned w = (unsigned)a.eitem[i];
    const int ie = (int)(w >> 5), l = (int)(w & 31u) + 1;
    if (a.mem_istot && a.isact > a.mem_istot[a.elem_mem[ie]]) return;
    const int nel = a.nel, nkn = a.nkn;
    int kn[3];
    R b[3], c[3];
#pragma unroll
    for (int ii = 0; ii < 3; ii++) {
        kn[ii] = a.nen3v[3 * ie + ii];
        b[ii] = (R)a.geo[(size_t)ii * nel + ie];
        c[ii] = (R)a.geo[(size_t)(3 + ii) * nel + ie];
    }
    const R area = (R)a.geo[(size_t)6 * nel + ie];
    S v[NT][3];
#pragma unroll
    for (int t = 0; t < NT; t++)
#pragma unroll
        for (int ii = 0; ii < 3; ii++) {
            const S x = a.c[t][(size_t)(l - 1) * nkn + kn[ii]];
            v[t][ii] = (a.clip[t] && x < (S)1.e-15) ? (S)0 : x;
        }
    S *eg = a.hbuf;
#pragma unroll
    for (int t = 0; t < NT; t++) {
        R ggx = 0, ggy = 0;
#pragma unroll
        for (int ii = 0; ii < 3; ii++) {
            ggx = ggx + (R)v[t][ii] * b[ii];
            ggy = ggy + (R)v[t][ii] * c[ii];
        }
        eg[egrad_at(nel, NT, t, 0, l, ie)] = ggx * area;
        eg[egrad_at(nel, NT, t, 1, l, ie)] = ggy * area;
    }
}

template <typename S, int NT>
__global__ __launch_bounds__(kThreads) void k_tvd_grad_node(ConzArgs<S> a) {
    using R = S;
    const int i = xcd_block() * kThreads + threadIdx.x;
    if (i >= a.n_nitem) return;
    const unsigned w = (unsigned)a.nitem[i];
    const int k = (int)(w >> 5), l = (int)(w & 31u) + 1;
    if (a.mem_istot && a.isact > a.mem_istot[a.node_mem[k]]) return;
    const int nel = a.nel, nkn = a.nkn;
    const S *eg = a.hbuf;
    // the first kIncWidth incident elements' data loaded at once (padding
    // slots hold element 0, a valid index), the rest one by one
    IncSlots inc;
    inc.load(a.inc_cnt, a.inc_el, nkn, k);
    int lv[kIncWidth];
    R ar[kIncWidth], gp[NT][2][kIncWidth];
#pragma unroll
    for (int s = 0; s < kIncWidth; s++) {
        const int ie = inc.pk[s] >> 2;
        lv[s] = a.ilhv[ie];
        ar[s] = (R)a.geo[(size_t)6 * nel + ie];
#pragma unroll
        for (int t = 0; t < NT; t++) {
            const bool has = lv[s] >= l;      // no product below the element's bottom
            gp[t][0][s] = has ? eg[egrad_at(nel, NT, t, 0, l, ie)] : (R)0;
            gp[t][1][s] = has ? eg[egrad_at(nel, NT, t, 1, l, ie)] : (R)0;
        }
    }
    R sx[NT], sy[NT], sa = 0;
#pragma unroll
    for (int t = 0; t < NT; t++) sx[t] = sy[t] = 0;
#pragma unroll
    for (int s = 0; s < kIncWidth; s++)
        if (s < inc.cnt && lv[s] >= l) {
            sa = sa + ar[s];
#pragma unroll
            for (int t = 0; t < NT; t++) {
                sx[t] = sx[t] + gp[t][0][s];
                sy[t] = sy[t] + gp[t][1][s];
            }
        }
    const int q0 = a.nkc_ptr[k], q1 = a.nkc_ptr[k + 1];
    for (int q = q0 + kIncWidth; q < q1; q++) {
        const int ie = a.nk_el[q] >> 2;
        if (a.ilhv[ie] < l) continue;
        sa = sa + (R)a.geo[(size_t)6 * nel + ie];
#pragma unroll
        for (int t = 0; t < NT; t++) {
            sx[t] = sx[t] + eg[egrad_at(nel, NT, t, 0, l, ie)];
            sy[t] = sy[t] + eg[egrad_at(nel, NT, t, 1, l, ie)];
        }
    }
#pragma unroll
    for (int t = 0; t < NT; t++) {
        a.grad[ngrad_at(nkn, NT, t, 0, l, k)] = sa > (R)0 ? sx[t] / sa : sx[t];
        a.grad[ngrad_at(nkn, NT, t, 1, l, k)] = sa > (R)0 ? sy[t] / sa : sy[t];
    }
}

// GRAD: itvd = 1 (the upstream node's gradient in place of the upwind point)
// HDIF: horizontal diffusion (rkpar > 0): the element's term
// rk3 hmed wdiff (concentration_omp.f90:560-570,727-735) is taken off the
// advective term before it is stored
template <typename S, int NT, bool GRAD, bool HDIF>
__global__ __launch_bounds__(kThreads) void k_conz_elem(ConzArgs<S> a) {
    using R = S;
    const int i = xcd_block() * kThreads + threadIdx.x;
    if (i >= a.n_eitem) return;
    const int n = a.n_eitem, nkn = a.nkn;
    const int *ri = a.rec_i;
    const double *rd = a.rec_d;
    if (a.mem_istot && a.isact > a.mem_istot[a.elem_mem[ri[(size_t)RI_IE * n + i]]]) return;
    const int l = ri[(size_t)RI_L * n + i], code = ri[(size_t)RI_CODE * n + i];
    int kn[3], so[3];
    double f[3];
#pragma unroll
    for (int ii = 0; ii < 3; ii++) {
        kn[ii] = ri[(size_t)(RI_KN + ii) * n + i];
        so[ii] = ri[(size_t)(RI_SO + ii) * n + i];
        f[ii] = rd[(size_t)(RD_F + ii) * n + i];
    }
    const int itot = code & 3, isum = (code >> 2) & 7;
    const bool tvd = (code >> 5) & 1;
    S cr[NT][3];
#pragma unroll
    for (int t = 0; t < NT; t++)
#pragma unroll
        for (int ii = 0; ii < 3; ii++) cr[t][ii] = a.c[t][(size_t)(l - 1) * nkn + kn[ii]];
    auto clip = [&](int t, S v) -> S { return (a.clip[t] && v < (S)1.e-15) ? (S)0 : v; };
    double fl[NT][3];
#pragma unroll
    for (int t = 0; t < NT; t++) fl[t][0] = fl[t][1] = fl[t][2] = 0.;
    if (!tvd) {
        // upwind rule (also the elements of a TVD run at an external
        // boundary); no flux when no vertex or all take inflow
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
            }
        }
    } else {
        // the two pairs' upwind values, then each tracer's limiter
        const int itot2 = itot - 1, itot1 = 2 - itot;
        const int ip = itot2 * (6 - isum) + itot1 * isum;
        const R fact = (R)(-itot2 + itot1);
        // itvd = 2: upwind nodes and weights; itvd = 1: the edge vector
        // (tw[k][0..1])
        int tkn[2][3] = {};
        double tw[2][3] = {};
        R alfa[2];
#pragma unroll
        for (int k = 0; k < 2; k++) {
#pragma unroll
            for (int q = 0; q < (GRAD ? 2 : 3); q++) {
                if (!GRAD) tkn[k][q] = ri[(size_t)(RI_TKN + 3 * k + q) * n + i];
                tw[k][q] = rd[(size_t)(RD_TW + 3 * k + q) * n + i];
            }
            alfa[k] = (R)rd[(size_t)(RD_ALFA + k) * n + i];
        }
        S cu[2][NT][3];
        if (!GRAD) {
#pragma unroll
            for (int k = 0; k < 2; k++)
                if ((code >> (12 + 8 * k)) & 1)
#pragma unroll
                    for (int t = 0; t < NT; t++)
#pragma unroll
                        for (int q = 0; q < 3; q++) cu[k][t][q] = a.c[t][(size_t)(l - 1) * nkn + tkn[k][q]];
        } else {
            // the gradient (x, y) of each tracer at each pair's upstream node
#pragma unroll
            for (int k = 0; k < 2; k++) {
                const int kc = sel3(kn, ((code >> (8 + 8 * k)) & 3) - 1);
#pragma unroll
                for (int t = 0; t < NT; t++) {
                    cu[k][t][0] = a.grad[ngrad_at(nkn, NT, t, 0, l, kc)];
                    cu[k][t][1] = a.grad[ngrad_at(nkn, NT, t, 1, l, kc)];
                }
            }
        }
#pragma unroll
        for (int k = 0; k < 2; k++) {
            const int ck = code >> (8 + 8 * k);
            const int ic = ck & 3, id = (ck >> 2) & 3;
            const bool pup = (ck >> 4) & 1;
            const int ii = k == 0 ? (ip == 1 ? 2 : 1) : (ip == 3 ? 2 : 3);
            const double fii = sel3(f, ii - 1);
#pragma unroll
            for (int t = 0; t < NT; t++) {
                const R conc = clip(t, sel3(cr[t], ic - 1)), cond = clip(t, sel3(cr[t], id - 1));
                R grad;
                if (GRAD) {
                    // itvd = 1: grad = 2 (g(kc) . edge), the edge vector in tw[k][0..1]
                    grad = (R)2. * ((R)cu[k][t][0] * (R)tw[k][0] + (R)cu[k][t][1] * (R)tw[k][1]);
                } else {
                    R conu = cond;
                    if (pup) {
                        double zh = 0.;
                        for (int q = 0; q < 3; q++) zh = zh + (double)clip(t, cu[k][t][q]) * tw[k][q];
                        conu = (R)zh;
                    }
                    grad = cond - conu;
                }
                R rf;
                if ((R)fabs((double)(conc - cond)) < (R)1.e-8) rf = -1;
                else rf = grad / (cond - conc) - (R)1;
                const double rfd = (double)rf;
                const R psi = (R)fmax(0., fmax(fmin(1., 2. * rfd), fmin(2., rfd)));
                const R conf = conc + (R)0.5 * psi * (cond - conc) * ((R)1 - alfa[k]);
                const R term = (R)((double)(fact * conf) * fii);
                for (int q = 0; q < 3; q++) {
                    if (q == ic - 1) fl[t][q] = fl[t][q] - (double)term;
                    if (q == id - 1) fl[t][q] = fl[t][q] + (double)term;
                }
            }
        }
    }
    double hd[NT][3];
#pragma unroll
    for (int t = 0; t < NT; t++) hd[t][0] = hd[t][1] = hd[t][2] = 0.;
    if (HDIF) {
        // rk3 = 3 rkpar difhv in the tracers' precision, hmed the element
        // layer's thickness at this substep, wdiff(ii) = sum over iii of
        // wdifhv(iii, ii) c(iii)
        const int ie = ri[(size_t)RI_IE * n + i], nel = a.nel;
        const size_t o = (size_t)(l - 1) * nel + ie;
        double rso = a.rso;
        if (a.mem_istot) rso = (a.isact - 1) / (double)a.mem_istot[a.elem_mem[ie]];
        const double rk3 = (double)((R)3 * (R)a.rkpar * a.difhv[o]);
        const double hmed = rso * (double)a.hden[o] + (1. - rso) * (double)a.hdeo[o];
        double wv[3][3];
#pragma unroll
        for (int ii = 0; ii < 3; ii++)
#pragma unroll
            for (int q = 0; q < 3; q++) wv[ii][q] = (double)a.wdif[(size_t)(3 * ii + q) * nel + ie];
#pragma unroll
        for (int t = 0; t < NT; t++) {
            double cl[3];
#pragma unroll
            for (int q = 0; q < 3; q++) cl[q] = (double)clip(t, cr[t][q]);
#pragma unroll
            for (int ii = 0; ii < 3; ii++) {
                double waux = 0.;
#pragma unroll
                for (int q = 0; q < 3; q++) waux = waux + wv[ii][q] * cl[q];
                hd[t][ii] = rk3 * hmed * waux;
            }
        }
    }
#pragma unroll
    for (int t = 0; t < NT; t++)
#pragma unroll
        for (int ii = 0; ii < 3; ii++)
            a.hbuf[((size_t)(l - 1) * NT + t) * a.nslot + so[ii]] = (S)(HDIF ? 3. * fl[t][ii] - hd[t][ii]
                                                                               : 3. * fl[t][ii]);
}

// ------------------------------------------------------------------ node phase
// The limited face value of the vertical TVD flux (vertical_flux_ie,
// tvd_admin.f90:930-1016): centre conc, downstream cond, upstream conu,
// Courant number alfa.
__device__ inline double tvd_face(double conc, double cond, double conu, double alfa) {
    double rf;
    if (fabs(conc - cond) < (double)1.e-8f) rf = -1.;
    else rf = (cond - conu) / (cond - conc) - 1.;
    const double psi = fmax(0., fmax(fmin(1., 2. * rf), fmin(2., rf)));
    return conc + 0.5 * psi * (cond - conc) * (1. - alfa);
}

// Node tiles (Mesh::conz_ntile, built at set-up): consecutive nodes, at most
// the node_tile setting (<= kNodeTileMax = the workgroup size) and at most
// kNodeItems node-layers, so that phase 1 is one pass of the workgroup
// whatever the depth of the tile's columns.
// EXT: the tracer's own terms (one tracer per launch): sinking in the
// vertical velocity, nudging in the explicit term (concentration_omp.f90:
// 479,697-701,727-735) and loads after the boundary flux (:816-834)
template <typename S, int NL, int NT, bool EXT>
__global__ __launch_bounds__(kThreads) void k_conz_node(ConzArgs<S> a) {
    using R = S;
    // phase-1 results by list position: right-hand sides of NT tracers and
    // the three diagonals of the column matrix
    __shared__ double cnP[NT][kNodeItems], cdP[kNodeItems], cloP[kNodeItems], chiP[kNodeItems];
    __shared__ int items[kNodeItems];                 // kk | l << 16
    constexpr int kWaves = kThreads / 64;
    // per wavefront and layer: its nodes that have the layer (bit per lane)
    // and the list position of the first of them
    __shared__ unsigned long long lmask[kWaves][NL + 1];
    __shared__ int lbase[kWaves][NL + 1];
    __shared__ int nitems_s;
    const int nkn = a.nkn;
    const int tile = xcd_block();
    const int k0 = a.ntile[tile], tn = a.ntile[tile + 1] - k0;
    const double ad = a.ad, aa = a.aa, adt = 1. - a.ad, aat = 1. - a.aa;
    const bool btvdv = a.itvdv > 0;
    const int nlv = min(a.nlv, NL);

    // The tile's node-layers that exist (l <= ilhkv of the node, members
    // still substepping), layer by layer with the nodes ascending in each:
    // phase 1 takes exactly these, and a layer's loads stay in node order.
    // Work-item kk owns node kk; each wavefront counts its nodes per layer
    // with a ballot, one work-item turns the counts into list positions.
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    int lmk = 0;
    if ((int)threadIdx.x < tn) {
        const int k = k0 + threadIdx.x;
        lmk = min(a.ilhkv[k], nlv);
        if (a.mem_istot && a.isact > a.mem_istot[a.node_mem[k]]) lmk = 0;
    }
    for (int l = 1; l <= nlv; l++) {
        const unsigned long long m = __ballot(lmk >= l);
        if (lane == 0) lmask[wv][l] = m;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int off = 0;
        for (int l = 1; l <= nlv; l++)
            for (int w = 0; w < kWaves; w++) {
                lbase[w][l] = off;
                off += __popcll(lmask[w][l]);
            }
        nitems_s = off;
    }
    __syncthreads();
    const int nitems = nitems_s;
    if (nitems > kNodeItems) {      // the tile table guarantees otherwise
        if (threadIdx.x == 0) atomicCAS(a.err, 0, ERR_CONZ);
        return;
    }
    for (int l = 1; l <= lmk; l++) items[lbase[wv][l] + __popcll(lmask[wv][l] & below)] = (int)threadIdx.x | (l << 16);
    __syncthreads();

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
#pragma unroll
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
        double wtop = (l == 1) ? 0. : (double)a.wln[(size_t)(l - 1) * nkn + k];
        double wbot = (double)a.wln[(size_t)l * nkn + k];
        double hfn = 0.;            // hold * fnudge of the tracer (EXT, NT = 1)
        if (EXT) {
            // the vertical velocity less the sinking, in the tracers' precision
            if (a.wsinkv) {
                if (l > 1)
                    wtop = (double)(R)(a.wln[(size_t)(l - 1) * nkn + k] -
                                       (R)a.wsink * a.wsinkv[(size_t)(l - 1) * nkn + k]);
                wbot = (double)(R)(a.wln[(size_t)l * nkn + k] - (R)a.wsink * a.wsinkv[(size_t)l * nkn + k]);
            }
            if (a.cobs) {
                const size_t o = (size_t)(l - 1) * nkn + k;
                const double fn = (double)a.robs * (double)a.rtauv[o] * ((double)a.cobs[o] - 1. * ckw[0][2]);
                hfn = hokw[2] * fn;
            }
        }
        // element terms of the three column truncations: bottom layer l
        // (v = 0), l+1 (v = 1) or deeper (v = 2; nothing below l+2 is read).
        // The truncation enters only through the layers below l: the
        // diffusion and vertical flux across the bottom interface exist
        // for v >= 1, and the upstream value beyond the next layer for v = 2.
        // Each distinct term is computed once, with the arithmetic of the
        // truncated column (vertical_flux_ie: zero at the column's ends).
        const double hom = hokw[1], hoc = hokw[2], hop = hokw[3];
        const double hnm = hnkw[1], hnc = hnkw[2], hnp = hnkw[3];
        const double pm = (l - 1 >= 1) ? 1. : 0.;
        const double hmotop = 2. * rvptop * pm / (hom + hoc);
        const double hmntop = 2. * rvptop * pm / (hnm + hnc);
        // [0]: no layer below l (pp = 0, the lower neighbour zero), [1]: one
        const double hmobot[2] = {2. * rvpbot * 0. / (hoc + 0.), 2. * rvpbot * 1. / (hoc + hop)};
        const double hmnbot[2] = {2. * rvpbot * 0. / (hnc + 0.), 2. * rvpbot * 1. / (hnc + hnp)};
        // matrix coefficients: [0] for v = 0, [1] for v >= 1 (the lower
        // diagonal does not depend on the truncation)
        double clcv[2], clpv[2], clm = 0.;
#pragma unroll
        for (int v = 0; v < 2; v++) {
            const double hmb = hmnbot[v];
            double clc = 0. + ad * (hmntop + hmb);
            clm = 0. - ad * hmntop;
            double clp = 0. - ad * hmb;
            double w = wtop;
            if (w >= 0.) clc += aa * w;
            else clm += aa * w;
            w = (v == 0) ? 0. : wbot;
            if (w > 0.) clp -= aa * w;
            else clc -= aa * w;
            clcv[v] = clc; clpv[v] = clp;
        }
        // Courant numbers of the two interfaces (tracer-independent)
        double alfa_t = 0., alfa_b = 0.;
        if (btvdv) {
            alfa_t = dt * fabs(wtop) / (0.5 * (hom + hoc));
            alfa_b = dt * fabs(wbot) / (0.5 * (hoc + hop));
        }
        // (one array per variant, indexed by the tracer only: a run-time
        // variant index would put the terms in scratch memory)
        double fd0[NT], fd1[NT], fw0[NT], fw1[NT], fw2[NT];
#pragma unroll
        for (int t = 0; t < NT; t++) {
            const double c0 = ckw[t][0], c1 = ckw[t][1], c2 = ckw[t][2], c3 = ckw[t][3], c4 = ckw[t][4];
            fd0[t] = adt * ((c2 - 0.) * hmobot[0] - (c1 - c2) * hmotop);
            fd1[t] = adt * ((c2 - c3) * hmobot[1] - (c1 - c2) * hmotop);
            // flux across the top interface (none at the surface): [0]
            // without a layer below l, [1] with one (they differ only in
            // the upstream value of a rising flow)
            double vt[2] = {0., 0.};
            if (l >= 2) {
                const double w = wtop;
                if (!btvdv) {
                    vt[0] = vt[1] = w * ((w > 0.) ? c2 : c1);
                } else if (w > 0.) {
                    vt[0] = w * tvd_face(c2, c1, c1, alfa_t);
                    vt[1] = w * tvd_face(c2, c1, c3, alfa_t);
                } else {
                    vt[0] = vt[1] = w * tvd_face(c1, c2, (l - 2 >= 1) ? c0 : c2, alfa_t);
                }
            }
            // flux across the bottom interface (none for v = 0): [0] for
            // v = 1, [1] for v = 2 (they differ only in the upstream value
            // of a rising flow)
            double vb[2];
            {
                const double w = wbot;
                if (!btvdv) {
                    vb[0] = vb[1] = w * ((w > 0.) ? c3 : c2);
                } else if (w > 0.) {
                    vb[0] = w * tvd_face(c3, c2, c2, alfa_b);
                    vb[1] = w * tvd_face(c3, c2, c4, alfa_b);
                } else {
                    vb[0] = vb[1] = w * tvd_face(c2, c3, (l - 1 >= 1) ? c1 : c3, alfa_b);
                }
            }
            fw0[t] = aat * (vt[0] - 0.);
            fw1[t] = aat * (vt[1] - vb[0]);
            fw2[t] = aat * (vt[1] - vb[1]);
        }
        // incident elements in colour order (conz3d_element's scatter)
        double cn[NT], cd = 0., clo = 0., chi = 0.;
#pragma unroll
        for (int t = 0; t < NT; t++) cn[t] = 0.;
        // One element's terms (bottom layer ilev, 4 * area aj4, horizontal
        // terms hf0 and hf1 of tracers 0 and 1).  A macro rather than a
        // lambda: the lambda's captures stayed in scratch memory.
        static_assert(NT <= 2, "two tracers per launch");
#define CONZ_ADD(ilev_, aj4_, hf0_, hf1_)                                                            \
    do {                                                                                           \
        const int ilev = (ilev_);                                                                  \
        if (ilev >= l) {                                                                           \
            const double aj4 = (aj4_), hfv[2] = {(hf0_), (hf1_)};                                  \
            const bool deep = ilev > l, deeper = ilev > l + 1; /* v >= 1, v == 2 */                \
            _Pragma("unroll") for (int t = 0; t < NT; t++) {                                       \
                const double fw = deeper ? fw2[t] : (deep ? fw1[t] : fw0[t]);                      \
                const double fd = deep ? fd1[t] : fd0[t];                                          \
                const double cexpl = aj4 * (hoc * ckw[t][2] + dt * (hfn + hfv[t] - fw - fd));      \
                cn[t] += cexpl;                                                                    \
            }                                                                                      \
            clo += aj4 * dt * clm;                                                                 \
            chi += aj4 * dt * (deep ? clpv[1] : clpv[0]);                                          \
            cd += aj4 * dt * (deep ? clcv[1] : clcv[0]) + aj4 * (1. + dt * 0.) * hnc;              \
        }                                                                                          \
    } while (0)
        // the first kIncWidth positions are read unconditionally (clamped
        // to the list), contiguous for the node; the rest one by one
        int slv[kIncWidth];
        double saj[kIncWidth], shf0[kIncWidth], shf1[kIncWidth];
        const size_t hl = (size_t)(l - 1) * NT * a.nslot;
#pragma unroll
        for (int s = 0; s < kIncWidth; s++) {
            const int q = min(q0 + s, a.nslot - 1);
            slv[s] = a.slot_lev[q];
            saj[s] = a.slot_aj4[q];
            shf0[s] = (double)a.hbuf[hl + q];
            shf1[s] = NT > 1 ? (double)a.hbuf[hl + (size_t)a.nslot + q] : 0.;
        }
#pragma unroll
        for (int s = 0; s < kIncWidth; s++)
            if (s < cnt) CONZ_ADD(slv[s], saj[s], shf0[s], shf1[s]);
        for (int q = q0 + kIncWidth; q < q0 + cnt; q++)
            CONZ_ADD(a.slot_lev[q], a.slot_aj4[q], (double)a.hbuf[hl + q],
                     NT > 1 ? (double)a.hbuf[hl + (size_t)a.nslot + q] : 0.);
#undef CONZ_ADD
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
        if (EXT && a.load) {
            // erosion, or deposition of at most what the node-layer holds
            const size_t o = (size_t)(l - 1) * nkn + k;
            double loading = (double)((R)a.rload * a.load[o]);
            if (loading < 0.) {
                double cload = 0.;
                if (cn[0] > 0.) {
                    cload = -dt * loading;
                    cload = cn[0] * (1. - exp(-cload / cn[0]));
                } else if (cn[0] < 0.) {
                    cn[0] = 0.;
                }
                if (cload > cn[0]) atomicCAS(a.err, 0, ERR_CONZ);
                loading = -cload / dt;
                if (a.rload > 0.f) a.load[o] = (S)(loading / (double)a.rload);
                cn[0] += dt * loading;
            } else if (loading > 0.) {
                cn[0] += dt * loading;
            }
        }
#pragma unroll
        for (int t = 0; t < NT; t++) cnP[t][it] = cn[t];
        cdP[it] = cd; cloP[it] = clo; chiP[it] = chi;
    }
    __syncthreads();

    // phase 2: the column solve of conz3d_nodes, one work-item per node; the
    // tracers share the column matrix and its elimination.  Layer l of node
    // kk sits at list position lstart[l] + (tile nodes before kk that have
    // a layer l).
    const int kk = threadIdx.x, k = k0 + kk;
    if (kk >= tn) return;
    if (a.mem_istot && a.isact > a.mem_istot[a.node_mem[k]]) return;
    const int lmaxk = a.ilhkv[k];
    auto pos = [&](int l0) -> int { return lbase[wv][l0 + 1] + __popcll(lmask[wv][l0 + 1] & below); };
    auto cn = [&](int t, int l0) -> double & { return cnP[t][pos(l0)]; };
    auto cd = [&](int l0) -> double & { return cdP[pos(l0)]; };
    auto clo = [&](int l0) -> double & { return cloP[pos(l0)]; };
    auto chi = [&](int l0) -> double & { return chiP[pos(l0)]; };
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

template <typename S>
static ConzArgs<S> conz_args(Model &M, const TracerSet &ts, int istot, int isact, void *hbuf, void *rec);

template <typename S, int NL, int NT>
static int conz_nl(Model &M, const TracerSet &ts, int istot, int isact, hipStream_t st, void *hbuf, void *rec,
                   bool prep) {
    if (!M.m.conz_eitem || !M.m.conz_ntile) return hip_check(M, hipErrorNotInitialized, "conz3d: no basin");
    if (!rec) return hip_check(M, hipErrorInvalidValue, "conz3d: no element-layer record buffer");
    const ConzArgs<S> a = conz_args<S>(M, ts, istot, isact, hbuf, rec);
    const int ge = (M.m.n_eitem + kThreads - 1) / kThreads;
    if (prep) {
        launch_timed(M, T_CONZ_PREP, k_conz_prep<S>, dim3(ge), dim3(kThreads), 0, st, a);
        HIPCHK(M, hipGetLastError());
    }
    if (M.p.rkpar > 0.f && !a.wdif) return hip_check(M, hipErrorNotInitialized, "conz3d: no diffusion weights");
    if (a.itvd == 1) {
        if (!M.m.conz_nitem) return hip_check(M, hipErrorNotInitialized, "conz3d: no node-layer list");
        if (!a.grad) return hip_check(M, hipErrorInvalidValue, "conz3d: no gradient buffer");
        const int gn = (M.m.n_nitem + kThreads - 1) / kThreads;
        launch_timed(M, T_TVD_GRAD, k_tvd_grad_elem<S, NT>, dim3(ge), dim3(kThreads), 0, st, a);
        launch_timed(M, T_TVD_GRAD, k_tvd_grad_node<S, NT>, dim3(gn), dim3(kThreads), 0, st, a);
        HIPCHK(M, hipGetLastError());
        if (a.wdif) launch_timed(M, T_CONZ_ELEM, k_conz_elem<S, NT, true, true>, dim3(ge), dim3(kThreads), 0, st, a);
        else launch_timed(M, T_CONZ_ELEM, k_conz_elem<S, NT, true, false>, dim3(ge), dim3(kThreads), 0, st, a);
    } else {
        if (a.wdif) launch_timed(M, T_CONZ_ELEM, k_conz_elem<S, NT, false, true>, dim3(ge), dim3(kThreads), 0, st, a);
        else launch_timed(M, T_CONZ_ELEM, k_conz_elem<S, NT, false, false>, dim3(ge), dim3(kThreads), 0, st, a);
    }
    HIPCHK(M, hipGetLastError());
    if (a.cobs || a.wsinkv || a.load) {
        if (NT != 1) return hip_check(M, hipErrorInvalidValue, "conz3d: tracer terms need one tracer per launch");
        launch_timed(M, T_CONZ_NODE, k_conz_node<S, NL, 1, true>, dim3(M.m.n_ntile), dim3(kThreads), 0, st, a);
    } else {
        launch_timed(M, T_CONZ_NODE, k_conz_node<S, NL, NT, false>, dim3(M.m.n_ntile), dim3(kThreads), 0, st, a);
    }
    HIPCHK(M, hipGetLastError());
    return 0;
}

// The arguments of substep isact of istot (zeroed first, so that their
// bytes, padding included, identify the launch).
template <typename S>
static ConzArgs<S> conz_args(Model &M, const TracerSet &ts, int istot, int isact, void *hbuf, void *rec) {
    ConzArgs<S> a;
    std::memset(&a, 0, sizeof a);
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
    for (int t = 0; t < ts.n && t < kMaxTracers; t++) {
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
    a.eitem = M.m.conz_eitem; a.n_eitem = M.m.n_eitem; a.ntile = M.m.conz_ntile;
    if (ts.n == 1 && tracer_has_terms(M, ts.itr[0])) {
        const int t = ts.itr[0];
        if (M.tr_robs[t] != 0.f && M.s.cobs[t] && M.s.rtauv[t]) {
            a.robs = M.tr_robs[t];
            a.cobs = (const S *)M.s.cobs[t];
            a.rtauv = (const S *)M.s.rtauv[t];
        }
        if (M.tr_wsink[t] != 0.f && M.s.wsinkv[t]) {
            a.wsink = M.tr_wsink[t];
            a.wsinkv = (const S *)M.s.wsinkv[t];
        }
        if (M.tr_rload[t] != 0.f && M.s.load[t]) {
            a.rload = M.tr_rload[t];
            a.load = (S *)M.s.load[t];
        }
    }
    if (M.p.rkpar > 0.f) {
        a.rkpar = M.p.rkpar;
        a.difhv = (const S *)M.s.difhv;
        a.wdif = (const S *)M.m.wdif;
    }
    a.nitem = M.m.conz_nitem; a.n_nitem = M.m.n_nitem;
    a.inc_cnt = M.m.inc_cnt; a.inc_el = M.m.inc_el; a.nk_el = M.m.nk_el;
    // the node gradients follow the element terms in the same buffer
    a.grad = (M.p.itvd == 1 && hbuf) ? (S *)hbuf + conz_hbuf_terms(M, ts.n) : nullptr;
    a.rec_i = (int *)rec;
    a.rec_d = rec ? (double *)((char *)rec + conz_rec_bytes(M.m.n_eitem) -
                               (size_t)RD_N * M.m.n_eitem * sizeof(double))
                  : nullptr;
    return a;
}

size_t conz_record_bytes(const Model &M) { return conz_rec_bytes(M.m.n_eitem); }

size_t conz_hbuf_terms(const Model &M, int ntr) { return (size_t)ntr * 3 * M.nlvdi * M.nel; }
size_t conz_hbuf_elems(const Model &M, int ntr) {
    return conz_hbuf_terms(M, ntr) + (size_t)ntr * 2 * M.nlvdi * M.nkn;
}

template <typename S, int NT>
static int conz_t(Model &M, const TracerSet &ts, int istot, int isact, hipStream_t st, void *hbuf, void *rec,
                  bool prep) {
    switch (M.nl_tmpl) {
    case 1: return conz_nl<S, 1, NT>(M, ts, istot, isact, st, hbuf, rec, prep);
    case 2: return conz_nl<S, 2, NT>(M, ts, istot, isact, st, hbuf, rec, prep);
    case 4: return conz_nl<S, 4, NT>(M, ts, istot, isact, st, hbuf, rec, prep);
    case 8: return conz_nl<S, 8, NT>(M, ts, istot, isact, st, hbuf, rec, prep);
    case 12: return conz_nl<S, 12, NT>(M, ts, istot, isact, st, hbuf, rec, prep);
    case 16: return conz_nl<S, 16, NT>(M, ts, istot, isact, st, hbuf, rec, prep);
    case 20: return conz_nl<S, 20, NT>(M, ts, istot, isact, st, hbuf, rec, prep);
    case 24: return conz_nl<S, 24, NT>(M, ts, istot, isact, st, hbuf, rec, prep);
    default: return conz_nl<S, 32, NT>(M, ts, istot, isact, st, hbuf, rec, prep);
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
constexpr int kBndoNb = 6;          // inner neighbours of a node loaded at once

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
    // the outflow test's normal velocity scal per node and layer
    // ([i * nlvdi + l - 1]): computed and kept by the first substep of a
    // step (the velocities and the wet/dry mask hold for the whole step),
    // read by the others; null: computed by every call
    S *scal;
    int nlvdi;
};

template <typename S>
__global__ __launch_bounds__(kBndoLanes) void k_bndo(BndoArgs<S> a) {
    using R = S;
    __shared__ R val[kBndoLanes];
    __shared__ int status[kBndoLanes];
    const int i = blockIdx.x;
    if (i >= a.nbndo) return;
    // the node's static data in one round of loads
    const int open = a.open[i], k = a.bk[i], lmax = a.lmax[i];
    if (!open) return;
    if (a.mem_istot && a.isact > a.mem_istot[a.node_mem[k]]) return;
    S *const cv = a.cv[blockIdx.y];
    const S *const rbc = a.rbc[blockIdx.y];
    const int lane = threadIdx.x, l = lane + 1;
    const int nkn = a.nkn, nel = a.nel;
    const R flag = (R)-999.;
    status[lane] = 0;
    if (l <= lmax) {
        const size_t o = (size_t)(l - 1) * nkn + k;
        R cur = cv[o];
        if (rbc) {
            const R rb = rbc[o];
            if (rb != flag) cur = rb;
        }
        R scal;
        if (a.scal && a.isact > 1) {
            scal = a.scal[(size_t)i * a.nlvdi + l - 1];
        } else {
            R vv = 0, up = 0, vp = 0;
            for (int q = a.eptr[i]; q < a.eptr[i + 1]; q++) {
                const int ie = a.el[q];
                if (a.iweg[ie] != 0 || l > a.elev[q]) continue;
                const R aj = (R)a.eaj[q];
                vv = vv + aj;
                up = up + aj * a.uln[(size_t)(l - 1) * nel + ie];
                vp = vp + aj * a.vln[(size_t)(l - 1) * nel + ie];
            }
            R u = up, v = vp;
            if (vv > (R)0) { u = u / vv; v = v / vv; }
            scal = a.bnx[i] * u + a.bny[i] * v;
            if (a.scal && blockIdx.y == 0) a.scal[(size_t)i * a.nlvdi + l - 1] = scal;
        }
        const bool binside = scal <= (R)0 || cur == flag;
        R bc = 0, tw = 0;
        int ntbc = 0;
        if (binside) {
            // the first kBndoNb inner neighbours' entries loaded at once, then
            // their values, then added in list order; more one by one
            const int j0 = a.bptr[i], nj = a.bptr[i + 1] - j0;
            int nlev[kBndoNb], nbk[kBndoNb];
            R w[kBndoNb], cvn[kBndoNb];
#pragma unroll
            for (int s = 0; s < kBndoNb; s++) {
                nlev[s] = 0; nbk[s] = 0; w[s] = 0;
                if (s < nj) { nlev[s] = a.nblev[j0 + s]; nbk[s] = a.bnb[j0 + s]; w[s] = a.bw[j0 + s]; }
            }
#pragma unroll
            for (int s = 0; s < kBndoNb; s++) {
                cvn[s] = 0;
                if (s < nj && l <= nlev[s]) cvn[s] = cv[(size_t)(l - 1) * nkn + nbk[s]];
            }
#pragma unroll
            for (int s = 0; s < kBndoNb; s++)
                if (s < nj && l <= nlev[s]) {
                    ntbc++;
                    tw = tw + w[s];
                    bc = bc + w[s] * cvn[s];
                }
            for (int j = j0 + kBndoNb; j < j0 + nj; j++) {
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
static BndoArgs<S> bndo_args(Model &M, const TracerSet &ts, int isact, void *scal);

template <typename S>
static int bndo_t(Model &M, const TracerSet &ts, hipStream_t st, int isact, void *scal) {
    if (M.m.nbndo == 0) return 0;
    if (M.nlv > kBndoLanes) return hip_check(M, hipErrorInvalidValue, "bndo: more layers than lanes");
    const BndoArgs<S> a = bndo_args<S>(M, ts, isact, scal);
    hipLaunchKernelGGL(k_bndo<S>, dim3(M.m.nbndo, ts.n), dim3(kBndoLanes), 0, st, a);
    HIPCHK(M, hipGetLastError());
    return 0;
}

// (zeroed first, like conz_args)
template <typename S>
static BndoArgs<S> bndo_args(Model &M, const TracerSet &ts, int isact, void *scal) {
    BndoArgs<S> a;
    std::memset(&a, 0, sizeof a);
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
    a.scal = (S *)scal;
    a.nlvdi = M.nlvdi;
    return a;
}

int launch_bndo(Model &M, const TracerSet &ts, hipStream_t st, int isact, void *scal) {
    return M.prec == SHYHIP_FP64 ? bndo_t<double>(M, ts, st, isact, scal) : bndo_t<float>(M, ts, st, isact, scal);
}

int launch_bndo(Model &M, int itr, hipStream_t st, int isact) {
    TracerSet ts;
    ts.n = 1;
    ts.itr[0] = itr;
    return launch_bndo(M, ts, st, isact, nullptr);
}

int launch_conz3d(Model &M, const TracerSet &ts, int istot, int isact, hipStream_t st, void *hbuf, void *rec,
                  bool prep) {
    if (ts.n == 2)
        return M.prec == SHYHIP_FP64 ? conz_t<double, 2>(M, ts, istot, isact, st, hbuf, rec, prep)
                                     : conz_t<float, 2>(M, ts, istot, isact, st, hbuf, rec, prep);
    return M.prec == SHYHIP_FP64 ? conz_t<double, 1>(M, ts, istot, isact, st, hbuf, rec, prep)
                                 : conz_t<float, 1>(M, ts, istot, isact, st, hbuf, rec, prep);
}

void conz3d_key(Model &M, const TracerSet &ts, int istot, void *hbuf, void *rec, void *bscal, std::string &key) {
    key_add(key, istot);
    key_add(key, M.nl_tmpl);
    key_add(key, M.prec);
    key_add(key, ts.n);
    key_add(key, M.m.nbndo);
    key_add(key, M.m.n_ntile);
    if (M.prec == SHYHIP_FP64) {
        key_add(key, conz_args<double>(M, ts, istot, 1, hbuf, rec));
        key_add(key, bndo_args<double>(M, ts, 1, bscal));
    } else {
        key_add(key, conz_args<float>(M, ts, istot, 1, hbuf, rec));
        key_add(key, bndo_args<float>(M, ts, 1, bscal));
    }
}

int launch_conz3d(Model &M, int itr, int istot, int isact, int clip_small, hipStream_t st, void *hbuf, void *rec,
                  const void *cbound) {
    TracerSet ts;
    ts.n = 1;
    ts.itr[0] = itr;
    ts.clip[0] = clip_small;
    ts.cbound[0] = cbound;
    return launch_conz3d(M, ts, istot, isact, st, hbuf, rec, true);
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
// 0 when it adds nothing.  The incident elements are taken in element order
// kStabBatch at a time: each batch's loads are issued together (the list
// entries, then every element's data and its vertices, then the vertices'
// vertical velocities), and the batch is then added in order.
constexpr int kStabBatch = 3;

template <typename S>
__device__ inline unsigned long long stab_ratio(int l, int k, int nkn, int nel, int nlv, const int *ilhv,
                                                const int *nk_ptr, const int *nk_el, const int *nen3v,
                                                const double *geo, double az, double ad, double aa, const S *utn,
                                                const S *uto, const S *vtn, const S *vto, const S *wln,
                                                const S *difv, const S *hdeo, const S *hdkn, const S *hdko,
                                                const S *mflux, float rkpar, const S *difhv, const S *wdif,
                                                float robs, const S *rtauv, float wsink, const S *wsinkv) {
    const double azt = 1. - az, adt = 1. - ad, aat = 1. - aa;
    double chigh = 0., cn = 0., cdiag = 0., clow = 0.;
    double co = 0.;
    if (mflux) {
        const S q = mflux[(size_t)(l - 1) * nkn + k];
        if (q > (S)0) co = (double)q;
    }
    // the node's own data (the layer below exists for elements deeper than l)
    const double dtop = l > 1 ? (double)difv[(size_t)(l - 1) * nkn + k] : 0.;
    const double dbot = l < nlv ? (double)difv[(size_t)l * nkn + k] : 0.;
    const double hn = (double)hdkn[(size_t)(l - 1) * nkn + k], ho = (double)hdko[(size_t)(l - 1) * nkn + k];
    const double rtau = rtauv ? (double)rtauv[(size_t)(l - 1) * nkn + k] : 0.;
    const int p0 = nk_ptr[k], p1 = nk_ptr[k + 1];
    for (int pb = p0; pb < p1; pb += kStabBatch) {
        int ie[kStabBatch], iv[kStabBatch], ilev[kStabBatch], kj[kStabBatch][3];
        double aj4[kStabBatch], us[kStabBatch], vs[kStabBatch], hdv[kStabBatch], hdt[kStabBatch], hdb[kStabBatch];
        double gb[kStabBatch][3], gc[kStabBatch][3], wt[kStabBatch][3], wb[kStabBatch][3];
        double rk3[kStabBatch], wd[kStabBatch];
#pragma unroll
        for (int b = 0; b < kStabBatch; b++) {
            const int e = nk_el[min(pb + b, p1 - 1)];
            ie[b] = e >> 2;
            iv[b] = e & 3;
        }
#pragma unroll
        for (int b = 0; b < kStabBatch; b++) {
            const int el = ie[b];
            const size_t le = (size_t)(l - 1) * nel + el;
            ilev[b] = ilhv[el];
            aj4[b] = 4. * geo[(size_t)6 * nel + el];
            us[b] = az * (double)utn[le] + azt * (double)uto[le];
            vs[b] = az * (double)vtn[le] + azt * (double)vto[le];
            hdv[b] = (double)hdeo[le];
            hdt[b] = l > 1 ? (double)hdeo[le - nel] : 0.;
            hdb[b] = l < nlv ? (double)hdeo[le + nel] : 0.;
            rk3[b] = 0.;
            wd[b] = 0.;
            if (wdif) {
                // horizontal diffusion: 3 rkpar difhv in the tracers'
                // precision and the vertex's own weight wdifhv(ii, ii)
                rk3[b] = (double)((S)3 * (S)rkpar * difhv[le]);
                wd[b] = (double)wdif[(size_t)(4 * iv[b]) * nel + el];
            }
#pragma unroll
            for (int jj = 0; jj < 3; jj++) {
                kj[b][jj] = nen3v[3 * el + jj];
                gb[b][jj] = geo[(size_t)jj * nel + el];
                gc[b][jj] = geo[(size_t)(3 + jj) * nel + el];
            }
        }
#pragma unroll
        for (int b = 0; b < kStabBatch; b++)
#pragma unroll
            for (int jj = 0; jj < 3; jj++) {
                wt[b][jj] = l > 1 ? (double)wln[(size_t)(l - 1) * nkn + kj[b][jj]] : 0.;
                wb[b][jj] = l < nlv ? (double)wln[(size_t)l * nkn + kj[b][jj]] : 0.;
                if (wsinkv) {
                    // less the tracer's sinking, in the tracers' precision
                    if (l > 1)
                        wt[b][jj] = (double)(S)(wln[(size_t)(l - 1) * nkn + kj[b][jj]] -
                                                (S)wsink * wsinkv[(size_t)(l - 1) * nkn + kj[b][jj]]);
                    if (l < nlv)
                        wb[b][jj] = (double)(S)(wln[(size_t)l * nkn + kj[b][jj]] -
                                                (S)wsink * wsinkv[(size_t)l * nkn + kj[b][jj]]);
                }
            }
#pragma unroll
        for (int b = 0; b < kStabBatch; b++) {
            if (pb + b >= p1 || l > ilev[b]) continue;
            const bool below = l < ilev[b];
            double coadv = 0., ciadv = 0., cadv = 0.;
#pragma unroll
            for (int jj = 0; jj < 3; jj++) {
                if (jj > iv[b]) break;
                const double f = us[b] * gb[b][jj] + vs[b] * gc[b][jj];
                double fwin = 0., fwout = 0.;
                const double w_t = wt[b][jj];
                const double w_b = below ? wb[b][jj] : 0.;
                if (w_t > 0.) fwout = fwout + aat * w_t;
                else fwin = fwin - aat * w_t;
                if (w_b > 0.) fwin = fwin + aat * w_b;
                else fwout = fwout - aat * w_b;
                const double chadv = aj4[b] * 3. * f;
                const double cviadv = aj4[b] * fwin, cvoadv = aj4[b] * fwout;
                if (chadv < 0.) coadv = cvoadv - chadv;
                else ciadv = cviadv + chadv;
                cadv = coadv > ciadv ? coadv : ciadv;
            }
            chigh = chigh + cadv;
            double hmtop = 0., hmbot = 0.;
            if (l > 1) hmtop = 2. * dtop / (hdt[b] + hdv[b]);
            if (below) hmbot = 2. * dbot / (hdv[b] + hdb[b]);
            cn = cn + aj4[b] * (adt * (hmtop + hmbot));
            clow = clow + aj4[b] * rk3[b] * ho * wd[b];
            if (rtauv) co = co + aj4[b] * ho * (double)robs * rtau;      // nudging
            cdiag = cdiag + aj4[b] * (ho < hn ? ho : hn);
        }
    }
    if (cdiag > 0.) {
        const double r = ((chigh + clow) + cn + co) / cdiag;
        if (r > 0.) return (unsigned long long)__double_as_longlong(r);
    }
    return 0ull;
}

// One work-item per entry of the node-layer list.
template <typename S>
__global__ void __launch_bounds__(kThreads) k_conz_stab(int nkn, int nel, int nlv, int n_nitem, const int *nitem,
                                                        const int *ilhv, const int *zbound, const int *nk_ptr,
                                                        const int *nk_el, const int *nen3v, const double *geo,
                                                        double az, double ad, double aa, const S *utn, const S *uto,
                                                        const S *vtn, const S *vto, const S *wln, const S *difv,
                                                        const S *hdeo, const S *hdkn, const S *hdko, const S *mflux,
                                                        float rkpar, const S *difhv, const S *wdif, float robs,
                                                        const S *rtauv, float wsink, const S *wsinkv,
                                                        const int *node_mem, unsigned long long *dmax) {
    const int i = xcd_block() * blockDim.x + threadIdx.x;
    unsigned long long val = 0ull;    // this node-layer's ratio as bits (0: none)
    int key = 0;                      // its member
    if (i < n_nitem) {
        const unsigned w = (unsigned)nitem[i];
        const int k = (int)(w >> 5), l = (int)(w & 31u) + 1;
        if (node_mem) key = node_mem[k];
        if (!zbound[k]) val = stab_ratio(l, k, nkn, nel, nlv, ilhv, nk_ptr, nk_el, nen3v, geo, az, ad, aa, utn, uto,
                                         vtn, vto, wln, difv, hdeo, hdkn, hdko, mflux, rkpar, difhv, wdif, robs,
                                         rtauv, wsink, wsinkv);
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
    if (M.p.rkpar > 0.f && !M.m.wdif) return hip_check(M, hipErrorNotInitialized, "scalar_stability: no diffusion weights");
    // the nudging and sinking of the tracer being transported, if any
    const int t = M.terms_tracer;
    float robs = 0.f, wsink = 0.f;
    const void *rtauv = nullptr, *wsinkv = nullptr;
    if (t >= 0 && t < M.ntr) {
        if (M.tr_robs[t] != 0.f && M.s.cobs[t] && M.s.rtauv[t]) { robs = M.tr_robs[t]; rtauv = M.s.rtauv[t]; }
        if (M.tr_wsink[t] != 0.f && M.s.wsinkv[t]) { wsink = M.tr_wsink[t]; wsinkv = M.s.wsinkv[t]; }
    }
    const int nmem = M.nmem_dev > 1 ? M.nmem_dev : 1;
    unsigned long long *w = nmem > 1 ? M.d_mem_stab : (unsigned long long *)(M.d_err + 8);
    const int n = M.m.n_nitem;
    hipLaunchKernelGGL(k_conz_stab<S>, dim3((n + kThreads - 1) / kThreads), dim3(kThreads), 0, M.stream, M.nkn,
                       M.nel, M.nlv, n, M.m.conz_nitem, M.m.ilhv, M.m.zbound, M.m.nk_ptr, M.m.nk_el, M.m.nen3v, M.m.geo,
                       (double)M.p.azpar, (double)M.p.adpar, (double)M.p.aapar, (const S *)M.s.utn,
                       (const S *)M.s.uto, (const S *)M.s.vtn, (const S *)M.s.vto, (const S *)M.s.wln,
                       (const S *)M.s.difv, (const S *)M.s.hdeo, (const S *)M.s.hdkn, (const S *)M.s.hdko,
                       M.has_mflux ? (const S *)M.s.mflux : (const S *)nullptr, M.p.rkpar,
                       (const S *)M.s.difhv, M.p.rkpar > 0.f ? (const S *)M.m.wdif : (const S *)nullptr,
                       robs, (const S *)rtauv, wsink, (const S *)wsinkv, nmem > 1 ? M.d_node_mem : nullptr, w);
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
