/*
 * shyoracle_conz.c -- 3D scalar advection-diffusion of the CPU oracle
 * (TEST INFRASTRUCTURE; see shyoracle.h).
 *
 * Restates src/shyfem/concentration_omp.f90 (conz3d_omp, conz3d_element,
 * conz3d_nodes), tvd_admin.f90 (tvd_grad_3d, tvd_fluxes, tvd_get_upwind_c,
 * vertical_flux_ie), evgeom.f90 (femintp_xi, xy2xi, adjust_xi),
 * bnd_routines.f90 make_scal_flux and the substep loop of
 * concentration.f90 scal3sh.  Element contributions are accumulated colour
 * by colour, in the colouring's element order, exactly as conz3d_omp does.
 * The scalar's own terms -- nudging (robs, rtauv, cobs), sinking (wsink,
 * wsinkv) and loads (rload, load) -- come with the state (oc_state).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "shyoracle.h"

#define EV(i, ie) m->ev[((i) - 1) + 19 * (ie)]
#define A3E(a, l, ie) (a)[((l) - 1) + nlvdi * (ie)]
#define A3K(a, l, k) (a)[((l) - 1) + nlvdi * (k)]
#define AW(a, l, k) (a)[(l) + (nlvdi + 1) * (k)]

static double dmax(double a, double b) { return a > b ? a : b; }
static double dmin(double a, double b) { return a < b ? a : b; }

/* adjust_xi: evgeom.f90:577-660 */
static int adjust_xi(double *xi)
{
    double xisum = 0., xiso = 0.;
    int is = 0, it = 0;
    for (int ii = 1; ii <= 3; ii++) {
        xiso += xi[ii - 1];
        if (xi[ii - 1] < 0.) xi[ii - 1] = 0.;
        if (xi[ii - 1] > 1.) xi[ii - 1] = 1.;
        if (xi[ii - 1] == 0.) { is += ii; it++; }
        xisum += xi[ii - 1];
    }
    if (it == 3 || fabs(xiso - 1.) > (double)1.e-8f) return 1;
    double xiadj = (xisum - 1.) / (3 - it);
    if (it == 2) {
        is = 6 - is;
        xi[is - 1] = 1.;
    } else {
        for (int ii = 1; ii <= 3; ii++) {
            if (ii != is) {
                xi[ii - 1] -= xiadj;
                if (xi[ii - 1] < 0.) xi[ii - 1] = 0.;
                if (xi[ii - 1] > 1.) xi[ii - 1] = 1.;
            }
        }
    }
    xiso = xi[0] + xi[1] + xi[2];
    if (fabs(xiso - 1.) > (double)1.e-8f) return 3;
    return 0;
}

/* tvd_get_upwind_c (tvd_admin.f90:608-669) with femintp_xi
 * (evgeom.f90:1413-1445): value of cv interpolated at the upwind point */
static real tvd_upwind(const oc_mesh *m, int ie, int l, int ic, int id, real cu, const real *cv)
{
    const int nlvdi = m->nlvdi;
    int idx = (id - 1) + 3 * (ic - 1) + 9 * ie;
    int ienew = m->tvd_ie[idx];
    if (ienew < 0) return cu;
    if (m->ilhv[ienew] < l) return cu;
    double x = m->tvd_x[idx], y = m->tvd_y[idx];
    double xi[3];
    for (int ii = 1; ii <= 3; ii++)
        xi[ii - 1] = EV(ii, ienew) + EV(3 + ii, ienew) * x + EV(6 + ii, ienew) * y;
    adjust_xi(xi);
    double zh = 0.;
    for (int ii = 1; ii <= 3; ii++) zh = zh + (double)A3K(cv, l, m->nen3v[3 * ienew + ii - 1]) * xi[ii - 1];
    return (real)zh;
}

/* tvd_upwind's value at layer l for every (id, ic, ie), out[(id-1) +
 * 3(ic-1) + 9 ie], 0 where there is no upwind element or id == ic (the
 * tests compare it with the reference's femintp_xi) */
void oc_tvd_upwind_table(const oc_mesh *m, int l, const real *cv, real *out)
{
    for (int ie = 0; ie < m->nel; ie++)
        for (int ic = 1; ic <= 3; ic++)
            for (int id = 1; id <= 3; id++)
                out[(id - 1) + 3 * (ic - 1) + 9 * ie] = id == ic ? (real)0 : tvd_upwind(m, ie, l, ic, id, (real)0, cv);
}

static int smartdelta(int a, int b)
{
    return (int)((float)((a + b) - abs(a - b)) / (float)((a + b) + abs(a - b)));
}

/* tvd_grad_3d: tvd_admin.f90:478-543 (itvd = 1).  The gradient of cc at
 * every node-layer: the mean of the gradients of the incident elements that
 * have the layer, weighted by ev(10), accumulated element by element in
 * ascending element number (the reference's ip_sort_elem order) in the
 * reference's real.  Node-layers below the node's bottom layer are zero. */
void oc_tvd_grad_3d(const oc_mesh *m, const real *cc, real *gx, real *gy)
{
    const int nlvdi = m->nlvdi;
    const size_t nn = (size_t)nlvdi * m->nkn;
    real *aux = (real *)calloc(nn, sizeof(real));
    memset(gx, 0, nn * sizeof(real));
    memset(gy, 0, nn * sizeof(real));
    for (int ie = 0; ie < m->nel; ie++) {
        const real area = (real)EV(10, ie);
        const int lmax = m->ilhv[ie];
        for (int l = 1; l <= lmax; l++) {
            real ggx = 0, ggy = 0;
            for (int ii = 1; ii <= 3; ii++) {
                const int k = m->nen3v[3 * ie + ii - 1];
                const real b = (real)EV(ii + 3, ie), c = (real)EV(ii + 6, ie);
                ggx = ggx + A3K(cc, l, k) * b;
                ggy = ggy + A3K(cc, l, k) * c;
                A3K(aux, l, k) = A3K(aux, l, k) + area;
            }
            for (int ii = 1; ii <= 3; ii++) {
                const int k = m->nen3v[3 * ie + ii - 1];
                A3K(gx, l, k) = A3K(gx, l, k) + ggx * area;
                A3K(gy, l, k) = A3K(gy, l, k) + ggy * area;
            }
        }
    }
    for (int k = 0; k < m->nkn; k++)
        for (int l = 1; l <= m->ilhkv[k]; l++) {
            const real area = A3K(aux, l, k);
            if (area > (real)0) {
                A3K(gx, l, k) = A3K(gx, l, k) / area;
                A3K(gy, l, k) = A3K(gy, l, k) / area;
            }
        }
    free(aux);
}

/* tvd_fluxes: tvd_admin.f90:673-823.  itvd = 2 (gx null): the upstream
 * value interpolated at the upwind point; itvd = 1: the gradient of the
 * upstream node kc projected on the edge, grad = 2 (gx dx + gy dy) */
static void tvd_fluxes(const oc_mesh *m, const oc_state *s, int ie, int l, int itot, int isum,
                       double dt, const double *cl /* cl[ii] at layer l */, const real *cv,
                       const real *gx, const real *gy, const double *f, double *fl)
{
    const int nlvdi = m->nlvdi;
    const real eps = (real)1.e-8;
    fl[0] = fl[1] = fl[2] = 0.;
    if (itot < 1 || itot > 2) return;
    int itot2 = itot - 1, itot1 = 2 - itot;
    real u = A3E(s->ulnv, l, ie), v = A3E(s->vlnv, l, ie);
    real aj = (real)(24 * EV(10, ie));
    int ip = isum;
    ip = itot2 * (6 - ip) + itot1 * ip;
    for (int ii = 1; ii <= 3; ii++) {
        if (ii == ip) continue;
        int ic = itot2 * ii + itot1 * ip;
        int id = itot2 * ip + itot1 * ii;
        real fact = (real)(-itot2 + itot1);
        real conc = (real)cl[ic - 1];
        real cond = (real)cl[id - 1];
        int iop = 6 - (id + ic);
        int tet1 = 1 + iop % 3;
        real dx = (real)(aj * EV(6 + iop, ie));
        dx = (real)(-2 * smartdelta(tet1, id)) * dx + dx;
        real dy = (real)(aj * EV(3 + iop, ie));
        dy = (real)(-2 * smartdelta(tet1, ic)) * dy + dy;
        real dis = (real)EV(16 + iop, ie);
        real vel = (real)fabs((double)(u * dx + v * dy)) / dis;
        real alfa = (real)((dt * vel) / dis);
        real grad;
        if (!gx) {
            real conu = tvd_upwind(m, ie, l, ic, id, cond, cv);
            grad = cond - conu;
        } else {
            const int kc = m->nen3v[3 * ie + ic - 1];
            const real gcx = A3K(gx, l, kc), gcy = A3K(gy, l, kc);
            grad = (real)2. * (gcx * dx + gcy * dy);
        }
        real rf;
        if ((real)fabs((double)(conc - cond)) < eps) rf = -1;
        else rf = grad / (cond - conc) - (real)1;
        real psi = (real)dmax(0., dmax(dmin(1., 2. * rf), dmin(2., rf)));
        real conf = conc + (real)0.5 * psi * (cond - conc) * ((real)1 - alfa);
        real term = (real)(fact * conf * f[ii - 1]);
        fl[ic - 1] = fl[ic - 1] - term;
        fl[id - 1] = fl[id - 1] + term;
    }
}

/* The horizontal advective flux of element ie at layer l
 * (concentration_omp.f90:640-688): the upwind rule for the vertex
 * projections f (itot of them negative, isum the sum of their vertex
 * numbers), replaced by the TVD fluxes when tvd is set (TVD run, no vertex
 * on an external boundary); cl holds the element's three values at l */
static void hflux(const oc_mesh *m, const oc_state *s, int ie, int l, int itot, int isum, double dt,
                  const double *cl, const real *cv, int tvd, const real *gx, const real *gy,
                  const double *f, double *fl)
{
    if (itot == 1) {
        for (int ii = 0; ii < 3; ii++) fl[ii] = f[ii] * cl[isum - 1];
    } else if (itot == 2) {
        int is2 = 6 - isum;
        for (int ii = 0; ii < 3; ii++) fl[ii] = f[ii] * cl[ii];
        fl[is2 - 1] = 0.;
        fl[is2 - 1] = -(fl[0] + fl[1] + fl[2]);
    } else {
        fl[0] = fl[1] = fl[2] = 0.;
    }
    if (tvd) tvd_fluxes(m, s, ie, l, itot, isum, dt, cl, cv, gx, gy, f, fl);
}

/* hflux of element ie (0-based) at layer l on the tracer cv as conz3d
 * computes it for a substep of length dt (the tests compare the TVD
 * variants element by element); gx, gy: the node gradients of cv
 * (oc_tvd_grad_3d) for itvd = 1, else unused */
void oc_conz_hflux(const oc_mesh *m, const oc_state *s, const oc_params *p, const real *cv, const real *gx,
                   const real *gy, int ie, int l, double dt, double *fl)
{
    const int nlvdi = m->nlvdi;
    const double az = p->azpar, azt = 1. - az;
    const double us = az * A3E(s->utlnv, l, ie) + azt * A3E(s->utlov, l, ie);
    const double vs = az * A3E(s->vtlnv, l, ie) + azt * A3E(s->vtlov, l, ie);
    int itot = 0, isum = 0, iext = 0;
    double f[3], cl[3];
    for (int ii = 1; ii <= 3; ii++) {
        const int k = m->nen3v[3 * ie + ii - 1];
        f[ii - 1] = us * EV(ii + 3, ie) + vs * EV(ii + 6, ie);
        if (f[ii - 1] < 0.) { itot++; isum += ii; }
        cl[ii - 1] = A3K(cv, l, k);
        if (m->iextb && m->iextb[k]) iext++;
    }
    const int tvd = p->itvd > 0 && iext == 0;
    hflux(m, s, ie, l, itot, isum, dt, cl, cv, tvd, p->itvd == 1 ? gx : NULL, p->itvd == 1 ? gy : NULL, f, fl);
}

/* conz3d_omp + conz3d_element + conz3d_nodes: concentration_omp.f90:65-886.
 * Returns 0, or 1 when a column has a zero diagonal. */
int oc_conz3d(const oc_mesh *m, const oc_state *s, const oc_params *p,
              real *cn1, const real *cbound, double ddt, int istot, int isact)
{
    const int nlvdi = m->nlvdi, nkn = m->nkn;
    const int itvd = p->itvd, itvdv = p->itvdv;
    double az = p->azpar, ad = p->adpar, aa = p->aapar;
    double azt = 1. - az, adt = 1. - ad, aat = 1. - aa;
    double rstot = istot;
    double rso = (isact - 1) / rstot, rsn = isact / rstot;
    double rsot = 1. - rso, rsnt = 1. - rsn;
    double dt = ddt / rstot;
    int btvd = itvd > 0, btvdv = itvdv > 0;
    size_t nn = (size_t)nlvdi * nkn;
    double *cn = (double *)calloc(nn, sizeof(double));
    double *cdiag = (double *)calloc(nn, sizeof(double));
    double *clow = (double *)calloc(nn, sizeof(double));
    double *chigh = (double *)calloc(nn, sizeof(double));
    /* itvd = 1: the node gradients of the tracer entering the substep
     * (tvd_grad_3d before conz3d_omp, concentration.f90:744) */
    real *gx = NULL, *gy = NULL;
    if (btvd && itvd == 1) {
        gx = (real *)malloc(nn * sizeof(real));
        gy = (real *)malloc(nn * sizeof(real));
        oc_tvd_grad_3d(m, cn1, gx, gy);
    }
    int L2 = nlvdi + 2;
#define H2(a, l, ii) (a)[(l) + L2 * ((ii) - 1)]
    /* colour by colour; inside a colour no two elements share a node, so
     * its elements may run in parallel (as conz3d_omp runs them) and every
     * node still receives its contributions in colour order */
    for (int ccol = 0; ccol < m->ncolor; ccol++) {
#pragma omp parallel
    {
    double *hold = (double *)malloc(sizeof(double) * 3 * L2);
    double *hnew = (double *)malloc(sizeof(double) * 3 * L2);
    double *cl = (double *)malloc(sizeof(double) * 3 * L2);
    double *wl = (double *)malloc(sizeof(double) * 3 * L2);
    double *vflux = (double *)malloc(sizeof(double) * 3 * L2);
    double *haver = (double *)malloc(sizeof(double) * L2);
    double *presentl = (double *)malloc(sizeof(double) * L2);
    double *clc = (double *)malloc(sizeof(double) * 3 * L2);
    double *clm = (double *)malloc(sizeof(double) * 3 * L2);
    double *clp = (double *)malloc(sizeof(double) * 3 * L2);
#pragma omp for schedule(static)
    for (int q = m->color_ptr[ccol]; q < m->color_ptr[ccol + 1]; q++) {
        int ie = m->color_elems[q];
        memset(hold, 0, sizeof(double) * 3 * L2);
        memset(hnew, 0, sizeof(double) * 3 * L2);
        memset(cl, 0, sizeof(double) * 3 * L2);
        memset(wl, 0, sizeof(double) * 3 * L2);
        memset(vflux, 0, sizeof(double) * 3 * L2);
        memset(haver, 0, sizeof(double) * L2);
        memset(presentl, 0, sizeof(double) * L2);
        memset(clc, 0, sizeof(double) * 3 * L2);
        memset(clm, 0, sizeof(double) * 3 * L2);
        memset(clp, 0, sizeof(double) * 3 * L2);
        int kn[3];
        double b[3], c[3];
        for (int ii = 0; ii < 3; ii++) {
            kn[ii] = m->nen3v[3 * ie + ii];
            b[ii] = EV(ii + 4, ie);
            c[ii] = EV(ii + 7, ie);
        }
        double aj = EV(10, ie), aj4 = 4. * aj;
        int ilevel = m->ilhv[ie], jlevel = 1;
        for (int l = jlevel; l <= ilevel; l++) {
            haver[l] = rso * A3E(s->hdenv, l, ie) + rsot * A3E(s->hdeov, l, ie);
            presentl[l] = 1.;
            for (int ii = 1; ii <= 3; ii++) {
                int k = kn[ii - 1];
                double hn = A3K(s->hdknv, l, k), ho = A3K(s->hdkov, l, k);
                H2(hold, l, ii) = rso * hn + rsot * ho;
                H2(hnew, l, ii) = rsn * hn + rsnt * ho;
                H2(cl, l, ii) = A3K(cn1, l, k);
                /* sinking enters the vertical velocity (concentration_omp.f90:479) */
                H2(wl, l, ii) = AW(s->wlnv, l, k) - (s->wsinkv ? s->wsink * AW(s->wsinkv, l, k) : (real)0);
            }
        }
        for (int ii = 1; ii <= 3; ii++) H2(wl, ilevel, ii) = 0.;
        /* vertical_flux_ie: tvd_admin.f90:930-1016 */
        for (int ii = 1; ii <= 3; ii++) {
            int lmin = jlevel, lmax = ilevel;
            for (int l = lmin; l <= lmax - 1; l++) {
                double w = H2(wl, l, ii);
                double conc = (w > 0.) ? H2(cl, l + 1, ii) : H2(cl, l, ii);
                double conf = conc;
                if (btvdv) {
                    double cond, conu;
                    if (w > 0.) {
                        cond = H2(cl, l, ii);
                        conu = cond;
                        if (l + 2 <= lmax) conu = H2(cl, l + 2, ii);
                    } else {
                        cond = H2(cl, l + 1, ii);
                        conu = cond;
                        if (l - 1 >= lmin) conu = H2(cl, l - 1, ii);
                    }
                    double hdis = 0.5 * (H2(hold, l, ii) + H2(hold, l + 1, ii));
                    double alfa = dt * fabs(w) / hdis;
                    double rf;
                    if (fabs(conc - cond) < (double)1.e-8f) rf = -1.;
                    else rf = (cond - conu) / (cond - conc) - 1.;
                    double psi = dmax(0., dmax(dmin(1., 2. * rf), dmin(2., rf)));
                    conf = conc + 0.5 * psi * (cond - conc) * (1. - alfa);
                }
                H2(vflux, l, ii) = w * conf;
            }
            H2(vflux, lmin - 1, ii) = 0.;
            H2(vflux, lmax, ii) = 0.;
        }
        int iext = 0;
        for (int ii = 0; ii < 3; ii++)
            if (m->iextb && m->iextb[kn[ii]]) iext++;
        for (int l = jlevel; l <= ilevel; l++) {
            double us = az * A3E(s->utlnv, l, ie) + azt * A3E(s->utlov, l, ie);
            double vs = az * A3E(s->vtlnv, l, ie) + azt * A3E(s->vtlov, l, ie);
            double rk3 = (real)3 * p->rkpar * A3E(s->difhv, l, ie);
            int itot = 0, isum = 0;
            double f[3], fw[3], fd[3], fl[3], wdiff[3];
            for (int ii = 1; ii <= 3; ii++) {
                int k = kn[ii - 1];
                f[ii - 1] = us * b[ii - 1] + vs * c[ii - 1];
                if (f[ii - 1] < 0.) { itot++; isum += ii; }
                double waux = 0.;
                for (int iii = 1; iii <= 3; iii++)
                    waux = waux + m->wdifhv[(iii - 1) + 3 * (ii - 1) + 9 * ie] * H2(cl, l, iii);
                wdiff[ii - 1] = waux;
                double rvptop = (real)(AW(s->difv, l - 1, k) + p->difmol);
                double rvpbot = (real)(AW(s->difv, l, k) + p->difmol);
                double hmotop = 2. * rvptop * presentl[l - 1] / (H2(hold, l - 1, ii) + H2(hold, l, ii));
                double hmobot = 2. * rvpbot * presentl[l + 1] / (H2(hold, l, ii) + H2(hold, l + 1, ii));
                double hmntop = 2. * rvptop * presentl[l - 1] / (H2(hnew, l - 1, ii) + H2(hnew, l, ii));
                double hmnbot = 2. * rvpbot * presentl[l + 1] / (H2(hnew, l, ii) + H2(hnew, l + 1, ii));
                fd[ii - 1] = adt * ((H2(cl, l, ii) - H2(cl, l + 1, ii)) * hmobot -
                                    (H2(cl, l - 1, ii) - H2(cl, l, ii)) * hmotop);
                H2(clc, l, ii) += ad * (hmntop + hmnbot);
                H2(clm, l, ii) -= ad * hmntop;
                H2(clp, l, ii) -= ad * hmnbot;
                double w = H2(wl, l - 1, ii);
                if (l == jlevel) w = 0.;
                if (w >= 0.) H2(clc, l, ii) += aa * w;
                else H2(clm, l, ii) += aa * w;
                w = H2(wl, l, ii);
                if (l == ilevel) w = 0.;
                if (w > 0.) H2(clp, l, ii) -= aa * w;
                else H2(clc, l, ii) -= aa * w;
                fw[ii - 1] = aat * (H2(vflux, l - 1, ii) - H2(vflux, l, ii));
            }
            {
                const double cll[3] = {H2(cl, l, 1), H2(cl, l, 2), H2(cl, l, 3)};
                hflux(m, s, ie, l, itot, isum, dt, cll, cn1, btvd && iext == 0, gx, gy, f, fl);
            }
            for (int ii = 1; ii <= 3; ii++) {
                int k = kn[ii - 1];
                double hmed = haver[l];
                /* nudging towards cobs, explicit (an = 0: ant = 1, finu = 0;
                 * concentration_omp.f90:203-208,697-701) */
                double fnudge = 0.;
                if (s->cobs && s->rtauv)
                    fnudge = s->robs * (double)A3K(s->rtauv, l, k) * ((double)A3K(s->cobs, l, k) - 1. * H2(cl, l, ii));
                double cexpl = aj4 * (H2(hold, l, ii) * H2(cl, l, ii) +
                                      dt * (H2(hold, l, ii) * fnudge + 3. * fl[ii - 1] - fw[ii - 1] -
                                            rk3 * hmed * wdiff[ii - 1] - fd[ii - 1]));
                double alow = aj4 * dt * H2(clm, l, ii);
                double ahigh = aj4 * dt * H2(clp, l, ii);
                double adiag = aj4 * dt * H2(clc, l, ii) + aj4 * (1. + dt * 0.) * H2(hnew, l, ii);
                A3K(cn, l, k) += cexpl;
                A3K(clow, l, k) += alow;
                A3K(chigh, l, k) += ahigh;
                A3K(cdiag, l, k) += adiag;
            }
        }
    }
    free(hold); free(hnew); free(cl); free(wl); free(vflux); free(haver); free(presentl);
    free(clc); free(clm); free(clp);
    }
    }
    /* conz3d_nodes: every node on its own */
    int ret = 0;
#pragma omp parallel for schedule(static) reduction(| : ret)
    for (int k = 0; k < nkn; k++) {
        int ilevel = m->ilhkv[k], jlevel = 1;
        for (int l = jlevel; l <= ilevel; l++) {
            double cconz = cbound ? A3K(cbound, l, k) : 0.;
            double qflux = s->mfluxv ? A3K(s->mfluxv, l, k) : 0.;
            if (qflux < 0. && m->ibnd && m->ibnd[k] > 0) cconz = A3K(cn1, l, k);
            double mflux = qflux * cconz;
            A3K(cn, l, k) += dt * mflux;
            /* loads: erosion or, negative, deposition bounded by what the
             * node holds (concentration_omp.f90:816-834) */
            if (s->load) {
                double loading = (double)(s->rload * A3K(s->load, l, k));
                if (loading == 0.) {
                } else if (loading < 0.) {
                    double cload = 0.;
                    if (A3K(cn, l, k) > 0.) {
                        cload = -dt * loading;
                        cload = A3K(cn, l, k) * (1. - exp(-cload / A3K(cn, l, k)));
                    } else if (A3K(cn, l, k) < 0.) {
                        A3K(cn, l, k) = 0.;
                    }
                    if (cload > A3K(cn, l, k)) ret |= 4;
                    loading = -cload / dt;
                    if (s->rload > 0) A3K(s->load, l, k) = (real)(loading / s->rload);
                    A3K(cn, l, k) += dt * loading;
                } else {
                    A3K(cn, l, k) += dt * loading;
                }
            }
        }
        if ((aa == 0. && ad == 0.) || m->nlv == 1) {
            for (int l = jlevel; l <= ilevel; l++)
                if (A3K(cdiag, l, k) != 0.) A3K(cn, l, k) = A3K(cn, l, k) / A3K(cdiag, l, k);
        } else {
            double aux = 1. / A3K(cdiag, jlevel, k);
            A3K(chigh, jlevel, k) *= aux;
            A3K(cn, jlevel, k) *= aux;
            for (int l = jlevel + 1; l <= ilevel; l++) {
                if (A3K(cdiag, l, k) == 0.) { ret = 1; break; }
                aux = 1. / (A3K(cdiag, l, k) - A3K(clow, l, k) * A3K(chigh, l - 1, k));
                A3K(chigh, l, k) *= aux;
                A3K(cn, l, k) = (A3K(cn, l, k) - A3K(clow, l, k) * A3K(cn, l - 1, k)) * aux;
            }
            for (int l = ilevel - 1; l >= jlevel; l--)
                A3K(cn, l, k) = A3K(cn, l, k) - A3K(cn, l + 1, k) * A3K(chigh, l, k);
        }
    }
    if (!ret)
        for (size_t i = 0; i < nn; i++) cn1[i] = (real)cn[i];
#undef H2
    free(cn); free(cdiag); free(clow); free(chigh);
    free(gx); free(gy);
    return ret;
}

/* make_scal_flux: bnd_routines.f90:1215-1303 (ssurf given, rqdsv = 0) */
static void make_scal_flux(const oc_mesh *m, const oc_state *s, const real *r3v, const real *scal,
                           real ssurf, real *sconz)
{
    const int nlvdi = m->nlvdi;
    for (int k = 0; k < m->nkn; k++) {
        int lmax = m->ilhkv[k], lmin = 1;
        real sflux_lmin = 0;
        for (int l = lmin; l <= lmax; l++) {
            real sflux = 0;
            A3K(sconz, l, k) = 0;
            real flux = s->mfluxv ? A3K(s->mfluxv, l, k) : (real)0;
            real conz = r3v ? A3K(r3v, l, k) : OC_FLAG;
            if (flux != 0 || conz != OC_FLAG) {
                if (conz <= (real)-990.) conz = A3K(scal, l, k);
                if (flux < 0) conz = A3K(scal, l, k);
                sflux = flux * conz;
                A3K(sconz, l, k) = conz;
            }
            if (l == lmin) sflux_lmin = sflux;
        }
        real conz = ssurf;
        if (ssurf <= (real)-990) conz = A3K(scal, lmin, k);
        sflux_lmin = sflux_lmin + (real)0 * conz;
        real mfl = s->mfluxv ? A3K(s->mfluxv, lmin, k) : (real)0;
        if (mfl != 0) A3K(sconz, lmin, k) = sflux_lmin / mfl;
    }
}

static int cmp_ll(const void *a, const void *b)
{
    long long x = *(const long long *)a, y = *(const long long *)b;
    return (x > y) - (x < y);
}

/* bndo_init: bndo.f90:210-470.  Open-boundary nodes in boundary order
 * (type 1 and 2 boundaries are "external"); the inward normal comes from the
 * neighbouring boundary nodes along the basin boundary (kantv, built from the
 * counter-clockwise node lists of connections.f90:594-612); the inner
 * neighbours are the non-boundary vertices of elements touching the node,
 * weighted by element area and normalised.  Returns nbndo or -1. */
int oc_bndo_init(const oc_mesh *m, const real *xgv, const real *ygv, int nb, const int *bnodes,
                 int *bndo_k, real *nx, real *ny, int *ptr, int *nbr, real *w, int maxnb)
{
    int nkn = m->nkn, nel = m->nel;
    int *iop = (int *)calloc((size_t)nkn, sizeof(int));
    int *knext = (int *)malloc(sizeof(int) * (size_t)nkn);
    int *klast = (int *)malloc(sizeof(int) * (size_t)nkn);
    for (int k = 0; k < nkn; k++) { knext[k] = -1; klast[k] = -1; }
    /* boundary edges (one element only) in element orientation */
    long long *key = (long long *)malloc(sizeof(long long) * 3 * (size_t)nel);
    for (int ie = 0; ie < nel; ie++)
        for (int ii = 0; ii < 3; ii++) {
            long long a = m->nen3v[3 * ie + ii], c = m->nen3v[3 * ie + (ii + 1) % 3];
            long long lo = a < c ? a : c, hi = a < c ? c : a;
            key[3 * ie + ii] = lo * nkn + hi;
        }
    long long *sorted = (long long *)malloc(sizeof(long long) * 3 * (size_t)nel);
    memcpy(sorted, key, sizeof(long long) * 3 * (size_t)nel);
    qsort(sorted, 3 * (size_t)nel, sizeof(long long), cmp_ll);
    for (int ie = 0; ie < nel; ie++)
        for (int ii = 0; ii < 3; ii++) {
            long long kk = key[3 * ie + ii];
            size_t lo = 0, hi = 3 * (size_t)nel;
            while (lo < hi) { size_t mid = (lo + hi) / 2; if (sorted[mid] < kk) lo = mid + 1; else hi = mid; }
            int cnt = 0;
            while (lo < 3 * (size_t)nel && sorted[lo] == kk) { cnt++; lo++; }
            if (cnt == 1) {
                int k1 = m->nen3v[3 * ie + ii], k2 = m->nen3v[3 * ie + (ii + 1) % 3];
                knext[k1] = k2;
                klast[k2] = k1;
            }
        }
    free(key); free(sorted);
    for (int i = 0; i < nb; i++) { bndo_k[i] = bnodes[i]; iop[bnodes[i]] = i + 1; }
    for (int i = 0; i < nb; i++) {
        int k = bnodes[i];
        int kn = knext[k], kl = klast[k];
        int inext = (kn >= 0) ? iop[kn] : 0, ilast = (kl >= 0) ? iop[kl] : 0;
        real dx, dy;
        if (inext > 0 && ilast > 0) { dx = xgv[kn] - xgv[kl]; dy = ygv[kn] - ygv[kl]; }
        else if (inext > 0) { dx = xgv[kn] - xgv[k]; dy = ygv[kn] - ygv[k]; }
        else if (ilast > 0) { dx = xgv[k] - xgv[kl]; dy = ygv[k] - ygv[kl]; }
        else { free(iop); free(knext); free(klast); return -1; }
        nx[i] = -dy;
        ny[i] = dx;
    }
    /* neighbour lists in element order (bndinsert) */
    int *cnt = (int *)calloc((size_t)nb, sizeof(int));
    for (int i = 0; i < nb * maxnb; i++) { nbr[i] = -1; w[i] = 0; }
    for (int ie = 0; ie < nel; ie++) {
        real area = (real)(12. * m->ev[9 + 19 * ie]);
        for (int ii = 0; ii < 3; ii++) {
            int ib = iop[m->nen3v[3 * ie + ii]];
            if (ib < 1) continue;
            for (int iii = 0; iii < 3; iii++) {
                int kn = m->nen3v[3 * ie + iii];
                if (iop[kn] > 0) continue;
                int *lst = nbr + (ib - 1) * maxnb;
                real *wl = w + (ib - 1) * maxnb;
                int j, found = 0;
                for (j = 0; j < cnt[ib - 1]; j++)
                    if (lst[j] == kn) { wl[j] = wl[j] + area; found = 1; break; }
                if (!found) {
                    if (cnt[ib - 1] >= maxnb) { free(iop); free(knext); free(klast); free(cnt); return -1; }
                    lst[cnt[ib - 1]] = kn;
                    wl[cnt[ib - 1]] = area;
                    cnt[ib - 1]++;
                }
            }
        }
    }
    /* compact into CSR and normalise */
    ptr[0] = 0;
    for (int i = 0; i < nb; i++) {
        real tot = 0;
        for (int j = 0; j < cnt[i]; j++) tot = tot + w[i * maxnb + j];
        for (int j = 0; j < cnt[i]; j++) {
            nbr[ptr[i] + j] = nbr[i * maxnb + j];
            w[ptr[i] + j] = (tot > 0) ? w[i * maxnb + j] / tot : w[i * maxnb + j];
        }
        ptr[i + 1] = ptr[i] + cnt[i];
    }
    free(iop); free(knext); free(klast); free(cnt);
    return nb;
}

/* uvtopr: transforms.f90:215-270 (nodal velocities of wet elements) */
void oc_uvtopr(const oc_mesh *m, const oc_state *s, real *uprv, real *vprv)
{
    const int nlvdi = m->nlvdi;
    size_t nn = (size_t)nlvdi * m->nkn;
    real *vv = (real *)calloc(nn, sizeof(real));
    memset(uprv, 0, sizeof(real) * nn);
    memset(vprv, 0, sizeof(real) * nn);
    for (int ie = 0; ie < m->nel; ie++) {
        if (s->iwegv[ie] != 0) continue;
        real aj = (real)EV(10, ie);
        for (int l = 1; l <= m->ilhv[ie]; l++)
            for (int ii = 0; ii < 3; ii++) {
                int k = m->nen3v[3 * ie + ii];
                A3K(vv, l, k) = A3K(vv, l, k) + aj;
                A3K(uprv, l, k) = A3K(uprv, l, k) + aj * A3E(s->ulnv, l, ie);
                A3K(vprv, l, k) = A3K(vprv, l, k) + aj * A3E(s->vlnv, l, ie);
            }
    }
    for (size_t i = 0; i < nn; i++)
        if (vv[i] > 0) { uprv[i] = uprv[i] / vv[i]; vprv[i] = vprv[i] / vv[i]; }
    free(vv);
}

/* bndo_setbc = bndo_impbc + bndo_adjbc (bndo.f90:620-890) for level (type 1)
 * boundaries: impose rbc where it is not the -999 flag, then on outflow
 * (n . u <= 0) or ambient flag take the weighted mean of the inner
 * neighbours.  Returns 0, or 1 on a node without neighbours. */
int oc_bndo_setbc(const oc_mesh *m, real *cv, const real *rbc, const real *uprv, const real *vprv)
{
    const int nlvdi = m->nlvdi;
    const real flag = (real)-999.;
    for (int i = 0; i < m->nbndo; i++) {
        int k = m->bndo_k[i];
        int ibc = m->ibnd[k];
        if (m->ibtyp[ibc] != 1) continue;
        int lmax = m->ilhkv[k];
        if (rbc)
            for (int l = 1; l <= lmax; l++) {
                real rb = A3K(rbc, l, k);
                if (rb != flag) A3K(cv, l, k) = rb;
            }
        real dx = m->bndo_nx[i], dy = m->bndo_ny[i];
        for (int l = 1; l <= lmax; l++) {
            real scal = dx * A3K(uprv, l, k) + dy * A3K(vprv, l, k);
            int bout = scal <= 0;
            int bamb = A3K(cv, l, k) == flag;
            real bc = 0, tweight = 0;
            int ntbc;
            if (bout || bamb) {
                ntbc = 0;
                for (int j = m->bndo_ptr[i]; j < m->bndo_ptr[i + 1]; j++) {
                    int ko = m->bndo_nb[j];
                    if (l <= m->ilhkv[ko]) {
                        ntbc++;
                        real weight = m->bndo_w[j];
                        tweight = tweight + weight;
                        bc = bc + weight * A3K(cv, l, ko);
                    }
                }
            } else {
                ntbc = m->bndo_ptr[i + 1] - m->bndo_ptr[i];
                tweight = 1;
                bc = A3K(cv, l, k);
            }
            if (ntbc > 0) A3K(cv, l, k) = bc / tweight;
            else if (l > 1) A3K(cv, l, k) = A3K(cv, l - 1, k);
            else return 1;
        }
    }
    return 0;
}

/* scalar_stability (stability.f90:220-260): scalar_compute_stability
 * (:117-177) runs conzstab (concentration.f90:1573-2097) for one second
 * (ddt = 1, istot = isact = 1) with difmol = 0 and without nudging or
 * sinking.  conzstab sums per node and layer the advective flux (the larger
 * of the last in- and out-flux seen over the element's vertices so far, as
 * the reference carries them from vertex to vertex), horizontal and explicit
 * vertical diffusion and point sources, divides by the node volume and takes
 * the maximum over the layers of the nodes that are not level boundaries,
 * scaled by 1/rstol.  Then rindex = dt * that and istot = 1 + rindex,
 * truncated.  Returns istot; *rindex receives rindex. */
int oc_scalar_stability(const oc_mesh *m, const oc_state *s, const oc_params *p, real dt, real rstol,
                        real *rindex)
{
    const int nlvdi = m->nlvdi, nkn = m->nkn;
    double az = p->azpar, ad = p->adpar, aa = p->aapar;
    double azt = 1. - az, adt = 1. - ad, aat = 1. - aa;
    const double ddt = 1.;
    const double rstot = 1., rso = 0. / rstot, rsn = 1. / rstot;
    const double rsot = 1. - rso, rsnt = 1. - rsn;
    const double dtt = ddt / rstot;
    const real difmol = 0, rkpar = p->rkpar;
    size_t nn = (size_t)nlvdi * nkn;
    double *cn = (double *)calloc(nn, sizeof(double));
    double *co = (double *)calloc(nn, sizeof(double));
    double *cdiag = (double *)calloc(nn, sizeof(double));
    double *clow = (double *)calloc(nn, sizeof(double));
    double *chigh = (double *)calloc(nn, sizeof(double));
    int L2 = nlvdi + 2;
    double *hdv = (double *)calloc((size_t)L2, sizeof(double));
    double *presentl = (double *)calloc((size_t)L2, sizeof(double));
    double *hold = (double *)calloc((size_t)3 * L2, sizeof(double));
    double *hnew = (double *)calloc((size_t)3 * L2, sizeof(double));
    double *wl = (double *)calloc((size_t)3 * L2, sizeof(double));
#define H2(a, l, ii) (a)[(l) + L2 * ((ii) - 1)]
    if (s->mfluxv)
        for (int k = 0; k < nkn; k++)
            for (int l = 1; l <= m->nlv; l++)
                if (A3K(s->mfluxv, l, k) > 0) A3K(co, l, k) = A3K(s->mfluxv, l, k);
    /* the layer arrays keep what earlier elements left below this one's
     * bottom, as the reference's do; `present` is zero there */
    for (int ie = 0; ie < m->nel; ie++) {
        int kn[3];
        double b[3], c[3];
        for (int ii = 0; ii < 3; ii++) {
            kn[ii] = m->nen3v[3 * ie + ii];
            b[ii] = EV(ii + 4, ie);
            c[ii] = EV(ii + 7, ie);
        }
        double aj = EV(10, ie), aj4 = 4. * aj;
        int ilevel = m->ilhv[ie], jlevel = 1;
        for (int l = jlevel; l <= ilevel; l++) {
            hdv[l] = A3E(s->hdeov, l, ie);
            presentl[l] = 1.;
            for (int ii = 1; ii <= 3; ii++) {
                int k = kn[ii - 1];
                double hn = A3K(s->hdknv, l, k), ho = A3K(s->hdkov, l, k);
                H2(hold, l, ii) = rso * hn + rsot * ho;
                H2(hnew, l, ii) = rsn * hn + rsnt * ho;
                H2(wl, l, ii) = AW(s->wlnv, l, k) - (s->wsinkv ? s->wsink * AW(s->wsinkv, l, k) : (real)0);
            }
        }
        for (int l = ilevel + 1; l <= m->nlv; l++) presentl[l] = 0.;
        for (int ii = 1; ii <= 3; ii++) H2(wl, ilevel, ii) = 0.;
        for (int l = jlevel; l <= ilevel; l++) {
            double us = az * A3E(s->utlnv, l, ie) + azt * A3E(s->utlov, l, ie);
            double vs = az * A3E(s->vtlnv, l, ie) + azt * A3E(s->vtlov, l, ie);
            double rk3 = (real)3 * rkpar * (s->difhv ? A3E(s->difhv, l, ie) : (real)0);
            double f[3], fwin[3], fwout[3], fd[3], wdiff[3];
            for (int ii = 1; ii <= 3; ii++) {
                int k = kn[ii - 1];
                wdiff[ii - 1] = m->wdifhv ? m->wdifhv[(ii - 1) + 3 * (ii - 1) + 9 * ie] : 0.;
                f[ii - 1] = us * b[ii - 1] + vs * c[ii - 1];
                fwin[ii - 1] = 0.;
                fwout[ii - 1] = 0.;
                double w = H2(wl, l - 1, ii);
                if (w > 0.) fwout[ii - 1] = fwout[ii - 1] + aat * w;
                else fwin[ii - 1] = fwin[ii - 1] - aat * w;
                w = H2(wl, l, ii);
                if (w > 0.) fwin[ii - 1] = fwin[ii - 1] + aat * w;
                else fwout[ii - 1] = fwout[ii - 1] - aat * w;
                double rvptop = (double)(AW(s->difv, l - 1, k) + difmol);
                double rvpbot = (double)(AW(s->difv, l, k) + difmol);
                double hmtop = 2. * rvptop * presentl[l - 1] / (hdv[l - 1] + hdv[l]);
                double hmbot = 2. * rvpbot * presentl[l + 1] / (hdv[l] + hdv[l + 1]);
                fd[ii - 1] = adt * (hmtop + hmbot);
            }
            double coadv = 0., ciadv = 0.;
            for (int ii = 1; ii <= 3; ii++) {
                int k = kn[ii - 1];
                double hmed = H2(hold, l, ii);
                double chdiff = dtt * aj4 * rk3 * hmed * wdiff[ii - 1];
                double cvdiff = dtt * aj4 * fd[ii - 1];
                double chadv = dtt * aj4 * 3. * f[ii - 1];
                double cviadv = dtt * aj4 * fwin[ii - 1];
                double cvoadv = dtt * aj4 * fwout[ii - 1];
                if (chadv < 0.) coadv = cvoadv - chadv;
                else ciadv = cviadv + chadv;
                double cadv = dmax(coadv, ciadv);
                A3K(chigh, l, k) += cadv;
                A3K(clow, l, k) += chdiff;
                A3K(cn, l, k) += cvdiff;
                if (s->rtauv) A3K(co, l, k) += dtt * aj4 * hmed * s->robs * (double)A3K(s->rtauv, l, k);
            }
        }
        for (int l = jlevel; l <= ilevel; l++)
            for (int ii = 1; ii <= 3; ii++) {
                int k = kn[ii - 1];
                double hmed = dmin(H2(hold, l, ii), H2(hnew, l, ii));
                A3K(cdiag, l, k) += aj4 * hmed;
            }
    }
    double stabind = 0.;
    for (int k = 0; k < nkn; k++) {
        int ibc = m->ibnd ? m->ibnd[k] : 0;
        if (ibc > 0 && m->ibtyp[ibc] == 1) continue;          /* is_zeta_bound */
        for (int l = 1; l <= m->ilhkv[k]; l++) {
            double voltot = A3K(cdiag, l, k);
            double flxtot = A3K(chigh, l, k) + A3K(clow, l, k) + A3K(cn, l, k) + A3K(co, l, k);
            if (voltot > 0.) stabind = dmax(stabind, flxtot / voltot);
        }
    }
    float raux = (float)stabind;                             /* shympi_max of a real */
    stabind = raux;
    real sindex = (real)(stabind / (double)rstol);
    real ri = dt * sindex;
    free(cn); free(co); free(cdiag); free(clow); free(chigh);
    free(hdv); free(presentl); free(hold); free(hnew); free(wl);
#undef H2
    *rindex = ri;
    return (int)((real)1 + ri);
}

/* scal3sh substep loop: concentration.f90:565-810.  istot <= 0: the
 * number of substeps from scalar_stability (rstol = 1, the STR default), as
 * the reference sets it (:678-700); more than istot_max = 1000 is the
 * reference's error stop (returns 3). */
int oc_scal3sh(const oc_mesh *m, oc_state *s, const oc_params *p, real *cnv,
               const real *rcv, int istot, int clip_small)
{
    if (istot <= 0) {
        real rindex;
        istot = oc_scalar_stability(m, s, p, (real)p->dt, (real)1, &rindex);
        if (istot > 1000) return 3;
    }
    size_t nn = (size_t)m->nlvdi * m->nkn;
    real *sbconz = (real *)calloc(nn, sizeof(real));
    int ret = 0;
    real ssurface = clip_small ? (real)0 : (real)-999.;
    real *uprv = (real *)malloc(sizeof(real) * nn), *vprv = (real *)malloc(sizeof(real) * nn);
    oc_uvtopr(m, s, uprv, vprv);                 /* make_prvel at the end of hydro */
    for (int isact = 1; isact <= istot; isact++) {
        make_scal_flux(m, s, rcv, cnv, ssurface, sbconz);
        if (clip_small)
            for (size_t i = 0; i < nn; i++)
                if (cnv[i] < (real)1.e-15) cnv[i] = 0;
        ret = oc_conz3d(m, s, p, cnv, sbconz, (double)(real)p->dt, istot, isact);
        if (ret) break;
        if (m->nbndo > 0 && oc_bndo_setbc(m, cnv, rcv, uprv, vprv)) { ret = 2; break; }
    }
    free(uprv); free(vprv);
    free(sbconz);
    return ret;
}

/* total tracer content sum_k sum_l c * vol (FP64 accumulation) */
double oc_massconc(const oc_mesh *m, const oc_state *s, const real *cnv, int mode)
{
    const int nlvdi = m->nlvdi;
    const real *hd = mode > 0 ? s->hdknv : s->hdkov;
    double total = 0.;
    for (int k = 0; k < m->nkn; k++)
        for (int l = 1; l <= m->ilhkv[k]; l++)
            total += (double)A3K(cnv, l, k) * (double)(A3K(hd, l, k) * A3K(m->areakv, l, k));
    return total;
}
