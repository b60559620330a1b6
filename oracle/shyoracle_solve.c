/*
 * shyoracle_solve.c -- zeta-system solver of the CPU oracle (TEST INFRASTRUCTURE).
 *
 * Restates the reference's SPARSKIT path used by spk_solve_system
 * (src/shyfem/sparskit_admin.f90:115-330): COO->CSR (the COO pattern of
 * coo_init_new is already row-sorted, so coocsr+csort is the identity on the
 * ordering), ILUT(lfil=5, droptol=1e-2) or ILU(0), and BiCGStab driven by
 * runrc with ipar(2)=1 -- which bisinit (spk_iters.f90:3433-3500) maps to
 * LEFT preconditioning -- and stopping test scheme 1 on the preconditioned
 * residual: ||M^-1 r|| <= rtol*||M^-1 r0|| + atol, at most itermax matvecs.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "shyoracle.h"

/* qsplit: spk_ilut.f90:2413-2470, 1-based a(1:n), ind(1:n) passed as a-1 */
static void qsplit(double *a, int *ind, int n, int ncut)
{
    int first = 1, last = n, mid;
    if (ncut < first || ncut > last) return;
    for (;;) {
        mid = first;
        double abskey = fabs(a[mid]);
        for (int j = first + 1; j <= last; j++) {
            if (fabs(a[j]) > abskey) {
                mid = mid + 1;
                double tmp = a[mid];
                int itmp = ind[mid];
                a[mid] = a[j];
                ind[mid] = ind[j];
                a[j] = tmp;
                ind[j] = itmp;
            }
        }
        double tmp = a[mid];
        a[mid] = a[first];
        a[first] = tmp;
        int itmp = ind[mid];
        ind[mid] = ind[first];
        ind[first] = itmp;
        if (mid == ncut) return;
        if (mid > ncut) last = mid - 1;
        else first = mid + 1;
    }
}

/* ilut: spk_ilut.f90:78-318.  Input CSR is 0-based; the MSR output (alu,jlu,ju)
 * uses the reference's 1-based convention stored from index 1 (alu[0] unused).
 * alu/jlu must hold iwk+1 entries, ju n+1.  Returns ierr. */
int oc_ilut(int n, const double *a, const int *ja, const int *ia, int lfil, double droptol,
            double *alu, int *jlu, int *ju, int iwk)
{
    if (lfil < 0) return -4;
    double *w = (double *)malloc(sizeof(double) * (size_t)(n + 2));
    int *jw = (int *)malloc(sizeof(int) * (size_t)(2 * n + 2));
    int ierr = 0;
    int ju0 = n + 2;
    jlu[1] = ju0;
    for (int j = 1; j <= n; j++) jw[n + j] = 0;
    for (int ii = 1; ii <= n; ii++) {
        int j1 = ia[ii - 1] + 1, j2 = ia[ii];
        double tnorm = 0.0;
        for (int k = j1; k <= j2; k++) tnorm += fabs(a[k - 1]);
        if (tnorm == 0.0) { ierr = -5; goto out; }
        tnorm = tnorm / (double)(float)(j2 - j1 + 1);
        int lenu = 1, lenl = 0;
        jw[ii] = ii;
        w[ii] = 0.0;
        jw[n + ii] = ii;
        for (int j = j1; j <= j2; j++) {
            int k = ja[j - 1] + 1;
            double t = a[j - 1];
            if (k < ii) {
                lenl++;
                jw[lenl] = k;
                w[lenl] = t;
                jw[n + k] = lenl;
            } else if (k == ii) {
                w[ii] = t;
            } else {
                lenu++;
                int jpos = ii + lenu - 1;
                jw[jpos] = k;
                w[jpos] = t;
                jw[n + k] = jpos;
            }
        }
        int jj = 0, len = 0;
        for (;;) {
            jj++;
            if (jj > lenl) break;
            int jrow = jw[jj], k = jj;
            for (int j = jj + 1; j <= lenl; j++)
                if (jw[j] < jrow) { jrow = jw[j]; k = j; }
            if (k != jj) {
                int j = jw[jj];
                jw[jj] = jw[k];
                jw[k] = j;
                jw[n + jrow] = jj;
                jw[n + j] = k;
                double s = w[jj];
                w[jj] = w[k];
                w[k] = s;
            }
            jw[n + jrow] = 0;
            double fact = w[jj] * alu[jrow];
            if (fabs(fact) <= droptol) continue;
            for (int kk = ju[jrow]; kk <= jlu[jrow + 1] - 1; kk++) {
                double s = fact * alu[kk];
                int j = jlu[kk];
                int jpos = jw[n + j];
                if (j >= ii) {
                    if (jpos == 0) {
                        lenu++;
                        if (lenu > n) { ierr = -1; goto out; }
                        int i = ii + lenu - 1;
                        jw[i] = j;
                        jw[n + j] = i;
                        w[i] = -s;
                    } else {
                        w[jpos] = w[jpos] - s;
                    }
                } else {
                    if (jpos == 0) {
                        lenl++;
                        if (lenl > n) { ierr = -1; goto out; }
                        jw[lenl] = j;
                        jw[n + j] = lenl;
                        w[lenl] = -s;
                    } else {
                        w[jpos] = w[jpos] - s;
                    }
                }
            }
            len++;
            w[len] = fact;
            jw[len] = jrow;
        }
        for (int k = 1; k <= lenu; k++) jw[n + jw[ii + k - 1]] = 0;
        lenl = len;
        len = lenl < lfil ? lenl : lfil;
        qsplit(w, jw, lenl, len);
        for (int k = 1; k <= len; k++) {
            if (ju0 > iwk) { ierr = -2; goto out; }
            alu[ju0] = w[k];
            jlu[ju0] = jw[k];
            ju0++;
        }
        ju[ii] = ju0;
        len = 0;
        for (int k = 1; k <= lenu - 1; k++) {
            if (fabs(w[ii + k]) > droptol * tnorm) {
                len++;
                w[ii + len] = w[ii + k];
                jw[ii + len] = jw[ii + k];
            }
        }
        lenu = len + 1;
        len = lenu < lfil ? lenu : lfil;
        qsplit(w + ii, jw + ii, lenu - 1, len);
        double t = fabs(w[ii]);
        if (len + ju0 > iwk) { ierr = -3; goto out; }
        for (int k = ii + 1; k <= ii + len - 1; k++) {
            jlu[ju0] = jw[k];
            alu[ju0] = w[k];
            t = t + fabs(w[k]);
            ju0++;
        }
        (void)t;
        if (w[ii] == 0.0) w[ii] = ((double)0.0001f + droptol) * tnorm;
        alu[ii] = 1.0 / w[ii];
        jlu[ii + 1] = ju0;
    }
out:
    free(w);
    free(jw);
    return ierr;
}

/* ilu0: spk_ilut.f90:1832-1950 (same 1-based MSR output as oc_ilut) */
int oc_ilu0(int n, const double *a, const int *ja, const int *ia, double *alu, int *jlu, int *ju)
{
    int *iw = (int *)calloc((size_t)n + 2, sizeof(int));
    int ju0 = n + 2, ierr = 0;
    jlu[1] = ju0;
    for (int ii = 1; ii <= n; ii++) {
        int js = ju0;
        for (int j = ia[ii - 1] + 1; j <= ia[ii]; j++) {
            int jcol = ja[j - 1] + 1;
            if (jcol == ii) {
                alu[ii] = a[j - 1];
                iw[jcol] = ii;
                ju[ii] = ju0;
            } else {
                alu[ju0] = a[j - 1];
                jlu[ju0] = jcol;
                iw[jcol] = ju0;
                ju0++;
            }
        }
        jlu[ii + 1] = ju0;
        int jf = ju0 - 1, jm = ju[ii] - 1;
        for (int j = js; j <= jm; j++) {
            int jrow = jlu[j];
            double tl = alu[j] * alu[jrow];
            alu[j] = tl;
            for (int jj = ju[jrow]; jj <= jlu[jrow + 1] - 1; jj++) {
                int jw = iw[jlu[jj]];
                if (jw != 0) alu[jw] = alu[jw] - tl * alu[jj];
            }
        }
        if (alu[ii] == 0.0) { ierr = ii; break; }
        alu[ii] = 1.0 / alu[ii];
        iw[ii] = 0;
        for (int i = js; i <= jf; i++) iw[jlu[i]] = 0;
    }
    free(iw);
    return ierr;
}

/* lusol: spk_ilut.f90:2309-2345; y, x are 0-based vectors */
void oc_lusol(int n, const double *y, double *x, const double *alu, const int *jlu, const int *ju)
{
    for (int i = 1; i <= n; i++) {
        double xi = y[i - 1];
        for (int k = jlu[i]; k <= ju[i] - 1; k++) xi = xi - alu[k] * x[jlu[k] - 1];
        x[i - 1] = xi;
    }
    for (int i = n; i >= 1; i--) {
        double xi = x[i - 1];
        for (int k = ju[i]; k <= jlu[i + 1] - 1; k++) xi = xi - alu[k] * x[jlu[k] - 1];
        x[i - 1] = alu[i] * xi;
    }
}

/* amux: spk_matvec.f90:74-116 (0-based CSR) */
void oc_amux(int n, const double *x, double *y, const double *a, const int *ja, const int *ia)
{
    for (int i = 0; i < n; i++) {
        double t = 0.0;
        for (int k = ia[i]; k < ia[i + 1]; k++) t = t + a[k] * x[ja[k]];
        y[i] = t;
    }
}

static double dot(int n, const double *x, const double *y)
{
    double t = 0.0;
    for (int i = 0; i < n; i++) t = t + x[i] * y[i];
    return t;
}

/* brkdn: spk_iters.f90 (alpha == 0 -> -3; not finite -> -9) */
static int brkdn(double alpha, int *ierr)
{
    if (alpha > 0.) {
        if (!(1. / alpha > 0.)) { *ierr = -9; return 1; }
    } else if (alpha < 0.) {
        if (!(1. / alpha < 0.)) { *ierr = -9; return 1; }
    } else if (alpha == 0.) {
        *ierr = -3;
        return 1;
    } else {
        *ierr = -9;
        return 1;
    }
    return 0;
}

/* bcgstab with lp=.true., rp=.false. (spk_iters.f90:1061-1334 as driven by
 * runrc, spk_itaux.f90:43-120), convergence scheme ipar(3)=1.  The reverse
 * communication is unrolled into straight-line code; ipar(7) counts matvecs.
 * Returns ipar(1) at exit: 0 converged, -1 too many iterations, -3/-9 breakdown. */
int oc_bcgstab(int n, const double *a, const int *ja, const int *ia,
               const double *alu, const int *jlu, const int *ju,
               const double *rhs, double *sol, double rtol, double atol, int itermax,
               int *iters)
{
    double *w = (double *)calloc((size_t)8 * n, sizeof(double));
    double *r0 = w, *r = w + n, *s = w + 2 * n, *t = w + 3 * n, *v = w + 4 * n, *p = w + 5 * n,
           *tmp = w + 6 * n, *dx = w + 7 * n;
    int nmv = 0, ierr = 0, conv = 0;
    double f3, f4, f5, f6, rho, alpha, omega;
    /* initial residual (label 10 and 20) */
    oc_amux(n, sol, r, a, ja, ia);
    nmv++;
    for (int i = 0; i < n; i++) r0[i] = rhs[i] - r[i];
    oc_lusol(n, r0, r, alu, jlu, ju); /* w(:,2) = M^-1 w(:,1) */
    for (int i = 0; i < n; i++) { r0[i] = r[i]; p[i] = r[i]; }
    rho = dot(n, r0, r0);
    f5 = sqrt(rho);
    f3 = f5;
    f4 = rtol * f3 + atol;
    f6 = f5;
    if (f5 <= f4) { conv = 1; goto done; }
    for (;;) {
        /* step (1): v = M^-1 A p */
        oc_amux(n, p, tmp, a, ja, ia);
        oc_lusol(n, tmp, v, alu, jlu, ju);
        nmv++;
        alpha = dot(n, r0, v);
        if (brkdn(alpha, &ierr)) goto done;
        alpha = rho / alpha;
        for (int i = 0; i < n; i++) s[i] = r[i] - alpha * v[i];
        /* step (4): t = M^-1 A s */
        oc_amux(n, s, tmp, a, ja, ia);
        oc_lusol(n, tmp, t, alu, jlu, ju);
        nmv++;
        omega = dot(n, t, t);
        if (brkdn(omega, &ierr)) goto done;
        omega = dot(n, t, s) / omega;
        if (brkdn(omega, &ierr)) goto done;
        for (int i = 0; i < n; i++) {
            tmp[i] = alpha * p[i] + omega * s[i];
            dx[i] = dx[i] + tmp[i];
            r[i] = s[i] - omega * t[i];
        }
        /* stopbis */
        if (itermax > 0 && nmv >= itermax) { ierr = -1; goto done; }
        f5 = sqrt(dot(n, r, r));
        f6 = f5;
        if (!(f6 > f4)) { conv = 1; goto done; }
        /* step (8) */
        double rho_old = rho;
        rho = dot(n, r, r0);
        double beta = rho * alpha / (omega * rho_old);
        for (int i = 0; i < n; i++) p[i] = r[i] + beta * (p[i] - omega * v[i]);
        if (brkdn(rho, &ierr)) goto done;
    }
done:
    /* tidycg */
    for (int i = 0; i < n; i++) sol[i] = sol[i] + dx[i];
    if (ierr == 0 && !conv) ierr = (f6 <= f4) ? 0 : -10;
    if (iters) *iters = nmv / 2;
    free(w);
    return ierr;
}

/* spk_solve_system (sparskit_admin.f90:115-330) for the 2D zeta system:
 * x holds the initial guess on entry (system_solve passes znv) and the
 * solution on return.  Returns ipar(1)-like status (0 ok). */
int oc_system_solve(const oc_mesh *m, const oc_params *p, const double *c2coo,
                    const double *rvec, double *x, int *iters)
{
    int n = m->nkn, nnz = m->nnz;
    int iwk = 2 * nnz;
    double *alu = (double *)malloc(sizeof(double) * (size_t)(iwk + n + 2));
    int *jlu = (int *)malloc(sizeof(int) * (size_t)(iwk + n + 2));
    int *ju = (int *)malloc(sizeof(int) * (size_t)(n + 2));
    int ierr;
    if (p->itpre == 1) ierr = oc_ilu0(n, c2coo, m->cols, m->ntg, alu, jlu, ju);
    else ierr = oc_ilut(n, c2coo, m->cols, m->ntg, p->lfil, p->droptol, alu, jlu, ju, iwk);
    if (ierr != 0) {
        free(alu); free(jlu); free(ju);
        return 1000 + (ierr < 0 ? -ierr : ierr);
    }
    ierr = oc_bcgstab(n, c2coo, m->cols, m->ntg, alu, jlu, ju, rvec, x, p->rtol, p->atol,
                      p->itermax, iters);
    free(alu); free(jlu); free(ju);
    return ierr;
}
