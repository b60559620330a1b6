/*
 * shyoracle_hydro.c -- the semi-implicit hydro step of the CPU oracle
 * (TEST INFRASTRUCTURE; see shyoracle.h).
 *
 * Restates src/shyfem/hydrodynamic.f90 (hydro, hydro_transports,
 * hydro_intern, hydro_zeta, hydro_transports_final, hydro_vertical),
 * chezy.f90 bottom_friction, wetdry.f90 (setweg, setuvd, setzev),
 * topo_update.f90 setnod, bnd_routines.f90 adjust_mass_flux,
 * elems_dealing.f90 (setdepth for z-layers, masscont), transforms.f90
 * (check_volume, baro2l, copy_uvz, ttov) and simsys_spkaux.f90 (system_*).
 *
 * Types follow the reference: `real` where the reference declares `real`
 * (f32 in liboracle_f32, f64 in liboracle_f64), double where it declares
 * double precision.  Terms that are identically zero for the supported
 * configurations (radiation stress, turbines, vts, nudging of zeta,
 * rdistv/rcomputev = 1, implicit non-linear advection) are omitted.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "shyoracle.h"

#define EV(i, ie) m->ev[((i) - 1) + 19 * (ie)]
#define A3E(a, l, ie) (a)[((l) - 1) + nlvdi * (ie)]
#define A3K(a, l, k) (a)[((l) - 1) + nlvdi * (k)]
#define AW(a, l, k) (a)[(l) + (nlvdi + 1) * (k)]
#define ZE(a, ii, ie) (a)[((ii) - 1) + 3 * (ie)]
#define DD(a, j, ie) (a)[((j) - 1) + 2 * nlvdi * (ie)]
#define OPT(a, i) ((a) ? (a)[i] : (real)0)

/* ---------------- bottom_friction: chezy.f90:239-398 ---------------- */
void oc_bottom_friction(const oc_mesh *m, oc_state *s, const oc_params *p)
{
    const int nlvdi = m->nlvdi;
    const real drittl = (real)1 / (real)3;
    for (int ie = 0; ie < m->nel; ie++) {
        int lmax = m->ilhv[ie];
        real uso = A3E(s->utlov, lmax, ie), vso = A3E(s->vtlov, lmax, ie);
        real uv = (real)sqrt((double)(uso * uso + vso * vso));
        real hzg = A3E(s->hdeov, lmax, ie);
        if (hzg < p->hzoff) hzg = p->hzoff;
        real rfric = m->czv[ie];
        real rr = 0, rcd = 0;
        switch (p->ireib) {
        case 0: rr = 0; break;
        case 1: rr = rfric; break;
        case 2: rcd = p->grav / ((rfric * rfric) * (real)pow((double)hzg, (double)drittl)); rr = rcd * uv / (hzg * hzg); break;
        case 3: rcd = p->grav / (rfric * rfric); rr = rcd * uv / (hzg * hzg); break;
        case 4: rr = rfric / hzg; break;
        case 5: rcd = rfric; rr = rcd * uv / (hzg * hzg); break;
        case 6: {
            /* rfric is z0: cdf (chezy.f90:402-419), kappa = 0.4; the log is
             * taken in double and rounded (the device does the same) */
            real cds = (real)0.4 / (real)log((double)((rfric + (real)0.5 * hzg) / rfric));
            rcd = cds * cds;
            rr = rcd * uv / (hzg * hzg);
            break;
        }
        case 7:
            if (rfric >= 1) rcd = p->grav / ((rfric * rfric) * (real)pow((double)hzg, (double)drittl));
            else rcd = rfric;
            rr = rcd * uv / (hzg * hzg);
            break;
        case 10:
            if (uv > p->uvmin * hzg) { rcd = rfric; rr = rcd * uv / (hzg * hzg); }
            else rr = (rfric * p->uvmin) / hzg;
            break;
        default: rr = 0; break;
        }
        s->rfricv[ie] = rr;
    }
}

/* ---------------- hydro_intern: hydrodynamic.f90:868-1416 ---------------- */
static int hydro_intern(const oc_mesh *m, oc_state *s, const oc_params *p, int ie,
                        double *smat, double *rvec, double *solv)
{
    const int nlvdi = m->nlvdi;
    const real az = p->azpar, af = p->afpar, at = p->atpar;
    const real dt = (real)p->dt;
    const real grav = p->grav, rowass = p->rowass;
    const int bcolin = p->iclin != 0;
    int ilevel = m->ilhv[ie], jlevel = 1;
    int ngl = 2 * ilevel;
    int b2d = (ngl == 2);
    const double drittl = OC_DRITTL;
    double b[3], c[3];
    double bz = 0., cz = 0., bpres = 0., cpres = 0., zm = 0., zmm = 0., taux = 0., tauy = 0.;
    const double rcomp = 1.;
    (void)az;
    for (int ii = 0; ii < 3; ii++) {
        int kk = m->nen3v[3 * ie + ii];
        b[ii] = EV(ii + 4, ie);
        c[ii] = EV(ii + 7, ie);
        double zz = (double)(ZE(s->zeov, ii + 1, ie) - OPT(s->zeqv, kk));
        zm += zz;
        zmm += ZE(s->zeov, ii + 1, ie);
        bz += zz * b[ii];
        cz += zz * c[ii];
        bpres += (double)OPT(s->ppv, kk) * b[ii];
        cpres += (double)OPT(s->ppv, kk) * c[ii];
        taux += OPT(s->tauxnv, kk);
        tauy += OPT(s->tauynv, kk);
    }
    zm = zm * drittl;
    zmm = zmm * drittl;
    taux = rcomp * taux * drittl;
    tauy = rcomp * tauy * drittl;
    real gammat = (real)(m->fcorv[ie] * 1.0);
    real gamma = af * dt * gammat;

    real hact[nlvdi + 2], rhact[nlvdi + 2], alev[nlvdi + 1];
    for (int i = 0; i < 5 * ngl; i++) smat[i] = 0.;
    for (int i = 0; i < 6 * nlvdi; i++) { rvec[i] = 0.; solv[i] = 0.; }
    hact[0] = 0;
    for (int l = 1; l <= ilevel; l++) hact[l] = A3E(s->hdeov, l, ie);
    hact[ilevel + 1] = 0;
    if (bcolin) hact[1] = (real)(hact[1] - zmm);
    for (int l = 0; l <= ilevel + 1; l++) rhact[l] = (hact[l] <= 0) ? (real)0 : (real)1 / hact[l];
    int k1 = m->nen3v[3 * ie], k2 = m->nen3v[3 * ie + 1], k3 = m->nen3v[3 * ie + 2];
    for (int l = jlevel - 1; l <= ilevel; l++) {
        double vis = p->vismol;
        vis = vis + (double)((AW(s->visv, l, k1) + AW(s->visv, l, k2) + AW(s->visv, l, k3)) / (real)3);
        alev[l] = (real)vis;
    }
#define SM(j, i) smat[((j) + 2) + 5 * ((i) - 1)]
    double s2d[6];
#define S2(j, i) s2d[((j) + 1) + 3 * ((i) - 1)]
    for (int i = 0; i < 6; i++) s2d[i] = 0.;
    for (int l = 1; l <= ilevel; l++) {
        int bfirst = (l == jlevel), blast = (l == ilevel);
        int lp = (l + 1 < ilevel) ? l + 1 : ilevel;
        int lm = (l - 1 > jlevel) ? l - 1 : jlevel;
        real uui = A3E(s->utlov, l, ie), uuip = A3E(s->utlov, lp, ie), uuim = A3E(s->utlov, lm, ie);
        real vvi = A3E(s->vtlov, l, ie), vvip = A3E(s->vtlov, lp, ie), vvim = A3E(s->vtlov, lm, ie);
        real hhi = hact[l], hhip = hact[l + 1], hhim = hact[l - 1];
        real rhp = 0, rhm = 0;
        if (hhi > 0) {
            if (hhip > 0) rhp = (real)2.0 * alev[l] / (hhi + hhip);
            if (hhim > 0) rhm = (real)2.0 * alev[l - 1] / (hhi + hhim);
        }
        real aus = 1;
        real aux = dt * at;
        real aa = aux * rhact[l] * (rhm + rhp);
        real aat = aus * rhact[l] * (rhm + rhp);
        real bb = aux * rhact[l + 1] * rhp;
        real bbt = aus * rhact[l + 1] * rhp;
        real cc = aux * rhact[l - 1] * rhm;
        real cct = aus * rhact[l - 1] * rhm;
        double ppx = 0., ppy = 0.;
        if (bfirst) { ppx = ppx - taux; ppy = ppy - tauy; }
        if (blast) {
            real rfric = s->rfricv[ie];
            aa = aa + dt * rfric;
            aat = aat + rfric;
        }
        const double uuadv = 0., uvadv = 0., vuadv = 0., vvadv = 0.;
        real xexpl = s->fxv ? A3E(s->fxv, l, ie) : (real)0;
        real yexpl = s->fyv ? A3E(s->fyv, l, ie) : (real)0;
        real presx = (real)(rcomp * bpres), presy = (real)(rcomp * cpres);
        real gravx = (real)(rcomp * grav * hhi * bz);
        real gravy = (real)(rcomp * grav * hhi * cz);
        real ppx_aux = aat * uui - bbt * uuip - cct * uuim - gammat * vvi + gravx + (hhi / rowass) * presx + xexpl;
        real ppy_aux = aat * vvi - bbt * vvip - cct * vvim + gammat * uui + gravy + (hhi / rowass) * presy + yexpl;
        ppx = ppx + ppx_aux;
        ppy = ppy + ppy_aux;
        int jv = l + l, ju = jv - 1;
        if (b2d) {
            S2(0, ju) = (double)((real)1 + aa) + uuadv;
            S2(0, jv) = (double)((real)1 + aa) + vvadv;
            S2(-1, jv) = (double)gamma + vuadv;
            S2(1, ju) = (double)(-gamma) + uvadv;
        } else {
            SM(0, ju) = (double)((real)1 + aa) + uuadv;
            SM(0, jv) = (double)((real)1 + aa) + vvadv;
            SM(-1, jv) = (double)gamma + vuadv;
            SM(1, ju) = (double)(-gamma) + uvadv;
        }
        if (!blast) { SM(2, ju) = -bb; SM(2, jv) = -bb; }
        if (!bfirst) { SM(-2, ju) = -cc; SM(-2, jv) = -cc; }
        rvec[ju - 1] = ppx;
        rvec[jv - 1] = ppy;
        rvec[ngl + ju - 1] = hhi;
        rvec[ngl + jv - 1] = 0.;
        rvec[2 * ngl + ju - 1] = 0.;
        rvec[2 * ngl + jv - 1] = hhi;
    }
    if (b2d) {
        oc_tria_multi(ngl, 3, s2d, rvec, solv);
    } else {
        if (oc_penta_fact(ngl, smat) != 0) return 1;
        oc_penta_solve(ngl, smat, rvec, solv);
        oc_penta_solve(ngl, smat, rvec + ngl, solv + ngl);
        oc_penta_solve(ngl, smat, rvec + 2 * ngl, solv + 2 * ngl);
    }
    int afix = 1 - (s->iuvfix ? s->iuvfix[ie] : 0);
    if (afix != 0) {
        real dtafix = dt * (real)afix;
        for (int l = jlevel; l <= ilevel; l++) {
            A3E(s->utlnv, l, ie) = (real)(A3E(s->utlov, l, ie) - dtafix * solv[2 * l - 2]);
            A3E(s->vtlnv, l, ie) = (real)(A3E(s->vtlov, l, ie) - dtafix * solv[2 * l - 1]);
        }
    }
    for (int l = 1; l <= ngl; l++) {
        DD(s->ddxv, l, ie) = solv[ngl + l - 1];
        DD(s->ddyv, l, ie) = solv[2 * ngl + l - 1];
    }
#undef SM
#undef S2
    return 0;
}

/* hydro_transports: hydrodynamic.f90:758-864 (bottom friction + columns) */
void oc_hydro_transports(const oc_mesh *m, oc_state *s, const oc_params *p)
{
    int nlvdi = m->nlvdi;
    oc_bottom_friction(m, s, p);
    /* element columns are independent (the reference's OMP PARALLEL DO,
     * hydrodynamic.f90:837-850) */
#pragma omp parallel
    {
        double *smat = (double *)malloc(sizeof(double) * (size_t)(10 * nlvdi + 10));
        double *rvec = (double *)malloc(sizeof(double) * (size_t)(6 * nlvdi + 6));
        double *solv = (double *)malloc(sizeof(double) * (size_t)(6 * nlvdi + 6));
#pragma omp for schedule(static)
        for (int ie = 0; ie < m->nel; ie++) hydro_intern(m, s, p, ie, smat, rvec, solv);
        free(smat); free(rvec); free(solv);
    }
}

/* ---------------- hydro_zeta: hydrodynamic.f90:488-754 ---------------- */
void oc_hydro_zeta(const oc_mesh *m, oc_state *s, const oc_params *p, double *c2coo, double *rvec2d)
{
    const int nlvdi = m->nlvdi;
    const double drittl = OC_DRITTL;
    double az = p->azpar, am = p->ampar;
    real dt = (real)p->dt;
    double ddt = dt;
    const double amatr[9] = {4., 0., 0., 0., 4., 0., 0., 0., 4.};
    memset(c2coo, 0, sizeof(double) * (size_t)m->nnz);           /* system_init */
    memset(rvec2d, 0, sizeof(double) * (size_t)m->nkn);
    for (int ie = 0; ie < m->nel; ie++) {
        int kn[3];
        double b[3], c[3], z[3], zndg[3], zm = 0.;
        for (int i = 0; i < 3; i++) {
            int kk = m->nen3v[3 * ie + i];
            kn[i] = kk;
            b[i] = EV(i + 4, ie);
            c[i] = EV(i + 7, ie);
            z[i] = ZE(s->zeov, i + 1, ie);
            zndg[i] = OPT(s->andgzv, kk);
            zm += z[i];
        }
        zm = zm * drittl;
        int ilevel = m->ilhv[ie], jlevel = 1;
        double aj = EV(10, ie);
        int afix = 1 - (s->iuvfix ? s->iuvfix[ie] : 0);
        double delta = ddt * ddt * az * am * (double)p->grav * afix;
        double dbb = 0., dbc = 0., dcb = 0., dcc = 0.;
        for (int l = jlevel; l <= ilevel; l++) {
            int jv = l + l, ju = jv - 1;
            dbb += DD(s->ddxv, ju, ie);
            dbc += DD(s->ddyv, ju, ie);
            dcb += DD(s->ddxv, jv, ie);
            dcc += DD(s->ddyv, jv, ie);
        }
        double uold = 0., vold = 0., uhat = 0., vhat = 0.;
        for (int l = jlevel; l <= ilevel; l++) {
            uold += A3E(s->utlov, l, ie);
            vold += A3E(s->vtlov, l, ie);
            uhat += A3E(s->utlnv, l, ie);
            vhat += A3E(s->vtlnv, l, ie);
        }
        double ut = az * uhat + (1. - az) * uold;
        double vt = az * vhat + (1. - az) * vold;
        real hia[3][3], hik[3];
        for (int n = 0; n < 3; n++) {
            for (int mm = 0; mm < 3; mm++) {
                double abn = b[n] * (b[mm] * dbb + c[mm] * dbc);
                double acn = c[n] * (b[mm] * dcb + c[mm] * dcc);
                double h11 = delta * (abn + acn);
                hia[n][mm] = (real)(aj * (amatr[n + 3 * mm] + 12. * h11));
            }
            double acu = hia[n][0] * z[0] + hia[n][1] * z[1] + hia[n][2] * z[2];
            double andg = 4. * aj * ddt * zndg[n];
            hik[n] = (real)(acu + andg + 12. * aj * ddt * (ut * b[n] + vt * c[n]));
        }
        for (int i = 0; i < 3; i++) {
            if (s->rzv[kn[i]] != OC_FLAG) {
                double rw = s->rzv[kn[i]];
                int j1 = (i + 1) % 3, j2 = (i + 2) % 3;
                hik[j1] = (real)(hik[j1] - rw * hia[j1][i]);
                hik[j2] = (real)(hik[j2] - rw * hia[j2][i]);
                hia[i][j1] = 0; hia[i][j2] = 0;
                hia[j1][i] = 0; hia[j2][i] = 0;
                hik[i] = (real)(rw * hia[i][i]);
            }
        }
        if (s->iwegv[ie] != 0) {
            double hh999 = aj * 12.;
            for (int n = 0; n < 3; n++) {
                for (int mm = 0; mm < 3; mm++) hia[n][mm] = (real)(hh999 * (b[n] * b[mm] + c[n] * c[mm]));
                hik[n] = 0;
            }
            for (int n = 0; n < 3; n++) {
                int iv = s->inodv[kn[n]];
                if (iv != 0 && iv != -2) {
                    for (int mm = 0; mm < 3; mm++) hia[n][mm] = 0;
                    hik[n] = 0;
                }
            }
        }
        /* system_assemble: simsys_spkaux.f90:533-593 */
        for (int i = 0; i < 3; i++) {
            for (int j = 0; j < 3; j++) {
                int kk = m->ijp[i + 3 * j + 9 * ie];
                if (kk >= 0) c2coo[kk] = c2coo[kk] + hia[i][j];
            }
            rvec2d[kn[i]] = rvec2d[kn[i]] + hik[i];
        }
    }
    /* system_add_rhs(dt,nkn,rqv): simsys_spkaux.f90:687-704 */
    if (s->rqv)
        for (int k = 0; k < m->nkn; k++) rvec2d[k] = rvec2d[k] + (double)(dt * s->rqv[k]);
}

/* ---------------- wet/dry: wetdry.f90:128-347 ---------------- */
void oc_setweg(const oc_mesh *m, oc_state *s, const oc_params *p, int iweich, int *iw)
{
    real hzmin = p->hzmin, hzoff = p->hzoff, hzon = p->hzon, volmin = p->volmin;
    int iwh = 0, nel = m->nel;
    const int iwait = 5;
    int bnewtime = s->dtime != s->dtime_old;
    if (iweich == -1) {
        for (int ie = 0; ie < nel; ie++) { s->iwegv[ie] = 0; s->iwetv[ie] = 0; }
    } else if (iweich == 0) {
        for (int ie = 0; ie < nel; ie++) {
            real aomega = (real)EV(10, ie);
            real zmin = hzmin + volmin / ((real)4 * aomega);
            if (zmin > hzoff) zmin = hzmin + (real)0.5 * (hzmin + hzoff);
            int iweg = 0;
            for (int ii = 1; ii <= 3; ii++) {
                real hzg = ZE(s->zenv, ii, ie) + ZE(m->hm3v, ii, ie);
                if (hzg < hzoff) iweg++;
                if (hzg < zmin) ZE(s->zenv, ii, ie) = zmin - ZE(m->hm3v, ii, ie);
            }
            s->iwegv[ie] = iweg;
            if (iweg == 0) s->iwetv[ie] = 1;
            if (iweg > 0) iwh++;
        }
    } else if (iweich == 1 || iweich == 2) {
        real hzlim = (iweich == 1) ? hzmin : hzoff;
        for (int ie = 0; ie < nel; ie++) {
            if (s->iwegv[ie] != 0) continue;
            int iweg = 0;
            for (int ii = 1; ii <= 3; ii++) {
                real hzg = s->znv[m->nen3v[3 * ie + ii - 1]] + ZE(m->hm3v, ii, ie);
                if (hzg < hzlim) iweg++;
            }
            if (iweg > 0) {
                if (s->iwegv[ie] == 0) { iwh++; s->iwetv[ie] = 0; }
                s->iwegv[ie] = iweg;
            }
        }
    } else if (iweich == 3) {
        for (int ie = 0; ie < nel; ie++) {
            if (s->iwegv[ie] <= 0) continue;
            int iweg = 0;
            real hztot = 0;
            for (int ii = 1; ii <= 3; ii++) {
                real hzg = s->znv[m->nen3v[3 * ie + ii - 1]] + ZE(m->hm3v, ii, ie);
                hztot = hztot + hzg;
                if (hzg < hzon) iweg++;
            }
            hztot = hztot / (real)3;
            int binclude = (hztot >= hzon) && (-s->iwetv[ie] > iwait);
            if (binclude) iweg = 0;
            if (iweg == 0) {
                if (s->iwegv[ie] > 0) { iwh++; s->iwetv[ie] = 0; }
                s->iwegv[ie] = 0;
            }
        }
    }
    for (int ie = 0; ie < nel; ie++) {
        int iweg = s->iwegv[ie], iwet = s->iwetv[ie];
        if (bnewtime || iwet == 0) {
            if (iweg == 0) s->iwetv[ie]++;
            else s->iwetv[ie]--;
        }
    }
    s->dtime_old = s->dtime;
    *iw = iwh;
}

/* set_dry: wetdry.f90:94-104 */
void oc_set_dry(const oc_mesh *m, oc_state *s, const oc_params *p)
{
    int iwa;
    oc_setweg(m, s, p, 2, &iwa);
    oc_setweg(m, s, p, 3, &iwa);
}

/* setnod: topo_update.f90:166-270.  Returns 0 or 1 (open-boundary node dry). */
int oc_setnod(const oc_mesh *m, oc_state *s)
{
    const double winmax = 359.99;
    double *winkv = (double *)calloc((size_t)m->nkn, sizeof(double));
    for (int ie = 0; ie < m->nel; ie++) {
        if (s->iwegv[ie] != 0) continue;
        for (int ii = 1; ii <= 3; ii++) {
            int k = m->nen3v[3 * ie + ii - 1];
            winkv[k] = winkv[k] + EV(10 + ii, ie);
        }
    }
    int ret = 0;
    for (int k = 0; k < m->nkn; k++) {
        if (winkv[k] > winmax) s->inodv[k] = 0;
        else if (winkv[k] == 0.) s->inodv[k] = -2;
        else s->inodv[k] = -1;
    }
    /* boundaries are visited in boundary order; a node belongs to one */
    for (int k = 0; k < m->nkn; k++) {
        int ibc = m->ibnd ? m->ibnd[k] : 0;
        if (ibc <= 0) continue;
        int ibtyp = m->ibtyp[ibc];
        if (ibtyp >= 3) {
            if (s->inodv[k] == -1) s->inodv[k] = ibc;
        } else if (ibtyp > 0) {
            if (s->inodv[k] != -1) ret = 1;
            s->inodv[k] = ibc;
        }
    }
    free(winkv);
    return ret;
}

/* adjust_mass_flux: bnd_routines.f90:1185-1212 */
static void adjust_mass_flux(const oc_mesh *m, oc_state *s)
{
    const int nlvdi = m->nlvdi;
    for (int k = 0; k < m->nkn; k++) {
        if (s->inodv[k] != -2) continue;
        if (s->mfluxv)
            for (int l = 1; l <= m->nlv; l++) A3K(s->mfluxv, l, k) = 0;
        if (s->rqv) s->rqv[k] = 0;
    }
}

/* ---------------- hydro_transports_final: hydrodynamic.f90:1468-1578 ---------------- */
void oc_hydro_transports_final(const oc_mesh *m, oc_state *s, const oc_params *p)
{
    const int nlvdi = m->nlvdi;
    real dt = (real)p->dt;
    double am = p->ampar;
    double beta = dt * p->grav * am;
    for (int ie = 0; ie < m->nel; ie++) {
        int ilevel = m->ilhv[ie], jlevel = 1;
        int afix = 1 - (s->iuvfix ? s->iuvfix[ie] : 0);
        double rfix = afix;
        double bz = 0., cz = 0.;
        for (int ii = 1; ii <= 3; ii++) {
            int kk = m->nen3v[3 * ie + ii - 1];
            double dz = (double)(s->znv[kk] - ZE(s->zeov, ii, ie));
            bz = bz + dz * EV(ii + 3, ie);
            cz = cz + dz * EV(ii + 6, ie);
        }
        for (int l = jlevel; l <= ilevel; l++) {
            int jv = l + l, ju = jv - 1;
            double du = beta * (DD(s->ddxv, ju, ie) * bz + DD(s->ddyv, ju, ie) * cz);
            double dv = beta * (DD(s->ddxv, jv, ie) * bz + DD(s->ddyv, jv, ie) * cz);
            A3E(s->utlnv, l, ie) = (real)(A3E(s->utlnv, l, ie) - du * rfix);
            A3E(s->vtlnv, l, ie) = (real)(A3E(s->vtlnv, l, ie) - dv * rfix);
        }
        double um = 0., vm = 0.;
        for (int l = jlevel; l <= ilevel; l++) {
            um = um + A3E(s->utlnv, l, ie);
            vm = vm + A3E(s->vtlnv, l, ie);
        }
        s->unv[ie] = (real)um;
        s->vnv[ie] = (real)vm;
    }
}

/* setzev: wetdry.f90:594-615 */
void oc_setzev(const oc_mesh *m, oc_state *s)
{
    for (int ie = 0; ie < m->nel; ie++)
        if (s->iwegv[ie] == 0)
            for (int ii = 1; ii <= 3; ii++) ZE(s->zenv, ii, ie) = s->znv[m->nen3v[3 * ie + ii - 1]];
}

/* setuvd: wetdry.f90:350-590 -- water level and barotropic velocity in
 * elements out of the system.  Returns 0, or 1/2 for the reference's
 * error stops (z computation / u,v computation). */
int oc_setuvd(const oc_mesh *m, oc_state *s, const oc_params *p)
{
    const real dt = (real)p->dt;
    const real hzmin = p->hzmin;
    const real az = p->azpar, azt = (real)1 - az;
    const real eps = (real)1.e-5;
    for (int ie = 0; ie < m->nel; ie++) {
        if (!(s->iwegv[ie] > 0)) continue;
        real uo = s->uov[ie], vo = s->vov[ie];
        real axdt = (real)3 * dt;
        real adt = (real)1 / axdt;
        double zm = 0., zmed, d1, d2, det;
        real z[3], b[3], c[3], zn[3], hm[3], u, v, zz;
        for (int ii = 0; ii < 3; ii++) {
            zn[ii] = ZE(s->zenv, ii + 1, ie);
            zm = zm + zn[ii];
            b[ii] = (real)EV(4 + ii, ie);
            c[ii] = (real)EV(7 + ii, ie);
            hm[ii] = ZE(m->hm3v, ii + 1, ie);
        }
        zmed = zm / 3.;
        int itot = 0, isum = 0;
        for (int ii = 0; ii < 3; ii++) {
            if (zmed + hm[ii] < hzmin) {
                z[ii] = hzmin - hm[ii];
                itot++;
                isum += ii + 1;
            } else {
                z[ii] = (real)zmed;
            }
            zm = zm - z[ii];
        }
        if (itot == 0) {
        } else if (itot == 1) {
            int i1 = (isum % 3) + 1, i2 = (i1 % 3) + 1;
            if (z[i1 - 1] + zm * 0.5 + hm[i1 - 1] < hzmin) {
                z[i1 - 1] = hzmin - hm[i1 - 1];
                z[i2 - 1] = (real)(3. * zmed - z[i1 - 1] - z[isum - 1]);
            } else if (z[i2 - 1] + zm * 0.5 + hm[i2 - 1] < hzmin) {
                z[i2 - 1] = hzmin - hm[i2 - 1];
                z[i1 - 1] = (real)(3. * zmed - z[i2 - 1] - z[isum - 1]);
            } else {
                z[i1 - 1] = (real)(z[i1 - 1] + 0.5 * zm);
                z[i2 - 1] = (real)(z[i2 - 1] + 0.5 * zm);
            }
        } else if (itot == 2) {
            int i1 = 6 - isum;
            z[i1 - 1] = (real)(z[i1 - 1] + zm);
        } else {
            return 1;
        }
        zm = 0.;
        for (int ii = 0; ii < 3; ii++) {
            zm = zm + z[ii];
            if (z[ii] + hm[ii] < hzmin - eps) return 1;
        }
        if (fabs(zm - 3. * zmed) > eps) return 1;
        d1 = (z[0] - zn[0]) * adt - azt * (b[0] * uo + c[0] * vo);
        d2 = (z[1] - zn[1]) * adt - azt * (b[1] * uo + c[1] * vo);
        det = (real)1 / (b[0] * c[1] - b[1] * c[0]);
        u = (real)(det * (c[1] * d1 - c[0] * d2));
        v = (real)(det * (-b[1] * d1 + b[0] * d2));
        itot = 0;
        for (int ii = 0; ii < 3; ii++) {
            zz = axdt * (b[ii] * (azt * uo + az * u) + c[ii] * (azt * vo + az * v)) + zn[ii];
            if (zz + hm[ii] < hzmin - eps) itot++;
        }
        if (itot > 0) {
            if (az == 0) return 2;
            d1 = (z[0] - zn[0]) * adt - azt * (b[0] * uo + c[0] * vo);
            d2 = (z[1] - zn[1]) * adt - azt * (b[1] * uo + c[1] * vo);
            det = (real)1 / (az * (b[0] * c[1] - b[1] * c[0]));
            u = (real)(det * (c[1] * d1 - c[0] * d2));
            v = (real)(det * (-b[1] * d1 + b[0] * d2));
        }
        itot = 0;
        zm = 0.;
        for (int ii = 0; ii < 3; ii++) {
            zz = axdt * (b[ii] * (azt * uo + az * u) + c[ii] * (azt * vo + az * v)) + zn[ii];
            if (zz + hm[ii] < hzmin - eps) itot++;
            ZE(s->zenv, ii + 1, ie) = zz;
            zm = zm + zz;
        }
        if (fabs(zm - 3. * zmed) > eps) return 1;
        if (itot > 0) return 2;
        s->unv[ie] = u;
        s->vnv[ie] = v;
    }
    return 0;
}

/* baro2l: transforms.f90:437-481 */
void oc_baro2l(const oc_mesh *m, oc_state *s)
{
    const int nlvdi = m->nlvdi;
    for (int ie = 0; ie < m->nel; ie++) {
        if (!(s->iwegv[ie] > 0)) continue;
        int ilevel = m->ilhv[ie], jlevel = 1;
        real htot = 0;
        for (int l = jlevel; l <= ilevel; l++) htot = htot + A3E(s->hdenv, l, ie);
        for (int l = jlevel; l <= ilevel; l++) {
            real weight = A3E(s->hdenv, l, ie) / htot;
            A3E(s->utlnv, l, ie) = s->unv[ie] * weight;
            A3E(s->vtlnv, l, ie) = s->vnv[ie] * weight;
        }
    }
}

/* setdepth for z-layers (nsigma = 0, no z-adaptive layers):
 * elems_dealing.f90:930-1156.  Returns 0, or 1 on non-positive thickness. */
int oc_setdepth(const oc_mesh *m, real *hdkn, real *hden, const real *zenv)
{
    const int nlvdi = m->nlvdi;
    const real hmin = 0;
    memset(hdkn, 0, sizeof(real) * (size_t)nlvdi * (size_t)m->nkn);
    memset(hden, 0, sizeof(real) * (size_t)nlvdi * (size_t)m->nel);
    for (int ie = 0; ie < m->nel; ie++) {
        real areael = (real)(12 * EV(10, ie));
        real areafv = areael / (real)3;
        int lmax = m->ilhv[ie], lmin = 1;
        real zmed = 0;
        for (int ii = 1; ii <= 3; ii++) {
            int k = m->nen3v[3 * ie + ii - 1];
            int lmink = 1;
            real htot = ZE(m->hm3v, ii, ie);
            real z = ZE(zenv, ii, ie);
            zmed = zmed + z;
            if (lmax > 0) {
                if (lmax == lmin) {
                    real hl = htot + z;
                    A3K(hdkn, lmin, k) = A3K(hdkn, lmin, k) + areafv * hl;
                } else {
                    int levmin = lmink;
                    for (int l = levmin; l <= lmax - 1; l++)
                        A3K(hdkn, l, k) = A3K(hdkn, l, k) + areafv * m->hldv[l - 1];
                    A3K(hdkn, lmink, k) = A3K(hdkn, lmink, k) + areafv * z;
                    real hlast = htot - m->hlv[lmax - 2];
                    if (hlast < 0) return 1;
                    A3K(hdkn, lmax, k) = A3K(hdkn, lmax, k) + areafv * hlast;
                }
            }
        }
        zmed = zmed / (real)3;
        real htot = m->hev[ie];
        if (lmax == lmin) {
            A3E(hden, lmin, ie) = htot + zmed;
        } else {
            for (int l = lmin; l <= lmax - 1; l++) A3E(hden, l, ie) = m->hldv[l - 1];
            A3E(hden, lmin, ie) = A3E(hden, lmin, ie) + zmed;
            real hlast = htot - m->hlv[lmax - 2];
            if (hlast < 0) return 1;
            A3E(hden, lmax, ie) = hlast;
        }
        for (int l = lmin; l <= lmax; l++)
            if (A3E(hden, l, ie) <= hmin) return 1;
    }
    for (int k = 0; k < m->nkn; k++) {
        int lmax = m->ilhkv[k];
        for (int l = 1; l <= lmax; l++) {
            real areafv = A3K(m->areakv, l, k);
            if (areafv > 0) A3K(hdkn, l, k) = A3K(hdkn, l, k) / areafv;
            else break;
        }
        for (int l = 1; l <= lmax; l++)
            if (A3K(hdkn, l, k) <= hmin) return 1;
    }
    return 0;
}

/* check_volume: transforms.f90:368-433 (first layer only) */
int oc_check_volume(const oc_mesh *m, const oc_state *s)
{
    const int nlvdi = m->nlvdi;
    for (int k = 0; k < m->nkn; k++)
        if (A3K(s->hdknv, 1, k) <= 0) return 1;
    for (int ie = 0; ie < m->nel; ie++)
        if (A3E(s->hdenv, 1, ie) <= 0) return 1;
    return 0;
}

/* is_zeta_bound: bndo.f90:591-612 (open boundary of type 1) */
static int is_zeta_bound(const oc_mesh *m, int k)
{
    int ibc = m->ibnd ? m->ibnd[k] : 0;
    return ibc > 0 && m->ibtyp[ibc] == 1;
}

/* hydro_vertical: hydrodynamic.f90:1582-1781 */
void oc_hydro_vertical(const oc_mesh *m, oc_state *s, const oc_params *p, real *dzeta)
{
    const int nlvdi = m->nlvdi, nkn = m->nkn;
    for (int k = 0; k < nkn; k++) dzeta[k] = 0;
    memset(s->wlnv, 0, sizeof(real) * (size_t)(nlvdi + 1) * (size_t)nkn);
    if (m->nlv == 1) return;
    real az = p->azpar, azt = (real)1 - az;
    real dt = (real)p->dt;
    real *vf = (real *)calloc((size_t)nlvdi * nkn, sizeof(real));
    real *va = (real *)calloc((size_t)nlvdi * nkn, sizeof(real));
    for (int ie = 0; ie < m->nel; ie++) {
        real aj = (real)(4. * EV(10, ie));
        int ilevel = m->ilhv[ie];
        for (int l = 1; l <= ilevel; l++) {
            for (int ii = 1; ii <= 3; ii++) {
                int kk = m->nen3v[3 * ie + ii - 1];
                real b = (real)EV(ii + 3, ie), c = (real)EV(ii + 6, ie);
                real ffn = A3E(s->utlnv, l, ie) * b + A3E(s->vtlnv, l, ie) * c;
                real ffo = A3E(s->utlov, l, ie) * b + A3E(s->vtlov, l, ie) * c;
                real ff = ffn * az + ffo * azt;
                A3K(vf, l, kk) = A3K(vf, l, kk) + (real)3 * aj * ff;
                A3K(va, l, kk) = A3K(va, l, kk) + aj;
            }
        }
    }
    for (int k = 0; k < nkn; k++) {
        int lmax = m->ilhkv[k], lmin = 1;
        AW(s->wlnv, lmax, k) = 0;
        for (int l = lmax; l >= lmin; l--) {
            real voln = A3K(s->hdknv, l, k) * A3K(m->areakv, l, k);
            real volo = A3K(s->hdkov, l, k) * A3K(m->areakv, l, k);
            real dvdt = (voln - volo) / dt;
            real q = s->mfluxv ? A3K(s->mfluxv, l, k) : (real)0;
            real wdiv = A3K(vf, l, k) + q;
            AW(s->wlnv, l - 1, k) = AW(s->wlnv, l, k) + wdiv - dvdt;
        }
        real dz = dt * AW(s->wlnv, lmin - 1, k) / A3K(va, lmin, k);
        AW(s->wlnv, lmin - 1, k) = 0;
        dzeta[k] = dz;
    }
    for (int k = 0; k < nkn; k++) {
        int lmax = m->ilhkv[k], lmin = 1;
        for (int l = lmin + 1; l <= lmax; l++) {
            real atop = A3K(va, l, k);
            if (atop > 0) AW(s->wlnv, l - 1, k) = AW(s->wlnv, l - 1, k) / atop;
        }
    }
    for (int k = 0; k < nkn; k++) {
        if (is_zeta_bound(m, k)) {
            for (int l = 0; l <= nlvdi; l++) AW(s->wlnv, l, k) = 0;
            dzeta[k] = 0;
        }
    }
    free(vf);
    free(va);
}

/* masscont: elems_dealing.f90:1191-1220 (FP64 sum of nodal volumes) */
double oc_masscont(const oc_mesh *m, const oc_state *s, int mode)
{
    const int nlvdi = m->nlvdi;
    const real *hd = mode > 0 ? s->hdknv : s->hdkov;
    double total = 0.;
    for (int k = 0; k < m->nkn; k++) {
        int nlev = m->ilhkv[k];
        for (int l = 1; l <= nlev; l++) {
            real vol = A3K(hd, l, k) * A3K(m->areakv, l, k);
            total = total + vol;
        }
    }
    return total;
}

/* ttov: transforms.f90:98-122 */
void oc_ttov(const oc_mesh *m, oc_state *s)
{
    size_t n = (size_t)m->nlvdi * m->nel;
    for (size_t i = 0; i < n; i++) {
        s->ulnv[i] = 0;
        s->vlnv[i] = 0;
        if (s->hdenv[i] > 0) {
            s->ulnv[i] = s->utlnv[i] / s->hdenv[i];
            s->vlnv[i] = s->vtlnv[i] / s->hdenv[i];
        }
    }
}

/* uvint: transforms.f90:336-362 */
static void uvint(const oc_mesh *m, oc_state *s)
{
    const int nlvdi = m->nlvdi;
    for (int ie = 0; ie < m->nel; ie++) {
        double u = 0., v = 0.;
        for (int l = 1; l <= m->ilhv[ie]; l++) {
            u = u + A3E(s->utlnv, l, ie);
            v = v + A3E(s->vtlnv, l, ie);
        }
        s->unv[ie] = (real)u;
        s->vnv[ie] = (real)v;
    }
}

/* copy_uvz (transforms.f90:523-552) and copy_depth (elems_dealing.f90:804-850) */
void oc_copy_uvz(const oc_mesh *m, oc_state *s)
{
    size_t ne3 = (size_t)m->nlvdi * m->nel, nw = (size_t)(m->nlvdi + 1) * m->nkn;
    memcpy(s->zov, s->znv, sizeof(real) * (size_t)m->nkn);
    memcpy(s->zeov, s->zenv, sizeof(real) * 3 * (size_t)m->nel);
    memcpy(s->utlov, s->utlnv, sizeof(real) * ne3);
    memcpy(s->vtlov, s->vtlnv, sizeof(real) * ne3);
    memcpy(s->wlov, s->wlnv, sizeof(real) * nw);
    memcpy(s->uov, s->unv, sizeof(real) * (size_t)m->nel);
    memcpy(s->vov, s->vnv, sizeof(real) * (size_t)m->nel);
}

void oc_copy_depth(const oc_mesh *m, oc_state *s)
{
    memcpy(s->hdkov, s->hdknv, sizeof(real) * (size_t)m->nlvdi * m->nkn);
    memcpy(s->hdeov, s->hdenv, sizeof(real) * (size_t)m->nlvdi * m->nel);
}

/* hydro: hydrodynamic.f90:284-484.  Returns 0 or an error code:
 * 1 setnod, 2 solver, 3 setuvd, 4 setdepth, 5 check_volume, 6 wet/dry loop. */
int oc_hydro(const oc_mesh *m, oc_state *s, const oc_params *p, oc_stats *st)
{
    int nkn = m->nkn;
    double *c2coo = (double *)malloc(sizeof(double) * (size_t)m->nnz);
    double *rvec = (double *)malloc(sizeof(double) * (size_t)nkn);
    double *x = (double *)malloc(sizeof(double) * (size_t)nkn);
    real *dzeta = (real *)malloc(sizeof(real) * (size_t)nkn);
    int ret = 0, iloop = 0, iw;
    st->total_solver_iters = 0;
    for (;;) {
        iloop++;
        if (iloop > 100) { ret = 6; goto out; }
        oc_hydro_transports(m, s, p);
        if (oc_setnod(m, s)) { ret = 1; goto out; }
        adjust_mass_flux(m, s);
        oc_hydro_zeta(m, s, p, c2coo, rvec);
        for (int k = 0; k < nkn; k++) x[k] = s->znv[k];
        int iters = 0;
        int ierr = oc_system_solve(m, p, c2coo, rvec, x, &iters);
        st->solver_iters = iters;
        st->solver_ierr = ierr;
        st->total_solver_iters += iters;
        if (ierr != 0) { ret = 2; goto out; }
        for (int k = 0; k < nkn; k++) s->znv[k] = (real)x[k];          /* system_get */
        oc_setweg(m, s, p, 1, &iw);
        if (iw == 0) break;
    }
    st->nloop = iloop;
    oc_hydro_transports_final(m, s, p);
    oc_setzev(m, s);
    if (oc_setuvd(m, s, p)) { ret = 3; goto out; }
    oc_baro2l(m, s);
    if (oc_setdepth(m, s->hdknv, s->hdenv, s->zenv)) { ret = 4; goto out; }
    if (oc_check_volume(m, s)) { ret = 5; goto out; }
    oc_hydro_vertical(m, s, p, dzeta);
    oc_ttov(m, s);                                                       /* compute_velocities */
    uvint(m, s);
out:
    st->err = ret;
    free(c2coo); free(rvec); free(x); free(dzeta);
    return ret;
}
