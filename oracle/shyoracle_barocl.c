/*
 * shyoracle_barocl.c -- CPU oracle for the baroclinic coupling of T/S to
 * the momentum equations, z-layers: the density anomaly from the UNESCO
 * 1980 equation of state (rhoset) and the baroclinic pressure-gradient term
 * (set_barocl_new_interface) that hydro adds to the explicit momentum terms.
 *
 * TEST INFRASTRUCTURE ONLY (see shyoracle.h).
 *
 * Pinned in tests/test_barocl_cpu.py: sigma_stp by the eight check values
 * the reference lists with its equation of state (eq_state.f90:126-136);
 * rhoset by the hydrostatic pressure at mid-layer (a column of uniform T/S
 * gives the density the equation of state gives there); the baroclinic term
 * by its closed form for a density anomaly that varies linearly in x and y
 * (gradient alpha): fxv(l) = g/rho0 * h_l * alpha * (depth of mid-layer l).
 */
#include <math.h>
#include <stddef.h>

#include "shyoracle.h"

/* sigma_stp, shyfem/eq_state.f90:154-235: UNESCO 1980 equation of state
 * (the published coefficients), evaluated in double precision:
 * sigma = rho - 1000 (kg/m3) for salinity s, temperature t (deg C) and
 * pressure p (bar). */
real oc_sigma_stp(real s, real t, real p) {
    const double t1 = t, t2 = t1 * t1, t3 = t2 * t1, t4 = t3 * t1, t5 = t4 * t1;
    const double s1 = s, s2 = s1 * s1, s32 = sqrt(s2 * s1);
    const double p1 = p, p2 = p1 * p1;
    /* pure water and the salinity terms at one standard atmosphere */
    const double rho0 = 999.842594 + 6.793952e-2 * t1 + -9.095290e-3 * t2 + 1.001685e-4 * t3 + -1.120083e-6 * t4 +
                        6.536332e-9 * t5 +
                        (8.24493e-1 + -4.0899e-3 * t1 + 7.6438e-5 * t2 + -8.2467e-7 * t3 + 5.3875e-9 * t4) * s1 +
                        (-5.72466e-3 + 1.0227e-4 * t1 + -1.6546e-6 * t2) * s32 + 4.8314e-4 * s2;
    /* secant bulk modulus */
    const double sbmk = 19652.21 + 148.4206 * t1 + -2.327105 * t2 + 1.360477e-2 * t3 + -5.155288e-5 * t4 +
                        (3.239908 + 1.43713e-3 * t1 + 1.16092e-4 * t2 + -5.77905e-7 * t3) * p1 +
                        (8.50935e-5 + -6.12293e-6 * t1 + 5.2787e-8 * t2) * p2 +
                        (54.6746 + -0.603459 * t1 + 1.09987e-2 * t2 + -6.1670e-5 * t3) * s1 +
                        (7.944e-2 + 1.6483e-2 * t1 + -5.3009e-4 * t2 + 1.91075e-4 * p1) * s32 +
                        (2.2838e-3 + -1.0981e-5 * t1 + -1.6078e-6 * t2) * p1 * s1 +
                        (-9.9348e-7 + 2.0816e-8 * t1 + 9.1697e-10 * t2) * p2 * s1;
    return (real)(rho0 / (1.0 - (double)p / sbmk) - 1000.0);
}

/* rhoset, shyfem/baroclinic.f90:654-720, z-layers: one pass down each node
 * column; the pressure at mid-layer (bar) is the barotropic rho0 g depth
 * plus the baroclinic part integrated from the anomaly of the previous pass,
 * and the new anomaly is sigma(S >= 0, T, p) - (rho0 - 1000). */
void oc_rhoset(const oc_mesh *m, const oc_params *p, const real *temp, const real *salt, real *rhov) {
    const real rho0 = p->rowass;
    const real sigma0 = rho0 - (real)1000.;
    for (int k = 0; k < m->nkn; k++) {
        real depth = 0, presbc = 0;
        const int lmax = m->ilhkv[k];
        for (int l = 1; l <= lmax; l++) {
            const size_t o = (size_t)(l - 1) + (size_t)m->nlvdi * k;
            const real hh = (real)0.5 * m->hldv[l - 1];
            depth = depth + hh;
            const real dpresc = rhov[o] * p->grav * hh;
            presbc = presbc + dpresc;
            const real presbt = rho0 * p->grav * depth;
            const real pres = (real)1.e-5 * (presbt + presbc);
            const real sl = salt[o] > (real)0. ? salt[o] : (real)0.;
            rhov[o] = oc_sigma_stp(sl, temp[o], pres) - sigma0;
            depth = depth + hh;
            presbc = presbc + dpresc;
        }
    }
}

/* set_barocl_new_interface, shyfem/explicit.f90:1226-1530, for z-layers
 * (no sigma or z-adaptive layers: interfaces are level, the sigma terms
 * vanish) and no distance/custom switches (ruseterm = 1): per element, the
 * anomaly gradient of each layer from the P1 gradients b, c, averaged onto
 * the interface above it (the first layer uses its own), integrated down
 * over the interface thicknesses, times g/rho0 and the layer thickness.
 * fxv/fyv receive the term (set_explicit starts them from zero). */
void oc_barocl(const oc_mesh *m, const oc_params *p, const real *rhov, real *fxv, real *fyv) {
    const double raux = (real)(p->grav / p->rowass);
    const real ruseterm = 1;
    for (int ie = 0; ie < m->nel; ie++) {
        double presbcx = 0., presbcy = 0., brup = 0., crup = 0., hhup = 0.;
        const int lmax = m->ilhv[ie];
        for (int l = 1; l <= lmax; l++) {
            const double hlayer = m->hldv[l - 1];
            const double hh = 0.5 * hlayer;
            const double hint = hh + hhup;
            double br = 0., cr = 0.;
            for (int ii = 0; ii < 3; ii++) {
                const int k = m->nen3v[3 * ie + ii];
                const real rhop = rhov[(size_t)(l - 1) + (size_t)m->nlvdi * k];
                br = br + m->ev[(3 + ii) + 19 * (size_t)ie] * rhop;
                cr = cr + m->ev[(6 + ii) + 19 * (size_t)ie] * rhop;
            }
            double brint, crint;
            if (l == 1) {
                brint = br;
                crint = cr;
            } else {
                brint = 0.5 * (br + brup);
                crint = 0.5 * (cr + crup);
            }
            brup = br;
            crup = cr;
            hhup = hh;
            const double dzdx = 0., dzdy = 0., psigma = 0.;
            presbcx = presbcx + hint * (brint - dzdx * psigma);
            presbcy = presbcy + hint * (crint - dzdy * psigma);
            const double xbcl = ruseterm * raux * hlayer * presbcx;
            const double ybcl = ruseterm * raux * hlayer * presbcy;
            const size_t o = (size_t)(l - 1) + (size_t)m->nlvdi * ie;
            fxv[o] = (real)(0.0 + xbcl);
            fyv[o] = (real)(0.0 + ybcl);
        }
    }
}
