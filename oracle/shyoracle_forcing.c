/*
 * shyoracle_forcing.c -- CPU oracle for the forcing step that precedes
 * hydro: time interpolation of boundary-level and wind records
 * (sp111(2), meteo_forcing_fem) and the wind-to-stress conversion.
 *
 * TEST INFRASTRUCTURE ONLY (see shyoracle.h).
 *
 * The record handling is restated as the reference runs it: a ring of nintp
 * records filled by iff_populate_records, advanced by
 * iff_read_and_interpolate while the check time tcomp is before the
 * requested time, and interpolated by rd_intp_neville with the points taken
 * in ring-slot order.  Pinned in tests/test_forcing_cpu.py by closed forms
 * (Lagrange interpolation reproduces polynomials of its degree, agrees with
 * an independent barycentric evaluation, returns the data at the nodes),
 * the window invariants (two records at or before t and two after for the
 * cubic window, one and one for the linear), and the stress formula
 * tau = (roluft/rowass) cd |w| w with its drag laws and speed limit.
 * The bits of the interpolated value depend on the ring-slot order of the
 * Neville points, which only the restated order fixes: parity unpinned at
 * the last-bit level (the device is compared with this restatement).
 *
 * set_mass_flux (discharges into the layers of their nodes) is pinned in
 * tests/test_mass_flux_cpu.py by what the routine must conserve and by
 * hand-computed cases: a column's parts sum to its discharge, parts follow
 * the layer volumes, levmin/levmax select the layers (negative: counted
 * from the bottom), two sources at one node add, rain goes to the top
 * layer, rqv is the column sum, and a bad layer range or a zero volume is
 * the error stop.
 */
#include <math.h>
#include <stddef.h>

#include "shyoracle.h"

/* rd_intp_neville, utils/generic/intp.f90:317-345: Neville's scheme in
 * double precision on real data */
double oc_intp_neville(int nintp, const double *xa, const float *ya, double x) {
    double p[8];
    const int n = nintp - 1;
    for (int i = 0; i < nintp; i++) p[i] = (double)ya[i];
    for (int k = 1; k <= n; k++)
        for (int i = n; i >= k; i--) {
            const double xl = xa[i - k], xh = xa[i];
            p[i] = ((x - xl) * p[i] - (x - xh) * p[i - 1]) / (xh - xl);
        }
    return p[n];
}

/* iff_populate_records, utils/shyutilmpi/intp_fem_file.f90:1047-1144: the
 * first record of the initial ring for start time dtime0.  Records are
 * skipped while the next record's time (+ the record spacing for the cubic
 * window) is before dtime0.  Returns the record index, or -1 when the
 * records do not cover dtime0 or are too few to fill the ring; nrec == 1
 * is a constant field (returns 0). */
int oc_iff_first_record(int nrec, const double *time, int nintp, double dtime0) {
    if (nrec < 1) return -1;
    if (nrec == 1) return 0;
    int j = 0;
    const double first = time[0];
    const double ddt = nintp == 4 ? time[1] - time[0] : 0.0;
    for (;;) {
        if (j + 1 >= nrec) return -1;                 /* goto 97 */
        if (time[j + 1] + ddt >= dtime0) break;
        j++;
    }
    if (dtime0 < first) return -1;                    /* goto 91 */
    if (dtime0 > time[j + 1] + ddt) return -1;
    if (j + nintp - 1 > nrec - 1) return -1;          /* goto 96 */
    return j;
}

/* the ring of one record series as the reference holds it */
typedef struct {
    int nintp, jstart;
    int ilast;         /* 1-based slot of the newest record */
    int jnext;         /* next record to read */
    int slot_rec[8];   /* record held in slot s (0-based) */
    int eof;
} oc_ring;

static double tcomp(const oc_ring *r, const double *time) {
    /* tcomp, intp_fem_file.f90:2824-2849 */
    const int n = r->nintp;
    int n1 = 1 + (r->ilast + n / 2) % n, n2 = n1;
    if (n % 2 != 0 && n > 1) n2 = n2 % n + 1;
    return 0.5 * (time[r->slot_rec[n1 - 1]] + time[r->slot_rec[n2 - 1]]);
}

/* the ring after iff_populate_records (slots 1..nintp = records jstart..) */
static void ring_init(oc_ring *r, int nintp, int jstart) {
    r->nintp = nintp;
    r->jstart = jstart;
    for (int s = 0; s < nintp; s++) r->slot_rec[s] = jstart + s;
    r->ilast = nintp;
    r->jnext = jstart + nintp;
    r->eof = 0;
}

/* iff_read_and_interpolate, intp_fem_file.f90:2452-2510: read records
 * into the ring while the check time is before t */
static void ring_advance(oc_ring *r, int nrec, const double *time, double t) {
    double tc = tcomp(r, time);
    while (tc < t) {
        if (r->jnext >= nrec) { r->eof = 1; break; }
        r->ilast = r->ilast % r->nintp + 1;
        r->slot_rec[r->ilast - 1] = r->jnext++;
        tc = tcomp(r, time);
    }
}

/* Values of one component of a record series at the times t[0..nt-1]
 * (non-decreasing), as a model run calls iff_read_and_interpolate +
 * iff_time_interpolate once per step (intp_fem_file.f90:2288-2416 and
 * 2514-2700 for one variable, one level).  vals is record-major:
 * vals[(rec*ncomp + comp)*nval + i]; out[k*nval + i].  A record value equal
 * to -999 in the window gives vdefault.  Also returns, per time, the newest
 * record of the ring (newest[k]) so the window can be checked.  Returns 0,
 * or 1 when a time lies outside the records of the ring (the reference's
 * "time out of range" stop). */
int oc_iff_series(int nrec, const double *time, const float *vals, int ncomp, int comp, int nval, int nintp,
                  double dtime0, float vdefault, int nt, const double *t, float *out, int *newest) {
    const float flag = -999.f;
    if (nrec == 1) {
        for (int k = 0; k < nt; k++) {
            for (int i = 0; i < nval; i++) {
                const float v = vals[(size_t)comp * nval + i];
                out[(size_t)k * nval + i] = v == flag ? vdefault : v;
            }
            if (newest) newest[k] = 0;
        }
        return 0;
    }
    const int j0 = oc_iff_first_record(nrec, time, nintp, dtime0);
    if (j0 < 0) return 1;
    oc_ring r;
    ring_init(&r, nintp, j0);
    for (int k = 0; k < nt; k++) {
        ring_advance(&r, nrec, time, t[k]);
        const int ifirst = r.ilast % nintp + 1;
        if (time[r.slot_rec[r.ilast - 1]] < t[k] || time[r.slot_rec[ifirst - 1]] > t[k]) return 1;
        double xa[8];
        for (int s = 0; s < nintp; s++) xa[s] = time[r.slot_rec[s]];
        for (int i = 0; i < nval; i++) {
            float ya[8];
            int bflag = 0;
            for (int s = 0; s < nintp; s++) {
                ya[s] = vals[((size_t)r.slot_rec[s] * ncomp + comp) * nval + i];
                if (ya[s] == flag) bflag = 1;
            }
            out[(size_t)k * nval + i] = bflag ? vdefault : (float)oc_intp_neville(nintp, xa, ya, t[k]);
        }
        if (newest) newest[k] = r.slot_rec[r.ilast - 1];
    }
    return 0;
}

/* get_drag, shyfem/meteo_forcing.f90:1580-1608 (itdrag 0, 1, 2) */
static float get_drag(int itdrag, float wxy, float cd) {
    if (itdrag == 1) return 0.001f * (0.63f + 0.066f * wxy);
    if (itdrag == 2) return wxy > 11.f ? 0.001f * (0.49f + 0.066f * wxy) : 0.001f * 1.2f;
    return cd;
}

/* meteo_convert_wind_data, shyfem/meteo_forcing.f90:864-1000, no ice
 * (fice = 1): iwtype 1 wind velocity -> stress per unit density,
 * iwtype 2 stress (N/m2) -> normalised stress and the equivalent wind.
 * wx/wy are updated in place (limited / derived wind), ws receives the
 * speed.  Returns 0, or 1 when the largest speed exceeds wsmax (the
 * reference's stop). */
int oc_convert_wind(int n, float *wx, float *wy, int iwtype, int itdrag, float dragco, float roluft, float rowass,
                    float wslim, float wsmax, float *tx, float *ty, float *ws) {
    const float fice = 1.f;
    const float wfact = iwtype == 2 ? 1.f / rowass : roluft / rowass;
    float wxymax = 0.f;
    if (iwtype == 2) {
        float cd = dragco;
        if (cd <= 0.f) cd = 2.5e-3f;
        for (int k = 0; k < n; k++) {
            tx[k] = fice * wfact * wx[k];
            ty[k] = fice * wfact * wy[k];
            const float txy = sqrtf(tx[k] * tx[k] + ty[k] * ty[k]);
            const float wspeed = sqrtf(txy / (cd * wfact * roluft));
            if (wspeed > wxymax) wxymax = wspeed;
            wx[k] = tx[k] / (cd * wspeed);
            wy[k] = ty[k] / (cd * wspeed);
            ws[k] = wspeed;
        }
    } else {
        for (int k = 0; k < n; k++) {
            const double dwxy = (double)wx[k] * (double)wx[k] + (double)wy[k] * (double)wy[k];
            ws[k] = (float)sqrt(dwxy);
        }
        for (int k = 0; k < n; k++) {
            const float wspeed = ws[k];
            if (wspeed > wxymax) wxymax = wspeed;
            const float cd = get_drag(itdrag, wspeed, dragco);
            const float wparam = fice * wfact * cd * wspeed;
            tx[k] = wparam * wx[k];
            ty[k] = wparam * wy[k];
        }
    }
    if (wslim > 0.f && wxymax > wslim) {
        for (int k = 0; k < n; k++) {
            const float wspeed = ws[k];
            if (wspeed <= wslim) continue;
            ws[k] = wslim;
            const float fact = wslim / wspeed;
            wx[k] = fact * wx[k];
            wy[k] = fact * wy[k];
            const float cd = get_drag(itdrag, ws[k], dragco);
            tx[k] = fice * wfact * cd * ws[k] * wx[k];
            ty[k] = fice * wfact * cd * ws[k] * wy[k];
        }
        wxymax = wslim;
    }
    return wxymax > wsmax ? 1 : 0;
}

/* set_mass_flux, bnd_routines.f90:1035-1183 (mode +1: volnode = hdknv *
 * areakv, elems_dealing.f90:741-800; jlhkv = 1 for z-layers).  Volumes are
 * reals summed in double, the share is double and every layer's part is
 * rounded to real, as the reference declares them. */
int oc_set_mass_flux(int nkn, int nlvdi, int nlv, const int *ilhkv, const real *hdknv, const real *areakv,
                     int nsrc, const int *node, const int *levmin, const int *levmax, const float *rflux,
                     const float *rqdsv, real *mfluxv, real *rqv) {
    for (size_t i = 0; i < (size_t)nlvdi * nkn; i++) mfluxv[i] = 0;
    for (int i = 0; i < nsrc; i++) {
        const int k = node[i];
        int lmax = ilhkv[k], lmin = 1;
        if (levmax[i] > 0 && levmax[i] < lmax) lmax = levmax[i];
        if (levmax[i] < 0 && lmax + 1 + levmax[i] < lmax) lmax = lmax + 1 + levmax[i];
        if (levmin[i] > 0 && levmin[i] > lmin) lmin = levmin[i];
        if (levmin[i] < 0 && lmax + 1 + levmin[i] > lmin) lmin = lmax + 1 + levmin[i];
        if (lmin > lmax || lmin < 1 || lmax > nlv) return 1;
        double dvols[256], dvoltot = 0.;
        for (int l = lmin; l <= lmax; l++) {
            const real vol = hdknv[(l - 1) + (size_t)nlvdi * k] * areakv[(l - 1) + (size_t)nlvdi * k];
            dvols[l - 1] = (double)vol;
            dvoltot += dvols[l - 1];
        }
        if (dvoltot <= 0.) return 1;
        for (int l = lmin; l <= lmax; l++) {
            const real part = (real)((double)rflux[i] * (dvols[l - 1] / dvoltot));
            mfluxv[(l - 1) + (size_t)nlvdi * k] += part;
        }
    }
    for (int k = 0; k < nkn; k++) {
        if (rqdsv) mfluxv[(size_t)nlvdi * k] += (real)rqdsv[k];
        double col = 0.;
        for (int l = 1; l <= ilhkv[k]; l++) col += (double)mfluxv[(l - 1) + (size_t)nlvdi * k];
        rqv[k] = (real)col;
    }
    return 0;
}
