/*
 * shyoracle_init.c -- init-time pieces of the CPU oracle (TEST INFRASTRUCTURE).
 *
 * Restates, in plain C, the reference's geometry (set_ev), zeta-matrix
 * pattern (construct_node_list / coo_init_new), element colouring
 * (domain_clusterization_ggu), nodal areas (set_area), TVD upwind tables
 * (tvd_upwind_init_elem) and the band solvers of penta.f90.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline use it.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "shyoracle.h"

#define EV(i, ie) ev[((i) - 1) + 19 * (ie)]

#ifdef _OPENMP
#include <omp.h>
#endif

/* host threads of the parallel loops (element columns, conz3d colours and
 * nodes); 1 = serial.  The results do not depend on it. */
int oc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

/* adjust_bc: evgeom.f90:1007-1042 -- make b (or c) sum to exactly zero */
static int adjust_bc(double *v1, double *v2, double *v3)
{
    double vv = *v1 + *v2 + *v3;
    if (vv == 0.) return 0;
    if (*v1 == 0. && *v2 == 0.) return 1;
    if (*v1 == 0. && *v3 == 0.) return 1;
    if (*v2 == 0. && *v3 == 0.) return 1;
    if (*v1 == 0.) {
        *v2 = *v2 - vv / 2.;
        *v3 = -*v2;
    } else if (*v2 == 0.) {
        *v3 = *v3 - vv / 2.;
        *v1 = -*v3;
    } else if (*v3 == 0.) {
        *v1 = *v1 - vv / 2.;
        *v2 = -*v1;
    } else {
        *v1 = *v1 - vv / 3.;
        *v2 = *v2 - vv / 3.;
        *v3 = *v3 - vv / 3.;
        *v3 = -*v1 - *v2;
    }
    return 0;
}

/* set_ev: evgeom.f90:167-349 (cartesian branch) */
void oc_set_ev(int nkn, int nel, const int *nen3v, const real *xgv, const real *ygv, double *ev)
{
    (void)nkn;
    const double one = 1., two = 2., four = 4., twofour = 24.;
    const double pi = four * atan(one);
    const double rad = 180. / pi;
    for (int ie = 0; ie < nel; ie++) {
        int k1 = nen3v[3 * ie], k2 = nen3v[3 * ie + 1], k3 = nen3v[3 * ie + 2];
        double x1 = xgv[k1], y1 = ygv[k1], x2 = xgv[k2], y2 = ygv[k2], x3 = xgv[k3], y3 = ygv[k3];
        double a1 = x2 * y3 - x3 * y2;
        double a2 = x3 * y1 - x1 * y3;
        double a3 = x1 * y2 - x2 * y1;
        double aj = (x2 - x1) * (y3 - y1) - (x3 - x1) * (y2 - y1);
        double aji = one / aj;
        double b1 = (y2 - y3) * aji, c1 = (x3 - x2) * aji;
        double b2 = (y3 - y1) * aji, c2 = (x1 - x3) * aji;
        double b3 = (y1 - y2) * aji, c3 = (x2 - x1) * aji;
        a1 *= aji; a2 *= aji; a3 *= aji;
        adjust_bc(&b1, &b2, &b3);
        adjust_bc(&c1, &c2, &c3);
        double ss1 = (x3 - x2) * (x3 - x2) + (y3 - y2) * (y3 - y2);
        double s1 = sqrt(ss1);
        double ss2 = (x3 - x1) * (x3 - x1) + (y3 - y1) * (y3 - y1);
        double s2 = sqrt(ss2);
        double ss3 = (x2 - x1) * (x2 - x1) + (y2 - y1) * (y2 - y1);
        double s3 = sqrt(ss3);
        double d1 = acos((ss2 + ss3 - ss1) / (two * s2 * s3)) * rad;
        double d2 = acos((ss1 + ss3 - ss2) / (two * s1 * s3)) * rad;
        double d3 = acos((ss1 + ss2 - ss3) / (two * s1 * s2)) * rad;
        double dd1 = aj * sqrt(b1 * b1 + c1 * c1);
        double dd2 = aj * sqrt(b2 * b2 + c2 * c2);
        double dd3 = aj * sqrt(b3 * b3 + c3 * c3);
        EV(1, ie) = a1; EV(2, ie) = a2; EV(3, ie) = a3;
        EV(4, ie) = b1; EV(5, ie) = b2; EV(6, ie) = b3;
        EV(7, ie) = c1; EV(8, ie) = c2; EV(9, ie) = c3;
        EV(10, ie) = aj / twofour;
        EV(11, ie) = d1; EV(12, ie) = d2; EV(13, ie) = d3;
        EV(14, ie) = dd1; EV(15, ie) = dd2; EV(16, ie) = dd3;
        EV(17, ie) = s1; EV(18, ie) = s2; EV(19, ie) = s3;
    }
}

static int cmp_int(const void *a, const void *b)
{
    int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

/* construct_node_list (coo_matrix.f90:330-404) + 2d part of coo_init_new
 * (coo_matrix.f90:51-172): row k = sorted unique {k} u edge neighbours;
 * ijp = COO slot of entry (i,j) of element ie.  Returns nnz, -1 on overflow. */
int oc_node_list(int nkn, int nel, const int *nen3v, int *ntg, int *cols, int maxnnz, int *ijp)
{
    int *cnt = (int *)calloc((size_t)nkn + 1, sizeof(int));
    for (int ie = 0; ie < nel; ie++)
        for (int ii = 0; ii < 3; ii++) cnt[nen3v[3 * ie + ii]] += 2;
    int *off = (int *)malloc(((size_t)nkn + 1) * sizeof(int));
    off[0] = 0;
    for (int k = 0; k < nkn; k++) off[k + 1] = off[k] + cnt[k] + 1;
    int *raw = (int *)malloc((size_t)off[nkn] * sizeof(int));
    int *fill = (int *)calloc((size_t)nkn, sizeof(int));
    for (int ie = 0; ie < nel; ie++) {
        for (int ii = 0; ii < 3; ii++) {
            int k1 = nen3v[3 * ie + ii], k2 = nen3v[3 * ie + (ii + 1) % 3];
            raw[off[k1] + fill[k1]++] = k2;
            raw[off[k2] + fill[k2]++] = k1;
        }
    }
    int nnz = 0;
    ntg[0] = 0;
    for (int k = 0; k < nkn; k++) {
        int *r = raw + off[k];
        int n = fill[k];
        r[n++] = k;
        qsort(r, (size_t)n, sizeof(int), cmp_int);
        int ip = 0;
        for (int i = 1; i < n; i++) {
            if (r[i] == r[ip]) continue;
            ip++;
            r[ip] = r[i];
        }
        n = ip + 1;
        if (nnz + n > maxnnz) { nnz = -1; break; }
        memcpy(cols + nnz, r, (size_t)n * sizeof(int));
        nnz += n;
        ntg[k + 1] = nnz;
    }
    if (nnz >= 0 && ijp) {
        for (int ie = 0; ie < nel; ie++) {
            for (int i = 0; i < 3; i++) {
                int ki = nen3v[3 * ie + i];
                for (int j = 0; j < 3; j++) {
                    int kj = nen3v[3 * ie + j];
                    int slot = -1;
                    for (int q = ntg[ki]; q < ntg[ki + 1]; q++)
                        if (cols[q] == kj) { slot = q; break; }
                    ijp[i + 3 * j + 9 * ie] = slot;
                }
            }
        }
    }
    free(cnt); free(off); free(raw); free(fill);
    return nnz;
}

/* ---------------- element colouring: subset.f90:371-711 ---------------- */

typedef struct {
    int nel;
    const int *nen3v;
    int *color;          /* [nel+1], 1-based element index */
    unsigned *colork;    /* [nkn] colour bitmask per node */
    int *subset_el;      /* [33] */
    int nsub;
} colctx;

static int may_color(colctx *c, int ie, int ic_from, int ic_to)
{
    if (c->color[ie] != ic_from) return 0;
    for (int ii = 0; ii < 3; ii++) {
        int k = c->nen3v[3 * (ie - 1) + ii];
        if ((c->colork[k] >> ic_to) & 1u) return 0;
    }
    return 1;
}

static void color_elem(colctx *c, int ie, int ic_new)
{
    int ic_old = c->color[ie];
    c->color[ie] = ic_new;
    for (int ii = 0; ii < 3; ii++) {
        int k = c->nen3v[3 * (ie - 1) + ii];
        unsigned col = c->colork[k];
        col &= ~(1u << ic_old);
        col |= (1u << ic_new);
        c->colork[k] = col;
    }
}

/* apply_one_color: subset.f90:594-622 (rotating sweep from ie_start) */
static int apply_one_color(colctx *c, int ic_old, int ic_new, int *ie_start, int nemax)
{
    int ncol = 0, n = 0, ie = *ie_start;
    for (;;) {
        n++;
        if (n > c->nel) break;
        if (ncol >= nemax) break;
        ie = ie + 1;
        if (ie > c->nel) ie = ie % c->nel;
        if (may_color(c, ie, ic_old, ic_new)) {
            ncol++;
            color_elem(c, ie, ic_new);
        }
    }
    *ie_start = ie;
    return ncol;
}

static int subset_filling(colctx *c)
{
    int nmax = 0;
    for (int i = 0; i <= 32; i++) c->subset_el[i] = 0;
    for (int ie = 1; ie <= c->nel; ie++) {
        int ic = c->color[ie];
        int n = ++c->subset_el[ic];
        if (n > nmax) nmax = n;
    }
    return nmax;
}

/* returns ncolor (>0) or -1 (more than 31 colours) / -2 (check_color failed).
 * color[ie] (0-based ie) receives the 1-based colour; color_elems lists the
 * elements colour by colour in ascending element order. */
int oc_color(int nkn, int nel, const int *nen3v, int *color_out, int *color_ptr, int *color_elems)
{
    colctx c;
    c.nel = nel;
    c.nen3v = nen3v;
    c.color = (int *)calloc((size_t)nel + 1, sizeof(int));
    c.colork = (unsigned *)calloc((size_t)nkn, sizeof(unsigned));
    int subset_el[33];
    c.subset_el = subset_el;
    /* compute_color :564-590 */
    int ic_new = 0, ie_start = 0, ncolor, ret = 0;
    for (;;) {
        ic_new++;
        if (ic_new > 31) { ret = -1; goto done; }
        int ncol = apply_one_color(&c, 0, ic_new, &ie_start, nel);
        if (ncol == 0) break;
    }
    ncolor = ic_new - 1;
    c.nsub = ncolor;
    subset_filling(&c);
    /* adjust_color :520-560 */
    {
        int naver = (int)lroundf((float)nel / (float)ncolor);
        int ie_st = 0;
        for (;;) {
            int ncmin = nel, ncmax = 0, icmin = 0, icmax = 0;
            for (int ic = 1; ic <= ncolor; ic++) {
                int n = subset_el[ic];
                if (n > ncmax) { ncmax = n; icmax = ic; }
                if (n < ncmin) { ncmin = n; icmin = ic; }
            }
            int nemax = ncmax - naver;
            if (naver - ncmin < nemax) nemax = naver - ncmin;
            int ncol = apply_one_color(&c, icmax, icmin, &ie_st, nemax);
            subset_filling(&c);
            if (ncol == 0) break;
        }
    }
    /* check_color :671-707 */
    {
        unsigned *ck = (unsigned *)malloc((size_t)nkn * sizeof(unsigned));
        memcpy(ck, c.colork, (size_t)nkn * sizeof(unsigned));
        for (int ie = 1; ie <= nel; ie++) {
            int ic = c.color[ie];
            for (int ii = 0; ii < 3; ii++) {
                int k = nen3v[3 * (ie - 1) + ii];
                if (((ck[k] >> ic) & 1u) == 0) ret = -2;
                ck[k] &= ~(1u << ic);
            }
        }
        for (int k = 0; k < nkn; k++)
            if (ck[k] != 0) ret = -2;
        free(ck);
        if (ret) goto done;
    }
    color_ptr[0] = 0;
    for (int ic = 1; ic <= ncolor; ic++) color_ptr[ic] = color_ptr[ic - 1] + subset_el[ic];
    {
        int *pos = (int *)calloc((size_t)ncolor + 1, sizeof(int));
        for (int ie = 1; ie <= nel; ie++) {
            int ic = c.color[ie];
            color_elems[color_ptr[ic - 1] + pos[ic]++] = ie - 1;
            color_out[ie - 1] = ic;
        }
        free(pos);
    }
    ret = ncolor;
done:
    free(c.color);
    free(c.colork);
    return ret;
}

/* diffweight: diffusion_admin.f90:182-340.  The horizontal diffusion
 * weights wdifhv(iii,ii,ie) at [(iii-1) + 3(ii-1) + 9 ie]: idtype 0 the
 * products b b + c c of the element's gradients; 1 the same with each
 * positive off-diagonal weight moved onto the third vertex of its column;
 * 2 -(2/3)/s^2 between the ends of each edge (s = ev(13+i), the length of
 * the edge opposite vertex i) and the diagonal the negated row sum.  The
 * checks of the reference (diagonal positive, off-diagonal not positive,
 * symmetric and zero row sums within 1e-3) fail with 1 for idtype 1 and
 * 2, as the reference stops; 0 otherwise. */
int oc_diffweight(int nel, const double *ev, int idtype, real *wdifhv)
{
    const double fact = (double)(2.f / 3.f), eps = (double)1.e-3f;
    int bad = 0;
    for (int ie = 0; ie < nel; ie++) {
        double b[3], c[3], bc[3][3];          /* bc[iii-1][ii-1] = bc(iii,ii) */
        for (int ii = 0; ii < 3; ii++) {
            b[ii] = ev[(3 + ii) + 19 * (size_t)ie];
            c[ii] = ev[(6 + ii) + 19 * (size_t)ie];
        }
        for (int ii = 0; ii < 3; ii++)
            for (int iii = 0; iii < 3; iii++) bc[iii][ii] = b[iii] * b[ii] + c[iii] * c[ii];
        if (idtype == 1) {
            for (int ii = 1; ii <= 3; ii++)
                for (int iii = 1; iii <= 3; iii++) {
                    const double w = bc[iii - 1][ii - 1];
                    if (w > 0. && iii != ii) {
                        const int i = 6 - ii - iii;
                        bc[iii - 1][ii - 1] = 0.;
                        bc[i - 1][ii - 1] = bc[i - 1][ii - 1] - w;
                    }
                }
        } else if (idtype == 2) {
            for (int i = 1; i <= 3; i++) {
                const int ia = 1 + i % 3, ib = 1 + (i + 1) % 3;
                const double s = ev[(12 + i) + 19 * (size_t)ie];
                double w = 1. / (s * s);
                w = w * fact;
                bc[ia - 1][ib - 1] = -w;
                bc[ib - 1][ia - 1] = -w;
                bc[i - 1][i - 1] = 0.;
            }
            for (int ii = 1; ii <= 3; ii++) {
                double w = 0.;
                for (int iii = 1; iii <= 3; iii++) w = w + bc[ii - 1][iii - 1];
                bc[ii - 1][ii - 1] = -w;
            }
        }
        int berror = 0;
        for (int ii = 1; ii <= 3; ii++) {
            double wacu = 0.;
            for (int iii = 1; iii <= 3; iii++) {
                const double w = bc[ii - 1][iii - 1];
                wacu = wacu + w;
                if (ii == iii) {
                    if (w <= 0) berror = 1;
                } else if (w > 0) {
                    berror = 1;
                }
                if (fabs(w - bc[iii - 1][ii - 1]) > eps) berror = 1;
                wdifhv[(ii - 1) + 3 * (iii - 1) + 9 * (size_t)ie] = (real)w;
            }
            if (fabs(wacu) > eps) berror = 1;
        }
        if (berror && idtype != 0) bad = 1;
    }
    return bad;
}

/* set_area: elems_dealing.f90:658-700 (areafv = 12*ev(10)/3 summed per layer) */
void oc_set_area(const oc_mesh *m, real *areakv)
{
    const double *ev = m->ev;
    int nlvdi = m->nlvdi;
    memset(areakv, 0, sizeof(real) * (size_t)nlvdi * (size_t)m->nkn);
    for (int ie = 0; ie < m->nel; ie++) {
        real areael = (real)(12. * EV(10, ie));
        real areafv = areael / (real)3;
        int nlev = m->ilhv[ie];
        for (int ii = 0; ii < 3; ii++) {
            int k = m->nen3v[3 * ie + ii];
            for (int l = 1; l <= nlev; l++) areakv[(l - 1) + nlvdi * k] += areafv;
        }
    }
}

/* in_element (src/utils/shyutil/find.f90:365-400) with intri
 * (src/utils/shyutil/regular.f90:1252-1282): the element's bounding box in
 * default reals, then each edge's outward test in double precision of
 * default-real differences; a point on an edge is inside. */
static int in_element_ref(const int *kn, const real *xgv, const real *ygv, float xp, float yp)
{
    float x[3], y[3];
    for (int ii = 0; ii < 3; ii++) { x[ii] = (float)xgv[kn[ii]]; y[ii] = (float)ygv[kn[ii]]; }
    float xmin = x[0], xmax = x[0], ymin = y[0], ymax = y[0];
    for (int ii = 1; ii < 3; ii++) {
        if (x[ii] < xmin) xmin = x[ii];
        if (x[ii] > xmax) xmax = x[ii];
        if (y[ii] < ymin) ymin = y[ii];
        if (y[ii] > ymax) ymax = y[ii];
    }
    if (!(xp >= xmin && xp <= xmax && yp >= ymin && yp <= ymax)) return 0;
    for (int k1 = 0; k1 < 3; k1++) {
        int k2 = (k1 + 1) % 3;
        double x21 = (double)(float)(x[k2] - x[k1]);
        double y21 = (double)(float)(y[k2] - y[k1]);
        double yn = x21, xn = -y21;
        double scal = (double)(float)(xp - x[k1]) * xn + (double)(float)(yp - y[k1]) * yn;
        if (scal < 0.) return 0;
    }
    return 1;
}

/* quad_tree_search (src/utils/shyutil/quad_tree.f90:313-397): of the
 * elements containing the point (in_element), the one with the highest
 * external number, here the element index (the basin's own numbering); -1
 * if none.  The quad tree only narrows the candidates: every element whose
 * bounding box holds the point is in the leaf the search reaches, so a
 * uniform grid of cells gives the same answer. */
static int find_elem(const oc_mesh *m, const real *xgv, const real *ygv,
                     const int *cell_ptr, const int *cell_el, double x0, double y0,
                     double dxc, double dyc, int ncx, int ncy, float px, float py)
{
    int cx = (int)floor(((double)px - x0) / dxc), cy = (int)floor(((double)py - y0) / dyc);
    if (cx < 0 || cy < 0 || cx >= ncx || cy >= ncy) return -1;
    int best = -1;
    int cell = cx + ncx * cy;
    for (int q = cell_ptr[cell]; q < cell_ptr[cell + 1]; q++) {
        int ie = cell_el[q];
        if (ie < best) continue;
        if (in_element_ref(m->nen3v + 3 * ie, xgv, ygv, px, py)) best = ie;
    }
    return best;
}

/* tvd_upwind_init_elem: tvd_admin.f90:361-474.  For vertex ic of element ie
 * and the other vertex id, the upwind point is 2*x(ic) - x(id), stored as a
 * default real, with the element quad_tree_search finds there (or -1). */
int oc_tvd_init(const oc_mesh *m, const real *xgv, const real *ygv, int *tvd_ie, real *tvd_x, real *tvd_y)
{
    int nkn = m->nkn, nel = m->nel;
    double xmin = 1e300, xmax = -1e300, ymin = 1e300, ymax = -1e300;
    for (int k = 0; k < nkn; k++) {
        if (xgv[k] < xmin) xmin = xgv[k];
        if (xgv[k] > xmax) xmax = xgv[k];
        if (ygv[k] < ymin) ymin = ygv[k];
        if (ygv[k] > ymax) ymax = ygv[k];
    }
    int ncx = (int)sqrt((double)nel) + 1, ncy = ncx;
    double dxc = (xmax - xmin) / ncx * (1 + 1e-9) + 1e-9, dyc = (ymax - ymin) / ncy * (1 + 1e-9) + 1e-9;
    int ncell = ncx * ncy;
    int *cnt = (int *)calloc((size_t)ncell + 1, sizeof(int));
    /* bounding boxes into cells */
    int *cell_ptr = (int *)calloc((size_t)ncell + 1, sizeof(int));
    for (int ie = 0; ie < nel; ie++) {
        const int *kn = m->nen3v + 3 * ie;
        double bx0 = 1e300, bx1 = -1e300, by0 = 1e300, by1 = -1e300;
        for (int ii = 0; ii < 3; ii++) {
            double x = xgv[kn[ii]], y = ygv[kn[ii]];
            if (x < bx0) bx0 = x; if (x > bx1) bx1 = x;
            if (y < by0) by0 = y; if (y > by1) by1 = y;
        }
        int cx0 = (int)floor((bx0 - xmin) / dxc), cx1 = (int)floor((bx1 - xmin) / dxc);
        int cy0 = (int)floor((by0 - ymin) / dyc), cy1 = (int)floor((by1 - ymin) / dyc);
        for (int cy = cy0; cy <= cy1; cy++)
            for (int cx = cx0; cx <= cx1; cx++) cnt[cx + ncx * cy]++;
    }
    for (int c = 0; c < ncell; c++) cell_ptr[c + 1] = cell_ptr[c] + cnt[c];
    int *cell_el = (int *)malloc((size_t)cell_ptr[ncell] * sizeof(int) + 1);
    memset(cnt, 0, sizeof(int) * (size_t)ncell);
    for (int ie = 0; ie < nel; ie++) {
        const int *kn = m->nen3v + 3 * ie;
        double bx0 = 1e300, bx1 = -1e300, by0 = 1e300, by1 = -1e300;
        for (int ii = 0; ii < 3; ii++) {
            double x = xgv[kn[ii]], y = ygv[kn[ii]];
            if (x < bx0) bx0 = x; if (x > bx1) bx1 = x;
            if (y < by0) by0 = y; if (y > by1) by1 = y;
        }
        int cx0 = (int)floor((bx0 - xmin) / dxc), cx1 = (int)floor((bx1 - xmin) / dxc);
        int cy0 = (int)floor((by0 - ymin) / dyc), cy1 = (int)floor((by1 - ymin) / dyc);
        for (int cy = cy0; cy <= cy1; cy++)
            for (int cx = cx0; cx <= cx1; cx++) {
                int c = cx + ncx * cy;
                cell_el[cell_ptr[c] + cnt[c]++] = ie;
            }
    }
    for (int ie = 0; ie < nel; ie++) {
        for (int ii = 1; ii <= 3; ii++) {
            int kc = m->nen3v[3 * ie + ii - 1];
            double xc = xgv[kc], yc = ygv[kc];
            for (int t = 0; t < 2; t++) {
                int j = (t == 0) ? (ii % 3) + 1 : ((ii + 1) % 3) + 1;
                int kd = m->nen3v[3 * ie + j - 1];
                double xd = xgv[kd], yd = ygv[kd];
                /* the point is a default real, as the reference stores it */
                float x = (float)(2 * xc - xd), y = (float)(2 * yc - yd);
                int ienew = find_elem(m, xgv, ygv, cell_ptr, cell_el, xmin, ymin, dxc, dyc, ncx, ncy, x, y);
                int idx = (j - 1) + 3 * (ii - 1) + 9 * ie;
                tvd_ie[idx] = ienew;
                tvd_x[idx] = x;
                tvd_y[idx] = y;
            }
        }
        /* diagonal entries (id == ic) are unused */
        for (int ii = 1; ii <= 3; ii++) {
            int idx = (ii - 1) + 3 * (ii - 1) + 9 * ie;
            tvd_ie[idx] = -1;
            tvd_x[idx] = 0;
            tvd_y[idx] = 0;
        }
    }
    free(cnt); free(cell_ptr); free(cell_el);
    return 0;
}

/* ---------------- band solvers: penta.f90 ---------------- */
#define V5(j, i) val[((j) + 2) + 5 * ((i) - 1)]

/* penta_fact: penta.f90:215-321.  Returns 0, or the row of a zero pivot. */
int oc_penta_fact(int n, double *val)
{
    if (n < 4) return -1;
    double *x = (double *)malloc(sizeof(double) * (size_t)(n + 2));
    double *y = (double *)malloc(sizeof(double) * (size_t)(n + 2));
    double *z = (double *)malloc(sizeof(double) * (size_t)(n + 2));
    const double zero = 0.;
    int i, ret = 0;
    double xx;
    z[1] = zero;
    i = 1;
    xx = V5(0, i);
    if (xx == zero) { ret = i; goto out; }
    x[i] = xx;
    y[i] = V5(1, i);
    z[i + 1] = V5(-1, i + 1) / xx;
    V5(-2, i + 2) = V5(-2, i + 2) / xx;
    i = 2;
    xx = V5(0, i) - y[i - 1] * z[i];
    if (xx == zero) { ret = i; goto out; }
    x[i] = xx;
    y[i] = V5(1, i) - z[i] * V5(2, i - 1);
    z[i + 1] = (V5(-1, i + 1) - V5(-2, i + 1) * y[i - 1]) / xx;
    V5(-2, i + 2) = V5(-2, i + 2) / xx;
    for (i = 3; i <= n - 2; i++) {
        xx = V5(0, i) - y[i - 1] * z[i] - V5(-2, i) * V5(2, i - 2);
        if (xx == zero) { ret = i; goto out; }
        x[i] = xx;
        y[i] = V5(1, i) - z[i] * V5(2, i - 1);
        z[i + 1] = (V5(-1, i + 1) - V5(-2, i + 1) * y[i - 1]) / xx;
        V5(-2, i + 2) = V5(-2, i + 2) / xx;
    }
    i = n - 1;
    xx = V5(0, i) - y[i - 1] * z[i] - V5(-2, i) * V5(2, i - 2);
    if (xx == zero) { ret = i; goto out; }
    x[i] = xx;
    y[i] = V5(1, i) - z[i] * V5(2, i - 1);
    z[i + 1] = (V5(-1, i + 1) - V5(-2, i + 1) * y[i - 1]) / xx;
    i = n;
    xx = V5(0, i) - y[i - 1] * z[i] - V5(-2, i) * V5(2, i - 2);
    if (xx == zero) { ret = i; goto out; }
    x[i] = xx;
    y[i] = zero;
    for (i = 1; i <= n; i++) {
        V5(0, i) = x[i];
        V5(1, i) = y[i];
        V5(-1, i) = z[i];
    }
out:
    free(x); free(y); free(z);
    return ret;
}

/* penta_solve: penta.f90:104-139 */
void oc_penta_solve(int n, const double *val, const double *r, double *s)
{
    double *g = (double *)malloc(sizeof(double) * (size_t)(n + 1));
    g[1] = r[0];
    g[2] = r[1] - V5(-1, 2) * g[1];
    for (int i = 3; i <= n; i++) g[i] = r[i - 1] - V5(-1, i) * g[i - 1] - V5(-2, i) * g[i - 2];
    s[n - 1] = g[n] / V5(0, n);
    s[n - 2] = (g[n - 1] - V5(1, n - 1) * s[n - 1]) / V5(0, n - 1);
    for (int i = n - 2; i >= 1; i--)
        s[i - 1] = (g[i] - V5(1, i) * s[i] - V5(2, i) * s[i + 1]) / V5(0, i);
    free(g);
}

/* tria_multi: penta.f90:332-368; v[(j-1)+3*(i-1)]: j=1 lower, 2 diag, 3 upper;
 * r and s are (n,m) column-major */
void oc_tria_multi(int n, int m, double *v, double *r, double *s)
{
#define V3(j, i) v[((j) - 1) + 3 * ((i) - 1)]
#define R(i, j) r[((i) - 1) + n * ((j) - 1)]
#define S(i, j) s[((i) - 1) + n * ((j) - 1)]
    for (int i = 2; i <= n; i++) {
        double aux = V3(1, i) / V3(2, i - 1);
        V3(2, i) = V3(2, i) - aux * V3(3, i - 1);
        for (int j = 1; j <= m; j++) R(i, j) = R(i, j) - aux * R(i - 1, j);
    }
    for (int j = 1; j <= m; j++) {
        S(n, j) = R(n, j) / V3(2, n);
        for (int i = n - 1; i >= 1; i--) S(i, j) = (R(i, j) - V3(3, i) * S(i + 1, j)) / V3(2, i);
    }
#undef V3
#undef R
#undef S
}
