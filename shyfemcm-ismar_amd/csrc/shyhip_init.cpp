// shyhip_init.cpp -- host-side, init-time preprocessing of the basin.
//
// These run once per basin on the CPU and produce the integer structures and
// FP64 geometry the device kernels consume:
//   host_set_ev        element geometry, src/utils/shyutil/evgeom.f90:167-349
//   host_node_list     zeta-matrix pattern + element slots,
//                      src/shyfem/coo_matrix.f90:51-172,330-404
//   host_color         greedy element colouring + balancing,
//                      src/shyfem/subset.f90:371-711
//   host_tvd_tables    upwind points of the itvd=2 scheme with their
//                      interpolation weights, src/shyfem/tvd_admin.f90:361-474,
//                      src/utils/shyutil/evgeom.f90:531-660
//   host_diffweight    horizontal diffusion weights of the scalars,
//                      src/shyfem/diffusion_admin.f90:182-340
//   host_block_schedule  block-ILU(0) pattern and level sets (no reference
//                      counterpart: replaces ILUT's sequential factorisation)
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "shyhip_internal.h"

namespace shy {

// fn(begin, end) over [0, n) split into contiguous ranges on up to 16 host
// threads (one when n is below `grain`); the ranges must write disjoint data
void parallel_for(int n, int grain, const std::function<void(int, int)> &fn) {
    const int hw = (int)std::max(1u, std::thread::hardware_concurrency());
    const int nt = std::max(1, std::min({16, hw, n / std::max(1, grain)}));
    if (nt <= 1) {
        fn(0, n);
        return;
    }
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; t++) {
        const int b = (int)((long long)n * t / nt), e = (int)((long long)n * (t + 1) / nt);
        pool.emplace_back([&fn, b, e] { fn(b, e); });
    }
    for (auto &th : pool) th.join();
}

// b and c coefficients are corrected so that each triple sums to zero
// (evgeom.f90:1007-1042)
static void zero_sum(double &v1, double &v2, double &v3) {
    const double vv = v1 + v2 + v3;
    if (vv == 0.) return;
    if ((v1 == 0. && v2 == 0.) || (v1 == 0. && v3 == 0.) || (v2 == 0. && v3 == 0.)) return;
    if (v1 == 0.) {
        v2 = v2 - vv / 2.;
        v3 = -v2;
    } else if (v2 == 0.) {
        v3 = v3 - vv / 2.;
        v1 = -v3;
    } else if (v3 == 0.) {
        v1 = v1 - vv / 2.;
        v2 = -v1;
    } else {
        v1 = v1 - vv / 3.;
        v2 = v2 - vv / 3.;
        v3 = v3 - vv / 3.;
        v3 = -v1 - v2;
    }
}

void host_set_ev(int nkn, int nel, const int *nen3v0, const float *xg, const float *yg, double *ev) {
    (void)nkn;
    const double pi = 4. * std::atan(1.);
    const double rad = 180. / pi;
    // every element on its own: ranges of them on the host's threads
    parallel_for(nel, 8192, [&](int e_begin, int e_end) {
    for (int ie = e_begin; ie < e_end; ie++) {
        const int *kn = nen3v0 + 3 * ie;
        const double x1 = xg[kn[0]], y1 = yg[kn[0]];
        const double x2 = xg[kn[1]], y2 = yg[kn[1]];
        const double x3 = xg[kn[2]], y3 = yg[kn[2]];
        const double aj = (x2 - x1) * (y3 - y1) - (x3 - x1) * (y2 - y1);
        const double aji = 1. / aj;
        double a[3] = {(x2 * y3 - x3 * y2) * aji, (x3 * y1 - x1 * y3) * aji, (x1 * y2 - x2 * y1) * aji};
        double b[3] = {(y2 - y3) * aji, (y3 - y1) * aji, (y1 - y2) * aji};
        double c[3] = {(x3 - x2) * aji, (x1 - x3) * aji, (x2 - x1) * aji};
        zero_sum(b[0], b[1], b[2]);
        zero_sum(c[0], c[1], c[2]);
        const double ss1 = (x3 - x2) * (x3 - x2) + (y3 - y2) * (y3 - y2);
        const double ss2 = (x3 - x1) * (x3 - x1) + (y3 - y1) * (y3 - y1);
        const double ss3 = (x2 - x1) * (x2 - x1) + (y2 - y1) * (y2 - y1);
        const double s1 = std::sqrt(ss1), s2 = std::sqrt(ss2), s3 = std::sqrt(ss3);
        double *e = ev + 19 * (size_t)ie;
        for (int i = 0; i < 3; i++) {
            e[i] = a[i];
            e[3 + i] = b[i];
            e[6 + i] = c[i];
            e[13 + i] = aj * std::sqrt(b[i] * b[i] + c[i] * c[i]);
        }
        e[9] = aj / 24.;
        e[10] = std::acos((ss2 + ss3 - ss1) / (2. * s2 * s3)) * rad;
        e[11] = std::acos((ss1 + ss3 - ss2) / (2. * s1 * s3)) * rad;
        e[12] = std::acos((ss1 + ss2 - ss3) / (2. * s1 * s2)) * rad;
        e[16] = s1;
        e[17] = s2;
        e[18] = s3;
    }
    });
}

int host_node_list(int nkn, int nel, const int *nen3v0, std::vector<int> &ntg, std::vector<int> &cols,
                   std::vector<int> &ijp) {
    // adjacency through element edges, plus the node itself, sorted and
    // unique: every edge end and the node in one array by node (counted
    // first), then each node's run sorted and compacted
    std::vector<int> cnt(nkn + 1, 1);                      // the node itself
    for (int ie = 0; ie < nel; ie++)
        for (int ii = 0; ii < 3; ii++) {
            cnt[nen3v0[3 * ie + ii]] += 1;
            cnt[nen3v0[3 * ie + (ii + 1) % 3]] += 1;
        }
    std::vector<int> start(nkn + 1, 0);
    for (int k = 0; k < nkn; k++) start[k + 1] = start[k] + cnt[k];
    std::vector<int> raw(start[nkn]), fill(start.begin(), start.end() - 1);
    for (int k = 0; k < nkn; k++) raw[fill[k]++] = k;
    for (int ie = 0; ie < nel; ie++)
        for (int ii = 0; ii < 3; ii++) {
            const int k1 = nen3v0[3 * ie + ii], k2 = nen3v0[3 * ie + (ii + 1) % 3];
            raw[fill[k1]++] = k2;
            raw[fill[k2]++] = k1;
        }
    std::vector<int> len(nkn);
    parallel_for(nkn, 8192, [&](int k_begin, int k_end) {
        for (int k = k_begin; k < k_end; k++) {
            int *a = raw.data() + start[k];
            std::sort(a, a + cnt[k]);
            len[k] = (int)(std::unique(a, a + cnt[k]) - a);
        }
    });
    ntg.assign(nkn + 1, 0);
    for (int k = 0; k < nkn; k++) ntg[k + 1] = ntg[k] + len[k];
    cols.resize(ntg[nkn]);
    for (int k = 0; k < nkn; k++) std::copy(raw.begin() + start[k], raw.begin() + start[k] + len[k], cols.begin() + ntg[k]);
    ijp.assign(9 * (size_t)nel, -1);
    parallel_for(nel, 8192, [&](int e_begin, int e_end) {
        for (int ie = e_begin; ie < e_end; ie++)
            for (int i = 0; i < 3; i++) {
                const int ki = nen3v0[3 * ie + i];
                for (int j = 0; j < 3; j++) {
                    const int kj = nen3v0[3 * ie + j];
                    auto first = cols.begin() + ntg[ki], last = cols.begin() + ntg[ki + 1];
                    auto it = std::lower_bound(first, last, kj);
                    if (it == last || *it != kj) continue;       // cannot happen: kj is a neighbour
                    ijp[9 * (size_t)ie + 3 * i + j] = (int)(it - cols.begin());
                }
            }
    });
    for (size_t q = 0; q < ijp.size(); q++)
        if (ijp[q] < 0) return -1;
    return (int)cols.size();
}

// Per element the 3x3 weights w(a, b) of horizontal diffusion (wdifhv):
// idtype 0 the products of the vertex gradients, grad(phi_a) . grad(phi_b);
// 1 those with each positive off-diagonal weight of a column b moved onto
// the column's third vertex; 2 the weight -(2/3)/s^2 between the two ends
// of each edge (s its length, ev 14-16) and on the diagonal the negated sum
// of the row.  The reference's 2/3 and tolerance are single-precision
// constants.  Out: w[(3 (b-1) + (a-1)) nel + ie] (column b contiguous for
// an element's three vertices).  Returns the first element (1-based) whose
// weights break the reference's checks -- diagonal positive, off-diagonal
// not positive, symmetric and zero row sums within the tolerance -- for
// idtype 1 and 2, where the reference stops; 0 when all pass.
int host_diffweight(int nel, const double *ev, int idtype, std::vector<double> &w) {
    const double two_thirds = (double)(2.f / 3.f), tol = (double)1.e-3f;
    w.assign(9 * (size_t)nel, 0.);
    int first_bad = 0;
    for (int ie = 0; ie < nel; ie++) {
        const double *e = ev + 19 * (size_t)ie;
        double m[3][3];                       // m[a][b] = w(a+1, b+1)
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) m[a][b] = e[3 + a] * e[3 + b] + e[6 + a] * e[6 + b];
        if (idtype == 1) {
            for (int b = 0; b < 3; b++)
                for (int a = 0; a < 3; a++) {
                    const double v = m[a][b];
                    if (a != b && v > 0.) {
                        m[a][b] = 0.;
                        m[3 - a - b][b] -= v;
                    }
                }
        } else if (idtype == 2) {
            for (int v = 0; v < 3; v++) {             // the edge opposite vertex v
                const int p = (v + 1) % 3, q = (v + 2) % 3;
                const double len = e[13 + v];
                const double x = (1. / (len * len)) * two_thirds;
                m[p][q] = -x;
                m[q][p] = -x;
                m[v][v] = 0.;
            }
            for (int a = 0; a < 3; a++) m[a][a] = -(m[a][0] + m[a][1] + m[a][2]);
        }
        bool bad = false;
        for (int a = 0; a < 3; a++) {
            double row = 0.;
            for (int b = 0; b < 3; b++) {
                row += m[a][b];
                if (a == b ? m[a][b] <= 0. : m[a][b] > 0.) bad = true;
                if (std::fabs(m[a][b] - m[b][a]) > tol) bad = true;
                w[(size_t)(3 * b + a) * nel + ie] = m[a][b];
            }
            if (std::fabs(row) > tol) bad = true;
        }
        if (bad && idtype != 0 && first_bad == 0) first_bad = ie + 1;
    }
    return first_bad;
}

// Greedy colouring in a rotating sweep; a node's colour bitmask forbids two
// elements of one colour sharing a node; then elements move from the largest
// to the smallest colour until no move is possible.
int host_color(int nkn, int nel, const int *nen3v0, std::vector<int> &color) {
    std::vector<int> col(nel + 1, 0);
    std::vector<uint32_t> nodemask(nkn, 0u);
    auto allowed = [&](int ie, int from, int to) {
        if (col[ie] != from) return false;
        const int *kn = nen3v0 + 3 * (ie - 1);
        for (int ii = 0; ii < 3; ii++)
            if ((nodemask[kn[ii]] >> to) & 1u) return false;
        return true;
    };
    auto recolor = [&](int ie, int to) {
        const int from = col[ie];
        col[ie] = to;
        const int *kn = nen3v0 + 3 * (ie - 1);
        for (int ii = 0; ii < 3; ii++) nodemask[kn[ii]] = (nodemask[kn[ii]] & ~(1u << from)) | (1u << to);
    };
    auto sweep = [&](int from, int to, int &start, int limit) {
        int moved = 0, ie = start;
        for (int n = 1; n <= nel && moved < limit; n++) {
            ie = ie + 1;
            if (ie > nel) ie = ie % nel;
            if (allowed(ie, from, to)) {
                moved++;
                recolor(ie, to);
            }
        }
        start = ie;
        return moved;
    };
    int ncolor = 0, start = 0;
    for (int c = 1;; c++) {
        if (c > 31) return -1;
        if (sweep(0, c, start, nel) == 0) {
            ncolor = c - 1;
            break;
        }
    }
    std::vector<int> count(ncolor + 1, 0);
    auto fill = [&]() {
        std::fill(count.begin(), count.end(), 0);
        for (int ie = 1; ie <= nel; ie++) count[col[ie]]++;
    };
    fill();
    const int naver = (int)std::lround((float)nel / (float)ncolor);
    int start2 = 0;
    for (;;) {
        int nmin = nel, nmax = 0, cmin = 0, cmax = 0;
        for (int c = 1; c <= ncolor; c++) {
            if (count[c] > nmax) { nmax = count[c]; cmax = c; }
            if (count[c] < nmin) { nmin = count[c]; cmin = c; }
        }
        const int limit = std::min(nmax - naver, naver - nmin);
        const int moved = sweep(cmax, cmin, start2, limit);
        fill();
        if (moved == 0) break;
    }
    color.assign(nel, 0);
    for (int ie = 1; ie <= nel; ie++) color[ie - 1] = col[ie];
    // every element coloured, no node shared within a colour
    std::vector<uint32_t> seen(nkn, 0u);
    for (int ie = 0; ie < nel; ie++) {
        const int c = color[ie];
        if (c < 1 || c > ncolor) return -2;
        for (int ii = 0; ii < 3; ii++) {
            uint32_t &m = seen[nen3v0[3 * ie + ii]];
            if ((m >> c) & 1u) return -2;
            m |= 1u << c;
        }
    }
    return ncolor;
}

// natural coordinates of point (x,y) in element ie, clipped and renormalised
// as adjust_xi does (evgeom.f90:577-660)
static bool natural_coords(const double *ev, int ie, double x, double y, double xi[3]) {
    const double *e = ev + 19 * (size_t)ie;
    double sum = 0., clipped = 0.;
    int nzero = 0, zsum = 0;
    for (int i = 0; i < 3; i++) {
        xi[i] = e[i] + e[3 + i] * x + e[6 + i] * y;
        sum += xi[i];
        xi[i] = std::min(1., std::max(0., xi[i]));
        if (xi[i] == 0.) {
            zsum += i + 1;
            nzero++;
        }
        clipped += xi[i];
    }
    if (nzero == 3 || std::fabs(sum - 1.) > (double)1.e-8f) return false;
    const double adj = (clipped - 1.) / (3 - nzero);
    if (nzero == 2) {
        xi[6 - zsum - 1] = 1.;
    } else {
        for (int i = 0; i < 3; i++)
            if (i + 1 != zsum) xi[i] = std::min(1., std::max(0., xi[i] - adj));
    }
    return std::fabs(xi[0] + xi[1] + xi[2] - 1.) <= (double)1.e-8f;
}

// For each ordered vertex pair (ic,id) of an element the upwind point is
// 2*x(ic) - x(id) (stored as a default real); the element containing it is
// found as quad_tree_search finds it (highest element number on shared edges)
// and the interpolation weights are stored.  Pair p = 2*(ic-1) + t enumerates id =
// mod(ic,3)+1 (t=0) and mod(ic+1,3)+1 (t=1).
void host_tvd_tables(int nkn, int nel, const int *nen3v0, const float *xg, const float *yg, const double *ev,
                     std::vector<int> &tel, std::vector<int> &tkn, std::vector<double> &tw) {
    double xmin = 1e300, xmax = -1e300, ymin = 1e300, ymax = -1e300;
    for (int k = 0; k < nkn; k++) {
        xmin = std::min(xmin, (double)xg[k]);
        xmax = std::max(xmax, (double)xg[k]);
        ymin = std::min(ymin, (double)yg[k]);
        ymax = std::max(ymax, (double)yg[k]);
    }
    const int nc = (int)std::sqrt((double)nel) + 1;
    const double dxc = (xmax - xmin) / nc * (1 + 1e-9) + 1e-9, dyc = (ymax - ymin) / nc * (1 + 1e-9) + 1e-9;
    // the elements of each cell (every cell its bounding box touches), in
    // ascending element number, as one array by cell: counted, then filled
    const size_t ncell = (size_t)nc * nc;
    std::vector<int> cbox(4 * (size_t)nel);
    for (int ie = 0; ie < nel; ie++) {
        const int *kn = nen3v0 + 3 * ie;
        double bx0 = 1e300, bx1 = -1e300, by0 = 1e300, by1 = -1e300;
        for (int ii = 0; ii < 3; ii++) {
            bx0 = std::min(bx0, (double)xg[kn[ii]]);
            bx1 = std::max(bx1, (double)xg[kn[ii]]);
            by0 = std::min(by0, (double)yg[kn[ii]]);
            by1 = std::max(by1, (double)yg[kn[ii]]);
        }
        cbox[4 * (size_t)ie] = (int)std::floor((bx0 - xmin) / dxc);
        cbox[4 * (size_t)ie + 1] = (int)std::floor((bx1 - xmin) / dxc);
        cbox[4 * (size_t)ie + 2] = (int)std::floor((by0 - ymin) / dyc);
        cbox[4 * (size_t)ie + 3] = (int)std::floor((by1 - ymin) / dyc);
    }
    std::vector<int> cell_ptr(ncell + 1, 0);
    for (int ie = 0; ie < nel; ie++)
        for (int cy = cbox[4 * (size_t)ie + 2]; cy <= cbox[4 * (size_t)ie + 3]; cy++)
            for (int cx = cbox[4 * (size_t)ie]; cx <= cbox[4 * (size_t)ie + 1]; cx++)
                cell_ptr[(size_t)cx + (size_t)nc * cy + 1]++;
    for (size_t c = 0; c < ncell; c++) cell_ptr[c + 1] += cell_ptr[c];
    std::vector<int> cell_el(cell_ptr[ncell]), cell_fill(cell_ptr.begin(), cell_ptr.end() - 1);
    for (int ie = 0; ie < nel; ie++)
        for (int cy = cbox[4 * (size_t)ie + 2]; cy <= cbox[4 * (size_t)ie + 3]; cy++)
            for (int cx = cbox[4 * (size_t)ie]; cx <= cbox[4 * (size_t)ie + 1]; cx++)
                cell_el[cell_fill[(size_t)cx + (size_t)nc * cy]++] = ie;
    // quad_tree_search (src/utils/shyutil/quad_tree.f90:313-397): of the
    // elements containing the point by in_element (find.f90:365-400: bounding
    // box in default reals, then intri's edge tests, regular.f90:1252-1282, in
    // double of default-real differences, edges inside), the one with the
    // highest number.  Every element whose bounding box holds the point is in
    // the point's cell, so the cells give the quad tree's answer.
    auto inside = [&](int ie, float xp, float yp) {
        const int *kn = nen3v0 + 3 * ie;
        float x[3], y[3];
        for (int ii = 0; ii < 3; ii++) { x[ii] = xg[kn[ii]]; y[ii] = yg[kn[ii]]; }
        const float x0 = std::min(x[0], std::min(x[1], x[2])), x1 = std::max(x[0], std::max(x[1], x[2]));
        const float y0 = std::min(y[0], std::min(y[1], y[2])), y1 = std::max(y[0], std::max(y[1], y[2]));
        if (!(xp >= x0 && xp <= x1 && yp >= y0 && yp <= y1)) return false;
        for (int k1 = 0; k1 < 3; k1++) {
            const int k2 = (k1 + 1) % 3;
            const double ex = (double)(float)(x[k2] - x[k1]), ey = (double)(float)(y[k2] - y[k1]);
            const double scal = (double)(float)(xp - x[k1]) * -ey + (double)(float)(yp - y[k1]) * ex;
            if (scal < 0.) return false;
        }
        return true;
    };
    auto locate = [&](float px, float py) {
        const int cx = (int)std::floor(((double)px - xmin) / dxc), cy = (int)std::floor(((double)py - ymin) / dyc);
        if (cx < 0 || cy < 0 || cx >= nc || cy >= nc) return -1;
        int best = -1;
        const size_t c = (size_t)cx + (size_t)nc * cy;
        for (int q = cell_ptr[c]; q < cell_ptr[c + 1]; q++) {
            const int ie = cell_el[q];
            if (ie > best && inside(ie, px, py)) best = ie;
        }
        return best;
    };
    tel.assign(6 * (size_t)nel, -1);
    tkn.assign(18 * (size_t)nel, 0);
    tw.assign(18 * (size_t)nel, 0.);
    // the elements are independent: ranges of them on the host's threads
    parallel_for(nel, 4096, [&](int e_begin, int e_end) {
    for (int ie = e_begin; ie < e_end; ie++) {
        for (int ic = 1; ic <= 3; ic++) {
            const int kc = nen3v0[3 * ie + ic - 1];
            for (int t = 0; t < 2; t++) {
                const int id = (t == 0) ? (ic % 3) + 1 : ((ic + 1) % 3) + 1;
                const int kd = nen3v0[3 * ie + id - 1];
                const float xu = (float)(2. * (double)xg[kc] - (double)xg[kd]);
                const float yu = (float)(2. * (double)yg[kc] - (double)yg[kd]);
                const int p = 2 * (ic - 1) + t;
                const int up = locate(xu, yu);
                if (up < 0) continue;
                double xi[3];
                if (!natural_coords(ev, up, (double)xu, (double)yu, xi)) continue;
                tel[(size_t)p * nel + ie] = up;
                for (int ii = 0; ii < 3; ii++) {
                    tkn[((size_t)ii * 6 + p) * nel + ie] = nen3v0[3 * up + ii];
                    tw[((size_t)ii * 6 + p) * nel + ie] = xi[ii];
                }
            }
        }
    }
    });
}

// Scalar open-boundary tables (bndo.f90:210-470): inward normal from the
// neighbouring open-boundary nodes along the basin boundary (the boundary
// neighbours of a node are the ends of its counter-clockwise node list,
// connections.f90:594-612), and the inner vertices of the elements touching
// the node with area weights, accumulated in element order and normalised.
// Weights and normals are rounded like the reference's default reals when
// the model runs in FP32.
int host_bndo_tables(int nkn, int nel, const int *nen3v0, const float *x, const float *y, const double *ev,
                     const std::vector<int> &bnodes, std::vector<double> &nx, std::vector<double> &ny,
                     std::vector<int> &ptr, std::vector<int> &nbr, std::vector<double> &w, bool f32) {
    auto rnd = [f32](double v) { return f32 ? (double)(float)v : v; };
    std::vector<long long> key(3 * (size_t)nel);
    for (int ie = 0; ie < nel; ie++)
        for (int ii = 0; ii < 3; ii++) {
            const long long a = nen3v0[3 * ie + ii], c = nen3v0[3 * ie + (ii + 1) % 3];
            key[3 * (size_t)ie + ii] = std::min(a, c) * nkn + std::max(a, c);
        }
    // how many elements share each edge: the edges by their lower node
    // (counted, then filled), each node's far ends sorted
    std::vector<int> eptr(nkn + 1, 0);
    for (size_t q = 0; q < key.size(); q++) eptr[key[q] / nkn + 1]++;
    for (int k = 0; k < nkn; k++) eptr[k + 1] += eptr[k];
    std::vector<int> efar(key.size()), efill(eptr.begin(), eptr.end() - 1);
    for (size_t q = 0; q < key.size(); q++) efar[efill[key[q] / nkn]++] = (int)(key[q] % nkn);
    parallel_for(nkn, 8192, [&](int k_begin, int k_end) {
        for (int k = k_begin; k < k_end; k++) std::sort(efar.begin() + eptr[k], efar.begin() + eptr[k + 1]);
    });
    auto shared_by = [&](long long kk) {
        const int lo = (int)(kk / nkn), far = (int)(kk % nkn);
        auto r = std::equal_range(efar.begin() + eptr[lo], efar.begin() + eptr[lo + 1], far);
        return r.second - r.first;
    };
    std::vector<int> knext(nkn, -1), klast(nkn, -1);
    for (int ie = 0; ie < nel; ie++)
        for (int ii = 0; ii < 3; ii++) {
            const long long kk = key[3 * (size_t)ie + ii];
            if (shared_by(kk) == 1) {
                knext[nen3v0[3 * ie + ii]] = nen3v0[3 * ie + (ii + 1) % 3];
                klast[nen3v0[3 * ie + (ii + 1) % 3]] = nen3v0[3 * ie + ii];
            }
        }
    const int nb = (int)bnodes.size();
    std::vector<int> iop(nkn, 0);
    for (int i = 0; i < nb; i++) iop[bnodes[i]] = i + 1;
    nx.assign(nb, 0.);
    ny.assign(nb, 0.);
    for (int i = 0; i < nb; i++) {
        const int k = bnodes[i], kn = knext[k], kl = klast[k];
        const bool hn = kn >= 0 && iop[kn] > 0, hl = kl >= 0 && iop[kl] > 0;
        double dx, dy;
        if (hn && hl) { dx = rnd((double)x[kn] - x[kl]); dy = rnd((double)y[kn] - y[kl]); }
        else if (hn) { dx = rnd((double)x[kn] - x[k]); dy = rnd((double)y[kn] - y[k]); }
        else if (hl) { dx = rnd((double)x[k] - x[kl]); dy = rnd((double)y[k] - y[kl]); }
        else return -1;                       // one-node open boundary
        nx[i] = -dy;
        ny[i] = dx;
    }
    std::vector<std::vector<int>> ln(nb);
    std::vector<std::vector<double>> lw(nb);
    for (int ie = 0; ie < nel; ie++) {
        const double area = rnd(12. * ev[19 * (size_t)ie + 9]);
        for (int ii = 0; ii < 3; ii++) {
            const int ib = iop[nen3v0[3 * ie + ii]];
            if (ib < 1) continue;
            for (int iii = 0; iii < 3; iii++) {
                const int kn = nen3v0[3 * ie + iii];
                if (iop[kn] > 0) continue;
                auto &L = ln[ib - 1];
                auto &W = lw[ib - 1];
                size_t j = 0;
                for (; j < L.size(); j++)
                    if (L[j] == kn) break;
                if (j < L.size()) W[j] = rnd(W[j] + area);
                else { L.push_back(kn); W.push_back(area); }
            }
        }
    }
    ptr.assign(nb + 1, 0);
    nbr.clear();
    w.clear();
    for (int i = 0; i < nb; i++) {
        double tot = 0.;
        for (double v : lw[i]) tot = rnd(tot + v);
        for (size_t j = 0; j < ln[i].size(); j++) {
            nbr.push_back(ln[i][j]);
            w.push_back(tot > 0. ? rnd(lw[i][j] / tot) : lw[i][j]);
        }
        ptr[i + 1] = (int)nbr.size();
    }
    return nb;
}

// Block-Jacobi ILU(0): rows are split into contiguous blocks of `brows`;
// couplings outside the diagonal block are dropped.  For each block the
// lower and upper triangular level sets are built for a one-workgroup
// level-scheduled factorisation and solve.
void host_block_schedule(int n, const std::vector<int> &rowptr, const std::vector<int> &cols, int brows,
                         BlockSchedule &bs) {
    bs.brows = brows;
    bs.nblk = (n + brows - 1) / brows;
    bs.rowptr.assign(n + 1, 0);
    bs.cols.clear();
    bs.src.clear();
    bs.diag.assign(n, -1);
    for (int i = 0; i < n; i++) {
        const int b0 = (i / brows) * brows, b1 = std::min(n, b0 + brows);
        for (int q = rowptr[i]; q < rowptr[i + 1]; q++) {
            const int j = cols[q];
            if (j < b0 || j >= b1) continue;
            if (j == i) bs.diag[i] = (int)bs.cols.size();
            bs.cols.push_back(j - b0);
            bs.src.push_back(q);
        }
        bs.rowptr[i + 1] = (int)bs.cols.size();
    }
    const int stride = brows + 1;
    bs.fwd_rows.assign(n, 0);
    bs.bwd_rows.assign(n, 0);
    bs.fwd_lptr.assign((size_t)bs.nblk * stride, 0);
    bs.bwd_lptr.assign((size_t)bs.nblk * stride, 0);
    bs.fwd_nlev.assign(bs.nblk, 0);
    bs.bwd_nlev.assign(bs.nblk, 0);
    // blocks are independent: ranges of them on the host's threads
    parallel_for(bs.nblk, 16, [&](int blk_begin, int blk_end) {
    std::vector<int> lev(brows);
    for (int b = blk_begin; b < blk_end; b++) {
        const int b0 = b * brows, nb = std::min(n, b0 + brows) - b0;
        for (int dir = 0; dir < 2; dir++) {
            int nlev = 0;
            for (int t = 0; t < nb; t++) {
                const int li = dir == 0 ? t : nb - 1 - t;
                const int i = b0 + li;
                int lv = 0;
                for (int q = bs.rowptr[i]; q < bs.rowptr[i + 1]; q++) {
                    const int lj = bs.cols[q];
                    if ((dir == 0 && lj < li) || (dir == 1 && lj > li)) lv = std::max(lv, lev[lj] + 1);
                }
                lev[li] = lv;
                nlev = std::max(nlev, lv + 1);
            }
            std::vector<int> cnt(nlev + 1, 0);
            for (int li = 0; li < nb; li++) cnt[lev[li] + 1]++;
            for (int l = 0; l < nlev; l++) cnt[l + 1] += cnt[l];
            int *lptr = (dir == 0 ? bs.fwd_lptr.data() : bs.bwd_lptr.data()) + (size_t)b * stride;
            int *rows = (dir == 0 ? bs.fwd_rows.data() : bs.bwd_rows.data()) + b0;
            for (int l = 0; l <= nlev; l++) lptr[l] = cnt[l];
            std::vector<int> pos(cnt.begin(), cnt.end() - 1);
            for (int li = 0; li < nb; li++) rows[pos[lev[li]]++] = li;
            (dir == 0 ? bs.fwd_nlev : bs.bwd_nlev)[b] = nlev;
        }
    }
    });
    // update pairs of the fine-grained factorisation, in ascending k: each
    // entry's pairs counted, then written at their offsets (rows in parallel)
    const size_t ne = bs.cols.size();
    bs.cp_ptr.assign(ne + 1, 0);
    bs.ent_jd.assign(ne, -1);
    auto row_pairs = [&](int i, bool write) {
        const int b0 = (i / brows) * brows, li = i - b0;
        for (int q = bs.rowptr[i]; q < bs.rowptr[i + 1]; q++) {
            const int lj = bs.cols[q];
            if (!write) bs.ent_jd[q] = lj < li ? bs.diag[b0 + lj] : (lj == li ? -2 - li : -1);
            const int m = std::min(li, lj);
            int c = write ? bs.cp_ptr[q] : 0;
            for (int qk = bs.rowptr[i]; qk < bs.rowptr[i + 1]; qk++) {
                const int lk = bs.cols[qk];
                if (lk >= m) break;
                const int r0 = bs.rowptr[b0 + lk], r1 = bs.rowptr[b0 + lk + 1];
                const int *hit = std::lower_bound(bs.cols.data() + r0, bs.cols.data() + r1, lj);
                if (hit != bs.cols.data() + r1 && *hit == lj) {
                    if (write) {
                        bs.cp_a[c] = qk;
                        bs.cp_b[c] = (int)(hit - bs.cols.data());
                    }
                    c++;
                }
            }
            if (!write) bs.cp_ptr[q + 1] = c;
        }
    };
    parallel_for(n, 8192, [&](int r_begin, int r_end) {
        for (int i = r_begin; i < r_end; i++) row_pairs(i, false);
    });
    for (size_t q = 0; q < ne; q++) bs.cp_ptr[q + 1] += bs.cp_ptr[q];
    bs.cp_a.assign(bs.cp_ptr[ne], 0);
    bs.cp_b.assign(bs.cp_ptr[ne], 0);
    parallel_for(n, 8192, [&](int r_begin, int r_end) {
        for (int i = r_begin; i < r_end; i++) row_pairs(i, true);
    });
    // entry levels: every entry an update pair or a pivot of q refers to
    // lies in an earlier row of its block or earlier in the same row, so one
    // pass in entry order sees them computed (blocks in parallel)
    bs.ent_lev.assign(ne, 0);
    bs.ent_nlev.assign(bs.nblk, 0);
    parallel_for(bs.nblk, 16, [&](int blk_begin, int blk_end) {
        for (int blk = blk_begin; blk < blk_end; blk++)
            for (int i = blk * brows; i < std::min(n, (blk + 1) * brows); i++)
                for (int q = bs.rowptr[i]; q < bs.rowptr[i + 1]; q++) {
                    int lv = 0;
                    for (int p = bs.cp_ptr[q]; p < bs.cp_ptr[q + 1]; p++)
                        lv = std::max(lv, std::max(bs.ent_lev[bs.cp_a[p]], bs.ent_lev[bs.cp_b[p]]) + 1);
                    if (bs.ent_jd[q] >= 0) lv = std::max(lv, bs.ent_lev[bs.ent_jd[q]] + 1);
                    bs.ent_lev[q] = lv;
                    bs.ent_nlev[blk] = std::max(bs.ent_nlev[blk], lv + 1);
                }
    });
}

}  // namespace shy
