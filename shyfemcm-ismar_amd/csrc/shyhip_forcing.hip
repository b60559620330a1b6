// shyhip_forcing.hip -- the forcing that precedes hydro in a time step
// (sp111(2), bnd_routines.f90:444-596, and meteo_force, meteo_admin.f90:196)
// evaluated on the device from record series held in HBM.
//
// A forcing file is a series of records (time, values).  The reference keeps
// a ring of nintp records (iff_populate_records, intp_fem_file.f90:1047-1144),
// reads the next record while the ring's check time is before the requested
// time (iff_read_and_interpolate, :2452-2510; tcomp, :2824-2849) and
// interpolates with Neville's scheme over the ring in slot order
// (iff_final_time_interpolate, :2514-2700; rd_intp_neville,
// utils/generic/intp.f90:317-345).  For non-decreasing times the ring's
// newest record has a closed form -- the first record j >= jstart+nintp-1
// whose check time time[j + 1 - nintp/2] is not before t (the last record at
// the end of the file) -- and slot q holds the record of the window with
// (j - jstart) mod nintp == q.  So every work-item finds its window by a
// binary search over the resident times and no ring state is kept: a whole
// step's forcing is one launch that needs nothing from the host but t.
//
// One launch: work-items [0, nb) are the open-boundary nodes (boundary level
// series, adjust_bound's zfact, bnd_routines.f90:1003-1031, into rzv), work-
// items [nb, nb + nw) the nodes of the wind series (meteo_convert_wind_data,
// meteo_forcing.f90:864-1000, into the stress per unit density).
#include <hip/hip_runtime.h>

#include <cmath>

#include "shyhip_internal.h"

namespace shy {

constexpr int kForceThreads = 256;
constexpr float kNoData = -999.f;     // record value meaning "no data" (the files' flag)

// newest record of the window for time t: the smallest j in [lo, hi] whose
// check time time[j + off] is not before t (hi when there is none).  The
// first guess assumes evenly spaced records and is confirmed by two loads
// issued together; a binary search over what remains covers the rest.
__device__ inline int window_last(const SeriesDev &s, double t) {
    const int off = 1 - s.nintp / 2;
    int lo = s.jstart + s.nintp - 1, hi = s.nrec - 1;
    const double t0 = s.time[lo + off], t1 = s.time[hi + off];
    if (t <= t0) return lo;
    if (t > t1) return hi;
    // here the answer lies in (lo, hi]
    int g = lo + (int)ceil((t - t0) / (t1 - t0) * (double)(hi - lo));
    g = min(hi, max(lo + 1, g));
    const double tg = s.time[g + off], tp = s.time[g - 1 + off];
    if (tg >= t && tp < t) return g;
    if (tg >= t) hi = g - 1;     // time[g-1+off] >= t as well
    else lo = g + 1;             // time[g+off] < t
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (s.time[mid + off] >= t) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

// component comp of value pos of the series at time t (vdefault 0 where a
// record of the window holds no data)
__device__ inline float series_value(const SeriesDev &s, int comp, int pos, double t) {
    const size_t stride = (size_t)s.ncomp * s.nval;
    if (s.nintp == 0) {
        const float v = s.val[(size_t)comp * s.nval + pos];
        return v == kNoData ? 0.f : v;
    }
    const int jl = window_last(s, t);
    double xa[4], p[4];
    bool none = false;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        if (q >= s.nintp) break;
        const int j = jl - (jl - s.jstart - q) % s.nintp;
        xa[q] = s.time[j];
        const float v = s.val[j * stride + (size_t)comp * s.nval + pos];
        none |= v == kNoData;
        p[q] = (double)v;
    }
    if (none) return 0.f;
    const int n = s.nintp - 1;
    for (int k = 1; k <= n; k++)
        for (int i = n; i >= k; i--) {
            const double xl = xa[i - k], xh = xa[i];
            p[i] = ((t - xl) * p[i] - (t - xh) * p[i - 1]) / (xh - xl);
        }
    return (float)p[n];
}

// get_drag, meteo_forcing.f90:1580-1608 (itdrag 0: the given coefficient)
__device__ inline float drag(int itdrag, float wxy, float cd) {
    if (itdrag == 1) return 0.001f * (0.63f + 0.066f * wxy);
    if (itdrag == 2) return wxy > 11.f ? 0.001f * (0.49f + 0.066f * wxy) : 0.001f * 1.2f;
    return cd;
}

// meteo_convert_wind_data for one node, no ice cover: iwtype 1 (wind
// velocity) or 2 (stress in N/m2); the speed limit wslim applies per node
// (the reference's gate on the largest speed only skips nodes at or below
// it).  Returns the speed after the limit.
__device__ inline float convert_wind(const WindPar &p, float wx, float wy, float &tx, float &ty) {
    const float fice = 1.f;
    float ws;
    if (p.iwtype == 2) {
        const float wfact = 1.f / p.rowass;
        const float cd = p.dragco > 0.f ? p.dragco : 2.5e-3f;
        tx = fice * wfact * wx;
        ty = fice * wfact * wy;
        const float txy = sqrtf(tx * tx + ty * ty);
        ws = sqrtf(txy / (cd * wfact * p.roluft));
        wx = tx / (cd * ws);
        wy = ty / (cd * ws);
    } else {
        const float wfact = p.roluft / p.rowass;
        ws = (float)sqrt((double)wx * (double)wx + (double)wy * (double)wy);
        const float cd = drag(p.itdrag, ws, p.dragco);
        const float wparam = fice * wfact * cd * ws;
        tx = wparam * wx;
        ty = wparam * wy;
    }
    if (p.wslim > 0.f && ws > p.wslim) {
        const float wfact = p.iwtype == 2 ? 1.f / p.rowass : p.roluft / p.rowass;
        const float fact = p.wslim / ws;
        ws = p.wslim;
        wx = fact * wx;
        wy = fact * wy;
        const float cd = drag(p.itdrag, ws, p.dragco);
        tx = fice * wfact * cd * ws * wx;
        ty = fice * wfact * cd * ws * wy;
    }
    return ws;
}

template <typename S>
__global__ __launch_bounds__(kForceThreads) void k_forcing(S *rzv, S *taux, S *tauy, int *err, const int *bnd_nodes,
                                                           const int *bnd_id, const int *bnd_pos, const int *ibtyp,
                                                           const SeriesDev *bser, const float *zfact, int nb,
                                                           SeriesDev wind, WindPar wp, int nw, double t) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nb) {
        const int b = bnd_id[i];
        if (ibtyp[b] != 1) return;
        const SeriesDev s = bser[b - 1];
        if (!s.time) return;
        const float rw = series_value(s, 0, s.nval == 1 ? 0 : bnd_pos[i], t);
        rzv[bnd_nodes[i]] = (S)(rw * zfact[b - 1]);
        return;
    }
    const int k = i - nb;
    if (k >= nw) return;
    const int pos = wind.nval == 1 ? 0 : k;
    const float wx = series_value(wind, 0, pos, t), wy = series_value(wind, 1, pos, t);
    float tx, ty;
    const float ws = convert_wind(wp, wx, wy, tx, ty);
    taux[k] = (S)tx;
    tauy[k] = (S)ty;
    if (ws > wp.wsmax) atomicCAS(err, 0, ERR_WIND);
}

template <typename S>
static int forcing_t(Model &M, double t) {
    Forcing &f = M.f;
    const int nb = f.nseries > 0 ? M.m.nbnd_nodes : 0;
    // a constant wind (one record) is converted on its first evaluation
    // only, as meteo_forcing_fem converts it on the first call
    // (meteo_forcing.f90:508-511)
    const int nw = (f.wind_on && !(f.wind.nrec == 1 && f.wind_applied)) ? M.nkn : 0;
    if (nb + nw == 0) return 0;
    const int g = (nb + nw + kForceThreads - 1) / kForceThreads;
    SeriesDev wind = f.wind;
    if (!f.wind_on) wind = SeriesDev{};
    hipLaunchKernelGGL(k_forcing<S>, dim3(g), dim3(kForceThreads), 0, M.stream, (S *)M.s.rzv, (S *)M.s.taux,
                       (S *)M.s.tauy, M.d_err, (const int *)M.m.bnd_nodes, (const int *)M.m.bnd_id,
                       (const int *)f.bnd_pos, (const int *)M.m.ibtyp, (const SeriesDev *)f.d_bser,
                       (const float *)f.d_zfact, nb, wind, f.wpar, nw, t);
    HIPCHK(M, hipGetLastError());
    if (nw > 0) f.wind_applied = true;
    return 0;
}

int launch_forcing(Model &M, double dtime) {
    return M.prec == SHYHIP_FP64 ? forcing_t<double>(M, dtime) : forcing_t<float>(M, dtime);
}

// set_mass_flux (bnd_routines.f90:1035-1183), one work-item per node.  The
// node's point sources, in the caller's (the reference's boundary loop)
// order, spread their discharge over layers lmin..lmax (levmin/levmax: > 0
// a layer, < 0 counted from the bottom) in proportion to the layer volumes
// hdknv*areakv (volnode with mode +1); the distributed source rqdsv goes to
// the top layer; rqv is the column sum, added in double.  Types follow the
// reference: volumes of the storage type summed in double, the share in
// double, each layer's part rounded to the storage type.
template <typename S>
__global__ void __launch_bounds__(kForceThreads)
k_mass_flux(S *__restrict__ mflux, S *__restrict__ rqv, int *__restrict__ err, const S *__restrict__ hdkn,
            const S *__restrict__ areakv, const int *__restrict__ ilhkv, const int *__restrict__ src_ptr,
            const int *__restrict__ src_lev, const int *__restrict__ src_idx, const float *__restrict__ src_q,
            const float *__restrict__ rqds, int nkn, int nlv) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nkn) return;
    const int lk = ilhkv[k];
    for (int l = 0; l < nlv; l++) mflux[(size_t)l * nkn + k] = (S)0;
    const int p0 = src_ptr ? src_ptr[k] : 0, p1 = src_ptr ? src_ptr[k + 1] : 0;
    for (int p = p0; p < p1; p++) {
        const int levmin = src_lev[2 * p], levmax = src_lev[2 * p + 1];
        int lmax = lk, lmin = 1;
        if (levmax > 0) lmax = min(lmax, levmax);
        if (levmax < 0) lmax = min(lmax, lmax + 1 + levmax);
        if (levmin > 0) lmin = max(lmin, levmin);
        if (levmin < 0) lmin = max(lmin, lmax + 1 + levmin);
        if (lmin > lmax || lmin < 1 || lmax > nlv) {
            atomicCAS(err, 0, ERR_MFLUX);
            return;
        }
        double voltot = 0.;
        for (int l = lmin - 1; l < lmax; l++) {
            const size_t o = (size_t)l * nkn + k;
            voltot += (double)(S)(hdkn[o] * areakv[o]);
        }
        if (voltot <= 0.) {
            atomicCAS(err, 0, ERR_MFLUX);
            return;
        }
        const float q = src_q[src_idx[p]];
        for (int l = lmin - 1; l < lmax; l++) {
            const size_t o = (size_t)l * nkn + k;
            const double share = (double)(S)(hdkn[o] * areakv[o]) / voltot;
            mflux[o] += (S)((double)q * share);
        }
    }
    if (rqds) mflux[k] += (S)rqds[k];
    double col = 0.;
    for (int l = 0; l < lk; l++) col += (double)mflux[(size_t)l * nkn + k];
    rqv[k] = (S)col;
}

template <typename S>
static int mass_flux_t(Model &M, bool with_rqds) {
    Forcing &f = M.f;
    const int g = (M.nkn + kForceThreads - 1) / kForceThreads;
    hipLaunchKernelGGL(k_mass_flux<S>, dim3(g), dim3(kForceThreads), 0, M.stream, (S *)M.s.mflux, (S *)M.s.rqv,
                       M.d_err, (const S *)M.s.hdkn, (const S *)M.m.areakv, (const int *)M.m.ilhkv,
                       (const int *)(f.nsrc > 0 ? f.src_ptr : nullptr), (const int *)f.src_lev,
                       (const int *)f.src_idx, (const float *)f.src_q, with_rqds ? (const float *)f.rqds : nullptr,
                       M.nkn, M.nlv);
    HIPCHK(M, hipGetLastError());
    return 0;
}

int launch_mass_flux(Model &M, bool with_rqds) {
    return M.prec == SHYHIP_FP64 ? mass_flux_t<double>(M, with_rqds) : mass_flux_t<float>(M, with_rqds);
}

}  // namespace shy
