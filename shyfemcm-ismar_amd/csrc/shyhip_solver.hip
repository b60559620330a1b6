// shyhip_solver.hip -- zeta system solve on gfx950.
//
// Preconditioner: block-Jacobi ILU(0).  Rows are split into contiguous blocks
// (<= 1024 rows, one workgroup each); each diagonal block is factorised in
// LDS with level scheduling (rows of one level in parallel, workgroup
// barriers between levels).  Every factor entry (i,j) is computed as
// a_ij - sum_k l_ik u_kj over the pattern, k ascending -- the updates of
// ilu0's IKJ loop (spk_ilut.f90:1832-1950) in their order, listed once on
// the host -- so a basin of at most one block of nodes gets exactly ILU(0);
// lusol :2309-2345 applies it.  The block factors are staged in LDS by every
// kernel that applies them.
//
// Krylov method: BiCGStab in the arrangement the reference uses -- runrc with
// ipar(2)=1, which bisinit maps to LEFT preconditioning (spk_iters.f90:
// 1061-1334, 3433-3500) -- with stopping test ||M^-1 r|| <= rtol*||M^-1 r0||
// + atol and at most itermax matvecs.  One iteration is two kernels: every
// vector update is fused into the SpMV/preconditioner kernel that consumes
// it (the update of x and r closing an iteration is done by the first kernel
// of the next, from five dot products of the second), and every dot product
// is written as per-workgroup partials that every workgroup of the next
// kernel reduces itself in the same fixed order -- no atomics or
// device-scope fences, visibility comes from the kernel boundary, and all
// workgroups (and all runs) see bitwise identical scalars.
// The FP32 path runs this in FP32 inside an FP64 iterative refinement whose
// control also lives on the device: every kernel of a refinement round
// returns at once when the FP64 residual test has passed, so the host
// enqueues the expected number of rounds and reads the flags once.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>

#include <algorithm>
#include <cmath>
#include <functional>

#include "shyhip_internal.h"

namespace shy {

// scalar slots; rho/alpha/omega are double-buffered by iteration parity so a
// kernel never overwrites a value that another workgroup of it still reads
enum { SC_TARGET = 0, SC_NORM0 = 1, SC_RHO = 2, SC_ALPHA = 4, SC_OMEGA = 6, SC_RNORM = 8, SC_RES64 = 9,
       SC_BNORM = 10 };
// partial-sum regions: [region][q][workgroup]
enum { PR_A = 0, PR_B = 1, PR_C = 2, PR_R = 3 };

template <typename V>
struct SolvArgs {
    int n, nblk, brows, mx, mxp, itermax, sweeps;
    double rtol, atol;
    const int *rowptr, *cols;
    const V *A;
    const int *brp, *bcol, *bsrc, *bdiag, *frows, *flptr, *fnlev, *brows_l, *blptr, *bnlev;
    const int *cp_ptr, *cp_a, *cp_b, *bjd;
    const int *pcol, *prp, *pdg;     // block pattern at a fixed stride per block
    const int *elp, *elr, *enlev;    // entry-level factorisation schedule
    const int4 *fdesc;               // per entry in level order: {entry, pairs begin, end, pivot}
    const int2 *fpair;               // update pairs {l_ik, u_kj} entries, relative to the block
    int edmax;
    const int *ell_col, *ell_len;    // ELL slices of the matrix rows
    const V *ell_a;
    // block factor rows slot-major ([k*n + i]: slot 0 the inverted pivot,
    // slots 1..nL the L entries, then the U entries; local columns) and per
    // row nL | cnt << 8; fs_val null: not used
    const int *fs_col, *fs_meta;
    V *fs_val;
    V *lu;                           // block factors at a fixed stride per block
    const double *aval;
    V *b, *x0, *r0, *r, *p0, *p1, *v0, *v1, *s, *t, *dx;
    double *part, *scal;
    int *ctrl;
    int nparts;      // stride of one partial-sum array
};

template <typename V>
static SolvArgs<V> solv_args(Model &M, const V *A, V *b, V *x0) {
    // zeroed first: the bytes, padding included, identify the launches
    SolvArgs<V> a;
    std::memset(&a, 0, sizeof a);
    a.n = M.nkn; a.nblk = M.sv.nblk; a.brows = M.sv.brows; a.mx = M.sv.max_blk_nnz; a.mxp = M.sv.max_blk_pairs;
    a.itermax = M.p.itermax;
    a.sweeps = M.p.tri_sweeps;
    a.rtol = M.p.rtol; a.atol = M.p.atol;
    a.rowptr = M.m.rowptr; a.cols = M.m.cols; a.A = A;
    a.brp = M.m.blk_rowptr; a.bcol = M.m.blk_cols; a.bsrc = M.m.blk_src; a.bdiag = M.m.blk_diag;
    a.frows = M.m.fwd_rows; a.flptr = M.m.fwd_lptr; a.fnlev = M.m.fwd_nlev;
    a.brows_l = M.m.bwd_rows; a.blptr = M.m.bwd_lptr; a.bnlev = M.m.bwd_nlev;
    a.cp_ptr = M.m.blk_cp_ptr; a.cp_a = M.m.blk_cp_a; a.cp_b = M.m.blk_cp_b; a.bjd = M.m.blk_jd;
    a.pcol = M.m.blk_pcol; a.prp = M.m.blk_prp; a.pdg = M.m.blk_pdg;
    a.elp = M.m.blk_elp; a.elr = M.m.blk_elr; a.enlev = M.m.blk_enlev; a.edmax = M.sv.max_elev;
    a.fdesc = (const int4 *)M.m.blk_fdesc; a.fpair = (const int2 *)M.m.blk_fpair;
    a.ell_col = M.m.ell_col; a.ell_len = M.m.ell_len; a.ell_a = (const V *)M.sv.ell_a;
    a.fs_col = M.m.fs_col; a.fs_meta = M.m.fs_meta;
    // factor rows in registers: one row per work-item and Jacobi sweeps
    a.fs_val = (M.p.tri_sweeps > 0 && M.sv.brows <= kThreads) ? (V *)M.sv.fs_val : nullptr;
    a.lu = (V *)M.sv.lu; a.aval = M.s.aval;
    a.b = b; a.x0 = x0;
    a.r0 = (V *)M.sv.r0; a.r = (V *)M.sv.r; a.p0 = (V *)M.sv.p0; a.p1 = (V *)M.sv.p1;
    a.v0 = (V *)M.sv.v0; a.v1 = (V *)M.sv.v1; a.s = (V *)M.sv.s; a.t = (V *)M.sv.t; a.dx = (V *)M.sv.dx;
    a.part = M.sv.part; a.scal = M.sv.scal; a.ctrl = M.sv.ctrl;
    a.nparts = M.sv.nparts;
    return a;
}

// ---------------------------------------------------------------- reductions
// Workgroup sums: butterfly within each 64-lane wavefront (every lane ends
// with the same total), then the wavefront totals added in a fixed order.
// The same data always give the same bits.
constexpr int kWaves = kThreads / 64;
constexpr int kRegion = 8;            // partial-sum slots per region

// v from the DPP-permuted lane (CTRL: a GFX9 DPP control code whose source
// lane always lies in the same row of 16, so every lane is written)
template <int CTRL>
__device__ inline double dpp_f64(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

__device__ inline double readlane_f64(double v, int lane) {
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), lane),
                            __builtin_amdgcn_readlane(__double2loint(v), lane));
}

// Sum over the 64 lanes, the same in every lane: exchanges within rows of
// 16 by DPP (lanes i^1, i^2 by quad permutes, then the mirrored halves of
// 8 and of 16 lanes, which hold equal values by then), then the four row
// sums read from lanes 0, 16, 32, 48.  Each exchange adds two values that
// the partner lane adds in the other order, so all lanes agree bitwise.
__device__ inline double wave_sum(double v) {
    v = v + dpp_f64<0xB1>(v);      // quad_perm [1,0,3,2]
    v = v + dpp_f64<0x4E>(v);      // quad_perm [2,3,0,1]
    v = v + dpp_f64<0x141>(v);     // row_half_mirror
    v = v + dpp_f64<0x140>(v);     // row_mirror
    return (readlane_f64(v, 0) + readlane_f64(v, 16)) + (readlane_f64(v, 32) + readlane_f64(v, 48));
}

template <int NQ>
__device__ inline void block_sum(double (&v)[NQ]) {
    __shared__ double red[NQ][kWaves];
    for (int q = 0; q < NQ; q++) v[q] = wave_sum(v[q]);
    if ((threadIdx.x & 63) == 0)
        for (int q = 0; q < NQ; q++) red[q][threadIdx.x >> 6] = v[q];
    __syncthreads();
    for (int q = 0; q < NQ; q++) {
        double t = red[q][0];
        for (int w = 1; w < kWaves; w++) t += red[q][w];
        v[q] = t;
    }
    __syncthreads();
}

// per-workgroup partials: part[(region*kRegion + q)*nparts + slot], slot =
// the workgroup's block (or its index)
template <int NQ>
__device__ inline void store_partials(double (&vals)[NQ], double *part, int region, int nparts, int slot) {
    block_sum<NQ>(vals);
    if (threadIdx.x == 0)
        for (int q = 0; q < NQ; q++) part[((size_t)region * kRegion + q) * nparts + slot] = vals[q];
}

// Block of a workgroup: xcd_block(w, n) gives each XCD a contiguous range of
// blocks, so the blocks a block couples to (its neighbours in node order)
// and its own factors are mostly produced on the same XCD and read from
// its L2.

// total of the partials of the previous kernel (n workgroups), computed by
// every workgroup in the same order; the result is returned to all threads
template <int NQ>
__device__ inline void sum_partials(const double *part, int region, int nparts, int n, double (&out)[NQ]) {
    for (int q = 0; q < NQ; q++) out[q] = 0.;
    for (int i = threadIdx.x; i < n; i += kThreads)
        for (int q = 0; q < NQ; q++) out[q] += part[((size_t)region * kRegion + q) * nparts + i];
    block_sum<NQ>(out);
}

// The same total with this work-item's first partials loaded early (the
// loads are issued at kernel start, the reduction done when needed); only
// the first nq of the NQ slots are read, the others are zero.
template <int NQ>
struct Partials {
    double v[NQ];
    __device__ inline void load(const double *part, int region, int nparts, int n, int nq) {
#pragma unroll
        for (int q = 0; q < NQ; q++)
            v[q] = (q < nq && (int)threadIdx.x < n) ? part[((size_t)region * kRegion + q) * nparts + threadIdx.x] : 0.;
    }
    __device__ inline void total(const double *part, int region, int nparts, int n, int nq, double (&out)[NQ]) {
#pragma unroll
        for (int q = 0; q < NQ; q++) out[q] = 0. + v[q];
        for (int i = threadIdx.x + kThreads; i < n; i += kThreads)
            for (int q = 0; q < nq; q++) out[q] += part[((size_t)region * kRegion + q) * nparts + i];
        block_sum<NQ>(out);
    }
};

__device__ inline bool breakdown(double x) {   // brkdn: zero or not finite
    if (x > 0.) return !(1. / x > 0.);
    if (x < 0.) return !(1. / x < 0.);
    return true;
}

// ---------------------------------------------------------------- block solve
// One diagonal block of the preconditioner: row starts, diagonal positions,
// local columns and factor values (diagonal stored inverted), either in
// global memory or staged in LDS.  Entry indices are relative to the arrays
// the view points at.
template <typename V>
struct BlockView {
    const int *rp, *dg, *col;
    const V *lu;
};

// LDS bytes of a kernel that stages its block: y and w (brows values each),
// the block's factor values and columns, row starts and diagonals
template <typename V>
static size_t staged_lds(int brows, int mx) {
    return (size_t)(2 * brows + mx) * sizeof(V) + (size_t)(mx + 2 * brows + 1) * sizeof(int);
}

// The block's factors, columns, row starts and diagonals at the block's
// fixed stride: load() issues the global loads into registers (up to
// kStageRegs entries per work-item; the rest are copied by store()), store()
// writes them to LDS and returns the view (visible after the caller's
// barrier).  Without STAGE the view points at global memory.
constexpr int kStageRegs = 8;

template <typename V, bool STAGE>
struct BlockStage {
    V lu[kStageRegs];
    int col[kStageRegs];
    int rp = 0, dg = 0, rpe = 0;
    __device__ inline void load(const SolvArgs<V> &a, int blk) {
        if (!STAGE) return;
        const size_t e0 = (size_t)blk * a.mx;
#pragma unroll
        for (int k = 0; k < kStageRegs; k++) {
            const int e = threadIdx.x + k * kThreads;
            if (e < a.mx) {
                lu[k] = a.lu[e0 + e];
                col[k] = a.pcol[e0 + e];
            }
        }
        const size_t r0 = (size_t)blk * (a.brows + 1);
        if ((int)threadIdx.x < a.brows) {
            rp = a.prp[r0 + threadIdx.x];
            dg = a.pdg[(size_t)blk * a.brows + threadIdx.x];
        }
        if (threadIdx.x == 0) rpe = a.prp[r0 + a.brows];
    }
    __device__ inline BlockView<V> store(const SolvArgs<V> &a, int blk, unsigned char *lds) const {
        const size_t e0 = (size_t)blk * a.mx, r0 = (size_t)blk * (a.brows + 1), d0 = (size_t)blk * a.brows;
        if (!STAGE) return BlockView<V>{a.prp + r0, a.pdg + d0, a.pcol + e0, a.lu + e0};
        V *llu = reinterpret_cast<V *>(lds);
        int *lrp = reinterpret_cast<int *>(llu + a.mx);
        int *ldg = lrp + a.brows + 1;
        int *lcol = ldg + a.brows;
#pragma unroll
        for (int k = 0; k < kStageRegs; k++) {
            const int e = threadIdx.x + k * kThreads;
            if (e < a.mx) {
                llu[e] = lu[k];
                lcol[e] = col[k];
            }
        }
        for (int e = threadIdx.x + kStageRegs * kThreads; e < a.mx; e += kThreads) {
            llu[e] = a.lu[e0 + e];
            lcol[e] = a.pcol[e0 + e];
        }
        if ((int)threadIdx.x < a.brows) {
            lrp[threadIdx.x] = rp;
            ldg[threadIdx.x] = dg;
        }
        for (int li = threadIdx.x + kThreads; li < a.brows; li += kThreads) {
            lrp[li] = a.prp[r0 + li];
            ldg[li] = a.pdg[d0 + li];
        }
        if (threadIdx.x == 0) lrp[a.brows] = rpe;
        return BlockView<V>{lrp, ldg, lcol, llu};
    }
};

// (LU)^-1 y for the local rows of block blk; y and w are LDS buffers of
// the block's rows, y holds the right-hand side (visible: the caller's
// barrier).  Returns the buffer that holds the result.
//   Exact: level-scheduled forward and backward substitution in lusol's
//   order, in place in y (rows of a level written by any work-item, a
//   barrier after each level).
//   a.sweeps = k > 0: each triangular factor applied by k Jacobi sweeps,
//   x <- y - L x from x = y, then x <- D^-1 (f - U x) from x = D^-1 f,
//   alternating between y and w (one barrier per sweep) with the
//   right-hand sides in registers; the result does not depend on the
//   order in which wavefronts run, and each work-item reads back only the
//   rows it wrote itself.
template <typename V>
__device__ inline V *block_lusol(const SolvArgs<V> &a, const BlockView<V> &B, int blk, int b0, int nb, V *y,
                                 V *w) {
    if (a.sweeps > 0) {
        constexpr int RPT = kMaxBlockRows / kThreads;   // rows per work-item
        V bl[RPT];
        V *src = y, *dst = w;
        int nr = 0;
        for (int li = threadIdx.x; li < nb; li += blockDim.x, nr++) bl[nr] = y[li];
        for (int sw = 1; sw < a.sweeps; sw++) {
            nr = 0;
            for (int li = threadIdx.x; li < nb; li += blockDim.x, nr++) {
                V acc = bl[nr];
                for (int q = B.rp[li]; q < B.dg[li]; q++) acc = acc - B.lu[q] * src[B.col[q]];
                dst[li] = acc;
            }
            __syncthreads();
            V *t = src; src = dst; dst = t;
        }
        // backward, right-hand side f = the forward result (own rows)
        nr = 0;
        for (int li = threadIdx.x; li < nb; li += blockDim.x, nr++) {
            bl[nr] = src[li];
            dst[li] = B.lu[B.dg[li]] * bl[nr];
        }
        __syncthreads();
        { V *t = src; src = dst; dst = t; }
        for (int sw = 1; sw < a.sweeps; sw++) {
            nr = 0;
            for (int li = threadIdx.x; li < nb; li += blockDim.x, nr++) {
                V acc = bl[nr];
                for (int q = B.dg[li] + 1; q < B.rp[li + 1]; q++) acc = acc - B.lu[q] * src[B.col[q]];
                dst[li] = B.lu[B.dg[li]] * acc;
            }
            __syncthreads();
            V *t = src; src = dst; dst = t;
        }
        return src;
    }
    const int stride = a.brows + 1;
    const int *lp = a.flptr + (size_t)blk * stride;
    const int nf = a.fnlev[blk];
    for (int lv = 0; lv < nf; lv++) {
        for (int t = lp[lv] + threadIdx.x; t < lp[lv + 1]; t += blockDim.x) {
            const int li = a.frows[b0 + t];
            V acc = y[li];
            for (int q = B.rp[li]; q < B.dg[li]; q++) acc = acc - B.lu[q] * y[B.col[q]];
            y[li] = acc;
        }
        __syncthreads();
    }
    const int *bp = a.blptr + (size_t)blk * stride;
    const int nbl = a.bnlev[blk];
    for (int lv = 0; lv < nbl; lv++) {
        for (int t = bp[lv] + threadIdx.x; t < bp[lv + 1]; t += blockDim.x) {
            const int li = a.brows_l[b0 + t];
            V acc = y[li];
            for (int q = B.dg[li] + 1; q < B.rp[li + 1]; q++) acc = acc - B.lu[q] * y[B.col[q]];
            y[li] = B.lu[B.dg[li]] * acc;
        }
        __syncthreads();
    }
    return y;
}

// One row of the block factors held in registers by the work-item that
// owns the row (blocks of at most one row per work-item): its inverted
// pivot (slot 0), L entries (slots 1..nl) and U entries (slots nl+1..cnt-1)
// with their local columns, loaded at kernel start from the slot-major copy,
// so that the Jacobi sweeps read only the neighbour values from LDS.  Same
// operations in the same order as block_lusol's sweeps.
constexpr int kFacWidth = kEllWidth;

template <typename V>
struct FacRow {
    V v[kFacWidth];
    int c[kFacWidth];
    int nl = 0, cnt = 0;
    __device__ __attribute__((always_inline)) void load(const SolvArgs<V> &a, int i, bool valid) {
        if (!valid) return;
        const int meta = a.fs_meta[i];
        nl = meta & 255;
        cnt = meta >> 8;
#pragma unroll
        for (int k = 0; k < kFacWidth; k++) {
            v[k] = a.fs_val[(size_t)k * a.n + i];
            c[k] = a.fs_col[(size_t)k * a.n + i];
        }
    }
    // (LU)^-1 y for this row, a.sweeps > 0 Jacobi sweeps per factor
    // alternating between y and w; returns the buffer holding the result
    __device__ __attribute__((always_inline)) V *lusol(int sweeps, bool mine, int li, V *y, V *w) const {
        V *src = y, *dst = w;
        V bl = mine ? y[li] : (V)0;
        for (int sw = 1; sw < sweeps; sw++) {
            if (mine) {
                V acc = bl;
#pragma unroll
                for (int k = 1; k < kFacWidth; k++)
                    if (k <= nl) acc = acc - v[k] * src[c[k]];
                dst[li] = acc;
            }
            __syncthreads();
            V *t = src; src = dst; dst = t;
        }
        const V dinv = v[0];
        if (mine) {
            bl = src[li];
            dst[li] = dinv * bl;
        }
        __syncthreads();
        { V *t = src; src = dst; dst = t; }
        for (int sw = 1; sw < sweeps; sw++) {
            if (mine) {
                V acc = bl;
#pragma unroll
                for (int k = 0; k < kFacWidth; k++)
                    if (k > nl && k < cnt) acc = acc - v[k] * src[c[k]];
                dst[li] = dinv * acc;
            }
            __syncthreads();
            V *t = src; src = dst; dst = t;
        }
        return src;
    }
};

// ILU(0) of each diagonal block (the pivot is stored inverted).  Row i,
// entries in column order:
//   l_ij = (a_ij - sum_{k<j} l_ik u_kj) * (1/u_jj)   (j < i)
//   u_ij =  a_ij - sum_{k<i} l_ik u_kj               (j >= i)
// with the (k ascending) update pairs of each entry.  Entries are computed
// level by level of the entry dependence graph (an entry's level is one
// more than that of every entry it reads), one work-item per entry, a
// barrier per level; each entry is the same expression in the same order
// as in a row-by-row factorisation.  Four wavefronts per block: staging
// and the wide early levels spread over 256 work-items (measured on the
// 50k-node lagoon, 256-row blocks: 56 us against 60 us with two
// wavefronts and 75 us with one).
// Per entry, in level order, one descriptor {entry, first pair, end of
// pairs, pivot entry (>= 0) / -1 (U) / -2 - row (diagonal)} and the pairs'
// two entries packed together, all relative to the block; with stage the
// descriptors and pairs are copied to LDS.
constexpr int kFactorThreads = 256;
// Warm start (sweeps > 0): the factors of the previous solve, whose matrix
// differs little, refined by synchronous sweeps over all entries at once
// (Chow and Patel, fine-grained parallel ILU, SIAM J. Sci. Comput. 37,
// 2015): every entry is recomputed from the current values of the others,
// in the same operations as above; after k sweeps every entry within
// dependence depth k is exact.  kWarmEntries bounds the entries a
// work-item keeps the matrix values of (larger blocks run cold).
constexpr int kWarmEntries = 8;

template <typename V>
__global__ __launch_bounds__(kFactorThreads) void k_bilu_factor(SolvArgs<V> a, int stage, int sweeps) {
    extern __shared__ double smem[];
    V *val = reinterpret_cast<V *>(smem);
    const int blk = blockIdx.x, b0 = blk * a.brows;
    const int nb = min(a.n, b0 + a.brows) - b0;
    const int q0 = a.brp[b0], q1 = a.brp[b0 + nb], ne = q1 - q0;
    const int p0 = a.cp_ptr[q0], np = a.cp_ptr[q1] - p0;
    const int nl = a.enlev[blk];
    const int *elp = a.elp + (size_t)blk * (a.edmax + 1);   // level starts
    const int4 *desc = a.fdesc + (size_t)blk * a.mx;         // entries by level
    const int2 *pairs = a.fpair + p0;
    // the lists start 16-byte aligned after the values (two value buffers
    // when warm)
    const size_t vreg = (a.mx * sizeof(V) + 15) & ~(size_t)15;
    unsigned char *lfree = reinterpret_cast<unsigned char *>(smem) + vreg * (sweeps > 0 ? 2 : 1);
    if (stage) {
        int4 *l_d = reinterpret_cast<int4 *>(lfree);
        int2 *l_p = reinterpret_cast<int2 *>(l_d + a.mx);
        int *l_e = reinterpret_cast<int *>(l_p + a.mxp);
        for (int t = threadIdx.x; t < ne; t += blockDim.x) l_d[t] = desc[t];
        for (int p = threadIdx.x; p < np; p += blockDim.x) l_p[p] = pairs[p];
        for (int t = threadIdx.x; t <= nl; t += blockDim.x) l_e[t] = elp[t];
        desc = l_d; pairs = l_p; elp = l_e;
    }
    if (sweeps > 0) {
        // this work-item's entries: descriptors and matrix values in registers
        V av[kWarmEntries];
        int4 dd[kWarmEntries];
        const V *prev = a.lu + (size_t)blk * a.mx;
        V *nxt = reinterpret_cast<V *>(reinterpret_cast<unsigned char *>(smem) + vreg);
#pragma unroll
        for (int k = 0; k < kWarmEntries; k++) {
            const int t = threadIdx.x + k * blockDim.x;
            if (t < ne) {
                dd[k] = desc[t];
                av[k] = (V)a.aval[a.bsrc[q0 + dd[k].x]];
            }
        }
        for (int e = threadIdx.x; e < ne; e += blockDim.x) val[e] = prev[e];
        __syncthreads();
        V *cur = val;
        for (int sw = 0; sw < sweeps; sw++) {
#pragma unroll
            for (int k = 0; k < kWarmEntries; k++) {
                const int t = threadIdx.x + k * blockDim.x;
                if (t >= ne) break;
                const int4 d = dd[k];
                V acc = av[k];
                for (int p = d.y; p < d.z; p++) {
                    const int2 pr = pairs[p];
                    acc = acc - cur[pr.x] * cur[pr.y];
                }
                V out;
                if (d.w >= 0) out = acc * cur[d.w];
                else if (d.w == -1) out = acc;
                else {
                    if (acc == (V)0) atomicCAS(a.ctrl + CT_FACTOR, 0, 1 + b0 + (-2 - d.w));
                    out = (V)1 / acc;
                }
                nxt[d.x] = out;
            }
            __syncthreads();
            V *tmp = cur; cur = nxt; nxt = tmp;
        }
        if (cur != val) {
            for (int e = threadIdx.x; e < ne; e += blockDim.x) val[e] = cur[e];
            __syncthreads();
        }
    } else {
        for (int q = q0 + threadIdx.x; q < q1; q += blockDim.x) val[q - q0] = (V)a.aval[a.bsrc[q]];
        __syncthreads();
    }
    for (int lv = 0; lv < (sweeps > 0 ? 0 : nl); lv++) {
        for (int t = elp[lv] + threadIdx.x; t < elp[lv + 1]; t += blockDim.x) {
            const int4 d = desc[t];
            V acc = val[d.x];
            for (int p = d.y; p < d.z; p++) {
                const int2 pr = pairs[p];
                acc = acc - val[pr.x] * val[pr.y];
            }
            if (d.w >= 0) {
                val[d.x] = acc * val[d.w];
            } else if (d.w == -1) {
                val[d.x] = acc;
            } else {
                if (acc == (V)0) atomicCAS(a.ctrl + CT_FACTOR, 0, 1 + b0 + (-2 - d.w));
                val[d.x] = (V)1 / acc;
            }
        }
        __syncthreads();
    }
    for (int e = threadIdx.x; e < ne; e += blockDim.x) a.lu[(size_t)blk * a.mx + e] = val[e];
    if (a.fs_val) {
        const int *rp = a.prp + (size_t)blk * (a.brows + 1);
        for (int li = threadIdx.x; li < nb; li += blockDim.x)
            for (int e = rp[li]; e < rp[li + 1]; e++) {
                const int pos = e - rp[li], nlr = a.fs_meta[b0 + li] & 255;
                const int slot = pos < nlr ? pos + 1 : (pos == nlr ? 0 : pos);
                a.fs_val[(size_t)slot * a.n + b0 + li] = val[e];
            }
    }
}

// The operands of one matrix row -- values and NV gathered vectors at the
// row's columns, and the vectors at the row itself -- held in registers.
// The row is read from its ELL slots (padding: column i, value 0), so every
// load is issued at kernel start without waiting for the row's start; the
// entries beyond kEllWidth are read from the CSR arrays where they are used.
template <typename V, int NV>
struct RowOps {
    V a[kEllWidth];
    V x[NV][kEllWidth];
    V own[NV];
    int len = 0;
    template <typename... P>
    __device__ inline void load(const SolvArgs<V> &s, int i, bool valid, P... vec) {
        const V *vs[NV] = {vec...};
        if (!valid) return;
        len = s.ell_len[i];
        int j[kEllWidth];
#pragma unroll
        for (int k = 0; k < kEllWidth; k++) {
            const size_t slot = (size_t)k * s.n + i;
            j[k] = s.ell_col[slot];
            a[k] = s.ell_a[slot];
        }
#pragma unroll
        for (int v = 0; v < NV; v++) own[v] = vs[v][i];
#pragma unroll
        for (int k = 0; k < kEllWidth; k++)
#pragma unroll
            for (int v = 0; v < NV; v++) x[v][k] = vs[v][j[k]];
    }
    // CSR range of the entries beyond the ELL slots (empty for most rows)
    __device__ inline void tail(const SolvArgs<V> &s, int i, int &q0, int &q1) const {
        q0 = q1 = 0;
        if (len > kEllWidth) {
            q1 = s.rowptr[i + 1];
            q0 = q1 - len + kEllWidth;
        }
    }
};

// r0 = M^-1 (b - A x0) (x0 = 0 when zero_x0), dx = 0; (r0,r0) into both
// slots of region C
template <typename V, bool STAGE>
__global__ __launch_bounds__(kThreads) void k_bicg_init(SolvArgs<V> a, int zero_x0) {
    if (a.ctrl[CT_REFCONV]) return;
    extern __shared__ double smem[];
    V *y = reinterpret_cast<V *>(smem), *w = y + a.brows;
    const int blk = xcd_block(blockIdx.x, gridDim.x), b0 = blk * a.brows;
    const int nb = min(a.n, b0 + a.brows) - b0;
    BlockStage<V, STAGE> st;
    st.load(a, blk);
    const BlockView<V> B = st.store(a, blk, reinterpret_cast<unsigned char *>(w + a.brows));
    for (int li = threadIdx.x; li < nb; li += blockDim.x) {
        const int i = b0 + li;
        V t = 0;
        if (!zero_x0)
            for (int q = a.rowptr[i]; q < a.rowptr[i + 1]; q++) t = t + a.A[q] * a.x0[a.cols[q]];
        y[li] = a.b[i] - t;
    }
    __syncthreads();
    const V *z = block_lusol(a, B, blk, b0, nb, y, w);
    double loc[2] = {0., 0.};
    for (int li = threadIdx.x; li < nb; li += blockDim.x) {
        const int i = b0 + li;
        const V r = z[li];
        a.r0[i] = r;
        a.dx[i] = 0;
        loc[0] += (double)r * (double)r;
    }
    loc[1] = loc[0];
    store_partials<2>(loc, a.part, PR_C, a.nparts, blk);
    if (blk == 0 && threadIdx.x == 0) {
        a.ctrl[CT_STOPIT] = 0; a.ctrl[CT_CONV] = 0; a.ctrl[CT_FAIL] = 0; a.ctrl[CT_MAXIT] = 0; a.ctrl[CT_NMV] = 1;
    }
}

// Iteration it, first kernel.  For it > 0 it completes iteration it-1 from
// the partial sums of k_bicg_b: omega = (t,s)/(t,t), x += alpha p + omega s,
// r = s - omega t, with |r|^2 = (s,s) - 2 omega (t,s) + omega^2 (t,t) and
// rho = (r0,s) - omega (r0,t); then the stopping test (stopbis), beta, the
// new direction p = r + beta (p - omega v) (neighbour rows formed on the
// fly), v = M^-1 A p and the partial (r0,v).  A stop decided by this kernel
// (convergence, iteration limit) is stamped with it+1 in CT_STOPIT, so that
// its own workgroups still finish the update of x; breakdowns are found by
// every workgroup before any update.  With blocks of at most one row per
// work-item, every load that depends on no other load (control words,
// scalars, the block's factors, the row's operands, the partial sums) is
// issued before the first test.
template <typename V, int MODE>
__global__ __launch_bounds__(kThreads) void k_bicg_a(SolvArgs<V> a, int it) {
    constexpr bool STAGE = MODE == 1;      // block factors staged in LDS (0: read from global memory)
    extern __shared__ double smem[];
    V *y = reinterpret_cast<V *>(smem), *w = y + a.brows;
    const int par = it & 1, pp = par ^ 1;
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    const int blk = xcd_block(blockIdx.x, gridDim.x), b0 = blk * a.brows;
    const int nb = min(a.n, b0 + a.brows) - b0;
    const V *pold = pp ? a.p1 : a.p0;     // direction of iteration it-1 (it = 0: unused)
    const V *vold = pp ? a.v1 : a.v0;
    V *pnew = par ? a.p1 : a.p0;
    V *vnew = par ? a.v1 : a.v0;
    const bool pre = a.brows <= kThreads;
    const int i = b0 + threadIdx.x;
    const bool mine = pre && (int)threadIdx.x < nb;
    const int nq = it == 0 ? 2 : 5, region = it == 0 ? PR_C : PR_B;
    // independent loads first
    const int stop = a.ctrl[CT_STOPIT], fail = a.ctrl[CT_FAIL], refconv = a.ctrl[CT_REFCONV];
    const double alpha_prev = a.scal[SC_ALPHA + pp], target_prev = a.scal[SC_TARGET], rho_prev = a.scal[SC_RHO + pp];
    constexpr bool regs = MODE == 2;       // factor rows in registers, no block staging
    BlockStage<V, STAGE> st;
    FacRow<V> fr;
    if constexpr (regs) fr.load(a, i, mine);
    else st.load(a, blk);
    Partials<5> ps;
    ps.load(a.part, region, a.nparts, a.nblk, nq);
    RowOps<V, 4> ro;
    V r0_i = 0, dx_i = 0;
    if (pre) {
        if (it == 0) ro.load(a, i, mine, a.r0, a.r0, a.r0, a.r0);
        else ro.load(a, i, mine, a.s, a.t, pold, vold);
        if (mine) {
            r0_i = a.r0[i];
            if (it > 0) dx_i = a.dx[i];
        }
    }
    if (fail || refconv || (stop != 0 && stop <= it)) return;
    const BlockView<V> B = regs ? BlockView<V>{} : st.store(a, blk, reinterpret_cast<unsigned char *>(w + a.brows));
    double q[5];   // it = 0: (r0,r0) (r0,r0); else (t,t) (t,s) (s,s) (r0,s) (r0,t)
    ps.total(a.part, region, a.nparts, a.nblk, nq, q);
    double rn, rho, omega = 0., alpha = 0.;
    if (it == 0) {
        rn = sqrt(q[0]);
        rho = q[1];
    } else {
        if (breakdown(q[0])) {
            if (lead) a.ctrl[CT_FAIL] = -3;
            return;
        }
        omega = q[1] / q[0];
        if (breakdown(omega)) {
            if (lead) a.ctrl[CT_FAIL] = -3;
            return;
        }
        alpha = alpha_prev;
        rn = sqrt(fmax(0., q[2] - 2. * omega * q[1] + omega * omega * q[0]));
        rho = q[3] - omega * q[4];
        const V va = (V)alpha, vo = (V)omega;
        if (pre) {
            if (mine) {
                a.dx[i] = dx_i + (va * ro.own[2] + vo * ro.own[0]);
                a.r[i] = ro.own[0] - vo * ro.own[1];
            }
        } else {
            for (int li = threadIdx.x; li < nb; li += blockDim.x) {
                const int k = b0 + li;
                a.dx[k] = a.dx[k] + (va * pold[k] + vo * a.s[k]);
                a.r[k] = a.s[k] - vo * a.t[k];
            }
        }
        if (lead) { a.scal[SC_OMEGA + pp] = omega; a.ctrl[CT_ITS] += 1; }
    }
    double target;
    if (it == 0) {
        target = a.rtol * rn + a.atol;
        if (lead) { a.scal[SC_TARGET] = target; a.scal[SC_NORM0] = rn; }
    } else {
        target = target_prev;
    }
    if (lead) { a.scal[SC_RNORM] = rn; a.ctrl[CT_NMV] = 1 + 2 * it; }
    if (!(rn > target)) {
        if (lead) { a.ctrl[CT_STOPIT] = it + 1; a.ctrl[CT_CONV] = 1; }
        return;
    }
    if (it > 0 && a.itermax > 0 && 1 + 2 * it >= a.itermax) {
        if (lead) { a.ctrl[CT_STOPIT] = it + 1; a.ctrl[CT_MAXIT] = 1; }
        return;
    }
    double beta = 0.;
    if (it > 0) {
        if (breakdown(rho)) {
            if (lead) a.ctrl[CT_FAIL] = -3;
            return;
        }
        beta = rho * alpha / (omega * rho_prev);
    }
    if (lead) a.scal[SC_RHO + par] = rho;
    const V vb = (V)beta, vo = (V)omega;
    if (pre) {
        if (mine) {
            V t = 0;
            int q0, q1;
            ro.tail(a, i, q0, q1);
            if (it == 0) {
#pragma unroll
                for (int k = 0; k < kEllWidth; k++)
                    if (k < ro.len) t = t + ro.a[k] * ro.x[0][k];
                for (int q = q0; q < q1; q++) t = t + a.A[q] * a.r0[a.cols[q]];
                pnew[i] = ro.own[0];
            } else {
#pragma unroll
                for (int k = 0; k < kEllWidth; k++) {
                    if (k < ro.len) {
                        const V rj = ro.x[0][k] - vo * ro.x[1][k];
                        const V pj = rj + vb * (ro.x[2][k] - vo * ro.x[3][k]);
                        t = t + ro.a[k] * pj;
                    }
                }
                for (int q = q0; q < q1; q++) {
                    const int j = a.cols[q];
                    const V rj = a.s[j] - vo * a.t[j];
                    const V pj = rj + vb * (pold[j] - vo * vold[j]);
                    t = t + a.A[q] * pj;
                }
                const V ri = ro.own[0] - vo * ro.own[1];
                pnew[i] = ri + vb * (ro.own[2] - vo * ro.own[3]);
            }
            y[threadIdx.x] = t;
        }
    } else if (it == 0) {
        for (int li = threadIdx.x; li < nb; li += blockDim.x) {
            const int k = b0 + li;
            V t = 0;
            for (int q = a.rowptr[k]; q < a.rowptr[k + 1]; q++) t = t + a.A[q] * a.r0[a.cols[q]];
            y[li] = t;
            pnew[k] = a.r0[k];
        }
    } else {
        for (int li = threadIdx.x; li < nb; li += blockDim.x) {
            const int k = b0 + li;
            V t = 0;
            for (int q = a.rowptr[k]; q < a.rowptr[k + 1]; q++) {
                const int j = a.cols[q];
                const V rj = a.s[j] - vo * a.t[j];
                const V pj = rj + vb * (pold[j] - vo * vold[j]);
                t = t + a.A[q] * pj;
            }
            y[li] = t;
            const V ri = a.s[k] - vo * a.t[k];
            pnew[k] = ri + vb * (pold[k] - vo * vold[k]);
        }
    }
    __syncthreads();
    const V *z = regs ? fr.lusol(a.sweeps, mine, threadIdx.x, y, w) : block_lusol(a, B, blk, b0, nb, y, w);
    double loc[1] = {0.};
    if (pre) {
        if (mine) {
            vnew[i] = z[threadIdx.x];
            loc[0] = (double)r0_i * (double)z[threadIdx.x];
        }
    } else {
        for (int li = threadIdx.x; li < nb; li += blockDim.x) {
            const int k = b0 + li;
            vnew[k] = z[li];
            loc[0] += (double)a.r0[k] * (double)z[li];
        }
    }
    store_partials<1>(loc, a.part, PR_A, a.nparts, blk);
}

// Iteration it, second kernel: alpha = rho / (r0,v); s = r - alpha v
// (fused); t = M^-1 A s; partials (t,t) (t,s) (s,s) (r0,s) (r0,t).
// Independent loads are issued first, as in k_bicg_a.
template <typename V, int MODE>
__global__ __launch_bounds__(kThreads) void k_bicg_b(SolvArgs<V> a, int it) {
    constexpr bool STAGE = MODE == 1;      // block factors staged in LDS (0: read from global memory)
    extern __shared__ double smem[];
    V *y = reinterpret_cast<V *>(smem), *w = y + a.brows;
    const int par = it & 1;
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    const V *vnew = par ? a.v1 : a.v0;
    const V *rv = it == 0 ? a.r0 : a.r;
    const int blk = xcd_block(blockIdx.x, gridDim.x), b0 = blk * a.brows;
    const int nb = min(a.n, b0 + a.brows) - b0;
    const bool pre = a.brows <= kThreads;
    const int i = b0 + threadIdx.x;
    const bool mine = pre && (int)threadIdx.x < nb;
    const int stop = a.ctrl[CT_STOPIT], fail = a.ctrl[CT_FAIL], refconv = a.ctrl[CT_REFCONV];
    const double rho = a.scal[SC_RHO + par];
    constexpr bool regs = MODE == 2;
    BlockStage<V, STAGE> st;
    FacRow<V> fr;
    if constexpr (regs) fr.load(a, i, mine);
    else st.load(a, blk);
    Partials<1> ps;
    ps.load(a.part, PR_A, a.nparts, a.nblk, 1);
    RowOps<V, 2> ro;
    V r0_i = 0;
    if (pre) {
        ro.load(a, i, mine, rv, vnew);
        if (mine) r0_i = a.r0[i];
    }
    if (stop || fail || refconv) return;
    const BlockView<V> B = regs ? BlockView<V>{} : st.store(a, blk, reinterpret_cast<unsigned char *>(w + a.brows));
    double d[1];
    ps.total(a.part, PR_A, a.nparts, a.nblk, 1, d);
    if (breakdown(d[0])) {
        if (lead) a.ctrl[CT_FAIL] = -3;
        return;
    }
    const double alpha = rho / d[0];
    if (lead) a.scal[SC_ALPHA + par] = alpha;
    const V va = (V)alpha;
    V s_i = 0;
    if (pre) {
        if (mine) {
            V t = 0;
            int q0, q1;
            ro.tail(a, i, q0, q1);
#pragma unroll
            for (int k = 0; k < kEllWidth; k++)
                if (k < ro.len) t = t + ro.a[k] * (ro.x[0][k] - va * ro.x[1][k]);
            for (int q = q0; q < q1; q++) {
                const int j = a.cols[q];
                t = t + a.A[q] * (rv[j] - va * vnew[j]);
            }
            y[threadIdx.x] = t;
            s_i = ro.own[0] - va * ro.own[1];
            a.s[i] = s_i;
        }
    } else {
        for (int li = threadIdx.x; li < nb; li += blockDim.x) {
            const int k = b0 + li;
            V t = 0;
            for (int q = a.rowptr[k]; q < a.rowptr[k + 1]; q++) {
                const int j = a.cols[q];
                const V sj = rv[j] - va * vnew[j];
                t = t + a.A[q] * sj;
            }
            y[li] = t;
            a.s[k] = rv[k] - va * vnew[k];
        }
    }
    __syncthreads();
    const V *z = regs ? fr.lusol(a.sweeps, mine, threadIdx.x, y, w) : block_lusol(a, B, blk, b0, nb, y, w);
    double loc[5] = {0., 0., 0., 0., 0.};
    if (pre) {
        if (mine) {
            const double tv = (double)z[threadIdx.x], sv = (double)s_i, r0 = (double)r0_i;
            a.t[i] = z[threadIdx.x];
            loc[0] = tv * tv;
            loc[1] = tv * sv;
            loc[2] = sv * sv;
            loc[3] = r0 * sv;
            loc[4] = r0 * tv;
        }
    } else {
        for (int li = threadIdx.x; li < nb; li += blockDim.x) {
            const int k = b0 + li;
            const double tv = (double)z[li], sv = (double)a.s[k], r0 = (double)a.r0[k];
            a.t[k] = z[li];
            loc[0] += tv * tv;
            loc[1] += tv * sv;
            loc[2] += sv * sv;
            loc[3] += r0 * sv;
            loc[4] += r0 * tv;
        }
    }
    store_partials<5>(loc, a.part, PR_B, a.nparts, blk);
}

// x = x0 + dx (tidycg) into the FP64 solution; with add, x += dx unless the
// refinement has converged
template <typename V>
__global__ void k_bicg_finish(const V *x0, const V *dx, double *x, int n, int add, const int *ctrl) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || (add && ctrl[CT_REFCONV]) || (!add && !ctrl[CT_CONV])) return;
    if (add) x[i] = x[i] + (double)dx[i];
    else x[i] = (double)(x0[i] + dx[i]);
}

// the solve's copies of the assembled matrix: values in the ELL slots of
// the iteration kernels (precision V), and for the FP32 path the FP32 CSR
template <typename V>
__global__ void k_prep_matrix(const double *aval, const int *csr_ell, V *ell_a, float *a32, int nnz) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nnz) return;
    const double v = aval[q];
    if (a32) a32[q] = (float)v;
    const int slot = csr_ell[q];
    if (slot >= 0) ell_a[slot] = (V)v;
}

// initial guess x = zn; clears the factorisation flag of this solve
template <typename S>
__global__ void k_guess(const S *zn, double *x, int n, int *ctrl) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) { ctrl[CT_FACTOR] = 0; ctrl[CT_ITS] = 0; ctrl[CT_NREF] = 0; ctrl[CT_REFCONV] = 0; }
    if (i < n) x[i] = (double)zn[i];
}

// FP32 path: initial guess extrapolated linearly in time from the levels at
// the start of this step (zn) and of the previous one (zprev), x = 2 zn -
// zprev (zn alone without a history); zprev <- zn.  The refinement's target
// is absolute (a fraction of |b|), so a closer start saves rounds; the
// solution it converges to does not depend on the start.
__global__ void k_guess_extrap(const float *zn, float *zprev, double *x, int n, int *ctrl, int have_prev) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) { ctrl[CT_FACTOR] = 0; ctrl[CT_ITS] = 0; ctrl[CT_NREF] = 0; ctrl[CT_REFCONV] = 0; }
    if (i >= n) return;
    const double z = (double)zn[i];
    x[i] = have_prev ? 2. * z - (double)zprev[i] : z;
    zprev[i] = zn[i];
}

// FP64 residual r64 = b - A x with partial |r64|^2 and |b|^2 (refinement)
__global__ __launch_bounds__(kThreads) void k_resid64(const int *rowptr, const int *cols, const double *A,
                                                       const double *b, const double *x, double *r64, float *r32,
                                                       int n, double *part, int nparts, const int *ctrl) {
    if (ctrl[CT_REFCONV]) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double loc[2] = {0., 0.};
    if (i < n) {
        double t = 0.;
        for (int q = rowptr[i]; q < rowptr[i + 1]; q++) t = t + A[q] * x[cols[q]];
        const double r = b[i] - t;
        r64[i] = r;
        r32[i] = (float)r;
        loc[0] = r * r;
        loc[1] = b[i] * b[i];
    }
    store_partials<2>(loc, part, PR_R, nparts, blockIdx.x);
}

__global__ __launch_bounds__(kThreads) void k_resid_check(const double *part, int nparts, int n, double *scal,
                                                           int *ctrl, double tol) {
    if (ctrl[CT_REFCONV]) return;
    double rs[2];
    sum_partials<2>(part, PR_R, nparts, n, rs);
    if (threadIdx.x == 0) {
        // iterations of the inner solve just finished (round nref - 1),
        // negative when it stopped without converging (cut short)
        const int r = ctrl[CT_NREF];
        if (r >= 1 && r <= CT_NRITS) {
            const int its = (ctrl[CT_NMV] - 1) / 2;
            ctrl[CT_RITS + r - 1] = ctrl[CT_CONV] ? its : -its;
        }
        scal[SC_RES64] = sqrt(rs[0]);
        scal[SC_BNORM] = sqrt(rs[1]);
        const int conv = (sqrt(rs[0]) <= tol * sqrt(rs[1])) ? 1 : 0;
        ctrl[CT_REFCONV] = conv;
        if (!conv) ctrl[CT_NREF] += 1;
    }
}

// ---------------------------------------------------------------- host driver
// one read-back: the solver's control words (into h_pinned[16..]) and the
// device error word and counter d_err[0..1] (into h_pinned[0..1])
static int read_ctrl(Model &M, int *out, int n) {
    HIPCHK(M, hipMemcpyAsync(M.h_pinned + 16, M.sv.ctrl, sizeof(int) * n, hipMemcpyDeviceToHost, M.stream));
    HIPCHK(M, hipMemcpyAsync(M.h_pinned, M.d_err, 2 * sizeof(int), hipMemcpyDeviceToHost, M.stream));
    HIPCHK(M, hipStreamSynchronize(M.stream));
    for (int i = 0; i < n; i++) out[i] = M.h_pinned[16 + i];
    return 0;
}

static constexpr size_t kLdsLimit = 64 * 1024;
static constexpr size_t kLdsStatic = 3 * kThreads * sizeof(double);   // reduction buffers

// dynamic LDS of the kernels that apply the preconditioner: with the block
// staged, or y and w only (blocks too large to stage read factors from
// global memory)
template <typename V>
static bool stage_fits(const Model &M) {
    return staged_lds<V>(M.sv.brows, M.sv.max_blk_nnz) + kLdsStatic <= kLdsLimit;
}

template <typename V>
static size_t apply_lds(const Model &M) {
    return stage_fits<V>(M) ? staged_lds<V>(M.sv.brows, M.sv.max_blk_nnz) : 2 * (size_t)M.sv.brows * sizeof(V);
}

// block factors of the current matrix (exact block ILU(0)), every solve: the
// wet/dry state changes rows of the zeta matrix from one step to the next
// (dry-element rows, level boundaries), and factors of an earlier matrix
// then no longer precondition it.
// LDS of the factorisation: the block's values, and when they fit the
// descriptors, pairs and level starts
// sweeps of a warm start (four reach the iteration counts of exact factors
// on the 50k-node lagoon, two do not), and exact factorisations at least
// every kWarmRefresh solves
constexpr int kWarmSweeps = 4;
constexpr int kWarmRefresh = 16;

template <typename V>
static int factor(Model &M, SolvArgs<V> &a) {
    const int sweeps = (kWarmSweeps > 0 && M.sv.have_lu && M.sv.max_blk_nnz <= kWarmEntries * kFactorThreads)
                           ? kWarmSweeps : 0;
    M.sv.last_warm = sweeps > 0;
    if (sweeps == 0) M.sv.solves_since_exact = 0;
    const size_t vals = (((size_t)M.sv.max_blk_nnz * sizeof(V) + 15) & ~(size_t)15) * (sweeps > 0 ? 2 : 1);
    const size_t lists = (size_t)M.sv.max_blk_nnz * sizeof(int4) + (size_t)M.sv.max_blk_pairs * sizeof(int2) +
                         (size_t)(M.sv.max_elev + 1) * sizeof(int);
    const int stage = vals + lists <= kLdsLimit ? 1 : 0;
    launch_timed(M, T_FACTOR, k_bilu_factor<V>, dim3(M.sv.nblk), dim3(kFactorThreads), stage ? vals + lists : vals,
                 M.stream, a, stage, sweeps);
    HIPCHK(M, hipGetLastError());
    M.sv.have_lu = true;
    return 0;
}

// One iteration is two kernels (k_bicg_a, k_bicg_b).  Iterations are
// enqueued in chunks between host checks of the device flags;
// the first chunk is the iteration count of the previous solve of the same
// kind (`slot`), so a converging solve usually needs a single round trip.
// Kernels enqueued after convergence return immediately.  `tail` enqueues
// the work that follows a converged solve, switched on by its flag, before
// each check.
// MODE of the iteration kernels: 0 block factors read from global memory,
// 1 staged in LDS, 2 factor rows in registers (the init kernel stages)
template <typename V, int MODE>
static int bicgstab_t(Model &M, SolvArgs<V> &a, int *iters, int slot, int zero_x0,
                      const std::function<int()> &tail) {
    const int g = a.nblk;
    const size_t lds = apply_lds<V>(M);
    int it = 0, c[CT_WORDS] = {};
    const int maxit = M.p.itermax > 0 ? (M.p.itermax + 1) / 2 + 1 : 100000;
    int chunk = std::max(2, M.sv.last_its[slot] + 1);
    for (;;) {
        // the chunk's iterations (after the initialisation, for the first)
        // as one replayed launch sequence
        const int it0 = it;
        auto launches = [&]() -> int {
            if (it0 == 0)
                hipLaunchKernelGGL((k_bicg_init<V, MODE != 0>), dim3(g), dim3(kThreads), lds, M.stream, a, zero_x0);
            for (int q = 0; q < chunk; q++) {
                launch_timed(M, T_BICG_A, k_bicg_a<V, MODE>, dim3(g), dim3(kThreads), lds, M.stream, a, it0 + q);
                launch_timed(M, T_BICG_B, k_bicg_b<V, MODE>, dim3(g), dim3(kThreads), lds, M.stream, a, it0 + q);
            }
            HIPCHK(M, hipGetLastError());
            return 0;
        };
        std::string key("bicgstab");
        key_add(key, a);
        key_add(key, MODE);
        key_add(key, lds);
        key_add(key, it0);
        key_add(key, chunk);
        key_add(key, zero_x0);
        if (graph_run(M, M.stream, key, launches)) return 1;
        it += chunk;
        if (tail && tail()) return 1;
        if (read_ctrl(M, c, CT_WORDS)) return 1;
        if (c[CT_CONV] || c[CT_FAIL] || c[CT_MAXIT] || it >= maxit) break;
        chunk = 2;
    }
    if (iters) *iters = (c[CT_NMV] - 1) / 2;
    M.sv.last_its[slot] = (c[CT_NMV] - 1) / 2;
    if (c[CT_FACTOR] != 0) {
        M.err = "zeta solver: zero pivot in the block factorisation (row " + std::to_string(c[CT_FACTOR]) + ")";
        return 2;
    }
    if (c[CT_FAIL] != 0 || c[CT_MAXIT] != 0) {
        M.err = "zeta solver: BiCGStab failed (code " + std::to_string(c[CT_FAIL] ? c[CT_FAIL] : -1) + ")";
        return 2;
    }
    if (!c[CT_CONV]) {
        M.err = "zeta solver: no convergence";
        return 2;
    }
    return 0;
}

template <typename V>
static int iter_mode(const Model &M, const SolvArgs<V> &a) {
    if (!stage_fits<V>(M)) return 0;
    return (a.fs_val && a.brows <= kThreads) ? 2 : 1;
}

template <typename V>
static int bicgstab(Model &M, SolvArgs<V> &a, int *iters, int slot, int zero_x0, const std::function<int()> &tail) {
    switch (iter_mode(M, a)) {
    case 2: return bicgstab_t<V, 2>(M, a, iters, slot, zero_x0, tail);
    case 1: return bicgstab_t<V, 1>(M, a, iters, slot, zero_x0, tail);
    default: return bicgstab_t<V, 0>(M, a, iters, slot, zero_x0, tail);
    }
}

// one inner FP32 solve of the refinement, enqueued without host checks:
// init from a zero correction, then `chunk` iterations (no-ops once
// converged or once the refinement has converged)
template <typename V, int MODE>
static int enqueue_inner_t(Model &M, SolvArgs<V> &a, int chunk) {
    const int g = a.nblk;
    const size_t lds = apply_lds<V>(M);
    auto launches = [&]() -> int {
        hipLaunchKernelGGL((k_bicg_init<V, MODE != 0>), dim3(g), dim3(kThreads), lds, M.stream, a, 1);
        for (int it = 0; it < chunk; it++) {
            launch_timed(M, T_BICG_A, k_bicg_a<V, MODE>, dim3(g), dim3(kThreads), lds, M.stream, a, it);
            launch_timed(M, T_BICG_B, k_bicg_b<V, MODE>, dim3(g), dim3(kThreads), lds, M.stream, a, it);
        }
        HIPCHK(M, hipGetLastError());
        return 0;
    };
    // one replayed launch sequence per argument set and chunk
    std::string key("inner");
    key_add(key, a);
    key_add(key, MODE);
    key_add(key, lds);
    key_add(key, chunk);
    return graph_run(M, M.stream, key, launches);
}

template <typename V>
static int enqueue_inner(Model &M, SolvArgs<V> &a, int chunk) {
    switch (iter_mode(M, a)) {
    case 2: return enqueue_inner_t<V, 2>(M, a, chunk);
    case 1: return enqueue_inner_t<V, 1>(M, a, chunk);
    default: return enqueue_inner_t<V, 0>(M, a, chunk);
    }
}

int solver_setup(Model &M) { (void)M; return 0; }

// the words of one zeta solve, for a second attempt at it
__global__ void k_ctrl_restart(int *ctrl) {
    if (threadIdx.x == 0) {
        ctrl[CT_FACTOR] = 0; ctrl[CT_ITS] = 0; ctrl[CT_NREF] = 0; ctrl[CT_REFCONV] = 0;
    }
}

// one zeta solve; fed: the row assembly already wrote its inputs (or a
// first attempt did)
static int solve_once(Model &M, int *iters, int *refinements, int wetdry_bnew, bool fed) {
    const int n = M.nkn;
    const int g1 = (n + kThreads - 1) / kThreads;
    int it_total = 0, nref = 0;
    // system_get + setweg(1) behind the solve, gated by its convergence flag
    auto wetdry = [&](const int *gate, const double *dx) -> int {
        if (wetdry_bnew < 0) return 0;
        timer_begin(M, T_WETDRY);
        const int rc = launch_system_get_setweg1(M, wetdry_bnew, gate, dx);
        timer_end(M);
        return rc;
    };
    if (M.prec == SHYHIP_FP64) {
        // initial guess = current level (system_solve passes znv); the Krylov
        // x0 vector is the FP64 solution array itself
        if (!fed) {
            hipLaunchKernelGGL(k_guess<double>, dim3(g1), dim3(kThreads), 0, M.stream, (const double *)M.s.zn, M.s.x,
                               n, M.sv.ctrl);
            hipLaunchKernelGGL(k_prep_matrix<double>, dim3((M.nnz + kThreads - 1) / kThreads), dim3(kThreads), 0,
                               M.stream, (const double *)M.s.aval, (const int *)M.m.csr_ell, (double *)M.sv.ell_a,
                               (float *)nullptr, M.nnz);
        }
        auto a = solv_args<double>(M, M.s.aval, M.s.rhs, M.s.x);
        if (factor<double>(M, a)) return 1;
        int its = 0;
        auto tail = [&]() -> int {
            // x = x0 + dx: inside system_get when it follows, else on its own
            if (wetdry_bnew >= 0) return wetdry(M.sv.ctrl + CT_CONV, a.dx);
            hipLaunchKernelGGL(k_bicg_finish<double>, dim3(g1), dim3(kThreads), 0, M.stream, (const double *)M.s.x,
                               (const double *)a.dx, M.s.x, n, 0, (const int *)M.sv.ctrl);
            HIPCHK(M, hipGetLastError());
            return 0;
        };
        const int rc = bicgstab<double>(M, a, &its, 0, 0, tail);
        it_total = its;
        if (rc) return rc;
    } else {
        // FP32 BiCGStab on the FP32-rounded matrix inside FP64 refinement:
        // x <- x + A32^-1 (b - A64 x) until |b - A64 x| <= refine_rtol |b|;
        // each inner solve starts from a zero correction and runs at most
        // `chunk` iterations (a partial correction still refines).  The
        // rounds of the previous solve are enqueued at once; the flags are
        // read after them.
        float *a32 = (float *)M.sv.a32;
        if (!fed) {
            hipLaunchKernelGGL(k_guess_extrap, dim3(g1), dim3(kThreads), 0, M.stream, (const float *)M.s.zn,
                               (float *)M.sv.zprev, M.s.x, n, M.sv.ctrl, M.sv.have_prev ? 1 : 0);
            M.sv.have_prev = true;
            hipLaunchKernelGGL(k_prep_matrix<float>, dim3((M.nnz + kThreads - 1) / kThreads), dim3(kThreads), 0,
                               M.stream, (const double *)M.s.aval, (const int *)M.m.csr_ell, (float *)M.sv.ell_a, a32,
                               M.nnz);
        }
        float *b32 = (float *)M.sv.b, *x0 = (float *)M.sv.x0;
        auto a = solv_args<float>(M, a32, b32, x0);
        if (factor<float>(M, a)) return 1;
        a.rtol = M.p.inner_rtol;
        a.atol = 0.0;
        const int maxref = 40;
        const int itmax_inner = M.p.itermax > 0 ? (M.p.itermax + 1) / 2 : 500;
        // iterations enqueued for round r: that round's count in the previous
        // solve plus one (kernels beyond convergence return at once but
        // still cost a launch); a round cut short by its chunk leaves a
        // partial correction and the next round continues from it
        auto chunk_of = [&](int r) {
            const int prev = r < CT_NRITS ? M.sv.round_its[r] : 0;
            const int c0 = prev > 0 ? prev + 1 : std::max(4, M.sv.last_its[1] + 3);
            return std::min(itmax_inner, std::max(2, c0));
        };
        auto resid = [&]() -> int {
            hipLaunchKernelGGL(k_resid64, dim3(g1), dim3(kThreads), 0, M.stream, M.m.rowptr, M.m.cols,
                               (const double *)M.s.aval, (const double *)M.s.rhs, (const double *)M.s.x, M.sv.r64,
                               b32, n, M.sv.part, M.sv.nparts, (const int *)M.sv.ctrl);
            hipLaunchKernelGGL(k_resid_check, dim3(1), dim3(kThreads), 0, M.stream, (const double *)M.sv.part,
                               M.sv.nparts, g1, M.sv.scal, M.sv.ctrl, M.p.refine_rtol);
            HIPCHK(M, hipGetLastError());
            return 0;
        };
        // boost: the rounds enqueued so far did not reach the target (their
        // chunks, sized from the previous solve, cut them short or the system
        // changed), so each later round gets twice the iterations of the last
        int plan = std::min((int)CT_NRITS, std::max(1, M.sv.last_nref)), enq = 0, c[CT_WORDS] = {}, boost = 1;
        for (;;) {
            for (int r = 0; r < plan && enq < maxref; r++, enq++) {
                if (resid() || enqueue_inner<float>(M, a, std::min(itmax_inner, chunk_of(enq) * boost))) return 1;
                hipLaunchKernelGGL(k_bicg_finish<float>, dim3(g1), dim3(kThreads), 0, M.stream, (const float *)x0,
                                   (const float *)a.dx, M.s.x, n, 1, (const int *)M.sv.ctrl);
            }
            if (resid()) return 1;
            if (wetdry(M.sv.ctrl + CT_REFCONV, nullptr)) return 1;
            if (read_ctrl(M, c, CT_WORDS)) return 1;
            if (c[CT_FACTOR] != 0) {
                        M.err = "zeta solver: zero pivot in the block factorisation (row " + std::to_string(c[CT_FACTOR]) + ")";
                return 2;
            }
            if (c[CT_REFCONV]) break;
            if (enq >= maxref) {
                double sc[16] = {};
                (void)hipMemcpyAsync(sc, M.sv.scal, sizeof sc, hipMemcpyDeviceToHost, M.stream);
                (void)hipStreamSynchronize(M.stream);
                std::string its;
                for (int r = 0; r < std::min((int)CT_NRITS, c[CT_NREF]); r++) its += " " + std::to_string(c[CT_RITS + r]);
                char buf[160];
                std::snprintf(buf, sizeof buf, " (|b - Ax| %.3e, |b| %.3e after %d rounds, %d iterations; first rounds:",
                              sc[SC_RES64], sc[SC_BNORM], c[CT_NREF], c[CT_ITS]);
                M.err = std::string("zeta solver: FP64 refinement did not reach the residual target") + buf + its + ")";
                return 2;
            }
            plan = 1;
            boost = std::min(2 * boost, 64);
        }
        nref = c[CT_NREF];
        it_total = c[CT_ITS];
        M.sv.last_nref = nref;
        M.sv.last_its[1] = nref > 0 ? (it_total + nref - 1) / nref : 0;
        for (int r = 0; r < CT_NRITS; r++) {
            const int its_r = r < nref ? c[CT_RITS + r] : 0;
            // a round cut short by its chunk gets more next time
            M.sv.round_its[r] = its_r < 0 ? 2 - its_r : its_r;
        }
    }
    if (iters) *iters = it_total;
    if (refinements) *refinements = nref;
    return 0;
}

// The block factors are refined from the previous solve's (warm) except on
// the first solve and every kWarmRefresh-th, which factorise exactly.  A
// solve that fails with refined factors (a matrix that changed too much for
// the sweeps: rows switched by wetting and drying) is done again with exact
// factors from where it stopped, so the warm start can cost iterations but
// never a solve.
int solver_solve(Model &M, int *iters, int *refinements, int wetdry_bnew) {
    // inputs already written by the row assembly of this step's system?
    const bool fed = M.sv.fed;
    M.sv.fed = false;
    if (++M.sv.solves_since_exact >= kWarmRefresh) M.sv.have_lu = false;
    const int rc = solve_once(M, iters, refinements, wetdry_bnew, fed);
    if (rc != 2 || !M.sv.last_warm) return rc;
    M.solver_retries++;
    M.sv.have_lu = false;
    hipLaunchKernelGGL(k_ctrl_restart, dim3(1), dim3(64), 0, M.stream, M.sv.ctrl);
    HIPCHK(M, hipGetLastError());
    M.err.clear();
    return solve_once(M, iters, refinements, wetdry_bnew, true);
}

}  // namespace shy
