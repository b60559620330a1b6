// shyhip_internal.h -- device-side data model of libshyhip.
//
// Device layout (chosen for coalescing on CDNA, not the reference's):
//   element 3D   a[l*nel + ie]        (layer-major, element fastest)
//   node 3D      a[l*nkn + k]
//   interfaces   w[l*nkn + k], l = 0..nlvdi
//   vertex       z[ii*nel + ie]
//   geometry     geo[q*nel + ie], q = b1 b2 b3 c1 c2 c3 aj(=area/12)
// Host push/pull transposes from/to the reference's (l,ie) layout.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include <cstdint>
#include <functional>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/shyhip.h"

namespace shy {

constexpr float kFlag = -9988765.0f;                 // mkonst flag (const.f90:119)
constexpr double kDrittl = (double)(1.0f / 3.0f);    // default-real 1./3. in double
constexpr int kThreads = 256;
constexpr int kMaxBlockRows = 1024;                  // block-ILU(0) rows per workgroup (LDS)
constexpr int kEllWidth = 10;                        // matrix entries per row held in ELL slots
constexpr int kIncWidth = 10;                        // incident elements per node held in slots
// conz3d work division: the element-layer list is built in chunks of
// kConzChunk consecutive elements; a node tile holds at most kNodeTileMax
// nodes (one work-item each) and kNodeItems node-layers (one pass of a
// workgroup)
constexpr int kConzChunk = 64;
constexpr int kNodeTileMax = kThreads;
constexpr int kNodeItems = kThreads;

// error codes written by kernels into Model::d_err[0]
enum DevErr { ERR_NONE = 0, ERR_SETNOD = 1, ERR_SETUVD = 3, ERR_SETDEPTH = 4, ERR_VOLUME = 5,
              ERR_CONZ = 7, ERR_PENTA = 8, ERR_WIND = 9, ERR_MFLUX = 10 };

struct Mesh {                 // device copies of init-time data
    int *nen3v = nullptr;     // [3*nel] element-major (ie*3+ii), 0-based
    double *geo = nullptr;    // [7*nel]
    double *ang = nullptr;    // [3*nel] vertex angles (ev 11-13)
    double *dis = nullptr;    // [3*nel] edge lengths (ev 17-19)
    int *ilhv = nullptr, *ilhkv = nullptr;
    void *hm3v = nullptr;     // S [3*nel] vertex-major
    void *hev = nullptr, *hlv = nullptr, *hldv = nullptr, *areakv = nullptr;
    int *ibnd = nullptr, *zbound = nullptr, *iextb = nullptr, *ibtyp = nullptr;
    void *fcorv = nullptr, *czv = nullptr;
    int *nk_ptr = nullptr;    // node -> incident (element<<2 | vertex), element order
    int *nk_el = nullptr;
    int *nkc_el = nullptr;    // same lists in colour order
    // the first kIncWidth entries of both lists at [s*nkn + k] (padded with
    // 0, a valid element), and the list lengths, so that a node's element
    // loads can all be issued at once
    int *inc_cnt = nullptr, *inc_el = nullptr, *inc_elc = nullptr;
    // positions of the colour-order lists (nk_ptr / nkc_el): the position of
    // vertex ii of element ie ([ii][ie]), and each position's element bottom
    // layer and 4 * area (the conz3d element terms are stored by position, so
    // a node reads its incident elements' terms contiguously)
    int *slot_of = nullptr, *slot_lev = nullptr;
    double *slot_aj4 = nullptr;
    int *rowptr = nullptr, *cols = nullptr, *ijp = nullptr;   // zeta CSR; ijp[ie*9+i*3+j]
    int *bnd_nodes = nullptr; int *bnd_id = nullptr; int nbnd_nodes = 0;
    // block-ILU(0)
    int *blk_rowptr = nullptr, *blk_cols = nullptr, *blk_src = nullptr, *blk_diag = nullptr;
    int *blk_cp_ptr = nullptr, *blk_cp_a = nullptr, *blk_cp_b = nullptr, *blk_jd = nullptr;
    // the same blocks at a fixed stride per block (entries blk*mx.., rows
    // blk*brows.., row starts blk*(brows+1)..; indices relative to the block),
    // so a kernel can fetch its block without first reading where it starts
    int *blk_pcol = nullptr, *blk_prp = nullptr, *blk_pdg = nullptr;
    // entry-level factorisation schedule: level starts [blk*(max_elev+1) + lv]
    // and the block's entries by level [blk*max_blk_nnz + t]; levels per block
    int *blk_elp = nullptr, *blk_elr = nullptr, *blk_enlev = nullptr;
    // the same entries as descriptors {entry, pairs begin, end, pivot} (int4,
    // [blk*max_blk_nnz + t]) and the update pairs packed {a, b} (int2,
    // indexed like cp_a), all relative to the block
    int *blk_fdesc = nullptr, *blk_fpair = nullptr;
    // zeta matrix rows sliced ELL-style: the first kEllWidth entries of row i
    // at [k*nkn + i] (columns padded with i), the row length, and for each
    // CSR entry its slot (or -1 beyond the width)
    int *ell_col = nullptr, *ell_len = nullptr, *csr_ell = nullptr;
    // block factor rows slot-major: local columns [k*nkn + i] (padding: the
    // row itself) and nL | cnt << 8 per row (absent if a block row is longer
    // than kEllWidth)
    int *fs_col = nullptr, *fs_meta = nullptr;
    int *fwd_rows = nullptr, *fwd_lptr = nullptr, *fwd_nlev = nullptr;
    int *bwd_rows = nullptr, *bwd_lptr = nullptr, *bwd_nlev = nullptr;
    // TVD upwind tables: per element 6 ordered pairs (ic,id); [p*nel + ie]
    int *tvd_el = nullptr;    // upwind element or -1
    int *tvd_kn = nullptr;    // [3][6][nel] its nodes
    void *tvd_w = nullptr;    // double [3][6][nel] interpolation weights (adjusted xi)
    int *tvd_lev = nullptr;   // [6][nel] bottom layer of the upwind element (0 if none)
    int *elem_iext = nullptr; // [nel] vertices of the element flagged external (iextb)
    // conz3d work division (set_basin): the element-layer list, ie << 5 |
    // (l - 1), in chunks of kConzChunk elements taken layer by layer; node
    // tiles as first nodes [n_ntile + 1], built for at most ntile_tn nodes
    int *conz_eitem = nullptr, *conz_ntile = nullptr;
    int n_eitem = 0, n_ntile = 0, ntile_tn = 0;
    // the node-layer list, k << 5 | (l - 1), in chunks of kConzChunk nodes
    // taken layer by layer (scalar_stability)
    int *conz_nitem = nullptr;
    int n_nitem = 0;
    // scalar open boundary (bndo_init): nodes, inward normals, inner neighbours
    int nbndo = 0;
    int *bndo_k = nullptr, *bndo_ptr = nullptr, *bndo_nb = nullptr;
    void *bndo_nx = nullptr, *bndo_ny = nullptr, *bndo_w = nullptr;   // S
    // the same nodes' static data, resolved at set-up so that a launch reads
    // it directly: open (type 1) flag, bottom layer, incident elements in
    // element order with their bottom layers and areas, and the bottom
    // layer of each inner neighbour
    int *bndo_open = nullptr, *bndo_lmax = nullptr, *bndo_eptr = nullptr, *bndo_el = nullptr, *bndo_elev = nullptr;
    int *bndo_nblev = nullptr;
    double *bndo_eaj = nullptr;
    // horizontal diffusion weights of the scalars (host_diffweight), built
    // for wdif_idtype when a transport first runs with rkpar > 0
    void *wdif = nullptr;
    int wdif_idtype = -1;
};

struct State {
    void *zn = nullptr, *zo = nullptr, *zen = nullptr, *zeo = nullptr;
    void *utn = nullptr, *vtn = nullptr, *uto = nullptr, *vto = nullptr;
    void *uln = nullptr, *vln = nullptr;
    void *un = nullptr, *vn = nullptr, *uo = nullptr, *vo = nullptr;
    void *wln = nullptr, *wlo = nullptr;
    void *hdkn = nullptr, *hdko = nullptr, *hden = nullptr, *hdeo = nullptr;
    int *iweg = nullptr, *iwet = nullptr, *inod = nullptr;
    void *rfric = nullptr, *ddx = nullptr, *ddy = nullptr;
    void *visv = nullptr, *difv = nullptr, *difhv = nullptr;
    void *taux = nullptr, *tauy = nullptr, *rzv = nullptr, *rqv = nullptr, *mflux = nullptr;
    void *rhov = nullptr;     // density anomaly rho - rowass [l*nkn + k] (baroclinic runs)
    std::vector<void *> tracer;
    std::vector<void *> tracer_bc;   // boundary values per tracer (-999 = none)
    // each tracer's own terms (allocated when first pushed): nudging target
    // cobs and rate rtauv [l*nkn + k], sinking profile wsinkv [l*nkn + k],
    // l = 0..nlvdi, and loads [l*nkn + k] (updated by deposition)
    std::vector<void *> cobs, rtauv, wsinkv, load;
    double *zel = nullptr;    // [6*nel] dbb dbc dcb dcc ut vt (hydro_zeta inputs)
    double *aval = nullptr;   // [nnz] assembled zeta matrix (FP64, as c2coo)
    void *zrow = nullptr;     // S [12*nel] element matrices hia (9) and vectors hik (3)
    double *rhs = nullptr;    // [nkn] (as rvec2d)
    double *x = nullptr;      // [nkn] solution (FP64)
    void *hbuf = nullptr;     // conz3d horizontal terms [l][ii][ie]
    // optional forcing, allocated when first pushed (null: absent): node
    // atmospheric pressure (Pa), equilibrium tide (m), level nudging term
    // (m/s); fixed-transport elements and their transports [l*nel + ie]
    void *ppv = nullptr, *zeqv = nullptr, *andgz = nullptr;
    int *iuvfix = nullptr;
    void *utfix = nullptr, *vtfix = nullptr;
};

// zeta-solver control words (Solver::ctrl): iteration (+1) at which the
// current Krylov solve stopped, its convergence (same stamp) / breakdown /
// iteration-limit flags, its matvec count, iterations of all inner solves of
// this zeta solve, FP64 refinement converged, factorisation failure,
// refinements done; words CT_ITS..CT_NREF are cleared at the start of a solve
// CT_RITS.. hold the iterations of each refinement round (FP32 path)
enum { CT_STOPIT = 0, CT_CONV = 1, CT_FAIL = 2, CT_NMV = 3, CT_ITS = 4, CT_REFCONV = 5, CT_FACTOR = 6,
       CT_NREF = 7, CT_MAXIT = 8, CT_RITS = 9, CT_NRITS = 6, CT_WORDS = 16 };

struct Solver {               // Krylov work space (value type V = S of the model)
    int nblk = 0, brows = 0, max_blk_nnz = 0, max_blk_pairs = 0, nparts = 0, max_elev = 0;
    int last_its[8] = {};     // iteration count of the previous solve, per solve kind
    int last_nref = 0;        // FP64 refinement rounds of the previous solve
    int round_its[CT_NRITS] = {};   // iterations of each round of the previous solve
    void *lu = nullptr;       // block factors, fixed stride max_blk_nnz per block
    void *ell_a = nullptr;    // matrix values in the ELL slots (zero padding)
    void *fs_val = nullptr;   // block factor rows, slot-major (with fs_col)
    void *r0 = nullptr, *r = nullptr, *p0 = nullptr, *p1 = nullptr, *v0 = nullptr, *v1 = nullptr,
         *s = nullptr, *t = nullptr, *dx = nullptr, *b = nullptr, *x0 = nullptr;
    double *part = nullptr;   // [32*nparts] per-workgroup partial sums (4 regions x 8)
    double *scal = nullptr;   // [16] device scalars
    int *ctrl = nullptr;      // [16] control words of the solve (flags, counters)
    void *a32 = nullptr;      // FP32 copy of the matrix values (FP32 path)
    void *zprev = nullptr;    // FP32 path: level at the start of the previous step
    bool have_prev = false;   // zprev holds a level of this run's time series
    bool have_lu = false;     // lu holds the block factors of an earlier matrix of this pattern
    bool last_warm = false;   // the last factorisation refined earlier factors
    int solves_since_exact = 0;
    bool fed = false;         // the assembly already wrote the solve's inputs (ELL
                              // slots, FP32 copy, initial guess, cleared words)
    double *r64 = nullptr;    // FP64 residual (refinement)
};

// kernel groups timed with HIP events when timing is enabled
enum TimerId { T_SETDRY = 0, T_COPY, T_INTERN, T_SETNOD, T_ZETA, T_SOLVE, T_WETDRY, T_FINAL, T_SETUVD, T_DEPTH,
               T_VERT, T_TTOV, T_CONZ_ELEM, T_CONZ_NODE, T_BNDO, T_FACTOR, T_BICG_A, T_BICG_B, T_CONZ_PREP, T_TVD_GRAD, T_NTIMERS };

// A record series resident on the device (time interpolation of forcing
// files): times [nrec], values record-major [nrec][ncomp][nval]; nintp
// points per interpolation (0: one record, a constant field), the ring of
// records starting at record jstart; time null: no series.
struct SeriesDev {
    const double *time;
    const float *val;
    int nrec, nval, ncomp, nintp, jstart;
};
// meteo_convert_wind_data settings (wind forcing)
struct WindPar {
    int iwtype, itdrag;
    float dragco, roluft, rowass, wslim, wsmax;
};
struct Forcing {
    // boundary-level series per open boundary, its factor (zfact) and the
    // time range its ring covers
    std::vector<SeriesDev> bser;
    std::vector<float> zfact;
    std::vector<double> t_lo, t_hi;
    std::vector<void *> bbuf;          // device buffers of each boundary's series (2 per boundary)
    SeriesDev *d_bser = nullptr;       // [nbnd]
    float *d_zfact = nullptr;          // [nbnd]
    int *bnd_pos = nullptr;            // [nbnd_nodes] position of the node in its boundary's list
    int nseries = 0;                   // boundaries with a series
    // wind series (node values or one point) and conversion settings
    SeriesDev wind{};
    WindPar wpar{};
    double w_lo = 0., w_hi = 0.;
    void *wbuf[2] = {nullptr, nullptr};
    bool wind_on = false;
    bool wind_applied = false;         // a one-record (constant) wind is converted once
    // point sources of set_mass_flux: node k's sources are [src_ptr[k],
    // src_ptr[k+1]) in the caller's order; per source its layer limits
    // (levmin, levmax) and its index into the discharge array
    int nsrc = 0;
    int *src_ptr = nullptr, *src_lev = nullptr, *src_idx = nullptr;
    float *src_q = nullptr;            // [nsrc] discharges (rflux)
    float *rqds = nullptr;             // [nkn] distributed sources (rqdsv)
};

struct Model {
    int prec = SHYHIP_FP64;
    int nkn = 0, nel = 0, nlvdi = 0, nlv = 0, ntr = 0;
    int nl_tmpl = 1;          // kernel template depth (>= nlvdi)
    shyhip_params p{};
    hipStream_t stream = nullptr;
    Mesh m;
    State s;
    Solver sv;
    int nnz = 0, ncolor = 0, nbnd = 0;
    int *d_err = nullptr;     // [16] device error words + counters; [8..13] the
                              // scalar-stability maximum and result (64-bit words)
    int *h_pinned = nullptr;  // pinned host mirror for small readbacks
    double dtime = 0., dtime_old = -1.;
    bool has_coriolis = false, has_mflux = false, tvd_ready = false;
    // each tracer's nudging factor, sinking velocity and load factor
    // (conz3d_omp's robs, wsink, rload); the tracer whose terms the
    // stability index includes (-1: none)
    std::vector<float> tr_robs, tr_wsink, tr_rload;
    int terms_tracer = -1;
    // baroclinic coupling (ibarcl = 1): tracers holding T and S
    bool bcl_on = false;
    int bcl_temp = -1, bcl_salt = -1;
    int last_nloop = 0, last_iters = 0, last_refine = 0, last_istot = 1;
    long solver_retries = 0;   // zeta solves redone with exact factors after a warm start failed
    // members with their own scal3sh substeps (ensembles, set at set_basin
    // for more than one member): node/element -> member maps, each member's
    // istot of the current step, the stability words (3 per member) and
    // their pinned host mirror; mem_istot_on while a step's transport uses them
    int nmem_dev = 0;
    int *d_node_mem = nullptr, *d_elem_mem = nullptr, *d_mem_istot = nullptr;   // mem_istot: 2 slots of nmem
    int *cur_mem_istot = nullptr;    // the slot of the current step (pipelined steps alternate)
    unsigned long long *d_mem_stab = nullptr, *h_mem_stab = nullptr;
    std::vector<int> mem_istot;
    bool mem_istot_on = false;
    // independent basins held in one model: member m owns nodes
    // [mem_node[m], mem_node[m+1]) and elements [mem_elem[m], mem_elem[m+1])
    std::vector<int> mem_node, mem_elem;
    float *d_levels = nullptr, *h_levels = nullptr;   // boundary levels staging (nbnd)
    // tracers after the first run on streams of their own inside a step
    // (independent transports, shared read-only state), each with its own
    // scratch; created on first use
    std::vector<hipStream_t> tr_stream;
    std::vector<hipEvent_t> tr_done;
    hipEvent_t ev_hydro = nullptr;
    std::vector<void *> grp_hbuf;     // conz3d element terms of each tracer group
    // bndo_setbc's per-step outflow test of each tracer group, and of the
    // single-tracer scal3sh entry point (nbndo * nlvdi values; freed when
    // the boundaries change)
    std::vector<void *> grp_bscal;
    void *bscal1 = nullptr;
    // the step's conz3d element-layer records of each tracer group and of
    // the single-tracer entry points (conz_record_bytes each)
    std::vector<void *> grp_rec;
    void *rec1 = nullptr;
    std::string err;
    // host-side copies needed for init / plugin / checks
    std::vector<int> h_nen3v, h_rowptr, h_cols, h_ijp, h_color, h_ilhv, h_ilhkv, h_ibnd, h_ibtyp, h_iextb;
    std::vector<float> h_x, h_y;
    std::vector<double> h_ev;             // [19*nel] reference ev layout
    std::vector<double> h_coo, h_rvec;    // plugin accumulation (c2coo, rvec2d)
    std::vector<double> h_raux;           // explicit system: the diagonal (raux2d)
    bool sys_explicit = false;            // system_set_explicit was called
    void *d_cbound = nullptr;             // S [nlvdi*nkn] source concentrations of shyhip_conz3d
    size_t esize() const { return prec == SHYHIP_FP64 ? 8 : 4; }
    bool deferred = false;    // shyhip_step does not wait for the device at its end
    // Pipelined transport (deferred checks, no timing, no mass-flux field):
    // the tracers of step k are transported on streams of their own while
    // step k+1's hydro runs.  The arrays the transport reads are rotated over
    // three buffers (two for ulnv/vlnv and the wet/dry mask it sees), so a
    // step rewrites only buffers whose readers, two steps back, are done.
    bool pipe_live = false;            // a transport may still be running
    int pipe_k = 0;                    // pipelined step counter (parity)
    bool pipe_rec[2] = {false, false}; // events of parity p recorded
    hipStream_t tr0_stream = nullptr;  // tracer 0's stream when pipelined
    std::vector<hipEvent_t> pipe_ev[2];
    void *spare_ut = nullptr, *spare_vt = nullptr, *spare_wl = nullptr, *spare_hdk = nullptr, *spare_hde = nullptr;
    void *uln_b = nullptr, *vln_b = nullptr;
    int *iweg_tr[2] = {nullptr, nullptr};
    int *iweg_snap = nullptr;          // the element kernel of this step copies iwegv here
    const int *iweg_conz = nullptr;    // mask the open-boundary treatment reads (null: iwegv)
    hipEvent_t ev_levels = nullptr;   // the last upload from h_levels
    Forcing f;                         // forcing series on the device
    // timing
    bool timing = false;
    // launch sequences replayed from HIP graphs (graph_run): instantiated
    // graphs by the bytes of their launch arguments; on by default
    bool use_graphs = true;
    std::unordered_map<std::string, hipGraphExec_t> graphs;
    long graph_builds = 0, graph_replays = 0;
    struct TEv { int id; hipEvent_t a, b; };
    std::vector<TEv> tev;
    std::vector<int> tev_open;          // groups begun and not yet ended (nesting)
    double t_ms[T_NTIMERS] = {};
    long t_cnt[T_NTIMERS] = {};
};

// The incident elements of node k held in registers (element<<2 | vertex):
// slots s < cnt hold list entries s; padding slots hold element 0.  Entries
// from kIncWidth on are read from the CSR list (tail()).
// Workgroups are dispatched to the chip's 8 XCDs in turn (workgroup b to XCD
// b % 8), and each XCD has its own L2.  xcd_block() renumbers them so that
// each XCD takes one contiguous range of tiles: tiles that gather the same
// data (neighbouring elements and nodes) then share one L2 instead of
// fetching it into several.
// A bijection of [0, nb).
constexpr int kXcds = 8;
__device__ inline int xcd_block(int b, int nb) {
    const int x = b % kXcds, i = b / kXcds, per = nb / kXcds, rem = nb % kXcds;
    return x * per + min(x, rem) + i;
}
__device__ inline int xcd_block() { return xcd_block(blockIdx.x, gridDim.x); }

struct IncSlots {
    int pk[kIncWidth];
    int cnt = 0;
    __device__ inline void load(const int *inc_cnt, const int *inc_el, int nkn, int k) {
        cnt = inc_cnt[k];
#pragma unroll
        for (int s = 0; s < kIncWidth; s++) pk[s] = inc_el[(size_t)s * nkn + k];
    }
};

// event marks around a launch group (no-op unless timing is enabled)
void timer_begin(Model &M, int id);
void timer_end(Model &M);
int timer_flush(Model &M);

// Launch of a single kernel timed by events bound to its own dispatch when
// timing is enabled: the events carry the kernel's start and end
// timestamps, so the figure is the kernel's execution time, without the
// marker packets that timer_begin/timer_end put in the stream.
template <typename... KArgs, typename... Args>
void launch_timed(Model &M, int id, void (*kernel)(KArgs...), dim3 grid, dim3 block, size_t lds, hipStream_t st,
                  Args... args) {
    if (!M.timing) {
        hipLaunchKernelGGL(kernel, grid, block, lds, st, args...);
        return;
    }
    Model::TEv t{id, nullptr, nullptr};
    (void)hipEventCreate(&t.a);
    (void)hipEventCreate(&t.b);
    hipExtLaunchKernelGGL(kernel, grid, block, (std::uint32_t)lds, st, t.a, t.b, 0u, args...);
    M.tev.push_back(t);
}

// host-side mesh preprocessing (shyhip_init.cpp)
void host_set_ev(int nkn, int nel, const int *nen3v0, const float *x, const float *y, double *ev);
int host_node_list(int nkn, int nel, const int *nen3v0, std::vector<int> &ntg, std::vector<int> &cols,
                   std::vector<int> &ijp);
int host_color(int nkn, int nel, const int *nen3v0, std::vector<int> &color);
int host_diffweight(int nel, const double *ev, int idtype, std::vector<double> &w);
void parallel_for(int n, int grain, const std::function<void(int, int)> &fn);
int host_bndo_tables(int nkn, int nel, const int *nen3v0, const float *x, const float *y, const double *ev,
                     const std::vector<int> &bnodes, std::vector<double> &nx, std::vector<double> &ny,
                     std::vector<int> &ptr, std::vector<int> &nbr, std::vector<double> &w, bool f32);
void host_tvd_tables(int nkn, int nel, const int *nen3v0, const float *x, const float *y, const double *ev,
                     std::vector<int> &tel, std::vector<int> &tkn, std::vector<double> &tw);
struct BlockSchedule {
    int nblk = 0, brows = 0;
    std::vector<int> rowptr, cols, src, diag, fwd_rows, fwd_lptr, fwd_nlev, bwd_rows, bwd_lptr, bwd_nlev;
    // ILU(0) updates: for entry q of the block pattern, the pairs (q_ik, q_kj),
    // k < min(i, j) ascending, and ent_jd[q] = the diagonal entry of column j
    // for an L entry, -2 - (local row) for a diagonal, -1 for U
    std::vector<int> cp_ptr, cp_a, cp_b, ent_jd;
    // entry-level schedule of the factorisation: ent_lev[q] = 1 + the
    // highest level among the entries that q reads (0 if none); ent_nlev[b]
    // = levels of block b
    std::vector<int> ent_lev, ent_nlev;
};
void host_block_schedule(int n, const std::vector<int> &rowptr, const std::vector<int> &cols, int brows,
                         BlockSchedule &bs);

// kernel launchers (shyhip_hydro.hip, shyhip_solver.hip, shyhip_conz.hip)
int launch_set_dry(Model &M, int iweich, int bnewtime, int *changed);
int launch_setnod(Model &M);
int launch_hydro_intern(Model &M);
int launch_zeta_assemble(Model &M);
// system_get + setweg(1), enqueued only: gate (a solver flag, may be null)
// switches both off unless set; dx (may be null) is a correction added to
// the FP64 solution first; the changed-element count is left in d_err[1]
// for the caller's read-back
int launch_system_get_setweg1(Model &M, int bnewtime, const int *gate, const double *dx = nullptr);
int launch_transports_final(Model &M);
int launch_setzev_setuvd(Model &M);
int launch_setdepth(Model &M, bool old_level);
int launch_setdepth_node(Model &M);                 // node depths only (element depths done)
// transports_final + setzev/setuvd/baro2l + element depths + ttov/uvint, fused
int launch_elem_after_solve(Model &M);
int launch_hydro_vertical(Model &M);
int launch_ttov_uvint(Model &M);
int launch_init_levels(Model &M, float zconst);
int launch_copy_old(Model &M);
// set_dry + copy_uvz/copy_depth of a time step in one launch; rotate3:
// the arrays of the pipelined transport rotate over three buffers
int launch_step_start(Model &M, int bnewtime, bool rotate3 = false);
int launch_set_rzv(Model &M, int ibc, float level);
int launch_set_rzv_all(Model &M);
// sp111(2) and meteo_force on the device at time dtime: boundary levels of
// the boundaries with a series into rzv, the wind series into the stress
// (shyhip_forcing.hip); the host has checked that dtime is covered
int launch_forcing(Model &M, double dtime);
int launch_mass_flux(Model &M, bool with_rqds);
// rhoset (baroclinic.f90:654-720) niter times: the density anomaly of every
// node-layer from T, S and the mid-layer pressure (shyhip_hydro.hip)
int launch_rhoset(Model &M, int niter, hipStream_t st);
int launch_fill(Model &M, void *ptr, double value, size_t n);
int launch_volume(Model &M, int mode, double *vol);
// the allocation-free volume / tracer mass of caller device arrays
int noalloc_masscont(int dtype, int nkn, const int *ilhkv, const void *hdk, const void *areakv, const void *conc,
                     double *out, hipStream_t st, std::string &err);
int launch_tracer_mass(Model &M, int itr, double *mass);
int solver_setup(Model &M);
// Zeta solve.  wetdry_bnew >= 0: system_get + setweg(1) (bnewtime =
// wetdry_bnew) are enqueued behind the solve, switched on by its
// convergence flag, before each host read of the solver flags, so one
// read-back returns the solver state, d_err[0] and the changed-element
// count (in h_pinned[0..1]).
int solver_solve(Model &M, int *iters, int *refinements, int wetdry_bnew = -1);
// the tracers one conz3d launch transports together (1 or 2): slots, whether
// small values are clipped (not temperature), source concentrations (device,
// may be null)
// true when tracer itr has nudging, sinking or loads switched on with
// their fields pushed
bool tracer_has_terms(const Model &M, int itr);
struct TracerSet {
    int n = 1;
    int itr[2] = {0, 0};
    int clip[2] = {0, 0};
    const void *cbound[2] = {nullptr, nullptr};
};
// conz3d substep isact of istot: hbuf holds the element terms, rec the
// step's element-layer records (conz_record_bytes), computed first when prep
int launch_conz3d(Model &M, const TracerSet &ts, int istot, int isact, hipStream_t st, void *hbuf, void *rec,
                  bool prep);
size_t conz_record_bytes(const Model &M);
// the element-term scratch (hbuf) of a launch of ntr tracers, in values of
// the tracer type: the element terms (conz_hbuf_terms), then room for the
// node gradients of itvd = 1 (conz_hbuf_elems in all)
size_t conz_hbuf_terms(const Model &M, int ntr);
size_t conz_hbuf_elems(const Model &M, int ntr);
// the bytes that identify the launches of a tracer group's substeps 1..istot
// (conz3d and bndo_setbc), appended to key (a graph_run key)
void conz3d_key(Model &M, const TracerSet &ts, int istot, void *hbuf, void *rec, void *bscal, std::string &key);
// one substep of one tracer, its records computed first
int launch_conz3d(Model &M, int itr, int istot, int isact, int clip_small, hipStream_t st, void *hbuf, void *rec,
                  const void *cbound = nullptr);
int launch_bndo(Model &M, int itr, hipStream_t st, int isact = 1);
// bndo_setbc of substep isact; scal: the step's outflow-test buffer of the
// tracer group (nbndo * nlvdi storage values, written by substep 1, read by
// the later ones; null: recomputed by every call)
int launch_bndo(Model &M, const TracerSet &ts, hipStream_t st, int isact, void *scal);
// scalar_stability on the current state: istot and rindex of each member
// (one entry per member when the model holds several, else one)
// (synchronises)
int launch_scalar_stability(Model &M, float rstol, int *istot, double *rindex);

// helpers
int hip_check(Model &M, hipError_t e, const char *what);
// One lock for the operations that may synchronise the whole device --
// allocating and freeing device or pinned memory -- and for stream capture,
// shared by every model of the process: a capture on one model's stream must
// not overlap such an operation made by another thread for another model.
// The models' streams are non-blocking, so the copies of other threads on the
// default stream do not touch a capturing stream either.
std::recursive_mutex &device_lock();
// A fixed sequence of kernel launches on stream st, replayed from a HIP
// graph: key holds every argument of the launches (the bytes of their
// argument structures and grids); the first call with a key captures the
// launches on st and instantiates them, later calls launch the instantiated
// graph.  The launches must be kernels only (no memset, copy or host
// synchronisation).  Runs them directly when graphs are off or timing is on.
int graph_run(Model &M, hipStream_t st, const std::string &key, const std::function<int()> &launches);
template <typename T>
inline void key_add(std::string &key, const T &v) {
    key.append(reinterpret_cast<const char *>(&v), sizeof v);
}
#define HIPCHK(M, call) do { if (hip_check((M), (call), #call)) return 1; } while (0)

}  // namespace shy

struct shyhip_model {
    shy::Model M;
};
