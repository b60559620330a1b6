// shyhip_api.cpp -- C ABI of libshyhip.so (declared in include/shyhip.h).
//
// Owns the device mirrors of the model state, converts between the
// reference's host layout (Fortran (l,ie), layer fastest) and the device's
// layer-major layout, and sequences the kernels of one time step in the order
// of hydro (src/shyfem/hydrodynamic.f90:284-484) and scal3sh
// (src/shyfem/concentration.f90:565-810).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "shyhip_internal.h"

using namespace shy;

namespace shy {

int hip_check(Model &M, hipError_t e, const char *what) {
    if (e == hipSuccess) return 0;
    M.err = std::string("HIP error in ") + what + ": " + hipGetErrorString(e);
    return 1;
}

std::recursive_mutex &device_lock() {
    static std::recursive_mutex m;
    return m;
}

int graph_run(Model &M, hipStream_t st, const std::string &key, const std::function<int()> &launches) {
    if (!M.use_graphs || M.timing) return launches();
    auto it = M.graphs.find(key);
    if (it == M.graphs.end()) {
        std::lock_guard<std::recursive_mutex> lock(device_lock());
        // a bounded cache: the argument sets of a run are few (rotating
        // buffers, substep counts); start over if a run keeps making new ones
        if (M.graphs.size() >= 256) {
            for (auto &g : M.graphs) (void)hipGraphExecDestroy(g.second);
            M.graphs.clear();
        }
        HIPCHK(M, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        const int rc = launches();
        hipGraph_t graph = nullptr;
        const hipError_t e = hipStreamEndCapture(st, &graph);
        if (rc) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        HIPCHK(M, e);
        hipGraphExec_t exec = nullptr;
        const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        HIPCHK(M, ei);
        it = M.graphs.emplace(key, exec).first;
        M.graph_builds++;
    } else {
        M.graph_replays++;
    }
    HIPCHK(M, hipGraphLaunch(it->second, st));
    return 0;
}

void timer_begin(Model &M, int id) {
    if (!M.timing) return;
    Model::TEv t{id, nullptr, nullptr};
    (void)hipEventCreate(&t.a);
    (void)hipEventCreate(&t.b);
    (void)hipEventRecord(t.a, M.stream);
    M.tev_open.push_back((int)M.tev.size());
    M.tev.push_back(t);
}

void timer_end(Model &M) {
    if (!M.timing || M.tev_open.empty()) return;
    (void)hipEventRecord(M.tev[M.tev_open.back()].b, M.stream);
    M.tev_open.pop_back();
}

int timer_flush(Model &M) {
    if (M.tev.empty()) return 0;
    HIPCHK(M, hipStreamSynchronize(M.stream));
    for (auto &t : M.tev) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess) {
            M.t_ms[t.id] += ms;
            M.t_cnt[t.id] += 1;
        }
        (void)hipEventDestroy(t.a);
        (void)hipEventDestroy(t.b);
    }
    M.tev.clear();
    return 0;
}

// true when tracer itr has nudging, sinking or loads switched on with their
// fields pushed
bool tracer_has_terms(const Model &M, int itr) {
    if (itr < 0 || itr >= M.ntr) return false;
    return (M.tr_robs[itr] != 0.f && M.s.cobs[itr] && M.s.rtauv[itr]) ||
           (M.tr_wsink[itr] != 0.f && M.s.wsinkv[itr]) || (M.tr_rload[itr] != 0.f && M.s.load[itr]);
}

}  // namespace shy

namespace {

thread_local std::string g_err;   // errors before a handle exists

int fail(Model *M, const std::string &msg, int code = 1) {
    if (M) M->err = msg;
    else g_err = msg;
    return code;
}

int pipe_drain(Model &M);
void forcing_free_bnd(Model &M);
void forcing_free_wind(Model &M);
void forcing_free_sources(Model &M);

template <typename T>
int dalloc(Model &M, T **p, size_t n) {
    *p = nullptr;
    if (n == 0) n = 1;
    std::lock_guard<std::recursive_mutex> lock(device_lock());
    HIPCHK(M, hipMalloc((void **)p, n * sizeof(T)));
    HIPCHK(M, hipMemsetAsync(*p, 0, n * sizeof(T), M.stream));
    return 0;
}

int dalloc_s(Model &M, void **p, size_t n) {
    *p = nullptr;
    if (n == 0) n = 1;
    std::lock_guard<std::recursive_mutex> lock(device_lock());
    HIPCHK(M, hipMalloc(p, n * M.esize()));
    HIPCHK(M, hipMemsetAsync(*p, 0, n * M.esize(), M.stream));
    return 0;
}

template <typename T>
int upload(Model &M, T *d, const T *h, size_t n) {
    if (n == 0) return 0;
    HIPCHK(M, hipMemcpyAsync(d, h, n * sizeof(T), hipMemcpyHostToDevice, M.stream));
    HIPCHK(M, hipStreamSynchronize(M.stream));
    return 0;
}

// upload host floats/doubles into a device array of the model's storage type
int upload_s(Model &M, void *d, const double *h, size_t n) {
    if (M.prec == SHYHIP_FP64) return upload(M, (double *)d, h, n);
    std::vector<float> t(n);
    for (size_t i = 0; i < n; i++) t[i] = (float)h[i];
    return upload(M, (float *)d, t.data(), n);
}

void dfree(void *p) {
    std::lock_guard<std::recursive_mutex> lock(device_lock());
    if (p) (void)hipFree(p);
}

int nl_template(int nlvdi) {
    const int opts[] = {1, 2, 4, 8, 12, 16, 20, 24, 32};
    for (int o : opts)
        if (nlvdi <= o) return o;
    return -1;
}

// field descriptors for push/pull
enum Kind { K_ELEM3, K_NODE3, K_NODEW, K_VERT, K_ELEM2, K_NODE2, K_DD, K_IELEM, K_INODE, K_CSR, K_RHS };
struct Field {
    void **ptr;
    Kind kind;
    bool isint = false;
    bool isdouble_dev = false;  // device array always double
    bool lazy = false;          // optional: allocated at the first push
};

bool field_of(Model &M, int id, Field &f) {
    auto S = [&](void **p, Kind k) { f = Field{p, k, false, false}; return true; };
    switch (id) {
    case SHYHIP_ZNV: return S(&M.s.zn, K_NODE2);
    case SHYHIP_ZOV: return S(&M.s.zo, K_NODE2);
    case SHYHIP_ZENV: return S(&M.s.zen, K_VERT);
    case SHYHIP_ZEOV: return S(&M.s.zeo, K_VERT);
    case SHYHIP_UTLNV: return S(&M.s.utn, K_ELEM3);
    case SHYHIP_VTLNV: return S(&M.s.vtn, K_ELEM3);
    case SHYHIP_UTLOV: return S(&M.s.uto, K_ELEM3);
    case SHYHIP_VTLOV: return S(&M.s.vto, K_ELEM3);
    case SHYHIP_UNV: return S(&M.s.un, K_ELEM2);
    case SHYHIP_VNV: return S(&M.s.vn, K_ELEM2);
    case SHYHIP_UOV: return S(&M.s.uo, K_ELEM2);
    case SHYHIP_VOV: return S(&M.s.vo, K_ELEM2);
    case SHYHIP_ULNV: return S(&M.s.uln, K_ELEM3);
    case SHYHIP_VLNV: return S(&M.s.vln, K_ELEM3);
    case SHYHIP_WLNV: return S(&M.s.wln, K_NODEW);
    case SHYHIP_WLOV: return S(&M.s.wlo, K_NODEW);
    case SHYHIP_HDKNV: return S(&M.s.hdkn, K_NODE3);
    case SHYHIP_HDKOV: return S(&M.s.hdko, K_NODE3);
    case SHYHIP_HDENV: return S(&M.s.hden, K_ELEM3);
    case SHYHIP_HDEOV: return S(&M.s.hdeo, K_ELEM3);
    case SHYHIP_IWEGV: f = Field{(void **)&M.s.iweg, K_IELEM, true}; return true;
    case SHYHIP_IWETV: f = Field{(void **)&M.s.iwet, K_IELEM, true}; return true;
    case SHYHIP_INODV: f = Field{(void **)&M.s.inod, K_INODE, true}; return true;
    case SHYHIP_RFRICV: return S(&M.s.rfric, K_ELEM2);
    case SHYHIP_DDXV: return S(&M.s.ddx, K_DD);
    case SHYHIP_DDYV: return S(&M.s.ddy, K_DD);
    case SHYHIP_VISV: return S(&M.s.visv, K_NODEW);
    case SHYHIP_DIFV: return S(&M.s.difv, K_NODEW);
    case SHYHIP_DIFHV: return S(&M.s.difhv, K_ELEM3);
    case SHYHIP_TAUXNV: return S(&M.s.taux, K_NODE2);
    case SHYHIP_TAUYNV: return S(&M.s.tauy, K_NODE2);
    case SHYHIP_RZV: return S(&M.s.rzv, K_NODE2);
    case SHYHIP_RQV: return S(&M.s.rqv, K_NODE2);
    case SHYHIP_MFLUXV: return S(&M.s.mflux, K_NODE3);
    case SHYHIP_RHOV: return S(&M.s.rhov, K_NODE3);
    case SHYHIP_AREAKV: return S(&M.m.areakv, K_NODE3);
    case SHYHIP_CSR_VAL: f = Field{(void **)&M.s.aval, K_CSR, false, true}; return true;
    case SHYHIP_RHS: f = Field{(void **)&M.s.rhs, K_RHS, false, true}; return true;
    case SHYHIP_PPV: f = Field{&M.s.ppv, K_NODE2, false, false, true}; return true;
    case SHYHIP_ZEQV: f = Field{&M.s.zeqv, K_NODE2, false, false, true}; return true;
    case SHYHIP_ANDGZV: f = Field{&M.s.andgz, K_NODE2, false, false, true}; return true;
    case SHYHIP_IUVFIX: f = Field{(void **)&M.s.iuvfix, K_IELEM, true, false, true}; return true;
    case SHYHIP_UTFIX: f = Field{&M.s.utfix, K_ELEM3, false, false, true}; return true;
    case SHYHIP_VTFIX: f = Field{&M.s.vtfix, K_ELEM3, false, false, true}; return true;
    default:
        if (id >= SHYHIP_COBS0 && id < SHYHIP_COBS0 + M.ntr) {
            f = Field{&M.s.cobs[id - SHYHIP_COBS0], K_NODE3, false, false, true};
            return true;
        }
        if (id >= SHYHIP_RTAUV0 && id < SHYHIP_RTAUV0 + M.ntr) {
            f = Field{&M.s.rtauv[id - SHYHIP_RTAUV0], K_NODE3, false, false, true};
            return true;
        }
        if (id >= SHYHIP_WSINKV0 && id < SHYHIP_WSINKV0 + M.ntr) {
            f = Field{&M.s.wsinkv[id - SHYHIP_WSINKV0], K_NODEW, false, false, true};
            return true;
        }
        if (id >= SHYHIP_LOAD0 && id < SHYHIP_LOAD0 + M.ntr) {
            f = Field{&M.s.load[id - SHYHIP_LOAD0], K_NODE3, false, false, true};
            return true;
        }
        if (id >= SHYHIP_TRACER0 && id < SHYHIP_TRACER0 + M.ntr) return S(&M.s.tracer[id - SHYHIP_TRACER0], K_NODE3);
        if (id >= SHYHIP_TRACER_BC0 && id < SHYHIP_TRACER_BC0 + M.ntr)
            return S(&M.s.tracer_bc[id - SHYHIP_TRACER_BC0], K_NODE3);
        return false;
    }
}

// shape: host index h -> device index d
struct Shape {
    size_t n;
    int rows, cols;   // host (rows fastest) = device (cols fastest) transposition
};
Shape shape_of(Model &M, Kind k) {
    switch (k) {
    case K_ELEM3: return {(size_t)M.nlvdi * M.nel, M.nlvdi, M.nel};
    case K_NODE3: return {(size_t)M.nlvdi * M.nkn, M.nlvdi, M.nkn};
    case K_NODEW: return {(size_t)(M.nlvdi + 1) * M.nkn, M.nlvdi + 1, M.nkn};
    case K_VERT: return {(size_t)3 * M.nel, 3, M.nel};
    case K_DD: return {(size_t)2 * M.nlvdi * M.nel, 2 * M.nlvdi, M.nel};
    case K_ELEM2: case K_IELEM: return {(size_t)M.nel, 1, M.nel};
    case K_NODE2: case K_INODE: return {(size_t)M.nkn, 1, M.nkn};
    case K_CSR: return {(size_t)M.nnz, 1, M.nnz};
    case K_RHS: return {(size_t)M.nkn, 1, M.nkn};
    }
    return {0, 0, 0};
}

// wait for any transport still running (every entry point but the step
// and the forcing setters calls this first)
int pipe_drain(Model &M) {
    if (!M.pipe_live) return 0;
    if (M.tr0_stream) HIPCHK(M, hipStreamSynchronize(M.tr0_stream));
    for (hipStream_t st : M.tr_stream) HIPCHK(M, hipStreamSynchronize(st));
    M.pipe_live = false;
    M.pipe_rec[0] = M.pipe_rec[1] = false;
    return 0;
}

}  // namespace

extern "C" {

const char *shyhip_last_error(const shyhip_model *h) { return h ? h->M.err.c_str() : g_err.c_str(); }

int shyhip_device_count(int *n) {
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    *n = c;
    return 0;
}

int shyhip_create(shyhip_model **out, int precision, int nkn, int nel, int nlvdi, int nlv, int ntracer) {
    *out = nullptr;
    if (precision != SHYHIP_FP64 && precision != SHYHIP_FP32) return fail(nullptr, "precision must be FP64 or FP32");
    if (nkn <= 0 || nel <= 0 || nlvdi <= 0 || nlv <= 0 || nlv > nlvdi || ntracer < 0)
        return fail(nullptr, "invalid sizes");
    const int tmpl = nl_template(nlvdi);
    if (tmpl < 0) return fail(nullptr, "more than 32 layers are not supported");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(nullptr, "no HIP device available");
    auto *h = new shyhip_model();
    Model &M = h->M;
    M.prec = precision;
    M.nkn = nkn; M.nel = nel; M.nlvdi = nlvdi; M.nlv = nlv; M.ntr = ntracer;
    M.nl_tmpl = tmpl;
    {
        std::lock_guard<std::recursive_mutex> lock(device_lock());
        if (hip_check(M, hipStreamCreateWithFlags(&M.stream, hipStreamNonBlocking), "hipStreamCreateWithFlags")) {
            delete h;
            return fail(nullptr, "stream");
        }
        if (hip_check(M, hipHostMalloc((void **)&M.h_pinned, 64 * sizeof(int)), "hipHostMalloc")) {
            delete h;
            return fail(nullptr, "pinned");
        }
    }
    // default parameters (def_para.f90)
    shyhip_params &p = M.p;
    p.dt = 300.; p.azpar = 0.5f; p.ampar = 0.5f; p.afpar = 0.5f; p.atpar = 1.f; p.adpar = 1.f; p.aapar = 1.f;
    p.grav = 9.81f; p.rowass = 1025.f; p.vismol = 1.0e-6f; p.ireib = 0; p.uvmin = 0.2f;
    p.hzmin = 0.01f; p.hzoff = 0.05f; p.hzon = 0.10f; p.volmin = 1.f; p.iclin = 0;
    p.rtol = (double)1.0e-8f; p.atol = (double)1.0e-8f; p.itermax = 1000; p.refine_rtol = 1e-10;
    p.block_rows = 0; p.tri_sweeps = 0; p.inner_rtol = 1e-5; p.rkpar = 0.f; p.difmol = 1.0e-6f; p.itvd = 0; p.itvdv = 0;
    p.node_tile = 0; p.idtype = 2;
    const size_t ne = nel, nk = nkn, L = nlvdi;
    State &s = M.s;
    int rc = 0;
    rc |= dalloc_s(M, &s.zn, nk); rc |= dalloc_s(M, &s.zo, nk);
    rc |= dalloc_s(M, &s.zen, 3 * ne); rc |= dalloc_s(M, &s.zeo, 3 * ne);
    for (void **q : {&s.utn, &s.vtn, &s.uto, &s.vto, &s.uln, &s.vln, &s.hden, &s.hdeo, &s.difhv}) rc |= dalloc_s(M, q, L * ne);
    for (void **q : {&s.un, &s.vn, &s.uo, &s.vo, &s.rfric}) rc |= dalloc_s(M, q, ne);
    for (void **q : {&s.wln, &s.wlo, &s.visv, &s.difv}) rc |= dalloc_s(M, q, (L + 1) * nk);
    for (void **q : {&s.hdkn, &s.hdko, &s.mflux, &s.rhov}) rc |= dalloc_s(M, q, L * nk);
    for (void **q : {&s.taux, &s.tauy, &s.rzv, &s.rqv}) rc |= dalloc_s(M, q, nk);
    rc |= dalloc_s(M, &s.ddx, 2 * L * ne); rc |= dalloc_s(M, &s.ddy, 2 * L * ne);
    rc |= dalloc(M, &s.iweg, ne); rc |= dalloc(M, &s.iwet, ne); rc |= dalloc(M, &s.inod, nk);
    rc |= dalloc(M, &s.zel, 6 * ne);
    rc |= dalloc(M, &s.rhs, nk); rc |= dalloc(M, &s.x, nk);
    rc |= dalloc_s(M, &s.hbuf, conz_hbuf_elems(M, 1));
    s.tracer.assign(ntracer, nullptr);
    s.tracer_bc.assign(ntracer, nullptr);
    s.cobs.assign(ntracer, nullptr); s.rtauv.assign(ntracer, nullptr);
    s.wsinkv.assign(ntracer, nullptr); s.load.assign(ntracer, nullptr);
    M.tr_robs.assign(ntracer, 0.f); M.tr_wsink.assign(ntracer, 0.f); M.tr_rload.assign(ntracer, 0.f);
    for (int t = 0; t < ntracer; t++) {
        rc |= dalloc_s(M, &s.tracer[t], L * nk);
        rc |= dalloc_s(M, &s.tracer_bc[t], L * nk);
    }
    rc |= dalloc(M, &M.d_err, 16);
    if (rc) {
        std::string e = M.err;
        shyhip_destroy(h);
        return fail(nullptr, "device allocation failed: " + e);
    }
    if (launch_fill(M, s.rzv, (double)kFlag, nk)) return fail(&M, M.err);
    for (int t = 0; t < ntracer; t++)
        if (launch_fill(M, s.tracer_bc[t], -999., L * nk)) return fail(&M, M.err);
    *out = h;
    return 0;
}

int shyhip_destroy(shyhip_model *h) {
    if (!h) return 0;
    Model &M = h->M;
    if (M.stream) (void)hipStreamSynchronize(M.stream);
    (void)pipe_drain(M);
    {
        std::lock_guard<std::recursive_mutex> lock(device_lock());
        for (auto &g : M.graphs) (void)hipGraphExecDestroy(g.second);
    }
    M.graphs.clear();
    State &s = M.s;
    for (void *q : {s.zn, s.zo, s.zen, s.zeo, s.utn, s.vtn, s.uto, s.vto, s.uln, s.vln, s.un, s.vn, s.uo, s.vo,
                    s.wln, s.wlo, s.hdkn, s.hdko, s.hden, s.hdeo, s.rfric, s.ddx, s.ddy, s.visv, s.difv, s.difhv,
                    s.taux, s.tauy, s.rzv, s.rqv, s.mflux, s.rhov, s.hbuf})
        dfree(q);
    dfree(s.iweg); dfree(s.iwet); dfree(s.inod); dfree(s.zel); dfree(s.aval); dfree(s.zrow); dfree(s.rhs); dfree(s.x);
    for (void *q : {s.ppv, s.zeqv, s.andgz, (void *)s.iuvfix, s.utfix, s.vtfix}) dfree(q);
    for (void *t : s.tracer) dfree(t);
    for (void *t : s.tracer_bc) dfree(t);
    for (auto *v : {&s.cobs, &s.rtauv, &s.wsinkv, &s.load})
        for (void *t : *v) dfree(t);
    dfree(M.m.bndo_k); dfree(M.m.bndo_ptr); dfree(M.m.bndo_nb); dfree(M.m.bndo_nx); dfree(M.m.bndo_ny);
    dfree(M.m.bndo_w);
    for (void *q : {(void *)M.m.bndo_open, (void *)M.m.bndo_lmax, (void *)M.m.bndo_eptr, (void *)M.m.bndo_el,
                    (void *)M.m.bndo_elev, (void *)M.m.bndo_nblev, (void *)M.m.bndo_eaj})
        dfree(q);
    dfree(M.m.wdif);
    Mesh &m = M.m;
    for (void *q : {(void *)m.nen3v, (void *)m.geo, (void *)m.ang, (void *)m.dis, (void *)m.ilhv, (void *)m.ilhkv,
                    m.hm3v, m.hev, m.hlv, m.hldv, m.areakv, (void *)m.ibnd, (void *)m.zbound, (void *)m.iextb,
                    (void *)m.ibtyp, m.fcorv, m.czv, (void *)m.nk_ptr, (void *)m.nk_el, (void *)m.nkc_el,
                    (void *)m.inc_cnt, (void *)m.inc_el, (void *)m.inc_elc, (void *)m.slot_of, (void *)m.slot_lev,
                    (void *)m.slot_aj4,
                    (void *)m.rowptr, (void *)m.cols, (void *)m.ijp, (void *)m.bnd_nodes, (void *)m.bnd_id,
                    (void *)m.blk_rowptr, (void *)m.blk_cols, (void *)m.blk_src, (void *)m.blk_diag, (void *)m.blk_cp_ptr,
                    (void *)m.blk_cp_a, (void *)m.blk_cp_b, (void *)m.blk_jd, (void *)m.blk_pcol, (void *)m.blk_prp,
                    (void *)m.blk_pdg, (void *)m.fs_col, (void *)m.fs_meta, (void *)m.blk_fdesc, (void *)m.blk_fpair,
                    (void *)m.blk_elp, (void *)m.blk_elr, (void *)m.blk_enlev, (void *)m.ell_col, (void *)m.ell_len, (void *)m.csr_ell,
                    (void *)m.fwd_rows, (void *)m.fwd_lptr, (void *)m.fwd_nlev, (void *)m.bwd_rows,
                    (void *)m.bwd_lptr, (void *)m.bwd_nlev, (void *)m.tvd_el, (void *)m.tvd_kn, m.tvd_w,
                    (void *)m.tvd_lev, (void *)m.elem_iext, (void *)m.conz_eitem, (void *)m.conz_ntile, (void *)m.conz_nitem})
        dfree(q);
    Solver &v = M.sv;
    for (void *q : {v.lu, v.ell_a, v.fs_val, v.r0, v.r, v.p0, v.p1, v.v0, v.v1, v.s, v.t, v.dx, v.b, v.x0, (void *)v.part,
                    (void *)v.scal, (void *)v.ctrl, v.a32, (void *)v.r64, v.zprev})
        dfree(q);
    for (hipStream_t st : M.tr_stream) (void)hipStreamSynchronize(st);
    for (hipStream_t st : M.tr_stream) (void)hipStreamDestroy(st);
    for (hipEvent_t ev : M.tr_done) (void)hipEventDestroy(ev);
    if (M.ev_hydro) (void)hipEventDestroy(M.ev_hydro);
    if (M.ev_levels) (void)hipEventDestroy(M.ev_levels);
    if (M.tr0_stream) (void)hipStreamSynchronize(M.tr0_stream);
    if (M.tr0_stream) (void)hipStreamDestroy(M.tr0_stream);
    for (int p = 0; p < 2; p++)
        for (hipEvent_t ev : M.pipe_ev[p]) (void)hipEventDestroy(ev);
    for (void *q : {M.spare_ut, M.spare_vt, M.spare_wl, M.spare_hdk, M.spare_hde, M.uln_b, M.vln_b,
                    (void *)M.iweg_tr[0], (void *)M.iweg_tr[1]})
        dfree(q);
    for (void *q : M.grp_hbuf) dfree(q);
    for (void *q : M.grp_bscal) dfree(q);
    dfree(M.bscal1);
    for (void *q : M.grp_rec) dfree(q);
    dfree(M.rec1);
    forcing_free_bnd(M);
    forcing_free_wind(M);
    forcing_free_sources(M);
    dfree(M.d_err);
    dfree(M.d_cbound);
    dfree(M.d_node_mem); dfree(M.d_elem_mem); dfree(M.d_mem_istot); dfree(M.d_mem_stab);
    if (M.h_mem_stab) { std::lock_guard<std::recursive_mutex> lock(device_lock()); (void)hipHostFree(M.h_mem_stab); }
    dfree(M.d_levels);
    if (M.h_levels) { std::lock_guard<std::recursive_mutex> lock(device_lock()); (void)hipHostFree(M.h_levels); }
    if (M.h_pinned) { std::lock_guard<std::recursive_mutex> lock(device_lock()); (void)hipHostFree(M.h_pinned); }
    if (M.stream) (void)hipStreamDestroy(M.stream);
    delete h;
    return 0;
}

int shyhip_get_params(const shyhip_model *h, shyhip_params *p) {
    if (!h || !p) return fail(nullptr, "null argument");
    *p = h->M.p;
    return 0;
}

// the parameter checks of shyhip_set_params, against a model (M) or alone
static int check_params(Model *M, const shyhip_params *p) {
    if (p->dt <= 0.) return fail(M, "dt must be positive");
    if (p->itvdv > 0 && p->aapar != 0.f)
        return fail(M, "implicit vertical advection (aapar != 0) cannot be combined with itvdv > 0");
    if (p->ireib == 8 || p->ireib == 9)
        return fail(M, "ireib = 8/9 (roughness from the sediment or fluid-mud modules) is not supported by the device path");
    if (p->ireib < 0 || (p->ireib > 7 && p->ireib != 10))
        return fail(M, "unknown friction (ireib), as bottom_friction stops (chezy.f90:373-376)");
    if (!(p->rkpar >= 0.f)) return fail(M, "rkpar must be zero or positive");
    if (p->idtype < 0 || p->idtype > 2) return fail(M, "idtype must be 0, 1 or 2");
    if (p->itvd < 0 || p->itvd > 2 || p->itvdv < 0 || p->itvdv > 1) return fail(M, "itvd must be 0, 1 or 2, itvdv 0 or 1");
    if (p->block_rows < 0 || p->block_rows > kMaxBlockRows) return fail(M, "block_rows must be in 0..1024");
    if (p->tri_sweeps < 0 || p->tri_sweeps > 64) return fail(M, "tri_sweeps must be in 0..64");
    if (p->node_tile < 0 || p->node_tile > kNodeTileMax)
        return fail(M, "node_tile must be in 0..256");
    if (!(p->inner_rtol > 0. && p->inner_rtol < 1.)) return fail(M, "inner_rtol must be in (0,1)");
    return 0;
}

int shyhip_check_params(const shyhip_params *p) {
    if (!p) return fail(nullptr, "null argument");
    return check_params(nullptr, p);
}

int shyhip_set_params(shyhip_model *h, const shyhip_params *p) {
    if (!h || !p) return fail(nullptr, "null argument");
    Model &M = h->M;
    if (check_params(&M, p)) return 1;
    M.p = *p;
    return 0;
}

int shyhip_set_basin(shyhip_model *h, const int *nen3v, const float *xgv, const float *ygv, const float *hm3v,
                     const float *hev, const float *hlv, const float *hldv, const int *ilhv, const int *ilhkv) {
    if (!h) return fail(nullptr, "null handle");
    if (pipe_drain(h->M)) return 1;
    Model &M = h->M;
    const int nkn = M.nkn, nel = M.nel, L = M.nlvdi;
    M.h_nen3v.resize(3 * (size_t)nel);
    for (size_t i = 0; i < 3 * (size_t)nel; i++) {
        const int k = nen3v[i] - 1;
        if (k < 0 || k >= nkn) return fail(&M, "nen3v holds a node id outside 1..nkn");
        M.h_nen3v[i] = k;
    }
    for (int ie = 0; ie < nel; ie++)
        if (ilhv[ie] < 1 || ilhv[ie] > M.nlv) return fail(&M, "ilhv outside 1..nlv");
    for (int k = 0; k < nkn; k++)
        if (ilhkv[k] < 1 || ilhkv[k] > M.nlv) return fail(&M, "ilhkv outside 1..nlv");
    M.h_x.assign(xgv, xgv + nkn);
    M.h_y.assign(ygv, ygv + nkn);
    M.h_ilhv.assign(ilhv, ilhv + nel);
    M.h_ilhkv.assign(ilhkv, ilhkv + nkn);
    M.h_ev.assign(19 * (size_t)nel, 0.);
    host_set_ev(nkn, nel, M.h_nen3v.data(), xgv, ygv, M.h_ev.data());
    M.m.wdif_idtype = -1;                   // diffusion weights of the new geometry
    for (int ie = 0; ie < nel; ie++)
        if (!(M.h_ev[19 * (size_t)ie + 9] > 0.)) return fail(&M, "element " + std::to_string(ie + 1) + " is not counter-clockwise");
    M.nnz = host_node_list(nkn, nel, M.h_nen3v.data(), M.h_rowptr, M.h_cols, M.h_ijp);
    if (M.nnz <= 0) return fail(&M, "building the zeta-matrix pattern failed");
    if (M.mem_node.empty()) {
        M.mem_node = {0, nkn};
        M.mem_elem = {0, nel};
    }
    for (size_t m = 0; m + 1 < M.mem_node.size(); m++)
        for (int ie = M.mem_elem[m]; ie < M.mem_elem[m + 1]; ie++)
            for (int ii = 0; ii < 3; ii++) {
                const int k = M.h_nen3v[3 * (size_t)ie + ii];
                if (k < M.mem_node[m] || k >= M.mem_node[m + 1])
                    return fail(&M, "element " + std::to_string(ie + 1) + " uses a node of another member");
            }
    // maps for members that choose their own scal3sh substeps
    {
        const int nmem = (int)M.mem_node.size() - 1;
        dfree(M.d_node_mem); dfree(M.d_elem_mem); dfree(M.d_mem_istot); dfree(M.d_mem_stab);
        M.d_node_mem = M.d_elem_mem = M.d_mem_istot = nullptr;
        M.d_mem_stab = nullptr;
        if (M.h_mem_stab) { std::lock_guard<std::recursive_mutex> lock(device_lock()); (void)hipHostFree(M.h_mem_stab); }
        M.h_mem_stab = nullptr;
        M.nmem_dev = nmem;
        if (nmem > 1) {
            std::vector<int> nm(nkn), em(nel);
            for (int m = 0; m < nmem; m++) {
                for (int k = M.mem_node[m]; k < M.mem_node[m + 1]; k++) nm[k] = m;
                for (int ie = M.mem_elem[m]; ie < M.mem_elem[m + 1]; ie++) em[ie] = m;
            }
            int rc = dalloc(M, &M.d_node_mem, nkn) || upload(M, M.d_node_mem, nm.data(), nkn) ||
                     dalloc(M, &M.d_elem_mem, nel) || upload(M, M.d_elem_mem, em.data(), nel) ||
                     dalloc(M, &M.d_mem_istot, 2 * (size_t)nmem) || dalloc(M, &M.d_mem_stab, 3 * (size_t)nmem);
            if (rc) return 1;
            std::lock_guard<std::recursive_mutex> lock(device_lock());
            HIPCHK(M, hipHostMalloc((void **)&M.h_mem_stab, 2 * (size_t)nmem * sizeof(unsigned long long)));
            M.cur_mem_istot = M.d_mem_istot;
        }
    }
    // element colouring of every member on its own (domain_clusterization_ggu)
    // (members on the host's threads: each writes its own elements)
    M.h_color.assign(nel, 0);
    const int nmemb = (int)M.mem_node.size() - 1;
    std::vector<int> mem_nc(nmemb, 0);
    parallel_for(nmemb, 1, [&](int m_begin, int m_end) {
        for (int m = m_begin; m < m_end; m++) {
            const int k0 = M.mem_node[m], e0 = M.mem_elem[m], ne = M.mem_elem[m + 1] - e0;
            std::vector<int> loc(3 * (size_t)ne), col;
            for (size_t i = 0; i < loc.size(); i++) loc[i] = M.h_nen3v[3 * (size_t)e0 + i] - k0;
            mem_nc[m] = host_color(M.mem_node[m + 1] - k0, ne, loc.data(), col);
            if (mem_nc[m] > 0) std::copy(col.begin(), col.end(), M.h_color.begin() + e0);
        }
    });
    M.ncolor = 0;
    for (int m = 0; m < nmemb; m++) {
        if (mem_nc[m] <= 0) {
            M.ncolor = mem_nc[m];
            break;
        }
        M.ncolor = std::max(M.ncolor, mem_nc[m]);
    }
    if (M.ncolor <= 0) return fail(&M, "element colouring failed (more than 31 colours)");
    Mesh &m = M.m;
    int rc = 0;
    // geometry, SoA
    std::vector<double> geo(7 * (size_t)nel), ang(3 * (size_t)nel), dis(3 * (size_t)nel);
    for (int ie = 0; ie < nel; ie++) {
        const double *e = &M.h_ev[19 * (size_t)ie];
        for (int q = 0; q < 6; q++) geo[(size_t)q * nel + ie] = e[3 + q];
        geo[(size_t)6 * nel + ie] = e[9];
        for (int q = 0; q < 3; q++) {
            ang[(size_t)q * nel + ie] = e[10 + q];
            dis[(size_t)q * nel + ie] = e[16 + q];
        }
    }
    rc |= dalloc(M, &m.nen3v, 3 * (size_t)nel) || upload(M, m.nen3v, M.h_nen3v.data(), 3 * (size_t)nel);
    rc |= dalloc(M, &m.geo, geo.size()) || upload(M, m.geo, geo.data(), geo.size());
    rc |= dalloc(M, &m.ang, ang.size()) || upload(M, m.ang, ang.data(), ang.size());
    rc |= dalloc(M, &m.dis, dis.size()) || upload(M, m.dis, dis.data(), dis.size());
    rc |= dalloc(M, &m.ilhv, nel) || upload(M, m.ilhv, ilhv, nel);
    rc |= dalloc(M, &m.ilhkv, nkn) || upload(M, m.ilhkv, ilhkv, nkn);
    if (M.nlv > 32) return fail(&M, "at most 32 layers");
    // depths (vertex-major on device)
    std::vector<double> t(3 * (size_t)nel);
    for (int ie = 0; ie < nel; ie++)
        for (int ii = 0; ii < 3; ii++) t[(size_t)ii * nel + ie] = hm3v[3 * (size_t)ie + ii];
    rc |= dalloc_s(M, &m.hm3v, t.size()) || upload_s(M, m.hm3v, t.data(), t.size());
    std::vector<double> he(hev, hev + nel), hl(hlv, hlv + L), hd(hldv, hldv + L);
    rc |= dalloc_s(M, &m.hev, nel) || upload_s(M, m.hev, he.data(), nel);
    rc |= dalloc_s(M, &m.hlv, L) || upload_s(M, m.hlv, hl.data(), L);
    rc |= dalloc_s(M, &m.hldv, L) || upload_s(M, m.hldv, hd.data(), L);
    // set_area (elems_dealing.f90:658-700) in the storage precision, element order
    {
        std::vector<double> area((size_t)L * nkn, 0.);
        if (M.prec == SHYHIP_FP64) {
            for (int ie = 0; ie < nel; ie++) {
                const double areafv = (12. * M.h_ev[19 * (size_t)ie + 9]) / 3.;
                for (int ii = 0; ii < 3; ii++)
                    for (int l = 0; l < ilhv[ie]; l++) area[(size_t)l * nkn + M.h_nen3v[3 * ie + ii]] += areafv;
            }
        } else {
            std::vector<float> af((size_t)L * nkn, 0.f);
            for (int ie = 0; ie < nel; ie++) {
                const float areafv = (float)(12. * M.h_ev[19 * (size_t)ie + 9]) / 3.f;
                for (int ii = 0; ii < 3; ii++)
                    for (int l = 0; l < ilhv[ie]; l++) af[(size_t)l * nkn + M.h_nen3v[3 * ie + ii]] += areafv;
            }
            for (size_t i = 0; i < af.size(); i++) area[i] = af[i];
        }
        rc |= dalloc_s(M, &m.areakv, area.size()) || upload_s(M, m.areakv, area.data(), area.size());
    }
    // node -> element incidence: element order and colour order
    {
        std::vector<int> ptr(nkn + 1, 0), el(3 * (size_t)nel), elc(3 * (size_t)nel);
        for (int ie = 0; ie < nel; ie++)
            for (int ii = 0; ii < 3; ii++) ptr[M.h_nen3v[3 * ie + ii] + 1]++;
        for (int k = 0; k < nkn; k++) ptr[k + 1] += ptr[k];
        std::vector<int> pos(ptr.begin(), ptr.end() - 1), pos2 = pos;
        for (int ie = 0; ie < nel; ie++)
            for (int ii = 0; ii < 3; ii++) {
                const int k = M.h_nen3v[3 * ie + ii];
                el[pos[k]++] = (ie << 2) | ii;
            }
        // colour order: colours ascending, elements ascending within a colour
        std::vector<int> sof(3 * (size_t)nel), slev(3 * (size_t)nel);
        std::vector<double> saj4(3 * (size_t)nel);
        for (int c = 1; c <= M.ncolor; c++)
            for (int ie = 0; ie < nel; ie++) {
                if (M.h_color[ie] != c) continue;
                for (int ii = 0; ii < 3; ii++) {
                    const int k = M.h_nen3v[3 * ie + ii];
                    const int q = pos2[k]++;
                    elc[q] = (ie << 2) | ii;
                    sof[(size_t)ii * nel + ie] = q;
                    slev[q] = ilhv[ie];
                    saj4[q] = 4. * M.h_ev[19 * (size_t)ie + 9];
                }
            }
        rc |= dalloc(M, &m.slot_of, sof.size()) || upload(M, m.slot_of, sof.data(), sof.size());
        rc |= dalloc(M, &m.slot_lev, slev.size()) || upload(M, m.slot_lev, slev.data(), slev.size());
        rc |= dalloc(M, &m.slot_aj4, saj4.size()) || upload(M, m.slot_aj4, saj4.data(), saj4.size());
        rc |= dalloc(M, &m.nk_ptr, nkn + 1) || upload(M, m.nk_ptr, ptr.data(), ptr.size());
        rc |= dalloc(M, &m.nk_el, el.size()) || upload(M, m.nk_el, el.data(), el.size());
        rc |= dalloc(M, &m.nkc_el, elc.size()) || upload(M, m.nkc_el, elc.data(), elc.size());
        std::vector<int> icnt(nkn), iel((size_t)kIncWidth * nkn, 0), ielc((size_t)kIncWidth * nkn, 0);
        for (int k = 0; k < nkn; k++) {
            icnt[k] = ptr[k + 1] - ptr[k];
            for (int s = 0; s < kIncWidth && s < icnt[k]; s++) {
                iel[(size_t)s * nkn + k] = el[ptr[k] + s];
                ielc[(size_t)s * nkn + k] = elc[ptr[k] + s];
            }
        }
        rc |= dalloc(M, &m.inc_cnt, icnt.size()) || upload(M, m.inc_cnt, icnt.data(), icnt.size());
        rc |= dalloc(M, &m.inc_el, iel.size()) || upload(M, m.inc_el, iel.data(), iel.size());
        rc |= dalloc(M, &m.inc_elc, ielc.size()) || upload(M, m.inc_elc, ielc.data(), ielc.size());
    }
    // conz3d work division: the element-layer list (chunks of kConzChunk
    // elements, layer by layer) and the node tiles (at most node_tile nodes
    // and kNodeItems node-layers each)
    {
        std::vector<int> items;
        items.reserve((size_t)nel * 2);
        for (int c0 = 0; c0 < nel; c0 += kConzChunk) {
            const int c1 = std::min(nel, c0 + kConzChunk);
            int lmax = 0;
            for (int ie = c0; ie < c1; ie++) lmax = std::max(lmax, ilhv[ie]);
            for (int l = 1; l <= lmax; l++)
                for (int ie = c0; ie < c1; ie++)
                    if (ilhv[ie] >= l) items.push_back((ie << 5) | (l - 1));
        }
        if (nel >= (1 << 26)) return fail(&M, "at most 2^26 elements");
        const int tn = M.p.node_tile > 0 ? std::min(M.p.node_tile, kNodeTileMax) : kNodeTileMax;
        std::vector<int> tiles{0};
        int cnt = 0, nl = 0;
        for (int k = 0; k < nkn; k++) {
            if (cnt == tn || nl + ilhkv[k] > kNodeItems) {
                tiles.push_back(k);
                cnt = 0;
                nl = 0;
            }
            cnt++;
            nl += ilhkv[k];
        }
        tiles.push_back(nkn);
        std::vector<int> nitems;
        nitems.reserve((size_t)nkn * 2);
        for (int c0 = 0; c0 < nkn; c0 += kConzChunk) {
            const int c1 = std::min(nkn, c0 + kConzChunk);
            int lmax = 0;
            for (int k = c0; k < c1; k++) lmax = std::max(lmax, ilhkv[k]);
            for (int l = 1; l <= lmax; l++)
                for (int k = c0; k < c1; k++)
                    if (ilhkv[k] >= l) nitems.push_back((k << 5) | (l - 1));
        }
        if (nkn >= (1 << 26)) return fail(&M, "at most 2^26 nodes");
        m.n_nitem = (int)nitems.size();
        rc |= dalloc(M, &m.conz_nitem, nitems.size()) || upload(M, m.conz_nitem, nitems.data(), nitems.size());
        m.n_eitem = (int)items.size();
        m.n_ntile = (int)tiles.size() - 1;
        m.ntile_tn = tn;
        rc |= dalloc(M, &m.conz_eitem, items.size()) || upload(M, m.conz_eitem, items.data(), items.size());
        rc |= dalloc(M, &m.conz_ntile, tiles.size()) || upload(M, m.conz_ntile, tiles.data(), tiles.size());
    }
    // CSR pattern
    rc |= dalloc(M, &m.rowptr, nkn + 1) || upload(M, m.rowptr, M.h_rowptr.data(), nkn + 1);
    rc |= dalloc(M, &m.cols, M.nnz) || upload(M, m.cols, M.h_cols.data(), M.nnz);
    rc |= dalloc(M, &m.ijp, M.h_ijp.size()) || upload(M, m.ijp, M.h_ijp.data(), M.h_ijp.size());
    rc |= dalloc(M, &M.s.aval, M.nnz);
    rc |= dalloc_s(M, &M.s.zrow, (size_t)12 * nel);
    // block-ILU(0) schedule
    {
        int brows = M.p.block_rows > 0 ? M.p.block_rows : std::min(nkn, 256);
        if (nkn <= kMaxBlockRows && M.p.block_rows == 0) brows = nkn;   // one block: plain ILU(0)
        BlockSchedule bs;
        host_block_schedule(nkn, M.h_rowptr, M.h_cols, brows, bs);
        M.sv.nblk = bs.nblk;
        M.sv.brows = bs.brows;
        int mx = 0;
        for (int b = 0; b < bs.nblk; b++) {
            const int b0 = b * brows, b1 = std::min(nkn, b0 + brows);
            mx = std::max(mx, bs.rowptr[b1] - bs.rowptr[b0]);
        }
        M.sv.max_blk_nnz = mx;
        int mxp = 0;
        for (int b = 0; b < bs.nblk; b++) {
            const int b0 = b * brows, b1 = std::min(nkn, b0 + brows);
            mxp = std::max(mxp, bs.cp_ptr[bs.rowptr[b1]] - bs.cp_ptr[bs.rowptr[b0]]);
        }
        M.sv.max_blk_pairs = mxp;
        if ((size_t)mx * M.esize() > 64 * 1024) return fail(&M, "block too large for LDS; lower block_rows");
        rc |= dalloc(M, &m.blk_rowptr, bs.rowptr.size()) || upload(M, m.blk_rowptr, bs.rowptr.data(), bs.rowptr.size());
        rc |= dalloc(M, &m.blk_cols, bs.cols.size()) || upload(M, m.blk_cols, bs.cols.data(), bs.cols.size());
        rc |= dalloc(M, &m.blk_src, bs.src.size()) || upload(M, m.blk_src, bs.src.data(), bs.src.size());
        rc |= dalloc(M, &m.blk_diag, bs.diag.size()) || upload(M, m.blk_diag, bs.diag.data(), bs.diag.size());
        rc |= dalloc(M, &m.blk_cp_ptr, bs.cp_ptr.size()) || upload(M, m.blk_cp_ptr, bs.cp_ptr.data(), bs.cp_ptr.size());
        rc |= dalloc(M, &m.blk_cp_a, std::max<size_t>(1, bs.cp_a.size())) ||
              upload(M, m.blk_cp_a, bs.cp_a.data(), bs.cp_a.size());
        rc |= dalloc(M, &m.blk_cp_b, std::max<size_t>(1, bs.cp_b.size())) ||
              upload(M, m.blk_cp_b, bs.cp_b.data(), bs.cp_b.size());
        rc |= dalloc(M, &m.blk_jd, bs.ent_jd.size()) || upload(M, m.blk_jd, bs.ent_jd.data(), bs.ent_jd.size());
        rc |= dalloc(M, &m.fwd_rows, bs.fwd_rows.size()) || upload(M, m.fwd_rows, bs.fwd_rows.data(), bs.fwd_rows.size());
        rc |= dalloc(M, &m.fwd_lptr, bs.fwd_lptr.size()) || upload(M, m.fwd_lptr, bs.fwd_lptr.data(), bs.fwd_lptr.size());
        rc |= dalloc(M, &m.fwd_nlev, bs.fwd_nlev.size()) || upload(M, m.fwd_nlev, bs.fwd_nlev.data(), bs.fwd_nlev.size());
        rc |= dalloc(M, &m.bwd_rows, bs.bwd_rows.size()) || upload(M, m.bwd_rows, bs.bwd_rows.data(), bs.bwd_rows.size());
        rc |= dalloc(M, &m.bwd_lptr, bs.bwd_lptr.size()) || upload(M, m.bwd_lptr, bs.bwd_lptr.data(), bs.bwd_lptr.size());
        rc |= dalloc(M, &m.bwd_nlev, bs.bwd_nlev.size()) || upload(M, m.bwd_nlev, bs.bwd_nlev.data(), bs.bwd_nlev.size());
        for (int b = 0; b < bs.nblk; b++)
            for (int i = b * brows; i < std::min(nkn, (b + 1) * brows); i++)
                if (bs.diag[i] < 0) return fail(&M, "zeta pattern without diagonal");
        {
            // fixed-stride copies of the block pattern (relative indices)
            const int nbk = bs.nblk;
            std::vector<int> pcol((size_t)nbk * mx, 0), prp((size_t)nbk * (brows + 1), 0), pdg((size_t)nbk * brows, 0);
            for (int b = 0; b < nbk; b++) {
                const int b0 = b * brows, nb = std::min(nkn, b0 + brows) - b0;
                const int q0 = bs.rowptr[b0], ne = bs.rowptr[b0 + nb] - q0;
                for (int e = 0; e < ne; e++) pcol[(size_t)b * mx + e] = bs.cols[q0 + e];
                for (int li = 0; li <= brows; li++)
                    prp[(size_t)b * (brows + 1) + li] = li <= nb ? bs.rowptr[b0 + li] - q0 : ne;
                for (int li = 0; li < nb; li++) pdg[(size_t)b * brows + li] = bs.diag[b0 + li] - q0;
            }
            // entries by level, per block
            int dmax = 0;
            for (int b = 0; b < nbk; b++) dmax = std::max(dmax, bs.ent_nlev[b]);
            M.sv.max_elev = dmax;
            std::vector<int> elp((size_t)nbk * (dmax + 1), 0), elr((size_t)nbk * mx, 0);
            for (int b = 0; b < nbk; b++) {
                const int b0 = b * brows, nb = std::min(nkn, b0 + brows) - b0;
                const int q0 = bs.rowptr[b0], ne = bs.rowptr[b0 + nb] - q0, nl = bs.ent_nlev[b];
                int *lp = elp.data() + (size_t)b * (dmax + 1);
                for (int e = 0; e < ne; e++) lp[bs.ent_lev[q0 + e] + 1]++;
                for (int l = 0; l < dmax; l++) lp[l + 1] += lp[l];
                std::vector<int> pos(lp, lp + nl);
                for (int e = 0; e < ne; e++) elr[(size_t)b * mx + pos[bs.ent_lev[q0 + e]]++] = e;
            }
            // descriptors in level order and packed pairs, relative to the block
            std::vector<int> fdesc((size_t)nbk * mx * 4, 0), fpair(2 * bs.cp_a.size() + 2, 0);
            for (int b = 0; b < nbk; b++) {
                const int b0 = b * brows, nb = std::min(nkn, b0 + brows) - b0;
                const int q0 = bs.rowptr[b0], ne = bs.rowptr[b0 + nb] - q0, p0 = bs.cp_ptr[q0];
                for (int t = 0; t < ne; t++) {
                    const int e = elr[(size_t)b * mx + t], q = q0 + e;
                    int *d = &fdesc[((size_t)b * mx + t) * 4];
                    d[0] = e;
                    d[1] = bs.cp_ptr[q] - p0;
                    d[2] = bs.cp_ptr[q + 1] - p0;
                    d[3] = bs.ent_jd[q] >= 0 ? bs.ent_jd[q] - q0 : bs.ent_jd[q];
                }
                for (int p = p0; p < bs.cp_ptr[q0 + ne]; p++) {
                    fpair[2 * (size_t)p] = bs.cp_a[p] - q0;
                    fpair[2 * (size_t)p + 1] = bs.cp_b[p] - q0;
                }
            }
            rc |= dalloc(M, &m.blk_fdesc, fdesc.size()) || upload(M, m.blk_fdesc, fdesc.data(), fdesc.size());
            rc |= dalloc(M, &m.blk_fpair, fpair.size()) || upload(M, m.blk_fpair, fpair.data(), fpair.size());
            rc |= dalloc(M, &m.blk_elp, elp.size()) || upload(M, m.blk_elp, elp.data(), elp.size());
            rc |= dalloc(M, &m.blk_elr, elr.size()) || upload(M, m.blk_elr, elr.data(), elr.size());
            rc |= dalloc(M, &m.blk_enlev, bs.ent_nlev.size()) ||
                  upload(M, m.blk_enlev, bs.ent_nlev.data(), bs.ent_nlev.size());
            rc |= dalloc(M, &m.blk_pcol, pcol.size()) || upload(M, m.blk_pcol, pcol.data(), pcol.size());
            rc |= dalloc(M, &m.blk_prp, prp.size()) || upload(M, m.blk_prp, prp.data(), prp.size());
            rc |= dalloc(M, &m.blk_pdg, pdg.size()) || upload(M, m.blk_pdg, pdg.data(), pdg.size());
            // ELL slices of the matrix rows
            std::vector<int> ecol((size_t)kEllWidth * nkn), elen(nkn), cse(M.nnz, -1);
            for (int i = 0; i < nkn; i++) {
                const int q0 = M.h_rowptr[i], len = M.h_rowptr[i + 1] - q0;
                elen[i] = len;
                for (int k = 0; k < kEllWidth; k++) {
                    const size_t slot = (size_t)k * nkn + i;
                    ecol[slot] = k < len ? M.h_cols[q0 + k] : i;
                    if (k < len) cse[q0 + k] = (int)slot;
                }
            }
            // block factor rows slot-major, when every block row fits the slots
            int maxrow = 0;
            for (int i = 0; i < nkn; i++) maxrow = std::max(maxrow, bs.rowptr[i + 1] - bs.rowptr[i]);
            if (maxrow <= kEllWidth) {
                std::vector<int> fcol((size_t)kEllWidth * nkn), fmeta(nkn);
                for (int i = 0; i < nkn; i++) {
                    // slot 0 the diagonal, slots 1..nl the L entries, then U
                    const int li = i % brows, q0 = bs.rowptr[i], cnt = bs.rowptr[i + 1] - q0;
                    int nl = 0;
                    for (int p = 0; p < cnt; p++) nl += bs.cols[q0 + p] < li;
                    for (int k = 0; k < kEllWidth; k++) fcol[(size_t)k * nkn + i] = li;
                    for (int p = 0; p < cnt; p++) {
                        const int slot = p < nl ? p + 1 : (p == nl ? 0 : p);
                        fcol[(size_t)slot * nkn + i] = bs.cols[q0 + p];
                    }
                    fmeta[i] = nl | (cnt << 8);
                }
                rc |= dalloc(M, &m.fs_col, fcol.size()) || upload(M, m.fs_col, fcol.data(), fcol.size());
                rc |= dalloc(M, &m.fs_meta, fmeta.size()) || upload(M, m.fs_meta, fmeta.data(), fmeta.size());
                rc |= dalloc_s(M, &M.sv.fs_val, (size_t)kEllWidth * nkn);
            }
            rc |= dalloc(M, &m.ell_col, ecol.size()) || upload(M, m.ell_col, ecol.data(), ecol.size());
            rc |= dalloc(M, &m.ell_len, elen.size()) || upload(M, m.ell_len, elen.data(), elen.size());
            rc |= dalloc(M, &m.csr_ell, cse.size()) || upload(M, m.csr_ell, cse.data(), cse.size());
        }
        Solver &v = M.sv;
        rc |= dalloc_s(M, &v.lu, (size_t)bs.nblk * mx);
        v.have_lu = false;
        rc |= dalloc_s(M, &v.ell_a, (size_t)kEllWidth * nkn);
        for (void **q : {&v.r0, &v.r, &v.p0, &v.p1, &v.v0, &v.v1, &v.s, &v.t, &v.dx, &v.b, &v.x0})
            rc |= dalloc_s(M, q, (size_t)nkn);
        const size_t nparts = std::max((size_t)bs.nblk, (size_t)(nkn + kThreads - 1) / kThreads) + 1;
        v.nparts = (int)nparts;
        rc |= dalloc(M, &v.part, 32 * nparts);        // 4 regions x 8 slots
        rc |= dalloc(M, &v.scal, 16);
        rc |= dalloc(M, &v.ctrl, 16);
        rc |= dalloc(M, &v.r64, (size_t)nkn);
        if (M.prec == SHYHIP_FP32) rc |= dalloc(M, (float **)&v.a32, (size_t)M.nnz);
        if (M.prec == SHYHIP_FP32) rc |= dalloc(M, (float **)&v.zprev, (size_t)nkn);
    }
    if (rc) return fail(&M, "basin upload failed: " + M.err);
    // defaults: no open boundary, no external flags
    std::vector<int> zeros(nkn, 0), ty(2, 0);
    rc |= dalloc(M, &m.ibnd, nkn) || dalloc(M, &m.zbound, nkn) || dalloc(M, &m.iextb, nkn);
    rc |= dalloc(M, &m.elem_iext, nel);
    rc |= dalloc(M, &m.ibtyp, 2);
    rc |= dalloc_s(M, &m.fcorv, nel) || dalloc_s(M, &m.czv, nel);
    HIPCHK(M, hipStreamSynchronize(M.stream));
    return rc;
}

int shyhip_set_boundaries(shyhip_model *h, int nbnd, const int *ibnd, const int *ibtyp, const int *iextb) {
    if (!h) return fail(nullptr, "null handle");
    if (pipe_drain(h->M)) return 1;
    Model &M = h->M;
    if (!M.m.ibnd) return fail(&M, "call shyhip_set_basin first");
    const int nkn = M.nkn;
    HIPCHK(M, hipStreamSynchronize(M.stream));
    forcing_free_bnd(M);                            // series belong to the previous boundaries
    for (void *q : M.grp_bscal) dfree(q);           // sized for the previous boundaries
    M.grp_bscal.clear();
    dfree(M.bscal1);
    M.bscal1 = nullptr;
    M.nbnd = nbnd;
    M.h_ibnd.assign(ibnd, ibnd + nkn);
    M.h_ibtyp.assign(ibtyp, ibtyp + nbnd + 1);
    M.h_iextb.assign(iextb ? iextb : ibnd, (iextb ? iextb : ibnd) + nkn);
    if (!iextb) std::fill(M.h_iextb.begin(), M.h_iextb.end(), 0);
    std::vector<int> zb(nkn, 0), nodes, ids;
    for (int k = 0; k < nkn; k++) {
        const int b = ibnd[k];
        if (b < 0 || b > nbnd) return fail(&M, "ibnd outside 0..nbnd");
        if (b > 0) {
            if (ibtyp[b] == 1) zb[k] = 1;
            nodes.push_back(k);
            ids.push_back(b);
        }
    }
    dfree(M.m.ibtyp);
    dfree(M.m.bnd_nodes);
    dfree(M.m.bnd_id);
    HIPCHK(M, hipStreamSynchronize(M.stream));     // no upload from h_levels still queued
    dfree(M.d_levels);
    if (M.h_levels) { std::lock_guard<std::recursive_mutex> lock(device_lock()); (void)hipHostFree(M.h_levels); }
    M.d_levels = nullptr;
    M.h_levels = nullptr;
    if (nbnd > 0) {
        if (dalloc(M, &M.d_levels, nbnd)) return 1;
        std::lock_guard<std::recursive_mutex> lock(device_lock());
        HIPCHK(M, hipHostMalloc((void **)&M.h_levels, nbnd * sizeof(float)));
    }
    M.m.nbnd_nodes = (int)nodes.size();
    int rc = 0;
    rc |= upload(M, M.m.ibnd, ibnd, nkn) || upload(M, M.m.zbound, zb.data(), nkn) ||
          upload(M, M.m.iextb, M.h_iextb.data(), nkn);
    {
        std::vector<int> ext(M.nel, 0);
        for (int ie = 0; ie < M.nel; ie++)
            for (int ii = 0; ii < 3; ii++) ext[ie] += M.h_iextb[M.h_nen3v[3 * (size_t)ie + ii]] != 0;
        rc |= upload(M, M.m.elem_iext, ext.data(), ext.size());
    }
    rc |= dalloc(M, &M.m.ibtyp, nbnd + 1) || upload(M, M.m.ibtyp, ibtyp, nbnd + 1);
    rc |= dalloc(M, &M.m.bnd_nodes, nodes.size()) || upload(M, M.m.bnd_nodes, nodes.data(), nodes.size());
    rc |= dalloc(M, &M.m.bnd_id, ids.size()) || upload(M, M.m.bnd_id, ids.data(), ids.size());
    if (rc) return rc;
    // scalar open-boundary tables for the external (type 1 and 2) boundaries
    std::vector<int> bnodes;
    for (int k = 0; k < nkn; k++) {
        const int b = ibnd[k];
        if (b > 0 && (ibtyp[b] == 1 || ibtyp[b] == 2)) bnodes.push_back(k);
    }
    std::vector<double> bnx, bny, bw;
    std::vector<int> bptr, bnb;
    if (host_bndo_tables(nkn, M.nel, M.h_nen3v.data(), M.h_x.data(), M.h_y.data(), M.h_ev.data(), bnodes, bnx, bny,
                         bptr, bnb, bw, M.prec == SHYHIP_FP32) < 0)
        return fail(&M, "open boundary with a single node is not permitted");
    dfree(M.m.bndo_k); dfree(M.m.bndo_ptr); dfree(M.m.bndo_nb); dfree(M.m.bndo_nx); dfree(M.m.bndo_ny);
    dfree(M.m.bndo_w);
    for (void *q : {(void *)M.m.bndo_open, (void *)M.m.bndo_lmax, (void *)M.m.bndo_eptr, (void *)M.m.bndo_el,
                    (void *)M.m.bndo_elev, (void *)M.m.bndo_nblev, (void *)M.m.bndo_eaj})
        dfree(q);
    M.m.nbndo = (int)bnodes.size();
    {
        // static per-node data of the open-boundary treatment
        std::vector<int> open, lmax, eptr(1, 0), el, elev, nblev;
        std::vector<double> eaj;
        for (int k : bnodes) {
            open.push_back(ibtyp[ibnd[k]] == 1 ? 1 : 0);
            lmax.push_back(M.h_ilhkv[k]);
        }
        std::vector<std::vector<int>> inc(bnodes.size());
        std::vector<int> slot(nkn, -1);
        for (size_t i = 0; i < bnodes.size(); i++) slot[bnodes[i]] = (int)i;
        for (int ie = 0; ie < M.nel; ie++)              // element order
            for (int ii = 0; ii < 3; ii++) {
                const int s = slot[M.h_nen3v[3 * (size_t)ie + ii]];
                if (s >= 0) inc[s].push_back(ie);
            }
        for (auto &v : inc) {
            for (int ie : v) {
                el.push_back(ie);
                elev.push_back(M.h_ilhv[ie]);
                eaj.push_back(M.h_ev[19 * (size_t)ie + 9]);
            }
            eptr.push_back((int)el.size());
        }
        for (int ko : bnb) nblev.push_back(M.h_ilhkv[ko]);
        auto up_i = [&](int **d, const std::vector<int> &h) {
            return dalloc(M, d, std::max<size_t>(1, h.size())) || (!h.empty() && upload(M, *d, h.data(), h.size()));
        };
        rc |= up_i(&M.m.bndo_open, open) || up_i(&M.m.bndo_lmax, lmax) || up_i(&M.m.bndo_eptr, eptr) ||
              up_i(&M.m.bndo_el, el) || up_i(&M.m.bndo_elev, elev) || up_i(&M.m.bndo_nblev, nblev);
        rc |= dalloc(M, &M.m.bndo_eaj, std::max<size_t>(1, eaj.size())) ||
              (!eaj.empty() && upload(M, M.m.bndo_eaj, eaj.data(), eaj.size()));
    }
    rc |= dalloc(M, &M.m.bndo_k, bnodes.size()) || upload(M, M.m.bndo_k, bnodes.data(), bnodes.size());
    rc |= dalloc(M, &M.m.bndo_ptr, bptr.size()) || upload(M, M.m.bndo_ptr, bptr.data(), bptr.size());
    rc |= dalloc(M, &M.m.bndo_nb, bnb.size()) || upload(M, M.m.bndo_nb, bnb.data(), bnb.size());
    rc |= dalloc_s(M, &M.m.bndo_nx, bnx.size()) || upload_s(M, M.m.bndo_nx, bnx.data(), bnx.size());
    rc |= dalloc_s(M, &M.m.bndo_ny, bny.size()) || upload_s(M, M.m.bndo_ny, bny.data(), bny.size());
    rc |= dalloc_s(M, &M.m.bndo_w, bw.size()) || upload_s(M, M.m.bndo_w, bw.data(), bw.size());
    return rc;
}

int shyhip_set_coefficients(shyhip_model *h, const float *fcorv, const float *czv) {
    if (!h) return fail(nullptr, "null handle");
    if (pipe_drain(h->M)) return 1;
    Model &M = h->M;
    if (!M.m.fcorv) return fail(&M, "call shyhip_set_basin first");
    std::vector<double> t(M.nel, 0.);
    const bool had = M.has_coriolis;
    M.has_coriolis = false;
    if (fcorv) {
        for (int i = 0; i < M.nel; i++) {
            t[i] = fcorv[i];
            if (fcorv[i] != 0.f) M.has_coriolis = true;
        }
    }
    // without Coriolis the cross terms of ddxv/ddyv are not written: clear
    // what a coupled step may have left there
    if (had && !M.has_coriolis)
        for (void *q : {M.s.ddx, M.s.ddy})
            if (launch_fill(M, q, 0., 2 * (size_t)M.nlvdi * M.nel)) return 1;
    if (upload_s(M, M.m.fcorv, t.data(), M.nel)) return 1;
    std::fill(t.begin(), t.end(), 0.);
    if (czv)
        for (int i = 0; i < M.nel; i++) t[i] = czv[i];
    return upload_s(M, M.m.czv, t.data(), M.nel);
}

int shyhip_tvd_init(shyhip_model *h) {
    if (!h) return fail(nullptr, "null handle");
    if (pipe_drain(h->M)) return 1;
    Model &M = h->M;
    // upwind tables per member (a member's upwind points lie in its own elements)
    const size_t nel = M.nel;
    std::vector<int> tel(6 * nel, -1), tkn(18 * nel, 0);
    std::vector<double> tw(18 * nel, 0.);
    for (size_t m = 0; m + 1 < M.mem_node.size(); m++) {
        const int k0 = M.mem_node[m], e0 = M.mem_elem[m];
        const int nk = M.mem_node[m + 1] - k0, ne = M.mem_elem[m + 1] - e0;
        std::vector<int> loc(3 * (size_t)ne), ltel, ltkn;
        std::vector<double> ltw;
        for (size_t i = 0; i < loc.size(); i++) loc[i] = M.h_nen3v[3 * (size_t)e0 + i] - k0;
        host_tvd_tables(nk, ne, loc.data(), M.h_x.data() + k0, M.h_y.data() + k0, M.h_ev.data() + 19 * (size_t)e0,
                        ltel, ltkn, ltw);
        for (int p = 0; p < 6; p++)
            for (int ie = 0; ie < ne; ie++) {
                const int up = ltel[(size_t)p * ne + ie];
                tel[(size_t)p * nel + e0 + ie] = up < 0 ? -1 : up + e0;
                for (int q = 0; q < 3; q++) {
                    tkn[((size_t)q * 6 + p) * nel + e0 + ie] = ltkn[((size_t)q * 6 + p) * ne + ie] + k0;
                    tw[((size_t)q * 6 + p) * nel + e0 + ie] = ltw[((size_t)q * 6 + p) * ne + ie];
                }
            }
    }
    std::vector<int> tlev(6 * nel, 0);
    for (size_t q = 0; q < tel.size(); q++) tlev[q] = tel[q] >= 0 ? M.h_ilhv[tel[q]] : 0;
    dfree(M.m.tvd_el); dfree(M.m.tvd_kn); dfree(M.m.tvd_w); dfree(M.m.tvd_lev);
    int rc = 0;
    rc |= dalloc(M, &M.m.tvd_el, tel.size()) || upload(M, M.m.tvd_el, tel.data(), tel.size());
    rc |= dalloc(M, &M.m.tvd_lev, tlev.size()) || upload(M, M.m.tvd_lev, tlev.data(), tlev.size());
    rc |= dalloc(M, &M.m.tvd_kn, tkn.size()) || upload(M, M.m.tvd_kn, tkn.data(), tkn.size());
    double *w = nullptr;
    rc |= dalloc(M, &w, tw.size()) || upload(M, w, tw.data(), tw.size());
    M.m.tvd_w = w;
    M.tvd_ready = rc == 0;
    return rc;
}

int shyhip_init_state(shyhip_model *h, float zconst) {
    if (!h) return fail(nullptr, "null handle");
    if (pipe_drain(h->M)) return 1;
    Model &M = h->M;
    M.sv.have_prev = false;
    HIPCHK(M, hipMemsetAsync(M.d_err, 0, 8 * sizeof(int), M.stream));
    for (void *q : {M.s.utn, M.s.vtn, M.s.uto, M.s.vto})
        if (launch_fill(M, q, 0., (size_t)M.nlvdi * M.nel)) return 1;
    if (launch_init_levels(M, zconst)) return 1;
    // copy_depth; hydro_vertical (zero transports, equal depths -> w = 0);
    // compute_velocities; copy_uvz; copy_depth
    if (launch_copy_old(M)) return 1;            // exchange: old <- new
    if (launch_setdepth(M, false)) return 1;     // new level again (old == new)
    if (launch_hydro_vertical(M)) return 1;
    if (launch_ttov_uvint(M)) return 1;
    if (launch_copy_old(M)) return 1;
    if (launch_setdepth(M, false)) return 1;
    if (launch_hydro_vertical(M)) return 1;
    if (launch_ttov_uvint(M)) return 1;
    M.dtime = 0.;
    M.dtime_old = 0.;
    HIPCHK(M, hipMemcpyAsync(M.h_pinned, M.d_err, sizeof(int), hipMemcpyDeviceToHost, M.stream));
    HIPCHK(M, hipStreamSynchronize(M.stream));
    if (M.h_pinned[0]) return fail(&M, "initialisation failed (device error " + std::to_string(M.h_pinned[0]) + ")");
    return 0;
}

int shyhip_push(shyhip_model *h, int field, const void *host, int dtype, size_t n) {
    if (!h) return fail(nullptr, "null handle");
    if (pipe_drain(h->M)) return 1;
    Model &M = h->M;
    Field f;
    if (!field_of(M, field, f)) return fail(&M, "unknown field id " + std::to_string(field));
    const Shape sh = shape_of(M, f.kind);
    if (n != sh.n) return fail(&M, "push: size mismatch for field " + std::to_string(field));
    if (f.kind == K_NODE3 && field == SHYHIP_MFLUXV) M.has_mflux = true;
    // a pushed level starts a new level history (the FP32 solve's
    // extrapolated start), and a pushed level, matrix or right-hand side is
    // the next solve's input; tracers, coefficients and other fields leave both
    if (field == SHYHIP_ZNV || field == SHYHIP_ZOV || field == SHYHIP_ZENV || field == SHYHIP_ZEOV)
        M.sv.have_prev = false;
    if (field == SHYHIP_ZNV || field == SHYHIP_CSR_VAL || field == SHYHIP_RHS) M.sv.fed = false;
    if (f.lazy && !*f.ptr) {
        const int rc = f.isint ? dalloc(M, (int **)f.ptr, sh.n) : dalloc_s(M, f.ptr, sh.n);
        if (rc) return 1;
    }
    if (f.isint) {
        if (dtype != SHYHIP_I32) return fail(&M, "integer field needs SHYHIP_I32");
        return upload(M, (int *)*f.ptr, (const int *)host, n);
    }
    std::vector<double> t(n);
    for (int r = 0; r < sh.rows; r++)
        for (int c = 0; c < sh.cols; c++) {
            const size_t hi = (size_t)c * sh.rows + r, di = (size_t)r * sh.cols + c;
            t[di] = dtype == SHYHIP_F64 ? ((const double *)host)[hi] : (double)((const float *)host)[hi];
        }
    if (f.isdouble_dev) return upload(M, (double *)*f.ptr, t.data(), n);
    return upload_s(M, *f.ptr, t.data(), n);
}

int shyhip_pull(shyhip_model *h, int field, void *host, int dtype, size_t n) {
    if (!h) return fail(nullptr, "null handle");
    if (pipe_drain(h->M)) return 1;
    Model &M = h->M;
    Field f;
    if (!field_of(M, field, f)) return fail(&M, "unknown field id " + std::to_string(field));
    const Shape sh = shape_of(M, f.kind);
    if (n != sh.n) return fail(&M, "pull: size mismatch for field " + std::to_string(field));
    if (f.lazy && !*f.ptr) return fail(&M, "pull: field " + std::to_string(field) + " was never pushed");
    HIPCHK(M, hipStreamSynchronize(M.stream));
    if (f.isint) {
        if (dtype != SHYHIP_I32) return fail(&M, "integer field needs SHYHIP_I32");
        HIPCHK(M, hipMemcpyAsync(host, *f.ptr, n * sizeof(int), hipMemcpyDeviceToHost, M.stream));
        HIPCHK(M, hipStreamSynchronize(M.stream));
        return 0;
    }
    std::vector<double> t(n);
    if (f.isdouble_dev || M.prec == SHYHIP_FP64) {
        HIPCHK(M, hipMemcpyAsync(t.data(), *f.ptr, n * sizeof(double), hipMemcpyDeviceToHost, M.stream));
        HIPCHK(M, hipStreamSynchronize(M.stream));
    } else {
        std::vector<float> tf(n);
        HIPCHK(M, hipMemcpyAsync(tf.data(), *f.ptr, n * sizeof(float), hipMemcpyDeviceToHost, M.stream));
        HIPCHK(M, hipStreamSynchronize(M.stream));
        for (size_t i = 0; i < n; i++) t[i] = tf[i];
    }
    for (int r = 0; r < sh.rows; r++)
        for (int c = 0; c < sh.cols; c++) {
            const size_t hi = (size_t)c * sh.rows + r, di = (size_t)r * sh.cols + c;
            if (dtype == SHYHIP_F64) ((double *)host)[hi] = t[di];
            else ((float *)host)[hi] = (float)t[di];
        }
    return 0;
}

int shyhip_set_boundary_level(shyhip_model *h, int ibc, float level) {
    if (!h) return fail(nullptr, "null handle");
    return launch_set_rzv(h->M, ibc, level);
}

int shyhip_set_boundary_levels(shyhip_model *h, int n, const float *levels) {
    if (!h) return fail(nullptr, "null handle");
    Model &M = h->M;
    if (n != M.nbnd || !M.d_levels) return fail(&M, "set_boundary_levels: n must equal nbnd");
    // the staging buffer is rewritten only after its previous upload ran
    // (steps may be queued ahead of the device)
    if (!M.ev_levels) HIPCHK(M, hipEventCreateWithFlags(&M.ev_levels, hipEventDisableTiming));
    else HIPCHK(M, hipEventSynchronize(M.ev_levels));
    std::copy(levels, levels + n, M.h_levels);
    HIPCHK(M, hipMemcpyAsync(M.d_levels, M.h_levels, n * sizeof(float), hipMemcpyHostToDevice, M.stream));
    HIPCHK(M, hipEventRecord(M.ev_levels, M.stream));
    return launch_set_rzv_all(M);
}

int shyhip_set_members(shyhip_model *h, int nmem, const int *node_start, const int *elem_start) {
    if (!h) return fail(nullptr, "null handle");
    if (pipe_drain(h->M)) return 1;
    Model &M = h->M;
    if (M.m.nen3v) return fail(&M, "set_members must precede set_basin");
    if (nmem < 1 || node_start[0] != 0 || elem_start[0] != 0 || node_start[nmem] != M.nkn ||
        elem_start[nmem] != M.nel)
        return fail(&M, "set_members: ranges must start at 0 and end at nkn / nel");
    for (int m = 0; m < nmem; m++)
        if (node_start[m + 1] <= node_start[m] || elem_start[m + 1] <= elem_start[m])
            return fail(&M, "set_members: empty or decreasing member range");
    M.mem_node.assign(node_start, node_start + nmem + 1);
    M.mem_elem.assign(elem_start, elem_start + nmem + 1);
    return 0;
}

// The state a member carries from step to step, copied device to device
// from member sm of src into member dm of dst (same basin: node, element
// and layer counts, precision, tracers).  Node arrays [l*nkn + k] and
// element arrays [l*nel + ie] are copied layer by layer with one strided
// copy each; 2D arrays are one layer.
int shyhip_copy_member_state(shyhip_model *dst, int dm, shyhip_model *src, int sm) {
    if (!dst || !src) return fail(nullptr, "null handle");
    Model &D = dst->M, &S = src->M;
    if (pipe_drain(D) || pipe_drain(S)) return 1;
    const int nd = (int)D.mem_node.size() - 1, ns = (int)S.mem_node.size() - 1;
    if (dm < 0 || dm >= nd || sm < 0 || sm >= ns) return fail(&D, "copy_member_state: member out of range");
    const int kn = D.mem_node[dm + 1] - D.mem_node[dm], ke = D.mem_elem[dm + 1] - D.mem_elem[dm];
    if (kn != S.mem_node[sm + 1] - S.mem_node[sm] || ke != S.mem_elem[sm + 1] - S.mem_elem[sm] ||
        D.prec != S.prec || D.nlvdi != S.nlvdi || D.ntr != S.ntr)
        return fail(&D, "copy_member_state: the members' basins, precisions or tracers differ");
    HIPCHK(S, hipStreamSynchronize(S.stream));
    const size_t es = D.esize();
    const int L = D.nlvdi;
    // (array in dst, array in src, element size, node (0) or element (1) array, layers)
    struct F { void *d; const void *s; size_t sz; int elem; int nl; };
    const State &a = D.s, &b = S.s;
    std::vector<F> f = {
        {a.zn, b.zn, es, 0, 1}, {a.zo, b.zo, es, 0, 1}, {a.zen, b.zen, es, 1, 3}, {a.zeo, b.zeo, es, 1, 3},
        {a.utn, b.utn, es, 1, L}, {a.vtn, b.vtn, es, 1, L}, {a.uto, b.uto, es, 1, L}, {a.vto, b.vto, es, 1, L},
        {a.uln, b.uln, es, 1, L}, {a.vln, b.vln, es, 1, L}, {a.un, b.un, es, 1, 1}, {a.vn, b.vn, es, 1, 1},
        {a.uo, b.uo, es, 1, 1}, {a.vo, b.vo, es, 1, 1}, {a.wln, b.wln, es, 0, L + 1}, {a.wlo, b.wlo, es, 0, L + 1},
        {a.hdkn, b.hdkn, es, 0, L}, {a.hdko, b.hdko, es, 0, L}, {a.hden, b.hden, es, 1, L},
        {a.hdeo, b.hdeo, es, 1, L}, {a.iweg, b.iweg, 4, 1, 1}, {a.iwet, b.iwet, 4, 1, 1}, {a.inod, b.inod, 4, 0, 1},
        {a.rfric, b.rfric, es, 1, 1}, {a.visv, b.visv, es, 0, L + 1}, {a.difv, b.difv, es, 0, L + 1},
        {a.taux, b.taux, es, 0, 1}, {a.tauy, b.tauy, es, 0, 1}, {a.rqv, b.rqv, es, 0, 1}, {a.rhov, b.rhov, es, 0, L},
        {a.difhv, b.difhv, es, 1, L}, {a.mflux, b.mflux, es, 0, L},
    };
    for (int t = 0; t < D.ntr; t++) {
        f.push_back({a.tracer[t], b.tracer[t], es, 0, L});
        f.push_back({a.tracer_bc[t], b.tracer_bc[t], es, 0, L});
    }
    for (const F &x : f) {
        if (!x.d || !x.s) continue;
        const size_t w = (size_t)(x.elem ? ke : kn) * x.sz;
        const size_t dp = (size_t)(x.elem ? D.nel : D.nkn) * x.sz, sp = (size_t)(x.elem ? S.nel : S.nkn) * x.sz;
        const size_t d0 = (size_t)(x.elem ? D.mem_elem[dm] : D.mem_node[dm]) * x.sz;
        const size_t s0 = (size_t)(x.elem ? S.mem_elem[sm] : S.mem_node[sm]) * x.sz;
        // zenv/zeov are vertex-major [v*nel + ie]: three "layers" of elements
        HIPCHK(D, hipMemcpy2DAsync((char *)x.d + d0, dp, (const char *)x.s + s0, sp, w, x.nl, hipMemcpyDeviceToDevice,
                                   D.stream));
    }
    HIPCHK(D, hipStreamSynchronize(D.stream));
    return 0;
}

int shyhip_set_wind_uniform(shyhip_model *h, float taux, float tauy) {
    if (!h) return fail(nullptr, "null handle");
    Model &M = h->M;
    if (launch_fill(M, M.s.taux, taux, M.nkn)) return 1;
    return launch_fill(M, M.s.tauy, tauy, M.nkn);
}

int shyhip_set_member_wind(shyhip_model *h, int m, float taux, float tauy) {
    if (!h) return fail(nullptr, "null handle");
    Model &M = h->M;
    if (m < 0 || m + 1 >= (int)M.mem_node.size()) return fail(&M, "set_member_wind: member out of range");
    const size_t k0 = (size_t)M.mem_node[m], n = (size_t)(M.mem_node[m + 1] - M.mem_node[m]);
    if (launch_fill(M, (char *)M.s.taux + k0 * M.esize(), taux, n)) return 1;
    return launch_fill(M, (char *)M.s.tauy + k0 * M.esize(), tauy, n);
}

// ------------------------------------------------------------ forcing series
}  // extern "C"

namespace {

// first record of the initial ring for start time dtime0 (the rule of
// iff_populate_records, intp_fem_file.f90:1065-1102): records are passed
// while the next one (plus the record spacing for the cubic window) is
// before dtime0; -1 when the records cannot serve dtime0
int first_record(int nrec, const double *time, int nintp, double dtime0) {
    const double ddt = nintp == 4 ? time[1] - time[0] : 0.0;
    int j = 0;
    while (time[j + 1] + ddt < dtime0) {
        if (++j + 1 >= nrec) return -1;
    }
    if (dtime0 < time[0] || j + nintp > nrec) return -1;
    return j;
}

// validates a series' times and interpolation order; fills jstart, nintp,
// and the time range [lo, hi] the ring covers
int series_setup(Model &M, const char *what, int nrec, const double *time, int intpol, double dtime0, SeriesDev &s,
                 double &lo, double &hi) {
    if (nrec < 1 || !time) return fail(&M, std::string(what) + ": no records");
    for (int r = 1; r < nrec; r++)
        if (!(time[r] > time[r - 1])) return fail(&M, std::string(what) + ": record times not increasing");
    if (nrec == 1) {
        s.nintp = 0;
        s.jstart = 0;
        lo = -HUGE_VAL;
        hi = HUGE_VAL;
        return 0;
    }
    if (intpol != 2 && intpol != 4) return fail(&M, std::string(what) + ": interpolation must use 2 or 4 records");
    if (nrec < intpol) return fail(&M, std::string(what) + ": fewer records than the interpolation uses");
    s.nintp = intpol;
    s.jstart = first_record(nrec, time, intpol, dtime0);
    if (s.jstart < 0) return fail(&M, std::string(what) + ": the records do not cover the start time");
    lo = time[s.jstart];
    hi = time[nrec - 1];
    return 0;
}

void forcing_free_bnd(Model &M) {
    Forcing &f = M.f;
    for (void *q : f.bbuf) dfree(q);
    for (void *q : {(void *)f.d_bser, (void *)f.d_zfact, (void *)f.bnd_pos}) dfree(q);
    f.bbuf.clear(); f.bser.clear(); f.zfact.clear(); f.t_lo.clear(); f.t_hi.clear();
    f.d_bser = nullptr; f.d_zfact = nullptr; f.bnd_pos = nullptr;
    f.nseries = 0;
}

void forcing_free_sources(Model &M) {
    Forcing &f = M.f;
    dfree(f.src_ptr); dfree(f.src_lev); dfree(f.src_idx); dfree(f.src_q); dfree(f.rqds);
    f.src_ptr = f.src_lev = f.src_idx = nullptr;
    f.src_q = f.rqds = nullptr;
    f.nsrc = 0;
}

void forcing_free_wind(Model &M) {
    Forcing &f = M.f;
    dfree(f.wbuf[0]); dfree(f.wbuf[1]);
    f.wbuf[0] = f.wbuf[1] = nullptr;
    f.wind = SeriesDev{};
    f.wind_on = false;
    f.wind_applied = false;
}

// every series must cover dtime (the reference stops otherwise,
// iff_time_interpolate, intp_fem_file.f90:2344-2352)
int forcing_check(Model &M, double t) {
    Forcing &f = M.f;
    for (size_t b = 0; b < f.bser.size(); b++)
        if (f.bser[b].time && (t < f.t_lo[b] || t > f.t_hi[b]))
            return fail(&M, "boundary " + std::to_string(b + 1) + ": time " + std::to_string(t) +
                                " outside the records of its series", 2);
    if (f.wind_on && (t < f.w_lo || t > f.w_hi))
        return fail(&M, "wind: time " + std::to_string(t) + " outside the records of its series", 2);
    return 0;
}

}  // namespace

extern "C" {

int shyhip_set_boundary_series(shyhip_model *h, int ibc, int nrec, const double *time, const float *val, int nval,
                               int intpol, float zfact, double dtime0) {
    if (!h) return fail(nullptr, "null handle");
    if (pipe_drain(h->M)) return 1;
    Model &M = h->M;
    if (ibc < 1 || ibc > M.nbnd) return fail(&M, "set_boundary_series: boundary out of range");
    if (M.h_ibtyp[ibc] != 1) return fail(&M, "set_boundary_series: only level boundaries (ibtyp 1) take a series");
    Forcing &f = M.f;
    const int nbn = M.m.nbnd_nodes;
    if (f.bser.empty()) {
        // per-boundary tables and each node's position in its boundary
        f.bser.assign(M.nbnd, SeriesDev{});
        f.zfact.assign(M.nbnd, 1.f);
        f.t_lo.assign(M.nbnd, 0.);
        f.t_hi.assign(M.nbnd, 0.);
        f.bbuf.assign(2 * (size_t)M.nbnd, nullptr);
        std::vector<int> pos, cnt(M.nbnd + 1, 0);
        for (int k = 0; k < M.nkn; k++)
            if (M.h_ibnd[k] > 0) pos.push_back(cnt[M.h_ibnd[k]]++);
        if (dalloc(M, &f.d_bser, M.nbnd) || dalloc(M, &f.d_zfact, M.nbnd) || dalloc(M, &f.bnd_pos, nbn) ||
            upload(M, f.bnd_pos, pos.data(), pos.size()))
            return 1;
    }
    int nk = 0;
    for (int k = 0; k < M.nkn; k++) nk += M.h_ibnd[k] == ibc;
    if (nval != 1 && nval != nk)
        return fail(&M, "set_boundary_series: nval must be 1 or the boundary's node count (" + std::to_string(nk) + ")");
    SeriesDev s{};
    double lo, hi;
    if (series_setup(M, "set_boundary_series", nrec, time, intpol > 0 ? intpol : 4, dtime0, s, lo, hi)) return 1;
    if (!val) return fail(&M, "set_boundary_series: no values");
    HIPCHK(M, hipStreamSynchronize(M.stream));     // no queued step still reads the old series
    const int b = ibc - 1;
    dfree(f.bbuf[2 * b]); dfree(f.bbuf[2 * b + 1]);
    f.bbuf[2 * b] = f.bbuf[2 * b + 1] = nullptr;
    double *dt_ = nullptr;
    float *dv = nullptr;
    if (dalloc(M, &dt_, nrec) || dalloc(M, &dv, (size_t)nrec * nval)) return 1;
    f.bbuf[2 * b] = dt_;
    f.bbuf[2 * b + 1] = dv;
    if (upload(M, dt_, time, nrec) || upload(M, dv, val, (size_t)nrec * nval)) return 1;
    s.time = dt_; s.val = dv; s.nrec = nrec; s.nval = nval; s.ncomp = 1;
    if (!f.bser[b].time) f.nseries++;
    f.bser[b] = s;
    f.zfact[b] = zfact;
    f.t_lo[b] = lo;
    f.t_hi[b] = hi;
    return upload(M, f.d_bser, f.bser.data(), f.bser.size()) || upload(M, f.d_zfact, f.zfact.data(), f.zfact.size());
}

int shyhip_set_wind_series(shyhip_model *h, int nrec, const double *time, const float *wx, const float *wy, int nval,
                           int intpol, double dtime0, const shyhip_wind_params *wp) {
    if (!h) return fail(nullptr, "null handle");
    if (pipe_drain(h->M)) return 1;
    Model &M = h->M;
    HIPCHK(M, hipStreamSynchronize(M.stream));     // no queued step still reads the old series
    forcing_free_wind(M);
    if (nrec == 0) return 0;
    if (!wp) return fail(&M, "set_wind_series: no parameters");
    if (wp->iwtype != 1 && wp->iwtype != 2) return fail(&M, "set_wind_series: iwtype must be 1 (wind) or 2 (stress)");
    if (wp->itdrag < 0 || wp->itdrag > 2) return fail(&M, "set_wind_series: itdrag must be 0, 1 or 2");
    if (!(wp->roluft > 0.f) || !(wp->wsmax > 0.f)) return fail(&M, "set_wind_series: roluft and wsmax must be > 0");
    if (wp->iwtype == 1 && wp->itdrag == 0 && !(wp->dragco > 0.f))
        return fail(&M, "set_wind_series: dragco must be > 0");
    if (nval != 1 && nval != M.nkn) return fail(&M, "set_wind_series: nval must be 1 or nkn");
    if (!wx || !wy) return fail(&M, "set_wind_series: no values");
    Forcing &f = M.f;
    SeriesDev s{};
    if (series_setup(M, "set_wind_series", nrec, time, intpol > 0 ? intpol : 2, dtime0, s, f.w_lo, f.w_hi)) return 1;
    std::vector<float> v(2 * (size_t)nrec * nval);
    for (int r = 0; r < nrec; r++) {
        std::copy(wx + (size_t)r * nval, wx + (size_t)(r + 1) * nval, v.begin() + 2 * (size_t)r * nval);
        std::copy(wy + (size_t)r * nval, wy + (size_t)(r + 1) * nval, v.begin() + (2 * (size_t)r + 1) * nval);
    }
    double *dt_ = nullptr;
    float *dv = nullptr;
    if (dalloc(M, &dt_, nrec) || dalloc(M, &dv, v.size())) return 1;
    f.wbuf[0] = dt_;
    f.wbuf[1] = dv;
    if (upload(M, dt_, time, nrec) || upload(M, dv, v.data(), v.size())) return 1;
    s.time = dt_; s.val = dv; s.nrec = nrec; s.nval = nval; s.ncomp = 2;
    f.wind = s;
    f.wpar = WindPar{wp->iwtype, wp->itdrag, wp->dragco, wp->roluft, M.p.rowass, wp->wslim, wp->wsmax};
    f.wind_on = true;
    return 0;
}

int shyhip_set_baroclinic(shyhip_model *h, int ibarcl, int itemp, int isalt) {
    if (!h) return fail(nullptr, "null handle");
    if (pipe_drain(h->M)) return 1;
    Model &M = h->M;
    if (ibarcl == 0) {
        M.bcl_on = false;
        return 0;
    }
    if (ibarcl != 1) return fail(&M, "set_baroclinic: ibarcl must be 0 or 1 (prognostic T/S)");
    if (itemp < 0 || itemp >= M.ntr || isalt < 0 || isalt >= M.ntr || itemp == isalt)
        return fail(&M, "set_baroclinic: itemp/isalt must be two different tracers of the model");
    M.bcl_on = true;
    M.bcl_temp = itemp;
    M.bcl_salt = isalt;
    // barocl(0): rhov = 0, then rhoset three times (baroclinic.f90:398-404)
    if (launch_fill(M, M.s.rhov, 0., (size_t)M.nlvdi * M.nkn)) return 1;
    if (launch_rhoset(M, 3, M.stream)) return 1;
    HIPCHK(M, hipStreamSynchronize(M.stream));
    return 0;
}

int shyhip_rhoset(shyhip_model *h, int niter) {
    if (!h) return fail(nullptr, "null handle");
    if (pipe_drain(h->M)) return 1;
    Model &M = h->M;
    if (!M.bcl_on) return fail(&M, "rhoset: call shyhip_set_baroclinic first");
    if (niter < 1) return fail(&M, "rhoset: niter must be >= 1");
    return launch_rhoset(M, niter, M.stream);
}

int shyhip_forcing(shyhip_model *h, double dtime) {
    if (!h) return fail(nullptr, "null handle");
    Model &M = h->M;
    if (forcing_check(M, dtime)) return 2;
    return launch_forcing(M, dtime);
}

static int check_dev_err(Model &M, const char *where);

int shyhip_set_sources(shyhip_model *h, int nsrc, const int *node, const int *levmin, const int *levmax) {
    if (!h) return fail(nullptr, "null handle");
    if (pipe_drain(h->M)) return 1;
    Model &M = h->M;
    if (nsrc < 0 || (nsrc > 0 && (!node || !levmin || !levmax))) return fail(&M, "set_sources: bad arguments");
    for (int i = 0; i < nsrc; i++)
        if (node[i] < 1 || node[i] > M.nkn)
            return fail(&M, "set_sources: node " + std::to_string(node[i]) + " out of range");
    forcing_free_sources(M);
    Forcing &f = M.f;
    if (dalloc(M, &f.rqds, (size_t)M.nkn)) return 1;
    if (nsrc == 0) return 0;
    // the sources of each node, in the caller's order (a counting sort)
    std::vector<int> ptr(M.nkn + 1, 0), idx(nsrc), lev(2 * (size_t)nsrc);
    for (int i = 0; i < nsrc; i++) ptr[node[i]]++;
    for (int k = 0; k < M.nkn; k++) ptr[k + 1] += ptr[k];
    std::vector<int> fill(ptr.begin(), ptr.end() - 1);
    for (int i = 0; i < nsrc; i++) {
        const int p = fill[node[i] - 1]++;
        idx[p] = i;
        lev[2 * p] = levmin[i];
        lev[2 * p + 1] = levmax[i];
    }
    if (dalloc(M, &f.src_ptr, ptr.size()) || dalloc(M, &f.src_lev, lev.size()) || dalloc(M, &f.src_idx, idx.size()) ||
        dalloc(M, &f.src_q, (size_t)nsrc))
        return 1;
    if (upload(M, f.src_ptr, ptr.data(), ptr.size()) || upload(M, f.src_lev, lev.data(), lev.size()) ||
        upload(M, f.src_idx, idx.data(), idx.size()))
        return 1;
    f.nsrc = nsrc;
    return 0;
}

int shyhip_set_mass_flux(shyhip_model *h, const float *rflux, const float *rqdsv) {
    if (!h) return fail(nullptr, "null handle");
    if (pipe_drain(h->M)) return 1;
    Model &M = h->M;
    Forcing &f = M.f;
    if (!f.rqds) return fail(&M, "set_mass_flux: call shyhip_set_sources first");
    if (f.nsrc > 0 && !rflux) return fail(&M, "set_mass_flux: rflux is needed for the sources");
    if (f.nsrc > 0 && upload(M, f.src_q, rflux, (size_t)f.nsrc)) return 1;
    if (rqdsv && upload(M, f.rqds, rqdsv, (size_t)M.nkn)) return 1;
    M.has_mflux = true;
    if (launch_mass_flux(M, rqdsv != nullptr)) return 1;
    return check_dev_err(M, "set_mass_flux");
}

int shyhip_set_dry(shyhip_model *h) {
    if (!h) return fail(nullptr, "null handle");
    if (pipe_drain(h->M)) return 1;
    Model &M = h->M;
    const int bnew = M.dtime != M.dtime_old;
    if (launch_set_dry(M, 2, bnew, nullptr)) return 1;
    M.dtime_old = M.dtime;
    if (launch_set_dry(M, 3, 0, nullptr)) return 1;
    return 0;
}

int shyhip_copy_uvz(shyhip_model *h) {
    if (!h) return fail(nullptr, "null handle");
    if (pipe_drain(h->M)) return 1;
    return launch_copy_old(h->M);
}

int shyhip_hydro_transports(shyhip_model *h) {
    if (!h) return fail(nullptr, "null handle");
    return pipe_drain(h->M) || launch_hydro_intern(h->M);
}
int shyhip_hydro_zeta(shyhip_model *h) {
    if (!h) return fail(nullptr, "null handle");
    return pipe_drain(h->M) || launch_zeta_assemble(h->M);
}
int shyhip_hydro_transports_final(shyhip_model *h) {
    if (!h) return fail(nullptr, "null handle");
    return pipe_drain(h->M) || launch_transports_final(h->M);
}
int shyhip_hydro_vertical(shyhip_model *h) {
    if (!h) return fail(nullptr, "null handle");
    return pipe_drain(h->M) || launch_hydro_vertical(h->M);
}

int shyhip_system_solve_device(shyhip_model *h, int *iters) {
    if (!h) return fail(nullptr, "null handle");
    if (pipe_drain(h->M)) return 1;
    Model &M = h->M;
    int ref = 0;
    const int rc = solver_solve(M, iters, &ref);
    M.last_refine = ref;
    return rc;
}

// device error word d_err[0]: set by kernels (first error wins), cleared
// when it is reported
static int report_dev_err(Model &M, int e, const char *where) {
    if (!e) return 0;
    HIPCHK(M, hipMemsetAsync(M.d_err, 0, sizeof(int), M.stream));
    HIPCHK(M, hipStreamSynchronize(M.stream));
    static const char *names[] = {"", "setnod: open boundary node out of the system", "", "setuvd: z or u/v computation",
                                  "setdepth: non-positive layer thickness", "check_volume: negative volume", "",
                                  "conz3d: zero diagonal in a column", "penta: singular vertical column",
                                  "meteo_convert_wind_data: wind speed above wsmax",
                                  "set_mass_flux: source layers out of bounds or zero volume"};
    return fail(&M, std::string(where) + ": " + (e >= 0 && e <= 10 ? names[e] : "device error"), 3);
}

static int check_dev_err(Model &M, const char *where) {
    HIPCHK(M, hipMemcpyAsync(M.h_pinned, M.d_err, sizeof(int), hipMemcpyDeviceToHost, M.stream));
    HIPCHK(M, hipStreamSynchronize(M.stream));
    return report_dev_err(M, M.h_pinned[0], where);
}

// hydro; with final_check the device error word is read at the end (a
// caller that synchronises later anyway, shyhip_step, checks it there)
static int hydro_impl(Model &M, double dtime, bool final_check) {
    M.dtime = dtime;
    int iloop = 0, changed = 0, its_total = 0;
    timer_begin(M, T_INTERN);
    if (launch_hydro_intern(M)) return 1;          // hydro_transports (invariant over the wet/dry loop)
    timer_end(M);
    for (;;) {
        iloop++;
        if (iloop > 100) return fail(&M, "hydro: wet/dry loop did not settle", 3);
        timer_begin(M, T_SETNOD);
        if (launch_setnod(M)) return 1;             // setnod + adjust_mass_flux
        timer_end(M);
        timer_begin(M, T_ZETA);
        if (launch_zeta_assemble(M)) return 1;      // system_init + hydro_zeta + system_add_rhs
        timer_end(M);
        int its = 0;
        // the solve, then system_get + setweg(1) behind it: one host
        // read-back returns the solver flags, the device error word and the
        // number of elements whose wet/dry state changed
        const int bnew = M.dtime != M.dtime_old;
        timer_begin(M, T_SOLVE);
        int ref = 0;
        const int rc = solver_solve(M, &its, &ref, bnew);
        M.last_refine = ref;
        timer_end(M);
        its_total += its;
        if (rc) return rc;
        changed = M.h_pinned[1];
        M.dtime_old = M.dtime;
        if (changed == 0) break;
    }
    M.last_nloop = iloop;
    M.last_iters = its_total;
    if (report_dev_err(M, M.h_pinned[0], "hydro loop")) return 3;   // read with the setweg counter
    // one element kernel: transports_final, setzev + setuvd + baro2l, the
    // element depths of make_new_depth, compute_velocities (ulnv/vlnv,
    // unv/vnv) -- then the node depths (+ check_volume) and hydro_vertical
    timer_begin(M, T_FINAL);
    if (launch_elem_after_solve(M)) return 1;
    timer_end(M);
    timer_begin(M, T_DEPTH);
    if (launch_setdepth_node(M)) return 1;
    timer_end(M);
    timer_begin(M, T_VERT);
    if (launch_hydro_vertical(M)) return 1;
    timer_end(M);
    return final_check ? check_dev_err(M, "hydro") : 0;
}

int shyhip_hydro(shyhip_model *h, double dtime) {
    if (!h) return fail(nullptr, "null handle");
    if (pipe_drain(h->M)) return 1;
    return hydro_impl(h->M, dtime, true);
}

// The horizontal diffusion weights for the current idtype, built when a
// transport or stability check first runs with rkpar > 0 (host_diffweight,
// diffusion_admin.f90:182-340; the reference stops on weights that break
// its checks, and so does this)
static int ensure_wdif(Model &M) {
    if (!(M.p.rkpar > 0.f)) return 0;
    if (M.m.wdif && M.m.wdif_idtype == M.p.idtype) return 0;
    if (M.h_ev.size() != 19 * (size_t)M.nel) return fail(&M, "horizontal diffusion needs the basin (shyhip_set_basin)");
    std::vector<double> w;
    const int bad = host_diffweight(M.nel, M.h_ev.data(), M.p.idtype, w);
    if (bad)
        return fail(&M, "diffweight: error in the weights of element " + std::to_string(bad) + " (idtype " +
                            std::to_string(M.p.idtype) + ")");
    if (!M.m.wdif && dalloc_s(M, &M.m.wdif, w.size())) return 1;
    if (upload_s(M, M.m.wdif, w.data(), w.size())) return 1;
    M.m.wdif_idtype = M.p.idtype;
    return 0;
}

// istot of scal3sh: the caller's (1..1000), or for istot <= 0 the
// reference's choice, scalar_stability with rstol = 1 (the STR default),
// stopping above istot_max = 1000 (concentration.f90:678-700)
static int resolve_istot(Model &M, int *istot) {
    M.mem_istot_on = false;
    if (*istot > 1000) return fail(&M, "istot must be at most 1000");
    if (*istot >= 1) {
        M.last_istot = *istot;
        return 0;
    }
    // the members of an ensemble choose their own istot, as each does alone;
    // the step runs the largest, and members past their own last substep
    // are left alone by the kernels
    const int nmem = M.nmem_dev > 1 ? M.nmem_dev : 1;
    if (nmem > 1) M.cur_mem_istot = M.d_mem_istot + (size_t)(M.pipe_k & 1) * nmem;
    std::vector<int> ist(nmem);
    std::vector<double> ri(nmem);
    if (ensure_wdif(M) || launch_scalar_stability(M, 1.0f, ist.data(), ri.data())) return 1;
    int mx = 1, mn = 1 << 30;
    for (int m = 0; m < nmem; m++) {
        if (ist[m] > 1000 || ist[m] < 1)
            return fail(&M, "scal3sh: istot index too high (" + std::to_string(ist[m]) + ", sindex " +
                                std::to_string(ri[m]) + (nmem > 1 ? ", member " + std::to_string(m) : "") + ")", 3);
        mx = std::max(mx, ist[m]);
        mn = std::min(mn, ist[m]);
    }
    *istot = mx;
    M.mem_istot = ist;
    M.mem_istot_on = nmem > 1 && mn != mx;
    M.last_istot = mx;
    return 0;
}

int shyhip_member_istot(shyhip_model *h, int nmem, int *istot) {
    if (!h || !istot) return fail(nullptr, "null argument");
    const Model &M = h->M;
    const int n = M.nmem_dev > 1 ? M.nmem_dev : 1;
    if (nmem != n) return fail(&h->M, "member_istot: the model has " + std::to_string(n) + " members");
    for (int m = 0; m < n; m++) istot[m] = (M.mem_istot.size() == (size_t)n) ? M.mem_istot[m] : M.last_istot;
    return 0;
}

int shyhip_masscont_noalloc(int dtype, int nkn, const int *ilhkv, const void *hdk, const void *areakv,
                            const void *conc, double *out, void *stream) {
    std::string err;
    if (noalloc_masscont(dtype, nkn, ilhkv, hdk, areakv, conc, out, (hipStream_t)stream, err))
        return fail(nullptr, "masscont_noalloc: " + err);
    return 0;
}

int shyhip_get_sizes(const shyhip_model *h, int *nkn, int *nel, int *nlvdi, int *nlv, int *ntracer) {
    if (!h) return fail(nullptr, "null handle");
    const Model &M = h->M;
    if (nkn) *nkn = M.nkn;
    if (nel) *nel = M.nel;
    if (nlvdi) *nlvdi = M.nlvdi;
    if (nlv) *nlv = M.nlv;
    if (ntracer) *ntracer = M.ntr;
    return 0;
}

int shyhip_get_retries(shyhip_model *h, long *retries) {
    if (!h || !retries) return fail(nullptr, "null argument");
    *retries = h->M.solver_retries;
    return 0;
}

int shyhip_get_istot(shyhip_model *h, int *istot) {
    if (!h || !istot) return fail(nullptr, "null argument");
    *istot = h->M.last_istot;
    return 0;
}

int shyhip_scalar_stability(shyhip_model *h, float rstol, int *istot, double *rindex) {
    if (!h || !istot || !rindex) return fail(nullptr, "null argument");
    Model &M = h->M;
    if (pipe_drain(M)) return 1;
    if (!(rstol > 0.f)) return fail(&M, "rstol must be positive");
    const int nmem = M.nmem_dev > 1 ? M.nmem_dev : 1;
    if (nmem > 1) M.cur_mem_istot = M.d_mem_istot + (size_t)(M.pipe_k & 1) * nmem;
    std::vector<int> ist(nmem);
    std::vector<double> ri(nmem);
    if (ensure_wdif(M) || launch_scalar_stability(M, rstol, ist.data(), ri.data())) return 1;
    *istot = *std::max_element(ist.begin(), ist.end());
    *rindex = *std::max_element(ri.begin(), ri.end());
    return 0;
}

// scal3sh of tracer itr on stream st with conz3d scratch hbuf
static int scal3sh_on(Model &M, const TracerSet &ts, int istot, hipStream_t st, void *hbuf, void *rec, void *bscal) {
    for (int q = 0; q < ts.n; q++)
        if (ts.itr[q] < 0 || ts.itr[q] >= M.ntr) return fail(&M, "tracer index out of range");
    if (istot < 1 || istot > 1000) return fail(&M, "istot must be in 1..1000");
    if (M.p.itvd == 2 && !M.tvd_ready) return fail(&M, "itvd = 2 needs shyhip_tvd_init");
    if (ensure_wdif(M)) return 1;
    auto substeps = [&]() -> int {
        for (int isact = 1; isact <= istot; isact++) {
            // the step's element-layer records before the first substep
            if (launch_conz3d(M, ts, istot, isact, st, hbuf, rec, isact == 1)) return 1;
            timer_begin(M, T_BNDO);
            if (launch_bndo(M, ts, st, isact, bscal)) return 1;       // bndo_setbc
            timer_end(M);
        }
        return 0;
    };
    if (!M.use_graphs || M.timing) return substeps();
    std::string key("scal3sh");
    conz3d_key(M, ts, istot, hbuf, rec, bscal, key);
    return graph_run(M, st, key, substeps);
}

// The tracers of a step are transported in groups: pairs (0,1), (2,3), ...
// and a last single one; each group's launches carry both its tracers.
// Slot 0 (temperature) is not clipped, the others are (scal3sh's
// 'temp' / non-'temp' rule).
static int ngroups(const Model &M) { return (M.ntr + 1) / 2; }
static TracerSet group_of(const Model &M, int g) {
    TracerSet ts;
    const int t0 = 2 * g;
    ts.n = t0 + 1 < M.ntr ? 2 : 1;
    for (int q = 0; q < ts.n; q++) {
        ts.itr[q] = t0 + q;
        ts.clip[q] = t0 + q == 0 ? 0 : 1;
    }
    return ts;
}
// the element-term scratch of each group (sized for its tracers; created once)
static int group_buffers(Model &M) {
    if ((int)M.grp_hbuf.size() != ngroups(M))
        for (int g = 0; g < ngroups(M); g++) {
            void *hb = nullptr;
            if (dalloc_s(M, &hb, conz_hbuf_elems(M, group_of(M, g).n))) return 1;
            M.grp_hbuf.push_back(hb);
        }
    if ((int)M.grp_rec.size() != ngroups(M))
        for (int g = 0; g < ngroups(M); g++) {
            char *rb = nullptr;
            if (dalloc(M, &rb, conz_record_bytes(M))) return 1;
            M.grp_rec.push_back(rb);
        }
    if ((int)M.grp_bscal.size() != ngroups(M))
        for (int g = 0; g < ngroups(M); g++) {
            void *bs = nullptr;
            if (dalloc_s(M, &bs, (size_t)std::max(1, M.m.nbndo) * M.nlvdi)) return 1;
            M.grp_bscal.push_back(bs);
        }
    return 0;
}

int shyhip_conz3d(shyhip_model *h, int itr, const float *cbound, int istot, int isact) {
    if (!h) return fail(nullptr, "null handle");
    if (pipe_drain(h->M)) return 1;
    Model &M = h->M;
    if (itr < 0 || itr >= M.ntr) return fail(&M, "tracer index out of range");
    if (istot < 1 || istot > 1000 || isact < 1 || isact > istot) return fail(&M, "conz3d: need 1 <= isact <= istot <= 1000");
    if (M.p.itvd == 2 && !M.tvd_ready) return fail(&M, "itvd = 2 needs shyhip_tvd_init");
    if (ensure_wdif(M)) return 1;
    void *cb = nullptr;
    if (cbound) {
        if (!M.d_cbound && dalloc_s(M, &M.d_cbound, (size_t)M.nlvdi * M.nkn)) return 1;
        std::vector<double> t((size_t)M.nlvdi * M.nkn);
        for (int k = 0; k < M.nkn; k++)
            for (int l = 0; l < M.nlvdi; l++) t[(size_t)l * M.nkn + k] = cbound[(size_t)k * M.nlvdi + l];
        if (upload_s(M, M.d_cbound, t.data(), t.size())) return 1;
        cb = M.d_cbound;
    }
    M.mem_istot_on = false;
    if (!M.rec1 && dalloc(M, (char **)&M.rec1, conz_record_bytes(M))) return 1;
    if (launch_conz3d(M, itr, istot, isact, 0, M.stream, M.s.hbuf, M.rec1, cb)) return 1;
    return check_dev_err(M, "conz3d");
}

int shyhip_bndo_setbc(shyhip_model *h, int itr) {
    if (!h) return fail(nullptr, "null handle");
    if (pipe_drain(h->M)) return 1;
    Model &M = h->M;
    if (itr < 0 || itr >= M.ntr) return fail(&M, "tracer index out of range");
    M.mem_istot_on = false;
    if (launch_bndo(M, itr, M.stream)) return 1;
    return check_dev_err(M, "bndo_setbc");
}

int shyhip_set_tracer_terms(shyhip_model *h, int itr, float robs, float wsink, float rload) {
    if (!h) return fail(nullptr, "null handle");
    Model &M = h->M;
    if (pipe_drain(M)) return 1;
    if (itr < 0 || itr >= M.ntr) return fail(&M, "tracer index out of range");
    if (!std::isfinite(robs) || !std::isfinite(wsink) || !std::isfinite(rload))
        return fail(&M, "tracer terms must be finite");
    M.tr_robs[itr] = robs;
    M.tr_wsink[itr] = wsink;
    M.tr_rload[itr] = rload;
    return 0;
}

int shyhip_scal3sh(shyhip_model *h, int itr, int istot, int clip_small) {
    if (!h) return fail(nullptr, "null handle");
    if (pipe_drain(h->M)) return 1;
    if (itr < 0 || itr >= h->M.ntr) return fail(&h->M, "tracer index out of range");
    h->M.terms_tracer = tracer_has_terms(h->M, itr) ? itr : -1;
    const int rc = resolve_istot(h->M, &istot);
    h->M.terms_tracer = -1;
    if (rc) return 1;
    TracerSet ts;
    ts.itr[0] = itr;
    ts.clip[0] = clip_small;
    if (!h->M.bscal1 && dalloc_s(h->M, &h->M.bscal1, (size_t)std::max(1, h->M.m.nbndo) * h->M.nlvdi)) return 1;
    if (!h->M.rec1 && dalloc(h->M, (char **)&h->M.rec1, conz_record_bytes(h->M))) return 1;
    return scal3sh_on(h->M, ts, istot, h->M.stream, h->M.s.hbuf, h->M.rec1, h->M.bscal1);
}

// streams and events of the tracer groups after the first (created once)
static int tracer_streams(Model &M) {
    if (group_buffers(M)) return 1;
    if (!M.tr_stream.empty() || ngroups(M) < 2) return 0;
    if (!M.ev_hydro) HIPCHK(M, hipEventCreateWithFlags(&M.ev_hydro, hipEventDisableTiming));
    for (int g = 1; g < ngroups(M); g++) {
        hipStream_t st = nullptr;
        hipEvent_t ev = nullptr;
        HIPCHK(M, hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        M.tr_stream.push_back(st);
        HIPCHK(M, hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        M.tr_done.push_back(ev);
    }
    return 0;
}

// buffers, streams and events of the pipelined transport (created once)
static int pipe_setup(Model &M) {
    if (M.tr0_stream) return 0;
    if (tracer_streams(M)) return 1;
    if (!M.ev_hydro) HIPCHK(M, hipEventCreateWithFlags(&M.ev_hydro, hipEventDisableTiming));
    HIPCHK(M, hipStreamCreateWithFlags(&M.tr0_stream, hipStreamNonBlocking));
    for (int p = 0; p < 2; p++)
        for (int g = 0; g < ngroups(M); g++) {
            hipEvent_t ev = nullptr;
            HIPCHK(M, hipEventCreateWithFlags(&ev, hipEventDisableTiming));
            M.pipe_ev[p].push_back(ev);
        }
    const size_t L = M.nlvdi, ne = M.nel, nk = M.nkn;
    int rc = 0;
    rc |= dalloc_s(M, &M.spare_ut, L * ne) || dalloc_s(M, &M.spare_vt, L * ne) || dalloc_s(M, &M.spare_hde, L * ne);
    rc |= dalloc_s(M, &M.spare_wl, (L + 1) * nk) || dalloc_s(M, &M.spare_hdk, L * nk);
    rc |= dalloc_s(M, &M.uln_b, L * ne) || dalloc_s(M, &M.vln_b, L * ne);
    rc |= dalloc(M, &M.iweg_tr[0], ne) || dalloc(M, &M.iweg_tr[1], ne);
    return rc;
}


int shyhip_step(shyhip_model *h, double dtime, int istot) {
    if (!h) return fail(nullptr, "null handle");
    Model &M = h->M;
    // the forcing series must cover dtime; checked before the step starts
    if ((M.f.nseries > 0 || M.f.wind_on) && forcing_check(M, dtime)) return 2;
    // tracers with their own terms (nudging, sinking, loads) are each
    // transported alone with their own substeps, as scal3sh does per tracer
    bool terms = false;
    for (int t = 0; t < M.ntr; t++) terms = terms || tracer_has_terms(M, t);
    // the transport overlaps the next step's hydro unless that hydro needs
    // this step's density (baroclinic runs)
    const bool pipe = M.deferred && !M.timing && !M.has_mflux && M.ntr >= 1 && !M.bcl_on && !terms;
    int par = 0;
    if (!pipe) {
        if (pipe_drain(M)) return 1;
    } else {
        if (pipe_setup(M)) return 1;
        par = M.pipe_k & 1;
        // the transport of two steps back read the buffers this step rewrites
        if (M.pipe_rec[par])
            for (hipEvent_t ev : M.pipe_ev[par]) HIPCHK(M, hipStreamWaitEvent(M.stream, ev, 0));
    }
    M.iweg_snap = pipe ? M.iweg_tr[par] : nullptr;
    // set_dry (setweg 2 and 3) and copy_uvz/copy_depth in one launch
    timer_begin(M, T_SETDRY);
    const int bnew = M.dtime != M.dtime_old;
    if (launch_step_start(M, bnew, pipe)) return 1;
    M.dtime_old = M.dtime;
    timer_end(M);
    // sp111(2): boundary levels and wind of the installed series at dtime
    if ((M.f.nseries > 0 || M.f.wind_on) && launch_forcing(M, dtime)) return 1;
    if (hydro_impl(M, dtime, false)) return M.err.empty() ? 1 : 3;
    M.iweg_snap = nullptr;
    if (terms) {
        if (tracer_streams(M)) return 1;
        for (int t = 0; t < M.ntr; t++) {
            int ist = istot;
            M.terms_tracer = tracer_has_terms(M, t) ? t : -1;
            const int rc = resolve_istot(M, &ist);
            M.terms_tracer = -1;
            if (rc) return 1;
            TracerSet ts;
            ts.itr[0] = t;
            ts.clip[0] = t == 0 ? 0 : 1;
            if (scal3sh_on(M, ts, ist, M.stream, M.grp_hbuf[t / 2], M.grp_rec[t / 2], M.grp_bscal[t / 2])) return 1;
        }
        if (launch_rhoset(M, 1, M.stream)) return 1;
        if (timer_flush(M)) return 1;
        if (M.deferred && !M.timing) return 0;
        return check_dev_err(M, "step");
    }
    // one istot for all tracers: conzstab does not depend on the tracer
    if (M.ntr > 0 && resolve_istot(M, &istot)) return 1;
    if (pipe) {
        // every tracer group on its own stream after this step's hydro; the model's
        // stream goes on with the next step
        HIPCHK(M, hipEventRecord(M.ev_hydro, M.stream));
        M.iweg_conz = M.iweg_tr[par];
        for (int g = 0; g < ngroups(M); g++) {
            hipStream_t st = g == 0 ? M.tr0_stream : M.tr_stream[g - 1];
            HIPCHK(M, hipStreamWaitEvent(st, M.ev_hydro, 0));
            if (scal3sh_on(M, group_of(M, g), istot, st, M.grp_hbuf[g], M.grp_rec[g], M.grp_bscal[g])) { M.iweg_conz = nullptr; return 1; }
            HIPCHK(M, hipEventRecord(M.pipe_ev[par][g], st));
        }
        M.iweg_conz = nullptr;
        M.pipe_rec[par] = true;
        M.pipe_live = true;
        M.pipe_k++;
        return 0;
    }
    if (M.ntr > 0) {
        if (tracer_streams(M)) return 1;
        if (ngroups(M) > 1 && !M.timing) {
            // the groups are transported concurrently: each after hydro, on
            // its own stream, joined back into the model's stream
            HIPCHK(M, hipEventRecord(M.ev_hydro, M.stream));
            for (int g = 1; g < ngroups(M); g++) HIPCHK(M, hipStreamWaitEvent(M.tr_stream[g - 1], M.ev_hydro, 0));
            if (scal3sh_on(M, group_of(M, 0), istot, M.stream, M.grp_hbuf[0], M.grp_rec[0], M.grp_bscal[0])) return 1;
            for (int g = 1; g < ngroups(M); g++) {
                if (scal3sh_on(M, group_of(M, g), istot, M.tr_stream[g - 1], M.grp_hbuf[g], M.grp_rec[g], M.grp_bscal[g])) return 1;
                HIPCHK(M, hipEventRecord(M.tr_done[g - 1], M.tr_stream[g - 1]));
                HIPCHK(M, hipStreamWaitEvent(M.stream, M.tr_done[g - 1], 0));
            }
        } else {
            for (int g = 0; g < ngroups(M); g++)
                if (scal3sh_on(M, group_of(M, g), istot, M.stream, M.grp_hbuf[g], M.grp_rec[g], M.grp_bscal[g])) return 1;
        }
    }
    // barocl: the density of the transported T/S (rhoset_shell(1),
    // baroclinic.f90:542-548) for the next step's baroclinic term
    if (launch_rhoset(M, 1, M.stream)) return 1;
    if (timer_flush(M)) return 1;
    if (M.deferred && !M.timing) return 0;
    return check_dev_err(M, "step");
}

int shyhip_set_deferred_checks(shyhip_model *h, int on) {
    if (!h) return fail(nullptr, "null handle");
    if (pipe_drain(h->M)) return 1;
    h->M.deferred = on != 0;
    return 0;
}

int shyhip_set_graphs(shyhip_model *h, int on) {
    if (!h) return fail(nullptr, "null handle");
    if (pipe_drain(h->M)) return 1;
    HIPCHK(h->M, hipStreamSynchronize(h->M.stream));
    h->M.use_graphs = on != 0;
    return 0;
}

int shyhip_graph_stats(shyhip_model *h, long *builds, long *replays) {
    if (!h || !builds || !replays) return fail(h ? &h->M : nullptr, "null argument");
    *builds = h->M.graph_builds;
    *replays = h->M.graph_replays;
    return 0;
}

int shyhip_set_timing(shyhip_model *h, int on) {
    if (!h) return fail(nullptr, "null handle");
    if (pipe_drain(h->M)) return 1;
    Model &M = h->M;
    if (timer_flush(M)) return 1;
    M.timing = on != 0;
    for (int i = 0; i < T_NTIMERS; i++) { M.t_ms[i] = 0.; M.t_cnt[i] = 0; }
    return 0;
}

int shyhip_get_timing(shyhip_model *h, int id, const char **name, double *total_ms, long *count) {
    static const char *names[T_NTIMERS] = {"set_dry", "copy_uvz", "hydro_intern", "setnod", "zeta_assemble",
                                           "zeta_solve", "system_get_setweg", "transports_final", "setuvd_baro2l",
                                           "setdepth", "hydro_vertical", "ttov_uvint", "conz3d_elem", "conz3d_node",
                                           "bndo_setbc", "bilu_factor", "bicg_a", "bicg_b", "conz3d_prep",
                                           "tvd_grad"};
    if (!h) return fail(nullptr, "null handle");
    if (id < 0 || id >= T_NTIMERS) return fail(&h->M, "timer id out of range");
    if (timer_flush(h->M)) return 1;
    if (name) *name = names[id];
    if (total_ms) *total_ms = h->M.t_ms[id];
    if (count) *count = h->M.t_cnt[id];
    return 0;
}

int shyhip_sync(shyhip_model *h) {
    if (!h) return fail(nullptr, "null handle");
    if (pipe_drain(h->M)) return 1;
    HIPCHK(h->M, hipStreamSynchronize(h->M.stream));
    return check_dev_err(h->M, "sync");
}

int shyhip_volume(shyhip_model *h, int mode, double *vol) {
    if (!h) return fail(nullptr, "null handle");
    if (pipe_drain(h->M)) return 1;
    return launch_volume(h->M, mode, vol);
}

int shyhip_tracer_mass(shyhip_model *h, int itr, double *mass) {
    if (!h) return fail(nullptr, "null handle");
    if (pipe_drain(h->M)) return 1;
    if (itr < 0 || itr >= h->M.ntr) return fail(&h->M, "tracer index out of range");
    return launch_tracer_mass(h->M, itr, mass);
}

int shyhip_stats(shyhip_model *h, int *nloop, int *solver_iters, int *refinements) {
    if (!h) return fail(nullptr, "null handle");
    if (nloop) *nloop = h->M.last_nloop;
    if (solver_iters) *solver_iters = h->M.last_iters;
    if (refinements) *refinements = h->M.last_refine;
    return 0;
}

int shyhip_get_pattern(shyhip_model *h, int *nnz, int *ntg, int *cols, int *ncolor, int *color) {
    if (!h) return fail(nullptr, "null handle");
    if (pipe_drain(h->M)) return 1;
    Model &M = h->M;
    if (nnz) *nnz = M.nnz;
    if (ncolor) *ncolor = M.ncolor;
    if (ntg) std::copy(M.h_rowptr.begin(), M.h_rowptr.end(), ntg);
    if (cols) std::copy(M.h_cols.begin(), M.h_cols.end(), cols);
    if (color) std::copy(M.h_color.begin(), M.h_color.end(), color);
    return 0;
}

// ---- zeta-system plugin (simsys_spkaux.f90 interface) ----------------------
int shyhip_system_init(shyhip_model *h) {
    if (!h) return fail(nullptr, "null handle");
    if (pipe_drain(h->M)) return 1;
    Model &M = h->M;
    M.h_coo.assign(M.nnz, 0.);
    M.h_rvec.assign(M.nkn, 0.);
    M.h_raux.assign(M.nkn, 0.);
    return 0;
}

int shyhip_system_set_explicit(shyhip_model *h) {
    if (!h) return fail(nullptr, "null handle");
    h->M.sys_explicit = true;
    return 0;
}

int shyhip_system_assemble(shyhip_model *h, int ie, const int *kn, const float *hia, const float *hik) {
    if (!h) return fail(nullptr, "null handle");
    if (pipe_drain(h->M)) return 1;
    Model &M = h->M;
    if (ie < 1 || ie > M.nel) return fail(&M, "system_assemble: element out of range");
    if (M.h_coo.size() != (size_t)M.nnz) return fail(&M, "system_assemble before system_init");
    const int e = ie - 1;
    if (M.sys_explicit) {
        // the lumped system: diagonal and right-hand side only; an
        // off-diagonal entry is the reference's error stop (:561-592)
        for (int i = 0; i < 3; i++) {
            const int k = kn[i] - 1;
            if (k < 0 || k >= M.nkn) return fail(&M, "system_assemble: node out of range");
            M.h_raux[k] = M.h_raux[k] + (double)hia[i + 3 * i];
            M.h_rvec[k] = M.h_rvec[k] + (double)hik[i];
            for (int j = 0; j < 3; j++)
                if (i != j && hia[i + 3 * j] != 0.f)
                    return fail(&M, "system_assemble: non diag elems /= 0 (element " + std::to_string(ie) + ")");
        }
        return 0;
    }
    for (int i = 0; i < 3; i++) {
        const int k = kn[i] - 1;
        if (k != M.h_nen3v[3 * e + i]) return fail(&M, "system_assemble: kn does not match nen3v");
        for (int j = 0; j < 3; j++) {
            const int slot = M.h_ijp[9 * (size_t)e + 3 * i + j];
            M.h_coo[slot] = M.h_coo[slot] + (double)hia[i + 3 * j];   // hia(i,j), column-major
        }
        M.h_rvec[k] = M.h_rvec[k] + (double)hik[i];
    }
    return 0;
}

int shyhip_system_add_rhs(shyhip_model *h, float dt, int n, const float *array) {
    if (!h) return fail(nullptr, "null handle");
    if (pipe_drain(h->M)) return 1;
    Model &M = h->M;
    if (n != M.nkn || M.h_rvec.size() != (size_t)n) return fail(&M, "system_add_rhs: size mismatch");
    for (int k = 0; k < n; k++) M.h_rvec[k] = M.h_rvec[k] + (double)(dt * array[k]);
    return 0;
}

int shyhip_system_solve(shyhip_model *h, int n, const float *z) {
    if (!h) return fail(nullptr, "null handle");
    if (pipe_drain(h->M)) return 1;
    Model &M = h->M;
    if (n != M.nkn || M.h_coo.size() != (size_t)M.nnz) return fail(&M, "system_solve: size mismatch or no assembly");
    if (M.sys_explicit) {
        // explicit: rvec2d / raux2d (simsys_spkaux.f90:475-479)
        std::vector<double> x(n);
        for (int k = 0; k < n; k++) x[k] = M.h_rvec[k] / M.h_raux[k];
        return upload(M, M.s.x, x.data(), n);
    }
    M.sv.fed = false;      // the system comes from the host: the solve prepares its own inputs
    if (upload(M, M.s.aval, M.h_coo.data(), M.nnz) || upload(M, M.s.rhs, M.h_rvec.data(), n)) return 1;
    std::vector<double> zz(z, z + n);
    if (upload_s(M, M.s.zn, zz.data(), n)) return 1;     // initial guess
    int its = 0;
    return shyhip_system_solve_device(h, &its);
}

int shyhip_system_get(shyhip_model *h, int n, float *z) {
    if (!h) return fail(nullptr, "null handle");
    if (pipe_drain(h->M)) return 1;
    Model &M = h->M;
    if (n != M.nkn) return fail(&M, "system_get: size mismatch");
    std::vector<double> x(n);
    HIPCHK(M, hipStreamSynchronize(M.stream));
    HIPCHK(M, hipMemcpyAsync(x.data(), M.s.x, n * sizeof(double), hipMemcpyDeviceToHost, M.stream));
    HIPCHK(M, hipStreamSynchronize(M.stream));
    for (int k = 0; k < n; k++) z[k] = (float)x[k];
    return 0;
}

}  // extern "C"
