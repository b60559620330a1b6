/*
 * shyoracle.h -- CPU oracle for the SHYFEM semi-implicit hydro step and the
 * 3D scalar transport.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm (SHYFEM 8_1_1, Fortran), used by tests/, __graft_entry__.smoke()
 * and the cpu_baseline leg of bench.py as the checker.  The product library
 * (libshyhip.so) never links, loads or calls it.
 *
 * Build: the same source is compiled twice,
 *   liboracle_f32.so : OC_REAL=float   -- the reference's own precision policy
 *                      (f32 state, f64 element locals, f64 solve)
 *   liboracle_f64.so : OC_REAL=double  -- the FP64 oracle (all state in f64)
 *
 * Array layouts follow the reference (Fortran column-major, layer fastest):
 *   element 3D  a(l,ie)      -> a[(l-1) + nlvdi*ie]      l = 1..nlvdi
 *   node 3D     a(l,k)       -> a[(l-1) + nlvdi*k]
 *   interfaces  w(0:nlvdi,k) -> w[l + (nlvdi+1)*k]       l = 0..nlvdi
 *   vertex      z(ii,ie)     -> z[(ii-1) + 3*ie]
 *   ev(i,ie)                 -> ev[(i-1) + 19*ie]
 *   ddxv(j,ie)  j=1..2*nlvdi -> ddxv[(j-1) + 2*nlvdi*ie]
 * Node and element indices are 0-based in C.
 *
 * "Parity unpinned" functions (see the per-function comments): none of the
 * reference's own tests hold vectors for them; they are pinned by invariants
 * and closed forms in tests/ (see the test files named in each comment).
 */
#ifndef SHYORACLE_H
#define SHYORACLE_H

#ifndef OC_REAL
#define OC_REAL float
#endif
typedef OC_REAL real;

#ifdef __cplusplus
extern "C" {
#endif

/* reference constants: mkonst flag (const.f90:119), 1./3. as a default-real
 * literal stored into a double (hydrodynamic.f90:517-518) */
#define OC_FLAG (-9988765.0f)
#define OC_DRITTL ((double)(1.0f / 3.0f))

typedef struct {
    int nkn, nel, nlvdi, nlv;
    const int *nen3v;      /* [3*nel] 0-based */
    const double *ev;      /* [19*nel] */
    const int *ilhv;       /* [nel]  */
    const int *ilhkv;      /* [nkn]  */
    const real *hm3v;      /* [3*nel] */
    const real *hev;       /* [nel]  */
    const real *hlv;       /* [nlvdi] bottom depth of layer l */
    const real *hldv;      /* [nlvdi] thickness of layer l */
    const real *areakv;    /* [nlvdi*nkn] */
    const int *ibnd;       /* [nkn] open-boundary number (0 = none) */
    const int *ibtyp;      /* [nbnd+1] type of boundary b (index b) */
    const int *iextb;      /* [nkn] 1 if node is on the external boundary */
    const real *fcorv;     /* [nel] Coriolis parameter */
    const real *czv;       /* [nel] friction parameter */
    const real *wdifhv;    /* [9*nel] horizontal diffusion weights (iii,ii,ie) */
    /* CSR pattern of the zeta system (coo_init_new) */
    const int *ntg;        /* [nkn+1] */
    const int *cols;       /* [nnz]  */
    const int *ijp;        /* [9*nel] slot of (i,j,ie): ijp[i + 3*j + 9*ie] */
    int nnz;
    /* element colouring (domain_clusterization_ggu) */
    int ncolor;
    const int *color_ptr;  /* [ncolor+1] */
    const int *color_elems;/* [nel] */
    /* TVD upwind tables, index (id,ic,ie) -> (id-1) + 3*(ic-1) + 9*ie */
    const int *tvd_ie;     /* upwind element (0-based) or -1 */
    const real *tvd_x, *tvd_y;
    /* scalar open-boundary tables (bndo_init) */
    int nbndo;
    const int *bndo_k;     /* [nbndo] open-boundary nodes */
    const real *bndo_nx, *bndo_ny;   /* [nbndo] inward normal */
    const int *bndo_ptr;   /* [nbndo+1] */
    const int *bndo_nb;    /* inner neighbour nodes */
    const real *bndo_w;    /* their weights */
} oc_mesh;

typedef struct {
    real *znv, *zov, *zenv, *zeov;
    real *utlnv, *vtlnv, *utlov, *vtlov;
    real *unv, *vnv, *uov, *vov;
    real *ulnv, *vlnv;
    real *wlnv, *wlov;
    real *hdknv, *hdkov, *hdenv, *hdeov;
    int *iwegv, *iwetv, *inodv;
    real *rfricv;
    double *ddxv, *ddyv;
    real *visv, *difv;           /* [(nlvdi+1)*nkn] */
    real *difhv;                 /* [nlvdi*nel] */
    real *fxv, *fyv;             /* may be NULL (zero) */
    real *tauxnv, *tauynv;       /* may be NULL */
    real *ppv, *zeqv, *andgzv;   /* may be NULL */
    int *iuvfix;                 /* may be NULL */
    real *rzv;                   /* [nkn] prescribed level or OC_FLAG */
    real *rqv;                   /* [nkn] may be NULL */
    real *mfluxv;                /* [nlvdi*nkn] may be NULL */
    double dtime, dtime_old;     /* setweg time bookkeeping */
    /* the terms of the scalar being transported (conz3d_omp's robs, cobs,
     * rtauv, wsink, wsinkv, rload, load; zero / NULL: none) */
    real robs, wsink, rload;
    real *cobs, *rtauv;          /* [nlvdi*nkn] */
    real *wsinkv;                /* [(nlvdi+1)*nkn] */
    real *load;                  /* [nlvdi*nkn], updated by deposition */
} oc_state;

typedef struct {
    double dt;
    real azpar, ampar, afpar, atpar, adpar, aapar;
    real grav, rowass, vismol;
    int ireib; real uvmin;
    real hzmin, hzoff, hzon, volmin;
    int iclin;
    /* zeta solver */
    int itpre;             /* 1 = ilu0, 2 = ilut (reference default) */
    int lfil; double droptol;
    double rtol, atol; int itermax;
    /* scalars */
    real rkpar, difmol;
    int itvd, itvdv;
} oc_params;

typedef struct {
    int nloop;             /* wet/dry iterations of the last hydro call */
    int solver_iters;      /* BiCGStab matvecs/2 of the last solve */
    int solver_ierr;
    int total_solver_iters;
    int err;               /* 0 ok, else error code (see oc_errstr) */
    char msg[160];
} oc_stats;

int  oc_set_threads(int n);
/* geometry / topology (init) */
void oc_set_ev(int nkn, int nel, const int *nen3v, const real *xgv, const real *ygv, double *ev);
int  oc_diffweight(int nel, const double *ev, int idtype, real *wdifhv);
int  oc_node_list(int nkn, int nel, const int *nen3v, int *ntg, int *cols, int maxnnz, int *ijp);
int  oc_color(int nkn, int nel, const int *nen3v, int *color, int *color_ptr, int *color_elems);
void oc_set_area(const oc_mesh *m, real *areakv);
int  oc_tvd_init(const oc_mesh *m, const real *xgv, const real *ygv, int *tvd_ie, real *tvd_x, real *tvd_y);
void oc_tvd_upwind_table(const oc_mesh *m, int l, const real *cv, real *out);
void oc_tvd_grad_3d(const oc_mesh *m, const real *cc, real *gx, real *gy);
int  oc_scalar_stability(const oc_mesh *m, const oc_state *s, const oc_params *p, real dt, real rstol,
                         real *rindex);

/* band solvers (penta.f90) */
int  oc_penta_fact(int n, double *val);           /* val[(j+2) + 5*(i-1)], j=-2..2 */
void oc_penta_solve(int n, const double *val, const double *r, double *s);
void oc_tria_multi(int n, int m, double *v, double *r, double *s); /* v[(j-1)+3*(i-1)] */

/* sparse solver (SPARSKIT restatement) */
int  oc_ilut(int n, const double *a, const int *ja, const int *ia, int lfil, double droptol,
             double *alu, int *jlu, int *ju, int iwk);
int  oc_ilu0(int n, const double *a, const int *ja, const int *ia, double *alu, int *jlu, int *ju);
void oc_lusol(int n, const double *y, double *x, const double *alu, const int *jlu, const int *ju);
void oc_amux(int n, const double *x, double *y, const double *a, const int *ja, const int *ia);
int  oc_bcgstab(int n, const double *a, const int *ja, const int *ia,
                const double *alu, const int *jlu, const int *ju,
                const double *rhs, double *sol, double rtol, double atol, int itermax,
                int *iters);
int  oc_system_solve(const oc_mesh *m, const oc_params *p, const double *c2coo,
                     const double *rvec, double *x, int *iters);

/* hydro step pieces */
void oc_setweg(const oc_mesh *m, oc_state *s, const oc_params *p, int iweich, int *iw);
void oc_set_dry(const oc_mesh *m, oc_state *s, const oc_params *p);
int  oc_setnod(const oc_mesh *m, oc_state *s);
void oc_copy_uvz(const oc_mesh *m, oc_state *s);
void oc_copy_depth(const oc_mesh *m, oc_state *s);
void oc_bottom_friction(const oc_mesh *m, oc_state *s, const oc_params *p);
void oc_hydro_transports(const oc_mesh *m, oc_state *s, const oc_params *p);
void oc_hydro_zeta(const oc_mesh *m, oc_state *s, const oc_params *p, double *c2coo, double *rvec);
void oc_hydro_transports_final(const oc_mesh *m, oc_state *s, const oc_params *p);
void oc_setzev(const oc_mesh *m, oc_state *s);
int  oc_setuvd(const oc_mesh *m, oc_state *s, const oc_params *p);
void oc_baro2l(const oc_mesh *m, oc_state *s);
int  oc_setdepth(const oc_mesh *m, real *hdkn, real *hden, const real *zenv);
int  oc_check_volume(const oc_mesh *m, const oc_state *s);
void oc_hydro_vertical(const oc_mesh *m, oc_state *s, const oc_params *p, real *dzeta);
double oc_masscont(const oc_mesh *m, const oc_state *s, int mode);
void oc_ttov(const oc_mesh *m, oc_state *s);
int  oc_hydro(const oc_mesh *m, oc_state *s, const oc_params *p, oc_stats *st);

/* scalar transport */
void oc_conz_hflux(const oc_mesh *m, const oc_state *s, const oc_params *p, const real *cv, const real *gx,
                   const real *gy, int ie, int l, double dt, double *fl);
int  oc_conz3d(const oc_mesh *m, const oc_state *s, const oc_params *p,
               real *cn1, const real *cbound, double ddt, int istot, int isact);
int  oc_scal3sh(const oc_mesh *m, oc_state *s, const oc_params *p, real *cnv,
                const real *rcv, int istot, int clip_small);   /* rcv: boundary values or NULL */
double oc_massconc(const oc_mesh *m, const oc_state *s, const real *cnv, int mode);
int  oc_bndo_init(const oc_mesh *m, const real *xgv, const real *ygv, int nb, const int *bnodes,
                  int *bndo_k, real *nx, real *ny, int *ptr, int *nbr, real *w, int maxnb);
void oc_uvtopr(const oc_mesh *m, const oc_state *s, real *uprv, real *vprv);
int  oc_bndo_setbc(const oc_mesh *m, real *cv, const real *rbc, const real *uprv, const real *vprv);

/* baroclinic coupling, z-layers (shyoracle_barocl.c) */
real oc_sigma_stp(real s, real t, real p);
void oc_rhoset(const oc_mesh *m, const oc_params *p, const real *temp, const real *salt, real *rhov);
void oc_barocl(const oc_mesh *m, const oc_params *p, const real *rhov, real *fxv, real *fyv);

/* forcing (shyoracle_forcing.c): record series and wind conversion, always
 * in the reference's own types (f32 data, f64 times and interpolation) */
double oc_intp_neville(int nintp, const double *xa, const float *ya, double x);
int  oc_iff_first_record(int nrec, const double *time, int nintp, double dtime0);
int  oc_iff_series(int nrec, const double *time, const float *vals, int ncomp, int comp, int nval, int nintp,
                   double dtime0, float vdefault, int nt, const double *t, float *out, int *newest);
int  oc_convert_wind(int n, float *wx, float *wy, int iwtype, int itdrag, float dragco, float roluft, float rowass,
                     float wslim, float wsmax, float *tx, float *ty, float *ws);
/* set_mass_flux (shyoracle_forcing.c): sources in the reference's loop
 * order, node 0-based; rqdsv may be NULL; returns 1 on the error stop */
int  oc_set_mass_flux(int nkn, int nlvdi, int nlv, const int *ilhkv, const real *hdknv, const real *areakv,
                      int nsrc, const int *node, const int *levmin, const int *levmax, const float *rflux,
                      const float *rqdsv, real *mfluxv, real *rqv);

#ifdef __cplusplus
}
#endif
#endif
