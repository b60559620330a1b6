/*
 * shyhip.h -- C ABI of libshyhip.so, the MI355X (gfx950) implementation of
 * SHYFEM's semi-implicit hydrodynamic time step and 3D scalar transport.
 *
 * The library owns device-resident copies of the model state.  Host arrays
 * crossing this boundary use the reference's own layouts and types, so a
 * Fortran caller passes its module arrays directly (c_loc):
 *   real    = float (32-bit); double precision = double
 *   element 3D  a(nlvdi,nel), node 3D a(nlvdi,nkn), interfaces a(0:nlvdi,nkn)
 *   vertex      zenv(3,nel),  connectivity nen3v(3,nel) with 1-based node ids
 * All entry points return 0 on success and a non-zero code otherwise; the
 * message is available from shyhip_last_error().  Handles are independent:
 * calls on different handles may run concurrently from different host
 * threads (each handle has its own HIP stream and scratch); calls on one
 * handle must not overlap.
 *
 * Reference interfaces replaced (SHYFEM 8_1_1, paths relative to its root):
 *   shyhip_hydro               hydro                src/shyfem/hydrodynamic.f90:284
 *   shyhip_hydro_transports    hydro_transports     src/shyfem/hydrodynamic.f90:758
 *   shyhip_hydro_zeta          hydro_zeta(vqv)      src/shyfem/hydrodynamic.f90:488
 *   shyhip_hydro_transports_final                   src/shyfem/hydrodynamic.f90:1468
 *   shyhip_hydro_vertical      hydro_vertical(dzeta) src/shyfem/hydrodynamic.f90:1582
 *   shyhip_set_dry             set_dry              src/shyfem/wetdry.f90:94
 *   shyhip_copy_uvz            copy_uvz + copy_depth src/shyfem/transforms.f90:523,
 *                                                    src/shyfem/elems_dealing.f90:843
 *   shyhip_scal3sh             scal3sh -> conz3d_omp src/shyfem/concentration.f90:565,
 *                                                    src/shyfem/concentration_omp.f90:65
 *   shyhip_system_*            system_initialize/init/assemble/add_rhs/solve/get/finalize
 *                                                    src/shyfem/simsys_spkaux.f90:85-721
 *   shyhip_volume              vol_mass -> masscont src/shyfem/check.f90:647,
 *                                                    src/shyfem/elems_dealing.f90:1191
 */
#ifndef SHYHIP_H
#define SHYHIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct shyhip_model shyhip_model;

/* precision of the device path */
enum { SHYHIP_FP64 = 0,   /* all state and the zeta solve in FP64 */
       SHYHIP_FP32 = 1 }; /* FP32 state and transport, FP32 Krylov inside FP64 refinement */

/* element-column solver variant chosen by shyhip_set_params */
typedef struct {
    double dt;                                  /* idt                          */
    float azpar, ampar, afpar, atpar, adpar, aapar;
    float grav, rowass, vismol;
    int ireib; float uvmin;                     /* bottom friction (chezy.f90)  */
    float hzmin, hzoff, hzon, volmin;           /* wet/dry thresholds           */
    int iclin;
    /* zeta solver */
    double rtol, atol;                          /* BiCGStab tolerances (1e-8)   */
    int itermax;                                /* max matvecs (1000)           */
    double refine_rtol;                         /* FP32 path: FP64 refinement residual target */
    int block_rows;                             /* block-ILU(0) block size (<= 1024), 0 = auto */
    int tri_sweeps;                             /* 0: exact level-scheduled triangular solves;
                                                   k > 0: k Jacobi sweeps per triangular factor */
    double inner_rtol;                          /* FP32 path: relative tolerance of the inner solve */
    /* scalars */
    float rkpar, difmol;                        /* horizontal diffusion factor of the scalars
                                                   (times the difhv field; 0 = none), vertical
                                                   molecular diffusivity */
    int itvd, itvdv;                            /* horizontal TVD: 0 none, 1 gradients at the
                                                   nodes (tvd_grad_3d, on the device each
                                                   substep), 2 upwind points (needs
                                                   shyhip_tvd_init); vertical TVD 0 or 1 */
    /* transport kernels */
    int node_tile;                              /* most nodes per workgroup of the conz3d node
                                                   phase (a workgroup also holds at most 256
                                                   node-layers): 0 = auto (256), else 1..256;
                                                   read by shyhip_set_basin, like block_rows */
    int idtype;                                 /* weights of horizontal diffusion (diffweight):
                                                   0 gradient products, 1 the same without
                                                   positive off-diagonals, 2 (default) edge
                                                   lengths */
} shyhip_params;

/* field ids for shyhip_push / shyhip_pull */
enum {
    SHYHIP_ZNV = 1, SHYHIP_ZOV, SHYHIP_ZENV, SHYHIP_ZEOV,
    SHYHIP_UTLNV, SHYHIP_VTLNV, SHYHIP_UTLOV, SHYHIP_VTLOV,
    SHYHIP_UNV, SHYHIP_VNV, SHYHIP_UOV, SHYHIP_VOV,
    SHYHIP_ULNV, SHYHIP_VLNV,
    SHYHIP_WLNV, SHYHIP_WLOV,
    SHYHIP_HDKNV, SHYHIP_HDKOV, SHYHIP_HDENV, SHYHIP_HDEOV,
    SHYHIP_IWEGV, SHYHIP_IWETV, SHYHIP_INODV,
    SHYHIP_RFRICV, SHYHIP_DDXV, SHYHIP_DDYV,
    SHYHIP_VISV, SHYHIP_DIFV, SHYHIP_DIFHV,
    SHYHIP_TAUXNV, SHYHIP_TAUYNV, SHYHIP_RZV, SHYHIP_RQV, SHYHIP_MFLUXV,
    SHYHIP_AREAKV, SHYHIP_CSR_VAL, SHYHIP_RHS,
    SHYHIP_RHOV,               /* density anomaly rho - rowass (rhov), node 3D          */
    /* optional forcing of hydro, absent until first pushed (a pull before
     * then fails): atmospheric pressure ppv (Pa) and equilibrium tide zeqv
     * (m) at the nodes (hydro_intern, hydrodynamic.f90:1025-1033,1235),
     * the level nudging term andgzv (m/s, hydro_zeta :612,679); elements
     * with fixed transports iuvfix (element, I32, 0/1: hydro_intern skips
     * them, :1383-1392, and their momentum leaves the zeta matrix and the
     * final transports, :627-629,1525) and the transports they take each
     * step (utfix/vtfix, element 3D, what bclfix writes into utlnv/vtlnv,
     * bcvel.f90:332-338; not pushed: the previous step's) */
    SHYHIP_PPV, SHYHIP_ZEQV, SHYHIP_ANDGZV, SHYHIP_IUVFIX, SHYHIP_UTFIX, SHYHIP_VTFIX,
    SHYHIP_TRACER0 = 100,      /* tracer itr: SHYHIP_TRACER0 + itr, node 3D            */
    SHYHIP_TRACER_BC0 = 200,   /* its open-boundary value (-999 = none), node 3D        */
    /* tracer itr's own terms (conz3d_omp's cobs, rtauv, wsinkv, load),
     * absent until first pushed; see shyhip_set_tracer_terms */
    SHYHIP_COBS0 = 300,        /* nudging target, node 3D                               */
    SHYHIP_RTAUV0 = 400,       /* nudging rate (1/s), node 3D                           */
    SHYHIP_WSINKV0 = 500,      /* sinking profile, node interfaces 0..nlvdi             */
    SHYHIP_LOAD0 = 600         /* load, node 3D (deposition writes what it took)        */
};
/* dtype of host buffers */
enum { SHYHIP_F32 = 0, SHYHIP_F64 = 1, SHYHIP_I32 = 2 };

const char *shyhip_last_error(const shyhip_model *h);
int  shyhip_device_count(int *n);

/* life cycle ------------------------------------------------------------ */
int  shyhip_create(shyhip_model **h, int precision, int nkn, int nel, int nlvdi, int nlv, int ntracer);
int  shyhip_destroy(shyhip_model *h);
/* the sizes the model was created with (any pointer may be NULL) */
int  shyhip_get_sizes(const shyhip_model *h, int *nkn, int *nel, int *nlvdi, int *nlv, int *ntracer);
/* the handle's parameters (defaults after create: the reference's STR
 * defaults, default_parms in para.f90 / def_para.f90) */
int  shyhip_get_params(const shyhip_model *h, shyhip_params *p);
int  shyhip_set_params(shyhip_model *h, const shyhip_params *p);
/* the same checks without a model and without a GPU (the text is in
 * shyhip_last_error(NULL)): a run the device path cannot do is refused,
 * as the reference stops on an unknown option (e.g. bottom_friction,
 * chezy.f90:373-376, for ireib outside 0-10) */
int  shyhip_check_params(const shyhip_params *p);

/* several independent basins in one model (an ensemble): member m owns the
 * nodes node_start[m] .. node_start[m+1]-1 and the elements
 * elem_start[m] .. elem_start[m+1]-1 (0-based, node_start[nmem] = nkn,
 * elem_start[nmem] = nel); its elements use only its nodes.  Colouring and
 * upwind tables are built per member, so each member's element order and
 * upwind points are those it has alone.  Call before shyhip_set_basin;
 * without it the model is one member.  No counterpart in the reference,
 * which runs one basin per process. */
int  shyhip_set_members(shyhip_model *h, int nmem, const int *node_start, const int *elem_start);
/* No counterpart in the reference (the serving loop's continuous batching:
 * runs that join or leave a batch move into a new union model).  Copies
 * the state member sm of src carries from step to step -- levels,
 * transports, velocities, depths, wet/dry masks, friction, viscosity and
 * diffusivity, wind stress, discharges, density, tracers and their
 * boundary values -- into member dm of dst, device to device.  The two
 * members must hold the same basin (node, element and layer counts),
 * precision and tracers.  Waits for both models' work. */
int  shyhip_copy_member_state(shyhip_model *dst, int dm, shyhip_model *src, int sm);
/* No counterpart in the reference: a uniform wind stress on the nodes of
 * member m only (each run of a batch has its own wind). */
int  shyhip_set_member_wind(shyhip_model *h, int m, float taux, float tauy);

/* basin and layers (init time): nen3v(3,nel) 1-based, xgv/ygv(nkn),
 * hm3v(3,nel), hev(nel), hlv/hldv(nlvdi), ilhv(nel), ilhkv(nkn).
 * Builds geometry (set_ev), the zeta-matrix pattern (coo_init_new), the
 * element colouring (domain_clusterization_ggu), nodal areas (set_area) and
 * the block-ILU(0) level schedule, and uploads everything to the device. */
int  shyhip_set_basin(shyhip_model *h, const int *nen3v, const float *xgv, const float *ygv,
                      const float *hm3v, const float *hev, const float *hlv, const float *hldv,
                      const int *ilhv, const int *ilhkv);
/* ibnd(nkn): open-boundary number of each node (0 none); ibtyp(nbnd+1):
 * type of boundary b at index b; iextb(nkn): node on the external boundary */
int  shyhip_set_boundaries(shyhip_model *h, int nbnd, const int *ibnd, const int *ibtyp, const int *iextb);
/* fcorv(nel), czv(nel); any pointer may be NULL (zero) */
int  shyhip_set_coefficients(shyhip_model *h, const float *fcorv, const float *czv);
/* TVD (itvd = 2) upwind tables, built on the host from the basin */
int  shyhip_tvd_init(shyhip_model *h);

/* the reference's start-up sequence for a constant initial level:
 * setweg(-1), setnod, init_z, setweg(0), setznv, make_new_depth,
 * copy_depth, zero transports, hydro_vertical, copy_uvz, copy_depth */
int  shyhip_init_state(shyhip_model *h, float zconst);

/* state transfer in the reference's layout; n = number of values */
int  shyhip_push(shyhip_model *h, int field, const void *host, int dtype, size_t n);
int  shyhip_pull(shyhip_model *h, int field, void *host, int dtype, size_t n);

/* forcing ---------------------------------------------------------------- */
/* level of open boundary ibc (sp111 for ibtyp = 1 boundaries) */
int  shyhip_set_boundary_level(shyhip_model *h, int ibc, float level);
/* levels of all open boundaries at once: levels[b-1] for boundary b = 1..n,
 * n = nbnd (applied to the type 1 boundaries) */
int  shyhip_set_boundary_levels(shyhip_model *h, int n, const float *levels);
/* uniform wind stress per unit density (meteo_forcing.f90:973) */
int  shyhip_set_wind_uniform(shyhip_model *h, float taux, float tauy);

/* Forcing series held on the device.  Replaces the forcing files that
 * sp111(2) reads for each step (bnd_routines.f90:444-596, the boundary
 * files opened at :330-356) and meteo_forcing_fem reads for the wind
 * (meteo_forcing.f90:301-555), with their time interpolation
 * (iff_read_and_interpolate + iff_time_interpolate, intp_fem_file.f90:
 * 2288-2510, rd_intp_neville, utils/generic/intp.f90:317-345): records
 * time[0..nrec-1] (strictly increasing, seconds as dtime), interpolated with
 * intpol = 2 or 4 records (0: the reference's default, 4 for a level
 * boundary, 2 for the wind), the first ring chosen for start time dtime0 as
 * iff_populate_records does (:1047-1144).  One record = a constant field.
 * A record value of -999 inside the window gives 0.  Once installed, every
 * shyhip_step evaluates the series at its dtime on the device before hydro
 * (as sp111(2) precedes hydro, shyfem_subs.f90:592-604); times must not
 * decrease between steps, and a time outside the records is an error
 * (status 2), as the reference stops (intp_fem_file.f90:2344-2352).
 *
 * Level boundary ibc (ibtyp 1): val[r*nval + i], nval = 1 (one level for the
 * whole boundary) or the boundary's node count (its nodes in ascending node
 * order); the level is multiplied by zfact (adjust_bound,
 * bnd_routines.f90:1003-1031) and written to rzv. */
int  shyhip_set_boundary_series(shyhip_model *h, int ibc, int nrec, const double *time, const float *val, int nval,
                                int intpol, float zfact, double dtime0);
/* wind conversion settings (meteo_set_wind_data, meteo_forcing.f90:700-802;
 * STR defaults def_para.f90:782-799: iwtype 1, itdrag 0, dragco 2.5e-3,
 * roluft 1.225, wslim -1, wsmax 50); rowass is the model parameter */
typedef struct {
    int iwtype;      /* 1: wind velocity (m/s), 2: stress (N/m2)                       */
    int itdrag;      /* 0: dragco, 1: Smith and Banke (1975), 2: Large and Pond (1981)   */
    float dragco, roluft, wslim, wsmax;
} shyhip_wind_params;
/* Wind series: wx/wy[r*nval + i], nval = 1 (one point for the whole basin)
 * or nkn (node values); each step converts it to the stress per unit density
 * in tauxnv/tauynv (meteo_convert_wind_data, meteo_forcing.f90:864-1000, no
 * ice cover).  A speed above wsmax is reported as a device error at the next
 * check.  nrec = 0 removes the series. */
int  shyhip_set_wind_series(shyhip_model *h, int nrec, const double *time, const float *wx, const float *wy, int nval,
                            int intpol, double dtime0, const shyhip_wind_params *wp);
/* sp111(2) + meteo_force on the device at time dtime, outside a step */
int  shyhip_forcing(shyhip_model *h, double dtime);

/* output in the SHY format ----------------------------------------------------
 * (io_shy.f90, format version 14; the file SHYFEM's post-processing reads).
 * shyhip_shy_open writes the header (shy_write_header, io_shy.f90:1550-1606)
 * from host arrays in the reference's layouts: nen3v(3,nel) 1-based,
 * hm3v(3,nel), external numbers ipev/ipv and area codes iarv/iarnv (null:
 * 1..n and 0), simpar(4) = hzmin, hzoff, nzadapt, dtaver (null: zeros).
 * ftype 1: hydro (npr 3, nvar 4), 2: scalars on nodes (npr 1).
 * shyhip_shy_write_record writes one data record (shy_write_record,
 * :1610-1681): c(nlvddi, n*m), a 3D record (lmax > 1) packed over each
 * node's/element's own layers.  Status 0 = ok; shyhip_shy_error gives the
 * text.  Host code only: no GPU is needed. */
typedef struct shyhip_shyfile shyhip_shyfile;
int  shyhip_shy_open(const char *path, int ftype, int nkn, int nel, int npr, int nlv, int nvar, int date, int time,
                     const char *title, const char *femver, const float *simpar, const int *nen3v, const int *ipev,
                     const int *ipv, const int *iarv, const int *iarnv, const float *xgv, const float *ygv,
                     const float *hm3v, const float *hlv, const int *ilhv, const int *ilhkv, shyhip_shyfile **f);
int  shyhip_shy_write_record(shyhip_shyfile *f, double dtime, int ivar, int belem, int n, int m, int lmax, int nlvddi,
                             const float *c);
int  shyhip_shy_close(shyhip_shyfile *f);
const char *shyhip_shy_error(const shyhip_shyfile *f);
/* one hydro time record of the model's device state: znv, zenv, utlnv,
 * vtlnv (shy_write_hydro_records, shy_util.f90:1210-1245, as called by the
 * hydro output, ous_admin.f90:122-127) */
int  shyhip_shy_write_hydro(shyhip_model *h, shyhip_shyfile *f, double dtime);
/* one node 3D field of the model (e.g. SHYHIP_TRACER0 + itr) as a scalar
 * record with variable id ivar */
int  shyhip_shy_write_node3d(shyhip_model *h, shyhip_shyfile *f, double dtime, int ivar, int field);

/* debug dumps (shympi_debug, utils/shympi/mpi_debug.f90:114-146,367-415;
 * read and compared by check_shympi_debug): a time block is
 * shyhip_dbg_time, records of arrays (lmax, nn) in the reference's layout
 * (type 1 int32, 2 real, 3 double), shyhip_dbg_end_time.  Host code. */
typedef struct shyhip_dbgfile shyhip_dbgfile;
int  shyhip_dbg_open(const char *path, shyhip_dbgfile **d);
int  shyhip_dbg_time(shyhip_dbgfile *d, double dtime);
int  shyhip_dbg_record(shyhip_dbgfile *d, const char *name, int nn, int lmax, int type, const void *array);
int  shyhip_dbg_end_time(shyhip_dbgfile *d);
int  shyhip_dbg_close(shyhip_dbgfile *d);
const char *shyhip_dbg_error(const shyhip_dbgfile *d);
/* the records of the model's state inside a time block: znv, zenv, utlnv,
 * vtlnv, wlnv, hdknv, hdenv, iwegv (the arrays this path updates, as
 * shyfem_subs.f90:1210-1265 writes them) */
int  shyhip_dbg_write_state(shyhip_model *h, shyhip_dbgfile *d);

/* baroclinic coupling (z-layers) -------------------------------------------
 * ibarcl = 1 (barocl, baroclinic.f90:195-600; ibarcl = 0 switches it off):
 * tracers itemp and isalt are temperature and salinity.  Starts the density
 * anomaly as barocl(0) does (rhov = 0, rhoset three times, :398-404); then
 * every hydro adds the baroclinic pressure gradient to the explicit
 * momentum terms (set_explicit -> set_barocl_new_interface,
 * explicit.f90:193,1226-1530) and every shyhip_step ends with one rhoset
 * (rhoset_shell(1), :542-548) from the transported T/S. */
int  shyhip_set_baroclinic(shyhip_model *h, int ibarcl, int itemp, int isalt);
/* rhoset (baroclinic.f90:654-720) niter times: the density anomaly from
 * T, S and the mid-layer pressure, with the UNESCO 1980 equation of state
 * (sigma_stp, eq_state.f90:154-235) */
int  shyhip_rhoset(shyhip_model *h, int niter);

/* discharges (set_mass_flux, bnd_routines.f90:1035-1183) --------------------
 * The point sources: one entry per node of the flux (ibtyp 2) and inlet
 * (ibtyp 3) boundaries, in the reference's loop order (boundary by boundary,
 * kbnds order), with the boundary's levmin/levmax (get_bnd_ipar); node
 * 1-based.  nsrc = 0: distributed sources only. */
int  shyhip_set_sources(shyhip_model *h, int nsrc, const int *node, const int *levmin, const int *levmax);
/* set_mass_flux: mfluxv (per node and layer) and rqv on the device from the
 * discharge of every source (rflux(kindex) of sp111, bnd_routines.f90:
 * 532-544, m3/s, in shyhip_set_sources' order) and the distributed sources
 * rqdsv[nkn] (rain/evaporation, m3/s, NULL = none), using the current layer
 * volumes.  Call it where the reference calls set_mass_flux, before the step
 * (or before shyhip_hydro).  Errors (status 3): layer limits outside the
 * column, a zero source volume (the reference's error stop). */
int  shyhip_set_mass_flux(shyhip_model *h, const float *rflux, const float *rqdsv);

/* time step -------------------------------------------------------------- */
int  shyhip_set_dry(shyhip_model *h);
int  shyhip_copy_uvz(shyhip_model *h);                 /* copy_uvz + copy_depth */
int  shyhip_hydro(shyhip_model *h, double dtime);      /* hydro at time dtime   */
int  shyhip_hydro_transports(shyhip_model *h);
int  shyhip_hydro_zeta(shyhip_model *h);                /* assemble (uses rqv)   */
int  shyhip_system_solve_device(shyhip_model *h, int *iters);
int  shyhip_hydro_transports_final(shyhip_model *h);
int  shyhip_hydro_vertical(shyhip_model *h);
/* one full model step: set_dry, set_timestep, copy_uvz/copy_depth, hydro,
 * and scal3sh for every tracer with istot substeps */
int  shyhip_step(shyhip_model *h, double dtime, int istot);
/* scal3sh for tracer itr (0-based); clip_small = 1 for non-temperature.
 * istot <= 0: the number of substeps is chosen as the reference does,
 * by scalar_stability with rstol = 1 (concentration.f90:678-700; more than
 * 1000 is an error, status 3); shyhip_step does the same for all tracers. */
int  shyhip_scal3sh(shyhip_model *h, int itr, int istot, int clip_small);
/* conz3d_omp's own terms of tracer itr (concentration_omp.f90:65-151):
 * nudging robs * rtauv towards cobs (explicit, :697-701), sinking
 * wsink * wsinkv subtracted from the vertical velocity (:479) and loads
 * rload * load (erosion, or deposition bounded by the node's content,
 * :816-834); each acts once its field is pushed (SHYHIP_COBS0 ..
 * SHYHIP_LOAD0 + itr) and its factor is non-zero.  scalar_stability
 * includes the tracer's nudging and sinking, so shyhip_scal3sh and a step
 * transport such a tracer alone with its own substeps (the reference's
 * scal3sh per tracer); a step's other tracers keep theirs. */
int  shyhip_set_tracer_terms(shyhip_model *h, int itr, float robs, float wsink, float rload);
/* conz3d_omp (src/shyfem/concentration_omp.f90:65-340): one substep isact of
 * istot of tracer itr on the device state -- the element and node phases
 * only; the clipping of small values and the open-boundary treatment around
 * it are scal3sh's (shyhip_scal3sh).  cbound(nlvddi,nkn): the source
 * concentration of scal3sh's make_scal_flux (sbconz), NULL = zero. */
int  shyhip_conz3d(shyhip_model *h, int itr, const float *cbound, int istot, int isact);
/* the open-boundary treatment scal3sh applies after each substep
 * (bndo_setbc, src/shyfem/bndo.f90:620-890) on tracer itr, with the boundary
 * values of SHYHIP_TRACER_BC0 + itr */
int  shyhip_bndo_setbc(shyhip_model *h, int itr);
/* scalar_stability (src/shyfem/stability.f90:220-260, conzstab
 * src/shyfem/concentration.f90:1573-2097) on the current state: the
 * stability index of the scalar transport for the step dt, rindex, and
 * istot = 1 + rindex (an ensemble: the largest over its members) */
int  shyhip_scalar_stability(shyhip_model *h, float rstol, int *istot, double *rindex);
/* the number of substeps the last scal3sh or step ran (an ensemble: the
 * largest of its members') */
int  shyhip_get_istot(shyhip_model *h, int *istot);
/* an ensemble's members (shyhip_set_members) each choose their istot as they
 * do alone; the substeps each member ran in the last step (nmem = the
 * number of members) */
int  shyhip_member_istot(shyhip_model *h, int nmem, int *istot);
/* zeta solves since create that failed with warm-started block factors and
 * were done again with exact ones (the retry is counted, never hidden) */
int  shyhip_get_retries(shyhip_model *h, long *retries);
int  shyhip_sync(shyhip_model *h);
/* Deferred checks (no counterpart in the reference, which runs on the host
 * and stops at once): with on = 1, shyhip_step returns without waiting for
 * the device at its end, so the next step's launches queue behind it.  A
 * device error raised by a step's final phase (depths, transport) is then
 * returned by the next call that waits for the device: the next step's
 * solve, shyhip_sync, or a pull.  Off by default. */
int  shyhip_set_deferred_checks(shyhip_model *h, int on);
/* HIP graphs (no counterpart in the reference): each tracer group's
 * transport substeps (conz3d and bndo_setbc launches of a step) and each
 * chunk of zeta-solver iterations are captured once per distinct argument
 * set and replayed from the instantiated graph, one launch instead of
 * dozens.  On by default; the results are the same bits either way.
 * shyhip_graph_stats: graphs instantiated and replays since create. */
int  shyhip_set_graphs(shyhip_model *h, int on);
int  shyhip_graph_stats(shyhip_model *h, long *builds, long *replays);
/* kernel-group timing with HIP events on the handle's stream (off by
 * default; enabling it adds a synchronisation at the end of every step).
 * id = 0 .. 19: set_dry, copy_uvz, hydro_intern, setnod, zeta_assemble,
 * zeta_solve, system_get_setweg, transports_final, setuvd_baro2l, setdepth,
 * hydro_vertical, ttov_uvint, conz3d_elem, conz3d_node, bndo_setbc, inside
 * zeta_solve the single kernels bilu_factor, bicg_a, bicg_b, conz3d_prep
 * (the step's element-layer records of the transport) and tvd_grad (the
 * node gradients of itvd = 1) */
int  shyhip_set_timing(shyhip_model *h, int on);
int  shyhip_get_timing(shyhip_model *h, int id, const char **name, double *total_ms, long *count);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the FP64 volume (conc = NULL) or tracer mass of node 3D
 * device arrays in the library's layout a[l*nkn + k], element type dtype
 * (SHYHIP_F32 or SHYHIP_F64), summed exactly as shyhip_volume and
 * shyhip_tracer_mass sum them (same kernels and order, so the same bits),
 * into the device double *out, asynchronously on stream (a hipStream_t,
 * NULL = the default stream).  ilhkv: device int[nkn].  Allocates nothing:
 * the partial sums live in one static 512 KiB device array per GPU, so nkn
 * must be at most 16,777,216 (larger is an error); the array keeps the last
 * call's partials, and calls that overlap on one GPU race on it. */
int  shyhip_masscont_noalloc(int dtype, int nkn, const int *ilhkv, const void *hdk, const void *areakv,
                             const void *conc, double *out, void *stream);

/* diagnostics ------------------------------------------------------------ */
int  shyhip_volume(shyhip_model *h, int mode, double *vol);     /* FP64 total volume */
int  shyhip_tracer_mass(shyhip_model *h, int itr, double *mass);
int  shyhip_stats(shyhip_model *h, int *nloop, int *solver_iters, int *refinements);
/* integer structures built by shyhip_set_basin, for checking:
 * ntg(nkn+1), cols(nnz) 0-based; color(nel) 1-based colours */
int  shyhip_get_pattern(shyhip_model *h, int *nnz, int *ntg, int *cols, int *ncolor, int *color);

/* the zeta-system plugin (simsys_spkaux.f90): the stock hydro_zeta keeps
 * calling system_assemble(ie) with kn/hia/hik; the plugin accumulates on the
 * host in element order and solves on the device. */
int  shyhip_system_init(shyhip_model *h);
/* system_set_explicit (simsys_spkaux.f90:160-168): from now on the system
 * is the lumped one, diagonal and right-hand side only, and the solve is
 * their quotient (:475-479); an off-diagonal entry is an error */
int  shyhip_system_set_explicit(shyhip_model *h);
int  shyhip_system_assemble(shyhip_model *h, int ie, const int *kn, const float *hia, const float *hik);
int  shyhip_system_add_rhs(shyhip_model *h, float dt, int n, const float *array);
int  shyhip_system_solve(shyhip_model *h, int n, const float *z);
int  shyhip_system_get(shyhip_model *h, int n, float *z);

#ifdef __cplusplus
}
#endif
#endif
