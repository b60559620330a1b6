! shyhip_capi.f90 -- ISO_C_BINDING interface to libshyhip.so for Fortran
! host code (the C declarations are in include/shyhip.h).
!
! Arrays are passed in SHYFEM's own layouts: a(nlvdi,nel), a(nlvdi,nkn),
! zenv(3,nel), nen3v(3,nel) with 1-based node numbers.  Every function returns
! 0 on success; shyhip_check turns a non-zero status into an error stop with
! the library's message, the way the replaced routines stop
! (sparskit_admin.f90:322-325, penta.f90:284-287).

module shyhip_capi

  use iso_c_binding
  implicit none

  integer(c_int), parameter :: SHYHIP_FP64 = 0, SHYHIP_FP32 = 1
  integer(c_int), parameter :: SHYHIP_F32 = 0, SHYHIP_F64 = 1, SHYHIP_I32 = 2

  ! field ids of shyhip_push / shyhip_pull
  integer(c_int), parameter :: SHYHIP_ZNV = 1, SHYHIP_ZOV = 2, SHYHIP_ZENV = 3, SHYHIP_ZEOV = 4
  integer(c_int), parameter :: SHYHIP_UTLNV = 5, SHYHIP_VTLNV = 6, SHYHIP_UTLOV = 7, SHYHIP_VTLOV = 8
  integer(c_int), parameter :: SHYHIP_UNV = 9, SHYHIP_VNV = 10, SHYHIP_UOV = 11, SHYHIP_VOV = 12
  integer(c_int), parameter :: SHYHIP_ULNV = 13, SHYHIP_VLNV = 14, SHYHIP_WLNV = 15, SHYHIP_WLOV = 16
  integer(c_int), parameter :: SHYHIP_HDKNV = 17, SHYHIP_HDKOV = 18, SHYHIP_HDENV = 19, SHYHIP_HDEOV = 20
  integer(c_int), parameter :: SHYHIP_IWEGV = 21, SHYHIP_IWETV = 22, SHYHIP_INODV = 23
  integer(c_int), parameter :: SHYHIP_RFRICV = 24, SHYHIP_DDXV = 25, SHYHIP_DDYV = 26
  integer(c_int), parameter :: SHYHIP_VISV = 27, SHYHIP_DIFV = 28, SHYHIP_DIFHV = 29
  integer(c_int), parameter :: SHYHIP_TAUXNV = 30, SHYHIP_TAUYNV = 31, SHYHIP_RZV = 32
  integer(c_int), parameter :: SHYHIP_RQV = 33, SHYHIP_MFLUXV = 34
  integer(c_int), parameter :: SHYHIP_AREAKV = 35, SHYHIP_CSR_VAL = 36, SHYHIP_RHS = 37
  integer(c_int), parameter :: SHYHIP_RHOV = 38
  integer(c_int), parameter :: SHYHIP_PPV = 39, SHYHIP_ZEQV = 40, SHYHIP_ANDGZV = 41, SHYHIP_IUVFIX = 42
  integer(c_int), parameter :: SHYHIP_UTFIX = 43, SHYHIP_VTFIX = 44
  integer(c_int), parameter :: SHYHIP_TRACER0 = 100, SHYHIP_TRACER_BC0 = 200
  integer(c_int), parameter :: SHYHIP_COBS0 = 300, SHYHIP_RTAUV0 = 400, SHYHIP_WSINKV0 = 500, SHYHIP_LOAD0 = 600

  type, bind(C) :: shyhip_params
    real(c_double) :: dt
    real(c_float)  :: azpar, ampar, afpar, atpar, adpar, aapar
    real(c_float)  :: grav, rowass, vismol
    integer(c_int) :: ireib
    real(c_float)  :: uvmin
    real(c_float)  :: hzmin, hzoff, hzon, volmin
    integer(c_int) :: iclin
    real(c_double) :: rtol, atol
    integer(c_int) :: itermax
    real(c_double) :: refine_rtol
    integer(c_int) :: block_rows
    integer(c_int) :: tri_sweeps
    real(c_double) :: inner_rtol
    real(c_float)  :: rkpar, difmol
    integer(c_int) :: itvd, itvdv
    integer(c_int) :: node_tile
    integer(c_int) :: idtype
  end type shyhip_params

  ! wind conversion settings of shyhip_set_wind_series
  type, bind(C) :: shyhip_wind_params
    integer(c_int) :: iwtype, itdrag
    real(c_float)  :: dragco, roluft, wslim, wsmax
  end type shyhip_wind_params

  interface

    function shyhip_last_error(h) bind(C, name="shyhip_last_error") result(msg)
      import :: c_ptr
      type(c_ptr), value :: h
      type(c_ptr) :: msg
    end function

    function shyhip_create(h, precision, nkn, nel, nlvdi, nlv, ntracer) bind(C, name="shyhip_create") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), intent(out) :: h
      integer(c_int), value :: precision, nkn, nel, nlvdi, nlv, ntracer
      integer(c_int) :: rc
    end function

    function shyhip_destroy(h) bind(C, name="shyhip_destroy") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: h
      integer(c_int) :: rc
    end function

    function shyhip_get_params(h, p) bind(C, name="shyhip_get_params") result(rc)
      import :: c_ptr, c_int, shyhip_params
      type(c_ptr), value :: h
      type(shyhip_params), intent(out) :: p
      integer(c_int) :: rc
    end function

    function shyhip_set_params(h, p) bind(C, name="shyhip_set_params") result(rc)
      import :: c_ptr, c_int, shyhip_params
      type(c_ptr), value :: h
      type(shyhip_params), intent(in) :: p
      integer(c_int) :: rc
    end function

    function shyhip_set_basin(h, nen3v, xgv, ygv, hm3v, hev, hlv, hldv, ilhv, ilhkv) &
        bind(C, name="shyhip_set_basin") result(rc)
      import :: c_ptr, c_int, c_float
      type(c_ptr), value :: h
      integer(c_int), intent(in) :: nen3v(*), ilhv(*), ilhkv(*)
      real(c_float), intent(in) :: xgv(*), ygv(*), hm3v(*), hev(*), hlv(*), hldv(*)
      integer(c_int) :: rc
    end function

    function shyhip_set_boundaries(h, nbnd, ibnd, ibtyp, iextb) bind(C, name="shyhip_set_boundaries") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: h
      integer(c_int), value :: nbnd
      integer(c_int), intent(in) :: ibnd(*), ibtyp(*), iextb(*)
      integer(c_int) :: rc
    end function

    function shyhip_set_coefficients(h, fcorv, czv) bind(C, name="shyhip_set_coefficients") result(rc)
      import :: c_ptr, c_int, c_float
      type(c_ptr), value :: h
      real(c_float), intent(in) :: fcorv(*), czv(*)
      integer(c_int) :: rc
    end function

    function shyhip_tvd_init(h) bind(C, name="shyhip_tvd_init") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: h
      integer(c_int) :: rc
    end function

    function shyhip_init_state(h, zconst) bind(C, name="shyhip_init_state") result(rc)
      import :: c_ptr, c_int, c_float
      type(c_ptr), value :: h
      real(c_float), value :: zconst
      integer(c_int) :: rc
    end function

    function shyhip_push(h, field, host, dtype, n) bind(C, name="shyhip_push") result(rc)
      import :: c_ptr, c_int, c_size_t
      type(c_ptr), value :: h
      integer(c_int), value :: field, dtype
      type(c_ptr), value :: host
      integer(c_size_t), value :: n
      integer(c_int) :: rc
    end function

    function shyhip_pull(h, field, host, dtype, n) bind(C, name="shyhip_pull") result(rc)
      import :: c_ptr, c_int, c_size_t
      type(c_ptr), value :: h
      integer(c_int), value :: field, dtype
      type(c_ptr), value :: host
      integer(c_size_t), value :: n
      integer(c_int) :: rc
    end function

    function shyhip_set_boundary_level(h, ibc, level) bind(C, name="shyhip_set_boundary_level") result(rc)
      import :: c_ptr, c_int, c_float
      type(c_ptr), value :: h
      integer(c_int), value :: ibc
      real(c_float), value :: level
      integer(c_int) :: rc
    end function

    function shyhip_set_wind_uniform(h, taux, tauy) bind(C, name="shyhip_set_wind_uniform") result(rc)
      import :: c_ptr, c_int, c_float
      type(c_ptr), value :: h
      real(c_float), value :: taux, tauy
      integer(c_int) :: rc
    end function

    ! the boundary file of level boundary ibc and the wind file, held on the
    ! device and evaluated by every step (sp111(2))
    function shyhip_set_boundary_series(h, ibc, nrec, time, val, nval, intpol, zfact, dtime0) &
        bind(C, name="shyhip_set_boundary_series") result(rc)
      import :: c_ptr, c_int, c_float, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: ibc, nrec, nval, intpol
      real(c_double), intent(in) :: time(*)
      real(c_float), intent(in) :: val(*)
      real(c_float), value :: zfact
      real(c_double), value :: dtime0
      integer(c_int) :: rc
    end function

    function shyhip_set_wind_series(h, nrec, time, wx, wy, nval, intpol, dtime0, wp) &
        bind(C, name="shyhip_set_wind_series") result(rc)
      import :: c_ptr, c_int, c_float, c_double, shyhip_wind_params
      type(c_ptr), value :: h
      integer(c_int), value :: nrec, nval, intpol
      real(c_double), intent(in) :: time(*)
      real(c_float), intent(in) :: wx(*), wy(*)
      real(c_double), value :: dtime0
      type(shyhip_wind_params), intent(in) :: wp
      integer(c_int) :: rc
    end function

    ! baroclinic coupling: tracers itemp/isalt are T and S (ibarcl 1)
    function shyhip_set_baroclinic(h, ibarcl, itemp, isalt) bind(C, name="shyhip_set_baroclinic") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: h
      integer(c_int), value :: ibarcl, itemp, isalt
      integer(c_int) :: rc
    end function

    function shyhip_rhoset(h, niter) bind(C, name="shyhip_rhoset") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: h
      integer(c_int), value :: niter
      integer(c_int) :: rc
    end function

    function shyhip_set_sources(h, nsrc, node, levmin, levmax) bind(C, name="shyhip_set_sources") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: h
      integer(c_int), value :: nsrc
      integer(c_int), intent(in) :: node(*), levmin(*), levmax(*)
      integer(c_int) :: rc
    end function

    function shyhip_set_mass_flux(h, rflux, rqdsv) bind(C, name="shyhip_set_mass_flux") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: h
      type(c_ptr), value :: rflux, rqdsv
      integer(c_int) :: rc
    end function

    function shyhip_forcing(h, dtime) bind(C, name="shyhip_forcing") result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      real(c_double), value :: dtime
      integer(c_int) :: rc
    end function

    function shyhip_set_dry(h) bind(C, name="shyhip_set_dry") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: h
      integer(c_int) :: rc
    end function

    function shyhip_copy_uvz(h) bind(C, name="shyhip_copy_uvz") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: h
      integer(c_int) :: rc
    end function

    function shyhip_hydro(h, dtime) bind(C, name="shyhip_hydro") result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      real(c_double), value :: dtime
      integer(c_int) :: rc
    end function

    function shyhip_scal3sh(h, itr, istot, clip_small) bind(C, name="shyhip_scal3sh") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: h
      integer(c_int), value :: itr, istot, clip_small
      integer(c_int) :: rc
    end function

    function shyhip_set_tracer_terms(h, itr, robs, wsink, rload) bind(C, name="shyhip_set_tracer_terms") result(rc)
      import :: c_ptr, c_int, c_float
      type(c_ptr), value :: h
      integer(c_int), value :: itr
      real(c_float), value :: robs, wsink, rload
      integer(c_int) :: rc
    end function

    function shyhip_step(h, dtime, istot) bind(C, name="shyhip_step") result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      real(c_double), value :: dtime
      integer(c_int), value :: istot
      integer(c_int) :: rc
    end function

    function shyhip_sync(h) bind(C, name="shyhip_sync") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: h
      integer(c_int) :: rc
    end function
    function shyhip_set_deferred_checks(h, on) bind(C, name="shyhip_set_deferred_checks") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: h
      integer(c_int), value :: on
      integer(c_int) :: rc
    end function

    function shyhip_volume(h, mode, vol) bind(C, name="shyhip_volume") result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: mode
      real(c_double), intent(out) :: vol
      integer(c_int) :: rc
    end function

    function shyhip_stats(h, nloop, solver_iters, refinements) bind(C, name="shyhip_stats") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: h
      integer(c_int), intent(out) :: nloop, solver_iters, refinements
      integer(c_int) :: rc
    end function

    ! the zeta-system plugin (simsys_spkaux.f90:85-721)
    function shyhip_system_init(h) bind(C, name="shyhip_system_init") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: h
      integer(c_int) :: rc
    end function

    function shyhip_system_set_explicit(h) bind(C, name="shyhip_system_set_explicit") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: h
      integer(c_int) :: rc
    end function

    function shyhip_conz3d(h, itr, cbound, istot, isact) bind(C, name="shyhip_conz3d") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: h
      integer(c_int), value :: itr
      type(c_ptr), value :: cbound
      integer(c_int), value :: istot, isact
      integer(c_int) :: rc
    end function

    function shyhip_bndo_setbc(h, itr) bind(C, name="shyhip_bndo_setbc") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: h
      integer(c_int), value :: itr
      integer(c_int) :: rc
    end function

    function shyhip_get_sizes(h, nkn, nel, nlvdi, nlv, ntracer) bind(C, name="shyhip_get_sizes") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: h
      integer(c_int), intent(out) :: nkn, nel, nlvdi, nlv, ntracer
      integer(c_int) :: rc
    end function

    function shyhip_get_istot(h, istot) bind(C, name="shyhip_get_istot") result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: h
      integer(c_int), intent(out) :: istot
      integer(c_int) :: rc
    end function

    function shyhip_scalar_stability(h, rstol, istot, rindex) bind(C, name="shyhip_scalar_stability") result(rc)
      import :: c_ptr, c_int, c_float, c_double
      type(c_ptr), value :: h
      real(c_float), value :: rstol
      integer(c_int), intent(out) :: istot
      real(c_double), intent(out) :: rindex
      integer(c_int) :: rc
    end function

    function shyhip_system_assemble(h, ie, kn, hia, hik) bind(C, name="shyhip_system_assemble") result(rc)
      import :: c_ptr, c_int, c_float
      type(c_ptr), value :: h
      integer(c_int), value :: ie
      integer(c_int), intent(in) :: kn(3)
      real(c_float), intent(in) :: hia(3,3), hik(3)
      integer(c_int) :: rc
    end function

    function shyhip_system_add_rhs(h, dt, n, array) bind(C, name="shyhip_system_add_rhs") result(rc)
      import :: c_ptr, c_int, c_float
      type(c_ptr), value :: h
      real(c_float), value :: dt
      integer(c_int), value :: n
      real(c_float), intent(in) :: array(*)
      integer(c_int) :: rc
    end function

    function shyhip_system_solve(h, n, z) bind(C, name="shyhip_system_solve") result(rc)
      import :: c_ptr, c_int, c_float
      type(c_ptr), value :: h
      integer(c_int), value :: n
      real(c_float), intent(in) :: z(*)
      integer(c_int) :: rc
    end function

    function shyhip_system_get(h, n, z) bind(C, name="shyhip_system_get") result(rc)
      import :: c_ptr, c_int, c_float
      type(c_ptr), value :: h
      integer(c_int), value :: n
      real(c_float), intent(out) :: z(*)
      integer(c_int) :: rc
    end function

    function c_strlen(s) bind(C, name="strlen") result(n)
      import :: c_ptr, c_size_t
      type(c_ptr), value :: s
      integer(c_size_t) :: n
    end function

  end interface

contains

  ! error stop with the library's message when rc /= 0
  subroutine shyhip_check(h, rc, where)
    type(c_ptr), intent(in) :: h
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    type(c_ptr) :: p
    integer :: n
    character(len=:), allocatable :: text
    integer :: i
    if (rc == 0) return
    p = shyhip_last_error(h)
    n = int(c_strlen(p))
    call c_f_pointer(p, msg, [n])
    allocate(character(len=n) :: text)
    do i = 1, n
      text(i:i) = msg(i)
    end do
    write(6, *) 'error stop ', where, ': ', text
    error stop 'error stop shyhip'
  end subroutine shyhip_check

  ! push/pull of real (32-bit) and integer module arrays of any shape
  subroutine shyhip_push_r4(h, field, a, n)
    type(c_ptr), intent(in) :: h
    integer(c_int), intent(in) :: field
    integer, intent(in) :: n
    real(c_float), intent(in), target :: a(n)
    call shyhip_check(h, shyhip_push(h, field, c_loc(a), SHYHIP_F32, int(n, c_size_t)), 'shyhip_push')
  end subroutine shyhip_push_r4

  subroutine shyhip_pull_r4(h, field, a, n)
    type(c_ptr), intent(in) :: h
    integer(c_int), intent(in) :: field
    integer, intent(in) :: n
    real(c_float), intent(out), target :: a(n)
    call shyhip_check(h, shyhip_pull(h, field, c_loc(a), SHYHIP_F32, int(n, c_size_t)), 'shyhip_pull')
  end subroutine shyhip_pull_r4

  subroutine shyhip_pull_i4(h, field, a, n)
    type(c_ptr), intent(in) :: h
    integer(c_int), intent(in) :: field
    integer, intent(in) :: n
    integer(c_int), intent(out), target :: a(n)
    call shyhip_check(h, shyhip_pull(h, field, c_loc(a), SHYHIP_I32, int(n, c_size_t)), 'shyhip_pull')
  end subroutine shyhip_pull_i4

end module shyhip_capi
