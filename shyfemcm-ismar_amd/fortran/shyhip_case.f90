! shyhip_case.f90 -- reads a basin case file and sets up a device model the
! way a SHYFEM host does after its initialisation (shyfem_subs.f90:392-453):
! parameters, basin, boundaries, coefficients, viscosity/diffusivity, TVD
! tables, the start state, the tracers with their open-boundary values and a
! uniform wind.  Shared by the Fortran host programs shyhip_run and
! shyhip_dropin_check.
!
! case file (stream, native byte order):
!   int32  nkn, nel, nlv, nbnd, nsteps, precision, ntracer, ireib, itvd, itvdv
!   real64 dt
!   real32 aapar, czv, visv, difv, taux, tauy
!   int32  nen3v(3,nel)  (1-based)
!   real32 xgv(nkn), ygv(nkn), hm3v(3,nel), hev(nel), hlv(nlv), hldv(nlv)
!   int32  ilhv(nel), ilhkv(nkn), ibnd(nkn), ibtyp(0:nbnd), iextb(nkn)
!   real32 level(nsteps)                  open-boundary level of boundary 1
!   real32 conc(nlv,nkn,ntracer)          initial tracers

module shyhip_case

  use iso_c_binding
  use shyhip_capi
  implicit none

  integer :: nkn, nel, nlv, nbnd, nsteps, ntracer
  real(c_double) :: dt
  integer(c_int), allocatable :: nen3v(:,:), ilhv(:), ilhkv(:), ibnd(:), ibtyp(:), iextb(:)
  real(c_float), allocatable :: level(:), conc(:,:,:)

contains

  subroutine case_open(path, h)
    character(*), intent(in) :: path
    type(c_ptr), intent(out) :: h
    integer(c_int) :: hdr(10), precision
    real(c_float) :: rpar(6)
    real(c_float), allocatable :: xgv(:), ygv(:), hm3v(:,:), hev(:), hlv(:), hldv(:)
    real(c_float), allocatable :: bc(:,:), visv(:,:), difv(:,:), fcorv(:), czv(:)
    type(shyhip_params) :: p
    integer :: u, i, k

    open(newunit=u, file=trim(path), access='stream', form='unformatted', status='old')
    read(u) hdr
    nkn = hdr(1); nel = hdr(2); nlv = hdr(3); nbnd = hdr(4); nsteps = hdr(5)
    precision = hdr(6); ntracer = hdr(7)
    read(u) dt
    read(u) rpar
    allocate(nen3v(3,nel), xgv(nkn), ygv(nkn), hm3v(3,nel), hev(nel), hlv(nlv), hldv(nlv))
    allocate(ilhv(nel), ilhkv(nkn), ibnd(nkn), ibtyp(0:nbnd), iextb(nkn), level(nsteps))
    allocate(conc(nlv,nkn,max(1,ntracer)))
    read(u) nen3v
    read(u) xgv, ygv, hm3v, hev, hlv, hldv
    read(u) ilhv, ilhkv, ibnd, ibtyp, iextb
    read(u) level
    if (ntracer > 0) read(u) conc(:,:,1:ntracer)
    close(u)

    call shyhip_check(c_null_ptr, shyhip_create(h, precision, nkn, nel, nlv, nlv, ntracer), 'shyhip_create')
    call shyhip_check(h, shyhip_get_params(h, p), 'shyhip_get_params')
    p%dt = dt
    p%ireib = hdr(8); p%itvd = hdr(9); p%itvdv = hdr(10)
    p%aapar = rpar(1)
    call shyhip_check(h, shyhip_set_params(h, p), 'shyhip_set_params')
    call shyhip_check(h, shyhip_set_basin(h, nen3v, xgv, ygv, hm3v, hev, hlv, hldv, ilhv, ilhkv), 'shyhip_set_basin')
    call shyhip_check(h, shyhip_set_boundaries(h, nbnd, ibnd, ibtyp, iextb), 'shyhip_set_boundaries')
    allocate(fcorv(nel), czv(nel))
    fcorv = 0.
    czv = rpar(2)
    call shyhip_check(h, shyhip_set_coefficients(h, fcorv, czv), 'shyhip_set_coefficients')
    allocate(visv(0:nlv,nkn), difv(0:nlv,nkn))
    visv = rpar(3)
    difv = rpar(4)
    call shyhip_push_r4(h, SHYHIP_VISV, visv, size(visv))
    call shyhip_push_r4(h, SHYHIP_DIFV, difv, size(difv))
    if (p%itvd == 2) call shyhip_check(h, shyhip_tvd_init(h), 'shyhip_tvd_init')
    call shyhip_check(h, shyhip_init_state(h, 0.), 'shyhip_init_state')

    ! tracers and their open-boundary values (the initial values on the
    ! boundary nodes, -999 elsewhere)
    allocate(bc(nlv,nkn))
    do i = 1, ntracer
      call shyhip_push_r4(h, SHYHIP_TRACER0 + i - 1, conc(:,:,i), nlv*nkn)
      bc = -999.
      do k = 1, nkn
        if (ibnd(k) > 0) bc(:,k) = conc(:,k,i)
      end do
      call shyhip_push_r4(h, SHYHIP_TRACER_BC0 + i - 1, bc, nlv*nkn)
    end do
    call shyhip_check(h, shyhip_set_wind_uniform(h, rpar(5), rpar(6)), 'shyhip_set_wind_uniform')
  end subroutine case_open

end module shyhip_case
