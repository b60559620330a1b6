! shyhip_run.f90 -- a Fortran host program that runs the hydro step and the
! scalar transport on the GPU through the C ABI, in the order of SHYFEM's
! time loop (shyfem_subs.f90:577-607): set_dry, set_timestep, copy_uvz and
! copy_depth, boundary levels (sp111), hydro, scal3sh per tracer.
!
! usage: shyhip_run <case file> <result file>
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
! result file (stream):
!   real32 znv(nkn), utlnv(nlv,nel), conc(nlv,nkn,ntracer)
!   int32  iwegv(nel)
!   real64 volume

program shyhip_run

  use iso_c_binding
  use shyhip_capi
  implicit none

  integer(c_int) :: hdr(10)
  integer :: nkn, nel, nlv, nbnd, nsteps, ntracer, i, k, it, n
  integer(c_int) :: precision
  real(c_double) :: dt, dtime, vol
  real(c_float) :: rpar(6)
  integer(c_int), allocatable :: nen3v(:,:), ilhv(:), ilhkv(:), ibnd(:), ibtyp(:), iextb(:), iwegv(:)
  real(c_float), allocatable :: xgv(:), ygv(:), hm3v(:,:), hev(:), hlv(:), hldv(:), level(:)
  real(c_float), allocatable :: conc(:,:,:), bc(:,:), visv(:,:), difv(:,:), fcorv(:), czv(:)
  real(c_float), allocatable :: znv(:), utlnv(:,:)
  type(c_ptr) :: h
  type(shyhip_params) :: p
  character(len=512) :: fin, fout
  integer :: u

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)

  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
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

  ! life cycle: create, parameters (defaults + case), basin, boundaries
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

  ! time loop
  dtime = 0.
  do it = 1, nsteps
    call shyhip_check(h, shyhip_set_dry(h), 'set_dry')
    dtime = dtime + dt
    call shyhip_check(h, shyhip_copy_uvz(h), 'copy_uvz')
    if (nbnd > 0) call shyhip_check(h, shyhip_set_boundary_level(h, 1, level(it)), 'sp111')
    call shyhip_check(h, shyhip_hydro(h, dtime), 'hydro')
    do i = 1, ntracer
      n = 1
      if (i == 1) n = 0
      call shyhip_check(h, shyhip_scal3sh(h, i - 1, 1, n), 'scal3sh')
    end do
  end do
  call shyhip_check(h, shyhip_sync(h), 'shyhip_sync')

  allocate(znv(nkn), utlnv(nlv,nel), iwegv(nel))
  call shyhip_pull_r4(h, SHYHIP_ZNV, znv, nkn)
  call shyhip_pull_r4(h, SHYHIP_UTLNV, utlnv, nlv*nel)
  do i = 1, ntracer
    call shyhip_pull_r4(h, SHYHIP_TRACER0 + i - 1, conc(:,:,i), nlv*nkn)
  end do
  call shyhip_pull_i4(h, SHYHIP_IWEGV, iwegv, nel)
  call shyhip_check(h, shyhip_volume(h, 1, vol), 'shyhip_volume')
  call shyhip_check(h, shyhip_destroy(h), 'shyhip_destroy')

  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) znv, utlnv
  if (ntracer > 0) write(u) conc(:,:,1:ntracer)
  write(u) iwegv, vol
  close(u)
  write(6, '(a,i0,a,es12.4)') 'shyhip_run: ', nsteps, ' steps, volume ', vol

end program shyhip_run
