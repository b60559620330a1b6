! shyhip_dropin_check.f90 -- calls the drop-in routines by the reference's
! names, as SHYFEM's own code calls them, so a maintainer (and the tests)
! can check that they link and compute what the C ABI computes.
!
! usage:
!   shyhip_dropin_check plugin <case> <systems> <result>
!     The zeta-system plugin (simsys_shyhip.f90) as hydro_zeta drives it
!     (hydrodynamic.f90:488-754): system_init, then per element kn/hia/hik
!     in mod_zeta_system and system_assemble(ie), system_add_rhs(dt, nkn,
!     rq), system_solve(nkn, z0), system_get(nkn, z).  With explicit /= 0,
!     system_set_explicit first.
!     systems (stream): int32 explicit; real32 dt; real32 hia(3,3,nel),
!     hik(3,nel), rq(nkn), z0(nkn).   result: real32 z(nkn).
!   shyhip_dropin_check transport <case> <istot> <result>
!     The time loop of shyhip_run, with each tracer transported as scal3sh
!     does it (concentration.f90:728-772), through conz3d_omp
!     (conz3d_shyhip.f90): per substep isact of istot, clip small values
!     (not for the first tracer, temperature), conz3d_omp, then the
!     open-boundary treatment (shyhip_bndo_setbc).  result: as shyhip_run's.

program shyhip_dropin_check

  use iso_c_binding
  use shyhip_capi
  use shyhip_case
  use shyhip_shyfem
  use mod_zeta_system, only : kn, hia, hik
  implicit none

  character(len=512) :: mode, fin, arg3, fout
  type(c_ptr) :: h
  integer :: u

  call get_command_argument(1, mode)
  call get_command_argument(2, fin)
  call get_command_argument(3, arg3)
  call get_command_argument(4, fout)
  call case_open(fin, h)
  call shyhip_shyfem_set(h, 0)

  select case (trim(mode))
  case ('plugin')
    call run_plugin()
  case ('transport')
    call run_transport()
  case default
    stop 'error stop shyhip_dropin_check: mode must be plugin or transport'
  end select
  call shyhip_check(h, shyhip_destroy(h), 'shyhip_destroy')

contains

  subroutine run_plugin()
    integer(c_int) :: iexpl
    real(c_float) :: sdt
    real(c_float), allocatable :: ehia(:,:,:), ehik(:,:), rq(:), z0(:), z(:)
    integer :: ie
    allocate(ehia(3,3,nel), ehik(3,nel), rq(nkn), z0(nkn), z(nkn))
    open(newunit=u, file=trim(arg3), access='stream', form='unformatted', status='old')
    read(u) iexpl, sdt
    read(u) ehia, ehik, rq, z0
    close(u)
    call system_initialize
    if (iexpl /= 0) call system_set_explicit
    call system_init
    do ie = 1, nel
      kn = nen3v(:,ie)
      hia = ehia(:,:,ie)
      hik = ehik(:,ie)
      call system_assemble(ie)
    end do
    call system_add_rhs(sdt, nkn, rq)
    call system_solve(nkn, z0)
    call system_get(nkn, z)
    call system_finalize
    open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
    write(u) z
    close(u)
  end subroutine run_plugin

  subroutine run_transport()
    integer :: istot, isact, i, it
    real(c_double) :: dtime, vol
    type(shyhip_params) :: p
    real(c_float), allocatable :: cn(:,:), co(:,:), cb(:,:), difhv(:,:), gx(:,:), gy(:,:), cobs(:,:), rtau(:,:)
    real(c_float), allocatable :: load(:,:), difv(:,:), wsinkv(:,:), znv(:), utlnv(:,:)
    integer(c_int), allocatable :: iwegv(:)
    read(arg3, *) istot
    call shyhip_check(h, shyhip_get_params(h, p), 'shyhip_get_params')
    allocate(cn(nlv,nkn), co(nlv,nkn), cb(nlv,nkn), difhv(nlv,nel), gx(nlv,nkn), gy(nlv,nkn))
    allocate(cobs(nlv,nkn), rtau(nlv,nkn), load(nlv,nkn), difv(0:nlv,nkn), wsinkv(0:nlv,nkn))
    cb = 0.; difhv = 0.; gx = 0.; gy = 0.; cobs = 0.; rtau = 0.; load = 0.; wsinkv = 0.
    call shyhip_pull_r4(h, SHYHIP_DIFV, difv, size(difv))
    dtime = 0.
    do it = 1, nsteps
      call shyhip_check(h, shyhip_set_dry(h), 'set_dry')
      dtime = dtime + dt
      call shyhip_check(h, shyhip_copy_uvz(h), 'copy_uvz')
      if (nbnd > 0) call shyhip_check(h, shyhip_set_boundary_level(h, 1, level(it)), 'sp111')
      call shyhip_check(h, shyhip_hydro(h, dtime), 'hydro')
      do i = 1, ntracer
        call shyhip_shyfem_set(h, i - 1)
        call shyhip_pull_r4(h, SHYHIP_TRACER0 + i - 1, cn, nlv*nkn)
        do isact = 1, istot
          if (i > 1) where (cn < 1.e-15) cn = 0.
          co = cn
          call conz3d_omp(cn, co, real(dt), p%rkpar, difhv, difv, p%difmol, cb, p%itvd, p%itvdv, gx, gy, &
                          cobs, 0., rtau, 0., wsinkv, 0., load, p%azpar, p%adpar, p%aapar, istot, isact, &
                          nlv, nlv)
          call shyhip_check(h, shyhip_bndo_setbc(h, i - 1), 'bndo_setbc')
          call shyhip_pull_r4(h, SHYHIP_TRACER0 + i - 1, cn, nlv*nkn)
        end do
      end do
    end do
    allocate(znv(nkn), utlnv(nlv,nel), iwegv(nel))
    call shyhip_pull_r4(h, SHYHIP_ZNV, znv, nkn)
    call shyhip_pull_r4(h, SHYHIP_UTLNV, utlnv, nlv*nel)
    do i = 1, ntracer
      call shyhip_pull_r4(h, SHYHIP_TRACER0 + i - 1, conc(:,:,i), nlv*nkn)
    end do
    call shyhip_pull_i4(h, SHYHIP_IWEGV, iwegv, nel)
    call shyhip_check(h, shyhip_volume(h, 1, vol), 'shyhip_volume')
    open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
    write(u) znv, utlnv
    if (ntracer > 0) write(u) conc(:,:,1:ntracer)
    write(u) iwegv, vol
    close(u)
  end subroutine run_transport

end program shyhip_dropin_check
