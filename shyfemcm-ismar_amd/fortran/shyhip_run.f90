! shyhip_run.f90 -- a Fortran host program that runs the hydro step and the
! scalar transport on the GPU through the C ABI, in the order of SHYFEM's
! time loop (shyfem_subs.f90:577-607): set_dry, set_timestep, copy_uvz and
! copy_depth, boundary levels (sp111), hydro, scal3sh per tracer.
!
! usage: shyhip_run <case file> <result file>
!
! case file: see shyhip_case.f90
! result file (stream):
!   real32 znv(nkn), utlnv(nlv,nel), conc(nlv,nkn,ntracer)
!   int32  iwegv(nel)
!   real64 volume

program shyhip_run

  use iso_c_binding
  use shyhip_capi
  use shyhip_case
  implicit none

  integer :: i, it, n
  real(c_double) :: dtime, vol
  integer(c_int), allocatable :: iwegv(:)
  real(c_float), allocatable :: znv(:), utlnv(:,:)
  type(c_ptr) :: h
  character(len=512) :: fin, fout
  integer :: u

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  call case_open(fin, h)

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
