! conz3d_shyhip.f90 -- conz3d_omp on the GPU, with the reference's name and
! argument list (src/shyfem/concentration_omp.f90:65-151), so that scal3sh
! (concentration.f90:748-760) calls it unchanged for each substep isact of
! istot.  The hydrodynamic state it transports with is the one the device
! model holds (after shyhip_hydro or a device step).  One call: push cn1 and
! difv into the model, run the element and node phases of substep isact on
! the device, pull cn1 back.  The model's tracer slot shy_conz_slot
! (shyhip_shyfem) holds the concentration meanwhile.
!
! Horizontal diffusion takes rkpar and difhv (pushed with each call) and the
! model's idtype; nudging (robs, cobs, rtauv), settling (wsink, wsinkv) and
! loads (rload, load) go to the slot's tracer terms, and the loads that
! deposition leaves come back in load.
! With itvd = 1 the device computes the node gradients of cn1 itself (the
! gradxv, gradyv that tvd_grad_3d computed from the same cn1 just before
! this call, concentration.f90:744), so gradxv and gradyv are not read.  azpar, adpar, aapar, difmol, itvd and itvdv are taken
! over into the model's parameters; ddt must be the model's time step.
! The reference runs the T and S calls in concurrent OpenMP tasks
! (baroclinic.f90:466-500); here they take turns on the one model.

subroutine conz3d_omp(cn1, co1, ddt, rkpar, difhv, difv, difmol, cbound, itvd, itvdv, gradxv, gradyv, &
                      cobs, robs, rtauv, wsink, wsinkv, rload, load, azpar, adpar, aapar, istot, isact, &
                      nlvddi, nlev)
  use iso_c_binding
  use shyhip_capi
  use shyhip_shyfem
  implicit none
  integer, intent(in) :: nlvddi, nlev, itvd, itvdv, istot, isact
  real, intent(in) :: difmol, robs, wsink, rload, ddt, rkpar
  real, intent(in) :: azpar, adpar, aapar
  real, intent(inout), target :: cn1(nlvddi,*)
  real, intent(in) :: co1(nlvddi,*)
  real, intent(in), target :: cbound(nlvddi,*)
  real, intent(in), target :: difhv(nlvddi,*)
  real, intent(in) :: gradxv(nlvddi,*), gradyv(nlvddi,*)
  real, intent(in), target :: cobs(nlvddi,*), rtauv(nlvddi,*)
  real, intent(inout), target :: load(nlvddi,*)
  real, intent(in), target :: difv(0:nlvddi,*)
  real, intent(in), target :: wsinkv(0:nlvddi,*)

  type(shyhip_params) :: p
  integer(c_int) :: nkn, nel, nlvdi, nlv, ntr
  logical :: change


!$omp critical (shyhip_conz3d)
  call shyhip_check(shy, shyhip_get_sizes(shy, nkn, nel, nlvdi, nlv, ntr), 'conz3d_omp')
  if (nlvddi /= nlvdi .or. nlev /= nlv) stop 'error stop conz3d_omp: layers differ from the device model'
  if (shy_conz_slot < 0 .or. shy_conz_slot >= ntr) stop 'error stop conz3d_omp: no tracer slot in the device model'
  call shyhip_check(shy, shyhip_get_params(shy, p), 'conz3d_omp')
  if (real(p%dt) /= ddt) stop 'error stop conz3d_omp: ddt differs from the time step of the device model'
  change = p%azpar /= azpar .or. p%adpar /= adpar .or. p%aapar /= aapar .or. p%difmol /= difmol &
           .or. p%itvd /= itvd .or. p%itvdv /= itvdv .or. p%rkpar /= rkpar
  if (change) then
    p%azpar = azpar
    p%adpar = adpar
    p%aapar = aapar
    p%difmol = difmol
    p%itvd = itvd
    p%itvdv = itvdv
    p%rkpar = rkpar
    call shyhip_check(shy, shyhip_set_params(shy, p), 'conz3d_omp')
  end if
  call shyhip_check(shy, shyhip_push(shy, SHYHIP_TRACER0 + shy_conz_slot, c_loc(cn1), SHYHIP_F32, &
                                     int(nlvddi, c_size_t) * nkn), 'conz3d_omp')
  call shyhip_check(shy, shyhip_push(shy, SHYHIP_DIFV, c_loc(difv), SHYHIP_F32, &
                                     int(nlvddi + 1, c_size_t) * nkn), 'conz3d_omp')
  if (rkpar /= 0.) call shyhip_check(shy, shyhip_push(shy, SHYHIP_DIFHV, c_loc(difhv), SHYHIP_F32, &
                                                      int(nlvddi, c_size_t) * nel), 'conz3d_omp')
  if (robs /= 0.) then
    call shyhip_check(shy, shyhip_push(shy, SHYHIP_COBS0 + shy_conz_slot, c_loc(cobs), SHYHIP_F32, &
                                       int(nlvddi, c_size_t) * nkn), 'conz3d_omp')
    call shyhip_check(shy, shyhip_push(shy, SHYHIP_RTAUV0 + shy_conz_slot, c_loc(rtauv), SHYHIP_F32, &
                                       int(nlvddi, c_size_t) * nkn), 'conz3d_omp')
  end if
  if (wsink /= 0.) call shyhip_check(shy, shyhip_push(shy, SHYHIP_WSINKV0 + shy_conz_slot, c_loc(wsinkv), &
                                                      SHYHIP_F32, int(nlvddi + 1, c_size_t) * nkn), 'conz3d_omp')
  if (rload /= 0.) call shyhip_check(shy, shyhip_push(shy, SHYHIP_LOAD0 + shy_conz_slot, c_loc(load), SHYHIP_F32, &
                                                      int(nlvddi, c_size_t) * nkn), 'conz3d_omp')
  call shyhip_check(shy, shyhip_set_tracer_terms(shy, shy_conz_slot, robs, wsink, rload), 'conz3d_omp')
  call shyhip_check(shy, shyhip_conz3d(shy, shy_conz_slot, c_loc(cbound), istot, isact), 'conz3d_omp')
  if (rload /= 0.) call shyhip_check(shy, shyhip_pull(shy, SHYHIP_LOAD0 + shy_conz_slot, c_loc(load), SHYHIP_F32, &
                                                      int(nlvddi, c_size_t) * nkn), 'conz3d_omp')
  call shyhip_check(shy, shyhip_pull(shy, SHYHIP_TRACER0 + shy_conz_slot, c_loc(cn1), SHYHIP_F32, &
                                     int(nlvddi, c_size_t) * nkn), 'conz3d_omp')
!$omp end critical (shyhip_conz3d)
end subroutine conz3d_omp
