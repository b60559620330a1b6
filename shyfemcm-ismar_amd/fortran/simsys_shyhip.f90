! simsys_shyhip.f90 -- the system_* solver plugin of SHYFEM on the GPU.
!
! Replaces src/shyfem/simsys_spkaux.f90 (SPARSKIT) at the link-time seam
! the reference already has for its solvers (SOLVER in src/shyfem/Makefile:
! 104-123): the stock hydro_zeta keeps assembling element by element through
! system_assemble(ie) with kn/hia/hik of mod_zeta_system, and the solve runs
! on the device (block ILU(0) + BiCGStab; FP32 inside FP64 refinement on the
! FP32 path).  Each plugin defines mod_zeta_system itself, as
! simsys_spkaux.f90:75-82 and simsys_petsc.f90:124-149 do; its variables
! keep the reference's names and types because hydro_zeta writes them.
! Errors end the run with the library's message (shyhip_check), as the
! reference's solver stops.  The 3D variants serve only the non-hydrostatic
! module and stop if called.

module mod_zeta_system
  implicit none
  integer kn(3)
  real, target :: hia(3,3)
  real, target :: hik(3)
  character*80 :: solver_type = 'shyhip'
end module mod_zeta_system

subroutine system_initialize
  ! the pattern and the solver's schedule were built by shyhip_set_basin
  use shyhip_shyfem
  use shyhip_capi
  implicit none
  if (.not. c_associated(shy)) stop 'error stop system_initialize: no device model (shyhip_shyfem_set)'
end subroutine system_initialize

subroutine system_init
  use shyhip_shyfem
  use shyhip_capi
  implicit none
  call shyhip_check(shy, shyhip_system_init(shy), 'system_init')
end subroutine system_init

subroutine system_set_explicit
  use shyhip_shyfem
  use shyhip_capi
  implicit none
  call shyhip_check(shy, shyhip_system_set_explicit(shy), 'system_set_explicit')
end subroutine system_set_explicit

subroutine system_assemble(ie)
  use mod_zeta_system, only : kn, hia, hik
  use shyhip_shyfem
  use shyhip_capi
  implicit none
  integer ie
  call shyhip_check(shy, shyhip_system_assemble(shy, ie, kn, hia, hik), 'system_assemble')
end subroutine system_assemble

subroutine system_add_rhs(dt, n, array)
  use shyhip_shyfem
  use shyhip_capi
  implicit none
  real dt
  integer n
  real array(n)
  call shyhip_check(shy, shyhip_system_add_rhs(shy, dt, n, array), 'system_add_rhs')
end subroutine system_add_rhs

subroutine system_solve(n, z)
  ! z is the initial guess, not the solution (system_get returns that)
  use shyhip_shyfem
  use shyhip_capi
  implicit none
  integer, intent(in) :: n
  real, intent(in) :: z(n)
  call shyhip_check(shy, shyhip_system_solve(shy, n, z), 'system_solve')
end subroutine system_solve

subroutine system_get(n, z)
  use shyhip_shyfem
  use shyhip_capi
  implicit none
  integer n
  real z(n)
  call shyhip_check(shy, shyhip_system_get(shy, n, z), 'system_get')
end subroutine system_get

subroutine system_finalize
  implicit none
end subroutine system_finalize

subroutine system_solve_3d(n, nlvdi, nlv, z)
  implicit none
  integer n, nlvdi, nlv
  real z(nlvdi, n)
  stop 'error stop system_solve_3d: the non-hydrostatic system is not on the device'
end subroutine system_solve_3d

subroutine system_assemble_3d(ie, l, nlv, kn, mass, rhs)
  implicit none
  integer ie, l, nlv
  integer kn(3)
  real mass(-1:1,3,3)
  real rhs(3)
  stop 'error stop system_assemble_3d: the non-hydrostatic system is not on the device'
end subroutine system_assemble_3d

subroutine system_get_3d(n, nlvdi, nlv, z)
  implicit none
  integer n, nlvdi, nlv
  real z(nlvdi, n)
  stop 'error stop system_get_3d: the non-hydrostatic system is not on the device'
end subroutine system_get_3d

subroutine system_adjust_matrix_3d
  implicit none
  stop 'error stop system_adjust_matrix_3d: the non-hydrostatic system is not on the device'
end subroutine system_adjust_matrix_3d
