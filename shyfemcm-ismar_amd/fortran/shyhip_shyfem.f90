! shyhip_shyfem.f90 -- the device model a SHYFEM build hands to the drop-in
! routines (simsys_shyhip.f90, conz3d_shyhip.f90).  The host creates the
! model through shyhip_capi after its initialisation (shyfem_subs.f90:
! 392-453) and stores it here with shyhip_shyfem_set.

module shyhip_shyfem

  use iso_c_binding
  implicit none

  type(c_ptr), save :: shy = c_null_ptr       ! the basin's model on the GPU
  integer, save :: shy_conz_slot = 0          ! tracer slot conz3d_omp works in

contains

  subroutine shyhip_shyfem_set(h, conz_slot)
    type(c_ptr), intent(in) :: h
    integer, intent(in), optional :: conz_slot
    shy = h
    if (present(conz_slot)) shy_conz_slot = conz_slot
  end subroutine shyhip_shyfem_set

end module shyhip_shyfem
