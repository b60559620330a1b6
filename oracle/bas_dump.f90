! bas_dump.f90 -- TEST INFRASTRUCTURE ONLY.
!
! Reads a basin file (.bas) through the reference's own reader (module
! basin, src/utils/shyutil/basin.f90, with src/utils/generic/file.f90;
! built into oracle/_ref by ref.mk) and prints what that reader returns as
! text, so the tests can check files written by the library against it.
!
!   bas_dump file.bas > dump.txt
!
! Output: "params nkn nel ngr mbw", "geom dcorbas dirnbas sphebas",
! "descr|...|", then one line per array (nen3v ipv ipev iarv xgv ygv hm3v).
program bas_dump
  use basin
  implicit none
  character(len=512) :: file
  integer, parameter :: iunit = 21     ! basin_test reads positive units only

  call get_command_argument(1, file)
  open (unit=iunit, file=trim(file), form='unformatted', status='old', action='read')
  call basin_read_by_unit(iunit)
  close (iunit)
  write (*, '(a,*(1x,i0))') 'params', nkn, nel, ngr, mbw
  write (*, '(a,2(1x,es16.8e3),1x,i0)') 'geom', dcorbas, dirnbas, sphebas
  write (*, '(a)') 'descr|' // descrr // '|'
  write (*, '(a,*(1x,i0))') 'nen3v', nen3v(:, 1:nel)
  write (*, '(a,*(1x,i0))') 'ipv', ipv(1:nkn)
  write (*, '(a,*(1x,i0))') 'ipev', ipev(1:nel)
  write (*, '(a,*(1x,i0))') 'iarv', iarv(1:nel)
  write (*, '(a,*(1x,es16.8e3))') 'xgv', xgv(1:nkn)
  write (*, '(a,*(1x,es16.8e3))') 'ygv', ygv(1:nkn)
  write (*, '(a,*(1x,es16.8e3))') 'hm3v', hm3v(:, 1:nel)
end program bas_dump
