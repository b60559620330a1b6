! shy_dump.f90 -- TEST INFRASTRUCTURE ONLY.
!
! Reads a SHY file through the reference's own reader (module shyfile,
! src/utils/shyutil/io_shy.f90, with src/utils/generic/to_linear.f90; built
! into oracle/_ref by ref.mk) and prints what that reader returns as text,
! so the tests can check files written by the library against it.
!
!   shy_dump file.shy > dump.txt
!
! Output: "params ftype nkn nel npr nlv nvar nfix date time", "simpar ...",
! "title|...|", "femver|...|", one line per header array, then for every
! data record "record dtime ivar n m lmax" and "values ..." (a 3D record
! over each node's or element's own layers).  A hydro file (ftype 1) holds
! per time record one node record and three element records
! (shy_write_hydro_records), a scalar file node records only: that tells the
! reader which records are on elements.
program shy_dump
  use shyfile
  implicit none
  character(len=512) :: file
  character(len=80) :: title, femver
  integer :: id, ierr, nkn, nel, npr, nlv, nvar, ftype, nfix, date, time
  integer :: ivar, n, m, lmax, i, irec, nmax
  real :: simpar(4)
  integer, allocatable :: nen3v(:, :), ipev(:), ipv(:), iarv(:), iarnv(:), ilhv(:), ilhkv(:), il(:)
  real, allocatable :: xgv(:), ygv(:), hm3v(:, :), hlv(:), c(:, :)
  double precision :: dtime
  logical :: belem

  call get_command_argument(1, file)
  id = shy_init_by_file(trim(file))
  if (id == 0) stop 'shy_dump: cannot open file'
  call shy_read_header(id, ierr)
  if (ierr /= 0) stop 'shy_dump: cannot read header'
  call shy_get_params(id, nkn, nel, npr, nlv, nvar)
  call shy_get_ftype(id, ftype)
  call shy_get_nfix(id, nfix)
  call shy_get_date(id, date, time)
  call shy_get_title(id, title)
  call shy_get_femver(id, femver)
  call shy_get_simpar(id, simpar)
  write (*, '(a,*(1x,i0))') 'params', ftype, nkn, nel, npr, nlv, nvar, nfix, date, time
  write (*, '(a,*(1x,es16.8e3))') 'simpar', simpar
  write (*, '(a)') 'title|' // title // '|'
  write (*, '(a)') 'femver|' // femver // '|'

  allocate (nen3v(3, nel), ipev(nel), ipv(nkn), iarv(nel), iarnv(nkn), ilhv(nel), ilhkv(nkn))
  allocate (xgv(nkn), ygv(nkn), hm3v(3, nel), hlv(nlv))
  call shy_get_elemindex(id, nen3v)
  call shy_get_extnumbers(id, ipev, ipv)
  call shy_get_areacode(id, iarv, iarnv)
  call shy_get_coords(id, xgv, ygv)
  call shy_get_depth(id, hm3v)
  call shy_get_layers(id, hlv)
  call shy_get_layerindex(id, ilhv, ilhkv)
  write (*, '(a,*(1x,i0))') 'nen3v', nen3v
  write (*, '(a,*(1x,i0))') 'ipev', ipev
  write (*, '(a,*(1x,i0))') 'ipv', ipv
  write (*, '(a,*(1x,i0))') 'iarv', iarv
  write (*, '(a,*(1x,i0))') 'iarnv', iarnv
  write (*, '(a,*(1x,es16.8e3))') 'xgv', xgv
  write (*, '(a,*(1x,es16.8e3))') 'ygv', ygv
  write (*, '(a,*(1x,es16.8e3))') 'hm3v', hm3v
  write (*, '(a,*(1x,es16.8e3))') 'hlv', hlv
  write (*, '(a,*(1x,i0))') 'ilhv', ilhv
  write (*, '(a,*(1x,i0))') 'ilhkv', ilhkv

  nmax = max(nkn, 3 * nel)
  allocate (c(nlv, nmax))
  irec = 0
  do
    belem = .false.
    if (ftype == 1) belem = mod(irec, 4) /= 0
    call shy_read_record(id, dtime, belem, ivar, n, m, lmax, nlv, c, ierr)
    if (ierr /= 0) exit
    irec = irec + 1
    write (*, '(a,1x,es24.16e3,*(1x,i0))') 'record', dtime, ivar, n, m, lmax
    if (lmax <= 1) then
      write (*, '(a,*(1x,es16.8e3))') 'values', c(1, 1:n * m)
    else
      allocate (il(n))
      if (belem) then
        il = ilhv
      else
        il = ilhkv
      end if
      if (nfix > 0) il = nfix
      write (*, '(a)', advance='no') 'values'
      do i = 1, n
        write (*, '(*(1x,es16.8e3))', advance='no') c(1:min(lmax, il(i)), i)
      end do
      write (*, '(a)') ''
      deallocate (il)
    end if
  end do
  write (*, '(a,1x,i0)') 'end', irec
end program shy_dump
