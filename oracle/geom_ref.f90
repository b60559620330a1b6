! geom_ref.f90 -- TEST INFRASTRUCTURE ONLY.
!
! Runs the reference's own init-time geometry routines on a basin given in
! a binary file, and writes what they return, so the tests can check the C
! oracle's restatements bit for bit:
!   set_ev                       (src/utils/shyutil/evgeom.f90:164-345)
!   domain_clusterization_ggu    (src/shyfem/subset.f90:371-711)
!   quad_tree_search             (src/utils/shyutil/quad_tree.f90:313-397),
!     called at the mirrored points of tvd_upwind_init_elem
!     (src/shyfem/tvd_admin.f90:361-474; the point arithmetic of that
!     routine is repeated here, as it needs the model's modules)
!   femintp_xi                   (src/utils/shyutil/evgeom.f90:1413-1445),
!     the upwind value of tvd_get_upwind_c (tvd_admin.f90:608-669)
! Built by ref.mk with the reference's sources into oracle/_ref/geom_ref.
!
!   geom_ref in.bin out.bin
!
! in.bin (stream, little endian): nkn, nel, ngr (int32); xgv(nkn),
!   ygv(nkn) (real); nen3v(3,nel) (int32, 1-based); cv(nkn) (real).
! out.bin: ev(19,nel) (double); ncolor (int32); color(nel) (int32);
!   ietvd(3,3,nel) (int32, 0 = none); xtvd(3,3,nel), ytvd(3,3,nel) (real);
!   cu(3,3,nel) (real, femintp_xi of cv at the upwind point, 0 if none).
program geom_ref
  use basin
  use evgeom
  use mod_subset
  use mod_quad_tree
  implicit none
  character(len=512) :: fin, fout
  integer :: nk, ne, ng, ie, ii, j, k, ieq, ic, n, t
  real, allocatable :: cv(:)
  integer, allocatable :: color(:), ietvd(:,:,:)
  real, allocatable :: xtvd(:,:,:), ytvd(:,:,:), cu(:,:,:)
  double precision :: xc, yc, xd, yd, xu, yu
  real :: x, y, c(3), val

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open (unit=21, file=trim(fin), access='stream', form='unformatted', status='old', action='read')
  read (21) nk, ne, ng
  call basin_init(nk, ne)
  ngr = ng
  sphebas = 0
  read (21) xgv(1:nk)
  read (21) ygv(1:nk)
  read (21) nen3v(:, 1:ne)
  allocate (cv(nk))
  read (21) cv
  close (21)
  do ie = 1, ne
    ipev(ie) = ie
  end do
  do k = 1, nk
    ipv(k) = k
  end do
  iarv = 0
  iarnv = 0
  hm3v = 0.

  call set_ev

  call domain_clusterization_ggu
  allocate (color(ne))
  color = 0
  do ic = 1, subset_num
    do n = 1, subset_el(ic)
      color(indipendent_subset(n, ic)) = ic
    end do
  end do

  call quad_tree_initialize
  allocate (ietvd(3, 3, ne), xtvd(3, 3, ne), ytvd(3, 3, ne), cu(3, 3, ne))
  ietvd = 0
  xtvd = 0.
  ytvd = 0.
  cu = 0.
  do ie = 1, ne
    do ii = 1, 3
      k = nen3v(ii, ie)
      xc = xgv(k)
      yc = ygv(k)
      do t = 1, 2
        if (t == 1) then
          j = mod(ii, 3) + 1
        else
          j = mod(ii + 1, 3) + 1
        end if
        k = nen3v(j, ie)
        xd = xgv(k)
        yd = ygv(k)
        xu = 2*xc - xd
        yu = 2*yc - yd
        x = xu
        y = yu
        call quad_tree_search(x, y, ieq)
        xtvd(j, ii, ie) = x
        ytvd(j, ii, ie) = y
        if (ieq > 0) then
          ietvd(j, ii, ie) = ieq
          c(1) = cv(nen3v(1, ieq))
          c(2) = cv(nen3v(2, ieq))
          c(3) = cv(nen3v(3, ieq))
          call femintp_xi(ieq, c, x, y, val)
          cu(j, ii, ie) = val
        end if
      end do
    end do
  end do

  open (unit=22, file=trim(fout), access='stream', form='unformatted', status='replace', action='write')
  write (22) ev(:, 1:ne)
  write (22) subset_num
  write (22) color
  write (22) ietvd
  write (22) xtvd
  write (22) ytvd
  write (22) cu
  close (22)
end program geom_ref
