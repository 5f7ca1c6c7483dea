! lh_dropin.f90 -- external subroutines with the reference's names AND
! argument lists, so that the reference's call sites compile and link
! against them unchanged, forwarding to liblhgpu:
!
!   LUsolU_W(u,rhsu,w,rhsw,a,grid,iband,myid)      tridLU_3D.f90:55
!   LUsolV(u,rhsu,a,grid,iband,myid)               tridLU_3D.f90:135
!   LUsolP(x,myid,iband,nystart,nyend)             tridLU_3D.f90:192
!   LUsol0(x,nystart,nyend)                        tridLU_3D.f90:228
!   LU_buildP(nystart,nyend,myid,iband,a)          tridLU_3D.f90:333
!   modes_to_planes_UVP(xPL,x,grid,myid,status,ierr)   FOU3D.f90:1000
!   planes_to_modes_UVP(x,xPL,grid,myid,status,ierr)   FOU3D.f90:2295
!   planes_to_modes_NUVP(x,xPL,grid,myid,status,ierr)  FOU3D.f90:2772
!   four_to_phys_u(u1,u2,u3,iband), four_to_phys_du(du,iband),
!   phys_to_four_du(duPL,iband)                    FOU3D.f90:738-784, 2253
!   four_to_phys_N(du,iband), phys_to_four_N(duPL,iband)  FOU3D.f90:786, 2274
!   cft(c,ise,isa,m,iopt,aux), cfti(ngal,wsac1)    cft_buff.f:1, 71
!   rft(c,isa,m,iopt,aux), rfti(mgal,wsaf1)        rft_buff.f:24, 1
!   boundary                                       boundary_post_coll.f90:1
!
! They read the host's module `declaration` (the reference's own, or
! declaration_min.f90 here) for the band layout, the grid sizes and kRK, and
! the device context set up by lh_host (lh_host_setup).  Arrays are the
! reference's: band arrays of type(cfield) with f(jlim(1,grid,b):jlim(2,
! grid,b), columns_num(b,myid)), planes real(8)(igal, kgal, j).
!
! Layout accepted: one rank (myid as the host passes it, normally 0) and one
! x/z resolution at every height -- the long pencils of the middle band hold
! every (kx, kz) column in the order proc_lims_columns builds
! (start.f90:1600-1624) and the short-pencil bands are empty.  Any other
! layout stops with a message.  Arguments the device does not need are
! accepted and left alone: the LU workspace `a` (the operators are built and
! factored on the device), `status` (no MPI); `ierr` is set to 0.  The
! reference's routines stop on errors (littleharsh.f90:85); these do too.
module lh_dropin_support
  use iso_c_binding
  use lhgpu
  use declaration
  use lh_host, only: lhctx, lh_check, staging, dropin_chan, hN => N, nyb
  implicit none
contains

  ! rows and columns of band `iband` on grid `grid`; active = the band holds
  ! columns.  Checks the layout against the device channel.
  subroutine band_layout(iband, myid, grid, jlo, jhi, ncol, active)
    integer, intent(in) :: iband, myid, grid
    integer, intent(out) :: jlo, jhi, ncol
    logical, intent(out) :: active
    integer :: c, nxh, ny, ch
    ncol = columns_num(iband, myid)
    active = ncol > 0
    jlo = jlim(1, grid, iband)
    jhi = jlim(2, grid, iband)
    if (.not. active) return
    ch = dropin_chan + 1
    ny = nyb(ch)
    if (N(1, iband) /= hN(1, ch) .or. N(2, iband) /= hN(2, ch)) &
      error stop 'lh_dropin: band resolution differs from the device channel (multi-resolution bands)'
    nxh = N(1, iband) / 2 + 1
    if (ncol /= nxh * N(2, iband)) &
      error stop 'lh_dropin: a band must hold every (kx,kz) column (one rank, uniform resolution)'
    do c = 1, ncol
      if (columns_i(c, iband, myid) /= mod(c - 1, nxh) .or. columns_k(c, iband, myid) /= (c - 1) / nxh + 1) &
        error stop 'lh_dropin: columns not in proc_lims_columns order'
    end do
    select case (grid)
    case (ugrid)
      if (jlo /= 0 .or. jhi /= ny) error stop 'lh_dropin: u rows must be 0..ny'
    case (vgrid)
      if (jlo /= 0 .or. jhi /= ny - 1) error stop 'lh_dropin: v rows must be 0..ny-1'
    case default
      if (jlo /= 1 .or. jhi /= ny - 1) error stop 'lh_dropin: p rows must be 1..ny-1'
    end select
  end subroutine band_layout

  ! copy `bytes` from the host to offset `off` of the staging buffer, sized
  ! for `total` bytes (every part of one call's staging asks for the same
  ! total, so an earlier part is never moved)
  function put(h, bytes, off, total) result(d)
    type(c_ptr), intent(in) :: h
    integer(c_size_t), intent(in) :: bytes, off, total
    type(c_ptr) :: d
    d = staging(total, off)
    call lh_check(lh_copy_to_device(d, h, bytes), 'lh_copy_to_device')
  end function put

  subroutine get(h, d, bytes)
    type(c_ptr), intent(in) :: h, d
    integer(c_size_t), intent(in) :: bytes
    call lh_check(lh_copy_to_host(h, d, bytes), 'lh_copy_to_host')
  end subroutine get

  ! one plane in place through the device plane transforms; nfour = the
  ! Fourier modes N1 of the plane's band, n1, n2 the plane's grid
  subroutine plane_in_place(pl, nwords, nfour, n1, n2, dir, ngrid)
    real(8), intent(inout), target :: pl(*)
    integer, intent(in) :: nwords, nfour, n1, n2, dir
    logical, intent(in) :: ngrid
    integer(c_size_t) :: b
    type(c_ptr) :: d
    b = int(nwords, c_size_t) * 8_c_size_t
    d = put(c_loc(pl), b, 0_c_size_t, b)
    if (ngrid) then
      if (dir > 0) then
        call lh_check(lh_four_to_phys_N_planes(int(n1, c_int), int(n2, c_int), d, 1_c_int, c_null_ptr), 'four_to_phys_N')
      else
        call lh_check(lh_phys_to_four_N_planes(int(n1, c_int), int(n2, c_int), d, 1_c_int, c_null_ptr), 'phys_to_four_N')
      end if
    else if (dir > 0) then
      call lh_check(lh_four_to_phys_planes(int(nfour, c_int), int(n1, c_int), int(n2, c_int), d, 1_c_int, &
                    c_null_ptr), 'four_to_phys')
    else
      call lh_check(lh_phys_to_four_planes(int(nfour, c_int), int(n1, c_int), int(n2, c_int), d, 1_c_int, &
                    c_null_ptr), 'phys_to_four')
    end if
    call get(c_loc(pl), d, b)
  end subroutine plane_in_place

  ! rows jlo..jhi of band columns x (leading dimension ld, first stored row
  ! jfirst) <-> planes xPL(igal, kgal, jlo..jhi); dir > 0: to the planes
  subroutine transpose_rows(xf, ld, jfirst, ncol, pl, jlo, jhi, iband, dir)
    complex(8), intent(inout), target :: xf(ld, ncol)
    real(8), intent(inout), target :: pl(igal, kgal, jlo:jhi)
    integer, intent(in) :: ld, jfirst, ncol, jlo, jhi, iband, dir
    complex(8), allocatable, target :: tmp(:,:)
    integer :: nrows
    integer(c_size_t) :: bc, bp
    type(c_ptr) :: dc, dp
    nrows = jhi - jlo + 1
    bc = int(nrows, c_size_t) * int(ncol, c_size_t) * 16_c_size_t
    bp = int(nrows, c_size_t) * int(igal, c_size_t) * int(kgal, c_size_t) * 8_c_size_t
    if (igal /= Ngal(1, iband) + 2 .or. kgal /= Ngal(2, iband)) error stop 'lh_dropin: igal, kgal do not match Ngal'
    allocate(tmp(jlo:jhi, ncol))
    dc = staging(bc + bp, 0_c_size_t)
    dp = staging(bc + bp, bc)
    if (dir > 0) then
      tmp = xf(jlo - jfirst + 1:jhi - jfirst + 1, :)
      call lh_check(lh_copy_to_device(dc, c_loc(tmp), bc), 'lh_copy_to_device')
      call lh_check(lh_modes_to_planes(int(N(1, iband), c_int), int(N(2, iband), c_int), int(Ngal(1, iband), c_int), &
                    int(Ngal(2, iband), c_int), dc, int(nrows, c_int), dp, c_null_ptr), 'lh_modes_to_planes')
      call get(c_loc(pl), dp, bp)
    else
      call lh_check(lh_copy_to_device(dp, c_loc(pl), bp), 'lh_copy_to_device')
      call lh_check(lh_planes_to_modes(int(N(1, iband), c_int), int(N(2, iband), c_int), int(Ngal(1, iband), c_int), &
                    int(Ngal(2, iband), c_int), dp, int(nrows, c_int), dc, c_null_ptr), 'lh_planes_to_modes')
      call get(c_loc(tmp), dc, bc)
      xf(jlo - jfirst + 1:jhi - jfirst + 1, :) = tmp
    end if
    deallocate(tmp)
  end subroutine transpose_rows

end module lh_dropin_support

! ---------------------------------------------------------------------------
! tridLU_3D.f90:55 -- (I - beta (D2 + k^2)) u = rhsu and w, band iband
subroutine LUsolU_W(u, rhsu, w, rhsw, a, grid, iband, myid)
  use lh_dropin_support
  implicit none
  integer :: grid, iband, myid
  type(cfield) :: u(sband:eband), w(sband:eband), rhsu(sband:eband), rhsw(sband:eband)
  real(8) :: a(3, *)
  integer :: jlo, jhi, ncol
  logical :: active
  integer(c_size_t) :: b
  type(c_ptr) :: d1, d2, d3, d4
  call band_layout(iband, myid, grid, jlo, jhi, ncol, active)
  if (.not. active) return
  if (grid /= ugrid) error stop 'LUsolU_W: grid must be ugrid'
  b = int(jhi - jlo + 1, c_size_t) * int(ncol, c_size_t) * 16_c_size_t
  d1 = put(c_loc(rhsu(iband)%f), b, 0_c_size_t, 4 * b)
  d2 = put(c_loc(rhsw(iband)%f), b, b, 4 * b)
  d3 = staging(4 * b, 2 * b)
  d4 = staging(4 * b, 3 * b)
  call lh_check(lh_lusol_uw(lhctx, int(dropin_chan, c_int), int(kRK, c_int), d1, d2, d3, d4), 'LUsolU_W')
  call get(c_loc(u(iband)%f), d3, b)
  call get(c_loc(w(iband)%f), d4, b)
end subroutine LUsolU_W

! tridLU_3D.f90:135 -- the v-grid operator, walls v = 0
subroutine LUsolV(u, rhsu, a, grid, iband, myid)
  use lh_dropin_support
  implicit none
  integer :: grid, iband, myid
  type(cfield) :: u(sband:eband), rhsu(sband:eband)
  real(8) :: a(3, *)
  integer :: jlo, jhi, ncol
  logical :: active
  integer(c_size_t) :: b
  type(c_ptr) :: d1, d2
  call band_layout(iband, myid, grid, jlo, jhi, ncol, active)
  if (.not. active) return
  if (grid /= vgrid) error stop 'LUsolV: grid must be vgrid'
  b = int(jhi - jlo + 1, c_size_t) * int(ncol, c_size_t) * 16_c_size_t
  d1 = put(c_loc(rhsu(iband)%f), b, 0_c_size_t, 2 * b)
  d2 = staging(2 * b, b)
  call lh_check(lh_lusol_v(lhctx, int(dropin_chan, c_int), int(kRK, c_int), d1, d2), 'LUsolV')
  call get(c_loc(u(iband)%f), d2, b)
end subroutine LUsolV

! tridLU_3D.f90:192 -- the Poisson operator of LU_buildP, in place
subroutine LUsolP(x, myid, iband, nystart, nyend)
  use lh_dropin_support
  implicit none
  integer :: myid, iband, nystart, nyend
  complex(8), target :: x(nystart:nyend, *)
  integer :: jlo, jhi, ncol
  logical :: active
  integer(c_size_t) :: b
  type(c_ptr) :: d
  call band_layout(iband, myid, pgrid, jlo, jhi, ncol, active)
  if (.not. active) return
  if (nystart /= jlo .or. nyend /= jhi) error stop 'LUsolP: rows must be the band''s p rows'
  b = int(jhi - jlo + 1, c_size_t) * int(ncol, c_size_t) * 16_c_size_t
  d = put(c_loc(x(nystart, 1)), b, 0_c_size_t, b)
  call lh_check(lh_lusol_p(lhctx, int(dropin_chan, c_int), d), 'LUsolP')
  call get(c_loc(x(nystart, 1)), d, b)
end subroutine LUsolP

! tridLU_3D.f90:228 -- the real (0,0) U system of substep kRK, in place
subroutine LUsol0(x, nystart, nyend)
  use lh_dropin_support
  implicit none
  integer :: nystart, nyend
  real(8), target :: x(nystart:nyend)
  integer(c_size_t) :: b
  type(c_ptr) :: d
  if (nystart /= 0 .or. nyend /= nyb(dropin_chan + 1)) error stop 'LUsol0: rows must be 0..ny'
  b = int(nyend - nystart + 1, c_size_t) * 8_c_size_t
  d = put(c_loc(x), b, 0_c_size_t, b)
  call lh_check(lh_lusol0(lhctx, int(dropin_chan, c_int), int(kRK, c_int), d), 'LUsol0')
  call get(c_loc(x), d, b)
end subroutine LUsol0

! tridLU_3D.f90:333 -- the device builds and factors the Poisson operator
! when the context is created, and LUsolP reads those factors: `a` is not
! needed and is left alone; the rows are checked.
subroutine LU_buildP(nystart, nyend, myid, iband, a)
  use lh_dropin_support
  implicit none
  integer :: nystart, nyend, myid, iband
  type(rfield_dg) :: a(sband:eband)
  integer :: jlo, jhi, ncol
  logical :: active
  call band_layout(iband, myid, pgrid, jlo, jhi, ncol, active)
  if (active .and. (nystart /= jlo .or. nyend /= jhi)) error stop 'LU_buildP: rows must be the band''s p rows'
end subroutine LU_buildP

! ---------------------------------------------------------------------------
! FOU3D.f90:1000 -- band columns to planes (one rank: every plane is local)
subroutine modes_to_planes_UVP(xPL, x, grid, myid, status, ierr)
  use lh_dropin_support
  implicit none
  integer :: grid, myid, ierr
  integer :: status(*)
  type(cfield) :: x(sband:eband)
  real(8), target :: xPL(igal, kgal, jgal(grid, 1) - 1:jgal(grid, 2) + 1)
  integer :: ib, jlo, jhi, ncol
  logical :: active
  do ib = sband, eband
    call band_layout(ib, myid, grid, jlo, jhi, ncol, active)
    if (.not. active) cycle
    if (jlo < jgal(grid, 1) - 1 .or. jhi > jgal(grid, 2) + 1) error stop 'modes_to_planes_UVP: planes miss rows'
    call transpose_rows(x(ib)%f, jhi - jlo + 1, jlo, ncol, xPL(:, :, jlo:jhi), jlo, jhi, ib, 1)
  end do
  ierr = 0
end subroutine modes_to_planes_UVP

! FOU3D.f90:2295 -- planes to band columns, every row of the band
subroutine planes_to_modes_UVP(x, xPL, grid, myid, status, ierr)
  use lh_dropin_support
  implicit none
  integer :: grid, myid, ierr
  integer :: status(*)
  type(cfield) :: x(sband:eband)
  real(8), target :: xPL(igal, kgal, jgal(grid, 1) - 1:jgal(grid, 2) + 1)
  integer :: ib, jlo, jhi, ncol
  logical :: active
  do ib = sband, eband
    call band_layout(ib, myid, grid, jlo, jhi, ncol, active)
    if (.not. active) cycle
    if (jlo < jgal(grid, 1) - 1 .or. jhi > jgal(grid, 2) + 1) error stop 'planes_to_modes_UVP: planes miss rows'
    call transpose_rows(x(ib)%f, jhi - jlo + 1, jlo, ncol, xPL(:, :, jlo:jhi), jlo, jhi, ib, -1)
  end do
  ierr = 0
end subroutine planes_to_modes_UVP

! FOU3D.f90:2772 -- the nonlinear planes (interior rows only) to columns
subroutine planes_to_modes_NUVP(x, xPL, grid, myid, status, ierr)
  use lh_dropin_support
  implicit none
  integer :: grid, myid, ierr
  integer :: status(*)
  type(cfield) :: x(sband:eband)
  real(8), target :: xPL(igal, kgal, jgal(grid, 1):jgal(grid, 2))
  integer :: ib, jlo, jhi, ncol, lb
  logical :: active
  do ib = sband, eband
    call band_layout(ib, myid, grid, jlo, jhi, ncol, active)
    if (.not. active) cycle
    lb = lbound(x(ib)%f, 1)
    if (jlo + 1 < jgal(grid, 1) .or. jhi - 1 > jgal(grid, 2)) error stop 'planes_to_modes_NUVP: planes miss rows'
    call transpose_rows(x(ib)%f, size(x(ib)%f, 1), lb, ncol, xPL(:, :, jlo + 1:jhi - 1), jlo + 1, jhi - 1, ib, -1)
  end do
  ierr = 0
end subroutine planes_to_modes_NUVP

! ---------------------------------------------------------------------------
! FOU3D.f90:738-784, 2253-2272, 786-804, 2274-2293 -- one plane in place
subroutine four_to_phys_u(u1, u2, u3, iband)
  use lh_dropin_support
  implicit none
  integer :: iband
  real(8) :: u1(*), u2(*), u3(*)
  integer :: nw
  nw = (Ngal(1, iband) + 2) * Ngal(2, iband)
  call plane_in_place(u1, nw, N(1, iband), Ngal(1, iband), Ngal(2, iband), 1, .false.)
  call plane_in_place(u2, nw, N(1, iband), Ngal(1, iband), Ngal(2, iband), 1, .false.)
  call plane_in_place(u3, nw, N(1, iband), Ngal(1, iband), Ngal(2, iband), 1, .false.)
end subroutine four_to_phys_u

subroutine four_to_phys_du(du, iband)
  use lh_dropin_support
  implicit none
  integer :: iband
  real(8) :: du(*)
  call plane_in_place(du, (Ngal(1, iband) + 2) * Ngal(2, iband), N(1, iband), Ngal(1, iband), Ngal(2, iband), 1, .false.)
end subroutine four_to_phys_du

subroutine phys_to_four_du(duPL, iband)
  use lh_dropin_support
  implicit none
  integer :: iband
  real(8) :: duPL(*)
  call plane_in_place(duPL, (Ngal(1, iband) + 2) * Ngal(2, iband), N(1, iband), Ngal(1, iband), Ngal(2, iband), -1, &
                      .false.)
end subroutine phys_to_four_du

subroutine four_to_phys_N(du, iband)
  use lh_dropin_support
  implicit none
  integer :: iband
  real(8) :: du(*)
  call plane_in_place(du, (N(1, iband) + 2) * N(2, iband), N(1, iband), N(1, iband), N(2, iband), 1, .true.)
end subroutine four_to_phys_N

subroutine phys_to_four_N(duPL, iband)
  use lh_dropin_support
  implicit none
  integer :: iband
  real(8) :: duPL(*)
  call plane_in_place(duPL, (N(1, iband) + 2) * N(2, iband), N(1, iband), N(1, iband), N(2, iband), -1, .true.)
end subroutine phys_to_four_N

! ---------------------------------------------------------------------------
! cft_buff.f:71 / rft_buff.f:1 -- the length and 1/length in aux(1:2), as
! the reference's readers of aux expect; the device keeps its own twiddles,
! so the rest of the work area is only cleared
subroutine cfti(ngal, wsac1)
  implicit none
  integer :: ngal
  real(8) :: wsac1(10 * ngal + 19)
  wsac1 = 0d0
  wsac1(1) = ngal
  wsac1(2) = 1d0 / ngal
end subroutine cfti

subroutine rfti(mgal, wsaf1)
  implicit none
  integer :: mgal
  real(8) :: wsaf1(2 * mgal + 18)
  wsaf1 = 0d0
  wsaf1(1) = mgal
  wsaf1(2) = 1d0 / mgal
end subroutine rfti

! cft_buff.f:1 -- m complex vectors, element stride ise, vector stride isa
subroutine cft(c, ise, isa, m, iopt, aux)
  use lh_dropin_support
  implicit none
  integer :: ise, isa, m, iopt
  real(8), target :: c(*)
  real(8) :: aux(*)
  integer :: nlen
  integer(c_size_t) :: span, b
  type(c_ptr) :: d
  nlen = nint(aux(1))
  span = int(m - 1, c_size_t) * int(isa, c_size_t) + int(nlen - 1, c_size_t) * int(ise, c_size_t) + 2_c_size_t
  b = span * 8_c_size_t
  d = put(c_loc(c), b, 0_c_size_t, b)
  call lh_check(lh_cft(d, int(ise, c_int), int(isa, c_int), int(m, c_int), int(iopt, c_int), int(nlen, c_int), &
                c_null_ptr), 'cft')
  call get(c_loc(c), d, b)
end subroutine cft

! rft_buff.f:24 -- m real vectors c(1:n+2, j), isa apart
subroutine rft(c, isa, m, iopt, aux)
  use lh_dropin_support
  implicit none
  integer :: isa, m, iopt
  real(8), target :: c(*)
  real(8) :: aux(*)
  integer :: nlen
  integer(c_size_t) :: span, b
  type(c_ptr) :: d
  nlen = nint(aux(1))
  span = int(m - 1, c_size_t) * int(isa, c_size_t) + int(nlen + 2, c_size_t)
  b = span * 8_c_size_t
  d = put(c_loc(c), b, 0_c_size_t, b)
  call lh_check(lh_rft(d, int(isa, c_int), int(m, c_int), int(iopt, c_int), int(nlen, c_int), c_null_ptr), 'rft')
  call get(c_loc(c), d, b)
end subroutine rft

! ---------------------------------------------------------------------------
! boundary_post_coll.f90:1 -- the wall geometry to the device.  This build
! models textured walls by the planeBC mask the host's proc_lims_columns
! builds (start.f90:1776-1829; 1 no-slip, 2 free-shear), the same on both
! walls, applied in the Helmholtz solves and the ghost rows; geometry_type
! 0 is a smooth channel.  The immersed-boundary roughness of
! geometry_type 2 (inactive in the reference's step, FOU3D.f90:412-423) is
! not part of this build and stops.
subroutine boundary
  use lh_dropin_support
  implicit none
  integer, allocatable, target :: mask(:,:)
  integer :: n1, n2, i, k, ch
  ch = dropin_chan + 1
  if (geometry_type == 0) then
    call lh_check(lh_set_texture(lhctx, int(dropin_chan, c_int), c_null_ptr, c_null_ptr), 'boundary')
    return
  end if
  if (geometry_type /= 1) error stop 'boundary: immersed-boundary roughness (geometry_type 2) is not built'
  if (.not. associated(planeBC)) error stop 'boundary: planeBC is not built (proc_lims_columns)'
  n1 = hN(1, ch)
  n2 = hN(2, ch)
  allocate(mask(0:n1 - 1, 0:n2 - 1))         ! index z*N1 + x
  do k = 0, n2 - 1
    do i = 0, n1 - 1
      mask(i, k) = planeBC(i + 1, k + 1)
    end do
  end do
  call lh_check(lh_set_texture(lhctx, int(dropin_chan, c_int), c_loc(mask), c_loc(mask)), 'boundary')
  deallocate(mask)
end subroutine boundary
