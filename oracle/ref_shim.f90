! ref_shim.f90 -- bind(C) entry points that load the reference's module state
! (module `declaration`, declaration.f90) and call the reference's OWN
! tridiagonal routines (tridLU_3D.f90:55-537), compiled unmodified from the
! reference sources by oracle/build_ref.sh.
!
! TEST INFRASTRUCTURE ONLY: used by tests/ to pin the oracle's solves
! against the reference code.  One channel, one rank (myid = 0), all columns
! in the middle band, the layout proc_lims_columns builds for equal band
! resolutions (start.f90:1574-1600, 1680-1706).
module lh_ref_shim
  use iso_c_binding
  use declaration
  implicit none
  integer :: s_ncol = 0, s_ny = 0
contains

  subroutine ref_setup(n1, n2, ny, alp_in, bet_in, re_in, dt_in, yu_in, yv_in, &
                       dthdyu_in, dthdyv_in, dyu2i_in, dyv2i_in, gw_in) bind(C, name="ref_setup")
    integer(c_int), value :: n1, n2, ny
    real(c_double), value :: alp_in, bet_in, re_in, dt_in
    real(c_double), intent(in) :: yu_in(0:ny), yv_in(0:ny-1), dthdyu_in(0:ny), dthdyv_in(0:ny-1)
    real(c_double), intent(in) :: dyu2i_in(3,0:ny), dyv2i_in(3,0:ny-1), gw_in(2)
    integer :: ncol, c, i, k

    ncol = (n1/2+1)*n2
    s_ncol = ncol
    s_ny = ny
    alp = alp_in; bet = bet_in; Re = re_in; dt = dt_in
    ddthetavi = 1d0
    if (associated(jlim)) deallocate(jlim)
    allocate(jlim(2,3,nband))
    jlim = 0
    jlim(1,ugrid,midband) = 0;  jlim(2,ugrid,midband) = ny
    jlim(1,vgrid,midband) = 0;  jlim(2,vgrid,midband) = ny-1
    jlim(1,pgrid,midband) = 1;  jlim(2,pgrid,midband) = ny-1
    if (associated(columns_num)) deallocate(columns_num, columns_i, columns_k)
    allocate(columns_num(nband,0:0), columns_i(ncol,nband,0:0), columns_k(ncol,nband,0:0))
    columns_num = 0
    columns_num(midband,0) = ncol
    c = 0
    do k = 1, n2
      do i = 0, n1/2
        c = c+1
        columns_i(c,midband,0) = i
        columns_k(c,midband,0) = k
      end do
    end do
    if (associated(k2F_x)) deallocate(k2F_x, k2F_z)
    allocate(k2F_x(0:n1/2), k2F_z(1:n2))
    do i = 0, n1/2
      k2F_x(i) = -(alp*i)**2
    end do
    do k = 1, n2/2
      k2F_z(k) = -(bet*(k-1))**2
    end do
    do k = n2/2+1, n2
      k2F_z(k) = -(bet*(n2-k+1))**2
    end do
    if (associated(yu)) deallocate(yu, yv, dthdyu, dthdyv, dyu2i, dyv2i)
    allocate(yu(0:ny), yv(0:ny-1), dthdyu(0:ny), dthdyv(0:ny-1), dyu2i(3,0:ny), dyv2i(3,0:ny-1))
    yu = yu_in; yv = yv_in; dthdyu = dthdyu_in; dthdyv = dthdyv_in
    dyu2i = dyu2i_in; dyv2i = dyv2i_in
    if (associated(gridweighting)) deallocate(gridweighting)
    allocate(gridweighting(nband,2))
    gridweighting = 0d0
    gridweighting(midband,1) = gw_in(1)
    gridweighting(midband,2) = gw_in(2)
    aRK = (/ 4d0/15d0, 1d0/15d0, 1d0/6d0 /)
    bRK = aRK
    gRK = (/ 8d0/15d0, 5d0/12d0, 3d0/4d0 /)
    cRK = (/ 0d0, -17d0/60d0, -5d0/12d0 /)
    dRK = (/ 8d0/15d0, 2d0/15d0, 1d0/3d0 /)
    if (allocated(DG)) deallocate(DG)
    allocate(DG(sband:eband))
    do i = sband, eband
      allocate(DG(i)%f_dg(3, 1:ny-1, max(columns_num(i,0),1)))
      DG(i)%f_dg = 0d0
    end do
    call LU_buildP(1, ny-1, 0, midband, DG)
  end subroutine

  subroutine ref_lusol_uw(krk_in, u, rhsu, w, rhsw) bind(C, name="ref_lusol_uw")
    integer(c_int), value :: krk_in
    complex(c_double_complex), target, intent(inout) :: u(0:s_ny, s_ncol), w(0:s_ny, s_ncol)
    complex(c_double_complex), target, intent(inout) :: rhsu(0:s_ny, s_ncol), rhsw(0:s_ny, s_ncol)
    type(cfield) :: ut(sband:eband), wt(sband:eband), rut(sband:eband), rwt(sband:eband)
    real(8), allocatable :: a(:,:)
    kRK = krk_in
    ut(midband)%f => u
    wt(midband)%f => w
    rut(midband)%f => rhsu
    rwt(midband)%f => rhsw
    allocate(a(3,0:s_ny))
    call LUsolU_W(ut, rut, wt, rwt, a, ugrid, midband, 0)
    deallocate(a)
  end subroutine

  subroutine ref_lusol_v(krk_in, v, rhsv) bind(C, name="ref_lusol_v")
    integer(c_int), value :: krk_in
    complex(c_double_complex), target, intent(inout) :: v(0:s_ny-1, s_ncol), rhsv(0:s_ny-1, s_ncol)
    type(cfield) :: vt(sband:eband), rvt(sband:eband)
    real(8), allocatable :: a(:,:)
    kRK = krk_in
    vt(midband)%f => v
    rvt(midband)%f => rhsv
    allocate(a(3,0:s_ny-1))
    call LUsolV(vt, rvt, a, vgrid, midband, 0)
    deallocate(a)
  end subroutine

  subroutine ref_lusol_p(x) bind(C, name="ref_lusol_p")
    complex(c_double_complex), intent(inout) :: x(1:s_ny-1, s_ncol)
    call LUsolP(x, 0, midband, 1, s_ny-1)
  end subroutine

  subroutine ref_lusol0(krk_in, x) bind(C, name="ref_lusol0")
    integer(c_int), value :: krk_in
    real(c_double), intent(inout) :: x(0:s_ny)
    kRK = krk_in
    call LUsol0(x, 0, s_ny)
  end subroutine

  ! A banded channel (the reference's three blocks): column lists of the
  ! bands (ci, ck, 1-based ck, concatenated band by band), their row limits
  ! (lims(1:2, grid, band), grid = vgrid 1, ugrid 2, pgrid 3 as in
  ! declaration.f90:15-17) and the edge weights gw(side, band) of
  ! start.f90:539-544; the Poisson factors of every band by the reference's
  ! LU_buildP (called as start.f90:656-658 does, per band).
  subroutine ref_setup_bands(n1, n2, ny, alp_in, bet_in, re_in, dt_in, yu_in, yv_in, &
                             dthdyu_in, dthdyv_in, dyu2i_in, dyv2i_in, gw_in, lims, ncb, ci, ck) &
                             bind(C, name="ref_setup_bands")
    integer(c_int), value :: n1, n2, ny
    real(c_double), value :: alp_in, bet_in, re_in, dt_in
    real(c_double), intent(in) :: yu_in(0:ny), yv_in(0:ny-1), dthdyu_in(0:ny), dthdyv_in(0:ny-1)
    real(c_double), intent(in) :: dyu2i_in(3,0:ny), dyv2i_in(3,0:ny-1), gw_in(2,3)
    integer(c_int), intent(in) :: lims(2,3,3), ncb(3), ci(*), ck(*)
    integer :: b, c, q, i, k, mx

    alp = alp_in; bet = bet_in; Re = re_in; dt = dt_in
    ddthetavi = 1d0
    s_ny = ny
    if (associated(jlim)) deallocate(jlim)
    allocate(jlim(2,3,nband))
    jlim = lims
    mx = max(maxval(ncb), 1)
    if (associated(columns_num)) deallocate(columns_num, columns_i, columns_k)
    allocate(columns_num(nband,0:0), columns_i(mx,nband,0:0), columns_k(mx,nband,0:0))
    columns_i = 0
    columns_k = 1
    q = 0
    do b = 1, nband
      columns_num(b,0) = ncb(b)
      do c = 1, ncb(b)
        q = q+1
        columns_i(c,b,0) = ci(q)
        columns_k(c,b,0) = ck(q)
      end do
    end do
    if (associated(k2F_x)) deallocate(k2F_x, k2F_z)
    allocate(k2F_x(0:n1/2), k2F_z(1:n2))
    do i = 0, n1/2
      k2F_x(i) = -(alp*i)**2
    end do
    do k = 1, n2/2
      k2F_z(k) = -(bet*(k-1))**2
    end do
    do k = n2/2+1, n2
      k2F_z(k) = -(bet*(n2-k+1))**2
    end do
    if (associated(yu)) deallocate(yu, yv, dthdyu, dthdyv, dyu2i, dyv2i)
    allocate(yu(0:ny), yv(0:ny-1), dthdyu(0:ny), dthdyv(0:ny-1), dyu2i(3,0:ny), dyv2i(3,0:ny-1))
    yu = yu_in; yv = yv_in; dthdyu = dthdyu_in; dthdyv = dthdyv_in
    dyu2i = dyu2i_in; dyv2i = dyv2i_in
    if (associated(gridweighting)) deallocate(gridweighting)
    allocate(gridweighting(nband,2))
    do b = 1, nband
      gridweighting(b,1) = gw_in(1,b)
      gridweighting(b,2) = gw_in(2,b)
    end do
    aRK = (/ 4d0/15d0, 1d0/15d0, 1d0/6d0 /)
    bRK = aRK
    gRK = (/ 8d0/15d0, 5d0/12d0, 3d0/4d0 /)
    cRK = (/ 0d0, -17d0/60d0, -5d0/12d0 /)
    dRK = (/ 8d0/15d0, 2d0/15d0, 1d0/3d0 /)
    if (allocated(DG)) deallocate(DG)
    allocate(DG(sband:eband))
    do b = sband, eband
      allocate(DG(b)%f_dg(3, jlim(1,pgrid,b):jlim(2,pgrid,b), max(columns_num(b,0),1)))
      DG(b)%f_dg = 0d0
      call LU_buildP(jlim(1,pgrid,b), jlim(2,pgrid,b), 0, b, DG)
    end do
  end subroutine

  subroutine ref_lusol_uw_b(ib, krk_in, u, rhsu, w, rhsw) bind(C, name="ref_lusol_uw_b")
    integer(c_int), value :: ib, krk_in
    complex(c_double_complex), target, intent(inout) :: u(jlim(1,ugrid,ib):jlim(2,ugrid,ib), columns_num(ib,0))
    complex(c_double_complex), target, intent(inout) :: w(jlim(1,ugrid,ib):jlim(2,ugrid,ib), columns_num(ib,0))
    complex(c_double_complex), target, intent(inout) :: rhsu(jlim(1,ugrid,ib):jlim(2,ugrid,ib), columns_num(ib,0))
    complex(c_double_complex), target, intent(inout) :: rhsw(jlim(1,ugrid,ib):jlim(2,ugrid,ib), columns_num(ib,0))
    type(cfield) :: ut(sband:eband), wt(sband:eband), rut(sband:eband), rwt(sband:eband)
    real(8), allocatable :: a(:,:)
    kRK = krk_in
    ut(ib)%f => u
    wt(ib)%f => w
    rut(ib)%f => rhsu
    rwt(ib)%f => rhsw
    allocate(a(3,jlim(1,ugrid,ib):jlim(2,ugrid,ib)))
    call LUsolU_W(ut, rut, wt, rwt, a, ugrid, ib, 0)
    deallocate(a)
  end subroutine

  subroutine ref_lusol_v_b(ib, krk_in, v, rhsv) bind(C, name="ref_lusol_v_b")
    integer(c_int), value :: ib, krk_in
    complex(c_double_complex), target, intent(inout) :: v(jlim(1,vgrid,ib):jlim(2,vgrid,ib), columns_num(ib,0))
    complex(c_double_complex), target, intent(inout) :: rhsv(jlim(1,vgrid,ib):jlim(2,vgrid,ib), columns_num(ib,0))
    type(cfield) :: vt(sband:eband), rvt(sband:eband)
    real(8), allocatable :: a(:,:)
    kRK = krk_in
    vt(ib)%f => v
    rvt(ib)%f => rhsv
    allocate(a(3,jlim(1,vgrid,ib):jlim(2,vgrid,ib)))
    call LUsolV(vt, rvt, a, vgrid, ib, 0)
    deallocate(a)
  end subroutine

  subroutine ref_lusol_p_b(ib, x) bind(C, name="ref_lusol_p_b")
    integer(c_int), value :: ib
    complex(c_double_complex), intent(inout) :: x(jlim(1,pgrid,ib):jlim(2,pgrid,ib), columns_num(ib,0))
    call LUsolP(x, 0, ib, jlim(1,pgrid,ib), jlim(2,pgrid,ib))
  end subroutine

end module
