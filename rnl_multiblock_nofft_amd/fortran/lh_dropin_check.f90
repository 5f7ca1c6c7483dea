! lh_dropin_check -- calls the routines of lh_dropin.f90 the way the
! reference's call sites do (implicit interfaces, band arrays of
! type(cfield), the reference's argument lists), so that a test can compare
! what they return with an independent implementation:
!   solveU_W / solveV's loops over the bands (littleharsh.f90:822-826,
!   903-906): LUsolU_W(u,rhsu,w,rhsw,a,ugrid,iband,myid), LUsolV(...);
!   solveP (littleharsh.f90:623): LUsolP(psi(iband)%f,myid,iband,jlim..);
!   flowrate_corr (littleharsh.f90:495): LUsol0(u11,0,ny);
!   start (start.f90:657): LU_buildP(jlim(1,pgrid,b),jlim(2,pgrid,b),myid,b,DG);
!   nonlinear (FOU3D.f90:116-121, 252-260): modes_to_planes_UVP,
!   planes_to_modes_UVP, planes_to_modes_NUVP with (...,grid,myid,status,ierr);
!   ops_in_planes2 (FOU3D.f90:2116-2127): four_to_phys_u, phys_to_four_du;
!   FOU3D.f90:757-762, 2267-2268: cft / rft on a plane after cfti / rfti
!   (start.f90:400-403);
!   getbounds (start.f90:1227-1245): boundary, then one RK3 step of the
!   textured channel.
!
!   lh_dropin_check dropin.nml
!
! Files are raw FP64 streams as in lh_shim_check; planebc is N1 x N2 int32
! (planeBC(1:N1, 1:N2)).
program lh_dropin_check
  use iso_c_binding
  use lhgpu
  use declaration
  use lh_host, only: lh_host_setup, lh_host_finish, lhctx, lh_check, dropin_chan
  implicit none
  integer :: N1 = 16, N2 = 16, ny = 33, kstage = 1, grid_kind = 1
  real(8) :: alp = 1d0, bet = 2d0, Reb = 4200d0, gamma = 2d0, dtin = 1d-3
  character(len=1024) :: dir = '.', nmlfile
  namelist /check/ N1, N2, ny, alp, bet, Reb, grid_kind, gamma, dtin, kstage, dir
  type(lh_config) :: cfg(1)
  type(cfield) :: u(sband:eband), w(sband:eband), v(sband:eband), ru(sband:eband), rw(sband:eband)
  type(cfield) :: rv(sband:eband), psi(sband:eband), x(sband:eband), xn(sband:eband)
  type(rfield_dg) :: DG(sband:eband)
  real(8), allocatable :: a(:,:), u11(:), xPL(:,:,:), xPLn(:,:,:), pl(:,:), buffC(:), buffR(:)
  real(8), allocatable :: p1(:,:), p2(:,:), p3(:,:)
  complex(8), allocatable, target :: fld(:,:)
  integer :: status(8), ierr, myid, ib, ncol, nxh, c, q, j, Ng1, Ng2, un
  character(len=2) :: fn(4) = ['u1', 'u2', 'u3', 'p ']
  integer :: fid(4), jl(4), jh(4)

  call get_command_argument(1, nmlfile)
  open(newunit=un, file=trim(nmlfile), status='old', action='read')
  read(un, nml=check)
  close(un)
  Ng1 = 3 * N1 / 2
  Ng2 = 3 * N2 / 2
  nxh = N1 / 2 + 1
  ncol = nxh * N2
  myid = 0

  ! the host's module state for one rank, one resolution at every height
  ! (what start / proc_lims_columns / proc_lims_planes leave behind)
  np = 1
  allocate(N(4, 0:nband + 1), Ngal(2, nband), jlim(2, 3, nband))
  N(1, :) = N1
  N(2, :) = N2
  Ngal(1, :) = Ng1
  Ngal(2, :) = Ng2
  jlim = 0
  jlim(:, ugrid, midband) = [0, ny]
  jlim(:, vgrid, midband) = [0, ny - 1]
  jlim(:, pgrid, midband) = [1, ny - 1]
  allocate(columns_num(nband, 0:0), columns_i(ncol, nband, 0:0), columns_k(ncol, nband, 0:0))
  columns_num = 0
  columns_num(midband, 0) = ncol
  do c = 1, ncol
    columns_i(c, midband, 0) = mod(c - 1, nxh)
    columns_k(c, midband, 0) = (c - 1) / nxh + 1
  end do
  jgal(ugrid, :) = [1, ny - 1]
  jgal(vgrid, :) = [1, ny - 2]
  jgal(pgrid, :) = [2, ny - 2]
  igal = Ng1 + 2
  kgal = Ng2
  kRK = kstage
  dt = dtin
  Re = Reb
  cfg(1) = lh_config(N1, N2, Ng1, Ng2, ny, alp, bet, Re, grid_kind, 10d0, 5, gamma, 0d0, 0d0, 0, 0d0, dt, &
                     LH_PREC_FP64, 0, 0.5d0)
  call lh_host_setup(cfg)
  dropin_chan = 0

  ! band arrays as the reference allocates them (empty short-pencil bands)
  do ib = sband, eband
    q = columns_num(ib, myid)
    allocate(u(ib)%f(jlim(1, ugrid, ib):jlim(2, ugrid, ib), q), w(ib)%f(jlim(1, ugrid, ib):jlim(2, ugrid, ib), q))
    allocate(ru(ib)%f(jlim(1, ugrid, ib):jlim(2, ugrid, ib), q), rw(ib)%f(jlim(1, ugrid, ib):jlim(2, ugrid, ib), q))
    allocate(v(ib)%f(jlim(1, vgrid, ib):jlim(2, vgrid, ib), q), rv(ib)%f(jlim(1, vgrid, ib):jlim(2, vgrid, ib), q))
    allocate(psi(ib)%f(jlim(1, pgrid, ib):jlim(2, pgrid, ib), q))
    allocate(x(ib)%f(jlim(1, ugrid, ib):jlim(2, ugrid, ib), q))
    allocate(xn(ib)%f(jlim(1, ugrid, ib) + 1:jlim(2, ugrid, ib) - 1, q))
    allocate(DG(ib)%f_dg(3, jlim(1, pgrid, ib):jlim(2, pgrid, ib), q))
  end do
  allocate(a(3, 0:ny))

  call rdc('rhsu', ru(midband)%f, size(ru(midband)%f))
  call rdc('rhsw', rw(midband)%f, size(rw(midband)%f))
  call rdc('rhsv', rv(midband)%f, size(rv(midband)%f))
  do ib = sband, eband
    call LUsolU_W(u, ru, w, rw, a, ugrid, ib, myid)
  end do
  do ib = sband, eband
    call LUsolV(v, rv, a, vgrid, ib, myid)
  end do
  call wrc('sol_u', u(midband)%f, size(u(midband)%f))
  call wrc('sol_w', w(midband)%f, size(w(midband)%f))
  call wrc('sol_v', v(midband)%f, size(v(midband)%f))

  do ib = sband, eband
    call LU_buildP(jlim(1, pgrid, ib), jlim(2, pgrid, ib), myid, ib, DG)
  end do
  call rdc('rhsp', psi(midband)%f, size(psi(midband)%f))
  do ib = sband, eband
    if (columns_num(ib, myid) > 0) call LUsolP(psi(ib)%f, myid, ib, jlim(1, pgrid, ib), jlim(2, pgrid, ib))
  end do
  call wrc('sol_p', psi(midband)%f, size(psi(midband)%f))
  allocate(u11(0:ny))
  call rd('rhs0', u11, size(u11))
  call LUsol0(u11, 0, ny)
  call wr('sol_0', u11, size(u11))

  ! transposes on the u grid: columns -> planes, planes -> columns
  allocate(xPL(igal, kgal, jgal(ugrid, 1) - 1:jgal(ugrid, 2) + 1), xPLn(igal, kgal, jgal(ugrid, 1):jgal(ugrid, 2)))
  call rdc('colsu', x(midband)%f, size(x(midband)%f))
  xPL = 0d0
  call modes_to_planes_UVP(xPL, x, ugrid, myid, status, ierr)
  call wr('plu_out', xPL, size(xPL))
  call rd('plu', xPL, size(xPL))
  call planes_to_modes_UVP(x, xPL, ugrid, myid, status, ierr)
  call wrc('colsu_out', x(midband)%f, size(x(midband)%f))
  xPLn = xPL(:, :, jgal(ugrid, 1):jgal(ugrid, 2))
  xn(midband)%f = (0d0, 0d0)
  call planes_to_modes_NUVP(xn, xPLn, ugrid, myid, status, ierr)
  call wrc('colsn_out', xn(midband)%f, size(xn(midband)%f))

  ! plane transforms as ops_in_planes2 calls them, and cft / rft
  allocate(p1(igal, kgal), p2(igal, kgal), p3(igal, kgal), pl(igal, kgal))
  call rd('plane1', p1, size(p1))
  call rd('plane2', p2, size(p2))
  call rd('plane3', p3, size(p3))
  call four_to_phys_u(p1, p2, p3, midband)
  call wr('phys1', p1, size(p1))
  call wr('phys2', p2, size(p2))
  call wr('phys3', p3, size(p3))
  call phys_to_four_du(p1, midband)
  call wr('four1', p1, size(p1))
  allocate(buffC(10 * Ng2 + 19), buffR(2 * Ng1 + 18))
  call cfti(Ng2, buffC)
  call rfti(Ng1, buffR)
  call rd('plane1', pl, size(pl))
  call cft(pl, Ngal(1, midband) + 2, 2, (N(1, midband) + 2) / 2, 1, buffC)
  call wr('cft_inv', pl, size(pl))
  call rft(pl, Ngal(1, midband) + 2, Ngal(2, midband), 1, buffR)
  call wr('rft_inv', pl, size(pl))
  call rft(pl, Ngal(1, midband) + 2, Ngal(2, midband), -1, buffR)
  call wr('rft_fwd', pl, size(pl))
  call cft(pl, Ngal(1, midband) + 2, 2, (N(1, midband) + 2) / 2, -1, buffC)
  call wr('cft_fwd', pl, size(pl))

  ! wall geometry from planeBC, then one step of the textured channel
  allocate(planeBC(0:N1 + 2, N2))
  planeBC = 2
  call rdi('planebc', planeBC(1:N1, :), N1 * N2)
  geometry_type = 1
  call boundary
  fid = [LH_U1, LH_U2, LH_U3, LH_P]
  jl = [0, 0, 0, 1]
  jh = [ny, ny - 1, ny, ny - 1]
  do q = 1, 4
    allocate(fld(jl(q):jh(q), ncol))
    call rdc('f_' // trim(fn(q)), fld, size(fld))
    call lh_check(lh_field_upload(lhctx, 0_c_int, int(fid(q), c_int), fld), 'lh_field_upload')
    deallocate(fld)
  end do
  call lh_check(lh_init_flowrate(lhctx), 'lh_init_flowrate')
  call lh_check(lh_step(lhctx, 1_c_int), 'lh_step')
  do q = 1, 4
    allocate(fld(jl(q):jh(q), ncol))
    call lh_check(lh_field_download(lhctx, 0_c_int, int(fid(q), c_int), fld), 'lh_field_download')
    call wrc('step_' // trim(fn(q)), fld, size(fld))
    deallocate(fld)
  end do
  j = ierr
  call lh_host_finish()
  write(*, '(a)') 'lh_dropin_check: done'

contains

  subroutine rd(name, arr, n)
    character(*), intent(in) :: name
    integer, intent(in) :: n
    real(8), intent(out) :: arr(n)
    integer :: h
    open(newunit=h, file=trim(dir) // '/' // name, access='stream', form='unformatted', status='old', action='read')
    read(h) arr
    close(h)
  end subroutine rd

  subroutine rdi(name, arr, n)
    character(*), intent(in) :: name
    integer, intent(in) :: n
    integer, intent(out) :: arr(n)
    integer :: h
    open(newunit=h, file=trim(dir) // '/' // name, access='stream', form='unformatted', status='old', action='read')
    read(h) arr
    close(h)
  end subroutine rdi

  subroutine wr(name, arr, n)
    character(*), intent(in) :: name
    integer, intent(in) :: n
    real(8), intent(in) :: arr(n)
    integer :: h
    open(newunit=h, file=trim(dir) // '/' // name, access='stream', form='unformatted', status='replace', &
         action='write')
    write(h) arr
    close(h)
  end subroutine wr

  subroutine rdc(name, arr, n)
    character(*), intent(in) :: name
    integer, intent(in) :: n
    complex(8), intent(out) :: arr(n)
    integer :: h
    open(newunit=h, file=trim(dir) // '/' // name, access='stream', form='unformatted', status='old', action='read')
    read(h) arr
    close(h)
  end subroutine rdc

  subroutine wrc(name, arr, n)
    character(*), intent(in) :: name
    integer, intent(in) :: n
    complex(8), intent(in) :: arr(n)
    integer :: h
    open(newunit=h, file=trim(dir) // '/' // name, access='stream', form='unformatted', status='replace', &
         action='write')
    write(h) arr
    close(h)
  end subroutine wrc

end program lh_dropin_check
