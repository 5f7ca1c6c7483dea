! lh_host.f90 -- Fortran host layer over liblhgpu: the reference's subroutine
! names and argument meanings, forwarding to the MI355X kernels.
!
! Module state mirrors what the reference keeps in `module declaration`
! (declaration.f90): N(1:2,iband), Ngal(1:2,iband), the substep index kRK,
! the number of bands.  Here one "band" is one channel of the device
! context (the reference's bands split one channel in y across MPI ranks;
! on one GPU a channel is never split, so every band index selects a whole
! channel and several bands form an ensemble).
!
! Routines with the reference's names:
!   four_to_phys_u, four_to_phys_du, phys_to_four_du   (FOU3D.f90:738-784,2253)
!   four_to_phys_N, phys_to_four_N                      (FOU3D.f90:786-804,2274-2293)
!   LUsolU_W, LUsolV, LUsolP, LUsol0                    (tridLU_3D.f90:55,135,192,228)
!   modes_to_planes_UVP, planes_to_modes_UVP/_NUVP      (FOU3D.f90:1000,2295,2772)
!   read_in_field, record_out_field                     (start.f90:2444, FOU3D.f90:2856)
!   write_stats_file                                    (stats.f90:254-441)
!   hist_open, hist_write                               (start.f90:238-253, FOU3D.f90:3040-3070)
!   nonlinRead                                          (start.f90:2983-3069)
! The LU factor argument `a`, `grid` and `myid` of LUsolU_W/LUsolV are
! dropped: the factors of the fixed-dt operators live on the device (built
! once at setup, as LU_buildP is in the reference), and there is one rank.
module lh_host
  use iso_c_binding
  use lhgpu
  implicit none

  type(c_ptr), save :: lhctx = c_null_ptr
  integer, save :: nband = 0
  integer, allocatable, save :: N(:,:), Ngal(:,:), nyb(:)
  integer, save :: kRK = 1
  ! the device channel (0-based) that the reference-signature routines of
  ! lh_dropin.f90 act on: the reference's bands of one uniform-resolution
  ! channel all map to it
  integer, save :: dropin_chan = 0

  ! device staging for the plane and solve shims
  type(c_ptr), save :: dstage = c_null_ptr
  integer(c_size_t), save :: dstage_bytes = 0

  ! header of a field file, kept so that record_out writes what read_in read
  type :: field_header
    real(8) :: t = 0d0, Re = 0d0, alp = 1d0, bet = 2d0, mpgx = 0d0
    integer :: nband = 1, iter = 0
    integer, allocatable :: Nb(:,:)
    real(8), allocatable :: y(:), dthdy(:)
    real(8) :: dthetavi = 1d0
    ! as read: the grid code and y of each plane record (yrow(jlo:jhi))
    integer :: grid = 0
    real(8), allocatable :: yrow(:)
  end type field_header

contains

  ! ---------------------------------------------------------------------
  ! set-up / tear-down (replaces start's per-band setup, start.f90:36-691)
  subroutine lh_host_setup(cfgs)
    type(lh_config), intent(in) :: cfgs(:)
    integer :: b
    nband = size(cfgs)
    allocate(N(2, nband), Ngal(2, nband), nyb(nband))
    do b = 1, nband
      N(1, b) = cfgs(b)%N1;  N(2, b) = cfgs(b)%N2
      Ngal(1, b) = cfgs(b)%Ngal1;  Ngal(2, b) = cfgs(b)%Ngal2
      nyb(b) = cfgs(b)%ny
    end do
    call lh_check(lh_ctx_create(cfgs, int(nband, c_int), lhctx), 'lh_ctx_create')
  end subroutine lh_host_setup

  subroutine lh_host_finish()
    if (c_associated(dstage)) call lh_check(lh_dev_free(dstage), 'lh_dev_free')
    dstage = c_null_ptr
    dstage_bytes = 0
    if (c_associated(lhctx)) call lh_check(lh_ctx_destroy(lhctx), 'lh_ctx_destroy')
    lhctx = c_null_ptr
    if (allocated(N)) deallocate(N, Ngal, nyb)
    nband = 0
  end subroutine lh_host_finish

  ! device staging buffer of at least `bytes`; returns the pointer `off`
  ! bytes into it
  function staging(bytes, off) result(p)
    integer(c_size_t), intent(in) :: bytes, off
    type(c_ptr) :: p
    integer(c_intptr_t) :: a
    if (bytes > dstage_bytes) then
      if (c_associated(dstage)) call lh_check(lh_dev_free(dstage), 'lh_dev_free')
      call lh_check(lh_dev_alloc(bytes, dstage), 'lh_dev_alloc')
      dstage_bytes = bytes
    end if
    a = transfer(dstage, a) + int(off, c_intptr_t)
    p = transfer(a, p)
  end function staging

  ! ---------------------------------------------------------------------
  ! FOU3D.f90:766 four_to_phys_du: oddball row Ngal2/2+1 zeroed, inverse z
  ! transform of the retained x-modes, inverse x transform; in place
  subroutine four_to_phys_du(du, iband)
    integer, intent(in) :: iband
    real(8), intent(inout), target :: du(Ngal(1, iband) + 2, Ngal(2, iband))
    call plane_transform(du, iband, 1)
  end subroutine four_to_phys_du

  ! FOU3D.f90:738 four_to_phys_u: the same for three fields
  subroutine four_to_phys_u(u1, u2, u3, iband)
    integer, intent(in) :: iband
    real(8), intent(inout), target :: u1(Ngal(1, iband) + 2, Ngal(2, iband))
    real(8), intent(inout), target :: u2(Ngal(1, iband) + 2, Ngal(2, iband))
    real(8), intent(inout), target :: u3(Ngal(1, iband) + 2, Ngal(2, iband))
    call plane_transform(u1, iband, 1)
    call plane_transform(u2, iband, 1)
    call plane_transform(u3, iband, 1)
  end subroutine four_to_phys_u

  ! FOU3D.f90:2253 phys_to_four_du: forward x and z transforms (1/n each),
  ! oddball row zeroed; in place
  subroutine phys_to_four_du(duPL, iband)
    integer, intent(in) :: iband
    real(8), intent(inout), target :: duPL(Ngal(1, iband) + 2, Ngal(2, iband))
    call plane_transform(duPL, iband, -1)
  end subroutine phys_to_four_du

  subroutine plane_transform(pl, iband, dir)
    integer, intent(in) :: iband, dir
    real(8), intent(inout), target :: pl(Ngal(1, iband) + 2, Ngal(2, iband))
    integer(c_size_t) :: bytes
    type(c_ptr) :: d
    bytes = int(size(pl), c_size_t) * 8_c_size_t
    d = staging(bytes, 0_c_size_t)
    call lh_check(lh_copy_to_device(d, c_loc(pl), bytes), 'lh_copy_to_device')
    if (dir > 0) then
      call lh_check(lh_four_to_phys_planes(int(N(1, iband), c_int), int(Ngal(1, iband), c_int), &
                    int(Ngal(2, iband), c_int), d, 1_c_int, c_null_ptr), 'lh_four_to_phys_planes')
    else
      call lh_check(lh_phys_to_four_planes(int(N(1, iband), c_int), int(Ngal(1, iband), c_int), &
                    int(Ngal(2, iband), c_int), d, 1_c_int, c_null_ptr), 'lh_phys_to_four_planes')
    end if
    call lh_check(lh_copy_to_host(c_loc(pl), d, bytes), 'lh_copy_to_host')
  end subroutine plane_transform

  ! FOU3D.f90:786 four_to_phys_N / :2274 phys_to_four_N: the same transforms
  ! on the N1 x N2 grid, planes (N1+2, N2), no oddball-row zeroing
  subroutine four_to_phys_N(du, iband)
    integer, intent(in) :: iband
    real(8), intent(inout), target :: du(N(1, iband) + 2, N(2, iband))
    call plane_transform_N(du, iband, 1)
  end subroutine four_to_phys_N

  subroutine phys_to_four_N(du, iband)
    integer, intent(in) :: iband
    real(8), intent(inout), target :: du(N(1, iband) + 2, N(2, iband))
    call plane_transform_N(du, iband, -1)
  end subroutine phys_to_four_N

  subroutine plane_transform_N(pl, iband, dir)
    integer, intent(in) :: iband, dir
    real(8), intent(inout), target :: pl(N(1, iband) + 2, N(2, iband))
    integer(c_size_t) :: bytes
    type(c_ptr) :: d
    bytes = int(size(pl), c_size_t) * 8_c_size_t
    d = staging(bytes, 0_c_size_t)
    call lh_check(lh_copy_to_device(d, c_loc(pl), bytes), 'lh_copy_to_device')
    if (dir > 0) then
      call lh_check(lh_four_to_phys_N_planes(int(N(1, iband), c_int), int(N(2, iband), c_int), d, 1_c_int, &
                    c_null_ptr), 'lh_four_to_phys_N_planes')
    else
      call lh_check(lh_phys_to_four_N_planes(int(N(1, iband), c_int), int(N(2, iband), c_int), d, 1_c_int, &
                    c_null_ptr), 'lh_phys_to_four_N_planes')
    end if
    call lh_check(lh_copy_to_host(c_loc(pl), d, bytes), 'lh_copy_to_host')
  end subroutine plane_transform_N

  ! ---------------------------------------------------------------------
  ! Column <-> plane transposes (FOU3D.f90:1000 modes_to_planes_UVP, :2295
  ! planes_to_modes_UVP, :2772 planes_to_modes_NUVP) for one band: x holds
  ! rows jlo:jhi of the grid (1 = v: 0..ny-1, 2 = u: 0..ny, 3 = p: 1..ny-1),
  ! xPL the same rows as (Ngal1+2, Ngal2) planes.  modes_to_planes_UVP zeroes
  ! the planes outside the stored modes; planes_to_modes_NUVP fills only the
  ! rows inside the walls (jlo+1..jhi-1), as the reference's does.
  subroutine grid_rows(grid, iband, jlo, jhi)
    integer, intent(in) :: grid, iband
    integer, intent(out) :: jlo, jhi
    select case (grid)
    case (1)
      jlo = 0; jhi = nyb(iband) - 1
    case (2)
      jlo = 0; jhi = nyb(iband)
    case default
      jlo = 1; jhi = nyb(iband) - 1
    end select
  end subroutine grid_rows

  subroutine modes_to_planes_UVP(xPL, x, grid, iband)
    integer, intent(in) :: grid, iband
    real(8), intent(out), target :: xPL(Ngal(1, iband) + 2, Ngal(2, iband), *)
    complex(8), intent(in), target :: x(*)
    call transpose_band(xPL, x, grid, iband, 1, 0)
  end subroutine modes_to_planes_UVP

  subroutine planes_to_modes_UVP(x, xPL, grid, iband)
    integer, intent(in) :: grid, iband
    complex(8), intent(inout), target :: x(*)
    real(8), intent(in), target :: xPL(Ngal(1, iband) + 2, Ngal(2, iband), *)
    call transpose_band(xPL, x, grid, iband, -1, 0)
  end subroutine planes_to_modes_UVP

  subroutine planes_to_modes_NUVP(x, xPL, grid, iband)
    integer, intent(in) :: grid, iband
    complex(8), intent(inout), target :: x(*)
    real(8), intent(in), target :: xPL(Ngal(1, iband) + 2, Ngal(2, iband), *)
    call transpose_band(xPL, x, grid, iband, -1, 1)
  end subroutine planes_to_modes_NUVP

  subroutine transpose_band(xPL, x, grid, iband, dir, inner)
    integer, intent(in) :: grid, iband, dir, inner
    ! written or read according to dir
    real(8), target :: xPL(Ngal(1, iband) + 2, Ngal(2, iband), *)
    complex(8), target :: x(*)
    integer :: jlo, jhi, nrows, ncol, q
    integer(c_size_t) :: bc, bp
    type(c_ptr) :: dc, dp
    complex(8), allocatable, target :: tmp(:,:)
    call grid_rows(grid, iband, jlo, jhi)
    nrows = jhi - jlo + 1
    ncol = (N(1, iband) / 2 + 1) * N(2, iband)
    bc = int(nrows, c_size_t) * int(ncol, c_size_t) * 16_c_size_t
    bp = int(nrows, c_size_t) * int(Ngal(1, iband) + 2, c_size_t) * int(Ngal(2, iband), c_size_t) * 8_c_size_t
    dc = staging(bc + bp, 0_c_size_t)
    dp = staging(bc + bp, bc)
    if (dir > 0) then
      call lh_check(lh_copy_to_device(dc, c_loc(x), bc), 'lh_copy_to_device')
      call lh_check(lh_modes_to_planes(int(N(1, iband), c_int), int(N(2, iband), c_int), &
                    int(Ngal(1, iband), c_int), int(Ngal(2, iband), c_int), dc, int(nrows, c_int), dp, &
                    c_null_ptr), 'lh_modes_to_planes')
      call lh_check(lh_copy_to_host(c_loc(xPL), dp, bp), 'lh_copy_to_host')
    else
      call lh_check(lh_copy_to_device(dp, c_loc(xPL), bp), 'lh_copy_to_device')
      call lh_check(lh_planes_to_modes(int(N(1, iband), c_int), int(N(2, iband), c_int), &
                    int(Ngal(1, iband), c_int), int(Ngal(2, iband), c_int), dp, int(nrows, c_int), dc, &
                    c_null_ptr), 'lh_planes_to_modes')
      if (inner == 0) then
        call lh_check(lh_copy_to_host(c_loc(x), dc, bc), 'lh_copy_to_host')
      else
        ! only the rows inside the walls change
        allocate(tmp(jlo:jhi, ncol))
        call lh_check(lh_copy_to_host(c_loc(tmp), dc, bc), 'lh_copy_to_host')
        do q = 1, ncol
          x((q - 1) * nrows + 2:(q - 1) * nrows + nrows - 1) = tmp(jlo + 1:jhi - 1, q)
        end do
        deallocate(tmp)
      end if
    end if
  end subroutine transpose_band

  ! ---------------------------------------------------------------------
  ! tridLU_3D.f90:55 LUsolU_W: (I - beta (D2 + k^2)) u = rhsu and the same
  ! operator for w, for substep kRK, all (kx,kz) columns of band iband.
  ! Arrays in the reference layout u(0:ny, ncol).
  subroutine LUsolU_W(u, rhsu, w, rhsw, iband)
    integer, intent(in) :: iband
    complex(8), intent(out), target :: u(0:nyb(iband), (N(1, iband) / 2 + 1) * N(2, iband))
    complex(8), intent(in), target :: rhsu(0:nyb(iband), (N(1, iband) / 2 + 1) * N(2, iband))
    complex(8), intent(out), target :: w(0:nyb(iband), (N(1, iband) / 2 + 1) * N(2, iband))
    complex(8), intent(in), target :: rhsw(0:nyb(iband), (N(1, iband) / 2 + 1) * N(2, iband))
    integer(c_size_t) :: b
    type(c_ptr) :: d1, d2, d3, d4
    b = int(size(u), c_size_t) * 16_c_size_t
    d1 = staging(4 * b, 0_c_size_t)
    d2 = staging(4 * b, b)
    d3 = staging(4 * b, 2 * b)
    d4 = staging(4 * b, 3 * b)
    call lh_check(lh_copy_to_device(d1, c_loc(rhsu), b), 'lh_copy_to_device')
    call lh_check(lh_copy_to_device(d2, c_loc(rhsw), b), 'lh_copy_to_device')
    call lh_check(lh_lusol_uw(lhctx, int(iband - 1, c_int), int(kRK, c_int), d1, d2, d3, d4), 'lh_lusol_uw')
    call lh_check(lh_copy_to_host(c_loc(u), d3, b), 'lh_copy_to_host')
    call lh_check(lh_copy_to_host(c_loc(w), d4, b), 'lh_copy_to_host')
  end subroutine LUsolU_W

  ! tridLU_3D.f90:135 LUsolV: the v-grid operator, rows 0:ny-1 (walls v = 0)
  subroutine LUsolV(u, rhsu, iband)
    integer, intent(in) :: iband
    complex(8), intent(out), target :: u(0:nyb(iband) - 1, (N(1, iband) / 2 + 1) * N(2, iband))
    complex(8), intent(in), target :: rhsu(0:nyb(iband) - 1, (N(1, iband) / 2 + 1) * N(2, iband))
    integer(c_size_t) :: b
    type(c_ptr) :: d1, d2
    b = int(size(u), c_size_t) * 16_c_size_t
    d1 = staging(2 * b, 0_c_size_t)
    d2 = staging(2 * b, b)
    call lh_check(lh_copy_to_device(d1, c_loc(rhsu), b), 'lh_copy_to_device')
    call lh_check(lh_lusol_v(lhctx, int(iband - 1, c_int), int(kRK, c_int), d1, d2), 'lh_lusol_v')
    call lh_check(lh_copy_to_host(c_loc(u), d2, b), 'lh_copy_to_host')
  end subroutine LUsolV

  ! tridLU_3D.f90:192 LUsolP: the Poisson operator of LU_buildP (with the
  ! mean-mode pin) solved in place on x(1:ny-1, ncol)
  subroutine LUsolP(x, iband)
    integer, intent(in) :: iband
    complex(8), intent(inout), target :: x(1:nyb(iband) - 1, (N(1, iband) / 2 + 1) * N(2, iband))
    integer(c_size_t) :: b
    type(c_ptr) :: d
    b = int(size(x), c_size_t) * 16_c_size_t
    d = staging(b, 0_c_size_t)
    call lh_check(lh_copy_to_device(d, c_loc(x), b), 'lh_copy_to_device')
    call lh_check(lh_lusol_p(lhctx, int(iband - 1, c_int), d), 'lh_lusol_p')
    call lh_check(lh_copy_to_host(c_loc(x), d, b), 'lh_copy_to_host')
  end subroutine LUsolP

  ! tridLU_3D.f90:228 LUsol0: the real (0,0) U system of substep kRK, in
  ! place on x(0:ny)
  subroutine LUsol0(x, iband)
    integer, intent(in) :: iband
    real(8), intent(inout), target :: x(0:nyb(iband))
    integer(c_size_t) :: b
    type(c_ptr) :: d
    b = int(size(x), c_size_t) * 8_c_size_t
    d = staging(b, 0_c_size_t)
    call lh_check(lh_copy_to_device(d, c_loc(x), b), 'lh_copy_to_device')
    call lh_check(lh_lusol0(lhctx, int(iband - 1, c_int), int(kRK, c_int), d), 'lh_lusol0')
    call lh_check(lh_copy_to_host(c_loc(x), d, b), 'lh_copy_to_host')
  end subroutine LUsol0

  ! ---------------------------------------------------------------------
  ! Field files in the format record_out writes and read_in reads
  ! (FOU3D.f90:2856-2920, start.f90:2444-2882): unformatted sequential
  ! records with 4-byte markers.  Record 1: t,Re,alp,bet,mpgx,nband,iter +
  ! 88 int; record 2: N(4,0:nband+1); record 3: y, dthetavi, dthdy; then one
  ! record per y row: j, grid, nx, nz, y(j), plane(nx,nz) with plane(2i+1,k),
  ! plane(2i+2,k) the real/imaginary parts of x-mode i, z-index k.
  ! Handled with stream access and explicit markers so that no
  ! compiler-specific record convention is assumed.

  ! read the file into f(jlo:jhi, ncol) (rows as present in the file)
  subroutine read_in_field(fname, hdr, f, jlo, jhi)
    character(*), intent(in) :: fname
    type(field_header), intent(out) :: hdr
    complex(8), allocatable, intent(out) :: f(:,:)
    integer, intent(out) :: jlo, jhi
    integer :: u, ios, m, m2, L, j, grid, nx, nz, nrec, q
    integer(8) :: pos, fsize
    real(8) :: yj
    real(8), allocatable :: pl(:,:)
    integer, allocatable :: js(:)
    open(newunit=u, file=fname, access='stream', form='unformatted', status='old', action='read', iostat=ios)
    if (ios /= 0) then
      write(*, '(a,a)') 'read_in_field: cannot open ', trim(fname)
      error stop 1
    end if
    inquire(unit=u, size=fsize)
    pos = 1
    read(u, pos=pos) m
    read(u) hdr%t, hdr%Re, hdr%alp, hdr%bet, hdr%mpgx, hdr%nband, hdr%iter
    pos = pos + 8 + m
    read(u, pos=pos) m
    allocate(hdr%Nb(4, 0:hdr%nband + 1))
    read(u) hdr%Nb
    pos = pos + 8 + m
    read(u, pos=pos) m
    L = (m / 8 - 1) / 2
    allocate(hdr%y(L), hdr%dthdy(L))
    read(u) hdr%y, hdr%dthetavi, hdr%dthdy
    pos = pos + 8 + m
    ! first pass: count rows and find j range
    nrec = 0
    jlo = huge(jlo); jhi = -huge(jhi)
    do while (pos < fsize)
      read(u, pos=pos) m
      read(u) j, grid, nx, nz
      nrec = nrec + 1
      jlo = min(jlo, j); jhi = max(jhi, j)
      pos = pos + 8 + m
    end do
    if (jhi - jlo + 1 /= nrec) then
      write(*, '(a)') 'read_in_field: rows are not a contiguous range'
      error stop 1
    end if
    allocate(f(jlo:jhi, (nx / 2) * nz), pl(nx, nz), js(nrec), hdr%yrow(jlo:jhi))
    pos = 1
    do q = 1, 3
      read(u, pos=pos) m
      pos = pos + 8 + m
    end do
    do q = 1, nrec
      read(u, pos=pos) m
      read(u) j, grid, nx, nz, yj, pl
      hdr%grid = grid
      hdr%yrow(j) = yj
      read(u) m2
      if (m2 /= m) then
        write(*, '(a)') 'read_in_field: corrupt record marker'
        error stop 1
      end if
      f(j, :) = reshape(cmplx(pl(1::2, :), pl(2::2, :), kind=8), [(nx / 2) * nz])
      pos = pos + 8 + m
    end do
    close(u)
  end subroutine read_in_field

  subroutine record_out_field(fname, hdr, f, jlo, jhi, grid, nx, nz, yrow)
    character(*), intent(in) :: fname
    type(field_header), intent(in) :: hdr
    integer, intent(in) :: jlo, jhi, grid, nx, nz
    complex(8), intent(in) :: f(jlo:jhi, (nx / 2) * nz)
    real(8), intent(in) :: yrow(jlo:jhi)
    integer :: u, m, j
    real(8) :: pl(nx, nz)
    complex(8) :: c(nx / 2, nz)
    integer, parameter :: zeros(88) = 0
    open(newunit=u, file=fname, access='stream', form='unformatted', status='replace', action='write')
    m = 5 * 8 + 2 * 4 + 88 * 4
    write(u) m, hdr%t, hdr%Re, hdr%alp, hdr%bet, hdr%mpgx, hdr%nband, hdr%iter, zeros, m
    m = 4 * size(hdr%Nb)
    write(u) m, hdr%Nb, m
    m = 8 * (2 * size(hdr%y) + 1)
    write(u) m, hdr%y, hdr%dthetavi, hdr%dthdy, m
    do j = jlo, jhi
      c = reshape(f(j, :), [nx / 2, nz])
      pl(1::2, :) = real(c, 8)
      pl(2::2, :) = aimag(c)
      m = 4 * 4 + 8 + 8 * nx * nz
      write(u) m, j, grid, nx, nz, yrow(j), pl, m
    end do
    close(u)
  end subroutine record_out_field

  ! ---------------------------------------------------------------------
  ! Statistics file of write_stats (stats.f90:254-441) for one process,
  ! from the device accumulators of band `iband`: header, Ngal, yu/dthetai/
  ! dthdyu, yv/dthetai/dthdyv, istat, then Um U2m Vm V2m Wm W2m Pm P2m UVm
  ! UWm VWm wxm wx2m on their rows (u 0..ny, v 0..ny-1, p 1..ny-1).  The
  ! tile-conditional records of the reference's wall bands are not written.
  ! Like write_stats, the accumulators are reset afterwards (all bands).
  subroutine write_stats_file(fname, hdr, ngal_arr, iband)
    character(*), intent(in) :: fname
    type(field_header), intent(in) :: hdr
    integer, intent(in) :: ngal_arr(:,:), iband
    real(8), allocatable :: prof(:,:), yu(:), yv(:), du(:), dv(:)
    integer(c_int) :: istat
    integer :: u, m, q, ny, j0, nr
    integer, parameter :: zeros(88) = 0
    ! grid of each profile: 0 u (rows 0..ny), 1 v (0..ny-1), 2 p (1..ny-1)
    integer, parameter :: grid_of(0:12) = [0, 0, 1, 1, 0, 0, 2, 2, 0, 0, 1, 1, 1]
    ny = nyb(iband)
    allocate(prof(0:ny, 0:LH_ST_N - 1), yu(0:ny), yv(0:ny - 1), du(0:ny), dv(0:ny - 1))
    call lh_check(lh_stats_get(lhctx, int(iband - 1, c_int), prof, istat), 'lh_stats_get')
    call lh_check(lh_get_grid(lhctx, int(iband - 1, c_int), yu, yv, du, dv), 'lh_get_grid')
    open(newunit=u, file=fname, access='stream', form='unformatted', status='replace', action='write')
    m = 5 * 8 + 2 * 4 + 88 * 4
    write(u) m, hdr%t, hdr%Re, hdr%alp, hdr%bet, hdr%mpgx, hdr%nband, hdr%iter, zeros, m
    m = 4 * size(ngal_arr)
    write(u) m, ngal_arr, m
    m = 8 * (2 * (ny + 1) + 1)
    write(u) m, yu, hdr%dthetavi, du, m
    m = 8 * (2 * ny + 1)
    write(u) m, yv, hdr%dthetavi, dv, m
    m = 4
    write(u) m, int(istat), m
    do q = 0, LH_ST_N - 1
      select case (grid_of(q))
      case (0)
        j0 = 0; nr = ny + 1
      case (1)
        j0 = 0; nr = ny
      case default
        j0 = 1; nr = ny - 1
      end select
      m = 8 * nr
      write(u) m, prof(j0:j0 + nr - 1, q), m
    end do
    close(u)
    call lh_check(lh_stats_reset(lhctx), 'lh_stats_reset')
  end subroutine write_stats_file

  ! ---------------------------------------------------------------------
  ! History file (start.f90:238-253): opened for appending when it exists,
  ! else created with the records (t, Re, alp, bet, mpgx, nband) and
  ! N(4, 0:nband+1); hist_write appends record_out's line (FOU3D.f90:
  ! 3063-3068) from the device's history reduction (lh_history).
  subroutine hist_open(fname, t, Re, alp, bet, mpgx, Nb, u)
    character(*), intent(in) :: fname
    real(8), intent(in) :: t, Re, alp, bet, mpgx
    integer, intent(in) :: Nb(:,:)
    integer, intent(out) :: u
    logical :: ex
    integer :: m
    inquire(file=fname, exist=ex)
    if (ex) then
      open(newunit=u, file=fname, access='stream', form='unformatted', status='old', position='append', &
           action='write')
    else
      open(newunit=u, file=fname, access='stream', form='unformatted', status='new', action='write')
      m = 5 * 8 + 4
      write(u) m, t, Re, alp, bet, mpgx, size(Nb, 2) - 2, m
      m = 4 * size(Nb)
      write(u) m, Nb, m
    end if
  end subroutine hist_open

  subroutine hist_write(u, flag_ctpress, iband)
    integer, intent(in) :: u, flag_ctpress, iband
    real(c_double) :: h(12)
    integer :: m
    call lh_check(lh_history(lhctx, int(iband - 1, c_int), h), 'lh_history')
    ! h = t, dtv, dtc, dt, err, Qx, QxT, mpgx, dgx, Umax, Uslip, iter
    if (flag_ctpress == 0) then
      m = 2 * 4 + 11 * 8
      write(u) m, flag_ctpress, nint(h(12)), h(1:11), m
    else
      m = 2 * 4 + 9 * 8
      write(u) m, flag_ctpress, nint(h(12)), h(1:6), h(8), h(10:11), m
    end if
    flush(u)
  end subroutine hist_write

  ! ---------------------------------------------------------------------
  ! start.f90:2983 nonlinRead: the triad lists of band iband from the files
  ! trim(dirlist)//trim(heading)//'NN.dat' (NN = 01.., stream access: nx, nz,
  ! j, then 9 x (type, length, int8 list(length,4))), with the reference's
  ! plane mapping (u planes above the centre use index j-1, FOU3D.f90:
  ! 1811-1815; planes above the centre read file Ngal3+1-j, start.f90:
  ! 3025-3030, Ngal3 = ny-2), handed to the device (lh_set_triads).  A plane
  ! whose file is absent gets no triads, and so does every plane when
  ! `missing_ok` is false and a file is absent: then the routine stops.
  subroutine nonlinRead(dirlist, heading, iband, missing_ok)
    character(*), intent(in) :: dirlist, heading
    integer, intent(in) :: iband
    logical, intent(in) :: missing_ok
    type :: rowbuf
      integer(c_signed_char), allocatable :: b(:)
    end type rowbuf
    type(rowbuf) :: per(9)
    integer :: ny, g3, nfile, q, j, jj, u, ios, t, nt, hdr(3), th(2), nrow, mid
    integer(c_int), allocatable :: lens(:), mu(:), mv(:), have(:)
    integer(c_signed_char), allocatable :: rows(:), tmp(:)
    character(len=2) :: ext
    character(len=1024) :: fname
    logical :: ex
    ny = nyb(iband)
    g3 = ny - 2
    mid = (g3 + 1) / 2
    nfile = mid
    allocate(lens(9 * nfile), mu(ny - 1), mv(ny - 2), have(nfile))
    lens = 0
    have = 0
    allocate(rows(0))
    nrow = 0
    do q = 1, nfile
      write(ext, '(i2.2)') q
      fname = trim(dirlist) // trim(heading) // ext // '.dat'
      inquire(file=trim(fname), exist=ex)
      if (.not. ex) then
        if (.not. missing_ok) then
          write(*, '(a,a)') 'nonlinRead: missing ', trim(fname)
          error stop 1
        end if
        cycle
      end if
      have(q) = 1
      open(newunit=u, file=trim(fname), access='stream', form='unformatted', status='old', action='read', &
           iostat=ios)
      read(u) hdr
      if (hdr(1) /= N(1, iband) / 2 - 1 .or. hdr(2) /= N(2, iband) / 2 - 1 .or. hdr(3) /= q) then
        write(*, '(a,a,3i6)') 'nonlinRead: header does not match the run: ', trim(fname), hdr
        error stop 1
      end if
      do t = 1, 9
        read(u) th
        nt = th(2)
        lens(9 * (q - 1) + th(1)) = nt
        if (allocated(per(th(1))%b)) deallocate(per(th(1))%b)
        allocate(per(th(1))%b(4 * nt))
        read(u) per(th(1))%b
      end do
      close(u)
      ! rows of a file in type order 1..9
      do t = 1, 9
        allocate(tmp(nrow + size(per(t)%b)))
        tmp(1:nrow) = rows(1:nrow)
        tmp(nrow + 1:) = per(t)%b
        nrow = size(tmp)
        call move_alloc(tmp, rows)
      end do
    end do
    do j = 1, ny - 1
      jj = j
      if (j > mid) jj = j - 1
      if (jj > mid) jj = g3 + 1 - jj
      mu(j) = merge(jj - 1, -1, have(jj) == 1)
    end do
    do j = 1, ny - 2
      jj = j
      if (jj > mid) jj = g3 + 1 - jj
      mv(j) = merge(jj - 1, -1, have(jj) == 1)
    end do
    if (nrow == 0) then
      deallocate(rows)
      allocate(rows(1))
    end if
    call lh_check(lh_set_triads(lhctx, int(iband - 1, c_int), int(nfile, c_int), lens, rows, mu, mv), &
                  'lh_set_triads')
  end subroutine nonlinRead

end module lh_host
