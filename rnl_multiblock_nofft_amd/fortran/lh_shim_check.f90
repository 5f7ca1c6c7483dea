! lh_shim_check -- drives the reference-named shims of lh_host on raw data so
! that a test can compare them with an independent implementation:
!   four_to_phys_du / phys_to_four_du on `nplanes` planes (Ngal1+2, Ngal2),
!   four_to_phys_N / phys_to_four_N on planes (N1+2, N2),
!   LUsolU_W, LUsolV, LUsolP and LUsol0 on right-hand sides in the reference
!   layout, and the transposes modes_to_planes_UVP, planes_to_modes_UVP and
!   planes_to_modes_NUVP on u-grid data.
!
!   lh_shim_check check.nml
!
! Files are raw little-endian FP64 streams (no record markers):
!   planes_in   nplanes * (Ngal1+2) * Ngal2
!   rhsu, rhsw  (0:ny, ncol) complex;  rhsv (0:ny-1, ncol) complex
!   planesN_in  nplanes * (N1+2) * N2
!   rhsp (1:ny-1, ncol) complex; rhs0 (0:ny) real
!   colsu (0:ny, ncol) complex; plu (Ngal1+2, Ngal2, 0:ny)
! Outputs: planes_phys (after four_to_phys_du), planes_four (after
! phys_to_four_du of planes_phys), planesN_phys, planesN_four, sol_u, sol_w,
! sol_v, sol_p, sol_0, plu_out (modes_to_planes_UVP of colsu), colsu_out
! (planes_to_modes_UVP of plu), colsn_out (colsu after planes_to_modes_NUVP
! of plu).
program lh_shim_check
  use iso_c_binding
  use lhgpu
  use lh_host
  implicit none
  integer :: N1 = 8, N2 = 8, ny = 17, grid_kind = 1, ppp = 5, flag_ctpress = 0
  integer :: nplanes = 1, kstage = 1
  real(8) :: alp = 1d0, bet = 2d0, Re = 4200d0, dyq = 10d0, gamma = 2d0, dt = 1d-3
  character(len=1024) :: dir = '.', nmlfile
  namelist /check/ N1, N2, ny, alp, bet, Re, grid_kind, dyq, ppp, gamma, flag_ctpress, dt, &
                   nplanes, kstage, dir
  type(lh_config) :: cfg(1)
  real(8), allocatable :: pl(:,:,:), plN(:,:,:), r0(:), plu(:,:,:), plo(:,:,:)
  complex(8), allocatable :: ru(:,:), rw(:,:), su(:,:), sw(:,:), rv(:,:), sv(:,:), rp(:,:)
  complex(8), allocatable :: cu(:,:), co(:,:)
  integer :: u, q, ncol

  call get_command_argument(1, nmlfile)
  open(newunit=u, file=trim(nmlfile), status='old', action='read')
  read(u, nml=check)
  close(u)
  cfg(1) = lh_config(N1, N2, 3 * N1 / 2, 3 * N2 / 2, ny, alp, bet, Re, grid_kind, dyq, ppp, &
                     gamma, 0d0, 0d0, flag_ctpress, 0d0, dt, LH_PREC_FP64, 0, 0.5d0)
  call lh_host_setup(cfg)
  ncol = (N1 / 2 + 1) * N2

  allocate(pl(Ngal(1, 1) + 2, Ngal(2, 1), nplanes))
  call rd(trim(dir) // '/planes_in', pl, size(pl))
  do q = 1, nplanes
    call four_to_phys_du(pl(:, :, q), 1)
  end do
  call wr(trim(dir) // '/planes_phys', pl, size(pl))
  do q = 1, nplanes
    call phys_to_four_du(pl(:, :, q), 1)
  end do
  call wr(trim(dir) // '/planes_four', pl, size(pl))

  kRK = kstage
  allocate(ru(0:ny, ncol), rw(0:ny, ncol), su(0:ny, ncol), sw(0:ny, ncol))
  allocate(rv(0:ny - 1, ncol), sv(0:ny - 1, ncol))
  call rdc(trim(dir) // '/rhsu', ru, size(ru))
  call rdc(trim(dir) // '/rhsw', rw, size(rw))
  call rdc(trim(dir) // '/rhsv', rv, size(rv))
  call LUsolU_W(su, ru, sw, rw, 1)
  call LUsolV(sv, rv, 1)
  call wrc(trim(dir) // '/sol_u', su, size(su))
  call wrc(trim(dir) // '/sol_w', sw, size(sw))
  call wrc(trim(dir) // '/sol_v', sv, size(sv))

  allocate(plN(N1 + 2, N2, nplanes))
  call rd(trim(dir) // '/planesN_in', plN, size(plN))
  do q = 1, nplanes
    call four_to_phys_N(plN(:, :, q), 1)
  end do
  call wr(trim(dir) // '/planesN_phys', plN, size(plN))
  do q = 1, nplanes
    call phys_to_four_N(plN(:, :, q), 1)
  end do
  call wr(trim(dir) // '/planesN_four', plN, size(plN))

  allocate(rp(1:ny - 1, ncol), r0(0:ny))
  call rdc(trim(dir) // '/rhsp', rp, size(rp))
  call LUsolP(rp, 1)
  call wrc(trim(dir) // '/sol_p', rp, size(rp))
  call rd(trim(dir) // '/rhs0', r0, size(r0))
  call LUsol0(r0, 1)
  call wr(trim(dir) // '/sol_0', r0, size(r0))

  allocate(cu(0:ny, ncol), co(0:ny, ncol))
  allocate(plu(Ngal(1, 1) + 2, Ngal(2, 1), 0:ny), plo(Ngal(1, 1) + 2, Ngal(2, 1), 0:ny))
  call rdc(trim(dir) // '/colsu', cu, size(cu))
  call rd(trim(dir) // '/plu', plu, size(plu))
  call modes_to_planes_UVP(plo, cu, 2, 1)
  call wr(trim(dir) // '/plu_out', plo, size(plo))
  call planes_to_modes_UVP(co, plu, 2, 1)
  call wrc(trim(dir) // '/colsu_out', co, size(co))
  call planes_to_modes_NUVP(cu, plu, 2, 1)
  call wrc(trim(dir) // '/colsn_out', cu, size(cu))
  call lh_host_finish()
  write(*, '(a)') 'lh_shim_check: done'

contains

  subroutine rd(name, a, n)
    character(*), intent(in) :: name
    integer, intent(in) :: n
    real(8), intent(out) :: a(n)
    integer :: v
    open(newunit=v, file=name, access='stream', form='unformatted', status='old', action='read')
    read(v) a
    close(v)
  end subroutine rd

  subroutine wr(name, a, n)
    character(*), intent(in) :: name
    integer, intent(in) :: n
    real(8), intent(in) :: a(n)
    integer :: v
    open(newunit=v, file=name, access='stream', form='unformatted', status='replace', action='write')
    write(v) a
    close(v)
  end subroutine wr

  subroutine rdc(name, a, n)
    character(*), intent(in) :: name
    integer, intent(in) :: n
    complex(8), intent(out) :: a(n)
    integer :: v
    open(newunit=v, file=name, access='stream', form='unformatted', status='old', action='read')
    read(v) a
    close(v)
  end subroutine rdc

  subroutine wrc(name, a, n)
    character(*), intent(in) :: name
    integer, intent(in) :: n
    complex(8), intent(in) :: a(n)
    integer :: v
    open(newunit=v, file=name, access='stream', form='unformatted', status='replace', action='write')
    write(v) a
    close(v)
  end subroutine wrc

end program lh_shim_check
