! lh_shim_check -- drives the reference-named shims of lh_host on raw data so
! that a test can compare them with an independent implementation:
!   four_to_phys_du / phys_to_four_du on `nplanes` planes (Ngal1+2, Ngal2),
!   LUsolU_W and LUsolV on right-hand sides in the reference layout.
!
!   lh_shim_check check.nml
!
! Files are raw little-endian FP64 streams (no record markers):
!   planes_in   nplanes * (Ngal1+2) * Ngal2
!   rhsu, rhsw  (0:ny, ncol) complex;  rhsv (0:ny-1, ncol) complex
! Outputs: planes_phys (after four_to_phys_du), planes_four (after
! phys_to_four_du of planes_phys), sol_u, sol_w, sol_v.
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
  real(8), allocatable :: pl(:,:,:)
  complex(8), allocatable :: ru(:,:), rw(:,:), su(:,:), sw(:,:), rv(:,:), sv(:,:)
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
