! lh_io_check -- reads a field file with read_in_field and writes it back
! with record_out_field (the record_out / read_in layout, FOU3D.f90:
! 2856-2920, start.f90:2444-2882), so that a test can compare the bytes.
! No device call is made.
!
!   lh_io_check in.dat out.dat
program lh_io_check
  use lh_host
  implicit none
  character(len=1024) :: fin, fout
  type(field_header) :: hdr
  complex(8), allocatable :: f(:,:)
  integer :: jlo, jhi, nx, nz
  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  call read_in_field(trim(fin), hdr, f, jlo, jhi)
  nz = hdr%Nb(2, 1)
  nx = 2 * (size(f, 2) / nz)
  call record_out_field(trim(fout), hdr, f, jlo, jhi, hdr%grid, nx, nz, hdr%yrow)
end program lh_io_check
