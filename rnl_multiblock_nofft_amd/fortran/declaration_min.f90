! declaration_min.f90 -- the part of the host program's module `declaration`
! (the reference's declaration.f90:1-228) that the drop-in routines of
! lh_dropin.f90 read, so that they build and are tested in this repository.
! A host that has its own `declaration` module compiles lh_dropin.f90
! against that one instead: these are the names, kinds and ranks the
! drop-in expects it to provide (band constants declaration.f90:8-17,
! index maps :24-37, grid sizes and RK state, the field types :137-139).
module declaration
  implicit none
  integer, parameter :: nband = 3
  integer, parameter :: sband = 1
  integer, parameter :: eband = nband
  integer, parameter :: botband = 1
  integer, parameter :: midband = (nband + 1) / 2
  integer, parameter :: topband = nband
  integer, parameter :: ugrid = 2   ! cell centres with ghost points: u, w
  integer, parameter :: vgrid = 1   ! faces: v
  integer, parameter :: pgrid = 3   ! cell centres without ghost points: p

  integer np
  integer, pointer :: N(:,:), Ngal(:,:)            ! N(1:4, 0:nband+1), Ngal(1:2, 1:nband)
  integer, pointer :: jlim(:,:,:)                  ! jlim(bottom/top, grid, band)
  integer jgal(3,2), igal, kgal                    ! this rank's planes and their shape
  integer, pointer :: columns_i(:,:,:), columns_k(:,:,:), columns_num(:,:)
  integer, pointer :: planeBC(:,:)                 ! planeBC(0:N1+2, N2): 1 no-slip, 2 free-shear
  integer geometry_type                            ! 0 smooth, 1 textured walls, 2 roughness
  integer kRK
  real(8) dt, Re

  type cfield
    complex(8), pointer :: f(:,:)
  end type cfield

  type rfield_dg
    real(8), pointer :: f_dg(:,:,:)
  end type rfield_dg
end module declaration
