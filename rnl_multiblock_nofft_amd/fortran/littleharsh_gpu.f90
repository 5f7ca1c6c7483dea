! littleharsh_gpu -- the reference's main program (littleharsh.f90:1-228)
! with the RK3 step on one MI355X: read the initial fields in the record_out
! format, run `nsteps` RK3 steps on the device (the main loop,
! littleharsh.f90:172-228, as a replayed HIP graph), write the fields back
! in the same format.
!
!   littleharsh_gpu run.nml
!
! run.nml holds one namelist &run with the channel (N1, N2, ny, alp, bet,
! Re, grid_kind, dyq, ppp, gamma, bslpu1, bslpu3, flag_ctpress, mpgx, dt,
! precision), nsteps, and the input/output file names fin(1:4), fout(1:4)
! of u1, u2, u3, p.  With nstat > 0 a statistics sample is taken after every
! nstat steps (the reference samples at the first substep of every nstat-th
! step, FOU3D.f90:58-63) and written to fstats at the end (write_stats,
! stats.f90:254-441).  With fhist set, the history file gets a record
! (record_out's, FOU3D.f90:3040-3070) at the start and after every nstat
! steps (or at the end).
program littleharsh_gpu
  use iso_c_binding
  use lhgpu
  use lh_host
  implicit none
  integer :: N1 = 32, N2 = 32, ny = 153, grid_kind = 1, ppp = 5, flag_ctpress = 0
  integer :: precision = 0, nsteps = 1, nstat = 0
  real(8) :: alp = 1d0, bet = 2d0, Re = 4200d0, dyq = 10d0, gamma = 2d0
  real(8) :: bslpu1 = 0d0, bslpu3 = 0d0, mpgx = 0d0, dt = 1d-3
  character(len=1024) :: fin(4) = '', fout(4) = '', fstats = '', fhist = '', nmlfile
  namelist /run/ N1, N2, ny, alp, bet, Re, grid_kind, dyq, ppp, gamma, bslpu1, bslpu3, &
                 flag_ctpress, mpgx, dt, precision, nsteps, fin, fout, nstat, fstats, fhist
  type(lh_config) :: cfg(1)
  type(field_header) :: hdr(4), shdr
  complex(8), allocatable :: f(:,:)
  real(8), allocatable :: yu(:), yv(:), du(:), dv(:), yrow(:)
  real(8) :: sc(8)
  integer :: u, q, jlo, jhi, ncol, grid, done, ngal_arr(4, 3), uh
  integer, parameter :: ids(4) = [LH_U1, LH_U2, LH_U3, LH_P]
  integer :: elo(4), ehi(4)

  call get_command_argument(1, nmlfile)
  open(newunit=u, file=trim(nmlfile), status='old', action='read')
  read(u, nml=run)
  close(u)

  cfg(1) = lh_config(N1, N2, 3 * N1 / 2, 3 * N2 / 2, ny, alp, bet, Re, grid_kind, dyq, ppp, &
                     gamma, bslpu1, bslpu3, flag_ctpress, mpgx, dt, precision, 0, 0.5d0)
  call lh_host_setup(cfg)
  ncol = (N1 / 2 + 1) * N2
  elo = [0, 0, 0, 1]
  ehi = [ny, ny - 1, ny, ny - 1]

  ! read_in (start.f90:2444): fields go straight to the device
  do q = 1, 4
    call read_in_field(trim(fin(q)), hdr(q), f, jlo, jhi)
    if (jlo /= elo(q) .or. jhi /= ehi(q) .or. size(f, 2) /= ncol) then
      write(*, '(a,a,a,4i6)') 'littleharsh_gpu: ', trim(fin(q)), ' has rows/columns ', &
          jlo, jhi, size(f, 2), ncol
      error stop 1
    end if
    call lh_check(lh_field_upload(lhctx, 0_c_int, int(ids(q), c_int), f), 'lh_field_upload')
    deallocate(f)
  end do
  call lh_check(lh_set_scalars(lhctx, 0_c_int, hdr(1)%t, mpgx, 0d0, int(hdr(1)%iter, c_int)), &
                'lh_set_scalars')
  call lh_check(lh_init_flowrate(lhctx), 'lh_init_flowrate')

  ! main loop (littleharsh.f90:172-228)
  if (len_trim(fhist) > 0) then
    call hist_open(trim(fhist), hdr(1)%t, Re, alp, bet, mpgx, hdr(1)%Nb, uh)
    call hist_write(uh, flag_ctpress, 1)
  end if
  if (nstat > 0) then
    call lh_check(lh_stats_reset(lhctx), 'lh_stats_reset')
    done = 0
    do while (done + nstat <= nsteps)
      call lh_check(lh_step(lhctx, int(nstat, c_int)), 'lh_step')
      call lh_check(lh_stats_accumulate(lhctx), 'lh_stats_accumulate')
      if (len_trim(fhist) > 0) call hist_write(uh, flag_ctpress, 1)
      done = done + nstat
    end do
    if (done < nsteps) call lh_check(lh_step(lhctx, int(nsteps - done, c_int)), 'lh_step')
  else
    call lh_check(lh_step(lhctx, int(nsteps, c_int)), 'lh_step')
    if (len_trim(fhist) > 0) call hist_write(uh, flag_ctpress, 1)
  end if
  if (len_trim(fhist) > 0) close(uh)
  call lh_check(lh_sync(lhctx), 'lh_sync')
  call lh_check(lh_get_scalars(lhctx, 0_c_int, sc), 'lh_get_scalars')
  if (nstat > 0 .and. len_trim(fstats) > 0) then
    ! Ngal(4, 0:nband+1) of the one band that holds the whole channel
    ngal_arr = 0
    ngal_arr(1:2, 2) = [3 * N1 / 2, 3 * N2 / 2]
    ngal_arr(3:4, 2) = [ny - 2, ny - 1]
    shdr%t = sc(1); shdr%iter = int(sc(8)); shdr%mpgx = sc(3)
    shdr%Re = Re; shdr%alp = alp; shdr%bet = bet; shdr%nband = 1
    call write_stats_file(trim(fstats), shdr, ngal_arr, 1)
  end if
  write(*, '(a,i8,a,es14.6,a,es22.14,a,es22.14)') 'iter', int(sc(8)), '  t', sc(1), &
      '  Qx', sc(5), '  mpgx', sc(3)

  ! record_out (FOU3D.f90:2856)
  allocate(yu(0:ny), yv(0:ny - 1), du(0:ny), dv(0:ny - 1))
  call lh_check(lh_get_grid(lhctx, 0_c_int, yu, yv, du, dv), 'lh_get_grid')
  do q = 1, 4
    allocate(f(elo(q):ehi(q), ncol), yrow(elo(q):ehi(q)))
    call lh_check(lh_field_download(lhctx, 0_c_int, int(ids(q), c_int), f), 'lh_field_download')
    if (q == 2) then
      yrow = yv(elo(q):ehi(q)); grid = 2
    else
      yrow = yu(elo(q):ehi(q)); grid = 1
    end if
    hdr(q)%t = sc(1); hdr(q)%iter = int(sc(8)); hdr(q)%mpgx = sc(3)
    call record_out_field(trim(fout(q)), hdr(q), f, elo(q), ehi(q), grid, N1 + 2, N2, yrow)
    deallocate(f, yrow)
  end do
  call lh_host_finish()
end program littleharsh_gpu
