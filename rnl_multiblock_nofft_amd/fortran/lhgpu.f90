! lhgpu.f90 -- ISO_C_BINDING interface to liblhgpu.so (include/lhgpu.h).
!
! The thin layer between the Fortran host and the MI355X kernels: one
! interface block per C entry point, the lh_config record, the field and
! error constants, and lh_check, which turns a non-zero status into
! `error stop` with the library's message (the reference stops on errors,
! e.g. littleharsh.f90:85).
module lhgpu
  use iso_c_binding
  implicit none
  private :: c_strlen

  integer(c_int), parameter :: LH_OK = 0, LH_EINVAL = -1, LH_EHIP = -2, &
                               LH_ENOMEM = -3, LH_ESTATE = -4
  integer(c_int), parameter :: LH_PREC_FP64 = 0, LH_PREC_MIXED = 1
  integer(c_int), parameter :: LH_U1 = 0, LH_U2 = 1, LH_U3 = 2, LH_P = 3, &
                               LH_NU1 = 4, LH_NU2 = 5, LH_NU3 = 6, LH_R1 = 7, &
                               LH_R2 = 8, LH_R3 = 9, LH_PSI = 10, LH_DIV = 11
  integer(c_int), parameter :: LH_PHASE_NONLINEAR = 1, LH_PHASE_RHS = 2, &
                               LH_PHASE_SOLVE = 3
  integer(c_int), parameter :: LH_TUNE_SOLVE_COLS = 1, LH_TUNE_ZINV_POINTS = 2, &
                               LH_TUNE_XPROD_POINTS = 3, LH_TUNE_ZFWD_POINTS = 4, &
                               LH_TUNE_TEX_SEPARABLE = 5, LH_TUNE_TEX_FUSED = 15, &
                               LH_TUNE_SHAPE_GROUPS = 16
  ! statistics profiles in write_stats' order (stats.f90:356-368), spectra
  integer(c_int), parameter :: LH_ST_UM = 0, LH_ST_U2M = 1, LH_ST_VM = 2, LH_ST_V2M = 3, &
                               LH_ST_WM = 4, LH_ST_W2M = 5, LH_ST_PM = 6, LH_ST_P2M = 7, &
                               LH_ST_UVM = 8, LH_ST_UWM = 9, LH_ST_VWM = 10, LH_ST_WXM = 11, &
                               LH_ST_WX2M = 12, LH_ST_N = 13
  integer(c_int), parameter :: LH_SP_U = 0, LH_SP_V = 1, LH_SP_W = 2, LH_SP_UV = 3, &
                               LH_SP_P = 4, LH_SP_N = 5
  integer(c_int), parameter :: LH_CL_NIN = 135, LH_CL_NHID = 128, LH_CL_NOUT = 3, &
                               LH_CL_NPARAM = 50819

  type, bind(C) :: lh_config
    integer(c_int) :: N1, N2, Ngal1, Ngal2, ny
    real(c_double) :: alp, bet, Re
    integer(c_int) :: grid_kind
    real(c_double) :: dyq
    integer(c_int) :: ppp
    real(c_double) :: gamma, bslpu1, bslpu3
    integer(c_int) :: flag_ctpress
    real(c_double) :: mpgx, dt
    integer(c_int) :: precision
    integer(c_int) :: adaptive_dt
    real(c_double) :: CFL
  end type lh_config

  interface
    function lh_last_error() bind(C, name='lh_last_error') result(p)
      import :: c_ptr
      type(c_ptr) :: p
    end function
    function lh_ctx_create(cfgs, nchan, ctx) bind(C, name='lh_ctx_create') result(rc)
      import :: lh_config, c_int, c_ptr
      type(lh_config), intent(in) :: cfgs(*)
      integer(c_int), value :: nchan
      type(c_ptr), intent(out) :: ctx
      integer(c_int) :: rc
    end function
    function lh_ctx_destroy(ctx) bind(C, name='lh_ctx_destroy') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function
    function lh_set_grid(ctx, ch, yu, yv, dthdyu, dthdyv) bind(C, name='lh_set_grid') result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: ch
      real(c_double), intent(in) :: yu(*), yv(*), dthdyu(*), dthdyv(*)
      integer(c_int) :: rc
    end function
    ! multi-resolution bands: nb of ygrid's band split, and the link of a
    ! banded channel's three channels (bottom wall band, middle, top)
    function lh_band_rows(mid, band_h, nb) bind(C, name='lh_band_rows') result(rc)
      import :: lh_config, c_int, c_double
      type(lh_config), intent(in) :: mid
      real(c_double), value :: band_h
      integer(c_int), intent(out) :: nb
      integer(c_int) :: rc
    end function
    function lh_set_bands(ctx, ch_bottom, ch_middle, ch_top, nb) bind(C, name='lh_set_bands') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), value :: ch_bottom, ch_middle, ch_top, nb
      integer(c_int) :: rc
    end function
    function lh_get_grid(ctx, ch, yu, yv, dthdyu, dthdyv) bind(C, name='lh_get_grid') result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: ch
      real(c_double), intent(out) :: yu(*), yv(*), dthdyu(*), dthdyv(*)
      integer(c_int) :: rc
    end function
    function lh_field_upload(ctx, ch, field, reim) bind(C, name='lh_field_upload') result(rc)
      import :: c_ptr, c_int, c_double_complex
      type(c_ptr), value :: ctx
      integer(c_int), value :: ch, field
      complex(c_double_complex), intent(in) :: reim(*)
      integer(c_int) :: rc
    end function
    function lh_field_download(ctx, ch, field, reim) bind(C, name='lh_field_download') result(rc)
      import :: c_ptr, c_int, c_double_complex
      type(c_ptr), value :: ctx
      integer(c_int), value :: ch, field
      complex(c_double_complex), intent(out) :: reim(*)
      integer(c_int) :: rc
    end function
    function lh_get_scalars(ctx, ch, out8) bind(C, name='lh_get_scalars') result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: ch
      real(c_double), intent(out) :: out8(8)
      integer(c_int) :: rc
    end function
    function lh_set_scalars(ctx, ch, t, mpgx, QxT, iter) bind(C, name='lh_set_scalars') result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: ch, iter
      real(c_double), value :: t, mpgx, QxT
      integer(c_int) :: rc
    end function
    function lh_init_flowrate(ctx) bind(C, name='lh_init_flowrate') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function
    function lh_substep(ctx, kRK) bind(C, name='lh_substep') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), value :: kRK
      integer(c_int) :: rc
    end function
    function lh_step(ctx, nsteps) bind(C, name='lh_step') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), value :: nsteps
      integer(c_int) :: rc
    end function
    function lh_sync(ctx) bind(C, name='lh_sync') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function
    function lh_time_kernels(ctx, nsub, ms_out6) bind(C, name='lh_time_kernels') result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nsub
      real(c_double), intent(out) :: ms_out6(6)
      integer(c_int) :: rc
    end function
    function lh_run_phase(ctx, phase, kRK) bind(C, name='lh_run_phase') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), value :: phase, kRK
      integer(c_int) :: rc
    end function
    function lh_four_to_phys_planes(N1, Ngal1, Ngal2, d_planes, nplanes, stream) &
        bind(C, name='lh_four_to_phys_planes') result(rc)
      import :: c_ptr, c_int
      integer(c_int), value :: N1, Ngal1, Ngal2, nplanes
      type(c_ptr), value :: d_planes, stream
      integer(c_int) :: rc
    end function
    function lh_phys_to_four_planes(N1, Ngal1, Ngal2, d_planes, nplanes, stream) &
        bind(C, name='lh_phys_to_four_planes') result(rc)
      import :: c_ptr, c_int
      integer(c_int), value :: N1, Ngal1, Ngal2, nplanes
      type(c_ptr), value :: d_planes, stream
      integer(c_int) :: rc
    end function
    function lh_plane_transform_noalloc(dir, N1, Ngal1, Ngal2, d_planes, nplanes, stream) &
        bind(C, name='lh_plane_transform_noalloc') result(rc)
      import :: c_ptr, c_int
      integer(c_int), value :: dir, N1, Ngal1, Ngal2, nplanes
      type(c_ptr), value :: d_planes, stream
      integer(c_int) :: rc
    end function
    function lh_cft(d_c, ise, isa, m, iopt, n, stream) bind(C, name='lh_cft') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: d_c, stream
      integer(c_int), value :: ise, isa, m, iopt, n
      integer(c_int) :: rc
    end function
    function lh_rft(d_c, isa, m, iopt, n, stream) bind(C, name='lh_rft') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: d_c, stream
      integer(c_int), value :: isa, m, iopt, n
      integer(c_int) :: rc
    end function
    function lh_lusol_uw(ctx, ch, kRK, d_rhsu, d_rhsw, d_u, d_w) bind(C, name='lh_lusol_uw') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx, d_rhsu, d_rhsw, d_u, d_w
      integer(c_int), value :: ch, kRK
      integer(c_int) :: rc
    end function
    function lh_lusol_v(ctx, ch, kRK, d_rhsv, d_v) bind(C, name='lh_lusol_v') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx, d_rhsv, d_v
      integer(c_int), value :: ch, kRK
      integer(c_int) :: rc
    end function
    function lh_dev_alloc(bytes, p) bind(C, name='lh_dev_alloc') result(rc)
      import :: c_ptr, c_int, c_size_t
      integer(c_size_t), value :: bytes
      type(c_ptr), intent(out) :: p
      integer(c_int) :: rc
    end function
    function lh_dev_free(p) bind(C, name='lh_dev_free') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: p
      integer(c_int) :: rc
    end function
    function lh_copy_to_device(d, h, bytes) bind(C, name='lh_copy_to_device') result(rc)
      import :: c_ptr, c_int, c_size_t
      type(c_ptr), value :: d, h
      integer(c_size_t), value :: bytes
      integer(c_int) :: rc
    end function
    function lh_copy_to_host(h, d, bytes) bind(C, name='lh_copy_to_host') result(rc)
      import :: c_ptr, c_int, c_size_t
      type(c_ptr), value :: h, d
      integer(c_size_t), value :: bytes
      integer(c_int) :: rc
    end function
    function lh_set_tuning(ctx, key, val) bind(C, name='lh_set_tuning') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), value :: key, val
      integer(c_int) :: rc
    end function
    function lh_get_tuning(ctx, key, val) bind(C, name='lh_get_tuning') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), value :: key
      integer(c_int), intent(out) :: val
      integer(c_int) :: rc
    end function
    function lh_autotune(ctx, reps) bind(C, name='lh_autotune') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), value :: reps
      integer(c_int) :: rc
    end function
    function lh_set_texture(ctx, ch, bottom, top) bind(C, name='lh_set_texture') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx, bottom, top
      integer(c_int), value :: ch
      integer(c_int) :: rc
    end function
    function lh_slip_stats_get(ctx, ch, wall, maps) bind(C, name='lh_slip_stats_get') result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: ch, wall
      real(c_double), intent(out) :: maps(*)
      integer(c_int) :: rc
    end function
    function lh_history(ctx, ch, out12) bind(C, name='lh_history') result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: ch
      real(c_double), intent(out) :: out12(12)
      integer(c_int) :: rc
    end function
    function lh_set_triads(ctx, ch, nfile, lens, rows, map_u, map_v) bind(C, name='lh_set_triads') result(rc)
      import :: c_ptr, c_int, c_signed_char
      type(c_ptr), value :: ctx
      integer(c_int), value :: ch, nfile
      integer(c_int), intent(in) :: lens(*), map_u(*), map_v(*)
      integer(c_signed_char), intent(in) :: rows(*)
      integer(c_int) :: rc
    end function
    function lh_lusol_p(ctx, ch, x) bind(C, name='lh_lusol_p') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx, x
      integer(c_int), value :: ch
      integer(c_int) :: rc
    end function
    function lh_lusol0(ctx, ch, kRK, x) bind(C, name='lh_lusol0') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx, x
      integer(c_int), value :: ch, kRK
      integer(c_int) :: rc
    end function
    function lh_four_to_phys_N_planes(N1, N2, planes, nplanes, stream) &
        bind(C, name='lh_four_to_phys_N_planes') result(rc)
      import :: c_ptr, c_int
      integer(c_int), value :: N1, N2, nplanes
      type(c_ptr), value :: planes, stream
      integer(c_int) :: rc
    end function
    function lh_phys_to_four_N_planes(N1, N2, planes, nplanes, stream) &
        bind(C, name='lh_phys_to_four_N_planes') result(rc)
      import :: c_ptr, c_int
      integer(c_int), value :: N1, N2, nplanes
      type(c_ptr), value :: planes, stream
      integer(c_int) :: rc
    end function
    function lh_modes_to_planes(N1, N2, Ngal1, Ngal2, cols, nrows, planes, stream) &
        bind(C, name='lh_modes_to_planes') result(rc)
      import :: c_ptr, c_int
      integer(c_int), value :: N1, N2, Ngal1, Ngal2, nrows
      type(c_ptr), value :: cols, planes, stream
      integer(c_int) :: rc
    end function
    function lh_planes_to_modes(N1, N2, Ngal1, Ngal2, planes, nrows, cols, stream) &
        bind(C, name='lh_planes_to_modes') result(rc)
      import :: c_ptr, c_int
      integer(c_int), value :: N1, N2, Ngal1, Ngal2, nrows
      type(c_ptr), value :: planes, cols, stream
      integer(c_int) :: rc
    end function
    function lh_stats_accumulate(ctx) bind(C, name='lh_stats_accumulate') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function
    function lh_stats_reset(ctx) bind(C, name='lh_stats_reset') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function
    ! prof(0:ny, 0:LH_ST_N-1): profile q in prof(:, q), rows by the
    ! reference's row index
    function lh_stats_get(ctx, ch, prof, istat) bind(C, name='lh_stats_get') result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: ch
      real(c_double), intent(out) :: prof(*)
      integer(c_int), intent(out) :: istat
      integer(c_int) :: rc
    end function
    ! sp(jlo:jhi, ncol), the layout of spU..spP (spectra.f90)
    function lh_spectrum_get(ctx, ch, which, sp) bind(C, name='lh_spectrum_get') result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: ch, which
      real(c_double), intent(out) :: sp(*)
      integer(c_int) :: rc
    end function
    ! learned closure of the nonlinear term (no counterpart in the reference):
    ! params LH_CL_NPARAM reals W1 b1 W2 b2 W3 b3 W4 b4, norm9 = mean(3),
    ! inv_scale(3), out_scale(3); device buffers as type(c_ptr)
    function lh_closure_set(ctx, params, norm9) bind(C, name='lh_closure_set') result(rc)
      import :: c_ptr, c_int, c_float
      type(c_ptr), value :: ctx
      real(c_float), intent(in) :: params(*), norm9(9)
      integer(c_int) :: rc
    end function
    function lh_closure_enable(ctx, ch, on) bind(C, name='lh_closure_enable') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), value :: ch, on
      integer(c_int) :: rc
    end function
    function lh_closure_eval(ctx, ch, d_phys, d_corr) bind(C, name='lh_closure_eval') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx, d_phys, d_corr
      integer(c_int), value :: ch
      integer(c_int) :: rc
    end function
    function lh_closure_sample(fine, chf, coarse, chc, d_phys, d_tau) bind(C, name='lh_closure_sample') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: fine, coarse, d_phys, d_tau
      integer(c_int), value :: chf, chc
      integer(c_int) :: rc
    end function
    function lh_closure_time(ctx, nrep, ms3) bind(C, name='lh_closure_time') result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nrep
      real(c_double), intent(out) :: ms3(3)
      integer(c_int) :: rc
    end function
    function lh_mlp_create(batch, m) bind(C, name='lh_mlp_create') result(rc)
      import :: c_ptr, c_int
      integer(c_int), value :: batch
      type(c_ptr), intent(out) :: m
      integer(c_int) :: rc
    end function
    function lh_mlp_destroy(m) bind(C, name='lh_mlp_destroy') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: m
      integer(c_int) :: rc
    end function
    function lh_mlp_set_params(m, params, norm9) bind(C, name='lh_mlp_set_params') result(rc)
      import :: c_ptr, c_int, c_float
      type(c_ptr), value :: m
      real(c_float), intent(in) :: params(*), norm9(9)
      integer(c_int) :: rc
    end function
    function lh_mlp_get_params(m, params) bind(C, name='lh_mlp_get_params') result(rc)
      import :: c_ptr, c_int, c_float
      type(c_ptr), value :: m
      real(c_float), intent(out) :: params(*)
      integer(c_int) :: rc
    end function
    function lh_mlp_set_batch(m, x, t) bind(C, name='lh_mlp_set_batch') result(rc)
      import :: c_ptr, c_int, c_float
      type(c_ptr), value :: m
      real(c_float), intent(in) :: x(*), t(*)
      integer(c_int) :: rc
    end function
    function lh_mlp_sample(m, N1, N2, ny, d_phys, d_tau, nsnap, seed) bind(C, name='lh_mlp_sample') result(rc)
      import :: c_ptr, c_int, c_int64_t
      type(c_ptr), value :: m, d_phys, d_tau
      integer(c_int), value :: N1, N2, ny, nsnap
      integer(c_int64_t), value :: seed
      integer(c_int) :: rc
    end function
    function lh_mlp_set_points(m, N1, N2, ny, d_phys, pts) bind(C, name='lh_mlp_set_points') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: m, d_phys
      integer(c_int), value :: N1, N2, ny
      integer(c_int), intent(in) :: pts(*)
      integer(c_int) :: rc
    end function
    function lh_mlp_forward(m, y) bind(C, name='lh_mlp_forward') result(rc)
      import :: c_ptr, c_int, c_float
      type(c_ptr), value :: m
      real(c_float), intent(out) :: y(*)
      integer(c_int) :: rc
    end function
    function lh_mlp_grad(m, loss) bind(C, name='lh_mlp_grad') result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: m
      real(c_double), intent(out) :: loss
      integer(c_int) :: rc
    end function
    function lh_mlp_get_grad(m, grad) bind(C, name='lh_mlp_get_grad') result(rc)
      import :: c_ptr, c_int, c_float
      type(c_ptr), value :: m
      real(c_float), intent(out) :: grad(*)
      integer(c_int) :: rc
    end function
    function lh_mlp_adam(m, lr, beta1, beta2, eps, wd) bind(C, name='lh_mlp_adam') result(rc)
      import :: c_ptr, c_int, c_float
      type(c_ptr), value :: m
      real(c_float), value :: lr, beta1, beta2, eps, wd
      integer(c_int) :: rc
    end function
    function lh_mlp_sync(m) bind(C, name='lh_mlp_sync') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: m
      integer(c_int) :: rc
    end function
    function c_strlen(s) bind(C, name='strlen') result(n)
      import :: c_ptr, c_size_t
      type(c_ptr), value :: s
      integer(c_size_t) :: n
    end function
  end interface

contains

  ! stop with the library's message when rc is not LH_OK
  subroutine lh_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    character(len=:), allocatable :: text
    type(c_ptr) :: p
    integer :: n, q
    if (rc == LH_OK) return
    p = lh_last_error()
    n = int(c_strlen(p))
    call c_f_pointer(p, msg, [n])
    allocate(character(len=n) :: text)
    do q = 1, n
      text(q:q) = msg(q)
    end do
    write(*, '(a,a,a,i0,a,a)') 'liblhgpu: ', where, ' failed (', rc, '): ', text
    error stop 1
  end subroutine lh_check

end module lhgpu
