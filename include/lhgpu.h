/*
 * lhgpu.h -- C-ABI of liblhgpu.so, the MI355X (gfx950) implementation of the
 * plain-FFT RK3 time step of LittleHarsh (channel-flow DNS).
 *
 * Plain pointers and sizes only.  Host arrays are passed in the reference's
 * Fortran layout `complex(8) f(jlim1:jlim2, ncol)` (y contiguous, then the
 * spectral column), which is what the Fortran shims hand over with c_loc.
 * Column c holds x-mode i = c % (N1/2+1) and z-index k = c / (N1/2+1), the
 * order proc_lims_columns builds for equal band resolutions
 * (start.f90:1574-1600).  Row ranges:
 *   LH_U1, LH_U3:   j = 0..ny     (ny+1 rows incl. ghost cells)
 *   LH_U2:          j = 0..ny-1   (v faces incl. walls)
 *   LH_P:           j = 1..ny-1   (pressure centres)
 * where ny is the number of v points wall to wall (input.in `ny`).
 *
 * Every entry point returns 0 on success and a negative LH_E* code on error;
 * lh_last_error() gives a message.  A context is single-threaded: calls on
 * one context must not overlap.
 */
#ifndef LHGPU_H
#define LHGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LH_OK 0
#define LH_EINVAL (-1)   /* invalid argument / shape */
#define LH_EHIP (-2)     /* HIP runtime error */
#define LH_ENOMEM (-3)   /* device allocation failed */
#define LH_ESTATE (-4)   /* call not valid in the context's state */

typedef struct lh_config {
  int N1, N2;          /* Fourier modes in x, z (N = 2*(Ngal/3)); even */
  int Ngal1, Ngal2;    /* physical (3/2-padded) grid; 3*2^k or 2^k */
  int ny;              /* v points wall to wall (input.in `ny`) */
  double alp, bet;     /* Lx = 2pi/alp, Lz = 2pi/bet */
  double Re;           /* 1/nu as the reference uses it (C1 = aRK/Re) */
  int grid_kind;       /* 0: polynomial y = a(j-q)+b(j-q)^p (ygrid); 1: tanh */
  double dyq;          /* polynomial grid: dymax/dymin */
  int ppp;             /* polynomial grid exponent */
  double gamma;        /* tanh grid: y = tanh(gamma(2j/(ny-1)-1))/tanh(gamma) */
  double bslpu1, bslpu3; /* Navier slip lengths of the ghost update */
  int flag_ctpress;    /* 0: constant flow rate; 1: constant pressure gradient */
  double mpgx;         /* mean pressure gradient (initial / prescribed) */
  double dt;           /* time step (the initial one when adaptive_dt) */
  int precision;       /* LH_PREC_FP64 or LH_PREC_MIXED */
  int adaptive_dt;     /* 0: dt fixed; 1: dt = min(2.5 dtv, CFL dtc) at every
                          first substep (FOU3D.f90:288-297, dtc_calc :699-736) */
  double CFL;          /* Courant number of the adaptive step */
} lh_config;

#define LH_PREC_FP64 0   /* everything in FP64 */
#define LH_PREC_MIXED 1  /* FFTs/products in FP32 (FP64 twiddles), solves FP64 */

/* field ids for upload/download */
#define LH_U1 0
#define LH_U2 1
#define LH_U3 2
#define LH_P 3
/* internal arrays, downloadable for diagnostics (rows as in the oracle):
 * LH_NU1/LH_NU3 rows 1..ny-1, LH_NU2 rows 1..ny-2, LH_DIV/LH_PSI rows 1..ny-1 */
#define LH_NU1 4
#define LH_NU2 5
#define LH_NU3 6
#define LH_R1 7    /* rhs u1 + dt*du1, rows 0..ny (ghost rows 0) */
#define LH_R2 8    /* rhs u2 + dt*du2, rows 0..ny-1 */
#define LH_R3 9
#define LH_PSI 10
#define LH_DIV 11

typedef struct lh_ctx lh_ctx;

const char *lh_last_error(void);

/* Create `nchan` channels (ensemble members) in one context.  cfgs[c] may
 * differ in every field except `precision`.  Replaces start()'s setup
 * (start.f90:36-691): y-grid (ygrid :693), wavenumbers (def_k :1079), RK3
 * coefficients (:419-433), grid weights (:537-559) and the Poisson operator
 * (LU_buildP, tridLU_3D.f90:333).  Fields start at zero. */
int lh_ctx_create(const lh_config *cfgs, int nchan, lh_ctx **out);
int lh_ctx_destroy(lh_ctx *ctx);

/* Replace channel `ch`'s y-grid (yu[0..ny], yv[0..ny-1], dthdyu[0..ny],
 * dthdyv[0..ny-1]) as the Fortran host's ygrid computed it; derived
 * coefficients (dyu2i, dyv2i, grid weights) are rebuilt. */
int lh_set_grid(lh_ctx *ctx, int ch, const double *yu, const double *yv,
                const double *dthdyu, const double *dthdyv);
int lh_get_grid(lh_ctx *ctx, int ch, double *yu, double *yv, double *dthdyu,
                double *dthdyv);

/* Host <-> device copies of one field of channel `ch`, reference layout,
 * re/im interleaved.  The device stores fields mode-major (each y row holds
 * all columns contiguously); the on-device LDS-tiled transpose replaces the
 * plane/column exchanges modes_to_planes_UVP / planes_to_modes_UVP
 * (FOU3D.f90:1000-1098, 2295-2442). */
int lh_field_upload(lh_ctx *ctx, int ch, int field, const double *reim);
int lh_field_download(lh_ctx *ctx, int ch, int field, double *reim);

/* scalars: out[0..7] = t, dt, mpgx, QxT, Qx, dtv, dtc, iter; dtc is the CFL
 * step of the last first substep (dtc_calc, FOU3D.f90:699-736), 0 before */
int lh_get_scalars(lh_ctx *ctx, int ch, double *out8);
int lh_set_scalars(lh_ctx *ctx, int ch, double t, double mpgx, double QxT, int iter);
/* The history record record_out writes (FOU3D.f90:3040-3070), reduced on the
 * device: out[0..11] = t, dtv, dtc, dt, err (max |div| over the pressure
 * rows, `error` littleharsh.f90:342-375), Qx (flowrateIm of the mean mode),
 * QxT, mpgx, dgx (the last flow-rate correction g, flowrate_corr
 * littleharsh.f90:476-509), Umax (maxvel, littleharsh.f90:511-527), Uslip
 * (the mean u1 interpolated to y = -1), iter. */
int lh_history(lh_ctx *ctx, int ch, double *out12);
/* QxT := flow rate of the current u1 mean mode (getini, start.f90:2166-2171;
 * flowrateIm littleharsh.f90:435) */
int lh_init_flowrate(lh_ctx *ctx);

/* One RK3 substep kRK in 1..3 for every channel (littleharsh.f90:176-215):
 * RHS0_u1/u2/u3, nonlinear (FOU3D.f90:41, with ops_in_planes2's plain FFT
 * products FOU3D.f90:2090-2251), solveU_W, solveV, meanflow_ctU, solveP,
 * v_corr. */
int lh_substep(lh_ctx *ctx, int kRK);
/* nsteps full RK3 steps (main loop littleharsh.f90:172-228), replayed from a
 * captured HIP graph after the first step. */
int lh_step(lh_ctx *ctx, int nsteps);
int lh_sync(lh_ctx *ctx);
/* Measurement: run `nsub` substeps (kRK cycling 1,2,3 from kRK = 1) with the
 * kernels of a substep launched one by one between HIP events on the
 * context's stream; ms_out[0..5] receive the summed milliseconds of the
 * z-inverse, x-product, z-forward, right-hand-side and column-solve kernels
 * and of the textured-wall kernels (0 without textures).  Advances the state
 * like lh_substep.  No counterpart in the reference. */
int lh_time_kernels(lh_ctx *ctx, int nsub, double *ms_out6);

/* Phases of a substep, for diagnostics and per-phase parity checks:
 * LH_PHASE_NONLINEAR: interp_u/v + modes_to_planes + ops_in_planes2 +
 *                     planes_to_modes (x/z part of N and the y-fluxes)
 * LH_PHASE_RHS:       der_y of the fluxes, RHS0_*, du update, rhs = u+dt*du
 * LH_PHASE_SOLVE:     solveU_W, solveV, meanflow_ctU, solveP, v_corr */
#define LH_PHASE_NONLINEAR 1
#define LH_PHASE_RHS 2
#define LH_PHASE_SOLVE 3
int lh_run_phase(lh_ctx *ctx, int phase, int kRK);

/* Stand-alone batched transforms on device planes, replacing four_to_phys_du
 * (FOU3D.f90:766) and phys_to_four_du (FOU3D.f90:2253) with their FFTPACK
 * conventions (cft_buff.f:1-65, rft_buff.f:24-90).  planes: nplanes device
 * planes of (Ngal1+2) x Ngal2 doubles, Fortran order, transformed in place. */
int lh_four_to_phys_planes(int N1, int Ngal1, int Ngal2, double *d_planes,
                           int nplanes, void *stream);
int lh_phys_to_four_planes(int N1, int Ngal1, int Ngal2, double *d_planes,
                           int nplanes, void *stream);

/* FFTPACK's vector transforms on device data, with the reference's
 * argument meaning: cft (cft_buff.f:1-65) transforms m complex vectors of
 * length n whose element j of vector k sits at d[k isa + j ise] (real) and
 * the next double (imaginary); rft (rft_buff.f:24-90) m real vectors of
 * length n at d + k isa (isa >= n + 2), packed on the Fourier side as
 * cc(k) = d[2k] + i d[2k+1], k = 0..n/2.  iopt < 0: forward, scaled by 1/n
 * (rft zeroes the imaginary parts of cc(0) and cc(n/2)); iopt >= 0:
 * inverse, unscaled (rft ignores those two imaginary parts and leaves
 * d[n], d[n+1] as they were).  n even, 2^a 3^b, at most 2048; the vectors
 * must not share elements.  n is what the reference keeps in aux(1) (cfti,
 * rfti). */
int lh_cft(double *d_c, int ise, int isa, int m, int iopt, int n, void *stream);
int lh_rft(double *d_c, int isa, int m, int iopt, int n, void *stream);

/* The two plane transforms above without any device allocation: dir = +1
 * is lh_four_to_phys_planes, dir = -1 lh_phys_to_four_planes, same
 * arguments and results.  No counterpart in the reference; it is for callers
 * that cannot allocate (code under HIP graph capture, tight loops, hosts that
 * forbid hipMalloc).  The twiddle tables live in a fixed 48 KiB device array
 * of the library, written on `stream` at every call: Ngal1, Ngal2 <= 1536
 * (LH_EINVAL beyond).  Calls that overlap on one GPU share that array and
 * race. */
int lh_plane_transform_noalloc(int dir, int N1, int Ngal1, int Ngal2, double *d_planes, int nplanes,
                               void *stream);

/* Stand-alone batched tridiagonal solves on device data in the reference
 * layout x(jlo:jhi, ncol) (re/im interleaved), replacing LUsolU_W / LUsolV
 * (tridLU_3D.f90:55,135) for one RK3 substep of channel `ch`: solves
 * (I - beta(D2 + k^2)) x = rhs with the reference's wall rows. */
int lh_lusol_uw(lh_ctx *ctx, int ch, int kRK, const double *d_rhsu,
                const double *d_rhsw, double *d_u, double *d_w);
int lh_lusol_v(lh_ctx *ctx, int ch, int kRK, const double *d_rhsv, double *d_v);
/* LUsolP (tridLU_3D.f90:192-226): the pressure Poisson system of LU_buildP
 * (tridLU_3D.f90:333-407, with the mean-mode pin of its first row) solved in
 * place on device data x(1:ny-1, ncol) of channel `ch`. */
int lh_lusol_p(lh_ctx *ctx, int ch, double *d_x);
/* LUsol0 (tridLU_3D.f90:228-254): the real mean-mode U system of LU_build0
 * (tridLU_3D.f90:409-435) for substep kRK and the current dt, in place on
 * device data x(0:ny) (real). */
int lh_lusol0(lh_ctx *ctx, int ch, int kRK, double *d_x);

/* N-grid plane transforms four_to_phys_N / phys_to_four_N (FOU3D.f90:
 * 786-804, 2274-2293): (N1+2) x N2 planes, no oddball-row zeroing. */
int lh_four_to_phys_N_planes(int N1, int N2, double *d_planes, int nplanes, void *stream);
int lh_phys_to_four_N_planes(int N1, int N2, double *d_planes, int nplanes, void *stream);

/* Column <-> plane transposes on device data, replacing
 * modes_to_planes_UVP (FOU3D.f90:1000-1098) and planes_to_modes_UVP /
 * planes_to_modes_NUVP (FOU3D.f90:2295-2442, 2772-2854) on one GPU (no MPI).
 * cols: x(jlo:jhi, ncol) complex (re/im interleaved), nrows = jhi-jlo+1;
 * planes: nrows planes of (Ngal1+2) x Ngal2 doubles, Fortran order, mode
 * (i, k) at (2i+1, krow) and (2i+2, krow), krow = k for k <= N2/2 and
 * k + Ngal2 - N2 above (1-based).  modes_to_planes zero-fills the rest of
 * each plane (the reference's callers zero the planes first).  Ngal = N
 * gives the N-grid planes of modes_to_planes_phys. */
int lh_modes_to_planes(int N1, int N2, int Ngal1, int Ngal2, const double *d_cols, int nrows,
                       double *d_planes, void *stream);
int lh_planes_to_modes(int N1, int N2, int Ngal1, int Ngal2, const double *d_planes, int nrows,
                       double *d_cols, void *stream);

/* Device memory for the device-pointer entry points above, for hosts without
 * HIP bindings of their own (the Fortran shims stage the reference's host
 * arrays through these).  No counterpart in the reference, whose arrays are
 * host memory.  Copies are synchronous. */
int lh_dev_alloc(size_t bytes, void **out);
int lh_dev_free(void *p);
int lh_copy_to_device(void *d, const void *h, size_t bytes);
int lh_copy_to_host(void *h, const void *d, size_t bytes);

/* Multi-resolution bands (the reference's nband = 3 blocks, declaration.f90:
 * 8-13; its MULTIBLOCK layout).  A banded channel is three channels of one
 * context: its middle band (the long pencils: that channel's N1 x N2 modes
 * over the whole height, ny v points) and its bottom and top wall bands (the
 * short pencils: the wall resolution's remaining modes, each over its own
 * rows, v rows 0..nb+1 and ny-2-nb..ny-1, created with ny = nb + 2,
 * flag_ctpress = 1, mpgx = 0, more or as many modes in x and z, the same
 * alp, bet, Re, dt and slip lengths).  lh_set_bands links them: the wall
 * bands take the middle band's grid rows (their band-edge weights follow,
 * start.f90:539-559), each band's column phases run over its own rows with
 * its own edge rows (RHS0, solveU_W, solveV, solveP, v_corr loop over
 * iband; LUsolU_W / LUsolV / LUsolP tridLU_3D.f90:55-226 with jlim,
 * start.f90:1677-1711), and every plane transforms at its band's resolution
 * holding every mode present at its row (modes_to_planes_UVP with dk,
 * FOU3D.f90:1000-1098; proc_lims_planes start.f90:1836-2142).  In a wall
 * band's arrays the long pencils' columns are not the band's: read the
 * middle band for those modes.  Fixed dt; no textures, triad lists or
 * closure on banded channels.  Replaces the band set-up of
 * proc_lims_columns/proc_lims_planes (start.f90:1490,1836).
 * lh_band_rows: nb = N(3,1) of ygrid's band split (start.f90:999-1011), the
 * first row j with yu(j) >= yu(0) + band_h on the grid of `mid`. */
int lh_band_rows(const lh_config *mid, double band_h, int *nb);
int lh_set_bands(lh_ctx *ctx, int ch_bottom, int ch_middle, int ch_top, int nb);

/* Tuning knobs of the launch shapes (no counterpart in the reference).
 * LH_TUNE_SOLVE_COLS: spectral columns per workgroup of the column solve;
 * 0 selects one lane per column streaming rows from HBM, 1..10 the form that
 * stages a workgroup's columns in LDS (must fit 160 KiB), -1 one column per
 * wave with its rows in the lanes' registers (ny < 384; every Thomas
 * recurrence a chunked scan over the wave, results agree to rounding).
 * LH_TUNE_{ZINV,XPROD,ZFWD}_POINTS: complex points per field that one
 * workgroup of the z-inverse, x-product and z-forward stages transforms
 * (whole sequences; the LDS they need must fit 160 KiB).
 * Changing a knob drops the captured step graph. */
#define LH_TUNE_SOLVE_COLS 1
#define LH_TUNE_ZINV_POINTS 2
#define LH_TUNE_XPROD_POINTS 3
#define LH_TUNE_ZFWD_POINTS 4
/* LH_TUNE_TEX_SEPARABLE (default 1): channels whose textured walls have
 * x-independent masks (streamwise ridges) iterate the wall sources per
 * x-mode with z transforms only; 0 forces the two-dimensional wall
 * transforms for every textured channel.  Results agree to rounding. */
#define LH_TUNE_TEX_SEPARABLE 5
/* LH_TUNE_SOLVE_SCAN (default 1): in the LDS form of the column solve, each
 * Thomas recurrence (forward elimination, back substitution) is split over
 * a group of lanes -- per-lane chunks composed as affine maps and scanned
 * across the group, then each chunk rerun from its true start value; 0 runs
 * every recurrence on one lane.  Results agree to rounding. */
#define LH_TUNE_SOLVE_SCAN 6
/* LH_TUNE_SOLVE_WAVES (default 1): waves per workgroup of the LDS column
 * solve (1, 2 or 4); more waves give each staged column more lanes for its
 * recurrences and more loads in flight.  The one-column-per-wave form runs
 * eight consecutive columns per workgroup, or two when this is 4 and the
 * columns have at most 192 rows (more workgroups for small channels). */
#define LH_TUNE_SOLVE_WAVES 7
/* LH_TUNE_XPROD_INPLACE (default 0): 1 runs the x-product stage's transforms
 * in place in one LDS buffer where every stage has at most one item per
 * thread (more workgroups per CU, more registers per thread).  Results are
 * identical. */
#define LH_TUNE_XPROD_INPLACE 8
/* LH_TUNE_ZFWD_INPLACE (default 0): the same for the z-forward stage. */
#define LH_TUNE_ZFWD_INPLACE 9
/* LH_TUNE_XPROD_THREADS (default 256): threads per workgroup of the
 * x-product stage, 128, 192 or 256. */
#define LH_TUNE_XPROD_THREADS 10
/* LH_TUNE_RHS_FUSED (default 1): the column solve of lh_step/lh_substep
 * computes the right-hand sides itself and r1, r2, r3 are not stored --
 * for contexts without textured walls in every solve form, with textured
 * walls in the one-column-per-wave form (LH_TUNE_SOLVE_COLS = -1), after a
 * pass that stores the near-wall rows of r1 and r3 the wall sources read.
 * (The LH_PHASE_RHS / LH_PHASE_SOLVE phases and lh_time_kernels keep the
 * separate right-hand-side kernel); 0 always runs that kernel.
 * lh_autotune sets it with the solve's shape, from the timings of both
 * forms. */
#define LH_TUNE_RHS_FUSED 11
/* LH_TUNE_GHOST_IN_ZINV (default 1): in lh_step's captured step, the
 * ghost-row pass of x-independent textured walls after substeps 1 and 2
 * runs as extra workgroups of the next substep's z-inverse transform (one
 * launch less per substep; the step's results are the same).  Channels with
 * N2 of 32 or 64 only; 0 keeps the separate launch. */
#define LH_TUNE_GHOST_IN_ZINV 12
/* LH_TUNE_ZINV_INPLACE (default 0): the same for the z-inverse stage; 2
 * runs it in place in workgroups of 512 threads (room for tiles of 8
 * x-modes at 384 points). */
#define LH_TUNE_ZINV_INPLACE 13
/* LH_TUNE_ZFWD_FIELDWISE (default 0): 1 or 2, the z-forward stage
 * transforms a plane's fields one after another in one LDS buffer, which
 * leaves room for tiles several x-modes wide (contiguous loads and stores),
 * in workgroups of 256 (1) or 512 (2) threads; used when every channel has
 * N2 = 2 Ngal2 / 3, the LH_TUNE_ZFWD_POINTS tile fits, and the channels
 * share one of the lengths 24, 48, 96, 192, 384 or all have lengths 24, 48
 * or 96; else the stage runs as before. */
#define LH_TUNE_ZFWD_FIELDWISE 14
/* LH_TUNE_TEX_FUSED (default 1): textured walls with x-varying masks run
 * each field's wall-source iteration in one workgroup with the wall planes
 * resident in LDS when they fit (160 KiB); 0 always runs the wall
 * transforms as separate launches (z transforms, then the x transforms and
 * the mask in one launch), as channels too large for LDS do.  Results agree
 * to rounding. */
#define LH_TUNE_TEX_FUSED 15
/* LH_TUNE_SHAPE_GROUPS (default 1): a context whose channels have several
 * plane shapes launches the nonlinear stages, the right-hand sides and the
 * solve once per group of channels of one shape (as banded contexts always
 * do): each launch runs its length's specialised code; 0 launches every
 * stage once over all channels.  Results are identical. */
#define LH_TUNE_SHAPE_GROUPS 16
int lh_set_tuning(lh_ctx *ctx, int key, int value);
int lh_get_tuning(lh_ctx *ctx, int key, int *value);
/* Autotuner: times each feasible setting of every knob on the context's own
 * channels (`reps` launches each, on the current state, which is saved
 * before and restored after) and keeps the fastest. */
int lh_autotune(lh_ctx *ctx, int reps);

/* Textured walls (configurations with superhydrophobic stripes or posts).
 * bottom / top: N1*N2 ints, index z*N1 + x on the N1 x N2 grid of the
 * channel's modes, holding the reference's planeBC codes (start.f90:
 * 1776-1829): 1 = no-slip, 2 = free-shear; NULL leaves that wall smooth.
 * The U/W Helmholtz solves then take the per-point wall rows of the
 * reference's textured LUsolU / LU_build (tridLU_3D.f90:753-1078,
 * 1284-1358; inactive there), solved exactly for the unfactored operator,
 * and the ghost rows of v_corr follow the mask point by point. */
int lh_set_texture(lh_ctx *ctx, int ch, const int *bottom, const int *top);

/* The restricted nonlinear path of the reference: products summed over
 * triad interaction lists in Fourier space (ops_in_planes, FOU3D.f90:
 * 1724-1987; nonlinInter :1989-2030) instead of the FFT products, for channel
 * `ch`.  Arguments as nonlinRead (start.f90:2983-3069) loads them: nfile
 * list files; lens[q*9 + t-1] the rows of list type t (1..9: uu uv uw vu vv
 * vw wu wv ww) in file q; rows the int8 lists concatenated in file, then
 * type order, each in the reference's list(len, 4) Fortran layout (all la,
 * ka, lp, kp); map_u[j-1] (u planes j = 1..ny-1) and map_v[j-1] (v planes
 * j = 1..ny-2) the file of each plane, -1 for none (the mean-mode terms
 * only).  Rows must satisfy |la|, |lp| <= N1/2-1, |ka|, |kp| <= N2/2-1,
 * 0 <= la+lp <= N1/2-1, |ka+kp| <= N2/2-1.  The channel must have a fixed
 * time step (the list path forms no physical planes for dtc_calc).
 * nfile = 0 returns the channel to the FFT path. */
int lh_set_triads(lh_ctx *ctx, int ch, int nfile, const int *lens, const signed char *rows,
                  const int *map_u, const int *map_v);

/* Running statistics and spectra, replacing stats (stats.f90:1-107) with
 * its gathers recvX (stats.f90:443-470) and spectra / buildsp / buildspUV
 * (spectra.f90:1-56).  lh_stats_accumulate adds one sample of the current
 * fields of every channel (the reference samples at the first substep of
 * every nstat-th step, FOU3D.f90:58-63, i.e. on the state between two
 * steps) and counts it (istat).  lh_stats_reset zeroes everything, as
 * write_stats / write_spect do after writing (stats.f90:409-440,
 * spectra.f90:101-105). */
int lh_stats_accumulate(lh_ctx *ctx);
int lh_stats_reset(lh_ctx *ctx);
/* Row profiles, prof[LH_ST_N][ny+1], profile q at prof[q*(ny+1) + j] for
 * the reference's row j; the order is write_stats' (stats.f90:356-368):
 * u-grid profiles (Um U2m Wm W2m UVm UWm) on rows 0..ny, v-grid profiles
 * (Vm V2m VWm wxm wx2m) on 0..ny-1, p profiles (Pm P2m) on 1..ny-1; other
 * entries are 0.  Sums over grid points, not yet divided by the sample or
 * point counts, as the reference writes them.  At the ghost / wall rows and
 * for p the reference's planes hold spectral coefficients, and the sums are
 * those of the stored real and imaginary parts, as there. */
#define LH_ST_UM 0
#define LH_ST_U2M 1
#define LH_ST_VM 2
#define LH_ST_V2M 3
#define LH_ST_WM 4
#define LH_ST_W2M 5
#define LH_ST_PM 6
#define LH_ST_P2M 7
#define LH_ST_UVM 8
#define LH_ST_UWM 9
#define LH_ST_VWM 10
#define LH_ST_WXM 11
#define LH_ST_WX2M 12
#define LH_ST_N 13
int lh_stats_get(lh_ctx *ctx, int ch, double *prof, int *istat);
/* One accumulated spectrum in the reference's column layout
 * sp(jlo:jhi, ncol) (spU, spV, spW, spUV, spP of spectra.f90): rows 0..ny
 * for U, W, UV, 0..ny-1 for V, 1..ny-1 for P. */
#define LH_SP_U 0
#define LH_SP_V 1
#define LH_SP_W 2
#define LH_SP_UV 3
#define LH_SP_P 4
#define LH_SP_N 5
int lh_spectrum_get(lh_ctx *ctx, int ch, int which, double *sp);
/* The wall maps of sl_stats (sl_stats2.f90:33-161), accumulated by
 * lh_stats_accumulate with the other statistics: for wall 0 (bottom, v row
 * 0) or 1 (top, v row ny-1), maps[4][(N1+2)*N2] = u1 and u3 interpolated to
 * the wall face and their wall-normal derivatives (the top one with the sign
 * flipped), as sums of spectral coefficients in the (N1+2) x N2 plane layout
 * of modes_to_planes_phys_lims_2 (bslip_u1_M, bslip_du1dy_M, bslip_u3_M,
 * bslip_du3dy_M).  write_sl_stats adds the two walls (stats.f90 of several
 * ranks: rank 0 the bottom, the last rank the top). */
int lh_slip_stats_get(lh_ctx *ctx, int ch, int wall, double *maps);

/* Learned closure of the nonlinear term.  No counterpart in the reference:
 * a coarse channel adds a network's estimate of what its grid misses.
 *
 * The network maps the 3x3x5 stencil of u1, u2, u3 around a point of the
 * channel's N1 x N2 physical grid (four_to_phys_N conventions) to a
 * correction of Nu1, Nu2, Nu3 at that point:
 *   135 inputs -> 128 -> 128 -> 128 -> 3, GELU (tanh form) after the hidden
 *   layers, bf16 weights and activations, fp32 accumulation.
 * Input f = field*45 + (dy+2)*9 + (dz+1)*3 + (dx+1): field 0..2 = u1, u2, u3
 * (row j of each field's own grid), dy = -2..2 rows (clamped to the field's
 * rows), dz, dx = -1..1 points (periodic), normalised as
 * (u - mean[field]) * inv_scale[field].  Output o times out_scale[o] is the
 * correction of Nu(o+1) at the point, applied on u rows 1..ny-1 (Nu1, Nu3)
 * and v rows 1..ny-2 (Nu2): the correction planes are transformed to modes
 * (phys_to_four_N conventions) and added to the nonlinear term of every mode
 * except the x-Nyquist (i = N1/2) and z-Nyquist (k = N2/2) ones.  Points are
 * the N1*N2*(ny-1) nodes of u rows 1..ny-1, index (j-1)*N1*N2 + z*N1 + x.
 *
 * params: LH_CL_NPARAM floats, row-major y = W x + b:
 *   W1[128][135] b1[128] W2[128][128] b2[128] W3[128][128] b3[128]
 *   W4[3][128] b4[3]
 * norm9: mean[3], inv_scale[3], out_scale[3].
 * The closure runs on channels of at most 64 x 64 modes. */
#define LH_CL_NIN 135
#define LH_CL_NHID 128
#define LH_CL_NOUT 3
#define LH_CL_NPARAM 50819
int lh_closure_set(lh_ctx *ctx, const float *params, const float *norm9);
/* on = 1: every substep of channel ch evaluates the network and adds its
 * correction (after the products, before the right-hand side).  on = 2: the
 * network runs at the first substep of each step only, on the state at the
 * start of the step; substeps 2 and 3 add the same correction (a third of
 * the network's cost; the correction is held over one step).  on = 0: off.
 * Drops the captured step graph. */
int lh_closure_enable(lh_ctx *ctx, int ch, int on);
/* On the current fields of channel ch: d_phys (device, floats
 * [3][ny+1][N2][N1]) gets u1, u2, u3 on the N1 x N2 grid (u2 rows 0..ny-1,
 * row ny zero); d_corr (device, same shape, or NULL) the network's
 * correction (rows outside the applied ranges zero). */
int lh_closure_eval(lh_ctx *ctx, int ch, float *d_phys, float *d_corr);
/* A training sample from channel chf of a fine context for channel chc of a
 * coarse one (same box, same y grid, at least the coarse modes): the fine
 * fields, box-filtered over a coarse cell (transfer function
 * sinc(pi i/N1c) sinc(pi m/N2c) for x-mode i, z wavenumber m; the coarse
 * z-Nyquist mode cleared) and restricted to the coarse modes, REPLACE the
 * coarse channel's u1, u2, u3, p.  Then, in the coarse layout,
 *   tau = filtered full nonlinear term of the fine channel
 *       - full nonlinear term of the coarse channel,
 * with Nyquist modes cleared, as the network's target.  d_phys receives the
 * coarse fields and d_tau the target, both as lh_closure_eval's d_phys
 * (device floats [3][ny+1][N2c][N1c]). */
int lh_closure_sample(lh_ctx *fine, int ch_fine, lh_ctx *coarse, int ch_coarse, float *d_phys, float *d_tau);
/* Average time (ms) of the closure's three parts over nrep launches on the
 * current fields: physical planes, network, correction to modes. */
int lh_closure_time(lh_ctx *ctx, int nrep, double *ms3);

/* Training of the closure network on one device (no counterpart in the
 * reference).  A trainer holds fp32 master parameters, AdamW moments and the
 * activations of a batch of `batch` points (a multiple of 64).  Forward and
 * backward passes are bf16 matrix-core GEMMs with fp32 accumulation; the
 * loss is the mean over points and outputs of (y - t)^2, t the normalised
 * target (tau / out_scale). */
typedef struct lh_mlp lh_mlp;
int lh_mlp_create(int batch, lh_mlp **out);
int lh_mlp_destroy(lh_mlp *m);
/* parameters as lh_closure_set; resets the AdamW state */
int lh_mlp_set_params(lh_mlp *m, const float *params, const float *norm9);
int lh_mlp_get_params(lh_mlp *m, float *params);
/* the batch from host arrays: x[batch][135] normalised inputs,
 * t[batch][3] normalised targets */
int lh_mlp_set_batch(lh_mlp *m, const float *x, const float *t);
/* a random batch from nsnap snapshots of lh_closure_sample (d_phys, d_tau:
 * nsnap consecutive [3][ny+1][N2][N1] blocks): sample n is point
 *   h mod (nsnap * (ny-1) * N1 * N2),
 *   h = splitmix64(seed ^ (n * 0xD1B54A32D192ED03)),
 * counted snapshot-major, then as lh_closure_eval's point index */
int lh_mlp_sample(lh_mlp *m, int N1, int N2, int ny, const float *d_phys, const float *d_tau, int nsnap,
                  uint64_t seed);
/* inputs at given points (host pts[batch]) of one physical-plane set, with
 * zero targets */
int lh_mlp_set_points(lh_mlp *m, int N1, int N2, int ny, const float *d_phys, const int *pts);
/* outputs of the current batch, y[batch][3] (normalised) */
int lh_mlp_forward(lh_mlp *m, float *y);
/* forward, loss and backward on the current batch; the gradient stays on
 * the device (lh_mlp_get_grad copies it out, LH_CL_NPARAM floats) */
int lh_mlp_grad(lh_mlp *m, double *loss);
int lh_mlp_get_grad(lh_mlp *m, float *grad);
/* one AdamW step with the last gradient (bias-corrected moments, decoupled
 * weight decay wd) */
int lh_mlp_adam(lh_mlp *m, float lr, float beta1, float beta2, float eps, float wd);
int lh_mlp_sync(lh_mlp *m);

#ifdef __cplusplus
}
#endif
#endif
