/*
 * lh_oracle.h -- CPU FP64 oracle for the plain-FFT RK3 step of LittleHarsh.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (liblhgpu.so) never links, loads or calls it.
 *
 * Field storage follows the reference's Fortran allocation of a band array
 * (`complex(8) f(jlim1:jlim2, ncol)`, littleharsh.f90:119-130): y is the
 * contiguous index, then the spectral column.  Column c (0-based) holds
 * x-mode i = c % (N1/2+1) and z-index k = c / (N1/2+1) (k = 0..N2-1, the
 * reference's columns_k - 1), the order proc_lims_columns builds for the
 * long pencils (start.f90:1574-1600).
 *
 * Row ranges (j is the reference's row index):
 *   u-grid (u1,u3,du1,du3):   j = 0 .. ny      (ny+1 rows incl. ghosts)
 *   v-grid (u2,du2):          j = 0 .. ny-1    (ny rows incl. walls)
 *   p-grid (p,psi,div):       j = 1 .. ny-1    (ny-1 rows)
 *   Nu1,Nu3:                  j = 1 .. ny-1
 *   Nu2:                      j = 1 .. ny-2
 * where ny is the number of v points wall to wall (input.in `ny`).
 */
#ifndef LH_ORACLE_H
#define LH_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lho_cfg {
  int N1, N2;          /* Fourier modes in x and z (N = 2*(Ngal/3)) */
  int Ngal1, Ngal2;    /* 3/2-padded physical grid */
  int ny;              /* v points wall to wall */
  double alp, bet;     /* Lx = 2pi/alp, Lz = 2pi/bet */
  double Re;           /* 1/nu, as the reference uses it */
  int grid_kind;       /* 0: polynomial y = a(j-q)+b(j-q)^p ; 1: tanh(gamma) */
  double dyq;          /* polynomial grid: dymax/dymin */
  int ppp;             /* polynomial grid exponent */
  double gamma;        /* tanh grid stretching factor */
  double bslpu1, bslpu3; /* Navier slip lengths used by the ghost update */
  int flag_ctpress;    /* 0: constant flow rate, 1: constant pressure gradient */
  double mpgx;         /* mean pressure gradient (initial / prescribed) */
  int fixed_dt;        /* 1: dt fixed; 0: dt = min(2.5 dtv, CFL dtc) at kRK=1 */
  double dt;
  double CFL;
  /* band edges (multi-resolution bands, start.f90:985-1060): bit 0 when the
   * first row of this column set is a band edge (the top short band's
   * bottom), bit 1 when the last row is (the bottom short band's top).  A
   * band edge keeps its solved ghost value: v_corr resets only the wall
   * ghosts (littleharsh.f90:992-1044).  0 for a whole channel. */
  int band_edge;
} lho_cfg;

typedef struct lho_state lho_state;

lho_state *lho_create(const lho_cfg *cfg);
void lho_destroy(lho_state *s);
int lho_set_threads(int n);

/* grid / wavenumber accessors (copies); arrays sized as documented */
int lho_nrows_u(const lho_state *s);   /* ny+1 */
int lho_ncol(const lho_state *s);      /* (N1/2+1)*N2 */
void lho_get_grid(const lho_state *s, double *yu, double *yv, double *dthdyu,
                  double *dthdyv, double *dyu2i, double *dyv2i);
/* dyu2i/dyv2i are 3 x rows, Fortran order dyu2i(1:3, j) */
void lho_set_grid(lho_state *s, const double *yu, const double *yv,
                  const double *dthdyu, const double *dthdyv);
void lho_get_scalars(const lho_state *s, double *out8);
/* out8 = {t, dt, mpgx, QxT, Qx, dtv, dtc, iter} */
void lho_set_scalars(lho_state *s, double t, double mpgx, double QxT, int iter);

/* field ids */
enum { LHO_U1 = 0, LHO_U2, LHO_U3, LHO_P, LHO_NU1, LHO_NU2, LHO_NU3,
       LHO_DU1, LHO_DU2, LHO_DU3, LHO_PSI, LHO_DIV, LHO_NFIELD };
/* copy a field in/out in the reference allocation layout (re,im interleaved) */
int lho_field_rows(const lho_state *s, int fid, int *jlo, int *jhi);
void lho_set_field(lho_state *s, int fid, const double *reim);
void lho_get_field(const lho_state *s, int fid, double *reim);
/* QxT from the current u1 mean mode (getini, start.f90:2166-2171) */
void lho_init_flowrate(lho_state *s);

/* RK3 pieces, littleharsh.f90:176-215 */
void lho_rhs0(lho_state *s, int kRK);
void lho_nonlinear(lho_state *s, int kRK);
void lho_solve_uw(lho_state *s, int kRK);
void lho_solve_v(lho_state *s, int kRK);
void lho_meanflow_ctU(lho_state *s, int kRK);
void lho_solve_p(lho_state *s, int kRK);
void lho_v_corr(lho_state *s, int kRK);
void lho_substep(lho_state *s, int kRK);
void lho_step(lho_state *s, int nsteps);
void lho_divergence(lho_state *s);

/* Plane transforms on one (Ngal1+2) x Ngal2 plane, Fortran layout, in place
 * (four_to_phys_du FOU3D.f90:766, phys_to_four_du FOU3D.f90:2253). */
void lho_four_to_phys_plane(const lho_state *s, double *plane);
void lho_phys_to_four_plane(const lho_state *s, double *plane);
/* four_to_phys_N / phys_to_four_N (FOU3D.f90:786-804, 2274-2293) on one
 * (N1+2) x N2 plane, in place */
void lho_four_to_phys_N_plane(const lho_state *s, double *plane);
void lho_phys_to_four_N_plane(const lho_state *s, double *plane);
/* FFTPACK-convention wrappers (cft_buff.f:1-65, rft_buff.f:24-90) */
void lho_cft(double *c, int n, int ise, int isa, int m, int iopt);
void lho_rft(double *c, int n, int isa, int m, int iopt);

/* Batched tridiagonal helpers on the reference's matrices (reference layout,
 * ncol x rows complex, re/im interleaved) */
void lho_lusol_helm(const lho_state *s, int kRK, int vgrid, const double *rhs, double *x);
void lho_lusolp(const lho_state *s, double *reim);
void lho_lusol0(const lho_state *s, int kRK, double *x); /* x(0:ny), LUsol0 */

/* Wall textures (planeBC codes per point of the N1 x N2 grid, index
 * z*N1 + x: 1 no-slip, 2 free-shear); NULL leaves that wall smooth.
 * Returns -1 on a bad code. */
int lho_set_texture(lho_state *s, const int *bottom, const int *top);
/* the textured U/W Helmholtz solve on arbitrary right-hand sides */
void lho_lusol_uw_tex(lho_state *s, int kRK, const double *rhs_u, const double *rhs_w, double *u, double *w);

/* Triad lists of the restricted nonlinear path (ops_in_planes,
 * FOU3D.f90:1724-1987; nonlinInter :1989-2030; nonlinRead, start.f90:
 * 2983-3069).  nfile list files; lens[q*9 + t-1] the rows of type t in file
 * q; rows the int8 lists concatenated (file, then type order), each in the
 * reference's (len, 4) Fortran layout; map_u[j-1], j = 1..ny-1, and
 * map_v[j-1], j = 1..ny-2, the file of each plane (-1: no triads).  Once
 * set, lho_nonlinear takes the list path.  nfile = 0 returns to the plain
 * path.  Returns -1 (and clears) on an entry outside the band. */
int lho_set_triads(lho_state *s, int nfile, const int *lens, const signed char *rows, const int *map_u,
                   const int *map_v);

/* One statistics sample of the current fields (stats.f90:1-107 and
 * spectra.f90:1-56), added to prof[13][ny+1] (profile q of write_stats'
 * order Um U2m Vm V2m Wm W2m Pm P2m UVm UWm VWm wxm wx2m at q*(ny+1) + j,
 * j the reference row) and sp[5][ncol][ny+1] (spU spV spW spUV spP at
 * w*ncol*(ny+1) + c*(ny+1) + j).  Physical planes are formed and summed as
 * the reference does. */
void lho_stats(lho_state *s, double *prof, double *sp);

/* sl_stats (sl_stats2.f90:33-161): adds one sample of the wall maps to
 * maps[2][4][(N1+2)*N2] (bottom / top wall; u1, du1dy, u3, du3dy at the wall
 * face, spectral coefficients in the (N1+2) x N2 plane layout, the top
 * gradient with the sign flipped as there). */
void lho_slstats(lho_state *s, double *maps);

/* ---- Multi-resolution bands (the reference's nband = 3 blocks,
 * declaration.f90:8-13).  A banded channel has long pencils, the modes of
 * the middle band's resolution (N1m x N2m), over the whole height, and short
 * pencils, the remaining modes of the wall bands' resolution (N1 x N2 of
 * the wall configuration), over the bottom band (v rows 0..nb+1) and the top
 * band (v rows ny-2-nb..ny-1), each with its own edge rows
 * (proc_lims_columns start.f90:1490-1774, jlim :1677-1711).  Planes of a
 * band transform at that band's resolution and hold every mode present at
 * their row (modes_to_planes_UVP FOU3D.f90:1000-1098 with dk,
 * proc_lims_planes start.f90:1836-2142).  Band q (0 bottom, 1 middle, 2
 * top) is an lho_state whose rows are the band's rows: the middle band the
 * whole channel at N1m x N2m, the wall bands sub-grids at N1 x N2 whose
 * long-pencil columns stay zero (those modes live in the middle band). */
typedef struct lhob_state lhob_state;
/* nb: the last v row of the bottom band, N(3,1) of ygrid (start.f90:
 * 999-1011); 1 <= nb and the two wall bands must not meet */
lhob_state *lhob_create(const lho_cfg *wall, int N1m, int N2m, int Ngal1m, int Ngal2m, int nb);
void lhob_destroy(lhob_state *b);
lho_state *lhob_band(lhob_state *b, int q);
/* global row of the band's local row 0 */
int lhob_row0(const lhob_state *b, int q);
/* 1 when column c of band q is a pencil of that band (a long pencil of the
 * middle band, a short pencil of a wall band) */
int lhob_owns(const lhob_state *b, int q, int c);
void lhob_init_flowrate(lhob_state *b);
void lhob_substep(lhob_state *b, int kRK);
void lhob_step(lhob_state *b, int nsteps);
/* the banded nonlinear term alone (interp, planes, products, fluxes, N),
 * as lho_nonlinear for one channel */
void lhob_nonlinear(lhob_state *b, int kRK);
#ifdef __cplusplus
}
#endif
#endif
