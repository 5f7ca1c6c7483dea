// lh_kernels.h -- host-side declarations of the kernel launchers.
#pragma once
#include <vector>

#include "lh_common.h"

namespace lh {

struct ChanDims {
  int N1, N2, Ng1, Ng2, ny;
};
// launch geometry covering every channel of a context
struct ShapeMax {
  int nchan = 0;
  int nslot = 0, ncol = 0, nyu = 0;
  std::vector<ChanDims> dims;
  // complex points per field that a workgroup of each nonlinear stage takes
  int e_zinv = 1024, e_xprod = 512, e_zfwd = 768;
  int zinv_tiles = 0, xprod_tiles = 0, zfwd_tiles = 0;
  size_t lds_zinv = 0, lds_xprod = 0, lds_zfwd = 0;  // in complex elements
  int solve_cols = 0;  // columns per workgroup of the LDS column solve (0: one lane per column, -1: one column per wave)
  int solve_scan = 1;  // LDS column solve: recurrences split over lane groups (0: one lane per sequence)
  int solve_waves = 1;  // LDS and wave column solves: waves per workgroup (1, 2 or 4)
  int xprod_ip = 0;     // x-product stage: in-place single-buffer stages where they fit
  int zfwd_ip = 0;      // z-forward stage: the same
  int zinv_ip = 0;      // z-inverse stage: the same
  int zfwd_fs = 0;      // z-forward stage: fields one after another in one LDS buffer, wider tiles
  int xprod_threads = 256;  // x-product stage: threads per workgroup (128, 192 or 256)
  int rhs_fused = 1;   // smooth-walled ensembles: right-hand sides computed inside the column solve
  int ghost_zinv = 1;  // in a captured step, ghost rows of x-independent textures done by k_zinv's extra workgroups
  int nphys = 0;       // N1*N2, largest over channels (texture mask kernels)
};
void shape_tiles(ShapeMax &m);
bool zfwd_fs_on(const ShapeMax &m);  // the field-by-field z-forward stage is in use
int zfwd_fs_len(const ShapeMax &m);  // its specialised length (0: mixed 24/48/96), -1 when off
int solve_cols_default(int nyu);
// rows per lane of the one-column-per-wave solve (solve_cols = -1): K with
// 64 K >= nyu, at most SOLVE_WAVE_KMAX; 0 when the column is too tall
constexpr int SOLVE_WAVE_KMAX = 6;
int solve_wave_rows(int nyu);
// its columns per workgroup for LH_TUNE_SOLVE_WAVES `waves` (8, or 2)
int solve_wave_cols(int nyu, int waves);
// waves (consecutive columns) per workgroup of that form, and its LDS: a
// staging tile of nw columns x S complex rows (S = nyu made odd, so that
// neighbouring columns' rows fall on different banks) and 9 row-coefficient
// arrays
__host__ __device__ inline int solve_wave_stride(int nyu) { return nyu | 1; }
__host__ __device__ inline size_t solve_wave_lds(int nyu, int nw) {
  return (size_t)nw * solve_wave_stride(nyu) * 16 + (size_t)9 * nyu * 8;
}
__host__ __device__ size_t solve_lds_bytes(int nyu, int C);

void shape_add(ShapeMax &m, int N1, int N2, int Ng1, int Ng2, int ny);
// kRK = 1 also records the CFL maxima of dtc_calc in the x-product stage;
// ghost: the z-inverse stage also runs the ghost-row pass of channels with
// x-independent textured walls (the one launch_tex_1d(..., 1, 1, 1, ...)
// makes; z lengths tex_ghost_small only); dtblk: the z-forward stage also
// runs the time-step update that launch_dt_update makes
hipError_t launch_nl_stage(char *arena, const Chan *d_chans, const ShapeMax &m, bool mixed, int stage, hipStream_t s,
                           int kRK = 1, int ghost = 0, int dtblk = 0);
hipError_t launch_nonlinear(char *arena, const Chan *d_chans, const ShapeMax &m, bool mixed, hipStream_t s,
                            int kRK = 1, int ghost = 0, int dtblk = 0);
// band = 1: only the near-wall rows of r1 and r3 that textured walls'
// source passes read, for a solve that computes the right-hand sides itself
hipError_t launch_rhs(char *arena, const Chan *d_chans, const ShapeMax &m, int kRK, hipStream_t s, int band = 0);
hipError_t launch_factors(char *arena, const Chan *d_chans, const ShapeMax &m, hipStream_t s);
hipError_t launch_dt_update(char *arena, const Chan *d_chans, const ShapeMax &m, hipStream_t s);
hipError_t launch_u11(char *arena, const Chan *d_chans, const ShapeMax &m, hipStream_t s);
// err, Qx, Umax, Uslip of every channel into out[ch*8 + 0..3] (device)
hipError_t launch_history(char *arena, const Chan *d_chans, const ShapeMax &m, double *out, hipStream_t s);
// fuse: compute the right-hand sides inside the solve instead of reading the
// r1, r2, r3 that launch_rhs stores (smooth-walled channels; with textured
// walls the one-column-per-wave form, C < 0, which reads the wall rows'
// sources, after launch_rhs with band = 1)
// keep = 0 lets the wave form skip storing psi and div (p still gets psi)
hipError_t launch_solve(char *arena, const Chan *d_chans, const ShapeMax &m, int kRK, hipStream_t s, int fuse = 0,
                        int keep = 1);
hipError_t launch_flowrate(char *arena, const Chan *d_chans, int nchan, int set_target, hipStream_t s);
// a banded channel: the channels of its bottom wall band, middle band and
// top wall band, and the top band's first row in the middle band's rows
struct BandLink {
  int cb, cm, ct, r0t;
};
// out = 0: the long pencils' rows u1, u2, u3, p of the wall bands from the
// middle band; out = 1: the wall bands' nonlinear terms and fluxes of those
// pencils back to the middle band
hipError_t launch_band_copy(char *arena, const Chan *d_chans, const BandLink *d_links, int nlinks, int ncol_mid,
                            int out, hipStream_t s);
hipError_t launch_to_rows(const double2 *src, double2 *dst, int ncol, int nrows, int jlo, int ldc,
                          hipStream_t s);
hipError_t launch_from_rows(const double2 *src, double2 *dst, int ncol, int nrows, int jlo, int ldc,
                            hipStream_t s);
// stand-alone plane transforms (reference plane layout); zero_nyq zeroes
// row Ngal2/2 as four_to_phys_du / phys_to_four_du do (the N-grid
// four_to_phys_N / phys_to_four_N do not)
hipError_t launch_plane_f2p(int N1, int Ng1, int Ng2, double *planes, int nplanes, const double2 *twz,
                            const double2 *twx, hipStream_t s, bool zero_nyq = true);
hipError_t launch_plane_p2f(int N1, int Ng1, int Ng2, double *planes, int nplanes, const double2 *twz,
                            const double2 *twx, hipStream_t s, bool zero_nyq = true);
hipError_t launch_cft(int n, int ise, int isa, int m, int iopt, double *c, const double2 *tw, hipStream_t s);
hipError_t launch_rft(int n, int isa, int m, int iopt, double *c, const double2 *tw, hipStream_t s);
// textured walls: channels of one (N1, N2) shape share a run of work planes
struct TexGroup {
  int N1, N2, nplanes;
  size_t o_texw, o_twz, o_twx;  // arena offsets: planes, N2- and N1-point twiddles
};
hipError_t launch_tex_tables(char *arena, const Chan *d_chans, const ShapeMax &m, hipStream_t s);
// adjacent-row values U of the smooth solution (before every source pass)
hipError_t launch_tex_adj(char *arena, const Chan *d_chans, const ShapeMax &m, int kRK, hipStream_t s);
hipError_t launch_tex_sources(char *arena, const Chan *d_chans, const ShapeMax &m,
                              const std::vector<TexGroup> &groups, int kRK, int nit, hipStream_t s);
hipError_t launch_tex_ghosts(char *arena, const Chan *d_chans, const ShapeMax &m,
                             const std::vector<TexGroup> &groups, hipStream_t s);
// the same with each field's two wall planes resident in one workgroup's LDS
size_t tex_fused_bytes(const std::vector<TexGroup> &groups);
hipError_t launch_tex_sources_fused(char *arena, const Chan *d_chans, const ShapeMax &m, size_t lds, int kRK,
                                    int nit, hipStream_t s);
hipError_t launch_tex_ghosts_fused(char *arena, const Chan *d_chans, const ShapeMax &m, size_t lds, hipStream_t s);
// walls with x-independent masks (Chan::tex1d): the wall iteration per
// x-mode; mode 0 sources (nit passes, after k_tex_adj), mode 1 ghost rows
hipError_t launch_tex_1d(char *arena, const Chan *d_chans, const ShapeMax &m, int kRK, int nit, int mode,
                         hipStream_t s);
// triad-list products and their derivatives for channels with Chan::tri
hipError_t launch_triads(char *arena, const Chan *d_chans, const ShapeMax &m, hipStream_t s);
// one statistics sample of every channel: row profiles and spectra added to
// the accumulators at sbase + Chan::o_stat / o_spec
hipError_t launch_stats(char *arena, const Chan *d_chans, const ShapeMax &m, char *sbase, hipStream_t s);
// stand-alone Helmholtz solves on reference-layout device data
hipError_t launch_lusol_ref2(char *arena, const Chan *d_chan, int kRK, int vgrid, const double2 *rhs_a,
                             const double2 *rhs_b, double2 *x_a, double2 *x_b, double *scr, int ncol,
                             hipStream_t s);

hipError_t launch_lusol_p_ref(char *arena, const Chan *d_chan, double2 *x, int ncol, hipStream_t s);
hipError_t launch_lusol0_ref(char *arena, const Chan *d_chan, int kRK, double *x, double *scr, hipStream_t s);
// column <-> plane transposes on reference-layout device data
hipError_t launch_modes_to_planes(int N1, int N2, int Ng1, int Ng2, const double2 *cols, int nrows, double *planes,
                                  hipStream_t s);
hipError_t launch_planes_to_modes(int N1, int N2, int Ng1, int Ng2, const double *planes, int nrows, double2 *cols,
                                  hipStream_t s);

}  // namespace lh
