// lh_common.h -- shared device-side definitions of liblhgpu (gfx950).
//
// Layout in HBM (one channel):
//   spectral fields  f[row][col]  double2, row stride ldc (ncol rounded up
//                    to 64), rows j as in the reference (u: 0..ny,
//                    v: 0..ny-1, p: 1..ny-1 stored at row j of a u-sized
//                    array).  One y row of every mode is contiguous, so a
//                    plane's spectral data is one contiguous slice (the
//                    plane transpose is free) and column-wise solves read
//                    row j of 64 consecutive columns in one coalesced access.
//   z-spectra        hz[slot][3][Ngal2][nxp]   (nonlinear pipeline stage 1)
//   x-half-spectra   gx[slot][4][Ngal2][nxp]   (stage 2)
//                    nxp = nxh rounded up to 16, so that tiles of 8 or 16
//                    x-modes are whole 128-byte lines
//   slot = j-1 for u-plane j = 1..ny-1, ny-1 + j-1 for v-plane j = 1..ny-2.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lh {

// spare rows after every field (and factor table) for branch-free blocked
// sweeps of the column solve (k_solve reads up to 3*SB rows past an end)
constexpr int PADROWS = 32;

// RK3 coefficients (Le & Moin 1991), start.f90:419-433
__host__ __device__ inline double rk_a(int k) { return k == 1 ? 4.0 / 15.0 : (k == 2 ? 1.0 / 15.0 : 1.0 / 6.0); }
__host__ __device__ inline double rk_b(int k) { return rk_a(k); }
__host__ __device__ inline double rk_g(int k) { return k == 1 ? 8.0 / 15.0 : (k == 2 ? 5.0 / 12.0 : 3.0 / 4.0); }
__host__ __device__ inline double rk_c(int k) { return k == 1 ? 0.0 : (k == 2 ? -17.0 / 60.0 : -5.0 / 12.0); }
__host__ __device__ inline double rk_d(int k) { return k == 1 ? 8.0 / 15.0 : (k == 2 ? 2.0 / 15.0 : 1.0 / 3.0); }

// per-channel device scalars (in HBM so that graph replays see updates)
// SC_QX1..+2: flow rate of the unit-forcing solution per kRK; SC_DTC the CFL
// time step of dtc_calc; SC_DTV the viscous one; SC_M1..M3 the bit patterns of
// the CFL maxima max|u1|, max|v dthdyv|, max|u3| (non-negative doubles
// compare as integers)
// SC_DGX: the last flow-rate correction g (dgx of flowrate_corr)
enum { SC_MPGX = 0, SC_QXT, SC_QX, SC_DT, SC_T, SC_QX1, SC_DTC = 8, SC_DTV, SC_M1, SC_M2, SC_M3, SC_DGX, SC_NSCAL = 16 };

// statistics profiles in the order write_stats writes them (stats.f90:356-368)
enum { ST_UM = 0, ST_U2M, ST_VM, ST_V2M, ST_WM, ST_W2M, ST_PM, ST_P2M, ST_UVM, ST_UWM, ST_VWM, ST_WXM, ST_WX2M, ST_N };
// spectra (spectra.f90:9-13): u, v, w, u.v_itp, p
enum { SP_U = 0, SP_V, SP_W, SP_UV, SP_P, SP_N };

// Per-channel descriptor.  Every device array lives in one context-wide
// arena; the descriptor holds byte offsets, and kernels derive typed pointers
// from the arena base they receive as an argument (so that the compiler sees
// global-memory pointers and emits global_load/store, not flat accesses).
struct Chan {
  int N1, N2, Ng1, Ng2;
  int ny, nyu, nyv, nxh, ncol, ldc;
  int nxp;  // row length of the hz/gx intermediates: nxh rounded up to 16
  // plane slots: u-planes j = pu0 .. pu0+nplu-1, then v-planes
  // j = pv0 .. pv0+nplv-1 (pu0 = pv0 = 1 for a whole channel; the middle
  // band of a banded channel transforms only its own band's planes)
  int nplu, nplv, nslot, pu0, pv0;
  // band edges of a wall band of a banded channel: bit 0 when the first row
  // is a band edge, bit 1 when the last row is; a band edge keeps its solved
  // ghost value (v_corr resets only wall ghosts, littleharsh.f90:992-1044)
  int band_edge;
  int flag_ctpress;
  int adaptive;  // dt = min(2.5 dtv, CFL dtc) at every kRK = 1 (else fixed)
  double CFL;
  double Re, alp, bet;
  double gw1, gw2, gwbc1, gwbc3, gwi1, gwi2;
  // grid / wavenumbers: yu,yv,dthdyu,dthdyv [rows], dyu2i/dyv2i [3*j+m],
  // kx,k2x [nxh], kz,k2z [N2], twz/twx exp(-2 pi i m/Ngal2), exp(-2 pi i m/Ngal1)
  size_t o_yu, o_yv, o_dthdyu, o_dthdyv, o_dyu2i, o_dyv2i, o_kx, o_kz, o_k2x, o_k2z, o_twz, o_twx;
  // state: fields [row][ldc] double2; N*o nonlinear terms of the previous
  // substep; r* Helmholtz right-hand sides; nl* x/z part of the new nonlinear
  // terms; fuv,fwv (v rows) and fvv (u rows) the y-fluxes
  size_t o_u1, o_u2, o_u3, o_p, o_N1o, o_N2o, o_N3o, o_r1, o_r2, o_r3, o_nl1, o_nl2, o_nl3;
  size_t o_fuv, o_fwv, o_fvv, o_psi, o_div;
  // scr: Thomas scratch [2][nyu][ldc]; u11: mean-mode unit-forcing solutions
  // [3][nyu]; hz/gx (or hzf/gxf in FP32): nonlinear pipeline intermediates;
  // scal: SC_* scalars
  size_t o_scr, o_u11, o_hz, o_gx, o_hzf, o_gxf, o_scal;
  // textured walls (tex bit 0 bottom, bit 1 top): mask[w][z*N1 + x] planeBC
  // codes (1 no-slip, 2 free-shear); wtab[(kRK-1)*2 + w][row][ldc] rows 1
  // (w = 0) and ny-1 (w = 1) of the inverse U/W Helmholtz operator; texu /
  // texw [f*2 + w][ncol] adjacent-row values and the wall-source work planes
  // (N-grid, (N1+2) x N2 doubles each); iscal[0..1] truncation depths of the
  // two inverse rows
  int tex;
  // 1 when every textured wall's mask is independent of x (streamwise
  // ridges): the wall iteration then decouples per x-mode (k_tex_1d)
  int tex1d;
  size_t o_mask, o_wtab, o_texu, o_texw, o_iscal;
  size_t o_twzN, o_twxN;  // N2- and N1-point twiddles of the N-grid
  size_t o_cfl;           // [slot][tile][2] CFL maxima of k_xprod's workgroups
  // [slot][2] the plane means of u1 and u3 (the (0,0) modes, interpolated on
  // v-planes) that the FP32 stages of mixed precision leave out
  size_t o_pmean;
  // statistics accumulators, byte offsets into the context's separate
  // statistics buffer (allocated on first use): prof[ST_*][nyu] row profiles,
  // spec[SP_*][nyu][ldc] per-mode spectra (doubles) and the wall maps of
  // sl_stats, slip[wall][4][ncol] (double2: u1, du1dy, u3, du3dy)
  size_t o_stat, o_spec, o_slip;
  // triad-list nonlinear path (tri = 1): file of each plane slot (-1: none),
  // per (file, type) the target-sorted rows (tri_off[(file*9 + type-1) *
  // (ncol+1) + c] .. +1 into tri_ent + tri_base[file*9 + type-1]; an entry
  // packs the passive column (bits 0-14), its conjugate flag (15), the active
  // column (16-30) and its flag (31)), and the product fields
  // tri_prod[slot][5][ncol]; separate device allocation
  int tri;
  const int *tri_map;
  const size_t *tri_base;
  const unsigned *tri_off, *tri_ent;
  double2 *tri_prod;
  // Thomas factors of the fixed-dt operators, [row][ldc] doubles each:
  // o_fac + ((kRK-1)*4 + m) * nyu*ldc*8 for m = 0 dU, 1 cU, 2 dV, 3 cV
  // (U/W and V Helmholtz, LU_dec order), then o_fac + 12*nyu*ldc*8 dP and
  // +13 eP: the Poisson operator eliminated from the top row down
  size_t o_fac;
};

// typed device pointers of one channel, derived from the arena base
struct View {
  const double *yu, *yv, *dthdyu, *dthdyv, *dyu2i, *dyv2i, *kx, *kz, *k2x, *k2z;
  const double2 *twz, *twx;
  double2 *u1, *u2, *u3, *p, *N1o, *N2o, *N3o, *r1, *r2, *r3, *nl1, *nl2, *nl3, *fuv, *fwv, *fvv, *psi, *div;
  double *scr, *u11, *scal;
  double2 *hz, *gx;
  float2 *hzf, *gxf;
  double *fac;
  int *mask, *iscal;
  double *wtab, *cfl, *pmean;
  double2 *texu, *texw;
  __host__ __device__ View(char *A, const Chan &c)
      : yu((const double *)(A + c.o_yu)), yv((const double *)(A + c.o_yv)),
        dthdyu((const double *)(A + c.o_dthdyu)), dthdyv((const double *)(A + c.o_dthdyv)),
        dyu2i((const double *)(A + c.o_dyu2i)), dyv2i((const double *)(A + c.o_dyv2i)),
        kx((const double *)(A + c.o_kx)), kz((const double *)(A + c.o_kz)), k2x((const double *)(A + c.o_k2x)),
        k2z((const double *)(A + c.o_k2z)), twz((const double2 *)(A + c.o_twz)),
        twx((const double2 *)(A + c.o_twx)), u1((double2 *)(A + c.o_u1)), u2((double2 *)(A + c.o_u2)),
        u3((double2 *)(A + c.o_u3)), p((double2 *)(A + c.o_p)), N1o((double2 *)(A + c.o_N1o)),
        N2o((double2 *)(A + c.o_N2o)), N3o((double2 *)(A + c.o_N3o)), r1((double2 *)(A + c.o_r1)),
        r2((double2 *)(A + c.o_r2)), r3((double2 *)(A + c.o_r3)), nl1((double2 *)(A + c.o_nl1)),
        nl2((double2 *)(A + c.o_nl2)), nl3((double2 *)(A + c.o_nl3)), fuv((double2 *)(A + c.o_fuv)),
        fwv((double2 *)(A + c.o_fwv)), fvv((double2 *)(A + c.o_fvv)), psi((double2 *)(A + c.o_psi)),
        div((double2 *)(A + c.o_div)), scr((double *)(A + c.o_scr)), u11((double *)(A + c.o_u11)),
        scal((double *)(A + c.o_scal)), hz((double2 *)(A + c.o_hz)), gx((double2 *)(A + c.o_gx)),
        hzf((float2 *)(A + c.o_hzf)), gxf((float2 *)(A + c.o_gxf)), fac((double *)(A + c.o_fac)),
        mask((int *)(A + c.o_mask)), iscal((int *)(A + c.o_iscal)), wtab((double *)(A + c.o_wtab)),
        cfl((double *)(A + c.o_cfl)), pmean((double *)(A + c.o_pmean)),
        texu((double2 *)(A + c.o_texu)), texw((double2 *)(A + c.o_texw)) {}
};

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cscale(double s, double2 a) { return make_double2(s * a.x, s * a.y); }
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
// (i*k)*a with the reference's complex(0,k) product: (0*ar - k*ai, 0*ai + k*ar)
__device__ __forceinline__ double2 cimul(double k, double2 a) { return make_double2(-k * a.y, k * a.x); }

__device__ __forceinline__ float2 fadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 fsub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 fmulc(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

}  // namespace lh
