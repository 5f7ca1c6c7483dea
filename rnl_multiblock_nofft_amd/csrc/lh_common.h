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
//   z-spectra        hz[slot][3][Ngal2][nxh]   (nonlinear pipeline stage 1)
//   x-half-spectra   gx[slot][4][Ngal2][nxh]   (stage 2)
//   slot = j-1 for u-plane j = 1..ny-1, ny-1 + j-1 for v-plane j = 1..ny-2.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lh {

// RK3 coefficients (Le & Moin 1991), start.f90:419-433
__host__ __device__ inline double rk_a(int k) { return k == 1 ? 4.0 / 15.0 : (k == 2 ? 1.0 / 15.0 : 1.0 / 6.0); }
__host__ __device__ inline double rk_b(int k) { return rk_a(k); }
__host__ __device__ inline double rk_g(int k) { return k == 1 ? 8.0 / 15.0 : (k == 2 ? 5.0 / 12.0 : 3.0 / 4.0); }
__host__ __device__ inline double rk_c(int k) { return k == 1 ? 0.0 : (k == 2 ? -17.0 / 60.0 : -5.0 / 12.0); }
__host__ __device__ inline double rk_d(int k) { return k == 1 ? 8.0 / 15.0 : (k == 2 ? 2.0 / 15.0 : 1.0 / 3.0); }

// per-channel device scalars (in HBM so that graph replays see updates)
enum { SC_MPGX = 0, SC_QXT, SC_QX, SC_DT, SC_T, SC_NSCAL = 8 };

struct Chan {
  int N1, N2, Ng1, Ng2;
  int ny, nyu, nyv, nxh, ncol, ldc;
  int nplu, nplv, nslot;
  int flag_ctpress;
  double Re, alp, bet;
  double gw1, gw2, gwbc1, gwbc3, gwi1, gwi2;
  // grid / wavenumbers
  const double *yu, *yv, *dthdyu, *dthdyv, *dyu2i, *dyv2i;  // dyu2i[3*j+m]
  const double *kx, *kz, *k2x, *k2z;                        // [nxh], [N2]
  const double2 *twz, *twx;  // exp(-2 pi i m / Ng2), exp(-2 pi i m / Ng1)
  // state
  double2 *u1, *u2, *u3, *p;
  double2 *N1o, *N2o, *N3o;  // nonlinear terms of the previous substep
  double2 *r1, *r2, *r3;     // right-hand sides of the Helmholtz solves
  double2 *nl1, *nl2, *nl3;  // x/z part of the new nonlinear terms
  double2 *fuv, *fwv, *fvv;  // y-fluxes uv, wv (v rows) and vv (u rows)
  double2 *psi, *div;
  double *scr;               // Thomas scratch [nyu][ldc]
  double *u11;               // unit-forcing mean-mode solution [nyu]
  double2 *hz, *gx;          // nonlinear pipeline intermediates (FP64 mode)
  float2 *hzf, *gxf;         // same, FP32 (mixed mode)
  double *scal;              // SC_* scalars
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
