// lh_standalone.hip -- stand-alone drop-in kernels on reference-layout data:
//   * plane transforms four_to_phys_du / phys_to_four_du (FOU3D.f90:766-784,
//     2253-2272) on (Ngal1+2) x Ngal2 Fortran-order planes, FFTPACK
//     conventions (cft_buff.f:1-65, rft_buff.f:24-90);
//   * the Helmholtz solves of LUsolU_W / LUsolV (tridLU_3D.f90:55-189), the
//     Poisson solve LUsolP (tridLU_3D.f90:192-226) and the mean-mode solve
//     LUsol0 (tridLU_3D.f90:228-254) on x(jlo:jhi, ncol) arrays;
//   * the column <-> plane transposes modes_to_planes_UVP /
//     planes_to_modes_UVP / planes_to_modes_NUVP (FOU3D.f90:1000-1098,
//     2295-2442, 2772-2854) as LDS-tiled transposes.
#include "lh_common.h"
#include "lh_fft.h"
#include "lh_kernels.h"

namespace lh {

static __host__ __device__ int plane_tile(int n) {
  int t = 1024 / n;
  return t < 1 ? 1 : t;
}

// Complex transforms of length Ng2 of nv vectors (cft, cft_buff.f:1-65):
// element z of vector i at P[z ise + i isa] (re) and the next double (im),
// P the plane blockIdx.y, pstride doubles apart.  The plane transforms call
// it for the retained x-modes i = 0..N1/2 (ise = Ngal1+2, isa = 2).
// dir = +1: inverse (with zero_nyq: zero row Ngal2/2 first); dir = -1:
// forward, scaled, then (zero_nyq) zero row Ngal2/2 of the whole plane.
// TI: vectors per workgroup (at most plane_tile(Ng2), which the LDS size
// the launches give assumes)
__global__ void __launch_bounds__(256) k_plane_z(double *planes, int nv, int Ng1, int Ng2, int ise, int isa,
                                                 size_t pstride, int dir, const double2 *twz, int zero_nyq, int TI) {
  const int igal = Ng1 + 2;
  double *P = planes + (size_t)blockIdx.y * pstride;
  const int i0 = blockIdx.x * TI;
  extern __shared__ double2 smem_d[];
  if (i0 < nv) {
    const int ni = min(TI, nv - i0);
    double2 *A = smem_d, *B = smem_d + TI * Ng2, *TW = B + TI * Ng2;
    stage_twiddles<double2>(TW, twz, Ng2);
    for (int t = threadIdx.x; t < ni * Ng2; t += blockDim.x) {
      int z = t / ni, ii = t - z * ni;
      const size_t at = (size_t)z * ise + (size_t)(i0 + ii) * isa;
      double2 v = make_double2(P[at], P[at + 1]);
      if (zero_nyq && dir > 0 && z == Ng2 / 2) v = make_double2(0, 0);
      A[ii * Ng2 + z] = v;
    }
    __syncthreads();
    double2 *R = fft_lds<double2>(A, B, Ng2, ni, Ng2, dir, TW);
    const double s = dir < 0 ? 1.0 / Ng2 : 1.0;
    for (int t = threadIdx.x; t < ni * Ng2; t += blockDim.x) {
      int z = t / ni, ii = t - z * ni;
      double2 v = R[ii * Ng2 + z];
      if (zero_nyq && dir < 0 && z == Ng2 / 2) v = make_double2(0, 0);
      const size_t at = (size_t)z * ise + (size_t)(i0 + ii) * isa;
      P[at] = v.x * s;
      P[at + 1] = v.y * s;
    }
  }
  // entries not transformed: row Ngal2/2 is zeroed (four_to_phys zeroes it
  // before, phys_to_four after the transforms)
  if (zero_nyq && blockIdx.x == 0) {
    __syncthreads();
    for (int ii = 2 * nv + threadIdx.x; ii < igal; ii += blockDim.x) P[(size_t)(Ng2 / 2) * igal + ii] = 0.0;
  }
}

// Real transforms of length Ng1 of Ng2 vectors (rft, rft_buff.f:24-90):
// vector z at P + z igal (igal >= Ng1 + 2 doubles apart), P the plane
// blockIdx.y, pstride doubles apart; two vectors share one complex FFT.
__global__ void __launch_bounds__(256) k_plane_x(double *planes, int Ng1, int Ng2, int igal, size_t pstride, int dir,
                                                 const double2 *twx) {
  const int h = Ng1 / 2;
  double *P = planes + (size_t)blockIdx.y * pstride;
  const int TZ = plane_tile(Ng1);  // pairs of rows per block
  const int z0 = blockIdx.x * 2 * TZ;
  if (z0 >= Ng2) return;
  const int nzr = min(2 * TZ, Ng2 - z0);  // rows in this block
  const int npair = (nzr + 1) / 2;
  extern __shared__ double2 smem_d[];
  double2 *A = smem_d, *B = smem_d + TZ * Ng1, *TW = B + TZ * Ng1;
  stage_twiddles<double2>(TW, twx, Ng1);
  if (dir > 0) {
    for (int t = threadIdx.x; t < npair * Ng1; t += blockDim.x) {
      int q = t / Ng1, m = t - q * Ng1;
      int za = z0 + 2 * q, zb = za + 1;
      double2 a = make_double2(0, 0), b = make_double2(0, 0);
      int mm = m <= h ? m : Ng1 - m;
      double sg = m <= h ? 1.0 : -1.0;
      const double *ra = P + (size_t)za * igal;
      a = make_double2(ra[2 * mm], sg * ra[2 * mm + 1]);
      if (mm == 0 || mm == h) a.y = 0.0;
      if (zb < z0 + nzr) {
        const double *rb = P + (size_t)zb * igal;
        b = make_double2(rb[2 * mm], sg * rb[2 * mm + 1]);
        if (mm == 0 || mm == h) b.y = 0.0;
      }
      A[q * Ng1 + m] = make_double2(a.x - b.y, a.y + b.x);
    }
    __syncthreads();
    double2 *R = fft_lds<double2>(A, B, Ng1, npair, Ng1, +1, TW);
    // keep the packed cc(n/2) entries where the reference leaves them
    for (int t = threadIdx.x; t < npair * Ng1; t += blockDim.x) {
      int q = t / Ng1, n = t - q * Ng1;
      int za = z0 + 2 * q, zb = za + 1;
      double2 v = R[q * Ng1 + n];
      P[(size_t)za * igal + n] = v.x;
      if (zb < z0 + nzr) P[(size_t)zb * igal + n] = v.y;
    }
  } else {
    for (int t = threadIdx.x; t < npair * Ng1; t += blockDim.x) {
      int q = t / Ng1, n = t - q * Ng1;
      int za = z0 + 2 * q, zb = za + 1;
      double a = P[(size_t)za * igal + n];
      double b = zb < z0 + nzr ? P[(size_t)zb * igal + n] : 0.0;
      A[q * Ng1 + n] = make_double2(a, b);
    }
    __syncthreads();
    double2 *R = fft_lds<double2>(A, B, Ng1, npair, Ng1, -1, TW);
    const double inv = 1.0 / Ng1;
    for (int t = threadIdx.x; t < npair * (h + 1); t += blockDim.x) {
      int q = t / (h + 1), m = t - q * (h + 1);
      int za = z0 + 2 * q, zb = za + 1;
      double2 Ym = R[q * Ng1 + m], Yc = R[q * Ng1 + (m == 0 ? 0 : Ng1 - m)];
      double2 Pa, Pb;
      if (m == 0 || m == h) {
        Pa = make_double2(0.5 * (Ym.x + Yc.x) * inv, 0.0);
        Pb = make_double2(0.5 * (Ym.y + Yc.y) * inv, 0.0);
      } else {
        Pa = make_double2(0.5 * (Ym.x + Yc.x) * inv, 0.5 * (Ym.y - Yc.y) * inv);
        Pb = make_double2(0.5 * (Ym.y + Yc.y) * inv, -0.5 * (Ym.x - Yc.x) * inv);
      }
      P[(size_t)za * igal + 2 * m] = Pa.x;
      P[(size_t)za * igal + 2 * m + 1] = Pa.y;
      if (zb < z0 + nzr) {
        P[(size_t)zb * igal + 2 * m] = Pb.x;
        P[(size_t)zb * igal + 2 * m + 1] = Pb.y;
      }
    }
  }
}

hipError_t launch_plane_f2p(int N1, int Ng1, int Ng2, double *planes, int nplanes, const double2 *twz,
                            const double2 *twx, hipStream_t s, bool zero_nyq) {
  const int igal = Ng1 + 2;
  const size_t ps = (size_t)igal * Ng2;
  int nv = N1 / 2 + 1, TI = plane_tile(Ng2), TZ = plane_tile(Ng1);
  dim3 gz((nv + TI - 1) / TI, nplanes), gx((Ng2 + 2 * TZ - 1) / (2 * TZ), nplanes);
  hipLaunchKernelGGL(k_plane_z, gz, dim3(256), (2 * TI + 1) * Ng2 * sizeof(double2), s, planes, nv, Ng1, Ng2, igal,
                     2, ps, +1, twz, (int)zero_nyq, TI);
  hipLaunchKernelGGL(k_plane_x, gx, dim3(256), (2 * TZ + 1) * Ng1 * sizeof(double2), s, planes, Ng1, Ng2, igal, ps,
                     +1, twx);
  return hipGetLastError();
}

hipError_t launch_plane_p2f(int N1, int Ng1, int Ng2, double *planes, int nplanes, const double2 *twz,
                            const double2 *twx, hipStream_t s, bool zero_nyq) {
  const int igal = Ng1 + 2;
  const size_t ps = (size_t)igal * Ng2;
  int nv = N1 / 2 + 1, TI = plane_tile(Ng2), TZ = plane_tile(Ng1);
  dim3 gz((nv + TI - 1) / TI, nplanes), gx((Ng2 + 2 * TZ - 1) / (2 * TZ), nplanes);
  hipLaunchKernelGGL(k_plane_x, gx, dim3(256), (2 * TZ + 1) * Ng1 * sizeof(double2), s, planes, Ng1, Ng2, igal, ps,
                     -1, twx);
  hipLaunchKernelGGL(k_plane_z, gz, dim3(256), (2 * TI + 1) * Ng2 * sizeof(double2), s, planes, nv, Ng1, Ng2, igal,
                     2, ps, -1, twz, (int)zero_nyq, TI);
  return hipGetLastError();
}

// cft(c, ise, isa, m, iopt, aux) (cft_buff.f:1-65) on device data: m complex
// vectors of length n, strides in doubles; iopt < 0 forward (scaled 1/n).
hipError_t launch_cft(int n, int ise, int isa, int m, int iopt, double *c, const double2 *tw, hipStream_t s) {
  const int TI = plane_tile(n);
  hipLaunchKernelGGL(k_plane_z, dim3((m + TI - 1) / TI, 1), dim3(256), (2 * TI + 1) * n * sizeof(double2), s, c, m,
                     0, n, ise, isa, (size_t)0, iopt < 0 ? -1 : +1, tw, 0, TI);
  return hipGetLastError();
}

// rft(c, isa, m, iopt, aux) (rft_buff.f:24-90) on device data: m real
// vectors of length n, isa >= n + 2 doubles apart, packed as cc(k) =
// c(2k+1) + i c(2k+2), k = 0..n/2; iopt < 0 forward (scaled 1/n, imaginary
// parts of cc(0) and cc(n/2) zeroed), else inverse (they are ignored and
// c(n+1), c(n+2) keep their values).
hipError_t launch_rft(int n, int isa, int m, int iopt, double *c, const double2 *tw, hipStream_t s) {
  const int TZ = plane_tile(n);
  hipLaunchKernelGGL(k_plane_x, dim3((m + 2 * TZ - 1) / (2 * TZ), 1), dim3(256), (2 * TZ + 1) * n * sizeof(double2), s,
                     c, n, m, isa, (size_t)0, iopt < 0 ? -1 : +1, tw);
  return hipGetLastError();
}

// LUsolU_W (vgrid = 0, two right-hand sides) / LUsolV (vgrid = 1, one)
__global__ void __launch_bounds__(64) k_lusol_ref(char *arena, const Chan *chp, int kRK, int vgrid, const double2 *ra,
                                                  const double2 *rb, double2 *xa, double2 *xb, double *scr,
                                                  int ncol) {
  const Chan ch = *chp;
  const View v(arena, ch);
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncol) return;
  const int i = c % ch.nxh, k = c / ch.nxh;
  const int n = vgrid ? ch.nyv : ch.nyu;
  const double *dy2 = vgrid ? v.dyv2i : v.dyu2i;
  const double beta = rk_b(kRK) * v.scal[SC_DT] / ch.Re;
  const double k2x = v.k2x[i], k2z = v.k2z[k];
  const size_t base = (size_t)c * n;
  double *sc = scr + base;
  double cprev = vgrid ? 0.0 : ch.gw1;
  double2 ya = ra[base], yb = rb ? rb[base] : make_double2(0, 0);
  xa[base] = ya;
  if (rb) xb[base] = yb;
  sc[0] = cprev;
  for (int j = 1; j < n; j++) {
    double a1, a2, a3;
    if (j <= n - 2) {
      a1 = -beta * dy2[3 * j];
      a2 = 1.0 - beta * ((dy2[3 * j + 1] + k2x) + k2z);
      a3 = -beta * dy2[3 * j + 2];
    } else {
      a1 = vgrid ? 0.0 : ch.gw2;
      a2 = 1.0;
      a3 = 0.0;
    }
    double d = 1.0 / (a2 - a1 * cprev);
    double2 r = ra[base + j];
    ya = make_double2((r.x - a1 * ya.x) * d, (r.y - a1 * ya.y) * d);
    xa[base + j] = ya;
    if (rb) {
      double2 q = rb[base + j];
      yb = make_double2((q.x - a1 * yb.x) * d, (q.y - a1 * yb.y) * d);
      xb[base + j] = yb;
    }
    cprev = a3 * d;
    sc[j] = cprev;
  }
  for (int j = n - 2; j >= 0; j--) {
    double cj = sc[j];
    double2 x = xa[base + j];
    ya = make_double2(x.x - cj * ya.x, x.y - cj * ya.y);
    xa[base + j] = ya;
    if (rb) {
      double2 w = xb[base + j];
      yb = make_double2(w.x - cj * yb.x, w.y - cj * yb.y);
      xb[base + j] = yb;
    }
  }
}

hipError_t launch_lusol_ref2(char *arena, const Chan *d_chan, int kRK, int vgrid, const double2 *rhs_a,
                             const double2 *rhs_b, double2 *x_a, double2 *x_b, double *scr, int ncol,
                             hipStream_t s) {
  hipLaunchKernelGGL(k_lusol_ref, dim3((ncol + 63) / 64), dim3(64), 0, s, arena, d_chan, kRK, vgrid, rhs_a, rhs_b,
                     x_a, x_b, scr, ncol);
  return hipGetLastError();
}

// LUsolP (tridLU_3D.f90:192-226) on x(1:ny-1, ncol): the Poisson operator of
// LU_buildP with the mean-mode pin, from the factors k_factors eliminated
// from the top row down (dP: inverse pivots, eP: multipliers); the upper
// off-diagonal is recomputed from the grid.  One lane per column.
__global__ void __launch_bounds__(64) k_lusol_p_ref(char *arena, const Chan *chp, double2 *x, int ncol) {
  const Chan ch = *chp;
  const View v(arena, ch);
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncol) return;
  const int ldc = ch.ldc, ns = 1, ne = ch.nyu - 2, n = ne - ns + 1;
  const size_t plane = (size_t)(ch.nyu + PADROWS) * ldc;
  const double *dP = v.fac + (size_t)12 * plane, *eP = dP + plane;
  double2 *col = x + (size_t)c * n - ns;  // col[j], j = ns..ne
  double2 y = col[ne];
  y = make_double2(y.x * dP[(size_t)ne * ldc + c], y.y * dP[(size_t)ne * ldc + c]);
  col[ne] = y;
  for (int q = ne - 1; q >= ns; q--) {
    const double a3 = (c == 0 && q == ns) ? 0.0 : v.dthdyu[q] * v.dthdyv[q];
    const double d = dP[(size_t)q * ldc + c];
    const double2 r = col[q];
    y = make_double2((r.x - a3 * y.x) * d, (r.y - a3 * y.y) * d);
    col[q] = y;
  }
  for (int q = ns + 1; q <= ne; q++) {
    const double e = eP[(size_t)q * ldc + c];
    const double2 r = col[q];
    y = make_double2(r.x - e * y.x, r.y - e * y.y);
    col[q] = y;
  }
}

// LUsol0 (tridLU_3D.f90:228-254) with LU_build0 / LU_dec0 (:409-435,
// :520-537): the real mean-mode U system of substep kRK on x(0:ny), in place.
__global__ void k_lusol0_ref(char *arena, const Chan *chp, int kRK, double *x, double *cs) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const Chan ch = *chp;
  const View v(arena, ch);
  const int n = ch.nyu;
  const double beta = rk_b(kRK) * v.scal[SC_DT] / ch.Re;
  double c = ch.gw1;  // row 0: x0 + gw1 x1, pivot 1
  cs[0] = c;
  for (int j = 1; j < n; j++) {
    double a1, a2, a3;
    if (j <= n - 2) {
      a1 = -beta * v.dyu2i[3 * j];
      a2 = 1.0 - beta * v.dyu2i[3 * j + 1];
      a3 = -beta * v.dyu2i[3 * j + 2];
    } else {
      a1 = ch.gw2;
      a2 = 1.0;
      a3 = 0.0;
    }
    const double d = 1.0 / (a2 - a1 * c);
    c = a3 * d;
    cs[j] = c;
    x[j] = (x[j] - a1 * x[j - 1]) * d;
  }
  for (int j = n - 2; j >= 0; j--) x[j] = x[j] - cs[j] * x[j + 1];
}

hipError_t launch_lusol_p_ref(char *arena, const Chan *d_chan, double2 *x, int ncol, hipStream_t s) {
  hipLaunchKernelGGL(k_lusol_p_ref, dim3((ncol + 63) / 64), dim3(64), 0, s, arena, d_chan, x, ncol);
  return hipGetLastError();
}

hipError_t launch_lusol0_ref(char *arena, const Chan *d_chan, int kRK, double *x, double *scr, hipStream_t s) {
  hipLaunchKernelGGL(k_lusol0_ref, dim3(1), dim3(64), 0, s, arena, d_chan, kRK, x, scr);
  return hipGetLastError();
}

// Column <-> plane transposes.  Columns: x(jlo:jhi, ncol) complex, rows
// contiguous (the reference's band arrays); planes: (Ngal1+2) x Ngal2
// doubles per row, Fortran order, mode (i, k) at xPL(2i+1, krow),
// xPL(2i+2, krow) with krow = k below N2/2 and k + Ngal2 - N2 above (negative
// z modes at the top, modes_to_planes_UVP's k - dk).  A workgroup moves a
// tile of 64 columns x 16 rows through LDS, so both sides are read and
// written in contiguous runs.
constexpr int TR_C = 64, TR_J = 16;

__device__ __forceinline__ size_t plane_pos(int c, int nxh, int N2, int Ng2, int igal) {
  const int i = c % nxh, k = c / nxh;
  const int krow = k < N2 / 2 ? k : k + Ng2 - N2;
  return (size_t)krow * igal + 2 * i;
}

__global__ void __launch_bounds__(256) k_modes_to_planes(int N1, int N2, int Ng1, int Ng2, const double2 *cols,
                                                         int nrows, double *planes) {
  __shared__ double2 tile[TR_C][TR_J + 1];
  const int nxh = N1 / 2 + 1, ncol = nxh * N2, igal = Ng1 + 2;
  const int c0 = blockIdx.x * TR_C, j0 = blockIdx.y * TR_J;
  for (int e = threadIdx.x; e < TR_C * TR_J; e += blockDim.x) {
    const int cl = e / TR_J, jl = e % TR_J, c = c0 + cl, j = j0 + jl;
    if (c < ncol && j < nrows) tile[cl][jl] = cols[(size_t)c * nrows + j];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < TR_C * TR_J; e += blockDim.x) {
    const int jl = e / TR_C, cl = e % TR_C, c = c0 + cl, j = j0 + jl;
    if (c < ncol && j < nrows)
      *(double2 *)(planes + (size_t)j * igal * Ng2 + plane_pos(c, nxh, N2, Ng2, igal)) = tile[cl][jl];
  }
}

__global__ void __launch_bounds__(256) k_planes_to_modes(int N1, int N2, int Ng1, int Ng2, const double *planes,
                                                         int nrows, double2 *cols) {
  __shared__ double2 tile[TR_C][TR_J + 1];
  const int nxh = N1 / 2 + 1, ncol = nxh * N2, igal = Ng1 + 2;
  const int c0 = blockIdx.x * TR_C, j0 = blockIdx.y * TR_J;
  for (int e = threadIdx.x; e < TR_C * TR_J; e += blockDim.x) {
    const int jl = e / TR_C, cl = e % TR_C, c = c0 + cl, j = j0 + jl;
    if (c < ncol && j < nrows)
      tile[cl][jl] = *(const double2 *)(planes + (size_t)j * igal * Ng2 + plane_pos(c, nxh, N2, Ng2, igal));
  }
  __syncthreads();
  for (int e = threadIdx.x; e < TR_C * TR_J; e += blockDim.x) {
    const int cl = e / TR_J, jl = e % TR_J, c = c0 + cl, j = j0 + jl;
    if (c < ncol && j < nrows) cols[(size_t)c * nrows + j] = tile[cl][jl];
  }
}

hipError_t launch_modes_to_planes(int N1, int N2, int Ng1, int Ng2, const double2 *cols, int nrows, double *planes,
                                  hipStream_t s) {
  const int ncol = (N1 / 2 + 1) * N2;
  hipError_t e = hipMemsetAsync(planes, 0, (size_t)nrows * (Ng1 + 2) * Ng2 * sizeof(double), s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_modes_to_planes, dim3((ncol + TR_C - 1) / TR_C, (nrows + TR_J - 1) / TR_J), dim3(256), 0, s,
                     N1, N2, Ng1, Ng2, cols, nrows, planes);
  return hipGetLastError();
}

hipError_t launch_planes_to_modes(int N1, int N2, int Ng1, int Ng2, const double *planes, int nrows, double2 *cols,
                                  hipStream_t s) {
  const int ncol = (N1 / 2 + 1) * N2;
  hipLaunchKernelGGL(k_planes_to_modes, dim3((ncol + TR_C - 1) / TR_C, (nrows + TR_J - 1) / TR_J), dim3(256), 0, s,
                     N1, N2, Ng1, Ng2, planes, nrows, cols);
  return hipGetLastError();
}

}  // namespace lh
