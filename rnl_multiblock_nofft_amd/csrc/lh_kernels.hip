// lh_kernels.hip -- the device side of one RK3 substep of LittleHarsh's plain
// FFT method on gfx950.  Five launches per substep, each covering every
// channel of an ensemble (blockIdx.z selects the channel):
//
//   k_zinv   interp_u/interp_v (FOU3D.f90:427-474) fused with the inverse z
//            transform of the retained x-modes (cft, four_to_phys_u
//            FOU3D.f90:738-764) for the 3 fields of every u- and v-plane
//   k_xprod  inverse x transform (rft), the physical products of
//            ops_in_planes2 (FOU3D.f90:2116-2127,2201-2210), forward x
//            transform (phys_to_four_du FOU3D.f90:2253-2272)
//   k_zfwd   forward z transform, dealiasing truncation and d/dx, d/dz
//            (der_x/der_z FOU3D.f90:476-593), written straight into the
//            mode-major spectral layout (planes_to_modes_* is free)
//   k_rhs    der_yu_h/der_yv_h of the fluxes (FOU3D.f90:636-676), RHS0_u1/
//            u2/u3 (littleharsh.f90:635-788), du update (FOU3D.f90:398-410)
//            and rhs = u + dt*du (solveU_W/solveV, littleharsh.f90:791-907)
//   k_solve  per column: LUsolU_W, LUsolV (tridLU_3D.f90:55-189), the
//            constant-flow-rate correction of mode (0,0) (flowrate_corr,
//            LUsol0, littleharsh.f90:476-509), divergence + LUsolP
//            (littleharsh.f90:581-633, tridLU_3D.f90:192-226,333-407) and
//            v_corr (littleharsh.f90:956-1052)
//
// The nonlinear stages run in FP64 (C = double2) or, in mixed mode, FP32
// (C = float2) with FP64 twiddles; k_rhs and k_solve are always FP64.
#include "lh_common.h"
#include "lh_fft.h"
#include "lh_kernels.h"

namespace lh {

// ------------------------------------------------------------------------
// tile sizes (shared with the host so LDS sizes agree)
__host__ __device__ int tile_zinv(int Ng2, int nxh) {
  int t = 2048 / Ng2;
  if (t < 1) t = 1;
  return t < nxh ? t : nxh;
}
__host__ __device__ int tile_xprod(int Ng1, int Ng2) {
  int t = 512 / Ng1;
  if (t < 1) t = 1;
  return t < Ng2 ? t : Ng2;
}
__host__ __device__ int tile_zfwd(int Ng2, int nxh) {
  int t = 512 / Ng2;
  if (t < 1) t = 1;
  return t < nxh ? t : nxh;
}

template <typename C> struct Inter;
template <> struct Inter<double2> {
  static __device__ double2 *hz(const Chan &c) { return c.hz; }
  static __device__ double2 *gx(const Chan &c) { return c.gx; }
};
template <> struct Inter<float2> {
  static __device__ float2 *hz(const Chan &c) { return c.hzf; }
  static __device__ float2 *gx(const Chan &c) { return c.gxf; }
};
template <typename C> __device__ __forceinline__ C from_d(double2 v) { return CT<C>::mk(v.x, v.y); }
template <typename C> __device__ __forceinline__ double2 to_d(C v) { return make_double2(v.x, v.y); }

// ------------------------------------------------------------------------
// K1: interpolation + inverse z transform
template <typename C>
__global__ void __launch_bounds__(256) k_zinv(const Chan *chans) {
  const Chan &ch = chans[blockIdx.z / 3];
  const int f = blockIdx.z % 3;
  const int slot = blockIdx.y;
  if (slot >= ch.nslot) return;
  const int Ng2 = ch.Ng2, nxh = ch.nxh, N2 = ch.N2, ldc = ch.ldc;
  const int TI = tile_zinv(Ng2, nxh);
  const int i0 = blockIdx.x * TI;
  if (i0 >= nxh) return;
  const int ni = min(TI, nxh - i0);
  extern __shared__ double2 smem_d[];
  C *A = reinterpret_cast<C *>(smem_d);
  C *B = A + TI * Ng2;
  const bool isU = slot < ch.nplu;
  const int j = isU ? slot + 1 : slot - ch.nplu + 1;

  for (int t = threadIdx.x; t < ni * Ng2; t += blockDim.x) A[t] = CT<C>::mk(0, 0);
  __syncthreads();
  // interpolation weights of this plane
  double wa = 0, wb = 0, wd = 1;
  const double2 *s0 = nullptr, *s1 = nullptr;
  if (isU) {
    if (f == 1) {  // interp_v: v faces -> u centre j
      wa = ch.yu[j] - ch.yv[j - 1];
      wb = ch.yv[j] - ch.yu[j];
      wd = ch.yv[j] - ch.yv[j - 1];
      s0 = ch.u2 + (size_t)j * ldc;
      s1 = ch.u2 + (size_t)(j - 1) * ldc;
    } else {
      s0 = (f == 0 ? ch.u1 : ch.u3) + (size_t)j * ldc;
    }
  } else {
    if (f == 1) {
      s0 = ch.u2 + (size_t)j * ldc;
    } else {  // interp_u: u centres -> v face j
      const double2 *u = f == 0 ? ch.u1 : ch.u3;
      wa = ch.yv[j] - ch.yu[j];
      wb = ch.yu[j + 1] - ch.yv[j];
      wd = ch.yu[j + 1] - ch.yu[j];
      s0 = u + (size_t)(j + 1) * ldc;
      s1 = u + (size_t)j * ldc;
    }
  }
  const int half = N2 / 2, dz = Ng2 - N2;
  for (int t = threadIdx.x; t < N2 * ni; t += blockDim.x) {
    int k = t / ni, ii = t - k * ni;
    int col = k * nxh + i0 + ii;
    double2 v = s0[col];
    if (s1) {
      double2 w = s1[col];
      v = make_double2((wa * v.x + wb * w.x) / wd, (wa * v.y + wb * w.y) / wd);
    }
    int z = k < half ? k : k + dz;
    A[ii * Ng2 + z] = from_d<C>(v);
  }
  __syncthreads();
  C *R = fft_lds<C>(A, B, Ng2, ni, Ng2, +1, ch.twz);
  C *hz = Inter<C>::hz(ch) + ((size_t)slot * 3 + f) * Ng2 * nxh;
  for (int t = threadIdx.x; t < ni * Ng2; t += blockDim.x) {
    int z = t / ni, ii = t - z * ni;
    hz[(size_t)z * nxh + i0 + ii] = R[ii * Ng2 + z];
  }
}

// ------------------------------------------------------------------------
// K2: inverse x transform, products, forward x transform
template <typename C>
__global__ void __launch_bounds__(256) k_xprod(const Chan *chans) {
  typedef typename CT<C>::R R_t;
  const Chan &ch = chans[blockIdx.z];
  const int slot = blockIdx.y;
  if (slot >= ch.nslot) return;
  const int Ng1 = ch.Ng1, Ng2 = ch.Ng2, nxh = ch.nxh, h1 = ch.N1 / 2;
  const int TZ = tile_xprod(Ng1, Ng2);
  const int z0 = blockIdx.x * TZ;
  if (z0 >= Ng2) return;
  const int nz = min(TZ, Ng2 - z0);
  extern __shared__ double2 smem_d[];
  C *A = reinterpret_cast<C *>(smem_d);
  C *B = A + 2 * TZ * Ng1;
  const bool isU = slot < ch.nplu;
  const C *hz = Inter<C>::hz(ch) + (size_t)slot * 3 * Ng2 * nxh;

  // Hermitian extension of the three half spectra, packed two per sequence:
  // seq 2zz = f0 + i f2, seq 2zz+1 = f1
  for (int t = threadIdx.x; t < nz * Ng1; t += blockDim.x) {
    int zz = t / Ng1, m = t - zz * Ng1;
    int z = z0 + zz;
    C F0 = CT<C>::mk(0, 0), F1 = F0, F2 = F0;
    if (m <= h1) {
      F0 = hz[((size_t)0 * Ng2 + z) * nxh + m];
      F1 = hz[((size_t)1 * Ng2 + z) * nxh + m];
      F2 = hz[((size_t)2 * Ng2 + z) * nxh + m];
      if (m == 0) {  // rft ignores imag(cc(0)) on the inverse transform
        F0.y = 0;
        F1.y = 0;
        F2.y = 0;
      }
    } else if (m >= Ng1 - h1) {
      int mm = Ng1 - m;
      F0 = hz[((size_t)0 * Ng2 + z) * nxh + mm];
      F1 = hz[((size_t)1 * Ng2 + z) * nxh + mm];
      F2 = hz[((size_t)2 * Ng2 + z) * nxh + mm];
      F0.y = -F0.y;
      F1.y = -F1.y;
      F2.y = -F2.y;
    }
    A[(2 * zz) * Ng1 + m] = CT<C>::mk(F0.x - F2.y, F0.y + F2.x);
    A[(2 * zz + 1) * Ng1 + m] = F1;
  }
  __syncthreads();
  C *X = fft_lds<C>(A, B, Ng1, 2 * nz, Ng1, +1, ch.twx);
  for (int t = threadIdx.x; t < nz * Ng1; t += blockDim.x) {
    int zz = t / Ng1, n = t - zz * Ng1;
    C s0 = X[(2 * zz) * Ng1 + n], s1 = X[(2 * zz + 1) * Ng1 + n];
    R_t a1 = s0.x, a3 = s0.y, a2 = s1.x;
    if (isU) {
      X[(2 * zz) * Ng1 + n] = CT<C>::mk(a1 * a1, a1 * a3);      // uu + i uw
      X[(2 * zz + 1) * Ng1 + n] = CT<C>::mk(a2 * a2, a3 * a3);  // vv + i ww
    } else {
      X[(2 * zz) * Ng1 + n] = CT<C>::mk(a1 * a2, a3 * a2);      // uv + i wv
      X[(2 * zz + 1) * Ng1 + n] = CT<C>::mk(0, 0);
    }
  }
  __syncthreads();
  C *Y = fft_lds<C>(X, X == A ? B : A, Ng1, 2 * nz, Ng1, -1, ch.twx);
  const int nseq = isU ? 2 : 1;
  const R_t inv = (R_t)(1.0 / Ng1);
  C *gx = Inter<C>::gx(ch) + (size_t)slot * 4 * Ng2 * nxh;
  for (int t = threadIdx.x; t < nz * nseq * (h1 + 1); t += blockDim.x) {
    int m = t % (h1 + 1);
    int r = t / (h1 + 1);
    int q = r % nseq, zz = r / nseq;
    const C *ys = Y + (2 * zz + q) * Ng1;
    C Ym = ys[m], Yc = ys[m == 0 ? 0 : Ng1 - m];
    C P, Q;
    if (m == 0) {
      P = CT<C>::mk(Ym.x * inv, 0);
      Q = CT<C>::mk(Ym.y * inv, 0);
    } else {
      // P = (Y_m + conj(Y_{n-m}))/2, Q = (Y_m - conj(Y_{n-m}))/(2i)
      P = CT<C>::mk((R_t)0.5 * (Ym.x + Yc.x) * inv, (R_t)0.5 * (Ym.y - Yc.y) * inv);
      Q = CT<C>::mk((R_t)0.5 * (Ym.y + Yc.y) * inv, -(R_t)0.5 * (Ym.x - Yc.x) * inv);
    }
    int z = z0 + zz;
    gx[((size_t)(2 * q) * Ng2 + z) * nxh + m] = P;
    gx[((size_t)(2 * q + 1) * Ng2 + z) * nxh + m] = Q;
  }
}

// ------------------------------------------------------------------------
// K3: forward z transform, truncation, d/dx + d/dz
template <typename C>
__global__ void __launch_bounds__(256) k_zfwd(const Chan *chans) {
  const Chan &ch = chans[blockIdx.z];
  const int slot = blockIdx.y;
  if (slot >= ch.nslot) return;
  const int Ng2 = ch.Ng2, nxh = ch.nxh, N2 = ch.N2, ldc = ch.ldc;
  const int TI = tile_zfwd(Ng2, nxh);
  const int i0 = blockIdx.x * TI;
  if (i0 >= nxh) return;
  const int ni = min(TI, nxh - i0);
  const bool isU = slot < ch.nplu;
  const int np = isU ? 4 : 2;
  const int j = isU ? slot + 1 : slot - ch.nplu + 1;
  extern __shared__ double2 smem_d[];
  C *A = reinterpret_cast<C *>(smem_d);
  C *B = A + 4 * TI * Ng2;
  const C *gx = Inter<C>::gx(ch) + (size_t)slot * 4 * Ng2 * nxh;
  for (int t = threadIdx.x; t < np * ni * Ng2; t += blockDim.x) {
    int p = t / (ni * Ng2);
    int rem = t - p * ni * Ng2;
    int z = rem / ni, ii = rem - z * ni;
    A[(p * ni + ii) * Ng2 + z] = gx[((size_t)p * Ng2 + z) * nxh + i0 + ii];
  }
  __syncthreads();
  C *R = fft_lds<C>(A, B, Ng2, np * ni, Ng2, -1, ch.twz);
  const double inv = 1.0 / Ng2;
  const int half = N2 / 2, dz = Ng2 - N2;
  for (int t = threadIdx.x; t < ni * N2; t += blockDim.x) {
    int k = t / ni, ii = t - k * ni;
    int i = i0 + ii;
    int col = k * nxh + i;
    int z = k < half ? k : k + dz;
    bool live = k != half;  // kz = -N2/2 is zeroed by der_x/der_z
    double kxi = ch.kx[i], kzk = ch.kz[k];
    size_t o = (size_t)j * ldc + col;
    double2 X0 = cscale(inv, to_d(R[(0 * ni + ii) * Ng2 + z]));
    double2 X1 = cscale(inv, to_d(R[(1 * ni + ii) * Ng2 + z]));
    if (isU) {
      double2 X2 = cscale(inv, to_d(R[(2 * ni + ii) * Ng2 + z]));
      double2 X3 = cscale(inv, to_d(R[(3 * ni + ii) * Ng2 + z]));
      ch.nl1[o] = live ? cadd(cimul(kxi, X0), cimul(kzk, X1)) : make_double2(0, 0);
      ch.nl3[o] = live ? cadd(cimul(kxi, X1), cimul(kzk, X3)) : make_double2(0, 0);
      ch.fvv[o] = X2;
    } else {
      ch.nl2[o] = live ? cadd(cimul(kxi, X0), cimul(kzk, X1)) : make_double2(0, 0);
      ch.fuv[o] = X0;
      ch.fwv[o] = X1;
    }
  }
}

// ------------------------------------------------------------------------
// K4: right-hand sides
__global__ void __launch_bounds__(256) k_rhs(const Chan *chans, int kRK) {
  const Chan &ch = chans[blockIdx.z];
  const int j = blockIdx.y;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ch.nyu || c >= ch.ncol) return;
  const int nxh = ch.nxh, ldc = ch.ldc, nyu = ch.nyu, nyv = ch.nyv;
  const int i = c % nxh, k = c / nxh;
  const double dt = ch.scal[SC_DT];
  const double C1 = rk_a(kRK) / ch.Re, C2 = -rk_c(kRK), Cg = -rk_g(kRK), dR = rk_d(kRK);
  const size_t o = (size_t)j * ldc + c;
  const double k2x = ch.k2x[i], k2z = ch.k2z[k];
  if (j >= 1 && j <= nyu - 2) {
    const double a0 = ch.dyu2i[3 * j], a1 = ch.dyu2i[3 * j + 1], a2 = ch.dyu2i[3 * j + 2];
    const double dd = (k2x + k2z) + a1;
    const double dyu = ch.dthdyu[j];
    {
      double2 um = ch.u1[o - ldc], u0 = ch.u1[o], up = ch.u1[o + ldc];
      double2 L = make_double2(a0 * um.x + dd * u0.x + a2 * up.x, a0 * um.y + dd * u0.y + a2 * up.y);
      double2 fm = ch.fuv[o - ldc], f0 = ch.fuv[o];
      double2 nl = ch.nl1[o];
      double2 Nn = make_double2(nl.x + (f0.x - fm.x) * dyu, nl.y + (f0.y - fm.y) * dyu);
      double2 No = ch.N1o[o];
      double2 pc = cimul(-dR * ch.kx[i], ch.p[o]);
      double2 du = make_double2(C1 * L.x + C2 * No.x + pc.x, C1 * L.y + C2 * No.y + pc.y);
      if (c == 0) du.x = du.x + (-dR * ch.scal[SC_MPGX]);
      du = make_double2(du.x + Cg * Nn.x, du.y + Cg * Nn.y);
      ch.r1[o] = make_double2(u0.x + dt * du.x, u0.y + dt * du.y);
      ch.N1o[o] = Nn;
    }
    {
      double2 um = ch.u3[o - ldc], u0 = ch.u3[o], up = ch.u3[o + ldc];
      double2 L = make_double2(a0 * um.x + dd * u0.x + a2 * up.x, a0 * um.y + dd * u0.y + a2 * up.y);
      double2 fm = ch.fwv[o - ldc], f0 = ch.fwv[o];
      double2 nl = ch.nl3[o];
      double2 Nn = make_double2(nl.x + (f0.x - fm.x) * dyu, nl.y + (f0.y - fm.y) * dyu);
      double2 No = ch.N3o[o];
      double2 pc = cimul(-dR * ch.kz[k], ch.p[o]);
      double2 du = make_double2(C1 * L.x + C2 * No.x + pc.x, C1 * L.y + C2 * No.y + pc.y);
      du = make_double2(du.x + Cg * Nn.x, du.y + Cg * Nn.y);
      ch.r3[o] = make_double2(u0.x + dt * du.x, u0.y + dt * du.y);
      ch.N3o[o] = Nn;
    }
  } else {
    ch.r1[o] = make_double2(0, 0);
    ch.r3[o] = make_double2(0, 0);
  }
  if (j >= 1 && j <= nyv - 2) {
    const double a0 = ch.dyv2i[3 * j], a1 = ch.dyv2i[3 * j + 1], a2 = ch.dyv2i[3 * j + 2];
    const double dd = (k2x + k2z) + a1;
    const double dyv = ch.dthdyv[j];
    double2 um = ch.u2[o - ldc], u0 = ch.u2[o], up = ch.u2[o + ldc];
    double2 L = make_double2(a0 * um.x + dd * u0.x + a2 * up.x, a0 * um.y + dd * u0.y + a2 * up.y);
    double2 f0 = ch.fvv[o], fp = ch.fvv[o + ldc];
    double2 nl = ch.nl2[o];
    double2 Nn = make_double2(nl.x + (fp.x - f0.x) * dyv, nl.y + (fp.y - f0.y) * dyv);
    double2 No = ch.N2o[o];
    double C3 = -dR * dyv;
    double2 p0 = ch.p[o], pp = ch.p[o + ldc];
    double2 du = make_double2(C1 * L.x + C2 * No.x + C3 * (pp.x - p0.x),
                              C1 * L.y + C2 * No.y + C3 * (pp.y - p0.y));
    du = make_double2(du.x + Cg * Nn.x, du.y + Cg * Nn.y);
    ch.r2[o] = make_double2(u0.x + dt * du.x, u0.y + dt * du.y);
    ch.N2o[o] = Nn;
  } else if (j == 0 || j == nyv - 1) {
    ch.r2[o] = make_double2(0, 0);
  }
}

// ------------------------------------------------------------------------
// K5: column solves.  One thread per spectral column, sequential in y.
__global__ void __launch_bounds__(64) k_solve(const Chan *chans, int kRK) {
  const Chan &ch = chans[blockIdx.z];
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ch.ncol) return;
  const int nxh = ch.nxh, ldc = ch.ldc, nyu = ch.nyu, nyv = ch.nyv;
  const int i = c % nxh, k = c / nxh;
  const double dt = ch.scal[SC_DT];
  const double beta = rk_b(kRK) * dt / ch.Re;
  const double k2x = ch.k2x[i], k2z = ch.k2z[k];
  double *scr = ch.scr;

  // ---- LUsolU_W: u1 and u3 share the matrix (LU_buildU + LU_dec)
  {
    double cprev = ch.gw1;  // a(3,0)*a(2,0) with a(2,0)=1
    double2 y1 = ch.r1[c], y3 = ch.r3[c];
    ch.u1[c] = y1;
    ch.u3[c] = y3;
    scr[c] = cprev;
    for (int j = 1; j < nyu; j++) {
      double a1, a2, a3;
      if (j <= nyu - 2) {
        a1 = -beta * ch.dyu2i[3 * j];
        a2 = 1.0 - beta * ((ch.dyu2i[3 * j + 1] + k2x) + k2z);
        a3 = -beta * ch.dyu2i[3 * j + 2];
      } else {
        a1 = ch.gw2;
        a2 = 1.0;
        a3 = 0.0;
      }
      double d = 1.0 / (a2 - a1 * cprev);
      size_t o = (size_t)j * ldc + c;
      double2 r1 = ch.r1[o], r3 = ch.r3[o];
      y1 = make_double2((r1.x - a1 * y1.x) * d, (r1.y - a1 * y1.y) * d);
      y3 = make_double2((r3.x - a1 * y3.x) * d, (r3.y - a1 * y3.y) * d);
      ch.u1[o] = y1;
      ch.u3[o] = y3;
      cprev = a3 * d;
      scr[o] = cprev;
    }
    for (int j = nyu - 2; j >= 0; j--) {
      size_t o = (size_t)j * ldc + c;
      double cj = scr[o];
      double2 x1 = ch.u1[o], x3 = ch.u3[o];
      y1 = make_double2(x1.x - cj * y1.x, x1.y - cj * y1.y);
      y3 = make_double2(x3.x - cj * y3.x, x3.y - cj * y3.y);
      ch.u1[o] = y1;
      ch.u3[o] = y3;
    }
  }
  // ---- LUsolV (LU_buildV: identity wall rows)
  {
    double cprev = 0.0;
    double2 y = ch.r2[c];
    ch.u2[c] = y;
    scr[c] = 0.0;
    for (int j = 1; j < nyv; j++) {
      double a1, a2, a3;
      if (j <= nyv - 2) {
        a1 = -beta * ch.dyv2i[3 * j];
        a2 = 1.0 - beta * ((ch.dyv2i[3 * j + 1] + k2x) + k2z);
        a3 = -beta * ch.dyv2i[3 * j + 2];
      } else {
        a1 = 0.0;
        a2 = 1.0;
        a3 = 0.0;
      }
      double d = 1.0 / (a2 - a1 * cprev);
      size_t o = (size_t)j * ldc + c;
      double2 r = ch.r2[o];
      y = make_double2((r.x - a1 * y.x) * d, (r.y - a1 * y.y) * d);
      ch.u2[o] = y;
      cprev = a3 * d;
      scr[o] = cprev;
    }
    for (int j = nyv - 2; j >= 0; j--) {
      size_t o = (size_t)j * ldc + c;
      double cj = scr[o];
      double2 x = ch.u2[o];
      y = make_double2(x.x - cj * y.x, x.y - cj * y.y);
      ch.u2[o] = y;
    }
  }
  // ---- meanflow_ctU / flowrate_corr on mode (0,0)
  if (c == 0 && ch.flag_ctpress == 0) {
    double *u11 = ch.u11;
    double bdt = -rk_d(kRK) * dt;
    double cprev = ch.gw1;
    double y = 0.0;  // u11(0) = 0
    u11[0] = y;
    scr[0] = cprev;
    for (int j = 1; j < nyu; j++) {
      double a1, a2, a3, rhs;
      if (j <= nyu - 2) {
        a1 = -beta * ch.dyu2i[3 * j];
        a2 = 1.0 - beta * ch.dyu2i[3 * j + 1];
        a3 = -beta * ch.dyu2i[3 * j + 2];
        rhs = bdt;
      } else {
        a1 = ch.gw2;
        a2 = 1.0;
        a3 = 0.0;
        rhs = 0.0;
      }
      double d = 1.0 / (a2 - a1 * cprev);
      y = (rhs - a1 * y) * d;
      u11[j] = y;
      cprev = a3 * d;
      scr[(size_t)j * ldc] = cprev;
    }
    for (int j = nyu - 2; j >= 0; j--) {
      y = u11[j] - scr[(size_t)j * ldc] * y;
      u11[j] = y;
    }
    double Qx1 = 0.0, Qx0 = 0.0;
    for (int j = 1; j <= nyu - 2; j++) Qx1 = Qx1 + (ch.yv[j] - ch.yv[j - 1]) * u11[j];
    for (int j = 1; j <= nyu - 2; j++) Qx0 = Qx0 + (ch.yv[j] - ch.yv[j - 1]) * ch.u1[(size_t)j * ldc].x;
    double g = (ch.scal[SC_QXT] - Qx0) / Qx1;
    ch.scal[SC_MPGX] = ch.scal[SC_MPGX] + g;
    for (int j = 0; j < nyu; j++) {
      size_t o = (size_t)j * ldc;
      ch.u1[o].x = ch.u1[o].x + g * u11[j];
    }
  }
  // ---- solveP: divergence, psi rhs, LUsolP with the LU_buildP operator
  {
    const double dtirk = (1.0 / dt) / (rk_a(kRK) + rk_b(kRK));
    const double kxi = ch.kx[i], kzk = ch.kz[k];
    const int ns = 1, ne = nyu - 2;
    double cprev = 0.0;
    double2 y = make_double2(0, 0);
    for (int j = ns; j <= ne; j++) {
      size_t o = (size_t)j * ldc + c;
      double2 a = cimul(kxi, ch.u1[o]), b = cimul(kzk, ch.u3[o]);
      double2 v0 = ch.u2[o], vm = ch.u2[o - ldc];
      double w = ch.dthdyu[j];
      double2 dv = make_double2((a.x + b.x) + w * (v0.x - vm.x), (a.y + b.y) + w * (v0.y - vm.y));
      double2 rhs = make_double2(dtirk * dv.x, dtirk * dv.y);
      if (c == 0 && j == ns) rhs = make_double2(0, 0);
      double a1, a2, a3;
      const double du_j = ch.dthdyu[j];
      if (c == 0 && j == ns) {
        double h = ch.yu[2] - ch.yu[1];
        a1 = 0.0;
        a2 = 1.0 / (h * h);
        a3 = 0.0;
      } else if (j == ns) {
        a1 = 0.0;
        a2 = ((du_j * (-ch.dthdyv[j])) + k2x) + k2z;
        a3 = du_j * ch.dthdyv[j];
      } else if (j == ne) {
        a1 = (-du_j) * (-ch.dthdyv[j - 1]);
        a2 = (((-du_j) * ch.dthdyv[j - 1]) + k2x) + k2z;
        a3 = 0.0;
      } else {
        a1 = (-du_j) * (-ch.dthdyv[j - 1]);
        a2 = ((((-du_j) * ch.dthdyv[j - 1]) + (du_j * (-ch.dthdyv[j]))) + k2x) + k2z;
        a3 = du_j * ch.dthdyv[j];
      }
      double d = 1.0 / (a2 - a1 * cprev);
      if (j == ns) d = 1.0 / a2;
      y = make_double2((rhs.x - a1 * y.x) * d, (rhs.y - a1 * y.y) * d);
      ch.psi[o] = y;
      cprev = a3 * d;
      scr[o] = cprev;
    }
    for (int j = ne - 1; j >= ns; j--) {
      size_t o = (size_t)j * ldc + c;
      double cj = scr[o];
      double2 x = ch.psi[o];
      y = make_double2(x.x - cj * y.x, x.y - cj * y.y);
      ch.psi[o] = y;
    }
  }
  // ---- p += psi, v_corr, divergence
  {
    const int ns = 1, ne = nyu - 2;
    const double dtrk = dt * (rk_a(kRK) + rk_b(kRK));
    const double dkx = dtrk * ch.kx[i], dkz = dtrk * ch.kz[k];
    for (int j = ns; j <= ne; j++) {
      size_t o = (size_t)j * ldc + c;
      double2 ps = ch.psi[o];
      double2 pp = ch.p[o];
      ch.p[o] = make_double2(pp.x + ps.x, pp.y + ps.y);
      double2 u = ch.u1[o], a = cimul(dkx, ps);
      ch.u1[o] = make_double2(u.x - a.x, u.y - a.y);
      double2 w = ch.u3[o], b = cimul(dkz, ps);
      ch.u3[o] = make_double2(w.x - b.x, w.y - b.y);
    }
    for (int j = 1; j <= nyv - 2; j++) {
      size_t o = (size_t)j * ldc + c;
      double2 ps = ch.psi[o], pp = ch.psi[o + ldc];
      double s = dtrk * ch.dthdyv[j];
      double2 v = ch.u2[o];
      ch.u2[o] = make_double2(v.x - s * (pp.x - ps.x), v.y - s * (pp.y - ps.y));
    }
    {
      double2 a = ch.u1[(size_t)1 * ldc + c], b = ch.u1[(size_t)(nyu - 2) * ldc + c];
      ch.u1[c] = cscale(ch.gwbc1, a);
      ch.u1[(size_t)(nyu - 1) * ldc + c] = cscale(ch.gwbc1, b);
      double2 a3 = ch.u3[(size_t)1 * ldc + c], b3 = ch.u3[(size_t)(nyu - 2) * ldc + c];
      ch.u3[c] = cscale(ch.gwbc3, a3);
      ch.u3[(size_t)(nyu - 1) * ldc + c] = cscale(ch.gwbc3, b3);
      ch.u2[c] = make_double2(0, 0);
      ch.u2[(size_t)(nyv - 1) * ldc + c] = make_double2(0, 0);
    }
    const double kxi = ch.kx[i], kzk = ch.kz[k];
    for (int j = ns; j <= ne; j++) {
      size_t o = (size_t)j * ldc + c;
      double2 a = cimul(kxi, ch.u1[o]), b = cimul(kzk, ch.u3[o]);
      double2 v0 = ch.u2[o], vm = ch.u2[o - ldc];
      double w = ch.dthdyu[j];
      ch.div[o] = make_double2((a.x + b.x) + w * (v0.x - vm.x), (a.y + b.y) + w * (v0.y - vm.y));
    }
  }
}

// ------------------------------------------------------------------------
// flow rate of the (0,0) mode of every channel (flowrateIm)
__global__ void k_flowrate(const Chan *chans, int nchan, int set_target) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nchan) return;
  const Chan &ch = chans[c];
  double Q = 0.0;
  for (int j = 1; j <= ch.nyu - 2; j++) Q = Q + (ch.yv[j] - ch.yv[j - 1]) * ch.u1[(size_t)j * ch.ldc].x;
  ch.scal[SC_QX] = Q;
  if (set_target && ch.flag_ctpress != 1) ch.scal[SC_QXT] = Q;
}

// ------------------------------------------------------------------------
// reference layout (ncol x nrows, y contiguous) <-> mode-major rows
__global__ void k_to_rows(const double2 *__restrict__ src, double2 *__restrict__ dst, int ncol,
                          int nrows, int jlo, int ldc) {
  __shared__ double2 tile[32][33];
  int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
  int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int q = ty; q < 32; q += 8) {
    int c = c0 + q, r = r0 + tx;
    if (c < ncol && r < nrows) tile[q][tx] = src[(size_t)c * nrows + r];
  }
  __syncthreads();
  for (int q = ty; q < 32; q += 8) {
    int r = r0 + q, c = c0 + tx;
    if (c < ncol && r < nrows) dst[(size_t)(jlo + r) * ldc + c] = tile[tx][q];
  }
}
__global__ void k_from_rows(const double2 *__restrict__ src, double2 *__restrict__ dst, int ncol,
                            int nrows, int jlo, int ldc) {
  __shared__ double2 tile[32][33];
  int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
  int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int q = ty; q < 32; q += 8) {
    int r = r0 + q, c = c0 + tx;
    if (c < ncol && r < nrows) tile[q][tx] = src[(size_t)(jlo + r) * ldc + c];
  }
  __syncthreads();
  for (int q = ty; q < 32; q += 8) {
    int c = c0 + q, r = r0 + tx;
    if (c < ncol && r < nrows) dst[(size_t)c * nrows + r] = tile[tx][q];
  }
}

// ------------------------------------------------------------------------
// launch wrappers
static size_t lds_zinv(int Ng2, int nxh, size_t es) { return 2 * (size_t)tile_zinv(Ng2, nxh) * Ng2 * es; }
static size_t lds_xprod(int Ng1, int Ng2, size_t es) { return 2 * 2 * (size_t)tile_xprod(Ng1, Ng2) * Ng1 * es; }
static size_t lds_zfwd(int Ng2, int nxh, size_t es) { return 2 * 4 * (size_t)tile_zfwd(Ng2, nxh) * Ng2 * es; }

hipError_t launch_nl_stage(const Chan *d_chans, const ShapeMax &m, bool mixed, int stage, hipStream_t s) {
  size_t es = mixed ? sizeof(float2) : sizeof(double2);
  dim3 b(256);
  if (stage == 0) {
    dim3 g(m.zinv_tiles, m.nslot, m.nchan * 3);
    if (mixed)
      hipLaunchKernelGGL(k_zinv<float2>, g, b, m.lds_zinv * es, s, d_chans);
    else
      hipLaunchKernelGGL(k_zinv<double2>, g, b, m.lds_zinv * es, s, d_chans);
  } else if (stage == 1) {
    dim3 g(m.xprod_tiles, m.nslot, m.nchan);
    if (mixed)
      hipLaunchKernelGGL(k_xprod<float2>, g, b, m.lds_xprod * es, s, d_chans);
    else
      hipLaunchKernelGGL(k_xprod<double2>, g, b, m.lds_xprod * es, s, d_chans);
  } else {
    dim3 g(m.zfwd_tiles, m.nslot, m.nchan);
    if (mixed)
      hipLaunchKernelGGL(k_zfwd<float2>, g, b, m.lds_zfwd * es, s, d_chans);
    else
      hipLaunchKernelGGL(k_zfwd<double2>, g, b, m.lds_zfwd * es, s, d_chans);
  }
  return hipGetLastError();
}

hipError_t launch_nonlinear(const Chan *d_chans, const ShapeMax &m, bool mixed, hipStream_t s) {
  for (int st = 0; st < 3; st++) {
    hipError_t e = launch_nl_stage(d_chans, m, mixed, st, s);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

void shape_add(ShapeMax &m, int N1, int N2, int Ng1, int Ng2, int ny) {
  int nxh = N1 / 2 + 1, ncol = nxh * N2;
  int nslot = (ny - 1) + (ny - 2);
  auto mx = [](int &a, int b) { if (b > a) a = b; };
  auto mxs = [](size_t &a, size_t b) { if (b > a) a = b; };
  mx(m.nslot, nslot);
  mx(m.zinv_tiles, (nxh + tile_zinv(Ng2, nxh) - 1) / tile_zinv(Ng2, nxh));
  mx(m.xprod_tiles, (Ng2 + tile_xprod(Ng1, Ng2) - 1) / tile_xprod(Ng1, Ng2));
  mx(m.zfwd_tiles, (nxh + tile_zfwd(Ng2, nxh) - 1) / tile_zfwd(Ng2, nxh));
  mx(m.ncol, ncol);
  mx(m.nyu, ny + 1);
  mxs(m.lds_zinv, lds_zinv(Ng2, nxh, 1));
  mxs(m.lds_xprod, lds_xprod(Ng1, Ng2, 1));
  mxs(m.lds_zfwd, lds_zfwd(Ng2, nxh, 1));
  m.nchan++;
}

hipError_t launch_rhs(const Chan *d_chans, const ShapeMax &m, int kRK, hipStream_t s) {
  dim3 g((m.ncol + 255) / 256, m.nyu, m.nchan);
  hipLaunchKernelGGL(k_rhs, g, dim3(256), 0, s, d_chans, kRK);
  return hipGetLastError();
}

hipError_t launch_solve(const Chan *d_chans, const ShapeMax &m, int kRK, hipStream_t s) {
  dim3 g((m.ncol + 63) / 64, 1, m.nchan);
  hipLaunchKernelGGL(k_solve, g, dim3(64), 0, s, d_chans, kRK);
  return hipGetLastError();
}

hipError_t launch_flowrate(const Chan *d_chans, int nchan, int set_target, hipStream_t s) {
  hipLaunchKernelGGL(k_flowrate, dim3((nchan + 63) / 64), dim3(64), 0, s, d_chans, nchan, set_target);
  return hipGetLastError();
}

hipError_t launch_to_rows(const double2 *src, double2 *dst, int ncol, int nrows, int jlo, int ldc,
                          hipStream_t s) {
  dim3 g((ncol + 31) / 32, (nrows + 31) / 32);
  hipLaunchKernelGGL(k_to_rows, g, dim3(256), 0, s, src, dst, ncol, nrows, jlo, ldc);
  return hipGetLastError();
}
hipError_t launch_from_rows(const double2 *src, double2 *dst, int ncol, int nrows, int jlo, int ldc,
                            hipStream_t s) {
  dim3 g((ncol + 31) / 32, (nrows + 31) / 32);
  hipLaunchKernelGGL(k_from_rows, g, dim3(256), 0, s, src, dst, ncol, nrows, jlo, ldc);
  return hipGetLastError();
}

}  // namespace lh
