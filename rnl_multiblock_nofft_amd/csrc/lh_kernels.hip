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
  static __device__ double2 *hz(const View &c) { return c.hz; }
  static __device__ double2 *gx(const View &c) { return c.gx; }
};
template <> struct Inter<float2> {
  static __device__ float2 *hz(const View &c) { return c.hzf; }
  static __device__ float2 *gx(const View &c) { return c.gxf; }
};
template <typename C> __device__ __forceinline__ C from_d(double2 v) { return CT<C>::mk(v.x, v.y); }
template <typename C> __device__ __forceinline__ double2 to_d(C v) { return make_double2(v.x, v.y); }

// ------------------------------------------------------------------------
// K1: interpolation + inverse z transform
template <typename C>
__global__ void __launch_bounds__(256) k_zinv(char *arena, const Chan *chans) {
  const Chan ch = chans[blockIdx.z / 3];
  const View v(arena, ch);
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
      wa = v.yu[j] - v.yv[j - 1];
      wb = v.yv[j] - v.yu[j];
      wd = v.yv[j] - v.yv[j - 1];
      s0 = v.u2 + (size_t)j * ldc;
      s1 = v.u2 + (size_t)(j - 1) * ldc;
    } else {
      s0 = (f == 0 ? v.u1 : v.u3) + (size_t)j * ldc;
    }
  } else {
    if (f == 1) {
      s0 = v.u2 + (size_t)j * ldc;
    } else {  // interp_u: u centres -> v face j
      const double2 *u = f == 0 ? v.u1 : v.u3;
      wa = v.yv[j] - v.yu[j];
      wb = v.yu[j + 1] - v.yv[j];
      wd = v.yu[j + 1] - v.yu[j];
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
  C *R = fft_lds<C>(A, B, Ng2, ni, Ng2, +1, v.twz);
  C *hz = Inter<C>::hz(v) + ((size_t)slot * 3 + f) * Ng2 * nxh;
  for (int t = threadIdx.x; t < ni * Ng2; t += blockDim.x) {
    int z = t / ni, ii = t - z * ni;
    hz[(size_t)z * nxh + i0 + ii] = R[ii * Ng2 + z];
  }
}

// ------------------------------------------------------------------------
// K2: inverse x transform, products, forward x transform
template <typename C>
__global__ void __launch_bounds__(256) k_xprod(char *arena, const Chan *chans) {
  typedef typename CT<C>::R R_t;
  const Chan ch = chans[blockIdx.z];
  const View v(arena, ch);
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
  const C *hz = Inter<C>::hz(v) + (size_t)slot * 3 * Ng2 * nxh;

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
  C *X = fft_lds<C>(A, B, Ng1, 2 * nz, Ng1, +1, v.twx);
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
  C *Y = fft_lds<C>(X, X == A ? B : A, Ng1, 2 * nz, Ng1, -1, v.twx);
  const int nseq = isU ? 2 : 1;
  const R_t inv = (R_t)(1.0 / Ng1);
  C *gx = Inter<C>::gx(v) + (size_t)slot * 4 * Ng2 * nxh;
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
__global__ void __launch_bounds__(256) k_zfwd(char *arena, const Chan *chans) {
  const Chan ch = chans[blockIdx.z];
  const View v(arena, ch);
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
  const C *gx = Inter<C>::gx(v) + (size_t)slot * 4 * Ng2 * nxh;
  for (int t = threadIdx.x; t < np * ni * Ng2; t += blockDim.x) {
    int p = t / (ni * Ng2);
    int rem = t - p * ni * Ng2;
    int z = rem / ni, ii = rem - z * ni;
    A[(p * ni + ii) * Ng2 + z] = gx[((size_t)p * Ng2 + z) * nxh + i0 + ii];
  }
  __syncthreads();
  C *R = fft_lds<C>(A, B, Ng2, np * ni, Ng2, -1, v.twz);
  const double inv = 1.0 / Ng2;
  const int half = N2 / 2, dz = Ng2 - N2;
  for (int t = threadIdx.x; t < ni * N2; t += blockDim.x) {
    int k = t / ni, ii = t - k * ni;
    int i = i0 + ii;
    int col = k * nxh + i;
    int z = k < half ? k : k + dz;
    bool live = k != half;  // kz = -N2/2 is zeroed by der_x/der_z
    double kxi = v.kx[i], kzk = v.kz[k];
    size_t o = (size_t)j * ldc + col;
    double2 X0 = cscale(inv, to_d(R[(0 * ni + ii) * Ng2 + z]));
    double2 X1 = cscale(inv, to_d(R[(1 * ni + ii) * Ng2 + z]));
    if (isU) {
      double2 X2 = cscale(inv, to_d(R[(2 * ni + ii) * Ng2 + z]));
      double2 X3 = cscale(inv, to_d(R[(3 * ni + ii) * Ng2 + z]));
      v.nl1[o] = live ? cadd(cimul(kxi, X0), cimul(kzk, X1)) : make_double2(0, 0);
      v.nl3[o] = live ? cadd(cimul(kxi, X1), cimul(kzk, X3)) : make_double2(0, 0);
      v.fvv[o] = X2;
    } else {
      v.nl2[o] = live ? cadd(cimul(kxi, X0), cimul(kzk, X1)) : make_double2(0, 0);
      v.fuv[o] = X0;
      v.fwv[o] = X1;
    }
  }
}

// ------------------------------------------------------------------------
// K4: right-hand sides
__global__ void __launch_bounds__(256) k_rhs(char *arena, const Chan *chans, int kRK) {
  const Chan ch = chans[blockIdx.z];
  const View v(arena, ch);
  const int j = blockIdx.y;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ch.nyu || c >= ch.ncol) return;
  const int nxh = ch.nxh, ldc = ch.ldc, nyu = ch.nyu, nyv = ch.nyv;
  const int i = c % nxh, k = c / nxh;
  const double dt = v.scal[SC_DT];
  const double C1 = rk_a(kRK) / ch.Re, C2 = -rk_c(kRK), Cg = -rk_g(kRK), dR = rk_d(kRK);
  const size_t o = (size_t)j * ldc + c;
  const double k2x = v.k2x[i], k2z = v.k2z[k];
  if (j >= 1 && j <= nyu - 2) {
    const double a0 = v.dyu2i[3 * j], a1 = v.dyu2i[3 * j + 1], a2 = v.dyu2i[3 * j + 2];
    const double dd = (k2x + k2z) + a1;
    const double dyu = v.dthdyu[j];
    {
      double2 um = v.u1[o - ldc], u0 = v.u1[o], up = v.u1[o + ldc];
      double2 L = make_double2(a0 * um.x + dd * u0.x + a2 * up.x, a0 * um.y + dd * u0.y + a2 * up.y);
      double2 fm = v.fuv[o - ldc], f0 = v.fuv[o];
      double2 nl = v.nl1[o];
      double2 Nn = make_double2(nl.x + (f0.x - fm.x) * dyu, nl.y + (f0.y - fm.y) * dyu);
      double2 No = v.N1o[o];
      double2 pc = cimul(-dR * v.kx[i], v.p[o]);
      double2 du = make_double2(C1 * L.x + C2 * No.x + pc.x, C1 * L.y + C2 * No.y + pc.y);
      if (c == 0) du.x = du.x + (-dR * v.scal[SC_MPGX]);
      du = make_double2(du.x + Cg * Nn.x, du.y + Cg * Nn.y);
      v.r1[o] = make_double2(u0.x + dt * du.x, u0.y + dt * du.y);
      v.N1o[o] = Nn;
    }
    {
      double2 um = v.u3[o - ldc], u0 = v.u3[o], up = v.u3[o + ldc];
      double2 L = make_double2(a0 * um.x + dd * u0.x + a2 * up.x, a0 * um.y + dd * u0.y + a2 * up.y);
      double2 fm = v.fwv[o - ldc], f0 = v.fwv[o];
      double2 nl = v.nl3[o];
      double2 Nn = make_double2(nl.x + (f0.x - fm.x) * dyu, nl.y + (f0.y - fm.y) * dyu);
      double2 No = v.N3o[o];
      double2 pc = cimul(-dR * v.kz[k], v.p[o]);
      double2 du = make_double2(C1 * L.x + C2 * No.x + pc.x, C1 * L.y + C2 * No.y + pc.y);
      du = make_double2(du.x + Cg * Nn.x, du.y + Cg * Nn.y);
      v.r3[o] = make_double2(u0.x + dt * du.x, u0.y + dt * du.y);
      v.N3o[o] = Nn;
    }
  } else {
    v.r1[o] = make_double2(0, 0);
    v.r3[o] = make_double2(0, 0);
  }
  if (j >= 1 && j <= nyv - 2) {
    const double a0 = v.dyv2i[3 * j], a1 = v.dyv2i[3 * j + 1], a2 = v.dyv2i[3 * j + 2];
    const double dd = (k2x + k2z) + a1;
    const double dyv = v.dthdyv[j];
    double2 um = v.u2[o - ldc], u0 = v.u2[o], up = v.u2[o + ldc];
    double2 L = make_double2(a0 * um.x + dd * u0.x + a2 * up.x, a0 * um.y + dd * u0.y + a2 * up.y);
    double2 f0 = v.fvv[o], fp = v.fvv[o + ldc];
    double2 nl = v.nl2[o];
    double2 Nn = make_double2(nl.x + (fp.x - f0.x) * dyv, nl.y + (fp.y - f0.y) * dyv);
    double2 No = v.N2o[o];
    double C3 = -dR * dyv;
    double2 p0 = v.p[o], pp = v.p[o + ldc];
    double2 du = make_double2(C1 * L.x + C2 * No.x + C3 * (pp.x - p0.x),
                              C1 * L.y + C2 * No.y + C3 * (pp.y - p0.y));
    du = make_double2(du.x + Cg * Nn.x, du.y + Cg * Nn.y);
    v.r2[o] = make_double2(u0.x + dt * du.x, u0.y + dt * du.y);
    v.N2o[o] = Nn;
  } else if (j == 0 || j == nyv - 1) {
    v.r2[o] = make_double2(0, 0);
  }
}

// ------------------------------------------------------------------------
// K5: column solves.  One thread per spectral column, three sweeps in y:
//   sweep 1 (up):   forward elimination of the U/W and V Helmholtz systems
//                   (LU_buildU/LU_buildV + LU_dec + forward substitution)
//   sweep 2 (down): back substitution of U/W/V, the divergence of the
//                   result, and the elimination of the Poisson system
//                   (LU_buildP operator) from the top row down
//   sweep 3 (up):   Poisson back substitution, p += psi, v_corr, the
//                   flow-rate correction of mode (0,0) and the divergence
// The recurrences are sequential (Thomas).  Latency is hidden by (a) staging
// the channel's grid coefficients in LDS once per block, and (b) reading the
// field rows in blocks of SB, the loads of block b+1 issued before the
// stores of block b (on gfx9-class GPUs one counter tracks loads and stores
// in order, so a load issued after a store also waits for that store).
#define SB 8

// LDS copy of the grid coefficients of one channel
struct SolveLDS {
  double *cu0, *cu1, *cu2, *cv0, *cv1, *cv2, *du, *dv, *yv;
};
__device__ SolveLDS stage_coeffs(double *sm, const View &v, int nyu, int nyv) {
  SolveLDS L;
  L.cu0 = sm;
  L.cu1 = L.cu0 + nyu;
  L.cu2 = L.cu1 + nyu;
  L.cv0 = L.cu2 + nyu;
  L.cv1 = L.cv0 + nyv;
  L.cv2 = L.cv1 + nyv;
  L.du = L.cv2 + nyv;
  L.dv = L.du + nyu;
  L.yv = L.dv + nyv;
  for (int j = threadIdx.x; j < nyu; j += blockDim.x) {
    L.cu0[j] = v.dyu2i[3 * j];
    L.cu1[j] = v.dyu2i[3 * j + 1];
    L.cu2[j] = v.dyu2i[3 * j + 2];
    L.du[j] = v.dthdyu[j];
  }
  for (int j = threadIdx.x; j < nyv; j += blockDim.x) {
    L.cv0[j] = v.dyv2i[3 * j];
    L.cv1[j] = v.dyv2i[3 * j + 1];
    L.cv2[j] = v.dyv2i[3 * j + 2];
    L.dv[j] = v.dthdyv[j];
    L.yv[j] = v.yv[j];
  }
  __syncthreads();
  return L;
}
__host__ __device__ inline size_t solve_lds_doubles(int nyu) { return 4 * (size_t)nyu + 5 * (size_t)(nyu - 1); }

struct Blk1 { double2 a[SB], b[SB], c[SB]; };
struct Blk2 { double2 a[SB], b[SB], c[SB]; double e[SB], f[SB]; };
struct Blk3 { double2 a[SB], b[SB], c[SB], d[SB], g[SB]; double e[SB]; };

__global__ void __launch_bounds__(64) k_solve(char *arena, const Chan *chans, int kRK) {
  extern __shared__ double smem_s[];
  const Chan ch = chans[blockIdx.z];
  const View v(arena, ch);
  const int nxh = ch.nxh, ldc = ch.ldc, nyu = ch.nyu, nyv = ch.nyv;
  if ((int)(blockIdx.x * blockDim.x) >= ch.ncol) return;
  const SolveLDS L = stage_coeffs(smem_s, v, nyu, nyv);
  const int c0 = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = c0 < ch.ncol;
  const int c = live ? c0 : ch.ncol - 1;  // idle lanes shadow a real column, never store
  const int i = c % nxh, k = c / nxh;
  const double dt = v.scal[SC_DT];
  const double beta = rk_b(kRK) * dt / ch.Re;
  const double k2x = v.k2x[i], k2z = v.k2z[k];
  const double kxi = v.kx[i], kzk = v.kz[k];
  double *__restrict__ scU = v.scr;
  double *__restrict__ scV = v.scr + (size_t)nyu * ldc;
  double2 *__restrict__ u1 = v.u1;
  double2 *__restrict__ u2 = v.u2;
  double2 *__restrict__ u3 = v.u3;
  const double2 *__restrict__ r1 = v.r1;
  const double2 *__restrict__ r2 = v.r2;
  const double2 *__restrict__ r3 = v.r3;
  double2 *__restrict__ psi = v.psi;
  double2 *__restrict__ pp = v.p;
  double2 *__restrict__ dvg = v.div;
  const int ns = 1, ne = nyu - 2;
  const double gw2 = ch.gw2;

  // ---- sweep 1: forward elimination, U/W rows 0..nyu-1, V rows 0..nyv-1
  double cU = ch.gw1, cV = 0.0;
  double2 y1 = r1[c], y3 = r3[c], y2 = r2[c];
  if (live) {
    u1[c] = y1;
    u3[c] = y3;
    u2[c] = y2;
    scU[c] = cU;
    scV[c] = cV;
  }
  auto load1 = [&](Blk1 &B, int jb) {
#pragma unroll
    for (int q = 0; q < SB; q++) {
      int j = min(jb + q, nyu - 1);
      size_t o = (size_t)j * ldc + c;
      B.a[q] = r1[o];
      B.b[q] = r3[o];
      B.c[q] = r2[(size_t)min(j, nyv - 1) * ldc + c];
    }
  };
  auto comp1 = [&](const Blk1 &B, int jb) {
#pragma unroll
    for (int q = 0; q < SB; q++) {
      int j = jb + q;
      if (j < nyu) {
        size_t o = (size_t)j * ldc + c;
        double a1, a2, a3;
        if (j <= nyu - 2) {
          a1 = -beta * L.cu0[j];
          a2 = 1.0 - beta * ((L.cu1[j] + k2x) + k2z);
          a3 = -beta * L.cu2[j];
        } else {
          a1 = gw2;
          a2 = 1.0;
          a3 = 0.0;
        }
        double d = 1.0 / (a2 - a1 * cU);
        y1 = make_double2((B.a[q].x - a1 * y1.x) * d, (B.a[q].y - a1 * y1.y) * d);
        y3 = make_double2((B.b[q].x - a1 * y3.x) * d, (B.b[q].y - a1 * y3.y) * d);
        cU = a3 * d;
        if (live) {
          u1[o] = y1;
          u3[o] = y3;
          scU[o] = cU;
        }
        if (j < nyv) {
          if (j <= nyv - 2) {
            a1 = -beta * L.cv0[j];
            a2 = 1.0 - beta * ((L.cv1[j] + k2x) + k2z);
            a3 = -beta * L.cv2[j];
          } else {
            a1 = 0.0;
            a2 = 1.0;
            a3 = 0.0;
          }
          d = 1.0 / (a2 - a1 * cV);
          y2 = make_double2((B.c[q].x - a1 * y2.x) * d, (B.c[q].y - a1 * y2.y) * d);
          cV = a3 * d;
          if (live) {
            u2[o] = y2;
            scV[o] = cV;
          }
        }
      }
    }
  };
  {
    Blk1 A, B;
    load1(A, 1);
    for (int jb = 1; jb < nyu; jb += 2 * SB) {
      load1(B, jb + SB);
      comp1(A, jb);
      load1(A, jb + 2 * SB);
      comp1(B, jb + SB);
    }
  }

  // ---- sweep 2: back substitution (rows nyu-2..0), divergence, Poisson
  //      elimination from row ne down to ns
  const double dtirk = (1.0 / dt) / (rk_a(kRK) + rk_b(kRK));
  const double yu21 = v.yu[2] - v.yu[1];
  double2 x1 = y1, x3 = y3;  // row nyu-1 (final)
  double2 x2 = y2;           // row nyv-1 = nyu-2 (final)
  double ep = 0.0;
  double2 yp = make_double2(0, 0);
  double Qx0 = 0.0;
  auto load2 = [&](Blk2 &B, int jb) {
#pragma unroll
    for (int q = 0; q < SB; q++) {
      int j = max(jb - q, 0);
      size_t o = (size_t)j * ldc + c;
      B.a[q] = u1[o];
      B.b[q] = u3[o];
      B.e[q] = scU[o];
      size_t ov = (size_t)min(j, nyv - 2) * ldc + c;
      B.c[q] = u2[ov];
      B.f[q] = scV[ov];
    }
  };
  auto comp2 = [&](const Blk2 &B, int jb) {
#pragma unroll
    for (int q = 0; q < SB; q++) {
      int j = jb - q;
      if (j >= 0) {
        size_t o = (size_t)j * ldc + c;
        double2 n1 = make_double2(B.a[q].x - B.e[q] * x1.x, B.a[q].y - B.e[q] * x1.y);
        double2 n3 = make_double2(B.b[q].x - B.e[q] * x3.x, B.b[q].y - B.e[q] * x3.y);
        double2 n2 = x2;
        if (j <= nyv - 2) n2 = make_double2(B.c[q].x - B.f[q] * x2.x, B.c[q].y - B.f[q] * x2.y);
        if (live) {
          u1[o] = n1;
          u3[o] = n3;
          if (j <= nyv - 2) u2[o] = n2;
        }
        const int qr = j + 1;  // pressure row whose divergence is now known
        if (qr >= ns && qr <= ne) {
          double2 a = cimul(kxi, x1), b = cimul(kzk, x3);
          double w = L.du[qr];
          double2 dv = make_double2((a.x + b.x) + w * (x2.x - n2.x), (a.y + b.y) + w * (x2.y - n2.y));
          double2 rhs = make_double2(dtirk * dv.x, dtirk * dv.y);
          double a1, a2, a3;
          const double du_j = w;
          if (c == 0 && qr == ns) {  // mean-mode pressure pin, tridLU_3D.f90:384-390
            rhs = make_double2(0, 0);
            a1 = 0.0;
            a2 = 1.0 / (yu21 * yu21);
            a3 = 0.0;
          } else if (qr == ns) {
            a1 = 0.0;
            a2 = ((du_j * (-L.dv[qr])) + k2x) + k2z;
            a3 = du_j * L.dv[qr];
          } else if (qr == ne) {
            a1 = (-du_j) * (-L.dv[qr - 1]);
            a2 = (((-du_j) * L.dv[qr - 1]) + k2x) + k2z;
            a3 = 0.0;
          } else {
            a1 = (-du_j) * (-L.dv[qr - 1]);
            a2 = ((((-du_j) * L.dv[qr - 1]) + (du_j * (-L.dv[qr]))) + k2x) + k2z;
            a3 = du_j * L.dv[qr];
          }
          if (c == 0) Qx0 = Qx0 + (L.yv[qr] - L.yv[qr - 1]) * x1.x;
          double d = qr == ne ? 1.0 / a2 : 1.0 / (a2 - a3 * ep);
          yp = make_double2((rhs.x - a3 * yp.x) * d, (rhs.y - a3 * yp.y) * d);
          ep = a1 * d;
          if (live) {
            size_t oq = (size_t)qr * ldc + c;
            psi[oq] = yp;
            scU[oq] = ep;  // row qr of scU was read one row earlier
          }
        }
        x1 = n1;
        x3 = n3;
        x2 = n2;
      }
    }
  };
  {
    Blk2 A, B;
    load2(A, nyu - 2);
    for (int jb = nyu - 2; jb >= 0; jb -= 2 * SB) {
      load2(B, jb - SB);
      comp2(A, jb);
      load2(A, jb - 2 * SB);
      comp2(B, jb - SB);
    }
  }
  // flow-rate correction of mode (0,0): flowrate_corr, littleharsh.f90:476-509
  double g = 0.0;
  if (c0 == 0 && ch.flag_ctpress == 0) {
    g = (v.scal[SC_QXT] - Qx0) / v.scal[SC_QX1 + kRK - 1];
    v.scal[SC_MPGX] = v.scal[SC_MPGX] + g;
  }
  const double *u11 = v.u11 + (size_t)(kRK - 1) * nyu;

  // ---- sweep 3: Poisson back substitution (rows ns..ne), p += psi, v_corr
  const double dtrk = dt * (rk_a(kRK) + rk_b(kRK));
  const double dkx = dtrk * kxi, dkz = dtrk * kzk;
  double2 pprev = make_double2(0, 0);                          // psi(j-1)
  double2 u1p = make_double2(0, 0), u3p = make_double2(0, 0);  // corrected u at j-1
  double2 vpp = make_double2(0, 0);                            // corrected v at j-2 (v(0) = 0)
  double2 u1first = make_double2(0, 0), u3first = make_double2(0, 0);
  auto load3 = [&](Blk3 &B, int jb) {
#pragma unroll
    for (int q = 0; q < SB; q++) {
      int j = min(jb + q, ne);
      size_t o = (size_t)j * ldc + c;
      B.a[q] = psi[o];
      B.e[q] = scU[o];
      B.b[q] = pp[o];
      B.c[q] = u1[o];
      B.d[q] = u3[o];
      B.g[q] = u2[o - ldc];
    }
  };
  auto comp3 = [&](const Blk3 &B, int jb) {
#pragma unroll
    for (int q = 0; q < SB; q++) {
      int j = jb + q;
      if (j <= ne) {
        size_t o = (size_t)j * ldc + c;
        double2 ps = j == ns ? B.a[q] : make_double2(B.a[q].x - B.e[q] * pprev.x, B.a[q].y - B.e[q] * pprev.y);
        double2 a = cimul(dkx, ps), b = cimul(dkz, ps);
        double2 n1 = make_double2(B.c[q].x - a.x, B.c[q].y - a.y);
        double2 n3 = make_double2(B.d[q].x - b.x, B.d[q].y - b.y);
        if (c0 == 0) n1.x = n1.x + g * u11[j];
        if (live) {
          psi[o] = ps;
          pp[o] = make_double2(B.b[q].x + ps.x, B.b[q].y + ps.y);
          u1[o] = n1;
          u3[o] = n3;
        }
        if (j == ns) {
          u1first = n1;
          u3first = n3;
        }
        if (j - 1 >= 1) {
          // v row j-1 and the divergence of pressure row j-1
          double sv = dtrk * L.dv[j - 1];
          double2 vv = make_double2(B.g[q].x - sv * (ps.x - pprev.x), B.g[q].y - sv * (ps.y - pprev.y));
          double2 da = cimul(kxi, u1p), db = cimul(kzk, u3p);
          double w = L.du[j - 1];
          double2 dd = make_double2((da.x + db.x) + w * (vv.x - vpp.x), (da.y + db.y) + w * (vv.y - vpp.y));
          if (live) {
            size_t ov = (size_t)(j - 1) * ldc + c;
            u2[ov] = vv;
            dvg[ov] = dd;
          }
          vpp = vv;
        }
        pprev = ps;
        u1p = n1;
        u3p = n3;
      }
    }
  };
  {
    Blk3 A, B;
    load3(A, ns);
    for (int jb = ns; jb <= ne; jb += 2 * SB) {
      load3(B, jb + SB);
      comp3(A, jb);
      load3(A, jb + 2 * SB);
      comp3(B, jb + SB);
    }
  }
  if (live) {
    // divergence of row ne: v(ne) is the upper wall
    double2 da = cimul(kxi, u1p), db = cimul(kzk, u3p);
    double w = L.du[ne];
    dvg[(size_t)ne * ldc + c] = make_double2((da.x + db.x) + w * (0.0 - vpp.x), (da.y + db.y) + w * (0.0 - vpp.y));
    // ghost cells and walls (v_corr, littleharsh.f90:992-1044)
    u1[c] = cscale(ch.gwbc1, u1first);
    u1[(size_t)(nyu - 1) * ldc + c] = cscale(ch.gwbc1, u1p);
    u3[c] = cscale(ch.gwbc3, u3first);
    u3[(size_t)(nyu - 1) * ldc + c] = cscale(ch.gwbc3, u3p);
    u2[c] = make_double2(0, 0);
    u2[(size_t)(nyv - 1) * ldc + c] = make_double2(0, 0);
  }
}

// ------------------------------------------------------------------------
// flow rate of the (0,0) mode of every channel (flowrateIm)
__global__ void k_flowrate(char *arena, const Chan *chans, int nchan, int set_target) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nchan) return;
  const Chan ch = chans[c];
  const View v(arena, ch);
  double Q = 0.0;
  for (int j = 1; j <= ch.nyu - 2; j++) Q = Q + (v.yv[j] - v.yv[j - 1]) * v.u1[(size_t)j * ch.ldc].x;
  v.scal[SC_QX] = Q;
  if (set_target && ch.flag_ctpress != 1) v.scal[SC_QXT] = Q;
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

hipError_t launch_nl_stage(char *arena, const Chan *d_chans, const ShapeMax &m, bool mixed, int stage, hipStream_t s) {
  size_t es = mixed ? sizeof(float2) : sizeof(double2);
  dim3 b(256);
  if (stage == 0) {
    dim3 g(m.zinv_tiles, m.nslot, m.nchan * 3);
    if (mixed)
      hipLaunchKernelGGL(k_zinv<float2>, g, b, m.lds_zinv * es, s, arena, d_chans);
    else
      hipLaunchKernelGGL(k_zinv<double2>, g, b, m.lds_zinv * es, s, arena, d_chans);
  } else if (stage == 1) {
    dim3 g(m.xprod_tiles, m.nslot, m.nchan);
    if (mixed)
      hipLaunchKernelGGL(k_xprod<float2>, g, b, m.lds_xprod * es, s, arena, d_chans);
    else
      hipLaunchKernelGGL(k_xprod<double2>, g, b, m.lds_xprod * es, s, arena, d_chans);
  } else {
    dim3 g(m.zfwd_tiles, m.nslot, m.nchan);
    if (mixed)
      hipLaunchKernelGGL(k_zfwd<float2>, g, b, m.lds_zfwd * es, s, arena, d_chans);
    else
      hipLaunchKernelGGL(k_zfwd<double2>, g, b, m.lds_zfwd * es, s, arena, d_chans);
  }
  return hipGetLastError();
}

hipError_t launch_nonlinear(char *arena, const Chan *d_chans, const ShapeMax &m, bool mixed, hipStream_t s) {
  for (int st = 0; st < 3; st++) {
    hipError_t e = launch_nl_stage(arena, d_chans, m, mixed, st, s);
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

hipError_t launch_rhs(char *arena, const Chan *d_chans, const ShapeMax &m, int kRK, hipStream_t s) {
  dim3 g((m.ncol + 255) / 256, m.nyu, m.nchan);
  hipLaunchKernelGGL(k_rhs, g, dim3(256), 0, s, arena, d_chans, kRK);
  return hipGetLastError();
}

hipError_t launch_solve(char *arena, const Chan *d_chans, const ShapeMax &m, int kRK, hipStream_t s) {
  dim3 g((m.ncol + 63) / 64, 1, m.nchan);
  hipLaunchKernelGGL(k_solve, g, dim3(64), solve_lds_doubles(m.nyu) * sizeof(double), s, arena, d_chans, kRK);
  return hipGetLastError();
}

hipError_t launch_flowrate(char *arena, const Chan *d_chans, int nchan, int set_target, hipStream_t s) {
  hipLaunchKernelGGL(k_flowrate, dim3((nchan + 63) / 64), dim3(64), 0, s, arena, d_chans, nchan, set_target);
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
