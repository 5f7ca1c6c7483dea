// lh_texture.hip -- textured walls (free-shear / no-slip point masks) for the
// U/W Helmholtz solves of the RK3 substep.
//
// The reference builds the mask planeBC (start.f90:1776-1829) and, in its
// commented-out LUsolU / LU_build (tridLU_3D.f90:753-1078, 1284-1358),
// replaces the wall row of the U/W systems point by point on the N1 x N2 grid
// (four_to_phys_N / phys_to_four_N): no-slip keeps u0 + gw1 u1 = 0,
// free-shear has u0 - u1 = 0 (top: u_ny - u_{ny-1} = 0 against
// u_ny + gw2 u_{ny-1} = 0).  Here that mode-coupled system is solved exactly
// for the unfactored operator A(k) as a boundary source on the wall rows:
//   A u = r + e_0 sb + e_ny st,  sb = T[Pb . T^-1 u_1],  st = T[Pt . T^-1 u_{ny-1}]
// with Pb = 1 + gw1, Pt = 1 + gw2 on free-shear points and 0 elsewhere, T the
// N-grid transform.  Writing u_1 = U_b + gbb sb + gbt st (U the smooth
// solution, g entries of A^-1), the sources are a fixed point, iterated with
// the column-independent part gbb(0) moved to the left:
//   sb <- T[Pb . T^-1(U_b + (gbb - gbb(0)) sb + gbt st) / (1 - gbb(0) Pb)]
// which contracts by max|gbb - gbb(0)| |Pb| (O(beta k^2) small) per pass.
// sb, st then become the right-hand sides of the wall rows and the ordinary
// column solve runs; afterwards the ghost rows follow the mask point by point
// (u0 = gwbc u1 on no-slip, u0 = u1 on free-shear points), as v_corr's
// ghost update (littleharsh.f90:992-1044) generalised to a mask, applied as
// u0 = gwbc u1 + T[(1 - gwbc) chi_fs . T^-1 u1]: the uniform part stays in
// Fourier space, because the N-grid transform cannot hold the x-Nyquist
// column i = N1/2 that the 3/2-padded step keeps.
//
// U_b and U_t come from rows 1 and ny-1 of A^-1 (tables built once), which
// decay exponentially away from their wall: only the first rows are summed.
#include <algorithm>

#include "lh_common.h"
#include "lh_fft.h"
#include "lh_kernels.h"
#include "lh_tex1d.h"

namespace lh {

// rows 1 and ny-1 of A^-1 per column and substep, from the LU factors:
// A^T z = e_m by U^T w = e_m, L^T z = w
__global__ void __launch_bounds__(64) k_tex_tables(char *arena, const Chan *chans) {
  const Chan ch = chans[blockIdx.z];
  const View v(arena, ch);
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ch.ncol) return;
  const int n = ch.nyu, ldc = ch.ldc;
  const size_t plane = (size_t)(n + PADROWS) * ldc, tab = (size_t)n * ldc;
  const double dt = v.scal[SC_DT];
  const double tol = 1e-18;
  int jb = 1, jt = 1;
  for (int kRK = 1; kRK <= 3; kRK++) {
    const double beta = rk_b(kRK) * dt / ch.Re;
    const double *dU = v.fac + (size_t)((kRK - 1) * 4) * plane, *cU = dU + plane;
    double *Wb = v.wtab + (size_t)((kRK - 1) * 2) * tab, *Wt = Wb + tab;
    auto a1 = [&](int j) { return j == 0 ? 0.0 : (j <= n - 2 ? -beta * v.dyu2i[3 * j] : ch.gw2); };
    // row 1: w_0 = 0, w_1 = 1, w_j = -c_{j-1} w_{j-1}
    double w = 0.0;
    Wb[c] = 0.0;
    for (int j = 1; j < n; j++) {
      w = j == 1 ? 1.0 : -cU[(size_t)(j - 1) * ldc + c] * w;
      Wb[(size_t)j * ldc + c] = w;
    }
    double z = Wb[(size_t)(n - 1) * ldc + c] * dU[(size_t)(n - 1) * ldc + c];
    Wb[(size_t)(n - 1) * ldc + c] = z;
    double zmax = fabs(z);
    for (int j = n - 2; j >= 0; j--) {
      z = (Wb[(size_t)j * ldc + c] - a1(j + 1) * z) * dU[(size_t)j * ldc + c];
      Wb[(size_t)j * ldc + c] = z;
      zmax = fmax(zmax, fabs(z));
    }
    for (int j = n - 1; j >= 0; j--)
      if (fabs(Wb[(size_t)j * ldc + c]) > tol * zmax) {
        jb = max(jb, j + 1);
        break;
      }
    // row n-2: w_{n-2} = 1, w_{n-1} = -c_{n-2}, zero below
    z = -cU[(size_t)(n - 2) * ldc + c] * dU[(size_t)(n - 1) * ldc + c];
    Wt[(size_t)(n - 1) * ldc + c] = z;
    zmax = fabs(z);
    for (int j = n - 2; j >= 0; j--) {
      z = ((j == n - 2 ? 1.0 : 0.0) - a1(j + 1) * z) * dU[(size_t)j * ldc + c];
      Wt[(size_t)j * ldc + c] = z;
      zmax = fmax(zmax, fabs(z));
    }
    for (int j = 0; j < n; j++)
      if (fabs(Wt[(size_t)j * ldc + c]) > tol * zmax) {
        jt = max(jt, n - j);
        break;
      }
  }
  atomicMax(&v.iscal[0], jb);
  atomicMax(&v.iscal[1], jt);
}

// adjacent-row values of the smooth solution, U = sum_j W[j] r_j over the
// rows where the inverse row is significant; grid (cols / 64, 2 walls,
// channels), 256 threads: 64 columns x 4 interleaved row subsets, each
// thread's loads unrolled so they are in flight together, partial sums
// added in a fixed order through LDS
__global__ void __launch_bounds__(256) k_tex_adj(char *arena, const Chan *chans, int kRK) {
  const Chan ch = chans[blockIdx.z];
  if (!ch.tex) return;
  const View v(arena, ch);
  const int cl = threadIdx.x & 63, part = threadIdx.x >> 6;
  const int c = blockIdx.x * 64 + cl;
  const bool live = c < ch.ncol;
  const int n = ch.nyu, ldc = ch.ldc, w = blockIdx.y;
  const size_t tab = (size_t)n * ldc;
  const double *W = v.wtab + (size_t)((kRK - 1) * 2 + w) * tab;
  const int J = w == 0 ? v.iscal[0] : v.iscal[1];
  const int j0 = w == 0 ? 0 : n - J, j1 = w == 0 ? J : n;
  double2 a = make_double2(0, 0), b = make_double2(0, 0);
  if (live) {
#pragma unroll 4
    for (int j = j0 + part; j < j1; j += 4) {
      const size_t o = (size_t)j * ldc + c;
      const double q = W[o];
      const double2 r1 = v.r1[o], r3 = v.r3[o];
      a = make_double2(a.x + q * r1.x, a.y + q * r1.y);
      b = make_double2(b.x + q * r3.x, b.y + q * r3.y);
    }
  }
  __shared__ double2 red[2][4][64];
  red[0][part][cl] = a;
  red[1][part][cl] = b;
  __syncthreads();
  if (part == 0 && live) {
    for (int q = 1; q < 4; q++) {
      a = cadd(a, red[0][q][cl]);
      b = cadd(b, red[1][q][cl]);
    }
    v.texu[(size_t)(0 * 2 + w) * ch.ncol + c] = a;
    v.texu[(size_t)(1 * 2 + w) * ch.ncol + c] = b;
  }
}

// texw <- U + (g - gbar) texw (iteration `it`; texw holds the previous
// sources, unused when it == 0)
__global__ void __launch_bounds__(256) k_tex_pre(char *arena, const Chan *chans, int kRK, int it) {
  const Chan ch = chans[blockIdx.z];
  if (!ch.tex || ch.tex1d) return;
  const View v(arena, ch);
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ch.ncol) return;
  const int n = ch.nyu, ldc = ch.ldc;
  const size_t tab = (size_t)n * ldc;
  const double *Wb = v.wtab + (size_t)((kRK - 1) * 2) * tab, *Wt = Wb + tab;
  const double gbb = Wb[c], gbt = Wb[(size_t)(n - 1) * ldc + c];
  const double gtb = Wt[c], gtt = Wt[(size_t)(n - 1) * ldc + c];
  const double gbb0 = Wb[0], gtt0 = Wt[(size_t)(n - 1) * ldc];
  const int nc = ch.ncol;
  for (int f = 0; f < 2; f++) {
    const double2 Ub = v.texu[(size_t)(f * 2) * nc + c], Ut = v.texu[(size_t)(f * 2 + 1) * nc + c];
    double2 sb = make_double2(0, 0), st = make_double2(0, 0);
    if (it > 0) {
      sb = v.texw[(size_t)(f * 2) * nc + c];
      st = v.texw[(size_t)(f * 2 + 1) * nc + c];
    }
    const double db = gbb - gbb0, dt_ = gtt - gtt0;
    v.texw[(size_t)(f * 2) * nc + c] =
        make_double2(Ub.x + db * sb.x + gbt * st.x, Ub.y + db * sb.y + gbt * st.y);
    v.texw[(size_t)(f * 2 + 1) * nc + c] =
        make_double2(Ut.x + gtb * sb.x + dt_ * st.x, Ut.y + gtb * sb.y + dt_ * st.y);
  }
}

// the pointwise mask of field f, wall w at one physical point (fs: the point
// is free-shear)
// mode 0: v <- P v / (1 - gbar P)   (wall sources)
// mode 1: v <- (1 - gwbc) v on free-shear points, 0 elsewhere (ghost rows)
struct TexMask {
  double g0, gw, gwbc;
  int mode;
  __device__ TexMask(const View &v, const Chan &ch, int kRK, int mode_, int f, int w) : mode(mode_) {
    g0 = 0.0;
    if (mode == 0) {
      const int n = ch.nyu, ldc = ch.ldc;
      const size_t tab = (size_t)n * ldc;
      const double *Wb = v.wtab + (size_t)((kRK - 1) * 2) * tab, *Wt = Wb + tab;
      g0 = w == 0 ? Wb[0] : Wt[(size_t)(n - 1) * ldc];
    }
    gw = w == 0 ? ch.gw1 : ch.gw2;
    gwbc = f == 0 ? ch.gwbc1 : ch.gwbc3;
  }
  __device__ __forceinline__ double operator()(double val, bool fs) const {
    if (mode == 0) {
      const double Pw = fs ? 1.0 + gw : 0.0;
      return Pw * val / (1.0 - g0 * Pw);
    }
    return fs ? (1.0 - gwbc) * val : 0.0;
  }
};

// pointwise mask on the physical N-grid planes of texw; grid (points, 4, ch)
__global__ void __launch_bounds__(256) k_tex_mask(char *arena, const Chan *chans, int kRK, int mode) {
  const Chan ch = chans[blockIdx.z];
  if (!ch.tex || ch.tex1d) return;
  const View v(arena, ch);
  const int N1 = ch.N1, N2 = ch.N2;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N1 * N2) return;
  const int p = blockIdx.y, f = p >> 1, w = p & 1;
  const int z = t / N1, x = t - z * N1;
  double *P = reinterpret_cast<double *>(v.texw + (size_t)p * ch.ncol);
  const TexMask mk(v, ch, kRK, mode, f, w);
  double &val = P[(size_t)z * (N1 + 2) + x];
  val = mk(val, v.mask[w * N1 * N2 + t] == 2);
}

// The x transforms of a mask pass and the mask in one launch: per pair of
// z rows of a wall plane, the inverse real x transform (rft packing two
// rows in one complex FFT, exactly as k_plane_x), the pointwise mask of
// k_tex_mask and the forward real x transform, the physical rows staying in
// LDS; the same arithmetic as the three launches, one HBM round trip and
// two launches less.  Grid (row-pair tiles, 4 planes, channels).
__host__ __device__ inline int tex_xmask_tile(int N1) {
  const int t = 1024 / N1;
  return t < 1 ? 1 : t;
}
__global__ void __launch_bounds__(256) k_tex_xmask(char *arena, const Chan *chans, int kRK, int mode) {
  const Chan ch = chans[blockIdx.z];
  if (!ch.tex || ch.tex1d) return;
  const View v(arena, ch);
  const int Ng1 = ch.N1, Ng2 = ch.N2, igal = Ng1 + 2, h = Ng1 / 2;
  const int TZ = tex_xmask_tile(Ng1);  // pairs of rows per workgroup
  const int z0 = blockIdx.x * 2 * TZ;
  if (z0 >= Ng2) return;
  const int p = blockIdx.y, f = p >> 1, w = p & 1;
  double *P = reinterpret_cast<double *>(v.texw + (size_t)p * ch.ncol);
  const int nzr = min(2 * TZ, Ng2 - z0), npair = (nzr + 1) / 2;
  extern __shared__ double2 smem_d[];
  double2 *A = smem_d, *B = A + TZ * Ng1, *TW = B + TZ * Ng1;
  stage_twiddles<double2>(TW, reinterpret_cast<const double2 *>(arena + ch.o_twxN), Ng1);
  for (int t = threadIdx.x; t < npair * Ng1; t += blockDim.x) {
    const int q = t / Ng1, m = t - q * Ng1;
    const int za = z0 + 2 * q, zb = za + 1;
    const int mm = m <= h ? m : Ng1 - m;
    const double sg = m <= h ? 1.0 : -1.0;
    const double *ra = P + (size_t)za * igal;
    double2 a = make_double2(ra[2 * mm], sg * ra[2 * mm + 1]), b = make_double2(0, 0);
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
  const TexMask mk(v, ch, kRK, mode, f, w);
  const int *fsm = v.mask + w * Ng1 * Ng2;
  for (int t = threadIdx.x; t < npair * Ng1; t += blockDim.x) {
    const int q = t / Ng1, n = t - q * Ng1;
    const int za = z0 + 2 * q, zb = za + 1;
    const double2 x = R[q * Ng1 + n];
    R[q * Ng1 + n] = make_double2(mk(x.x, fsm[za * Ng1 + n] == 2), zb < z0 + nzr ? mk(x.y, fsm[zb * Ng1 + n] == 2) : 0.0);
  }
  __syncthreads();
  double2 *Y = fft_lds<double2>(R, R == A ? B : A, Ng1, npair, Ng1, -1, TW);
  const double inv = 1.0 / Ng1;
  for (int t = threadIdx.x; t < npair * (h + 1); t += blockDim.x) {
    const int q = t / (h + 1), m = t - q * (h + 1);
    const int za = z0 + 2 * q, zb = za + 1;
    const double2 Ym = Y[q * Ng1 + m], Yc = Y[q * Ng1 + (m == 0 ? 0 : Ng1 - m)];
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

// mode 0: wall-row right-hand sides r_f(0), r_f(ny) <- sources;
// mode 1: ghost rows u_f(0), u_f(ny) (gwbc u1 from the column solve) += texw;
// only textured walls
__global__ void __launch_bounds__(256) k_tex_put(char *arena, const Chan *chans, int mode) {
  const Chan ch = chans[blockIdx.z];
  if (!ch.tex || ch.tex1d) return;
  const View v(arena, ch);
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ch.ncol) return;
  const int n = ch.nyu, ldc = ch.ldc, nc = ch.ncol;
  for (int f = 0; f < 2; f++) {
    double2 *dst = mode == 0 ? (f == 0 ? v.r1 : v.r3) : (f == 0 ? v.u1 : v.u3);
    const double2 wb = v.texw[(size_t)(f * 2) * nc + c], wt = v.texw[(size_t)(f * 2 + 1) * nc + c];
    const size_t ot = (size_t)(n - 1) * ldc + c;
    if (mode == 0) {
      if (ch.tex & 1) dst[c] = wb;
      if (ch.tex & 2) dst[ot] = wt;
    } else {
      if (ch.tex & 1) dst[c] = cadd(dst[c], wb);
      if (ch.tex & 2) dst[ot] = cadd(dst[ot], wt);
    }
  }
}

// texw <- rows 1 and ny-1 of u1, u3 (after the projection)
__global__ void __launch_bounds__(256) k_tex_get(char *arena, const Chan *chans) {
  const Chan ch = chans[blockIdx.z];
  if (!ch.tex || ch.tex1d) return;
  const View v(arena, ch);
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ch.ncol) return;
  const int n = ch.nyu, ldc = ch.ldc, nc = ch.ncol;
  for (int f = 0; f < 2; f++) {
    const double2 *src = f == 0 ? v.u1 : v.u3;
    v.texw[(size_t)(f * 2) * nc + c] = src[(size_t)1 * ldc + c];
    v.texw[(size_t)(f * 2 + 1) * nc + c] = src[(size_t)(n - 2) * ldc + c];
  }
}

hipError_t launch_tex_adj(char *arena, const Chan *d_chans, const ShapeMax &m, int kRK, hipStream_t s) {
  hipLaunchKernelGGL(k_tex_adj, dim3((m.ncol + 63) / 64, 2, m.nchan), dim3(256), 0, s, arena, d_chans, kRK);
  return hipGetLastError();
}

hipError_t launch_tex_tables(char *arena, const Chan *d_chans, const ShapeMax &m, hipStream_t s) {
  dim3 g((m.ncol + 63) / 64, 1, m.nchan);
  hipLaunchKernelGGL(k_tex_tables, g, dim3(64), 0, s, arena, d_chans);
  return hipGetLastError();
}

// The z transforms of a field's two wall planes with the Fourier-space work
// between passes fused in: one workgroup per (tile of x-modes, field,
// channel) holds both walls' sequences, since the source update couples the
// walls.  mode 0: texw = U (the first pass's update), inverse z transform;
// mode 1: forward z transform, the update of k_tex_pre for pass `it`,
// inverse z transform; mode 2: forward z transform, the sources into the
// wall rows of r (k_tex_put 0); mode 3: texw = rows 1 and ny-1 of u
// (k_tex_get), inverse z transform; mode 4: forward z transform, the ghost
// rows of u += texw (k_tex_put 1).  The arithmetic per value is that of the
// separate kernels, in the same order.
__host__ __device__ inline int tex_z_tile(int N2) {
  const int t = 512 / N2;
  return t < 1 ? 1 : t;
}
__global__ void __launch_bounds__(256) k_tex_z(char *arena, const Chan *chans, int kRK, int mode, int it) {
  const Chan ch = chans[blockIdx.z];
  if (!ch.tex || ch.tex1d) return;
  const View v(arena, ch);
  const int f = blockIdx.y, N2 = ch.N2, nxh = ch.nxh, nc = ch.ncol, igal = ch.N1 + 2;
  const int n = ch.nyu, ldc = ch.ldc;
  const int TI = tex_z_tile(N2), i0 = blockIdx.x * TI;
  if (i0 >= nxh) return;
  const int ni = min(TI, nxh - i0), nseq = 2 * ni;  // sequence q = w ni + ii
  extern __shared__ double2 smem_d[];
  double2 *A = smem_d, *B = A + 2 * TI * N2, *TW = B + 2 * TI * N2;
  stage_twiddles<double2>(TW, reinterpret_cast<const double2 *>(arena + ch.o_twzN), N2);
  double *P[2] = {reinterpret_cast<double *>(v.texw + (size_t)(f * 2) * nc),
                  reinterpret_cast<double *>(v.texw + (size_t)(f * 2 + 1) * nc)};
  double2 *R = A;
  if (mode == 0 || mode == 3) {  // Fourier-space values, then the inverse transform
    const double2 *src = f == 0 ? v.u1 : v.u3;
    for (int t = threadIdx.x; t < nseq * N2; t += blockDim.x) {
      const int q = t / N2, k = t - q * N2, w = q / ni, ii = q - w * ni;
      const int c = k * nxh + i0 + ii;
      A[q * N2 + k] = mode == 0 ? v.texu[(size_t)(f * 2 + w) * nc + c] : src[(size_t)(w == 0 ? 1 : n - 2) * ldc + c];
    }
    __syncthreads();
  } else {  // physical-z values, the forward transform
    for (int t = threadIdx.x; t < nseq * N2; t += blockDim.x) {
      const int q = t / N2, z = t - q * N2, w = q / ni, ii = q - w * ni;
      const double *pz = P[w] + (size_t)z * igal + 2 * (i0 + ii);
      A[q * N2 + z] = make_double2(pz[0], pz[1]);
    }
    __syncthreads();
    R = fft_lds<double2>(A, B, N2, nseq, N2, -1, TW);
    const double sc = 1.0 / N2;
    const size_t tab = (size_t)n * ldc;
    const double *Wb = v.wtab + (size_t)((kRK - 1) * 2) * tab, *Wt = Wb + tab;
    for (int t = threadIdx.x; t < ni * N2; t += blockDim.x) {
      const int ii = t / N2, k = t - ii * N2, c = k * nxh + i0 + ii;
      const double2 yb = R[ii * N2 + k], yt = R[(ni + ii) * N2 + k];
      const double2 sb = make_double2(yb.x * sc, yb.y * sc), st = make_double2(yt.x * sc, yt.y * sc);
      if (mode == 1) {
        const double gbb = Wb[c], gbt = Wb[(size_t)(n - 1) * ldc + c];
        const double gtb = Wt[c], gtt = Wt[(size_t)(n - 1) * ldc + c];
        const double gbb0 = Wb[0], gtt0 = Wt[(size_t)(n - 1) * ldc];
        const double2 Ub = v.texu[(size_t)(f * 2) * nc + c], Ut = v.texu[(size_t)(f * 2 + 1) * nc + c];
        const double db = gbb - gbb0, dt_ = gtt - gtt0;
        R[ii * N2 + k] = make_double2(Ub.x + db * sb.x + gbt * st.x, Ub.y + db * sb.y + gbt * st.y);
        R[(ni + ii) * N2 + k] = make_double2(Ut.x + gtb * sb.x + dt_ * st.x, Ut.y + gtb * sb.y + dt_ * st.y);
      } else {
        double2 *dst = mode == 2 ? (f == 0 ? v.r1 : v.r3) : (f == 0 ? v.u1 : v.u3);
        const size_t ot = (size_t)(n - 1) * ldc + c;
        if (mode == 2) {
          if (ch.tex & 1) dst[c] = sb;
          if (ch.tex & 2) dst[ot] = st;
        } else {
          if (ch.tex & 1) dst[c] = cadd(dst[c], sb);
          if (ch.tex & 2) dst[ot] = cadd(dst[ot], st);
        }
      }
    }
    if (mode != 1) return;
    __syncthreads();
  }
  // inverse z transform to the physical-z planes
  double2 *X = fft_lds<double2>(R, R == A ? B : A, N2, nseq, N2, +1, TW);
  for (int t = threadIdx.x; t < nseq * N2; t += blockDim.x) {
    const int q = t / N2, z = t - q * N2, w = q / ni, ii = q - w * ni;
    double *pz = P[w] + (size_t)z * igal + 2 * (i0 + ii);
    const double2 x = X[q * N2 + z];
    pz[0] = x.x;
    pz[1] = x.y;
  }
}
static hipError_t tex_z(char *arena, const Chan *d_chans, const ShapeMax &m, const std::vector<TexGroup> &groups,
                        int kRK, int mode, int it, hipStream_t s) {
  int gx = 1;
  size_t lds = 0;
  for (const TexGroup &g : groups) {
    const int TI = tex_z_tile(g.N2);
    gx = std::max(gx, (g.N1 / 2 + 1 + TI - 1) / TI);
    lds = std::max(lds, (size_t)(4 * TI + 1) * g.N2 * sizeof(double2));
  }
  hipLaunchKernelGGL(k_tex_z, dim3(gx, 2, m.nchan), dim3(256), lds, s, arena, d_chans, kRK, mode, it);
  return hipGetLastError();
}

// x transforms and mask of every textured channel's four wall planes
static hipError_t tex_xmask(char *arena, const Chan *d_chans, const ShapeMax &m, const std::vector<TexGroup> &groups,
                            int kRK, int mode, hipStream_t s) {
  int gx = 1;
  size_t lds = 0;
  for (const TexGroup &g : groups) {
    const int TZ = tex_xmask_tile(g.N1);
    gx = std::max(gx, (g.N2 + 2 * TZ - 1) / (2 * TZ));
    lds = std::max(lds, (size_t)(2 * TZ + 1) * g.N1 * sizeof(double2));
  }
  hipLaunchKernelGGL(k_tex_xmask, dim3(gx, 4, m.nchan), dim3(256), lds, s, arena, d_chans, kRK, mode);
  return hipGetLastError();
}

// nit passes: z inverse with the first update (k_tex_z 0), then per pass
// the x transforms and mask (k_tex_xmask) and either the next pass's update
// between the z transforms (k_tex_z 1) or, after the last, the sources into
// r (k_tex_z 2)
hipError_t launch_tex_sources(char *arena, const Chan *d_chans, const ShapeMax &m,
                              const std::vector<TexGroup> &groups, int kRK, int nit, hipStream_t s) {
  if (nit < 1) {  // no pass: the sources are U (as k_tex_pre 0 then k_tex_put 0)
    const dim3 b(256), gc((m.ncol + 255) / 256, 1, m.nchan);
    hipLaunchKernelGGL(k_tex_pre, gc, b, 0, s, arena, d_chans, kRK, 0);
    hipLaunchKernelGGL(k_tex_put, gc, b, 0, s, arena, d_chans, 0);
    return hipGetLastError();
  }
  hipError_t e = tex_z(arena, d_chans, m, groups, kRK, 0, 0, s);
  for (int it = 0; it < nit && e == hipSuccess; it++) {
    e = tex_xmask(arena, d_chans, m, groups, kRK, 0, s);
    if (e == hipSuccess) e = tex_z(arena, d_chans, m, groups, kRK, it + 1 < nit ? 1 : 2, it + 1, s);
  }
  return e;
}

hipError_t launch_tex_ghosts(char *arena, const Chan *d_chans, const ShapeMax &m,
                             const std::vector<TexGroup> &groups, hipStream_t s) {
  hipError_t e = tex_z(arena, d_chans, m, groups, 1, 3, 0, s);
  if (e == hipSuccess) e = tex_xmask(arena, d_chans, m, groups, 1, 1, s);
  if (e == hipSuccess) e = tex_z(arena, d_chans, m, groups, 1, 4, 0, s);
  return e;
}

// ------------------------------------------------------------------------
// The same iteration with one wall plane of one field resident in LDS:
// one workgroup per (field, wall, channel) runs every pass -- Fourier
// update, inverse z and x transforms, mask, forward transforms -- without
// leaving LDS.  The two walls are solved independently here: their coupling
// through the interior, gbt and gtb, is the product of ny pivots of order
// beta/dy^2 and vanishes in FP64 for every grid this solver takes (the
// oracle keeps it and checks that).  Layouts: Fourier plane [i][z] (z
// contiguous for the z transforms); physical rows in pairs [q][x] holding
// rows 2q + i*(2q+1).
// mode 0: wall source into the wall row of r (nit passes);
// mode 1: masked ghost row added to u (one pass).
__host__ __device__ inline size_t tex_fused_lds(int N1, int N2) {
  return (2 * (size_t)(N1 / 2 + 1) * N2 + N1 + N2) * sizeof(double2) + (size_t)N1 * N2;
}

__global__ void __launch_bounds__(1024) k_tex_fused(char *arena, const Chan *chans, int kRK, int nit, int mode) {
  const Chan ch = chans[blockIdx.z];
  const int f = blockIdx.x >> 1, w = blockIdx.x & 1;  // field (u1, u3), wall (bottom, top)
  if (!(ch.tex & (1 << w)) || ch.tex1d) return;
  const View v(arena, ch);
  const int N1 = ch.N1, N2 = ch.N2, nxh = ch.nxh, nc = ch.ncol, n = ch.nyu, ldc = ch.ldc;
  const int h = N1 / 2;
  const int nt = blockDim.x, tid = threadIdx.x;
  extern __shared__ double2 smem_d[];
  double2 *A = smem_d, *B = A + (size_t)nxh * N2;
  double2 *TZ = B + (size_t)nxh * N2, *TX = TZ + N2;
  unsigned char *FS = reinterpret_cast<unsigned char *>(TX + N1);  // 1 on free-shear points
  stage_twiddles<double2>(TZ, reinterpret_cast<const double2 *>(arena + ch.o_twzN), N2);
  stage_twiddles<double2>(TX, reinterpret_cast<const double2 *>(arena + ch.o_twxN), N1);
  {
    const int *mk = v.mask + (size_t)w * N1 * N2;
    for (int t = tid; t < N1 * N2; t += nt) FS[t] = mk[t] == 2;
  }
  const size_t tab = (size_t)n * ldc;
  const double *Wr = v.wtab + (size_t)((kRK - 1) * 2 + w) * tab;
  const int gj = w == 0 ? 0 : n - 1;  // g = W[gj]: response of this row to this wall's source
  const double g0 = Wr[(size_t)gj * ldc];
  const double gw = w == 0 ? ch.gw1 : ch.gw2;
  const double Pw = 1.0 + gw, cfs = mode == 0 ? Pw / (1.0 - g0 * Pw) : 1.0 - (f == 0 ? ch.gwbc1 : ch.gwbc3);
  const double2 *src = f == 0 ? v.u1 : v.u3;
  double2 *S = A;  // Fourier plane [i][z], the current source

  const int passes = mode == 0 ? nit : 1;
  for (int it = 0; it < passes; it++) {
    for (int c = tid; c < nc; c += nt) {
      const int i = c % nxh, k = c / nxh;
      const size_t o = (size_t)i * N2 + k;
      double2 val;
      if (mode == 0) {
        const double2 U = v.texu[(size_t)(f * 2 + w) * nc + c];
        const double dg = Wr[(size_t)gj * ldc + c] - g0;
        const double2 sg = it > 0 ? S[o] : make_double2(0, 0);
        val = make_double2(U.x + dg * sg.x, U.y + dg * sg.y);
      } else {
        val = src[(size_t)(w == 0 ? 1 : n - 2) * ldc + c];
      }
      S[o] = val;
    }
    __syncthreads();
    // inverse z transforms of the nxh columns
    double2 *R = fft_lds<double2>(S, S == A ? B : A, N2, nxh, N2, +1, TZ);
    double2 *X = R == A ? B : A;
    // Hermitian extension in x, two rows per complex sequence (rft inverse:
    // imaginary parts of modes 0 and N1/2 ignored)
    for (int t = tid; t < (N2 / 2) * N1; t += nt) {
      const int m = t % N1, q = t / N1;
      const int mm = m <= h ? m : N1 - m;
      const double sg = m <= h ? 1.0 : -1.0;
      double2 a = R[(size_t)mm * N2 + 2 * q], b = R[(size_t)mm * N2 + 2 * q + 1];
      a.y *= sg;
      b.y *= sg;
      if (mm == 0 || mm == h) {
        a.y = 0.0;
        b.y = 0.0;
      }
      X[t] = make_double2(a.x - b.y, a.y + b.x);
    }
    __syncthreads();
    double2 *Y = fft_lds<double2>(X, R, N1, N2 / 2, N1, +1, TX);
    // pointwise mask in physical space: re -> row 2q, im -> row 2q+1
    for (int t = tid; t < (N2 / 2) * N1; t += nt) {
      const int x = t % N1, q = t / N1;
      const double2 val = Y[t];
      const double ca = FS[(2 * q) * N1 + x] ? cfs : 0.0, cb = FS[(2 * q + 1) * N1 + x] ? cfs : 0.0;
      Y[t] = make_double2(ca * val.x, cb * val.y);
    }
    __syncthreads();
    double2 *Z = fft_lds<double2>(Y, Y == A ? B : A, N1, N2 / 2, N1, -1, TX);
    double2 *F = Z == A ? B : A;
    // separate the two rows' half spectra (1/N1; imaginary parts of modes 0
    // and N1/2 zero) into the [i][z] layout
    const double inv = 1.0 / N1;
    for (int t = tid; t < (N2 / 2) * (h + 1); t += nt) {
      const int m = t % (h + 1), q = t / (h + 1);
      const double2 Ym = Z[(size_t)q * N1 + m], Yc = Z[(size_t)q * N1 + (m == 0 ? 0 : N1 - m)];
      double2 Pa, Pb;
      if (m == 0 || m == h) {
        Pa = make_double2(0.5 * (Ym.x + Yc.x) * inv, 0.0);
        Pb = make_double2(0.5 * (Ym.y + Yc.y) * inv, 0.0);
      } else {
        Pa = make_double2(0.5 * (Ym.x + Yc.x) * inv, 0.5 * (Ym.y - Yc.y) * inv);
        Pb = make_double2(0.5 * (Ym.y + Yc.y) * inv, -0.5 * (Ym.x - Yc.x) * inv);
      }
      F[(size_t)m * N2 + 2 * q] = Pa;
      F[(size_t)m * N2 + 2 * q + 1] = Pb;
    }
    __syncthreads();
    double2 *G = fft_lds<double2>(F, Z, N2, nxh, N2, -1, TZ);
    const double invz = 1.0 / N2;
    for (int t = tid; t < nxh * N2; t += nt) G[t] = make_double2(G[t].x * invz, G[t].y * invz);
    __syncthreads();
    S = G;
  }
  // wall row: source into r (mode 0), or ghost added to u (mode 1)
  double2 *dst = (mode == 0 ? (f == 0 ? v.r1 : v.r3) : (f == 0 ? v.u1 : v.u3)) + (w == 0 ? 0 : (size_t)(n - 1) * ldc);
  for (int c = tid; c < nc; c += nt) {
    const int i = c % nxh, k = c / nxh;
    const double2 val = S[(size_t)i * N2 + k];
    dst[c] = mode == 0 ? val : cadd(dst[c], val);
  }
}

// ------------------------------------------------------------------------
// Walls whose mask does not depend on x (streamwise ridges).  The mask
// product then commutes with the x transforms: for x-mode i the N-grid pair
// T[m . T^-1 v] reduces to Tz[m(z) . Tz^-1 v_i], except that rft keeps only
// the real part of the z-transformed lines of i = 0 and i = N1/2, so those
// two take the real part before the product.  The wall iteration (the same
// passes as k_tex_fused) therefore decouples per x-mode: one wave per
// (x-mode, field, wall, channel), an N2-point transform pair per pass.
__global__ void __launch_bounds__(64) k_tex_1d(char *arena, const Chan *chans, int kRK, int nit, int mode) {
  const Chan ch = chans[blockIdx.z];
  const int i = blockIdx.x, f = blockIdx.y >> 1, w = blockIdx.y & 1;
  if (!ch.tex1d || !(ch.tex & (1 << w)) || i >= ch.nxh) return;
  tex_1d_block<8, false>(arena, ch, i, f, w, kRK, nit, mode);
}


static int shape_n2max(const ShapeMax &m) {
  int n2 = 0;
  for (const ChanDims &d : m.dims) n2 = std::max(n2, d.N2);
  return n2;
}
static int shape_nxhmax(const ShapeMax &m) {
  int h = 0;
  for (const ChanDims &d : m.dims) h = std::max(h, d.N1 / 2 + 1);
  return h;
}

hipError_t launch_tex_1d(char *arena, const Chan *d_chans, const ShapeMax &m, int kRK, int nit, int mode,
                         hipStream_t s) {
  const int n2 = shape_n2max(m);
  hipLaunchKernelGGL(k_tex_1d, dim3(shape_nxhmax(m), 4, m.nchan), dim3(64), tex_1d_lds(n2), s, arena, d_chans, kRK,
                     nit, mode);
  return hipGetLastError();
}

size_t tex_fused_bytes(const std::vector<TexGroup> &groups) {
  size_t b = 0;
  for (const TexGroup &g : groups) b = std::max(b, tex_fused_lds(g.N1, g.N2));
  return b;
}

hipError_t launch_tex_sources_fused(char *arena, const Chan *d_chans, const ShapeMax &m, size_t lds, int kRK,
                                    int nit, hipStream_t s) {
  hipLaunchKernelGGL(k_tex_fused, dim3(4, 1, m.nchan), dim3(1024), lds, s, arena, d_chans, kRK, nit, 0);
  return hipGetLastError();
}

hipError_t launch_tex_ghosts_fused(char *arena, const Chan *d_chans, const ShapeMax &m, size_t lds, hipStream_t s) {
  hipLaunchKernelGGL(k_tex_fused, dim3(4, 1, m.nchan), dim3(1024), lds, s, arena, d_chans, 1, 1, 1);
  return hipGetLastError();
}

}  // namespace lh
