// lh_stats.hip -- running statistics and spectra of the flow on the device.
//
// Replaces the reference's stats (stats.f90:1-107), its gathers recvX
// (stats.f90:443-470) and spectra/buildsp/buildspUV (spectra.f90:1-56).  The
// reference sums physical planes that the nonlinear step left behind:
//
//   u rows 1..ny-1   u1PL, u3PL, u2PL_itp in physical space on the
//                    Ngal1 x Ngal2 grid (ops_in_planes2, FOU3D.f90:2116)
//   v rows 1..ny-2   u2PL, u3PL_itp likewise (FOU3D.f90:2151)
//   u rows 0, ny     (ghosts) and v rows 0, ny-1 (walls): the planes still
//                    hold the spectral coefficients (modes_to_planes_UVP
//                    fills them, FOU3D.f90:1000-1098; the transform loops
//                    skip them), so the sums run over real and imaginary
//                    parts of the stored modes
//   p rows 1..ny-1   ppPL as modes_to_planes_UVP wrote it: spectral
//                    (FOU3D.f90:156-157)
//   wx, v rows       omega_x = -d(u2)/dz + d(u3)/dy on the N1 x N2 grid,
//                    physical (FOU3D.f90:302-327)
//
// Physical-plane sums are evaluated in spectral space with Parseval's
// identity, so no plane is transformed: for the unscaled inverse transforms
// of the reference (cft/rft), sum_grid f = n Re f(0,0) and
// sum_grid f g = n sum Re(F conj G) over the full Hermitian spectrum, n the
// number of grid points.  A stored column (i > 0) stands for the modes
// (i, k) and (-i, -k): weight 2.  The i = 0 column (and, on the N grid, the
// x-Nyquist column i = N1/2, of which rft keeps the real part) enters as its
// Hermitian part h(k) = (f(k) + conj f(-k)) / 2, because the real inverse
// transform in x keeps only the real part of the z-transformed line.  On the
// Ngal2 grid the stored z index N2/2 is the mode -N2/2, whose partner +N2/2
// is not stored: h = f/2 at both, i.e. weight 1/2.
//
// One workgroup per (row j, channel) reduces every quantity of that row;
// results are added to the per-row accumulators, so the result does not
// depend on scheduling.
#include "lh_kernels.h"

namespace lh {

namespace {

struct StatPt {
  double2 u1, u3, u2i;  // u row j: u1, u3 and u2 interpolated to the u row
  double2 v, u3i;       // v row j: u2 and u3 interpolated to the v row
  double2 wx;           // v row j: omega_x coefficient
  double2 p;            // p row j
};

// point values of column c at reference row j (interp_u / interp_v,
// FOU3D.f90:427-474; der_z_N FOU3D.f90:595-634; der_yv_h_wx FOU3D.f90:678-697)
__device__ __forceinline__ StatPt stat_point(const Chan &ch, const View &v, int j, int c) {
  StatPt s;
  const size_t L = ch.ldc;
  const int ny = ch.ny;
  const double2 z = make_double2(0.0, 0.0);
  s.u1 = v.u1[(size_t)j * L + c];
  s.u3 = v.u3[(size_t)j * L + c];
  // u2 at the u row: linear in the v rows around it; the ghost rows use the
  // grid weights of interp_v (gridweighting_interp, start.f90:553-559)
  auto itp_v = [&](int jj) -> double2 {
    const double2 a = v.u2[(size_t)jj * L + c], b = v.u2[(size_t)(jj - 1) * L + c];
    const double w1 = v.yu[jj] - v.yv[jj - 1], w2 = v.yv[jj] - v.yu[jj], d = v.yv[jj] - v.yv[jj - 1];
    return make_double2((w1 * a.x + w2 * b.x) / d, (w1 * a.y + w2 * b.y) / d);
  };
  if (j >= 1 && j <= ny - 1) {
    s.u2i = itp_v(j);
  } else if (j == 0) {
    const double2 o1 = itp_v(1), w = v.u2[c];
    s.u2i = make_double2(ch.gwi1 * (w.x - o1.x) + o1.x, ch.gwi1 * (w.y - o1.y) + o1.y);
  } else {
    const double2 o1 = itp_v(ny - 1), w = v.u2[(size_t)(ny - 1) * L + c];
    s.u2i = make_double2(ch.gwi2 * (w.x - o1.x) + o1.x, ch.gwi2 * (w.y - o1.y) + o1.y);
  }
  if (j <= ny - 1) {
    s.v = v.u2[(size_t)j * L + c];
    const double2 a = v.u3[(size_t)(j + 1) * L + c], b = s.u3;
    const double w1 = v.yv[j] - v.yu[j], w2 = v.yu[j + 1] - v.yv[j], d = v.yu[j + 1] - v.yu[j];
    s.u3i = make_double2((w1 * a.x + w2 * b.x) / d, (w1 * a.y + w2 * b.y) / d);
    const int kk = c / ch.nxh;
    const double2 dz = cimul(v.kz[kk], s.v);
    const double dy = v.dthdyv[j];
    s.wx = make_double2(-dz.x + (a.x - b.x) * dy, -dz.y + (a.y - b.y) * dy);
  } else {
    s.v = s.u3i = s.wx = z;
  }
  s.p = (j >= 1 && j <= ny - 1) ? v.p[(size_t)j * L + c] : z;
  return s;
}

__device__ __forceinline__ double dotc(double2 a, double2 b) { return a.x * b.x + a.y * b.y; }
// Hermitian part (a + conj ap) / 2
__device__ __forceinline__ double2 herm(double2 a, double2 ap) {
  return make_double2(0.5 * (a.x + ap.x), 0.5 * (a.y - ap.y));
}

enum {
  // spectral-layout sums: sum(re + im) and sum(re re' + im im') over stored modes
  A_U1, A_U1U1, A_U3, A_U3U3, A_U1U2I, A_U1U3, A_V, A_VV, A_VU3I, A_P, A_PP,
  // Parseval sums (without the grid-size factor)
  B_U1U1, B_U3U3, B_U1U2I, B_U1U3, B_VV, B_VU3I, B_WXWX,
  NACC
};

__global__ void __launch_bounds__(256) k_stats(char *arena, const Chan *chans, char *sbase) {
  const Chan ch = chans[blockIdx.y];
  const int j = blockIdx.x;
  if (j >= ch.nyu) return;
  View v(arena, ch);
  const int ny = ch.ny, nxh = ch.nxh, N2 = ch.N2, N1h = ch.N1 / 2;
  double acc[NACC];
#pragma unroll
  for (int q = 0; q < NACC; q++) acc[q] = 0.0;
  double *spec = (double *)(sbase + ch.o_spec);
  const size_t L = ch.ldc, plane = (size_t)ch.nyu * L;
  for (int c = threadIdx.x; c < ch.ncol; c += blockDim.x) {
    const StatPt s = stat_point(ch, v, j, c);
    const int i = c % nxh, kk = c / nxh;
    // spectral-layout sums
    acc[A_U1] += s.u1.x + s.u1.y;
    acc[A_U1U1] += dotc(s.u1, s.u1);
    acc[A_U3] += s.u3.x + s.u3.y;
    acc[A_U3U3] += dotc(s.u3, s.u3);
    acc[A_U1U2I] += dotc(s.u1, s.u2i);
    acc[A_U1U3] += dotc(s.u1, s.u3);
    acc[A_V] += s.v.x + s.v.y;
    acc[A_VV] += dotc(s.v, s.v);
    acc[A_VU3I] += dotc(s.v, s.u3i);
    acc[A_P] += s.p.x + s.p.y;
    acc[A_PP] += dotc(s.p, s.p);
    // Parseval sums on the Ngal grid
    if (i > 0) {
      acc[B_U1U1] += 2.0 * dotc(s.u1, s.u1);
      acc[B_U3U3] += 2.0 * dotc(s.u3, s.u3);
      acc[B_U1U2I] += 2.0 * dotc(s.u1, s.u2i);
      acc[B_U1U3] += 2.0 * dotc(s.u1, s.u3);
      acc[B_VV] += 2.0 * dotc(s.v, s.v);
      acc[B_VU3I] += 2.0 * dotc(s.v, s.u3i);
    } else if (kk == N2 / 2) {
      acc[B_U1U1] += 0.5 * dotc(s.u1, s.u1);
      acc[B_U3U3] += 0.5 * dotc(s.u3, s.u3);
      acc[B_U1U2I] += 0.5 * dotc(s.u1, s.u2i);
      acc[B_U1U3] += 0.5 * dotc(s.u1, s.u3);
      acc[B_VV] += 0.5 * dotc(s.v, s.v);
      acc[B_VU3I] += 0.5 * dotc(s.v, s.u3i);
    } else {
      const StatPt t = stat_point(ch, v, j, ((N2 - kk) % N2) * nxh);
      const double2 h1 = herm(s.u1, t.u1), h3 = herm(s.u3, t.u3), h2i = herm(s.u2i, t.u2i);
      const double2 hv = herm(s.v, t.v), h3i = herm(s.u3i, t.u3i);
      acc[B_U1U1] += dotc(h1, h1);
      acc[B_U3U3] += dotc(h3, h3);
      acc[B_U1U2I] += dotc(h1, h2i);
      acc[B_U1U3] += dotc(h1, h3);
      acc[B_VV] += dotc(hv, hv);
      acc[B_VU3I] += dotc(hv, h3i);
    }
    // omega_x on the N grid: the x-Nyquist column is real-projected too, and
    // the z index N2/2 is its own partner
    if (i > 0 && i < N1h) {
      acc[B_WXWX] += 2.0 * dotc(s.wx, s.wx);
    } else {
      const StatPt t = stat_point(ch, v, j, ((N2 - kk) % N2) * nxh + i);
      const double2 hw = herm(s.wx, t.wx);
      acc[B_WXWX] += dotc(hw, hw);
    }
    // spectra: sp += |f|^2 (buildsp) and re re' + im im' (buildspUV)
    double *su = spec + SP_U * plane + (size_t)j * L + c;
    *su = *su + s.u1.x * s.u1.x + s.u1.y * s.u1.y;
    double *sw = spec + SP_W * plane + (size_t)j * L + c;
    *sw = *sw + s.u3.x * s.u3.x + s.u3.y * s.u3.y;
    double *suv = spec + SP_UV * plane + (size_t)j * L + c;
    *suv = *suv + s.u1.x * s.u2i.x + s.u1.y * s.u2i.y;
    if (j <= ny - 1) {
      double *sv = spec + SP_V * plane + (size_t)j * L + c;
      *sv = *sv + s.v.x * s.v.x + s.v.y * s.v.y;
    }
    if (j >= 1 && j <= ny - 1) {
      double *sp = spec + SP_P * plane + (size_t)j * L + c;
      *sp = *sp + s.p.x * s.p.x + s.p.y * s.p.y;
    }
  }
  // workgroup reduction: 64-lane butterflies, then the 4 wave partials
  __shared__ double red[4][NACC];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NACC; q++) {
    double x = acc[q];
    for (int off = 32; off >= 1; off >>= 1) x += __shfl_xor(x, off);
    if (lane == 0) red[wave][q] = x;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double r[NACC];
#pragma unroll
  for (int q = 0; q < NACC; q++) r[q] = (red[0][q] + red[1][q]) + (red[2][q] + red[3][q]);
  double *prof = (double *)(sbase + ch.o_stat);
  const int n = ch.nyu;
  auto add = [&](int q, double x) { prof[(size_t)q * n + j] += x; };
  const double nG = (double)ch.Ng1 * ch.Ng2, nN = (double)ch.N1 * ch.N2;
  if (j >= 1 && j <= ny - 1) {  // physical u rows
    add(ST_UM, nG * v.u1[(size_t)j * L].x);
    add(ST_U2M, nG * r[B_U1U1]);
    add(ST_WM, nG * v.u3[(size_t)j * L].x);
    add(ST_W2M, nG * r[B_U3U3]);
    add(ST_UVM, nG * r[B_U1U2I]);
    add(ST_UWM, nG * r[B_U1U3]);
  } else {  // ghost rows: spectral planes
    add(ST_UM, r[A_U1]);
    add(ST_U2M, r[A_U1U1]);
    add(ST_WM, r[A_U3]);
    add(ST_W2M, r[A_U3U3]);
    add(ST_UVM, r[A_U1U2I]);
    add(ST_UWM, r[A_U1U3]);
  }
  if (j <= ny - 1) {
    if (j >= 1 && j <= ny - 2) {  // physical v rows
      add(ST_VM, nG * v.u2[(size_t)j * L].x);
      add(ST_V2M, nG * r[B_VV]);
      add(ST_VWM, nG * r[B_VU3I]);
    } else {  // wall rows: spectral planes
      add(ST_VM, r[A_V]);
      add(ST_V2M, r[A_VV]);
      add(ST_VWM, r[A_VU3I]);
    }
    // omega_x(0,0) = d(u3)/dy of the mean mode (kz = 0)
    const double2 a = v.u3[(size_t)(j + 1) * L], b = v.u3[(size_t)j * L];
    add(ST_WXM, nN * ((a.x - b.x) * v.dthdyv[j]));
    add(ST_WX2M, nN * r[B_WXWX]);
  }
  if (j >= 1 && j <= ny - 1) {
    add(ST_PM, r[A_P]);
    add(ST_P2M, r[A_PP]);
  }
}

}  // namespace

// sl_stats (sl_stats2.f90:33-161): u1, u3 interpolated to the wall faces
// (v rows 0 and ny-1, interp_u) and their wall-normal derivatives, the top
// one with the sign flipped, summed per mode; one thread per column
__global__ void __launch_bounds__(256) k_slstats(char *arena, const Chan *chans, char *sbase) {
  const Chan ch = chans[blockIdx.y];
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ch.ncol) return;
  const View v(arena, ch);
  const size_t L = ch.ldc;
  double2 *sl = (double2 *)(sbase + ch.o_slip);
  for (int w = 0; w < 2; w++) {
    const int j = w == 0 ? 0 : ch.nyv - 1;
    const double sg = w == 0 ? 1.0 : -1.0;
    const double w1 = v.yv[j] - v.yu[j], w2 = v.yu[j + 1] - v.yv[j], d = v.yu[j + 1] - v.yu[j];
    const double dy = v.dthdyv[j];
    for (int f = 0; f < 2; f++) {
      const double2 *u = f == 0 ? v.u1 : v.u3;
      const double2 a = u[(size_t)(j + 1) * L + c], b = u[(size_t)j * L + c];
      const double2 ui = make_double2((w1 * a.x + w2 * b.x) / d, (w1 * a.y + w2 * b.y) / d);
      const double2 du = make_double2(sg * ((a.x - b.x) * dy), sg * ((a.y - b.y) * dy));
      double2 *o = sl + ((size_t)(w * 4 + 2 * f) * ch.ncol + c);
      o[0] = cadd(o[0], ui);
      o[ch.ncol] = cadd(o[ch.ncol], du);
    }
  }
}

hipError_t launch_stats(char *arena, const Chan *d_chans, const ShapeMax &m, char *sbase, hipStream_t s) {
  hipLaunchKernelGGL(k_stats, dim3(m.nyu, m.nchan), dim3(256), 0, s, arena, d_chans, sbase);
  hipLaunchKernelGGL(k_slstats, dim3((m.ncol + 255) / 256, m.nchan), dim3(256), 0, s, arena, d_chans, sbase);
  return hipGetLastError();
}

}  // namespace lh
