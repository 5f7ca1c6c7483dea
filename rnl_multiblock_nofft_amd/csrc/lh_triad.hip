// lh_triad.hip -- the restricted nonlinear path: products summed over triad
// lists in Fourier space (ops_in_planes, FOU3D.f90:1724-1987; nonlinInter,
// FOU3D.f90:1989-2030), in place of the FFT stages for channels that carry
// lists.
//
// The host turns every list row (la, ka, lp, kp) and its z mirror into
// (target column, passive column + conjugate flag, active column +
// conjugate flag) and sorts them by target (compressed rows per target
// column), so that a lane group owns a target and sums its products without
// atomics, in a fixed order (eight lanes per target striding its
// entries, a butterfly sum).  k_triad stages the plane's two operand rows in
// LDS and writes one product field per (plane, list type); k_triad_fin forms
// the x/z derivatives and the y-flux rows the right-hand-side kernel reads,
// exactly where k_zfwd writes them on the FFT path.
#include "lh_common.h"
#include "lh_kernels.h"

namespace lh {

namespace {

// stored z index of wavenumber index k (negative k in the upper half; the
// index N2/2 stands for both -N2/2 and +N2/2, as kLkup does)
__device__ __forceinline__ int kk_of(int k, int N2) { return k >= 0 ? k : k + N2; }

// value of the full (Hermitian-extended) spectrum at (i, k) from a stored row
__device__ __forceinline__ double2 full_at(const double2 *row, int i, int k, int nxh, int N2) {
  if (i >= 0) return row[i + nxh * kk_of(k, N2)];
  const double2 a = row[-i + nxh * kk_of(-k, N2)];
  return make_double2(a.x, -a.y);
}

// row c of field fid on the plane: u planes (j = 1..ny-1) 1 = u1, 2 = u2
// interpolated to the u row (interp_v), 3 = u3; v planes (j = 1..ny-2)
// 1 = u1 and 3 = u3 interpolated to the v row (interp_u), 2 = u2
// (FOU3D.f90:427-474)
__device__ __forceinline__ double2 plane_val(const View &v, const Chan &ch, bool isU, int j, int fid, int c) {
  const size_t L = ch.ldc;
  if (isU) {
    if (fid == 1) return v.u1[(size_t)j * L + c];
    if (fid == 3) return v.u3[(size_t)j * L + c];
    const double2 a = v.u2[(size_t)j * L + c], b = v.u2[(size_t)(j - 1) * L + c];
    const double w1 = v.yu[j] - v.yv[j - 1], w2 = v.yv[j] - v.yu[j], d = v.yv[j] - v.yv[j - 1];
    return make_double2((w1 * a.x + w2 * b.x) / d, (w1 * a.y + w2 * b.y) / d);
  }
  if (fid == 2) return v.u2[(size_t)j * L + c];
  const double2 *f = fid == 1 ? v.u1 : v.u3;
  const double2 a = f[(size_t)(j + 1) * L + c], b = f[(size_t)j * L + c];
  const double w1 = v.yv[j] - v.yu[j], w2 = v.yu[j + 1] - v.yv[j], d = v.yu[j + 1] - v.yu[j];
  return make_double2((w1 * a.x + w2 * b.x) / d, (w1 * a.y + w2 * b.y) / d);
}

__device__ __forceinline__ double2 cm(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

}  // namespace

// one workgroup per (plane slot, list type of the plane, channel)
__global__ void __launch_bounds__(256) k_triad(char *arena, const Chan *chans) {
  const Chan ch = chans[blockIdx.z];
  if (!ch.tri) return;
  const int slot = blockIdx.x, ti = blockIdx.y;
  if (slot >= ch.nslot) return;
  const bool isU = slot < ch.nplu;
  if (!isU && ti >= 4) return;
  const View v(arena, ch);
  const int type = isU ? 2 * ti + 1 : 2 * ti + 2;
  const int j = isU ? ch.pu0 + slot : ch.pv0 + slot - ch.nplu;
  const int ncol = ch.ncol, nxh = ch.nxh, N1 = ch.N1, N2 = ch.N2;
  // passive / active fields of each type (FOU3D.f90:1821-1825, 1953-1956)
  //   u: 1 uu (u1,u1) 3 uw (u3,u1) 5 vv (v,v) 7 wu (u1,u3) 9 ww (u3,u3)
  //   v: 2 uv (v,u1) 4 vu (u1,v) 6 vw (u3,v) 8 wv (v,u3)
  const int fa_tab[10] = {0, 1, 2, 3, 1, 2, 3, 1, 2, 3};
  const int fp_tab[10] = {0, 1, 1, 1, 2, 2, 2, 3, 3, 3};
  const int fa = fa_tab[type], fp = fp_tab[type];
  extern __shared__ double2 smem_d[];
  double2 *A = smem_d, *P = A + ncol;
  for (int c = threadIdx.x; c < ncol; c += blockDim.x) {
    A[c] = plane_val(v, ch, isU, j, fa, c);
    P[c] = plane_val(v, ch, isU, j, fp, c);
  }
  __syncthreads();
  // the mean-mode target of the y fluxes vv, uv, wv: sum over all (i, k) of
  // x(i, k) y(-i, -k) (FOU3D.f90:1763-1778, 1892-1908)
  __shared__ double2 red[4];
  double2 zsum = make_double2(0, 0);
  const bool zeroth = type == 5 || type == 2 || type == 8;
  if (zeroth) {
    // vv: (v, v), k from -N2/2; uv: (u1, v), wv: (u3, v), k from -(N2/2-1)
    const double2 *X = type == 5 ? A : P, *Y = A;
    const int kneg = type == 5 ? N2 / 2 : N2 / 2 - 1, nk = kneg + N2 / 2, h1 = N1 / 2;
    double2 s = make_double2(0, 0);
    for (int q = threadIdx.x; q < (N1 + 1) * nk; q += blockDim.x) {
      const int i = q / nk - h1, k = q % nk - kneg;
      const double2 a = full_at(X, i, k, nxh, N2), b = full_at(Y, -i, -k, nxh, N2);
      s = make_double2(s.x + (a.x * b.x - a.y * b.y), s.y + (a.x * b.y + a.y * b.x));
    }
    for (int off = 32; off >= 1; off >>= 1) {
      s.x += __shfl_xor(s.x, off);
      s.y += __shfl_xor(s.y, off);
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    zsum = make_double2((red[0].x + red[1].x) + (red[2].x + red[3].x), (red[0].y + red[1].y) + (red[2].y + red[3].y));
  }
  // linear advection by the mean mode (FOU3D.f90:1784-1798, 1913-1925):
  // buff = a00 p + p00 a, with the real parts of the (0,0) entries; the
  // squares uu, vv, ww have buff = 2 a00 a
  const double a00 = A[0].x, p00 = P[0].x;
  const int q = ch.tri_map[slot];
  const unsigned *off = nullptr, *ent = nullptr;
  if (q >= 0) {
    const size_t lt = (size_t)q * 9 + (type - 1);
    off = ch.tri_off + lt * (ncol + 1);
    ent = ch.tri_ent + ch.tri_base[lt];
  }
  double2 *out = ch.tri_prod + ((size_t)slot * 5 + ti) * ncol;
  const bool square = type == 1 || type == 5 || type == 9;
  // eight lanes per target column, eight targets per wave: each load
  // instruction fetches eight contiguous 32-byte runs of the table, and the
  // partial sums meet in a three-level butterfly
  const int lane = threadIdx.x & 63, grp = lane >> 3, sub = lane & 7;
  const int wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
  unsigned *OFF = reinterpret_cast<unsigned *>(P + ncol);
  if (off)
    for (int c = threadIdx.x; c <= ncol; c += blockDim.x) OFF[c] = off[c];
  __syncthreads();
  auto term = [&](unsigned w) -> double2 {
    double2 a = A[w & 0x7fffu], b = P[(w >> 16) & 0x7fffu];
    if (w & 0x8000u) a.y = -a.y;
    if (w & 0x80000000u) b.y = -b.y;
    return cm(a, b);
  };
  for (int c0 = wave * 8; c0 < ncol; c0 += nwave * 8) {
    const int c = c0 + grp;
    double2 acc = make_double2(0, 0);
    if (off && c < ncol) {
      const unsigned e1 = OFF[c + 1];
      unsigned e = OFF[c] + sub;
      for (; e + 24 < e1; e += 32) {
        const unsigned w0 = ent[e], w1 = ent[e + 8], w2 = ent[e + 16], w3 = ent[e + 24];
        const double2 t0 = term(w0), t1 = term(w1), t2 = term(w2), t3 = term(w3);
        acc = make_double2(acc.x + ((t0.x + t1.x) + (t2.x + t3.x)), acc.y + ((t0.y + t1.y) + (t2.y + t3.y)));
      }
      for (; e < e1; e += 8) {
        const double2 t0 = term(ent[e]);
        acc = make_double2(acc.x + t0.x, acc.y + t0.y);
      }
    }
    for (int o = 4; o >= 1; o >>= 1) {
      acc.x += __shfl_xor(acc.x, o);
      acc.y += __shfl_xor(acc.y, o);
    }
    if (sub == 0 && c < ncol) {
      double2 x;
      if (square) {
        x = make_double2(2.0 * (a00 * A[c].x), 2.0 * (a00 * A[c].y));
      } else if (type == 3 || type == 2 || type == 6) {  // buff = p00 a + a00 p
        x = make_double2(p00 * A[c].x + a00 * P[c].x, p00 * A[c].y + a00 * P[c].y);
      } else {  // 7, 4, 8: buff = a00 p + p00 a
        x = make_double2(a00 * P[c].x + p00 * A[c].x, a00 * P[c].y + p00 * A[c].y);
      }
      if (c == 0 && zeroth) x = zsum;
      out[c] = make_double2(x.x + acc.x, x.y + acc.y);
    }
  }
}

// derivatives and flux rows (der_x + der_z, FOU3D.f90:476-593, with the
// z index N2/2 removed as they do); grid (column blocks, slots, channels)
__global__ void __launch_bounds__(256) k_triad_fin(char *arena, const Chan *chans) {
  const Chan ch = chans[blockIdx.z];
  if (!ch.tri) return;
  const int slot = blockIdx.y;
  if (slot >= ch.nslot) return;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ch.ncol) return;
  const View v(arena, ch);
  const bool isU = slot < ch.nplu;
  const int j = isU ? ch.pu0 + slot : ch.pv0 + slot - ch.nplu;
  const int ncol = ch.ncol, i = c % ch.nxh, k = c / ch.nxh;
  const bool live = k != ch.N2 / 2;
  const double kx = v.kx[i], kz = v.kz[k];
  const double2 *pr = ch.tri_prod + (size_t)slot * 5 * ncol + c;
  const size_t o = (size_t)j * ch.ldc + c;
  const double2 z = make_double2(0, 0);
  if (isU) {  // uu, uw, vv, wu, ww
    const double2 uu = pr[0], uw = pr[ncol], vv = pr[2 * ncol], wu = pr[3 * ncol], ww = pr[4 * ncol];
    v.nl1[o] = live ? cadd(cimul(kx, uu), cimul(kz, uw)) : z;
    v.nl3[o] = live ? cadd(cimul(kx, wu), cimul(kz, ww)) : z;
    v.fvv[o] = vv;
  } else {  // uv, vu, vw, wv
    const double2 uv = pr[0], vu = pr[ncol], vw = pr[2 * ncol], wv = pr[3 * ncol];
    v.nl2[o] = live ? cadd(cimul(kx, vu), cimul(kz, vw)) : z;
    v.fuv[o] = uv;
    v.fwv[o] = wv;
  }
}

hipError_t launch_triads(char *arena, const Chan *d_chans, const ShapeMax &m, hipStream_t s) {
  const size_t lds = 2 * (size_t)m.ncol * sizeof(double2) + ((size_t)m.ncol + 1) * sizeof(unsigned);
  hipLaunchKernelGGL(k_triad, dim3(m.nslot, 5, m.nchan), dim3(256), lds, s, arena, d_chans);
  hipLaunchKernelGGL(k_triad_fin, dim3((m.ncol + 255) / 256, m.nslot, m.nchan), dim3(256), 0, s, arena, d_chans);
  return hipGetLastError();
}

}  // namespace lh
