// lh_fft.h -- batched complex FFTs in LDS for gfx950 (Stockham autosort,
// mixed radix 3/4/2), executed cooperatively by all threads of a workgroup.
//
// These replace FFTPACK's cfftf/cfftb and rfftf/rfftb (cft_buff.f,
// rft_buff.f) as used by four_to_phys_* / phys_to_four_* (FOU3D.f90:738-804,
// 2253-2293).  Conventions: sign -1 is the forward transform
// sum_j x_j exp(-2 pi i jk/n), sign +1 the backward one; no scaling here.
// Every length the solver meets is 3^a * 2^b.
//
// Twiddles come from a table tw[m] = exp(-2 pi i m/n) that the calling
// workgroup stages in LDS once (stage_twiddles) from the FP64 table in HBM;
// the FP32 instantiation rounds each entry once to FP32.
#pragma once
#include "lh_common.h"

namespace lh {

template <typename C> struct CT;
template <> struct CT<double2> {
  typedef double R;
  static __device__ __forceinline__ double2 mk(double a, double b) { return make_double2(a, b); }
  static __device__ __forceinline__ double2 from(double2 w) { return w; }
};
template <> struct CT<float2> {
  typedef float R;
  static __device__ __forceinline__ float2 mk(float a, float b) { return make_float2(a, b); }
  static __device__ __forceinline__ float2 from(double2 w) { return make_float2((float)w.x, (float)w.y); }
};
// twiddle m of the table, conjugated for the backward transform
template <typename C> __device__ __forceinline__ C tw_at(const C *t, int m, int sign) {
  C w = t[m];
  return CT<C>::mk(w.x, sign > 0 ? -w.y : w.y);
}
// copy the FP64 table src[0..n) into LDS at dst (no barrier: the caller's
// next __syncthreads publishes it)
template <typename C> __device__ __forceinline__ void stage_twiddles(C *dst, const double2 *src, int n) {
  for (int m = threadIdx.x; m < n; m += blockDim.x) dst[m] = CT<C>::from(src[m]);
}

template <typename C> __device__ __forceinline__ C c_add(C a, C b) { return CT<C>::mk(a.x + b.x, a.y + b.y); }
template <typename C> __device__ __forceinline__ C c_sub(C a, C b) { return CT<C>::mk(a.x - b.x, a.y - b.y); }
template <typename C> __device__ __forceinline__ C c_mul(C a, C b) {
  return CT<C>::mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// number of Stockham stages and the radix of stage s for length n
__host__ __device__ inline int fft_radix(int n, int s) {
  int m = n, st = 0;
  while (m % 3 == 0) { if (st == s) return 3; m /= 3; st++; }
  while (m % 4 == 0) { if (st == s) return 4; m /= 4; st++; }
  while (m % 2 == 0) { if (st == s) return 2; m /= 2; st++; }
  return 0;
}

// One radix-R butterfly of a Stockham stage: reads x[jj + r*nbf] of one
// sequence, applies the twiddles of output group k = jj % Ns and writes
// y[idxD + r*Ns].
template <typename C, int R, int SIGN>
__device__ __forceinline__ void butterfly(const C *xs, C *ys, int jj, int nbf, int Ns, int k, int idxD, int span,
                                          const C *tw) {
  typedef typename CT<C>::R Rt;
  const Rt s3 = (Rt)0.86602540378443864676;  // sqrt(3)/2
  if (R == 4) {
    C v0 = xs[jj], v1 = xs[jj + nbf], v2 = xs[jj + 2 * nbf], v3 = xs[jj + 3 * nbf];
    if (Ns > 1) {
      int m = k * span;
      v1 = c_mul(v1, tw_at<C>(tw, m, SIGN));
      v2 = c_mul(v2, tw_at<C>(tw, 2 * m, SIGN));
      v3 = c_mul(v3, tw_at<C>(tw, 3 * m, SIGN));
    }
    C t0 = c_add(v0, v2), t1 = c_sub(v0, v2), t2 = c_add(v1, v3), d = c_sub(v1, v3);
    C t3 = SIGN < 0 ? CT<C>::mk(d.y, -d.x) : CT<C>::mk(-d.y, d.x);  // d * (sign*i)
    ys[idxD] = c_add(t0, t2);
    ys[idxD + Ns] = c_add(t1, t3);
    ys[idxD + 2 * Ns] = c_sub(t0, t2);
    ys[idxD + 3 * Ns] = c_sub(t1, t3);
  } else if (R == 3) {
    C v0 = xs[jj], v1 = xs[jj + nbf], v2 = xs[jj + 2 * nbf];
    if (Ns > 1) {
      int m = k * span;
      v1 = c_mul(v1, tw_at<C>(tw, m, SIGN));
      v2 = c_mul(v2, tw_at<C>(tw, 2 * m, SIGN));
    }
    C tt = c_add(v1, v2), d = c_sub(v1, v2);
    C mm = CT<C>::mk(v0.x - (Rt)0.5 * tt.x, v0.y - (Rt)0.5 * tt.y);
    // e = sign*i*s3*d
    C e = SIGN < 0 ? CT<C>::mk(s3 * d.y, -s3 * d.x) : CT<C>::mk(-s3 * d.y, s3 * d.x);
    ys[idxD] = c_add(v0, tt);
    ys[idxD + Ns] = c_add(mm, e);
    ys[idxD + 2 * Ns] = c_sub(mm, e);
  } else {
    C v0 = xs[jj], v1 = xs[jj + nbf];
    if (Ns > 1) v1 = c_mul(v1, tw_at<C>(tw, k * span, SIGN));
    ys[idxD] = c_add(v0, v1);
    ys[idxD + Ns] = c_sub(v0, v1);
  }
}

// In-LDS batched FFT of any length 3^a 4^b 2^c.  nb sequences of length n;
// element e of sequence b is at x[b*sstr + e].  y is scratch of the same
// extent, tw the staged twiddle table.  Returns the buffer holding the
// result (x or y).  Must be called by all threads of the block.
template <typename C, int SIGN>
__device__ __forceinline__ C *fft_lds_rt(C *x, C *y, int n, int nb, int sstr, const C *tw) {
  int Ns = 1;
  for (int st = 0;; st++) {
    int Rr = fft_radix(n, st);
    if (Rr == 0) break;
    int nbf = n / Rr;
    int span = n / (Ns * Rr);  // twiddle exponent unit
    int total = nb * nbf;
    for (int t = threadIdx.x; t < total; t += blockDim.x) {
      int b = t / nbf, jj = t - b * nbf;
      int k = jj % Ns;
      int idxD = (jj / Ns) * Ns * Rr + k;
      if (Rr == 4)
        butterfly<C, 4, SIGN>(x + b * sstr, y + b * sstr, jj, nbf, Ns, k, idxD, span, tw);
      else if (Rr == 3)
        butterfly<C, 3, SIGN>(x + b * sstr, y + b * sstr, jj, nbf, Ns, k, idxD, span, tw);
      else
        butterfly<C, 2, SIGN>(x + b * sstr, y + b * sstr, jj, nbf, Ns, k, idxD, span, tw);
    }
    __syncthreads();
    C *tmp = x;
    x = y;
    y = tmp;
    Ns *= Rr;
  }
  return x;
}

// The same with the length a compile-time constant N: the radix sequence
// (3s, then 4s, then a 2) unrolls and every index division is by a constant.
template <typename C, int N, int M, int Ns, int SIGN>
__device__ __forceinline__ C *fft_ct_run(C *x, C *y, int nb, int sstr, const C *tw) {
  if constexpr (M == 1) {
    return x;
  } else {
    constexpr int R = (M % 3 == 0) ? 3 : ((M % 4 == 0) ? 4 : 2);
    constexpr int nbf = N / R;
    constexpr int span = N / (Ns * R);
    const int total = nb * nbf;
    for (int t = threadIdx.x; t < total; t += blockDim.x) {
      const int b = t / nbf, jj = t - b * nbf;
      const int k = jj % Ns;
      const int idxD = (jj / Ns) * (Ns * R) + k;
      butterfly<C, R, SIGN>(x + b * sstr, y + b * sstr, jj, nbf, Ns, k, idxD, span, tw);
    }
    __syncthreads();
    return fft_ct_run<C, N, M / R, Ns * R, SIGN>(y, x, nb, sstr, tw);
  }
}

// Dispatch on the lengths the solver meets (Ngal = 3*2^k or 2^k); other
// lengths take the run-time-plan version.
template <typename C, int SIGN>
__device__ __forceinline__ C *fft_lds_s(C *x, C *y, int n, int nb, int sstr, const C *tw) {
  switch (n) {
    case 12: return fft_ct_run<C, 12, 12, 1, SIGN>(x, y, nb, sstr, tw);
    case 24: return fft_ct_run<C, 24, 24, 1, SIGN>(x, y, nb, sstr, tw);
    case 48: return fft_ct_run<C, 48, 48, 1, SIGN>(x, y, nb, sstr, tw);
    case 96: return fft_ct_run<C, 96, 96, 1, SIGN>(x, y, nb, sstr, tw);
    case 192: return fft_ct_run<C, 192, 192, 1, SIGN>(x, y, nb, sstr, tw);
    case 384: return fft_ct_run<C, 384, 384, 1, SIGN>(x, y, nb, sstr, tw);
    case 768: return fft_ct_run<C, 768, 768, 1, SIGN>(x, y, nb, sstr, tw);
    case 8: return fft_ct_run<C, 8, 8, 1, SIGN>(x, y, nb, sstr, tw);
    case 16: return fft_ct_run<C, 16, 16, 1, SIGN>(x, y, nb, sstr, tw);
    case 32: return fft_ct_run<C, 32, 32, 1, SIGN>(x, y, nb, sstr, tw);
    case 64: return fft_ct_run<C, 64, 64, 1, SIGN>(x, y, nb, sstr, tw);
    case 128: return fft_ct_run<C, 128, 128, 1, SIGN>(x, y, nb, sstr, tw);
    case 256: return fft_ct_run<C, 256, 256, 1, SIGN>(x, y, nb, sstr, tw);
    default: return fft_lds_rt<C, SIGN>(x, y, n, nb, sstr, tw);
  }
}

// sign -1: forward transform, +1: backward (unscaled).  Everything is
// inlined into the calling kernel: a called function would receive generic
// pointers and reach LDS through FLAT instructions.
template <typename C>
__device__ __forceinline__ C *fft_lds(C *x, C *y, int n, int nb, int sstr, int sign, const C *tw) {
  return sign < 0 ? fft_lds_s<C, -1>(x, y, n, nb, sstr, tw) : fft_lds_s<C, +1>(x, y, n, nb, sstr, tw);
}

}  // namespace lh
