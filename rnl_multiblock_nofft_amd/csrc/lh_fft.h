// lh_fft.h -- batched complex FFTs in LDS for gfx950 (Stockham autosort,
// mixed radix 3/4/2), executed cooperatively by all threads of a workgroup.
//
// These replace FFTPACK's cfftf/cfftb and rfftf/rfftb (cft_buff.f,
// rft_buff.f) as used by four_to_phys_* / phys_to_four_* (FOU3D.f90:738-804,
// 2253-2293).  Conventions: sign -1 is the forward transform
// sum_j x_j exp(-2 pi i jk/n), sign +1 the backward one; no scaling here.
// Every length the solver meets is 3^a * 2^b.
//
// Twiddles always come from an FP64 table tw[m] = exp(-2 pi i m/n) held in
// HBM (L2/L1 resident); the FP32 instantiation rounds them once to FP32.
#pragma once
#include "lh_common.h"

namespace lh {

template <typename C> struct CT;
template <> struct CT<double2> {
  typedef double R;
  static __device__ __forceinline__ double2 mk(double a, double b) { return make_double2(a, b); }
  static __device__ __forceinline__ double2 tw(const double2 *t, int m, int sign) {
    double2 w = t[m];
    return make_double2(w.x, sign > 0 ? -w.y : w.y);
  }
};
template <> struct CT<float2> {
  typedef float R;
  static __device__ __forceinline__ float2 mk(float a, float b) { return make_float2(a, b); }
  static __device__ __forceinline__ float2 tw(const double2 *t, int m, int sign) {
    double2 w = t[m];
    return make_float2((float)w.x, (float)(sign > 0 ? -w.y : w.y));
  }
};

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

// In-LDS batched FFT.  nb sequences of length n; element e of sequence b is
// at x[b*sstr + e].  y is scratch of the same extent.  Returns the buffer
// holding the result (x or y).  Must be called by all threads of the block.
template <typename C>
__device__ C *fft_lds(C *x, C *y, int n, int nb, int sstr, int sign, const double2 *tw) {
  typedef typename CT<C>::R R;
  const R s3 = (R)0.86602540378443864676;  // sqrt(3)/2
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
      const C *xs = x + b * sstr;
      C *ys = y + b * sstr;
      int idxD = (jj / Ns) * Ns * Rr + k;
      if (Rr == 4) {
        C v0 = xs[jj], v1 = xs[jj + nbf], v2 = xs[jj + 2 * nbf], v3 = xs[jj + 3 * nbf];
        if (Ns > 1) {
          int m = k * span;
          v1 = c_mul(v1, CT<C>::tw(tw, m, sign));
          v2 = c_mul(v2, CT<C>::tw(tw, 2 * m, sign));
          v3 = c_mul(v3, CT<C>::tw(tw, 3 * m, sign));
        }
        C t0 = c_add(v0, v2), t1 = c_sub(v0, v2), t2 = c_add(v1, v3), d = c_sub(v1, v3);
        C t3 = sign < 0 ? CT<C>::mk(d.y, -d.x) : CT<C>::mk(-d.y, d.x);  // d * (sign*i)
        ys[idxD] = c_add(t0, t2);
        ys[idxD + Ns] = c_add(t1, t3);
        ys[idxD + 2 * Ns] = c_sub(t0, t2);
        ys[idxD + 3 * Ns] = c_sub(t1, t3);
      } else if (Rr == 3) {
        C v0 = xs[jj], v1 = xs[jj + nbf], v2 = xs[jj + 2 * nbf];
        if (Ns > 1) {
          int m = k * span;
          v1 = c_mul(v1, CT<C>::tw(tw, m, sign));
          v2 = c_mul(v2, CT<C>::tw(tw, 2 * m, sign));
        }
        C tt = c_add(v1, v2), d = c_sub(v1, v2);
        C mm = CT<C>::mk(v0.x - (R)0.5 * tt.x, v0.y - (R)0.5 * tt.y);
        // e = sign*i*s3*d
        C e = sign < 0 ? CT<C>::mk(s3 * d.y, -s3 * d.x) : CT<C>::mk(-s3 * d.y, s3 * d.x);
        ys[idxD] = c_add(v0, tt);
        ys[idxD + Ns] = c_add(mm, e);
        ys[idxD + 2 * Ns] = c_sub(mm, e);
      } else {
        C v0 = xs[jj], v1 = xs[jj + nbf];
        if (Ns > 1) v1 = c_mul(v1, CT<C>::tw(tw, k * span, sign));
        ys[idxD] = c_add(v0, v1);
        ys[idxD + Ns] = c_sub(v0, v1);
      }
    }
    __syncthreads();
    C *tmp = x;
    x = y;
    y = tmp;
    Ns *= Rr;
  }
  return x;
}

}  // namespace lh
