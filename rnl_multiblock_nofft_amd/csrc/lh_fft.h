// lh_fft.h -- batched complex FFTs in LDS for gfx950 (Stockham autosort),
// executed cooperatively by all threads of a workgroup.  Each stage is one
// LDS round trip: a thread reads R points, twiddles them, runs an R-point
// DFT in registers (R = 6, 8, 12, 16 built from 2, 3, 4 with compile-time
// twiddles) and writes them back, so a 96-point transform takes two stages
// (12 x 8) instead of four radix-3/4/2 passes.
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
#include <type_traits>

#include "lh_common.h"

namespace lh {

template <typename C> struct CT;
template <> struct CT<double2> {
  typedef double R;
  static __host__ __device__ __forceinline__ double2 mk(double a, double b) { return make_double2(a, b); }
  static __host__ __device__ __forceinline__ double2 from(double2 w) { return w; }
};
template <> struct CT<float2> {
  typedef float R;
  static __host__ __device__ __forceinline__ float2 mk(float a, float b) { return make_float2(a, b); }
  static __host__ __device__ __forceinline__ float2 from(double2 w) { return make_float2((float)w.x, (float)w.y); }
};
// twiddle m of the table, conjugated for the backward transform
template <typename C> __host__ __device__ __forceinline__ C tw_at(const C *t, int m, int sign) {
  C w = t[m];
  return CT<C>::mk(w.x, sign > 0 ? -w.y : w.y);
}
// copy the FP64 table src[0..n) into LDS at dst (no barrier: the caller's
// next __syncthreads publishes it)
template <typename C> __device__ __forceinline__ void stage_twiddles(C *dst, const double2 *src, int n) {
  for (int m = threadIdx.x; m < n; m += blockDim.x) dst[m] = CT<C>::from(src[m]);
}

template <typename C> __host__ __device__ __forceinline__ C c_add(C a, C b) { return CT<C>::mk(a.x + b.x, a.y + b.y); }
template <typename C> __host__ __device__ __forceinline__ C c_sub(C a, C b) { return CT<C>::mk(a.x - b.x, a.y - b.y); }
template <typename C> __host__ __device__ __forceinline__ C c_mul(C a, C b) {
  return CT<C>::mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// cos(2 pi q / 48) for any integer q.  Every radix below divides 48, so
// these values cover all in-register twiddles; they are the exact values
// cos(7.5 q degrees) rounded to double.
__host__ __device__ constexpr double cos48(int q) {
  q %= 48;
  if (q < 0) q += 48;
  const bool neg = q > 12 && q < 36;
  const int r = q <= 12 ? q : (q < 24 ? 24 - q : (q <= 36 ? q - 24 : 48 - q));
  double c = 0.0;
  switch (r) {
    case 0: c = 1.0; break;
    case 1: c = 0.9914448613738104111446; break;
    case 2: c = 0.9659258262890682867497; break;
    case 3: c = 0.9238795325112867561282; break;
    case 4: c = 0.8660254037844386467637; break;
    case 5: c = 0.7933533402912351645798; break;
    case 6: c = 0.7071067811865475244008; break;
    case 7: c = 0.6087614290087206394161; break;
    case 8: c = 0.5; break;
    case 9: c = 0.3826834323650897717285; break;
    case 10: c = 0.2588190451025207623489; break;
    case 11: c = 0.1305261922200515915484; break;
    default: c = 0.0; break;
  }
  return neg ? -c : c;
}
// sin(2 pi q / 48) = cos(2 pi (q - 12) / 48)
__host__ __device__ constexpr double sin48(int q) { return cos48(q - 12); }

// compile-time loop: f(integral_constant<int, i>) for i = 0 .. N-1
template <int I, int N, typename F> __host__ __device__ __forceinline__ void sfor_(F &f) {
  if constexpr (I < N) {
    f(std::integral_constant<int, I>());
    sfor_<I + 1, N>(f);
  }
}
template <int N, typename F> __host__ __device__ __forceinline__ void sfor(F f) { sfor_<0, N>(f); }

// t / n for 0 <= t < 2^22 and 1 <= n < 2^22 (the index ranges of one
// workgroup): a float reciprocal and one correction step, ~10 VALU
// instructions against ~30 for the integer division sequence.  The float
// quotient is within (t/n) 2^-22 < 1 of t/n, so the truncated q is off by
// at most one and the remainder's sign and size correct it.
__device__ __forceinline__ int div_small(int t, int n) {
  int q = (int)((float)t * __builtin_amdgcn_rcpf((float)n));
  const int r = t - (int)__umul24((unsigned)q, (unsigned)n);
  q += (r >= n) - (r < 0);
  return q;
}

// t = hi * n + lo for t = t0, t0 + step, t0 + 2 step, ...: two divisions
// when the walk starts, none per step
struct Walk2 {
  int hi, lo, dhi, dlo, n;
  __device__ __forceinline__ Walk2(int t0, int step, int n_) : n(n_) {
    hi = div_small(t0, n);
    lo = t0 - hi * n;
    dhi = div_small(step, n);
    dlo = step - dhi * n;
  }
  __device__ __forceinline__ void next() {
    hi += dhi;
    lo += dlo;
    if (lo >= n) {
      lo -= n;
      hi++;
    }
  }
};

// t = (hi * nm + mid) * nl + lo for t = t0, t0 + step, ...
struct Walk3 {
  int hi, mid, lo, dhi, dmid, dlo, nm, nl;
  __device__ __forceinline__ Walk3(int t0, int step, int nl_, int nm_) : nm(nm_), nl(nl_) {
    int q = div_small(t0, nl);
    lo = t0 - q * nl;
    hi = div_small(q, nm);
    mid = q - hi * nm;
    q = div_small(step, nl);
    dlo = step - q * nl;
    dhi = div_small(q, nm);
    dmid = q - dhi * nm;
  }
  __device__ __forceinline__ void next() {
    lo += dlo;
    mid += dmid;
    hi += dhi;
    if (lo >= nl) {
      lo -= nl;
      mid++;
    }
    if (mid >= nm) {
      mid -= nm;
      hi++;
    }
  }
};

// a * exp(SIGN 2 pi i Q / 48), Q a compile-time constant; quarter turns are
// exact swaps
template <typename C, int SIGN, int Q> __host__ __device__ __forceinline__ C rot48(C a) {
  typedef typename CT<C>::R Rt;
  constexpr int q = ((Q % 48) + 48) % 48;
  if constexpr (q == 0) {
    return a;
  } else if constexpr (q == 24) {
    return CT<C>::mk(-a.x, -a.y);
  } else if constexpr (q == 12) {  // times SIGN i
    return SIGN > 0 ? CT<C>::mk(-a.y, a.x) : CT<C>::mk(a.y, -a.x);
  } else if constexpr (q == 36) {  // times -SIGN i
    return SIGN > 0 ? CT<C>::mk(a.y, -a.x) : CT<C>::mk(-a.y, a.x);
  } else {
    constexpr Rt c = (Rt)cos48(q), s = (Rt)(SIGN * sin48(q));
    return CT<C>::mk(a.x * c - a.y * s, a.x * s + a.y * c);
  }
}

// In-register DFT of length R (a divisor of 48), natural order in and out:
// v[k] <- sum_n v[n] exp(SIGN 2 pi i n k / R).  Lengths 2, 3, 4 directly;
// R = A B by the four-step split n = B n1 + n2, k = k1 + A k2:
//   A-point DFTs over n1, twiddle exp(SIGN 2 pi i n2 k1 / R), B-point DFTs
//   over n2.
template <typename C, int R, int SIGN> __host__ __device__ __forceinline__ void dft(C (&v)[R]) {
  typedef typename CT<C>::R Rt;
  static_assert(48 % R == 0, "in-register DFT lengths divide 48");
  if constexpr (R == 1) {
  } else if constexpr (R == 2) {
    const C a = v[0];
    v[0] = c_add(a, v[1]);
    v[1] = c_sub(a, v[1]);
  } else if constexpr (R == 3) {
    const Rt s3 = (Rt)0.86602540378443864676;  // sqrt(3)/2
    const C tt = c_add(v[1], v[2]), d = c_sub(v[1], v[2]);
    const C mm = CT<C>::mk(v[0].x - (Rt)0.5 * tt.x, v[0].y - (Rt)0.5 * tt.y);
    // e = SIGN i s3 d
    const C e = SIGN < 0 ? CT<C>::mk(s3 * d.y, -s3 * d.x) : CT<C>::mk(-s3 * d.y, s3 * d.x);
    v[0] = c_add(v[0], tt);
    v[1] = c_add(mm, e);
    v[2] = c_sub(mm, e);
  } else if constexpr (R == 4) {
    const C t0 = c_add(v[0], v[2]), t1 = c_sub(v[0], v[2]), t2 = c_add(v[1], v[3]), d = c_sub(v[1], v[3]);
    const C t3 = SIGN < 0 ? CT<C>::mk(d.y, -d.x) : CT<C>::mk(-d.y, d.x);  // d * (SIGN i)
    v[0] = c_add(t0, t2);
    v[1] = c_add(t1, t3);
    v[2] = c_sub(t0, t2);
    v[3] = c_sub(t1, t3);
  } else {
    constexpr int A = (R % 4 == 0) ? 4 : ((R % 3 == 0) ? 3 : 2);
    constexpr int B = R / A;
    C y[R];  // y[n2 * A + k1]
    sfor<B>([&](auto n2c) __attribute__((always_inline)) {
      constexpr int n2 = decltype(n2c)::value;
      C t[A];
      sfor<A>([&](auto n1c) __attribute__((always_inline)) { t[n1c] = v[B * n1c + n2]; });
      dft<C, A, SIGN>(t);
      sfor<A>([&](auto k1c) __attribute__((always_inline)) {
        constexpr int k1 = decltype(k1c)::value;
        y[n2 * A + k1] = rot48<C, SIGN, (48 / R) * n2 * k1>(t[k1]);
      });
    });
    sfor<A>([&](auto k1c) __attribute__((always_inline)) {
      constexpr int k1 = decltype(k1c)::value;
      C t[B];
      sfor<B>([&](auto n2c) __attribute__((always_inline)) { t[n2c] = y[n2c * A + k1]; });
      dft<C, B, SIGN>(t);
      sfor<B>([&](auto k2c) __attribute__((always_inline)) { v[k1 + A * k2c] = t[k2c]; });
    });
  }
}

// One radix-R item of a Stockham stage (Ns = product of the earlier
// radices): reads x[jj + r*nbf] of one sequence, applies the twiddles
// exp(SIGN 2 pi i r k / (Ns R)) of output group k = jj % Ns, transforms in
// registers and writes y[(jj / Ns) Ns R + k + r*Ns].
template <typename C, int R, int SIGN>
__host__ __device__ __forceinline__ void stage_item(const C *xs, C *ys, int jj, int nbf, int Ns, int span,
                                                    const C *tw) {
  C v[R];
#pragma unroll
  for (int r = 0; r < R; r++) v[r] = xs[jj + r * nbf];
  const int k = jj % Ns;
  if (Ns > 1) {
#pragma unroll
    for (int r = 1; r < R; r++) v[r] = c_mul(v[r], tw_at<C>(tw, r * k * span, SIGN));
  }
  dft<C, R, SIGN>(v);
  const int idxD = (jj / Ns) * Ns * R + k;
#pragma unroll
  for (int r = 0; r < R; r++) ys[idxD + r * Ns] = v[r];
}

// The radix of the first stage for a sequence of length n; the later stages
// follow the same rule on what is left (n / radix), except where
// stage_radix says otherwise.  Two stages up to 256 points (96 = 12 x 8,
// 192 = 12 x 16), three at 384 = 8 x 8 x 6 and 768; lengths outside the
// table take radices 3, then 4, then 2.
__host__ __device__ constexpr int plan_radix(int n) {
  switch (n) {
    case 2: return 2;
    case 3: return 3;
    case 4: return 4;
    case 6: return 6;
    case 8: return 8;
    case 12: return 12;
    case 16: return 16;
    case 24: return 6;
    case 32: return 8;
    case 48: return 6;
    case 64: return 8;
    case 96: return 12;
    case 128: return 16;
    case 192: return 12;
    case 256: return 16;
    case 384: return 8;
    case 768: return 12;
    default: return n % 3 == 0 ? 3 : (n % 4 == 0 ? 4 : (n % 2 == 0 ? 2 : 0));
  }
}

// The radix of the next stage of an n-point transform with m points left.
// 384 ends in radix 6 (8 x 8 x 6, not 8 x 6 x 8): the last stage is the
// one the x-product stage holds in registers for two sequences across the
// products, and six points per sequence instead of eight keep that kernel
// at four waves per SIMD.
__host__ __device__ constexpr int stage_radix(int n, int m) { return (n == 384 && m == 48) ? 8 : plan_radix(m); }

// In-LDS batched FFT of any length 3^a 2^b.  nb sequences of length n;
// element e of sequence b is at x[b*sstr + e].  y is scratch of the same
// extent, tw the staged twiddle table.  Returns the buffer holding the
// result (x or y).  Must be called by all threads of the block.
template <typename C, int SIGN>
__device__ __forceinline__ C *fft_lds_rt(C *x, C *y, int n, int nb, int sstr, const C *tw) {
  int Ns = 1;
  for (int m = n; m > 1;) {
    const int Rr = stage_radix(n, m);
    if (Rr == 0) break;  // lengths other than 2^a 3^b are rejected on the host
    const int nbf = n / Rr;
    const int span = n / (Ns * Rr);  // twiddle exponent unit
    const int total = nb * nbf;
    for (int t = threadIdx.x; t < total; t += blockDim.x) {
      const int b = t / nbf, jj = t - b * nbf;
      switch (Rr) {
        case 16: stage_item<C, 16, SIGN>(x + b * sstr, y + b * sstr, jj, nbf, Ns, span, tw); break;
        case 12: stage_item<C, 12, SIGN>(x + b * sstr, y + b * sstr, jj, nbf, Ns, span, tw); break;
        case 8: stage_item<C, 8, SIGN>(x + b * sstr, y + b * sstr, jj, nbf, Ns, span, tw); break;
        case 6: stage_item<C, 6, SIGN>(x + b * sstr, y + b * sstr, jj, nbf, Ns, span, tw); break;
        case 4: stage_item<C, 4, SIGN>(x + b * sstr, y + b * sstr, jj, nbf, Ns, span, tw); break;
        case 3: stage_item<C, 3, SIGN>(x + b * sstr, y + b * sstr, jj, nbf, Ns, span, tw); break;
        default: stage_item<C, 2, SIGN>(x + b * sstr, y + b * sstr, jj, nbf, Ns, span, tw); break;
      }
    }
    __syncthreads();
    C *tmp = x;
    x = y;
    y = tmp;
    Ns *= Rr;
    m /= Rr;
  }
  return x;
}

// The same with the length a compile-time constant N (M points of it left
// to split): the radix plan unrolls and every index division is by a
// constant.
template <typename C, int N, int M, int Ns, int SIGN>
__device__ __forceinline__ C *fft_ct_run(C *x, C *y, int nb, int sstr, const C *tw) {
  if constexpr (M == 1) {
    return x;
  } else {
    constexpr int R = stage_radix(N, M);
    static_assert(R > 1 && M % R == 0, "no radix plan");
    constexpr int nbf = N / R;
    constexpr int span = N / (Ns * R);
    const int total = nb * nbf;
    for (int t = threadIdx.x; t < total; t += blockDim.x) {
      const int b = t / nbf, jj = t - b * nbf;
      stage_item<C, R, SIGN>(x + b * sstr, y + b * sstr, jj, nbf, Ns, span, tw);
    }
    __syncthreads();
    return fft_ct_run<C, N, M / R, Ns * R, SIGN>(y, x, nb, sstr, tw);
  }
}

// ------------------------------------------------------------------------
// Batch-fastest variant for padded layouts.  Consecutive threads take the
// same item of consecutive sequences, and sequences lie seq_stride() apart,
// a stride whose residue modulo G = 256 / sizeof(C) (the lanes whose
// accesses one LDS cycle serves) spreads those lanes over distinct banks.
// Every read and write of a stage is then conflict-free once nb >= G; the
// twiddle reads are broadcasts (one address per lane group).

// LDS stride (elements) of nb sequences of length n with elements of es
// bytes: the smallest S >= n that is an odd multiple of rho, p the largest
// power of two <= min(nb, G).  rho = G / p would put the p sequences' same
// item on p lane groups G / p banks apart, filling the G lanes' banks; but
// when nb < G a group of G lanes also holds neighbouring items of the same
// sequences, and the measured best is half that spacing, rho = G / (2 p)
// (1 once p = G): then b S mod G, b = 0 .. p-1, are distinct multiples of
// rho, every other one of the 2 p such slots.  At 256 x 256 x 257: k_xprod 0.93 -> 0.89 ms (FP64, four
// sequences) and 0.57 -> 0.48 ms (FP32, four), z stages of eight FP32
// sequences 4% faster, FP64 eight-sequence stages unchanged; rho = G / (4 p)
// for four FP64 sequences was slower (0.94 ms).
__host__ __device__ inline int seq_stride(int n, int nb, int es) {
  const int G = 256 / es;
  int p = 1;
  while (2 * p <= nb && 2 * p <= G) p *= 2;
  if (p == 1) return n;
  const int rho = G / p > 1 ? G / p / 2 : 1;
  int S = (n + rho - 1) / rho * rho;
  if ((S / rho) % 2 == 0) S += rho;
  return S;
}

template <typename C, int N, int M, int Ns, int SIGN>
__device__ __forceinline__ C *fft_bf_run(C *x, C *y, int nb, int sstr, const C *tw) {
  if constexpr (M == 1) {
    return x;
  } else {
    constexpr int R = stage_radix(N, M);
    static_assert(R > 1 && M % R == 0, "no radix plan");
    constexpr int nbf = N / R;
    constexpr int span = N / (Ns * R);
    const int total = nb * nbf;
    Walk2 w(threadIdx.x, blockDim.x, nb);  // t = jj nb + b
    for (int t = threadIdx.x; t < total; t += blockDim.x, w.next())
      stage_item<C, R, SIGN>(x + w.lo * sstr, y + w.lo * sstr, w.hi, nbf, Ns, span, tw);
    __syncthreads();
    return fft_bf_run<C, N, M / R, Ns * R, SIGN>(y, x, nb, sstr, tw);
  }
}

template <typename C, int SIGN>
__device__ __forceinline__ C *fft_bf_rt(C *x, C *y, int n, int nb, int sstr, const C *tw) {
  int Ns = 1;
  for (int m = n; m > 1;) {
    const int Rr = stage_radix(n, m);
    if (Rr == 0) break;  // lengths other than 2^a 3^b are rejected on the host
    const int nbf = n / Rr;
    const int span = n / (Ns * Rr);
    const int total = nb * nbf;
    Walk2 w(threadIdx.x, blockDim.x, nb);  // t = jj nb + b
    for (int t = threadIdx.x; t < total; t += blockDim.x, w.next()) {
      const int jj = w.hi, b = w.lo;
      switch (Rr) {
        case 16: stage_item<C, 16, SIGN>(x + b * sstr, y + b * sstr, jj, nbf, Ns, span, tw); break;
        case 12: stage_item<C, 12, SIGN>(x + b * sstr, y + b * sstr, jj, nbf, Ns, span, tw); break;
        case 8: stage_item<C, 8, SIGN>(x + b * sstr, y + b * sstr, jj, nbf, Ns, span, tw); break;
        case 6: stage_item<C, 6, SIGN>(x + b * sstr, y + b * sstr, jj, nbf, Ns, span, tw); break;
        case 4: stage_item<C, 4, SIGN>(x + b * sstr, y + b * sstr, jj, nbf, Ns, span, tw); break;
        case 3: stage_item<C, 3, SIGN>(x + b * sstr, y + b * sstr, jj, nbf, Ns, span, tw); break;
        default: stage_item<C, 2, SIGN>(x + b * sstr, y + b * sstr, jj, nbf, Ns, span, tw); break;
      }
    }
    __syncthreads();
    C *tmp = x;
    x = y;
    y = tmp;
    Ns *= Rr;
    m /= Rr;
  }
  return x;
}

template <typename C, int SIGN>
__device__ __forceinline__ C *fft_bf_s(C *x, C *y, int n, int nb, int sstr, const C *tw) {
  switch (n) {
    case 12: return fft_bf_run<C, 12, 12, 1, SIGN>(x, y, nb, sstr, tw);
    case 24: return fft_bf_run<C, 24, 24, 1, SIGN>(x, y, nb, sstr, tw);
    case 48: return fft_bf_run<C, 48, 48, 1, SIGN>(x, y, nb, sstr, tw);
    case 96: return fft_bf_run<C, 96, 96, 1, SIGN>(x, y, nb, sstr, tw);
    case 192: return fft_bf_run<C, 192, 192, 1, SIGN>(x, y, nb, sstr, tw);
    case 384: return fft_bf_run<C, 384, 384, 1, SIGN>(x, y, nb, sstr, tw);
    default: return fft_bf_rt<C, SIGN>(x, y, n, nb, sstr, tw);
  }
}

// ------------------------------------------------------------------------
// Stages of a plan one at a time, for callers that fuse work into the last
// backward stage and the first forward stage.  plan_len(n) stages; stage s
// has radix plan_at(n, s) and follows stages with product plan_ns(n, s).
// The reversed plan (plan_at(n, L-1-s)) transforms the same lengths.
__host__ __device__ constexpr int plan_len(int n) {
  int k = 0;
  for (int m = n; m > 1; k++) m /= stage_radix(n, m);
  return k;
}
__host__ __device__ constexpr int plan_at(int n, int s) {
  int m = n;
  while (s-- > 0) m /= stage_radix(n, m);
  return stage_radix(n, m);
}
__host__ __device__ constexpr int plan_ns(int n, int s) {
  int p = 1;
  for (int i = 0; i < s; i++) p *= plan_at(n, i);
  return p;
}
// the same for the reversed plan
__host__ __device__ constexpr int rplan_at(int n, int s) { return plan_at(n, plan_len(n) - 1 - s); }
__host__ __device__ constexpr int rplan_ns(int n, int s) {
  int p = 1;
  for (int i = 0; i < s; i++) p *= rplan_at(n, i);
  return p;
}

// one batch-fastest stage (radix R after stages of product Ns) from x to y
template <typename C, int N, int R, int Ns, int SIGN>
__device__ __forceinline__ void bf_stage(const C *x, C *y, int nb, int sstr, const C *tw) {
  constexpr int nbf = N / R, span = N / (Ns * R);
  const int total = nb * nbf;
  Walk2 w(threadIdx.x, blockDim.x, nb);  // t = jj nb + b
  for (int t = threadIdx.x; t < total; t += blockDim.x, w.next())
    stage_item<C, R, SIGN>(x + w.lo * sstr, y + w.lo * sstr, w.hi, nbf, Ns, span, tw);
  __syncthreads();
}
// stages S0 .. S1-1 of the plan (REV: the reversed plan), ping-ponging
// between x and y; returns the buffer with the result
template <typename C, int N, int SIGN, bool REV, int S0, int S1>
__device__ __forceinline__ C *bf_stages(C *x, C *y, int nb, int sstr, const C *tw) {
  if constexpr (S0 >= S1) {
    return x;
  } else {
    constexpr int R = REV ? rplan_at(N, S0) : plan_at(N, S0);
    constexpr int Ns = REV ? rplan_ns(N, S0) : plan_ns(N, S0);
    bf_stage<C, N, R, Ns, SIGN>(x, y, nb, sstr, tw);
    return bf_stages<C, N, SIGN, REV, S0 + 1, S1>(y, x, nb, sstr, tw);
  }
}
// In-place variants for stages with at most one item per thread: each
// thread reads and transforms its item in registers, the workgroup
// synchronises, then every item is written back into the same buffer.
// One LDS buffer instead of two doubles the workgroups that fit a CU.
template <typename C, int R, int SIGN>
__device__ __forceinline__ void item_regs(const C *xs, int jj, int nbf, int Ns, int span, const C *tw, C (&v)[R]) {
#pragma unroll
  for (int r = 0; r < R; r++) v[r] = xs[jj + r * nbf];
  const int k = jj % Ns;
  if (Ns > 1) {
#pragma unroll
    for (int r = 1; r < R; r++) v[r] = c_mul(v[r], tw_at<C>(tw, r * k * span, SIGN));
  }
  dft<C, R, SIGN>(v);
}
template <typename C, int N, int R, int Ns, int SIGN>
__device__ __forceinline__ void bf_stage_ip(C *x, int nb, int sstr, const C *tw) {
  constexpr int nbf = N / R, span = N / (Ns * R);
  const int t = threadIdx.x, jj = div_small(t, nb), b = t - jj * nb;
  const bool has = t < nb * nbf;
  C v[R];
  if (has) item_regs<C, R, SIGN>(x + b * sstr, jj, nbf, Ns, span, tw, v);
  __syncthreads();
  if (has) {
    const int idxD = (jj / Ns) * Ns * R + jj % Ns;
#pragma unroll
    for (int r = 0; r < R; r++) x[b * sstr + idxD + r * Ns] = v[r];
  }
  __syncthreads();
}
template <typename C, int N, int SIGN, bool REV, int S0, int S1>
__device__ __forceinline__ void bf_stages_ip(C *x, int nb, int sstr, const C *tw) {
  if constexpr (S0 < S1) {
    constexpr int R = REV ? rplan_at(N, S0) : plan_at(N, S0);
    constexpr int Ns = REV ? rplan_ns(N, S0) : plan_ns(N, S0);
    bf_stage_ip<C, N, R, Ns, SIGN>(x, nb, sstr, tw);
    bf_stages_ip<C, N, SIGN, REV, S0 + 1, S1>(x, nb, sstr, tw);
  }
}
// items of the widest stage of N's plan per sequence (N / smallest radix)
__host__ __device__ constexpr int plan_max_items(int n) {
  int m = n, it = 0;
  while (m > 1) {
    const int r = stage_radix(n, m);
    it = n / r > it ? n / r : it;
    m /= r;
  }
  return it;
}

// The last stage of the plan for item jj (0 <= jj < N / R) of one sequence,
// kept in registers: v[r] = the transform at point jj + r N / R.  The first
// stage of the reversed plan reads exactly these points for its item jj:
// first_item_rev writes its result to y (ys[jj R + r]).
template <typename C, int N, int SIGN>
__device__ __forceinline__ void last_item(const C *xs, int jj, const C *tw, C (&v)[plan_at(N, plan_len(N) - 1)]) {
  constexpr int R = plan_at(N, plan_len(N) - 1), Ns = N / R;
#pragma unroll
  for (int r = 0; r < R; r++) v[r] = xs[jj + r * Ns];
  if constexpr (Ns > 1) {
#pragma unroll
    for (int r = 1; r < R; r++) v[r] = c_mul(v[r], tw_at<C>(tw, r * jj, SIGN));  // span = 1
  }
  dft<C, R, SIGN>(v);
}
template <typename C, int N, int SIGN>
__device__ __forceinline__ void first_item_rev(C (&v)[plan_at(N, plan_len(N) - 1)], int jj, C *ys) {
  constexpr int R = plan_at(N, plan_len(N) - 1);
  dft<C, R, SIGN>(v);
#pragma unroll
  for (int r = 0; r < R; r++) ys[jj * R + r] = v[r];
}

// sign -1 forward, +1 backward (unscaled); sequences sstr = seq_stride(n,
// nb, sizeof(C)) apart
template <typename C>
__device__ __forceinline__ C *fft_lds_bf(C *x, C *y, int n, int nb, int sstr, int sign, const C *tw) {
  return sign < 0 ? fft_bf_s<C, -1>(x, y, n, nb, sstr, tw) : fft_bf_s<C, +1>(x, y, n, nb, sstr, tw);
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
