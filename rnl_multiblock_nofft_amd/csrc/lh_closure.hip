// lh_closure.hip -- learned closure of the nonlinear term (gfx950).
//
// No counterpart in the reference.  A small network predicts, from the
// local 3x3x5 stencil of u1, u2, u3 on a coarse channel, the part of the
// nonlinear term that the coarse grid misses: the box-filtered nonlinear
// term of a fine run minus the coarse nonlinear term of the filtered fields.
//
//   k_cl_phys       three mode-major fields -> physical N1 x N2 planes (fp32)
//                   with the conventions of four_to_phys_N (FOU3D.f90:
//                   786-804): z backward CFFT, x backward real FFT, unscaled
//   k_cl_infer      the whole network per point: weights in LDS, hidden
//                   activations in registers, v_mfma_f32_16x16x32_bf16
//   k_cl_to_modes   correction planes -> modes (phys_to_four_N, FOU3D.f90:
//                   2274-2293, scaled 1/N), added to nl1, nl2, nl3
//   k_cl_restrict_fields, k_cl_tau   training samples from a fine channel
//   k_gemm + elementwise kernels + k_adam   training (forward, backward)
//
// Matrix instruction operands (16x16x32, bf16 in, fp32 out), lane l,
// g = l / 16, q = l % 16: A row q with eight k slots, B column q with the
// same eight k slots (the pairing of A and B slots is all the product
// needs), D rows 4g .. 4g+3 of column q.
#include <hip/hip_runtime.h>

#include "lh_closure.h"
#include "lh_common.h"
#include "lh_fft.h"

namespace lh {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;
typedef __attribute__((ext_vector_type(4))) unsigned u32x4;

__device__ __forceinline__ unsigned f2bf(float f) {
  // round to nearest even (finite inputs)
  const unsigned u = __float_as_uint(f);
  return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
}
__device__ __forceinline__ float bf2f(unsigned h) { return __uint_as_float(h << 16); }
__device__ __forceinline__ f32x4 mfma(const u32x4 &a, const u32x4 &b, const f32x4 &c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0,
                                                 0, 0);
}

typedef __attribute__((ext_vector_type(2))) __bf16 bf16x2;
typedef __attribute__((ext_vector_type(2))) float f32x2;
// two floats to packed bf16 (v_cvt_pk_bf16_f32, round to nearest even)
__device__ __forceinline__ unsigned pack_bf16(float a, float b) {
  const f32x2 v = {a, b};
  return __builtin_bit_cast(unsigned, __builtin_convertvector(v, bf16x2));
}

// GELU, tanh form: x * sigmoid(2u), u = sqrt(2/pi) (x + 0.044715 x^3), with
// the hardware exp2 and reciprocal (the result is rounded to bf16)
// exp(-2u) = exp2(x (C1 + C2 x^2)), C1 = -2 sqrt(2/pi) log2(e), C2 = 0.044715 C1
__device__ __forceinline__ float gelu_f(float x) {
  const float C1 = -2.3022082f, C2 = -0.10294324f;
  return x * __builtin_amdgcn_rcpf(1.f + __builtin_amdgcn_exp2f(x * __builtin_fmaf(C2, x * x, C1)));
}
__device__ __forceinline__ float dgelu_f(float x) {
  const float x2 = x * x;
  const float u = 0.7978845608028654f * x * (1.f + 0.044715f * x2);
  const float s = __builtin_amdgcn_rcpf(1.f + __expf(-2.f * u));
  const float du = 0.7978845608028654f * (1.f + 3.f * 0.044715f * x2);
  return s + 2.f * x * s * (1.f - s) * du;
}

// input f of the point (j, z, x): field f / 45, then (dy + 2) * 9 +
// (dz + 1) * 3 + dx + 1; rows clamped to the field's range, x and z periodic
__device__ __forceinline__ float cl_feature(const float *phys, int f, int j, int z, int x, int N1, int N2, int nyu,
                                            int nyv, const ClNorm &nm) {
  const int fld = f / 45, r = f - 45 * fld;
  const int r9 = r % 9;
  const int dy = r / 9 - 2, dz = r9 / 3 - 1, dx = r9 % 3 - 1;
  const int jmax = (fld == 1 ? nyv : nyu) - 1;
  int jj = j + dy;
  jj = jj < 0 ? 0 : (jj > jmax ? jmax : jj);
  int zz = z + dz, xx = x + dx;
  zz += zz < 0 ? N2 : 0;
  zz -= zz >= N2 ? N2 : 0;
  xx += xx < 0 ? N1 : 0;
  xx -= xx >= N1 ? N1 : 0;
  const float u = phys[((size_t)fld * nyu + jj) * (size_t)(N1 * N2) + (size_t)zz * N1 + xx];
  const float mu = fld == 0 ? nm.mean[0] : (fld == 1 ? nm.mean[1] : nm.mean[2]);
  const float is = fld == 0 ? nm.inv_scale[0] : (fld == 1 ? nm.inv_scale[1] : nm.inv_scale[2]);
  return (u - mu) * is;
}

// ------------------------------------------------------------------------
// physical planes of three fields: one workgroup per (row, field)
__global__ void __launch_bounds__(256) k_cl_phys(char *arena, const Chan *chp, ClSrc src, float *phys) {
  const Chan ch = *chp;
  const int j = blockIdx.x, f = blockIdx.y;
  const int N1 = ch.N1, N2 = ch.N2, nxh = ch.nxh, plane = N1 * N2;
  float *out = phys + ((size_t)f * ch.nyu + j) * plane;
  const int nrows = f == 0 ? src.nrows[0] : (f == 1 ? src.nrows[1] : src.nrows[2]);
  if (j >= nrows) {
    for (int e = threadIdx.x; e < plane; e += blockDim.x) out[e] = 0.f;
    return;
  }
  const double2 *fsrc = f == 0 ? src.f[0] : (f == 1 ? src.f[1] : src.f[2]);
  const double2 *row = fsrc + (size_t)j * ch.ldc;
  extern __shared__ float2 smem_f2[];
  float2 *A = smem_f2, *B = A + plane, *twz = B + plane, *twx = twz + N2;
  stage_twiddles<float2>(twz, (const double2 *)(arena + ch.o_twzN), N2);
  stage_twiddles<float2>(twx, (const double2 *)(arena + ch.o_twxN), N1);
  for (int c = threadIdx.x; c < nxh * N2; c += blockDim.x) {
    const int i = c % nxh, k = c / nxh;
    const double2 v = row[c];
    A[i * N2 + k] = make_float2((float)v.x, (float)v.y);
  }
  __syncthreads();
  float2 *R = fft_lds<float2>(A, B, N2, nxh, N2, +1, twz);
  float2 *O = R == A ? B : A;
  const int h = N1 / 2;
  for (int e = threadIdx.x; e < plane; e += blockDim.x) {
    const int z = e / N1, m = e - z * N1;
    float2 c;
    if (m == 0 || m == h) {
      c = make_float2(R[m * N2 + z].x, 0.f);
    } else if (m < h) {
      c = R[m * N2 + z];
    } else {
      const float2 cc = R[(N1 - m) * N2 + z];
      c = make_float2(cc.x, -cc.y);
    }
    O[e] = c;
  }
  __syncthreads();
  float2 *X = fft_lds<float2>(O, R, N1, N2, N1, +1, twx);
  for (int e = threadIdx.x; e < plane; e += blockDim.x) out[e] = X[e].x;
}

hipError_t launch_cl_phys(char *arena, const Chan *d_chan, const Chan &h, ClSrc src, float *phys, hipStream_t s) {
  const size_t lds = (size_t)(2 * h.N1 * h.N2 + h.N1 + h.N2) * sizeof(float2);
  hipLaunchKernelGGL(k_cl_phys, dim3(h.nyu, 3), dim3(256), lds, s, arena, d_chan, src, phys);
  return hipGetLastError();
}

// ------------------------------------------------------------------------
// the network on every point (j, z, x), j = 1..ny-1.  A workgroup of NW
// waves holds the weights in LDS; each wave takes 16 CL_NB points per pass,
// gathers their inputs into its own LDS rows (lanes over consecutive points,
// so each input is one coalesced load), then runs the four layers with the
// hidden activations in registers.  Eight waves of 16 points (two per SIMD)
// measured fastest: the network is bound by the VALU work of the GELUs.
constexpr int CL_XP = 168;  // LDS pitch of one point's inputs (bf16; 336 bytes)

// accumulators start from the bias of their rows (16b + 4g + r)
template <int NOB, int CL_NB>
__device__ __forceinline__ void cl_bias(f32x4 (&acc)[NOB][CL_NB], const float *bias, int g) {
#pragma unroll
  for (int b = 0; b < NOB; b++) {
    const f32x4 v = *(const f32x4 *)(bias + 16 * b + 4 * g);
#pragma unroll
    for (int n = 0; n < CL_NB; n++) acc[b][n] = v;
  }
}

// GELU of a hidden layer's accumulators (bias included), repacked as the
// next layer's B operand (k slots 0-3 from block 2s, 4-7 from block 2s+1)
template <int CL_NB>
__device__ __forceinline__ void cl_act(const f32x4 (&acc)[8][CL_NB], u32x4 (&hb)[CL_NB][4]) {
#pragma unroll
  for (int n = 0; n < CL_NB; n++)
#pragma unroll
    for (int s = 0; s < 4; s++) {
      unsigned w[4];
#pragma unroll
      for (int half = 0; half < 2; half++) {
        const f32x4 a = acc[2 * s + half][n];
        w[2 * half] = pack_bf16(gelu_f(a[0]), gelu_f(a[1]));
        w[2 * half + 1] = pack_bf16(gelu_f(a[2]), gelu_f(a[3]));
      }
      hb[n][s] = u32x4{w[0], w[1], w[2], w[3]};
    }
}

template <int CL_NB>
__device__ __forceinline__ void cl_hidden(const u32x4 *W, int f0, const float *bias, const u32x4 (&hin)[CL_NB][4],
                                          f32x4 (&acc)[8][CL_NB], int lane, int g) {
  cl_bias<8, CL_NB>(acc, bias, g);
#pragma unroll
  for (int s = 0; s < 4; s++) {
#pragma unroll
    for (int b = 0; b < 8; b++) {
      const u32x4 a = W[(f0 + b * 4 + s) * 64 + lane];
#pragma unroll
      for (int n = 0; n < CL_NB; n++) acc[b][n] = mfma(a, hin[n][s], acc[b][n]);
    }
    // at most one k-step of weight fragments in flight (register budget)
    __builtin_amdgcn_sched_barrier(0);
  }
}

template <int CL_NB, int NW>
__global__ void __launch_bounds__(64 * NW) k_cl_infer(const u32x4 *frags, const float *bias, ClNorm nm,
                                                  const float *phys, float *corr, int N1, int N2, int nyu, int nyv) {
  extern __shared__ u32x4 smem_w[];
  u32x4 *W = smem_w;
  float *Bs = (float *)(smem_w + CLF_N * 64);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, g = lane >> 4, q = lane & 15;
  uint16_t *Xs = (uint16_t *)(Bs + CL_NBIAS) + wv * 16 * CL_NB * CL_XP;
  for (int t = threadIdx.x; t < CLF_N * 64; t += blockDim.x) W[t] = frags[t];
  for (int t = threadIdx.x; t < CL_NBIAS; t += blockDim.x) Bs[t] = bias[t];
  const int plane = N1 * N2;
  const int npts = (nyu - 2) * plane;  // u rows 1..ny-1
  constexpr int CL_WPTS = 16 * CL_NB;
  __syncthreads();  // weights and biases staged
  // the waves run independently (no workgroup barrier below), so one wave's
  // activations overlap another's matrix work
  for (int base = (blockIdx.x * NW + wv) * CL_WPTS; base < npts; base += gridDim.x * NW * CL_WPTS) {
    {
      // inputs: lane l takes point l % PTS and every PARTS-th input from
      // l / PTS on; every input's decode is a compile-time constant of the
      // unrolled loop (selected by the part), all loads issue before the
      // stores
      constexpr int PTS = 16 * CL_NB, PARTS = 64 / PTS, NIT = (CL_NIN + PARTS - 1) / PARTS;
      const int pt = lane % PTS, part = lane / PTS;
      const int p = base + pt;
      const bool ok = p < npts;
      const int pp = ok ? p : npts - 1;
      const int j = 1 + pp / plane, r = pp - (j - 1) * plane, z = r / N1, x = r - z * N1;
      int ro[3][5], zo[3], xo[3];
#pragma unroll
      for (int fl = 0; fl < 3; fl++)
#pragma unroll
        for (int dy = 0; dy < 5; dy++) {
          const int jmax = (fl == 1 ? nyv : nyu) - 1;
          int jj = j + dy - 2;
          jj = jj < 0 ? 0 : (jj > jmax ? jmax : jj);
          ro[fl][dy] = (fl * nyu + jj) * plane;
        }
#pragma unroll
      for (int d = 0; d < 3; d++) {
        int zz = z + d - 1, xx = x + d - 1;
        zz += zz < 0 ? N2 : 0;
        zz -= zz >= N2 ? N2 : 0;
        xx += xx < 0 ? N1 : 0;
        xx -= xx >= N1 ? N1 : 0;
        zo[d] = zz * N1;
        xo[d] = xx;
      }
      float v[NIT];
#pragma unroll
      for (int it = 0; it < NIT; it++) {
        int a = ro[0][0] + zo[0] + xo[0];
#pragma unroll
        for (int q = 0; q < PARTS; q++) {
          const int f = PARTS * it + q;
          if (f < CL_NIN && part == q) a = ro[f / 45][(f % 45) / 9] + zo[(f % 9) / 3] + xo[f % 3];
        }
        v[it] = phys[a];
      }
      uint16_t *xr = Xs + pt * CL_XP;
#pragma unroll
      for (int it = 0; it < NIT; it++) {
        float m = 0.f, sc = 0.f;
#pragma unroll
        for (int q = 0; q < PARTS; q++) {
          const int f = PARTS * it + q;
          if (f < CL_NIN && part == q) {
            const int fl = f / 45;
            m = fl == 0 ? nm.mean[0] : (fl == 1 ? nm.mean[1] : nm.mean[2]);
            sc = fl == 0 ? nm.inv_scale[0] : (fl == 1 ? nm.inv_scale[1] : nm.inv_scale[2]);
          }
        }
        xr[PARTS * it + part] = (uint16_t)f2bf(ok ? (v[it] - m) * sc : 0.f);
      }
      for (int f = PARTS * NIT + part; f < CL_KIN; f += PARTS) xr[f] = 0;
    }
    // the wave's own LDS rows: its stores are ordered before its loads
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    f32x4 acc[8][CL_NB];
    cl_bias<8, CL_NB>(acc, Bs, g);
#pragma unroll
    for (int s = 0; s < 5; s++) {
      u32x4 xb[CL_NB];
#pragma unroll
      for (int n = 0; n < CL_NB; n++) xb[n] = *(const u32x4 *)(Xs + (16 * n + q) * CL_XP + 32 * s + 8 * g);
#pragma unroll
      for (int b = 0; b < 8; b++) {
        const u32x4 a = W[(CLF_L1 + b * 5 + s) * 64 + lane];
#pragma unroll
        for (int n = 0; n < CL_NB; n++) acc[b][n] = mfma(a, xb[n], acc[b][n]);
      }
      __builtin_amdgcn_sched_barrier(0);
    }
    u32x4 h[CL_NB][4];
    cl_act(acc, h);
    cl_hidden(W, CLF_L2, Bs + CL_NH, h, acc, lane, g);
    cl_act(acc, h);
    cl_hidden(W, CLF_L3, Bs + 2 * CL_NH, h, acc, lane, g);
    cl_act(acc, h);
    f32x4 o[1][CL_NB];
    cl_bias<1, CL_NB>(o, Bs + 3 * CL_NH, g);
#pragma unroll
    for (int s = 0; s < 4; s++) {
      const u32x4 a = W[(CLF_L4 + s) * 64 + lane];
#pragma unroll
      for (int n = 0; n < CL_NB; n++) o[0][n] = mfma(a, h[n][s], o[0][n]);
    }
    // outputs 0..2 are rows 0..2 of the last block: lanes with g = 0
    if (g == 0) {
#pragma unroll
      for (int n = 0; n < CL_NB; n++) {
        const int p = base + 16 * n + q;
        if (p >= npts) continue;
        const int j = 1 + p / plane, r = p - (j - 1) * plane;
        const size_t at = (size_t)j * plane + r;
        const size_t fs = (size_t)nyu * plane;
        corr[at] = o[0][n][0] * nm.out_scale[0];
        corr[fs + at] = o[0][n][1] * nm.out_scale[1];
        corr[2 * fs + at] = o[0][n][2] * nm.out_scale[2];
      }
    }
    // this pass's inputs are read before the next pass stores over them
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

hipError_t launch_cl_infer(const void *frags, const float *bias, ClNorm nm, const float *phys, float *corr, int N1,
                           int N2, int nyu, int nyv, hipStream_t s) {
  constexpr int NB = 1, NW = 8;
  const int npts = (nyu - 2) * N1 * N2;
  const int per_pass = NW * 16 * NB;
  int nb = (npts + per_pass - 1) / per_pass;
  if (nb > 256) nb = 256;
  const size_t lds = (size_t)CL_FRAG_BYTES + CL_NBIAS * sizeof(float) + (size_t)NW * 16 * NB * CL_XP * 2;
  hipLaunchKernelGGL((k_cl_infer<NB, NW>), dim3(nb), dim3(64 * NW), lds, s, (const u32x4 *)frags, bias, nm, phys, corr,
                     N1, N2, nyu, nyv);
  return hipGetLastError();
}

// ------------------------------------------------------------------------
// correction planes -> modes, added to the nonlinear term: one workgroup
// per (row, field); Nu1, Nu3 on u rows 1..ny-1, Nu2 on v rows 1..ny-2; the
// x-Nyquist (i = N1/2) and z-Nyquist (k = N2/2) modes are left alone
__global__ void __launch_bounds__(256) k_cl_to_modes(char *arena, const Chan *chp, const float *corr, double2 *store) {
  const Chan ch = *chp;
  const int j = blockIdx.x, f = blockIdx.y;
  if (j < 1 || j > (f == 1 ? ch.nyv : ch.nyu) - 2) return;
  const View v(arena, ch);
  const int N1 = ch.N1, N2 = ch.N2, nxh = ch.nxh, plane = N1 * N2;
  const float *in = corr + ((size_t)f * ch.nyu + j) * plane;
  extern __shared__ float2 smem_f2[];
  float2 *A = smem_f2, *B = A + plane, *twz = B + plane, *twx = twz + N2;
  stage_twiddles<float2>(twz, (const double2 *)(arena + ch.o_twzN), N2);
  stage_twiddles<float2>(twx, (const double2 *)(arena + ch.o_twxN), N1);
  for (int e = threadIdx.x; e < plane; e += blockDim.x) A[e] = make_float2(in[e], 0.f);
  __syncthreads();
  float2 *R = fft_lds<float2>(A, B, N1, N2, N1, -1, twx);
  float2 *O = R == A ? B : A;
  const float s1 = 1.f / (float)N1;
  for (int e = threadIdx.x; e < nxh * N2; e += blockDim.x) {
    const int i = e / N2, z = e - i * N2;
    const float2 c = R[z * N1 + i];
    O[e] = make_float2(c.x * s1, (i == 0 || i == N1 / 2) ? 0.f : c.y * s1);
  }
  __syncthreads();
  float2 *X = fft_lds<float2>(O, R, N2, nxh, N2, -1, twz);
  double2 *nl = f == 0 ? v.nl1 : (f == 1 ? v.nl2 : v.nl3);
  double2 *dst = nl + (size_t)j * ch.ldc;
  const double s2 = 1.0 / (double)N2;
  double2 *keep = store ? store + ((size_t)f * ch.nyu + j) * ch.ldc : nullptr;
  for (int c = threadIdx.x; c < ch.ncol; c += blockDim.x) {
    const int i = c % nxh, k = c / nxh;
    if (i == N1 / 2 || k == N2 / 2) continue;
    const float2 t = X[i * N2 + k];
    const double2 a = make_double2((double)t.x * s2, (double)t.y * s2);
    const double2 o = dst[c];
    dst[c] = make_double2(o.x + a.x, o.y + a.y);
    if (keep) keep[c] = a;
  }
}

// the stored correction modes added again (held closure, substeps 2 and 3):
// the same values and the same additions as k_cl_to_modes made
__global__ void __launch_bounds__(256) k_cl_add_modes(char *arena, const Chan *chp, const double2 *store) {
  const Chan ch = *chp;
  const int j = blockIdx.x, f = blockIdx.y;
  if (j < 1 || j > (f == 1 ? ch.nyv : ch.nyu) - 2) return;
  const View v(arena, ch);
  const int N1 = ch.N1, N2 = ch.N2, nxh = ch.nxh;
  double2 *nl = f == 0 ? v.nl1 : (f == 1 ? v.nl2 : v.nl3);
  double2 *dst = nl + (size_t)j * ch.ldc;
  const double2 *keep = store + ((size_t)f * ch.nyu + j) * ch.ldc;
  for (int c = threadIdx.x; c < ch.ncol; c += blockDim.x) {
    const int i = c % nxh, k = c / nxh;
    if (i == N1 / 2 || k == N2 / 2) continue;
    const double2 a = keep[c], o = dst[c];
    dst[c] = make_double2(o.x + a.x, o.y + a.y);
  }
}

hipError_t launch_cl_to_modes(char *arena, const Chan *d_chan, const Chan &h, const float *corr, hipStream_t s,
                              double2 *store) {
  const size_t lds = (size_t)(2 * h.N1 * h.N2 + h.N1 + h.N2) * sizeof(float2);
  hipLaunchKernelGGL(k_cl_to_modes, dim3(h.nyu, 3), dim3(256), lds, s, arena, d_chan, corr, store);
  return hipGetLastError();
}

hipError_t launch_cl_add_modes(char *arena, const Chan *d_chan, const Chan &h, const double2 *store, hipStream_t s) {
  hipLaunchKernelGGL(k_cl_add_modes, dim3(h.nyu, 3), dim3(256), 0, s, arena, d_chan, store);
  return hipGetLastError();
}

// ------------------------------------------------------------------------
// training samples: box filter of the coarse cell (Lx/N1c, Lz/N2c) in
// spectral form, G = sinc(pi i / N1c) sinc(pi m / N2c), restricted to the
// coarse modes; the coarse z-Nyquist mode is set to zero
__device__ __forceinline__ double cl_sinc(double a) { return a == 0.0 ? 1.0 : sin(a) / a; }
__device__ __forceinline__ int cl_fine_col(const Chan &F, const Chan &C, int c, double *G) {
  const int i = c % C.nxh, k = c / C.nxh;
  if (k == C.N2 / 2) return -1;
  const int m = k < C.N2 / 2 ? k : k - C.N2;
  const double pi = 3.14159265358979323846;
  *G = cl_sinc(pi * i / C.N1) * cl_sinc(pi * m / C.N2);
  const int kf = m >= 0 ? m : m + F.N2;
  return i + kf * F.nxh;
}

__global__ void __launch_bounds__(256) k_cl_restrict_fields(char *fa, const Chan *fp, char *ca, const Chan *cp) {
  const Chan F = *fp, C = *cp;
  const int c = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, f = blockIdx.z;
  if (c >= C.ncol) return;
  const int lo = f == 3 ? 1 : 0, hi = f == 1 ? C.nyv - 1 : (f == 3 ? C.nyu - 2 : C.nyu - 1);
  if (j < lo || j > hi) return;
  const View vf(fa, F), vc(ca, C);
  const double2 *src = f == 0 ? vf.u1 : (f == 1 ? vf.u2 : (f == 2 ? vf.u3 : vf.p));
  double2 *dst = f == 0 ? vc.u1 : (f == 1 ? vc.u2 : (f == 2 ? vc.u3 : vc.p));
  double G = 0.0;
  const int cf = cl_fine_col(F, C, c, &G);
  double2 val = make_double2(0.0, 0.0);
  if (cf >= 0) {
    const double2 u = src[(size_t)j * F.ldc + cf];
    val = make_double2(G * u.x, G * u.y);
  }
  dst[(size_t)j * C.ldc + c] = val;
}

hipError_t launch_cl_restrict_fields(char *fa, const Chan *d_fine, char *ca, const Chan *d_coarse, const Chan &hc,
                                     hipStream_t s) {
  hipLaunchKernelGGL(k_cl_restrict_fields, dim3((hc.ncol + 255) / 256, hc.nyu, 4), dim3(256), 0, s, fa, d_fine, ca,
                     d_coarse);
  return hipGetLastError();
}

// the full nonlinear term of a channel at (row j, column c) as the
// right-hand side uses it: Nu1 = nl1 + d/dy(uv), Nu3 = nl3 + d/dy(wv) on u
// rows 1..ny-1, Nu2 = nl2 + d/dy(vv) on v rows 1..ny-2 (der_yu_h, der_yv_h,
// FOU3D.f90:636-676)
__device__ __forceinline__ double2 cl_full_n(const View &v, const Chan &ch, int f, int j, int c) {
  const size_t o = (size_t)j * ch.ldc + c;
  if (f == 1) {
    const double2 nl = v.nl2[o], f0 = v.fvv[o], fp = v.fvv[o + ch.ldc];
    const double dyv = v.dthdyv[j];
    return make_double2(nl.x + (fp.x - f0.x) * dyv, nl.y + (fp.y - f0.y) * dyv);
  }
  const double2 *nl = f == 0 ? v.nl1 : v.nl3;
  const double2 *fl = f == 0 ? v.fuv : v.fwv;
  const double2 n = nl[o], f0 = fl[o], fm = fl[o - ch.ldc];
  const double dyu = v.dthdyu[j];
  return make_double2(n.x + (f0.x - fm.x) * dyu, n.y + (f0.y - fm.y) * dyu);
}

__global__ void __launch_bounds__(256) k_cl_tau(char *fa, const Chan *sp, char *ca, const Chan *cp, double2 *tau,
                                                int mode) {
  const Chan S = *sp, C = *cp;
  const int c = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, f = blockIdx.z;
  if (c >= C.ncol) return;
  double2 *t = tau + ((size_t)f * C.nyu + j) * C.ldc + c;
  const bool row_ok = j >= 1 && j <= (f == 1 ? C.nyv : C.nyu) - 2;
  if (mode == 0) {
    double G = 0.0;
    const int cf = cl_fine_col(S, C, c, &G);
    double2 val = make_double2(0.0, 0.0);
    if (row_ok && cf >= 0) {
      const View vs(fa, S);
      const double2 n = cl_full_n(vs, S, f, j, cf);
      val = make_double2(G * n.x, G * n.y);
    }
    *t = val;
  } else {
    const int i = c % C.nxh, k = c / C.nxh;
    if (!row_ok || i == C.N1 / 2 || k == C.N2 / 2) {
      *t = make_double2(0.0, 0.0);
      return;
    }
    const View vc(ca, C);
    const double2 n = cl_full_n(vc, C, f, j, c), o = *t;
    *t = make_double2(o.x - n.x, o.y - n.y);
  }
}

hipError_t launch_cl_tau(char *fa, const Chan *d_src, char *ca, const Chan *d_coarse, const Chan &hc, double2 *tau,
                         int mode, hipStream_t s) {
  hipLaunchKernelGGL(k_cl_tau, dim3((hc.ncol + 255) / 256, hc.nyu, 3), dim3(256), 0, s, fa, d_src, ca, d_coarse, tau,
                     mode);
  return hipGetLastError();
}

// ------------------------------------------------------------------------
// generic bf16 GEMM: a 256-thread workgroup computes a 64 x 64 tile of C,
// four waves of 32 x 32 (2 x 2 instruction blocks); K in steps of 32 staged
// through LDS as [row][k] for both operands (k contiguous, 16-byte reads)
constexpr int GLD = 40;  // LDS row pitch in bf16 (80 bytes)

// 64 rows x 32 k of X(r, k) = X[r*sr + k*sk] into T[r*GLD + k], zero outside
__device__ __forceinline__ void gemm_stage(const uint16_t *X, long sr, long sk, int R, int r0, int k0, int ke,
                                           uint16_t *T) {
  const int t = threadIdx.x;
  if (sk == 1) {
    const int r = t >> 2, kc = (t & 3) * 8;
    const int gr = r0 + r, gk = k0 + kc;
    u32x4 v = u32x4{0u, 0u, 0u, 0u};
    if (gr < R) {
      const uint16_t *src = X + (size_t)gr * sr + gk;
      if (gk + 8 <= ke && (((uintptr_t)src) & 15) == 0) {
        v = *(const u32x4 *)src;
      } else {
        unsigned e[8];
#pragma unroll
        for (int i = 0; i < 8; i++) e[i] = gk + i < ke ? src[i] : 0u;
        v = u32x4{e[0] | (e[1] << 16), e[2] | (e[3] << 16), e[4] | (e[5] << 16), e[6] | (e[7] << 16)};
      }
    }
    *(u32x4 *)(T + r * GLD + kc) = v;
  } else {
    const int k = t >> 3, rc = (t & 7) * 8;
    const int gk = k0 + k, gr = r0 + rc;
    unsigned e[8];
#pragma unroll
    for (int i = 0; i < 8; i++) e[i] = 0u;
    if (gk < ke) {
      const uint16_t *src = X + (size_t)gk * sk + gr;
      if (gr + 8 <= R && (((uintptr_t)src) & 15) == 0) {
        const u32x4 v = *(const u32x4 *)src;
#pragma unroll
        for (int i = 0; i < 4; i++) {
          e[2 * i] = v[i] & 0xffffu;
          e[2 * i + 1] = v[i] >> 16;
        }
      } else {
#pragma unroll
        for (int i = 0; i < 8; i++) e[i] = gr + i < R ? src[i] : 0u;
      }
    }
#pragma unroll
    for (int i = 0; i < 8; i++) T[(rc + i) * GLD + k] = (uint16_t)e[i];
  }
}

__global__ void __launch_bounds__(256) k_gemm(GemmP p) {
  __shared__ __attribute__((aligned(16))) uint16_t As[64 * GLD];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[64 * GLD];
  const int m0 = blockIdx.x * 64, n0 = blockIdx.y * 64;
  const int kb = blockIdx.z * p.kchunk;
  const int ke = min(p.K, kb + p.kchunk);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, g = lane >> 4, q = lane & 15;
  const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
  f32x4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = kb; k0 < ke; k0 += 32) {
    gemm_stage(p.A, p.sam, p.sak, p.M, m0, k0, ke, As);
    gemm_stage(p.B, p.sbn, p.sbk, p.N, n0, k0, ke, Bs);
    __syncthreads();
    u32x4 a[2], b[2];
#pragma unroll
    for (int i = 0; i < 2; i++) {
      a[i] = *(const u32x4 *)(As + (wm + 16 * i + q) * GLD + 8 * g);
      b[i] = *(const u32x4 *)(Bs + (wn + 16 * i + q) * GLD + 8 * g);
    }
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
      for (int jn = 0; jn < 2; jn++) acc[i][jn] = mfma(a[i], b[jn], acc[i][jn]);
    __syncthreads();
  }
  const bool split = gridDim.z > 1;
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int jn = 0; jn < 2; jn++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int m = m0 + wm + 16 * i + 4 * g + r, n = n0 + wn + 16 * jn + q;
        if (m >= p.M || n >= p.N) continue;
        float *c = p.C + (size_t)m * p.scm + (size_t)n * p.scn;
        const float val = acc[i][jn][r];
        if (split)
          atomicAdd(c, val);
        else if (p.accum)
          *c += val;
        else
          *c = val;
      }
}

hipError_t launch_gemm(const GemmP &p, hipStream_t s) {
  const int kc = p.kchunk > 0 ? p.kchunk : p.K;
  const int nz = (p.K + kc - 1) / kc;
  GemmP q = p;
  q.kchunk = kc;
  hipLaunchKernelGGL(k_gemm, dim3((p.M + 63) / 64, (p.N + 63) / 64, nz), dim3(256), 0, s, q);
  return hipGetLastError();
}

// ------------------------------------------------------------------------
// elementwise parts of training
__global__ void k_bias_gelu(float *Z, const float *b, uint16_t *H, int R, int B) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)R * B) return;
  const int r = (int)(e / B);
  const float z = Z[e] + b[r];
  Z[e] = z;
  H[e] = (uint16_t)f2bf(gelu_f(z));
}
hipError_t launch_bias_gelu(float *Z, const float *b, uint16_t *H, int R, int B, hipStream_t s) {
  const size_t n = (size_t)R * B;
  hipLaunchKernelGGL(k_bias_gelu, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, Z, b, H, R, B);
  return hipGetLastError();
}

__global__ void __launch_bounds__(256) k_out_loss(const float *Y, const float *b4, const float *T, uint16_t *dY,
                                                  double *loss, int B, float scale) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  double l = 0.0;
  if (n < B) {
    for (int o = 0; o < CL_MOUT; o++) {
      float d = 0.f;
      if (o < CL_NOUT) {
        d = Y[(size_t)o * B + n] + b4[o] - T[(size_t)o * B + n];
        l += (double)d * d;
      }
      dY[(size_t)o * B + n] = (uint16_t)f2bf(2.f * d * scale);
    }
  }
  __shared__ double red[256];
  red[threadIdx.x] = l;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(loss, red[0]);
}
hipError_t launch_out_loss(const float *Y, const float *b4, const float *T, uint16_t *dY, double *loss, int B,
                           float scale, hipStream_t s) {
  hipLaunchKernelGGL(k_out_loss, dim3((B + 255) / 256), dim3(256), 0, s, Y, b4, T, dY, loss, B, scale);
  return hipGetLastError();
}

__global__ void k_dgelu(const float *dH, const float *Z, uint16_t *dZ, int R, int B) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)R * B) return;
  dZ[e] = (uint16_t)f2bf(dH[e] * dgelu_f(Z[e]));
}
hipError_t launch_dgelu(const float *dH, const float *Z, uint16_t *dZ, int R, int B, hipStream_t s) {
  const size_t n = (size_t)R * B;
  hipLaunchKernelGGL(k_dgelu, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, dH, Z, dZ, R, B);
  return hipGetLastError();
}

// one workgroup per row
__global__ void __launch_bounds__(256) k_rowsum(const uint16_t *D, float *db, int B) {
  const int r = blockIdx.x;
  float a = 0.f;
  for (int n = threadIdx.x; n < B; n += blockDim.x) a += bf2f(D[(size_t)r * B + n]);
  __shared__ float red[256];
  red[threadIdx.x] = a;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) db[r] += red[0];
}
hipError_t launch_rowsum(const uint16_t *D, float *db, int R, int B, hipStream_t s) {
  hipLaunchKernelGGL(k_rowsum, dim3(R), dim3(256), 0, s, D, db, B);
  return hipGetLastError();
}

__global__ void k_adam(float *w, const float *g, float *m, float *v, int n, float lr, float b1, float b2, float eps,
                       float wd, float bc1, float bc2) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const float gr = g[e];
  const float mm = b1 * m[e] + (1.f - b1) * gr;
  const float vv = b2 * v[e] + (1.f - b2) * gr * gr;
  m[e] = mm;
  v[e] = vv;
  const float mh = mm / bc1, vh = vv / bc2;
  w[e] = w[e] - lr * (mh / (sqrtf(vh) + eps) + wd * w[e]);
}
hipError_t launch_adam(float *w, const float *g, float *m, float *v, int n, float lr, float b1, float b2, float eps,
                       float wd, float bc1, float bc2, hipStream_t s) {
  hipLaunchKernelGGL(k_adam, dim3((n + 255) / 256), dim3(256), 0, s, w, g, m, v, n, lr, b1, b2, eps, wd, bc1, bc2);
  return hipGetLastError();
}

__global__ void k_pack_bf16(const float *P, uint16_t *Wb) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= CLB_N) return;
  int src;
  if (e < CLB_W2)
    src = CLP_W1 + e;
  else if (e < CLB_W3)
    src = CLP_W2 + (e - CLB_W2);
  else if (e < CLB_W4)
    src = CLP_W3 + (e - CLB_W3);
  else
    src = CLP_W4 + (e - CLB_W4);
  Wb[e] = (uint16_t)f2bf(P[src]);
}
hipError_t launch_pack_bf16(const float *P, uint16_t *Wb, hipStream_t s) {
  hipLaunchKernelGGL(k_pack_bf16, dim3((CLB_N + 255) / 256), dim3(256), 0, s, P, Wb);
  return hipGetLastError();
}

// ------------------------------------------------------------------------
// training batches
__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

// sample n: h = splitmix64(seed ^ (n * 0xD1B54A32D192ED03)), point h mod
// (nsnap * (ny-1) * N1 * N2) in snapshot-major, then row, z, x order
__global__ void k_cl_batch(const float *phys, const float *tau, int nsnap, int N1, int N2, int nyu, int nyv, ClNorm nm,
                           uint64_t seed, uint16_t *X, float *T, int B) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= B) return;
  const int plane = N1 * N2;
  const uint64_t npts = (uint64_t)(nyu - 2) * plane;
  const uint64_t h = splitmix64(seed ^ ((uint64_t)n * 0xD1B54A32D192ED03ull)) % (npts * (uint64_t)nsnap);
  const int snap = (int)(h / npts);
  const int p = (int)(h - (uint64_t)snap * npts);
  const int j = 1 + p / plane, r = p % plane, z = r / N1, x = r % N1;
  const size_t ss = (size_t)3 * nyu * plane;
  const float *ph = phys + snap * ss;
  for (int f = 0; f < CL_KIN; f++)
    X[(size_t)f * B + n] = f < CL_NIN ? (uint16_t)f2bf(cl_feature(ph, f, j, z, x, N1, N2, nyu, nyv, nm)) : 0;
  const float *ta = tau + snap * ss + (size_t)j * plane + (size_t)z * N1 + x;
  for (int o = 0; o < CL_NOUT; o++) {
    const float os = o == 0 ? nm.out_scale[0] : (o == 1 ? nm.out_scale[1] : nm.out_scale[2]);
    T[(size_t)o * B + n] = ta[(size_t)o * nyu * plane] / os;
  }
}
hipError_t launch_cl_batch(const float *phys, const float *tau, int nsnap, int N1, int N2, int nyu, int nyv,
                           ClNorm nm, uint64_t seed, uint16_t *X, float *T, int B, hipStream_t s) {
  hipLaunchKernelGGL(k_cl_batch, dim3((B + 127) / 128), dim3(128), 0, s, phys, tau, nsnap, N1, N2, nyu, nyv, nm, seed,
                     X, T, B);
  return hipGetLastError();
}

__global__ void k_cl_features(const float *phys, int N1, int N2, int nyu, int nyv, ClNorm nm, const int *pts,
                              uint16_t *X, int B) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= B) return;
  const int plane = N1 * N2, p = pts[n];
  const int j = 1 + p / plane, r = p % plane, z = r / N1, x = r % N1;
  for (int f = 0; f < CL_KIN; f++)
    X[(size_t)f * B + n] = f < CL_NIN ? (uint16_t)f2bf(cl_feature(phys, f, j, z, x, N1, N2, nyu, nyv, nm)) : 0;
}
hipError_t launch_cl_features(const float *phys, int N1, int N2, int nyu, int nyv, ClNorm nm, const int *pts,
                              uint16_t *X, int B, hipStream_t s) {
  hipLaunchKernelGGL(k_cl_features, dim3((B + 127) / 128), dim3(128), 0, s, phys, N1, N2, nyu, nyv, nm, pts, X, B);
  return hipGetLastError();
}

}  // namespace lh
