// Per-x-mode wall pass of textured walls whose masks do not depend on x
// (k_tex_1d, and the ghost rows that k_zinv's extra workgroups complete).
// With such a mask the product T[P T^-1 (.)] acts on each x-mode i alone as
// a z transform of length N2, the mask and the transform back.
//   mode 0: the wall sources of the textured Helmholtz solve (tridLU_3D.f90
//           textured LU_build rows :1284-1358), nit fixed-point passes from
//           the no-slip values; written to the wall rows of r1 / r3.
//   mode 1: the ghost rows after the projection, u0 += T[(1 - gwbc) chi T^-1 u1]
//           (gridweighting_bc_u1/u3, start.f90:546-550, on free-shear points).
#pragma once
#include "lh_common.h"
#include "lh_fft.h"

namespace lh {

// the z lengths for which the pass is compiled into k_zinv: radix-8 plans,
// whose items need no more registers than k_zinv itself
__host__ __device__ inline bool tex_ghost_small(int n2) { return n2 == 32 || n2 == 64; }

template <int SIGN, bool SMALL>
__device__ __forceinline__ double2 *tex_fft(double2 *x, double2 *y, int n, const double2 *tw) {
  if constexpr (SMALL)
    return n == 32 ? fft_ct_run<double2, 32, 32, 1, SIGN>(x, y, 1, n, tw) : fft_ct_run<double2, 64, 64, 1, SIGN>(x, y, 1, n, tw);
  else
    return fft_lds_s<double2, SIGN>(x, y, n, 1, n, tw);
}

// LDS bytes of one pass: A, B, twiddles, U (N2 double2 each), mask and
// offsets (N2 doubles each)
__host__ __device__ inline size_t tex_1d_lds(int n2max) {
  return (size_t)4 * n2max * sizeof(double2) + (size_t)2 * n2max * sizeof(double);
}

// One (x-mode i, field f, wall w) of a channel; PQ values per thread of
// the ghost pass are read at the start (the rest at the end), SMALL selects
// the compiled z lengths.  Every thread of the workgroup calls it.
template <int PQ, bool SMALL>
__device__ __forceinline__ void tex_1d_block(char *arena, const Chan &ch, int i, int f, int w, int kRK, int nit,
                                             int mode) {
  const View v(arena, ch);
  const int N1 = ch.N1, N2 = ch.N2, nxh = ch.nxh, nc = ch.ncol, n = ch.nyu, ldc = ch.ldc;
  const int tid = threadIdx.x, nt = blockDim.x;
  // the ghost pass's input row and the rows it adds to, read first so that
  // these loads overlap the set-up below and are not a round trip after the
  // transforms (this workgroup alone writes the rows it adds to)
  const double2 *src = f == 0 ? v.u1 : v.u3;
  double2 *dst = (mode == 0 ? (f == 0 ? v.r1 : v.r3) : (f == 0 ? v.u1 : v.u3)) + (w == 0 ? 0 : (size_t)(n - 1) * ldc);
  double2 gin[PQ], pre[PQ];
#pragma unroll
  for (int q = 0; q < PQ; q++) {
    const int k = tid + q * nt;
    const bool ld = mode == 1 && k < N2;
    gin[q] = ld ? src[(size_t)(w == 0 ? 1 : n - 2) * ldc + i + nxh * k] : make_double2(0, 0);
    pre[q] = ld ? dst[i + nxh * k] : make_double2(0, 0);
  }
  extern __shared__ double2 smem_d[];
  double2 *A = smem_d, *B = A + N2, *TZ = B + N2;
  double *M = reinterpret_cast<double *>(TZ + N2);
  stage_twiddles<double2>(TZ, reinterpret_cast<const double2 *>(arena + ch.o_twzN), N2);
  const size_t tab = (size_t)n * ldc;
  const double *Wr = v.wtab + (size_t)((kRK - 1) * 2 + w) * tab;
  const int gj = w == 0 ? 0 : n - 1;
  const double g0 = Wr[(size_t)gj * ldc];
  const double gw = w == 0 ? ch.gw1 : ch.gw2;
  const double Pw = 1.0 + gw, cfs = mode == 0 ? Pw / (1.0 - g0 * Pw) : 1.0 - (f == 0 ? ch.gwbc1 : ch.gwbc3);
  {
    const int *mk = v.mask + (size_t)w * N1 * N2;
    // the mask carries the forward transform's 1/N2
    const double cfz = cfs * (1.0 / N2);
    for (int z = tid; z < N2; z += nt) M[z] = mk[(size_t)z * N1] == 2 ? cfz : 0.0;
  }
  const bool realpart = i == 0 || i == N1 / 2;
  // the no-slip values U and table offsets g - gbar of this x-mode's columns,
  // read once for every pass (each thread keeps its own k)
  double2 *Us = reinterpret_cast<double2 *>(M + N2);
  double *Ds = reinterpret_cast<double *>(Us + N2);
  if (mode == 0)
    for (int k = tid; k < N2; k += nt) {
      const int c = i + nxh * k;
      Us[k] = v.texu[(size_t)(f * 2 + w) * nc + c];
      Ds[k] = Wr[(size_t)gj * ldc + c] - g0;
    }
  double2 *S = A;
  const int passes = mode == 0 ? nit : 1;
  if (mode == 1) {  // the ghost pass transforms row 1 (or n-2) once
#pragma unroll
    for (int q = 0; q < PQ; q++)
      if (tid + q * nt < N2) A[tid + q * nt] = gin[q];
    for (int k = tid + PQ * nt; k < N2; k += nt) A[k] = src[(size_t)(w == 0 ? 1 : n - 2) * ldc + i + nxh * k];
  }
  for (int it = 0; it < passes; it++) {
    if (mode == 0)
      for (int k = tid; k < N2; k += nt) {
        const double2 U = Us[k];
        const double dg = Ds[k];
        const double2 sg = it > 0 ? S[k] : make_double2(0, 0);
        S[k] = make_double2(U.x + dg * sg.x, U.y + dg * sg.y);
      }
    __syncthreads();
    double2 *R = tex_fft<+1, SMALL>(S, S == A ? B : A, N2, TZ);
    for (int z = tid; z < N2; z += nt) {
      const double2 h = R[z];
      R[z] = make_double2(M[z] * h.x, realpart ? 0.0 : M[z] * h.y);
    }
    __syncthreads();
    S = tex_fft<-1, SMALL>(R, R == A ? B : A, N2, TZ);  // ends with a barrier
  }
#pragma unroll
  for (int q = 0; q < PQ; q++) {
    const int k = tid + q * nt;
    if (k < N2) dst[i + nxh * k] = mode == 0 ? S[k] : cadd(pre[q], S[k]);
  }
  for (int k = tid + PQ * nt; k < N2; k += nt) {
    const int c = i + nxh * k;
    dst[c] = mode == 0 ? S[k] : cadd(dst[c], S[k]);
  }
}


}  // namespace lh
