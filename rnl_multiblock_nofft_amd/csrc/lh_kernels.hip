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
#include <algorithm>
#include <type_traits>

#include "lh_common.h"
#include "lh_fft.h"
#include "lh_kernels.h"
#include "lh_tex1d.h"

namespace lh {

// ------------------------------------------------------------------------
// tile sizes (shared with the host so LDS sizes agree): a workgroup takes
// about E complex points per field, whole sequences only
__host__ __device__ inline int tile_of(int E, int n, int cap) {
  int t = E / n;
  if (t < 1) t = 1;
  return t < cap ? t : cap;
}
// the same for tiles of x-modes in the hz/gx rows: whole groups of 8 modes
// (128-byte lines of the padded rows) when the tile is at least 8 wide
__host__ __device__ inline int tile_of8(int E, int n, int cap) {
  int t = tile_of(E, n, cap);
  return t >= 8 && t < cap ? t / 8 * 8 : t;
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
// XCD-aware tile order.  Workgroups are dispatched round-robin over the 8
// XCDs, each with its own L2.  The launch pads ntile * nslot workgroups to a
// multiple of 8 and gives each XCD a contiguous run of (tile, slot) pairs,
// so neighbouring x-mode tiles of a plane, which share 128-byte lines of the
// spectral rows, meet in one L2.  Returns false for the padding.
__device__ __forceinline__ bool xcd_tile(int ntile, int nslot, int &tile, int &slot, int nblk = -1, int off = 0) {
  const int per = (nblk < 0 ? (int)gridDim.x : nblk) / 8;
  const int b = (int)blockIdx.x - off;  // blocks b with one residue mod 8 share an XCD
  const int L = (b % 8) * per + b / 8;
  if (L >= ntile * nslot) return false;
  tile = L % ntile;
  slot = L / ntile;
  return true;
}
static inline unsigned xcd_grid(int ntile, int nslot) { return (unsigned)((ntile * nslot + 7) / 8 * 8); }

// ------------------------------------------------------------------------
// K1 tail for the compiled lengths: every stage of the inverse z transform
// but the last through LDS, the last in registers per (item jj, x-mode ii),
// its points z = jj + r N / R stored straight to the hz rows (consecutive
// lanes take consecutive x-modes: coalesced)
template <typename C, int N>
__device__ __forceinline__ void zinv_tail(C *A, C *B, const C *TW, int S, int ni, int nxp, C *hz) {
  constexpr int L = plan_len(N), R = plan_at(N, L - 1), NB = N / R;
  const C *X = bf_stages<C, N, +1, false, 0, L - 1>(A, B, ni, S, TW);
  Walk2 w(threadIdx.x, blockDim.x, ni);  // t = jj ni + ii
  for (int t = threadIdx.x; t < ni * NB; t += blockDim.x, w.next()) {
    const int jj = w.hi, ii = w.lo;
    C o[R];
    last_item<C, N, +1>(X + ii * S, jj, TW, o);
#pragma unroll
    for (int r = 0; r < R; r++) hz[(size_t)(jj + r * NB) * nxp + ii] = o[r];
  }
}

// The same with every stage but the last in place in one LDS buffer
// (bf_stage_ip: at most one item per thread); the last stage only reads it.
template <typename C, int N>
__device__ __forceinline__ void zinv_tail_ip(C *A, const C *TW, int S, int ni, int nxp, C *hz) {
  constexpr int L = plan_len(N), R = plan_at(N, L - 1), NB = N / R;
  bf_stages_ip<C, N, +1, false, 0, L - 1>(A, ni, S, TW);
  Walk2 w(threadIdx.x, blockDim.x, ni);  // t = jj ni + ii
  for (int t = threadIdx.x; t < ni * NB; t += blockDim.x, w.next()) {
    const int jj = w.hi, ii = w.lo;
    C o[R];
    last_item<C, N, +1>(A + ii * S, jj, TW, o);
#pragma unroll
    for (int r = 0; r < R; r++) hz[(size_t)(jj + r * NB) * nxp + ii] = o[r];
  }
}
// K1 in place (one LDS buffer) for the compiled lengths when every stage has
// at most one item per thread: ti plan_max_items(N) <= nt threads
__host__ __device__ inline bool zinv_inplace(int N, int ti, int nt = 256) {
  switch (N) {
    case 24: case 48: case 96: case 384: return ti * plan_max_items(N) <= nt;
    default: return false;
  }
}

// ------------------------------------------------------------------------
// K1: interpolation + inverse z transform
// NL: the z length Ngal2 of every channel of the launch when the host
// specialises it (the stage then holds only that length's path, fewer
// registers), -1 for a launch whose channels mix the lengths 24, 48 and 96
// (the paths of those three only), 0 for any other mix; IP: the stages run in
// place where zinv_inplace allows (with NL the host has checked that they do);
// NT threads per workgroup (512 lets tiles of 8 x-modes, whole 128-byte
// lines of the spectral and hz rows, run in place at 384 points)
template <typename C, bool IP, int NL, int NT = 256>
__global__ void __launch_bounds__(NT) k_zinv(char *arena, const Chan *chans, int E, int ntile, int nslot, int nzinv,
                                              int nghost) {
  if ((int)blockIdx.x < nghost) {
    // workgroups before the transform grid (dispatched first, so that they
    // overlap it): the previous substep's ghost rows of textured walls with
    // x-independent masks (nothing here reads rows 0 and nyu-1, which they
    // write), one per (x-mode, wall) of u1 and u3
    const int f = blockIdx.z % 3, bx = blockIdx.x, i = bx >> 1, w = bx & 1;
    if (f == 2) return;
    const Chan ch = chans[blockIdx.z / 3];
    if (!ch.tex1d || !(ch.tex & (1 << w)) || i >= ch.nxh) return;
    tex_1d_block<2, true>(arena, ch, i, f, w, 1, 1, 1);
    return;
  }
  const Chan ch = chans[blockIdx.z / 3];
  if (ch.tri) return;  // triad-list channel: k_triad computes its products
  const View v(arena, ch);
  const int f = blockIdx.z % 3;
  int tile, slot;
  if (!xcd_tile(ntile, nslot, tile, slot, nzinv, nghost)) return;
  if (slot >= ch.nslot) return;
  // logical order u0 v0 u1 v1 ... (then the last u-plane): u-plane j and
  // v-planes j-1, j interpolate from the same spectral rows
  {
    const int nplv = ch.nslot - ch.nplu, m = min(ch.nplu, nplv);
    slot = slot < 2 * m ? ((slot & 1) ? ch.nplu + (slot >> 1) : (slot >> 1))
                        : (ch.nplu > nplv ? slot - m : ch.nplu + slot - m);
  }
  const int Ng2 = NL > 0 ? NL : ch.Ng2, nxh = ch.nxh, N2 = ch.N2, ldc = ch.ldc, nxp = ch.nxp;
  const int TI = tile_of8(E, Ng2, nxh);
  const int i0 = tile * TI;
  if (i0 >= nxh) return;
  const int ni = min(TI, nxh - i0);
  extern __shared__ double2 smem_d[];
  // ni sequences of Ng2 points, S apart (padded: conflict-free stages)
  const int S = seq_stride(Ng2, TI, sizeof(C));
  const bool inplace = IP && (NL > 0 ? true : zinv_inplace(Ng2, TI, NT));
  C *A = reinterpret_cast<C *>(smem_d);
  C *B = A + TI * S;
  C *TWs = inplace ? B : B + TI * S;
  const bool isU = slot < ch.nplu;
  const int j = isU ? ch.pu0 + slot : ch.pv0 + slot - ch.nplu;

  // FP64 in place at a specialised length: the twiddles are read from the
  // table in HBM (L1-resident, 6 KiB at 384 points) rather than staged, so
  // that the LDS of three 8-x-mode workgroups fits a CU (the host leaves
  // them out of the launch's LDS)
  constexpr bool tw_hbm = std::is_same<C, double2>::value && IP && NL > 0;
  const C *TW = tw_hbm ? reinterpret_cast<const C *>(v.twz) : TWs;
  if (!tw_hbm) stage_twiddles<C>(TWs, v.twz, Ng2);
  // only the zero-padding rows z = N2/2 .. Ngal2-N2/2-1 need clearing
  {
    const int half0 = ch.N2 / 2, npad = Ng2 - ch.N2;
    Walk2 w(threadIdx.x, blockDim.x, npad);  // t = ii npad + (z - half0)
    for (int t = threadIdx.x; t < ni * npad; t += blockDim.x, w.next()) A[w.hi * S + half0 + w.lo] = CT<C>::mk(0, 0);
  }
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
  // mixed precision: u1 and u3 enter the FP32 transforms without their plane
  // means (at[u] == 0 is the point z = 0 of the x-mode i0 + 0; with i0 == 0
  // that is the (0,0) mode)
  const bool split = sizeof(C) == 8 && f != 1 && i0 == 0;
  {
    constexpr int LU = 4;  // loads in flight per thread before the LDS stores
    const int total = N2 * ni;
    Walk2 wl(threadIdx.x, blockDim.x, ni);  // t = k ni + ii
    for (int t0 = threadIdx.x; t0 < total; t0 += LU * blockDim.x) {
      double2 va[LU], vb[LU];
      int at[LU];
#pragma unroll
      for (int u = 0; u < LU; u++, wl.next()) {
        const int k = wl.hi, ii = wl.lo;
        at[u] = -1;
        if (t0 + u * (int)blockDim.x >= total) continue;
        const int col = k * nxh + i0 + ii;
        at[u] = ii * S + (k < half ? k : k + dz);
        va[u] = s0[col];
        if (s1) vb[u] = s1[col];
      }
#pragma unroll
      for (int u = 0; u < LU; u++) {
        if (at[u] < 0) continue;
        double2 vv = va[u];
        if (s1) vv = make_double2((wa * vv.x + wb * vb[u].x) / wd, (wa * vv.y + wb * vb[u].y) / wd);
        if (split && at[u] == 0) {
          // the (0,0) mode of u1 or u3: the plane mean, kept in FP64 for the
          // z-forward stage; the FP32 stages carry the fluctuation
          v.pmean[2 * slot + (f >> 1)] = vv.x;
          vv = make_double2(0, 0);
        }
        A[at[u]] = from_d<C>(vv);
      }
    }
  }
  __syncthreads();
  C *hz = Inter<C>::hz(v) + ((size_t)slot * 3 + f) * Ng2 * nxp + i0;
#define LH_IP(n) \
  if constexpr (IP) zinv_tail_ip<C, n>(A, TW, S, ni, nxp, hz)
  if constexpr (NL == -1) {
    switch (inplace ? Ng2 : -Ng2) {
      case 24: LH_IP(24); break;
      case 48: LH_IP(48); break;
      case 96: LH_IP(96); break;
      case -24: zinv_tail<C, 24>(A, B, TW, S, ni, nxp, hz); break;
      case -48: zinv_tail<C, 48>(A, B, TW, S, ni, nxp, hz); break;
      default: zinv_tail<C, 96>(A, B, TW, S, ni, nxp, hz); break;
    }
  } else {
    switch (inplace ? Ng2 : -Ng2) {
      case 24: LH_IP(24); break;
      case 48: LH_IP(48); break;
      case 96: LH_IP(96); break;
      case 384: LH_IP(384); break;
      case -24: zinv_tail<C, 24>(A, B, TW, S, ni, nxp, hz); break;
      case -48: zinv_tail<C, 48>(A, B, TW, S, ni, nxp, hz); break;
      case -96: zinv_tail<C, 96>(A, B, TW, S, ni, nxp, hz); break;
      case -384: zinv_tail<C, 384>(A, B, TW, S, ni, nxp, hz); break;
      default: {
        C *R = fft_lds_bf<C>(A, B, Ng2, ni, S, +1, TW);
        Walk2 ws(threadIdx.x, blockDim.x, ni);  // t = z ni + ii
        for (int t = threadIdx.x; t < ni * Ng2; t += blockDim.x, ws.next())
          hz[(size_t)ws.hi * nxp + ws.lo] = R[ws.lo * S + ws.hi];
      }
    }
  }
#undef LH_IP
}

// ------------------------------------------------------------------------
// K2 core for lengths whose plan ends in a radix of at most 8: the inverse
// transform's stages but the last run through LDS; then, per (row zz, item
// jj), the last inverse stage of both sequences (u + i w, v), the products
// and the first stage of the forward transform's reversed plan run in
// registers (last_item / first_item_rev: the points one item produces are
// the points the reversed plan's first item reads); the remaining forward
// stages run through LDS.  Output sequence zz nseq + q.  Returns the buffer
// with the forward result.
// (CFL maxima: m1, m3 are the plane means of u1 and u3 that the mixed
// precision's FP32 points leave out, zero in FP64)
template <typename C, bool CFL, int N>
__device__ __forceinline__ C *xprod_fused(C *A, C *B, const C *TW, int S, int nz, bool isU, typename CT<C>::R &ma,
                                          typename CT<C>::R &mb, typename CT<C>::R m1, typename CT<C>::R m3) {
  typedef typename CT<C>::R R_t;
  constexpr int L = plan_len(N), R = plan_at(N, L - 1), NB = N / R;
  C *X = bf_stages<C, N, +1, false, 0, L - 1>(A, B, 2 * nz, S, TW);
  C *Yp = X == A ? B : A;
  const int total = nz * NB;
  Walk2 w(threadIdx.x, blockDim.x, nz);  // t = jj nz + zz
  for (int t = threadIdx.x; t < total; t += blockDim.x, w.next()) {
    const int jj = w.hi, zz = w.lo;
    C a[R], b[R];
    last_item<C, N, +1>(X + (2 * zz) * S, jj, TW, a);
    last_item<C, N, +1>(X + (2 * zz + 1) * S, jj, TW, b);
    if (isU) {
      C pp[R], qq[R];
#pragma unroll
      for (int r = 0; r < R; r++) {
        const R_t a1 = a[r].x, a3 = a[r].y, a2 = b[r].x;
        if (CFL) {
          ma = fmax(ma, fabs(a1 + m1));
          mb = fmax(mb, fabs(a3 + m3));
        }
        pp[r] = CT<C>::mk(a1 * a1, a1 * a3);  // uu + i uw
        qq[r] = CT<C>::mk(a2 * a2, a3 * a3);  // vv + i ww
      }
      first_item_rev<C, N, -1>(pp, jj, Yp + (2 * zz) * S);
      first_item_rev<C, N, -1>(qq, jj, Yp + (2 * zz + 1) * S);
    } else {
      C pp[R];
#pragma unroll
      for (int r = 0; r < R; r++) {
        const R_t a1 = a[r].x, a3 = a[r].y, a2 = b[r].x;
        if (CFL) ma = fmax(ma, fabs(a2));
        pp[r] = CT<C>::mk(a1 * a2, a3 * a2);  // uv + i wv
      }
      first_item_rev<C, N, -1>(pp, jj, Yp + zz * S);
    }
  }
  __syncthreads();
  return bf_stages<C, N, -1, true, 1, L>(Yp, X, (isU ? 2 : 1) * nz, S, TW);
}

// The same in one LDS buffer (bf_stage_ip): possible when every stage has at
// most one item per thread, i.e. 2 tz plan_max_items(N) <= nt (tz rows per
// workgroup of nt threads); the host then allocates one buffer.
__host__ __device__ inline bool xprod_inplace(int N, int tz, int nt) {
  switch (N) {
    case 24: case 48: case 96: case 384: return 2 * tz * plan_max_items(N) <= nt;
    default: return false;
  }
}
template <typename C, bool CFL, int N>
__device__ __forceinline__ C *xprod_fused_ip(C *A, const C *TW, int S, int nz, bool isU, typename CT<C>::R &ma,
                                             typename CT<C>::R &mb, typename CT<C>::R m1, typename CT<C>::R m3) {
  typedef typename CT<C>::R R_t;
  constexpr int L = plan_len(N), R = plan_at(N, L - 1), NB = N / R;
  bf_stages_ip<C, N, +1, false, 0, L - 1>(A, 2 * nz, S, TW);
  const int t = threadIdx.x, jj = div_small(t, nz), zz = t - jj * nz;
  const bool has = t < nz * NB;
  C pp[R], qq[R];
  if (has) {
    C a[R], b[R];
    last_item<C, N, +1>(A + (2 * zz) * S, jj, TW, a);
    last_item<C, N, +1>(A + (2 * zz + 1) * S, jj, TW, b);
#pragma unroll
    for (int r = 0; r < R; r++) {
      const R_t a1 = a[r].x, a3 = a[r].y, a2 = b[r].x;
      if (isU) {
        if (CFL) {
          ma = fmax(ma, fabs(a1 + m1));
          mb = fmax(mb, fabs(a3 + m3));
        }
        pp[r] = CT<C>::mk(a1 * a1, a1 * a3);  // uu + i uw
        qq[r] = CT<C>::mk(a2 * a2, a3 * a3);  // vv + i ww
      } else {
        if (CFL) ma = fmax(ma, fabs(a2));
        pp[r] = CT<C>::mk(a1 * a2, a3 * a2);  // uv + i wv
      }
    }
  }
  __syncthreads();
  if (has) {
    first_item_rev<C, N, -1>(pp, jj, A + (isU ? 2 * zz : zz) * S);
    if (isU) first_item_rev<C, N, -1>(qq, jj, A + (2 * zz + 1) * S);
  }
  __syncthreads();
  bf_stages_ip<C, N, -1, true, 1, L>(A, (isU ? 2 : 1) * nz, S, TW);
  return A;
}

// ------------------------------------------------------------------------
// K2: inverse x transform, products, forward x transform
// NL: the x length Ngal1 of every channel of the launch when the host
// specialises it (with IP the host also checks that the stages run in
// place), 0 for a launch whose channels differ
template <typename C, bool CFL, bool IP, int NL>
__global__ void __launch_bounds__(256) k_xprod(char *arena, const Chan *chans, int E) {
  typedef typename CT<C>::R R_t;
  const Chan ch = chans[blockIdx.z];
  if (ch.tri) return;
  const View v(arena, ch);
  const int slot = blockIdx.y;
  if (slot >= ch.nslot) return;
  const int Ng1 = NL ? NL : ch.Ng1, Ng2 = ch.Ng2, nxp = ch.nxp, h1 = ch.N1 / 2;
  const int TZ = tile_of(E, Ng1, Ng2);
  const int z0 = blockIdx.x * TZ;
  if (z0 >= Ng2) return;
  const int nz = min(TZ, Ng2 - z0);
  extern __shared__ double2 smem_d[];
  // 2 nz sequences of Ng1 points, S apart (padded: conflict-free stages);
  // one buffer when the stages can run in place
  const int S = seq_stride(Ng1, 2 * TZ, sizeof(C));
  const bool inplace = IP && (NL ? true : xprod_inplace(Ng1, TZ, blockDim.x));
  C *A = reinterpret_cast<C *>(smem_d);
  C *B = A + 2 * TZ * S;
  C *TW = inplace ? B : B + 2 * TZ * S;
  stage_twiddles<C>(TW, v.twx, Ng1);
  const bool isU = slot < ch.nplu;
  const C *hz = Inter<C>::hz(v) + (size_t)slot * 3 * Ng2 * nxp;

  // Hermitian extension of the three half spectra, packed two per sequence:
  // seq 2zz = f0 + i f2, seq 2zz+1 = f1.  Each retained mode m <= N1/2 is
  // read once and written at m and, conjugated, at its image Ng1 - m; the
  // modes between are zero.
  {
    const int nzero = Ng1 - 2 * h1 - 1;         // m = h1+1 .. Ng1-h1-1
    Walk2 wz(threadIdx.x, blockDim.x, nzero);   // t = seq nzero + q
    for (int t = threadIdx.x; t < 2 * nz * nzero; t += blockDim.x, wz.next())
      A[wz.hi * S + h1 + 1 + wz.lo] = CT<C>::mk(0, 0);
    // LU items per thread per batch, every load of a batch in flight before
    // the first LDS store
    constexpr int LU = 2;
    const int nm = h1 + 1, total = nz * nm;
    const size_t fs = (size_t)Ng2 * nxp;        // one field of the plane
    Walk2 wh(threadIdx.x, blockDim.x, nm);      // t = zz nm + m
    for (int t0 = threadIdx.x; t0 < total; t0 += LU * blockDim.x) {
      C F0[LU], F1[LU], F2[LU];
      int zl[LU], ml[LU];
#pragma unroll
      for (int u = 0; u < LU; u++, wh.next()) {
        zl[u] = -1;
        if (t0 + u * (int)blockDim.x >= total) continue;
        zl[u] = wh.hi;
        ml[u] = wh.lo;
        const C *r = hz + (size_t)(z0 + wh.hi) * nxp + wh.lo;
        F0[u] = r[0];
        F1[u] = r[fs];
        F2[u] = r[2 * fs];
      }
#pragma unroll
      for (int u = 0; u < LU; u++) {
        if (zl[u] < 0) continue;
        const int m = ml[u];
        C a = F0[u], b = F1[u], c = F2[u];
        if (m == 0) {  // rft ignores imag(cc(0)) on the inverse transform
          a.y = 0;
          b.y = 0;
          c.y = 0;
        }
        C *s0 = A + (2 * zl[u]) * S, *s1 = s0 + S;
        s0[m] = CT<C>::mk(a.x - c.y, a.y + c.x);
        s1[m] = b;
        if (m > 0) {
          s0[Ng1 - m] = CT<C>::mk(a.x + c.y, c.x - a.y);
          s1[Ng1 - m] = CT<C>::mk(b.x, -b.y);
        }
      }
    }
  }
  __syncthreads();
  const int nseq = isU ? 2 : 1;
  // CFL maxima of dtc_calc (FOU3D.f90:699-736), at the first substep only:
  // |u1|, |u3| on u-planes, |u2| on v-planes (times dthdyv(j) in k_dt_update)
  R_t ma = 0, mb = 0;
  const R_t m1 = sizeof(C) == 8 && isU ? (R_t)v.pmean[2 * slot] : 0,
            m3 = sizeof(C) == 8 && isU ? (R_t)v.pmean[2 * slot + 1] : 0;
  C *Y;
  // (the in-place path is its own instantiation: its registers would lower
  // the occupancy of the two-buffer path)
  switch (inplace ? Ng1 : -Ng1) {
#define LH_IP(n) Y = IP ? xprod_fused_ip<C, CFL, n>(A, TW, S, nz, isU, ma, mb, m1, m3) : nullptr
    case 24: LH_IP(24); break;
    case 48: LH_IP(48); break;
    case 96: LH_IP(96); break;
    case 384: LH_IP(384); break;
#undef LH_IP
    case -24: Y = xprod_fused<C, CFL, 24>(A, B, TW, S, nz, isU, ma, mb, m1, m3); break;
    case -48: Y = xprod_fused<C, CFL, 48>(A, B, TW, S, nz, isU, ma, mb, m1, m3); break;
    case -96: Y = xprod_fused<C, CFL, 96>(A, B, TW, S, nz, isU, ma, mb, m1, m3); break;
    case -384: Y = xprod_fused<C, CFL, 384>(A, B, TW, S, nz, isU, ma, mb, m1, m3); break;
    default: {
      C *X = fft_lds_bf<C>(A, B, Ng1, 2 * nz, S, +1, TW);
      C *Yp = X == A ? B : A;
      Walk2 wp(threadIdx.x, blockDim.x, Ng1);  // t = zz Ng1 + n
      for (int t = threadIdx.x; t < nz * Ng1; t += blockDim.x, wp.next()) {
        const int zz = wp.hi, n = wp.lo;
        const C s0 = X[(2 * zz) * S + n], s1 = X[(2 * zz + 1) * S + n];
        const R_t a1 = s0.x, a3 = s0.y, a2 = s1.x;
        if (CFL) {
          if (isU) {
            ma = fmax(ma, fabs(a1 + m1));
            mb = fmax(mb, fabs(a3 + m3));
          } else {
            ma = fmax(ma, fabs(a2));
          }
        }
        if (isU) {
          Yp[(2 * zz) * S + n] = CT<C>::mk(a1 * a1, a1 * a3);      // uu + i uw
          Yp[(2 * zz + 1) * S + n] = CT<C>::mk(a2 * a2, a3 * a3);  // vv + i ww
        } else {
          Yp[zz * S + n] = CT<C>::mk(a1 * a2, a3 * a2);  // uv + i wv
        }
      }
      __syncthreads();
      Y = fft_lds_bf<C>(Yp, X, Ng1, nseq * nz, S, -1, TW);
    }
  }
  // one (max a, max b) pair per workgroup, reduced by k_dt_update
  __shared__ double red[2][4];
  if (CFL) {
    double a = (double)ma, b = (double)mb;
    for (int off = 32; off >= 1; off >>= 1) {
      a = fmax(a, __shfl_xor(a, off));
      b = fmax(b, __shfl_xor(b, off));
    }
    if ((threadIdx.x & 63) == 0) {
      red[0][threadIdx.x >> 6] = a;
      red[1][threadIdx.x >> 6] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      const int nw = (blockDim.x + 63) >> 6;
      for (int q = 1; q < nw; q++) {
        a = fmax(a, red[0][q]);
        b = fmax(b, red[1][q]);
      }
      double *cm = v.cfl + 2 * ((size_t)slot * gridDim.x + blockIdx.x);
      cm[0] = a;
      cm[1] = b;
    }
  }
  const R_t inv = (R_t)(1.0 / Ng1);
  C *gx = Inter<C>::gx(v) + (size_t)slot * 4 * Ng2 * nxp;
  Walk2 wo(threadIdx.x, blockDim.x, h1 + 1);  // t = (zz nseq + q) (h1 + 1) + m
  for (int t = threadIdx.x; t < nz * nseq * (h1 + 1); t += blockDim.x, wo.next()) {
    const int m = wo.lo, r = wo.hi;
    const int q = nseq == 2 ? (r & 1) : 0, zz = nseq == 2 ? (r >> 1) : r;
    const C *ys = Y + r * S;
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
    gx[((size_t)(2 * q) * Ng2 + z) * nxp + m] = P;
    gx[((size_t)(2 * q + 1) * Ng2 + z) * nxp + m] = Q;
  }
}

// ------------------------------------------------------------------------
// Mixed precision: the FP32 stages transform u1 - U1 and u3 - U3, U1 and U3
// the plane means (k_zinv), so the products they form are the fluctuation
// products.  The rest of each product is linear in the fields' modes and is
// added here in FP64, per output mode (k, i) of the plane:
//   u-planes  uu: + 2 U1 u1'   uw: + U1 u3' + U3 u1'   ww: + 2 U3 u3'
//   v-planes  uv: + U1 u2      wv: + U3 u2
// (u' = the modes without (0,0); the constant U1^2, U1 U3, U3^2 sit in the
// (0,0) mode, which Nu1 and Nu3 multiply by kx = kz = 0).  The FP32 part is
// then small next to the products' leading terms, and the digits the FP32
// stages would lose on the mean flow are carried in FP64.
struct MeanLin {
  const double2 *a, *b;  // u-planes: u1, u3 at the output's offsets; v-planes: u2
  double U1, U3;
  bool on, isU;
  __device__ static MeanLin make(const View &v, const Chan &ch, int slot, bool on) {
    MeanLin m;
    m.on = on;
    m.isU = slot < ch.nplu;
    m.a = m.isU ? v.u1 : v.u2;
    m.b = v.u3;
    m.U1 = on ? v.pmean[2 * slot] : 0.0;
    m.U3 = on ? v.pmean[2 * slot + 1] : 0.0;
    return m;
  }
  // the linear term of product p (u: X0..X3, v: X0, X1) at field offset o;
  // c0: the output is the (0,0) mode
  __device__ __forceinline__ double2 term(int p, size_t o, bool c0) const {
    if (isU) {
      if (p == 2) return make_double2(0, 0);
      const double2 a1 = c0 ? make_double2(0, 0) : a[o], a3 = c0 ? make_double2(0, 0) : b[o];
      if (p == 0) return cscale(2 * U1, a1);
      if (p == 1) return cadd(cscale(U1, a3), cscale(U3, a1));
      return cscale(2 * U3, a3);
    }
    return cscale(p == 0 ? U1 : U3, a[o]);
  }
};

// ------------------------------------------------------------------------
// K3 outputs of one (kz index k, x-mode i0 + ii): u-planes Nu1 = ikx X0 +
// ikz X1, Nu3 = ikx X1 + ikz X3, fvv = X2; v-planes Nu2 = ikx X0 + ikz X1,
// fuv = X0, fwv = X1 (der_x/der_z FOU3D.f90:476-593; kz = -N2/2 zeroed)
struct ZOut {
  const double *kx, *kz;
  double2 *o0, *o1, *o2;  // u: nl1, nl3, fvv; v: nl2, fuv, fwv
  size_t row;             // j * ldc
  int nxh, i0, N2, Ng2;
  bool isU;
  MeanLin ml;
  __device__ __forceinline__ void put(int k, int ii, double2 X0, double2 X1, double2 X2, double2 X3) const {
    const int i = i0 + ii;
    const size_t o = row + (size_t)k * nxh + i;
    if (ml.on) {
      const bool c0 = k == 0 && i == 0;
      X0 = cadd(X0, ml.term(0, o, c0));
      X1 = cadd(X1, ml.term(1, o, c0));
      if (isU) X3 = cadd(X3, ml.term(3, o, c0));
    }
    const bool live = k != N2 / 2;
    const double kxi = kx[i], kzk = kz[k];
    if (isU) {
      o0[o] = live ? cadd(cimul(kxi, X0), cimul(kzk, X1)) : make_double2(0, 0);
      o1[o] = live ? cadd(cimul(kxi, X1), cimul(kzk, X3)) : make_double2(0, 0);
      o2[o] = X2;
    } else {
      o0[o] = live ? cadd(cimul(kxi, X0), cimul(kzk, X1)) : make_double2(0, 0);
      o1[o] = X0;
      o2[o] = X1;
    }
  }
};

// K3 in place (one LDS buffer) for the compiled lengths when every stage has
// at most one item per thread: 4 ti plan_max_items(N) <= 256
__host__ __device__ inline bool zfwd_inplace(int N, int ti) {
  switch (N) {
    case 24: case 48: case 96: case 192: case 384: return 4 * ti * plan_max_items(N) <= 256;
    default: return false;
  }
}

// ------------------------------------------------------------------------
// K3: forward z transform, truncation, d/dx + d/dz
// the step's time-step update of one channel (every thread of the
// workgroup calls it; red: 48 doubles of LDS)
__device__ __forceinline__ void dt_update_block(char *arena, const Chan &ch, int ntiles_max, int e_xprod, double *red) {
  View v(arena, ch);
  if (ch.tri) {  // no physical planes on the list path: dt stays fixed
    if (threadIdx.x == 0) v.scal[SC_T] = v.scal[SC_T] + v.scal[SC_DT];
    return;
  }
  double m1 = 0.0, m3 = 0.0;
  // k_xprod's tiles of this channel (stride: the launch's tile count)
  const int TZ = tile_of(e_xprod, ch.Ng1, ch.Ng2), nt = (ch.Ng2 + TZ - 1) / TZ;
  const double2 *cm = reinterpret_cast<const double2 *>(v.cfl);
  const int nu = ch.nplu * nt;
  Walk2 w(threadIdx.x, blockDim.x, nt);  // q = slot nt + tz
  // one workgroup reads every tile's pair: keep 16 loads in flight per
  // thread (the loop is bound by memory latency, not bandwidth)
#pragma unroll 16
  for (int q = threadIdx.x; q < nu; q += blockDim.x, w.next()) {
    const double2 c = cm[(size_t)w.hi * ntiles_max + w.lo];
    m1 = fmax(m1, c.x);
    m3 = fmax(m3, c.y);
  }
  for (int off = 32; off >= 1; off >>= 1) {
    m1 = fmax(m1, __shfl_xor(m1, off));
    m3 = fmax(m3, __shfl_xor(m3, off));
  }
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[wv] = m1;
    red[32 + wv] = m3;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
    for (int q = 1; q < nw; q++) {
      m1 = fmax(m1, red[q]);
      m3 = fmax(m3, red[32 + q]);
    }
    m1 = fmax(m1, red[0]);
    m3 = fmax(m3, red[32]);
    // The reference's third candidate 1/(u2mloc*dthetavi) (FOU3D.f90:726)
    // divides by dthetavi, which is declared (declaration.f90:76) but never
    // assigned (only ddthetavi is, start.f90:324): it is +Inf and never
    // binds.  Its history file confirms dtc = the u3 term.  Same here.
    const double dtc = fmin(1.0 / (ch.alp * (ch.N1 / 2) * m1), 1.0 / (ch.bet * (ch.N2 / 2) * m3));
    v.scal[SC_DTC] = dtc;
    if (ch.adaptive) v.scal[SC_DT] = fmin(2.5 * v.scal[SC_DTV], ch.CFL * dtc);
    v.scal[SC_T] = v.scal[SC_T] + v.scal[SC_DT];
  }
}

// NL as in k_zinv (with IP the host also checks that the stages run in place)
template <typename C, bool IP, int NL>
__global__ void __launch_bounds__(256) k_zfwd(char *arena, const Chan *chans, int E, int ntile, int nslot, int nzfwd,
                                              int ndt, int ntiles_xprod, int e_xprod) {
  extern __shared__ double2 smem_d[];
  if ((int)blockIdx.x < ndt) {
    // with ndt = 1 the first workgroup per channel (dispatched first, so that
    // it overlaps the transforms) makes the step's time-step update (kRK = 1)
    // from the CFL maxima of the x-product stage; nothing in this stage reads
    // dt or t
    dt_update_block(arena, chans[blockIdx.z], ntiles_xprod, e_xprod, reinterpret_cast<double *>(smem_d));
    return;
  }
  const Chan ch = chans[blockIdx.z];
  if (ch.tri) return;
  const View v(arena, ch);
  int tile, slot;
  if (!xcd_tile(ntile, nslot, tile, slot, nzfwd, ndt)) return;
  if (slot >= ch.nslot) return;
  const int Ng2 = NL ? NL : ch.Ng2, nxh = ch.nxh, N2 = ch.N2, ldc = ch.ldc, nxp = ch.nxp;
  const int TI = tile_of8(E, Ng2, nxh);
  const int i0 = tile * TI;
  if (i0 >= nxh) return;
  const int ni = min(TI, nxh - i0);
  const bool isU = slot < ch.nplu;
  const int np = isU ? 4 : 2;
  const int j = isU ? ch.pu0 + slot : ch.pv0 + slot - ch.nplu;
  // np ni sequences of Ng2 points, S apart (padded: conflict-free stages);
  // one buffer when the stages run in place
  const int S = seq_stride(Ng2, 4 * TI, sizeof(C));
  const bool inplace = IP && (NL ? true : zfwd_inplace(Ng2, TI));
  C *A = reinterpret_cast<C *>(smem_d);
  C *B = A + 4 * TI * S;
  C *TW = inplace ? B : B + 4 * TI * S;
  stage_twiddles<C>(TW, v.twz, Ng2);
  const C *gx = Inter<C>::gx(v) + (size_t)slot * 4 * Ng2 * nxp;
  {
    // loads in flight per thread before the LDS stores: batches pay in FP32;
    // in FP64 their registers cost more occupancy than they gain
    constexpr int LU = sizeof(C) == 16 ? 1 : 8;
    const int total = np * ni * Ng2;
    Walk3 wl(threadIdx.x, blockDim.x, ni, Ng2);  // t = (p Ng2 + z) ni + ii
    for (int t0 = threadIdx.x; t0 < total; t0 += LU * blockDim.x) {
      C g[LU];
      int at[LU];
#pragma unroll
      for (int u = 0; u < LU; u++, wl.next()) {
        const int p = wl.hi, z = wl.mid, ii = wl.lo;
        at[u] = -1;
        if (t0 + u * (int)blockDim.x >= total) continue;
        at[u] = (p * ni + ii) * S + z;
        g[u] = gx[((size_t)p * Ng2 + z) * nxp + i0 + ii];
      }
#pragma unroll
      for (int u = 0; u < LU; u++)
        if (at[u] >= 0) A[at[u]] = g[u];
    }
  }
  __syncthreads();
  ZOut zo{v.kx, v.kz, isU ? v.nl1 : v.nl2, isU ? v.nl3 : v.fuv, isU ? v.fvv : v.fwv, (size_t)j * ldc, nxh, i0,
          N2, Ng2, isU, MeanLin::make(v, ch, slot, sizeof(C) == 8)};
  // (a register-fused last stage, as in k_zinv, needs two sequences' points
  // live per item; at FP64 that cost occupancy more than it saved)
  C *R = A;
  bool done = false;
  if constexpr (IP) {
    if (inplace) {
      done = true;
      switch (Ng2) {
      case 24: bf_stages_ip<C, 24, -1, false, 0, plan_len(24)>(A, np * ni, S, TW); break;
      case 48: bf_stages_ip<C, 48, -1, false, 0, plan_len(48)>(A, np * ni, S, TW); break;
      case 96: bf_stages_ip<C, 96, -1, false, 0, plan_len(96)>(A, np * ni, S, TW); break;
      case 192: bf_stages_ip<C, 192, -1, false, 0, plan_len(192)>(A, np * ni, S, TW); break;
        default: bf_stages_ip<C, 384, -1, false, 0, plan_len(384)>(A, np * ni, S, TW); break;
      }
    }
  }
  if (!done) R = fft_lds_bf<C>(A, B, Ng2, np * ni, S, -1, TW);
  const double inv = 1.0 / Ng2;
  const int half = N2 / 2, dz = Ng2 - N2;
  Walk2 wo(threadIdx.x, blockDim.x, ni);  // t = k ni + ii
  for (int t = threadIdx.x; t < ni * N2; t += blockDim.x, wo.next()) {
    const int k = wo.hi, ii = wo.lo;
    const int z = k < half ? k : k + dz;
    double2 X[4];
#pragma unroll
    for (int p = 0; p < 4; p++)
      if (p < np) X[p] = cscale(inv, to_d(R[(p * ni + ii) * S + z]));
    zo.put(k, ii, X[0], X[1], np == 4 ? X[2] : X[0], np == 4 ? X[3] : X[1]);
  }
}

// ------------------------------------------------------------------------
// K3, field by field.  With one x-mode per tile (the tile k_zfwd can afford
// when it holds all four fields in LDS) every lane of a load reads 16 bytes
// of its own gx row (rows nxp points apart) and every output store is as
// scattered; those accesses, not the transforms, set k_zfwd's time.  Here a
// workgroup takes up to fs_ti_max(N) x-modes, so a z row is read, and an
// output row k written, as a run of TI consecutive points, and transforms
// the two (v-planes) or four (u-planes) fields one after another in a single
// in-place LDS buffer of TI sequences.  The outputs are sums over fields
// (Nu1 = ikx X0 + ikz X1, ...): each thread keeps the partial sums of the
// output points it owns in registers and stores an output once its last
// field is transformed.  Same arithmetic per output as k_zfwd (the sums are
// formed in the same order).  Lengths with a compiled plan only; the host
// checks zfwd_fs_fits for every channel (it includes N2 = 2 Ng2 / 3).
// (Issuing the next field's loads before the current transform was tried:
// the compiler keeps that register buffer in scratch memory, 0.62 -> 0.86 ms
// at 256 x 256 x 257.)
// widest tile for nt threads: one item per thread in every stage, at most
// six points of a field per thread
__host__ __device__ constexpr int fs_ti_max(int n, int nt) {
  if (n <= 0) return 1;
  return nt / plan_max_items(n) < 6 * nt / n ? nt / plan_max_items(n) : 6 * nt / n;
}
__host__ __device__ inline bool zfwd_fs_fits(int N, int N2, int ti, int nt) {
  switch (N) {
    case 24: case 48: case 96: case 192: case 384: return ti <= fs_ti_max(N, nt) && 3 * N2 == 2 * N;
    default: return false;
  }
}
// NP fields (4: u-plane X0..X3, 2: v-plane X0, X1) of one tile; o0..o2 as in
// ZOut (u: nl1, nl3, fvv; v: nl2, fuv, fwv), orow = j ldc + i0; NT threads
template <typename C, int NL, int NP, int NT>
__device__ __forceinline__ void zfwd_fs_run(const View &v, C *A, const C *TW, const C *gx, int ni, int S, int nxp,
                                            int nxh, int i0, double2 *o0, double2 *o1, double2 *o2, size_t orow,
                                            const MeanLin &ml) {
  constexpr int Ng2 = NL, N2 = 2 * NL / 3, TM = fs_ti_max(NL, NT);
  constexpr int LU = (TM * Ng2 + NT - 1) / NT, OU = (TM * N2 + NT - 1) / NT;  // loads, outputs per thread
  constexpr int half = N2 / 2, dz = Ng2 - N2;
  const int tid = threadIdx.x;
  // this thread's points of a field, z << 6 | ii for t = z ni + ii =
  // tid + u NT (ii < 64); points past the field's end take point 0 (same
  // value, same LDS place), so every load and store is unconditional
  int zi[LU];
#pragma unroll
  for (int u = 0; u < LU; u++) {
    const int t = tid + u * NT;
    const int z = t / ni, ii = t - z * ni;
    zi[u] = t < ni * Ng2 ? (z << 6 | ii) : 0;
  }
  const double inv = 1.0 / Ng2;
  double2 sq[OU];  // partial sums ikx X0 (u-planes then ikx X1)
  double2 lt[OU];  // mixed precision: the field mode of the next linear mean term
#pragma unroll
  for (int p = 0; p < NP; p++) {
    const C *gp = gx + (size_t)p * Ng2 * nxp;
    C g[LU];
#pragma unroll
    for (int u = 0; u < LU; u++) g[u] = gp[(zi[u] >> 6) * nxp + (zi[u] & 63)];
#pragma unroll
    for (int u = 0; u < LU; u++) A[(zi[u] & 63) * S + (zi[u] >> 6)] = g[u];
    __syncthreads();
    bf_stages_ip<C, NL, -1, false, 0, plan_len(NL)>(A, ni, S, TW);  // ends with a barrier
#pragma unroll
    for (int q = 0; q < OU; q++) {
      const int t = tid + q * NT;
      if (t < ni * N2) {
        const int k = t / ni, ii = t - k * ni;
        const size_t o = orow + (size_t)k * nxh + ii;
        double2 X = cscale(inv, to_d(A[ii * S + (k < half ? k : k + dz)]));
        if (ml.on) {
          // MeanLin::term with each field mode read once: lt holds u1 (u-planes,
          // fields 0-1) then u3 (fields 1-3), or u2 (v-planes)
          const bool c0 = NP == 4 && k == 0 && i0 + ii == 0;
          if (NP == 4) {
            if (p == 0) {
              lt[q] = c0 ? make_double2(0, 0) : ml.a[o];
              X = cadd(X, cscale(2 * ml.U1, lt[q]));
            } else if (p == 1) {
              const double2 a3 = c0 ? make_double2(0, 0) : ml.b[o];
              X = cadd(X, cadd(cscale(ml.U1, a3), cscale(ml.U3, lt[q])));
              lt[q] = a3;
            } else if (p == 3) {
              X = cadd(X, cscale(2 * ml.U3, lt[q]));
            }
          } else {
            if (p == 0) lt[q] = ml.a[o];
            X = cadd(X, cscale(p == 0 ? ml.U1 : ml.U3, lt[q]));
          }
        }
        const bool live = k != half;
        const double kxi = v.kx[i0 + ii], kzk = v.kz[k];
        if (NP == 4) {  // Nu1 = ikx X0 + ikz X1, Nu3 = ikx X1 + ikz X3, fvv = X2
          if (p == 0) sq[q] = cimul(kxi, X);
          if (p == 1) {
            o0[o] = live ? cadd(sq[q], cimul(kzk, X)) : make_double2(0, 0);
            sq[q] = cimul(kxi, X);
          }
          if (p == 2) o2[o] = X;
          if (p == 3) o1[o] = live ? cadd(sq[q], cimul(kzk, X)) : make_double2(0, 0);
        } else {  // Nu2 = ikx X0 + ikz X1, fuv = X0, fwv = X1
          if (p == 0) sq[q] = cimul(kxi, X);
          if (p == 1) o0[o] = live ? cadd(sq[q], cimul(kzk, X)) : make_double2(0, 0);
          (p == 0 ? o1 : o2)[o] = X;
        }
      }
    }
    if (p + 1 < NP) __syncthreads();  // every output read before the next field's stores
  }
}
// one (tile, slot) of a channel whose z length is NL
template <typename C, int NL, int NT>
__device__ __forceinline__ void zfwd_fs_tile(const View &v, const Chan &ch, int E, int tile, int slot) {
  const int nxh = ch.nxh, nxp = ch.nxp;
  const int TI = tile_of8(E, NL, nxh);
  const int i0 = tile * TI;
  if (i0 >= nxh) return;
  const int ni = min(TI, nxh - i0);
  const int S = seq_stride(NL, TI, sizeof(C));
  extern __shared__ double2 smem_d[];
  C *A = reinterpret_cast<C *>(smem_d);
  C *TW = A + TI * S;
  stage_twiddles<C>(TW, v.twz, NL);  // published by the first field's barrier
  const C *gx = Inter<C>::gx(v) + (size_t)slot * 4 * NL * nxp + i0;
  const MeanLin ml = MeanLin::make(v, ch, slot, sizeof(C) == 8);
  if (slot < ch.nplu)
    zfwd_fs_run<C, NL, 4, NT>(v, A, TW, gx, ni, S, nxp, nxh, i0, v.nl1, v.nl3, v.fvv,
                              (size_t)(ch.pu0 + slot) * ch.ldc + i0, ml);
  else
    zfwd_fs_run<C, NL, 2, NT>(v, A, TW, gx, ni, S, nxp, nxh, i0, v.nl2, v.fuv, v.fwv,
                              (size_t)(ch.pv0 + slot - ch.nplu) * ch.ldc + i0, ml);
}
// NL: the length every channel of the launch shares, or 0 for a launch
// mixing the lengths 24, 48 and 96 (each channel takes its own length's
// path; the registers are those of the largest of the three)
template <typename C, int NL, int NT>
__global__ void __launch_bounds__(NT) k_zfwd_fs(char *arena, const Chan *chans, int E, int ntile, int nslot,
                                                 int nzfwd, int ndt, int ntiles_xprod, int e_xprod) {
  extern __shared__ double2 smem_d[];
  if ((int)blockIdx.x < ndt) {  // the step's time-step update, as in k_zfwd
    dt_update_block(arena, chans[blockIdx.z], ntiles_xprod, e_xprod, reinterpret_cast<double *>(smem_d));
    return;
  }
  const Chan ch = chans[blockIdx.z];
  if (ch.tri) return;
  const View v(arena, ch);
  int tile, slot;
  if (!xcd_tile(ntile, nslot, tile, slot, nzfwd, ndt)) return;
  if (slot >= ch.nslot) return;
  if constexpr (NL != 0) {
    zfwd_fs_tile<C, NL, NT>(v, ch, E, tile, slot);
  } else {
    switch (ch.Ng2) {
      case 24: zfwd_fs_tile<C, 24, NT>(v, ch, E, tile, slot); break;
      case 48: zfwd_fs_tile<C, 48, NT>(v, ch, E, tile, slot); break;
      default: zfwd_fs_tile<C, 96, NT>(v, ch, E, tile, slot); break;
    }
  }
}

// ------------------------------------------------------------------------
// K4: right-hand sides.  rhs_point is the right-hand side of one point (row j,
// column c): r1, r3 on rows 0..nyu-1 and r2 on rows 0..nyv-1 (zero on the
// wall rows), and, when `put`, the new explicit terms N1o, N3o, N2o of the
// interior rows.  k_rhs stores r; the column solve of smooth-walled channels
// calls it in its staging phase instead (the values go straight to LDS).
// (rhs_u: r1 and r3 with N1o, N3o; rhs_v: r2 with N2o; rhs_point: both.
// With n1, n3 given, rhs_u returns the new N1o, N3o there instead of
// storing them, and whether the row is interior: a caller that computes
// several points defers the stores past all its loads, which the compiler
// may not move across a store to the same arena.)
__device__ __forceinline__ bool rhs_u(const View &v, const Chan &ch, int kRK, int j, int c, bool put, double2 &r1,
                                      double2 &r3, double2 *n1 = nullptr, double2 *n3 = nullptr) {
  const int nxh = ch.nxh, ldc = ch.ldc, nyu = ch.nyu;
  const int i = c % nxh, k = c / nxh;
  const double dt = v.scal[SC_DT];
  const double C1 = rk_a(kRK) / ch.Re, C2 = -rk_c(kRK), Cg = -rk_g(kRK), dR = rk_d(kRK);
  const size_t o = (size_t)j * ldc + c;
  const double k2x = v.k2x[i], k2z = v.k2z[k];
  r1 = r3 = make_double2(0, 0);
  if (j >= 1 && j <= nyu - 2) {
    const double a0 = v.dyu2i[3 * j], a1 = v.dyu2i[3 * j + 1], a2 = v.dyu2i[3 * j + 2];
    const double dd = (k2x + k2z) + a1;
    const double dyu = v.dthdyu[j];
    const double2 p0 = v.p[o];
    {
      double2 um = v.u1[o - ldc], u0 = v.u1[o], up = v.u1[o + ldc];
      double2 L = make_double2(a0 * um.x + dd * u0.x + a2 * up.x, a0 * um.y + dd * u0.y + a2 * up.y);
      double2 fm = v.fuv[o - ldc], f0 = v.fuv[o];
      double2 nl = v.nl1[o];
      double2 Nn = make_double2(nl.x + (f0.x - fm.x) * dyu, nl.y + (f0.y - fm.y) * dyu);
      double2 No = v.N1o[o];
      double2 pc = cimul(-dR * v.kx[i], p0);
      double2 du = make_double2(C1 * L.x + C2 * No.x + pc.x, C1 * L.y + C2 * No.y + pc.y);
      if (c == 0) du.x = du.x + (-dR * v.scal[SC_MPGX]);
      du = make_double2(du.x + Cg * Nn.x, du.y + Cg * Nn.y);
      r1 = make_double2(u0.x + dt * du.x, u0.y + dt * du.y);
      if (n1) *n1 = Nn;
      else if (put) v.N1o[o] = Nn;
    }
    {
      double2 um = v.u3[o - ldc], u0 = v.u3[o], up = v.u3[o + ldc];
      double2 L = make_double2(a0 * um.x + dd * u0.x + a2 * up.x, a0 * um.y + dd * u0.y + a2 * up.y);
      double2 fm = v.fwv[o - ldc], f0 = v.fwv[o];
      double2 nl = v.nl3[o];
      double2 Nn = make_double2(nl.x + (f0.x - fm.x) * dyu, nl.y + (f0.y - fm.y) * dyu);
      double2 No = v.N3o[o];
      double2 pc = cimul(-dR * v.kz[k], p0);
      double2 du = make_double2(C1 * L.x + C2 * No.x + pc.x, C1 * L.y + C2 * No.y + pc.y);
      du = make_double2(du.x + Cg * Nn.x, du.y + Cg * Nn.y);
      r3 = make_double2(u0.x + dt * du.x, u0.y + dt * du.y);
      if (n3) *n3 = Nn;
      else if (put) v.N3o[o] = Nn;
    }
    return true;
  }
  return false;
}
__device__ __forceinline__ bool rhs_v(const View &v, const Chan &ch, int kRK, int j, int c, bool put, double2 &r2,
                                      double2 *n2 = nullptr) {
  const int nxh = ch.nxh, ldc = ch.ldc, nyv = ch.nyv;
  const int i = c % nxh, k = c / nxh;
  const double dt = v.scal[SC_DT];
  const double C1 = rk_a(kRK) / ch.Re, C2 = -rk_c(kRK), Cg = -rk_g(kRK), dR = rk_d(kRK);
  const size_t o = (size_t)j * ldc + c;
  const double k2x = v.k2x[i], k2z = v.k2z[k];
  r2 = make_double2(0, 0);
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
    r2 = make_double2(u0.x + dt * du.x, u0.y + dt * du.y);
    if (n2) *n2 = Nn;
    else if (put) v.N2o[o] = Nn;
    return true;
  }
  return false;
}
__device__ __forceinline__ void rhs_point(const View &v, const Chan &ch, int kRK, int j, int c, bool put, double2 &r1,
                                          double2 &r3, double2 &r2) {
  double2 n1, n3, n2;
  const bool iu = rhs_u(v, ch, kRK, j, c, false, r1, r3, &n1, &n3);
  const bool iv = rhs_v(v, ch, kRK, j, c, false, r2, &n2);
  if (put) {  // every load above is issued before these stores
    const size_t o = (size_t)j * ch.ldc + c;
    if (iu) {
      v.N1o[o] = n1;
      v.N3o[o] = n3;
    }
    if (iv) v.N2o[o] = n2;
  }
}

// band = 1: only the rows of textured channels that the wall-source passes
// read (within the truncation depths iscal of either wall, wall rows
// included), r1 and r3 only, N*o untouched: the column solve that follows
// computes every right-hand side itself (rhs_fused with the wave form)
constexpr int RB_BAND = 16;
__global__ void __launch_bounds__(256) k_rhs(char *arena, const Chan *chans, int kRK, int gx, int gy, int band) {
  const Chan ch = chans[blockIdx.z];
  if (band && !ch.tex) return;
  const View v(arena, ch);
  // XCD-aware order: tiles (column block, row j) go round-robin to the 8
  // XCDs; hand each XCD a contiguous run so that rows j-1, j, j+1 of a column
  // block are read through one L2.  gridDim.x is gx*gy rounded up to 8.
  // Within a run the tiles go in groups of RG column blocks, rows outer:
  // tile (b, j+1) follows tile (b, j) RG tiles later, after ~RG x 52 KB of
  // other reads, so the rows it shares with (b, j) are still in the 4 MB L2
  // (row-major over all column blocks, that distance is a whole row of every
  // field, megabytes at 256 x 256, and the shared rows came from HBM again).
  constexpr int RG = 16;
  const int per = gridDim.x / 8;
  const int tile = (blockIdx.x % 8) * per + blockIdx.x / 8;
  if (tile >= gx * gy) return;
  const int cg = tile / (RG * gy), rem = tile - cg * RG * gy, w = min(RG, gx - cg * RG);
  const int j = rem / w;
  const int c = (cg * RG + rem - j * w) * blockDim.x + threadIdx.x;
  if (band) {
    // tile rows are slots: slot s < RB takes the rows s, s + RB, ... below
    // the bottom depth iscal[0]; slot RB + s the rows nyu-1-s, ... above
    // the top depth (gy = 2 RB: a few thousand workgroups at 256 x 256,
    // whatever the depths)
    if (c >= ch.ncol) return;
    const int n = ch.nyu, top = j >= RB_BAND, s0 = top ? j - RB_BAND : j, J = v.iscal[top];
    for (int q = s0; q < J && q < n; q += RB_BAND) {
      const int jj = top ? n - 1 - q : q;
      double2 r1, r3;
      rhs_u(v, ch, kRK, jj, c, false, r1, r3);
      const size_t o = (size_t)jj * ch.ldc + c;
      v.r1[o] = r1;
      v.r3[o] = r3;
    }
    return;
  }
  if (j >= ch.nyu || c >= ch.ncol) return;
  const size_t o = (size_t)j * ch.ldc + c;
  double2 r1, r3, r2;
  rhs_point(v, ch, kRK, j, c, true, r1, r3, r2);
  v.r1[o] = r1;
  v.r3[o] = r3;
  if (j < ch.nyv) v.r2[o] = r2;
}

// ------------------------------------------------------------------------
// Factor set-up (once per grid / dt): for every column the LU factors of
// LU_buildU/LU_buildV + LU_dec for the three substeps (tridLU_3D.f90:257-329,
// 437-474) and of LU_buildP (tridLU_3D.f90:333-407) eliminated from the top
// row down.  The operators depend on the grid, k^2 and bRK(kRK) dt/Re only.
__global__ void __launch_bounds__(64) k_factors(char *arena, const Chan *chans) {
  const Chan ch = chans[blockIdx.z];
  const View v(arena, ch);
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ch.ncol) return;
  const int nxh = ch.nxh, ldc = ch.ldc, nyu = ch.nyu, nyv = ch.nyv;
  const size_t plane = (size_t)(nyu + PADROWS) * ldc;
  const int i = c % nxh, k = c / nxh;
  const double k2x = v.k2x[i], k2z = v.k2z[k];
  const double dt = v.scal[SC_DT];
  for (int kRK = 1; kRK <= 3; kRK++) {
    const double beta = rk_b(kRK) * dt / ch.Re;
    double *dU = v.fac + (size_t)((kRK - 1) * 4 + 0) * plane, *cUt = dU + plane;
    double *dV = dU + 2 * plane, *cVt = dU + 3 * plane;
    double cU = ch.gw1, cV = 0.0;  // a(3,0)*a(2,0) with a(2,0) = 1
    dU[c] = 1.0;
    cUt[c] = cU;
    dV[c] = 1.0;
    cVt[c] = cV;
    for (int j = 1; j < nyu; j++) {
      double a1, a2, a3;
      if (j <= nyu - 2) {
        a1 = -beta * v.dyu2i[3 * j];
        a2 = 1.0 - beta * ((v.dyu2i[3 * j + 1] + k2x) + k2z);
        a3 = -beta * v.dyu2i[3 * j + 2];
      } else {
        a1 = ch.gw2;
        a2 = 1.0;
        a3 = 0.0;
      }
      double d = 1.0 / (a2 - a1 * cU);
      cU = a3 * d;
      dU[(size_t)j * ldc + c] = d;
      cUt[(size_t)j * ldc + c] = cU;
      if (j < nyv) {
        if (j <= nyv - 2) {
          a1 = -beta * v.dyv2i[3 * j];
          a2 = 1.0 - beta * ((v.dyv2i[3 * j + 1] + k2x) + k2z);
          a3 = -beta * v.dyv2i[3 * j + 2];
        } else {
          a1 = 0.0;
          a2 = 1.0;
          a3 = 0.0;
        }
        d = 1.0 / (a2 - a1 * cV);
        cV = a3 * d;
        dV[(size_t)j * ldc + c] = d;
        cVt[(size_t)j * ldc + c] = cV;
      }
    }
  }
  double *dP = v.fac + (size_t)12 * plane, *eP = dP + plane;
  const int ns = 1, ne = nyu - 2;
  double ep = 0.0;
  for (int q = ne; q >= ns; q--) {
    const double du_j = v.dthdyu[q];
    double a1, a2, a3;
    if (c == 0 && q == ns) {  // mean-mode pressure pin, tridLU_3D.f90:384-390
      double h = v.yu[2] - v.yu[1];
      a1 = 0.0;
      a2 = 1.0 / (h * h);
      a3 = 0.0;
    } else if (q == ns) {
      a1 = 0.0;
      a2 = ((du_j * (-v.dthdyv[q])) + k2x) + k2z;
      a3 = du_j * v.dthdyv[q];
    } else if (q == ne) {
      a1 = (-du_j) * (-v.dthdyv[q - 1]);
      a2 = (((-du_j) * v.dthdyv[q - 1]) + k2x) + k2z;
      a3 = 0.0;
    } else {
      a1 = (-du_j) * (-v.dthdyv[q - 1]);
      a2 = ((((-du_j) * v.dthdyv[q - 1]) + (du_j * (-v.dthdyv[q]))) + k2x) + k2z;
      a3 = du_j * v.dthdyv[q];
    }
    double d = q == ne ? 1.0 / a2 : 1.0 / (a2 - a3 * ep);
    ep = a1 * d;
    dP[(size_t)q * ldc + c] = d;
    eP[(size_t)q * ldc + c] = ep;
  }
}

// ------------------------------------------------------------------------
// K5: column solves.  One thread per spectral column, three sweeps in y,
// using the precomputed factors (so each row's recurrence is one or two
// FMAs):
//   sweep 1 (up):   forward substitution of the U/W and V Helmholtz systems
//   sweep 2 (down): back substitution of U/W/V, the divergence of the
//                   result, and the forward substitution of the Poisson
//                   system from the top row down
//   sweep 3 (up):   Poisson back substitution, p += psi, v_corr, the
//                   flow-rate correction of mode (0,0) and the divergence
// Field rows are read in blocks of SB, the loads of block b+1 issued before
// the stores of block b (on gfx9-class GPUs one counter tracks loads and
// stores in order, so a load issued after a store also waits for it).  The
// grid coefficients sit in LDS.
#define SB 8

struct SolveLDS {
  double *cu0, *cv0, *du, *dv, *yv;
};
__device__ SolveLDS stage_coeffs(double *sm, const View &v, int nyu, int nyv) {
  SolveLDS L;
  L.cu0 = sm;
  L.du = L.cu0 + nyu;
  L.cv0 = L.du + nyu;
  L.dv = L.cv0 + nyv;
  L.yv = L.dv + nyv;
  for (int j = threadIdx.x; j < nyu; j += blockDim.x) {
    L.cu0[j] = v.dyu2i[3 * j];
    L.du[j] = v.dthdyu[j];
  }
  for (int j = threadIdx.x; j < nyv; j += blockDim.x) {
    L.cv0[j] = v.dyv2i[3 * j];
    L.dv[j] = v.dthdyv[j];
    L.yv[j] = v.yv[j];
  }
  __syncthreads();
  return L;
}
__host__ __device__ inline size_t solve_lds_doubles(int nyu) { return 2 * (size_t)nyu + 3 * (size_t)(nyu - 1); }

// Row access through buffer resources: the array base lives in a scalar
// resource descriptor, the row offset in a scalar register and the lane's
// column in a vector register, so no vector instruction computes addresses.
typedef __attribute__((__vector_size__(4 * sizeof(int)))) int i32x4;
typedef __attribute__((__vector_size__(2 * sizeof(int)))) int i32x2;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t mkbuf(const void *p, size_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), 0, (int)(bytes < 0x7fffffffu ? bytes : 0x7fffffffu),
                                           0x00020000);
}
#define RB16(j) ((unsigned)(j) * rowb16)
#define RB8(j) ((unsigned)(j) * rowb8)
#define LDC(buf, j) __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(buf, oc, RB16(j), 0))
#define LDR(buf, j) __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(buf, orl, RB8(j), 0))
#define STC(buf, j, x) __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(i32x4, (double2)(x)), buf, os, RB16(j), 0)

__global__ void __launch_bounds__(64) k_solve(char *arena, const Chan *chans, int kRK) {
  const Chan ch = chans[blockIdx.z];
  const View v(arena, ch);
  const int nxh = ch.nxh, ldc = ch.ldc, nyu = ch.nyu, nyv = ch.nyv;
  if ((int)(blockIdx.x * blockDim.x) >= ch.ncol) return;
  const int c0 = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = c0 < ch.ncol;
  const int c = live ? c0 : ch.ncol - 1;  // idle lanes shadow a real column
  const unsigned oc = (unsigned)c * 16u, orl = (unsigned)c * 8u;
  const unsigned os = (unsigned)(live ? c : ch.ncol) * 16u;  // idle lanes store into column padding
  const int i = c % nxh, k = c / nxh;
  const double dt = v.scal[SC_DT];
  const double beta = rk_b(kRK) * dt / ch.Re;
  const double kxi = v.kx[i], kzk = v.kz[k];
  const size_t plane = (size_t)(nyu + PADROWS) * ldc;
  const unsigned rowb16 = (unsigned)ldc * 16u, rowb8 = (unsigned)ldc * 8u;
  const size_t fb16 = plane * 16, fb8 = plane * 8;
  const double *f0 = v.fac + (size_t)((kRK - 1) * 4) * plane;
  const auto dU = mkbuf(f0, fb8), cUt = mkbuf(f0 + plane, fb8), dV = mkbuf(f0 + 2 * plane, fb8),
             cVt = mkbuf(f0 + 3 * plane, fb8);
  const auto dP = mkbuf(v.fac + (size_t)12 * plane, fb8), eP = mkbuf(v.fac + (size_t)13 * plane, fb8);
  const double *__restrict__ cu0 = v.dyu2i;  // stride 3
  const double *__restrict__ cv0 = v.dyv2i;
  const double *__restrict__ du = v.dthdyu;
  const double *__restrict__ dvy = v.dthdyv;
  const double *__restrict__ yv = v.yv;
  const auto u1 = mkbuf(v.u1, fb16), u2 = mkbuf(v.u2, fb16), u3 = mkbuf(v.u3, fb16);
  const auto r1 = mkbuf(v.r1, fb16), r2 = mkbuf(v.r2, fb16), r3 = mkbuf(v.r3, fb16);
  const auto psi = mkbuf(v.psi, fb16), pp = mkbuf(v.p, fb16), dvg = mkbuf(v.div, fb16);
  const int ns = 1, ne = nyu - 2;
  const bool col0 = c0 == 0;

  // ---- sweep 1: forward substitution, U/W rows 1..nyu-1, V rows 1..nyv-1
  double2 y1 = LDC(r1, 0), y3 = LDC(r3, 0), y2 = LDC(r2, 0);  // row 0: a(2,0) = 1
  STC(u1, 0, y1);
  STC(u3, 0, y3);
  STC(u2, 0, y2);
  auto row1 = [&](int j, double2 R1, double2 R3, double2 R2, double d, double e, double a1, double b1) {
    y1 = make_double2((R1.x - a1 * y1.x) * d, (R1.y - a1 * y1.y) * d);
    y3 = make_double2((R3.x - a1 * y3.x) * d, (R3.y - a1 * y3.y) * d);
    y2 = make_double2((R2.x - b1 * y2.x) * e, (R2.y - b1 * y2.y) * e);
    STC(u1, j, y1);
    STC(u3, j, y3);
    STC(u2, j, y2);
  };
  int j = 1;
  const int jend1 = nyv - 2;  // last row where both systems are interior
  for (; j + SB - 1 <= jend1; j += SB) {
    double2 R1[SB], R3[SB], R2[SB];
    double D[SB], E[SB];
#pragma unroll
    for (int q = 0; q < SB; q++) {
      R1[q] = LDC(r1, j + q);
      R3[q] = LDC(r3, j + q);
      R2[q] = LDC(r2, j + q);
      D[q] = LDR(dU, j + q);
      E[q] = LDR(dV, j + q);
    }
#pragma unroll
    for (int q = 0; q < SB; q++)
      row1(j + q, R1[q], R3[q], R2[q], D[q], E[q], -beta * cu0[3 * (j + q)], -beta * cv0[3 * (j + q)]);
  }
  for (; j <= jend1; j++)
    row1(j, LDC(r1, j), LDC(r3, j), LDC(r2, j), LDR(dU, j), LDR(dV, j), -beta * cu0[3 * j], -beta * cv0[3 * j]);
  {
    // row nyv-1: V wall row (a1 = 0), U/W interior row nyu-2
    j = nyv - 1;
    row1(j, LDC(r1, j), LDC(r3, j), LDC(r2, j), LDR(dU, j), LDR(dV, j), -beta * cu0[3 * j], 0.0);
    // row nyu-1: U/W wall row (a1 = gw2)
    j = nyu - 1;
    const double d = LDR(dU, j), a1 = ch.gw2;
    const double2 R1 = LDC(r1, j), R3 = LDC(r3, j);
    y1 = make_double2((R1.x - a1 * y1.x) * d, (R1.y - a1 * y1.y) * d);
    y3 = make_double2((R3.x - a1 * y3.x) * d, (R3.y - a1 * y3.y) * d);
    STC(u1, j, y1);
    STC(u3, j, y3);
  }

  // ---- sweep 2: back substitution (rows nyu-2..0), divergence of pressure
  //      row j+1, Poisson forward substitution from row ne down to ns
  const double dtirk = (1.0 / dt) / (rk_a(kRK) + rk_b(kRK));
  double2 x1 = y1, x3 = y3;  // row nyu-1 (final)
  double2 x2 = y2;           // row nyv-1 = nyu-2 (final)
  double2 yp = make_double2(0, 0);
  double Qx0 = 0.0;
  {
    // row nyu-2: u1, u3 only (v row nyv-1 is final), no pressure row
    j = nyu - 2;
    const double cj = LDR(cUt, j);
    const double2 X1 = LDC(u1, j), X3 = LDC(u3, j);
    x1 = make_double2(X1.x - cj * x1.x, X1.y - cj * x1.y);
    x3 = make_double2(X3.x - cj * x3.x, X3.y - cj * x3.y);
    STC(u1, j, x1);
    STC(u3, j, x3);
  }
  // row j (0 <= j <= nyu-3) with pressure row qr = j+1; a3 is zero on row
  // ne and on the mean mode's pinned row (the caller passes dvq = 0)
  auto row2 = [&](int j, double2 X1, double2 X3, double2 X2, double cj, double cv, double d, double w,
                  double dvq, double dyv) {
    double2 n1 = make_double2(X1.x - cj * x1.x, X1.y - cj * x1.y);
    double2 n3 = make_double2(X3.x - cj * x3.x, X3.y - cj * x3.y);
    double2 n2 = make_double2(X2.x - cv * x2.x, X2.y - cv * x2.y);
    STC(u1, j, n1);
    STC(u3, j, n3);
    STC(u2, j, n2);
    double2 a = cimul(kxi, x1), b = cimul(kzk, x3);
    double2 dv = make_double2((a.x + b.x) + w * (x2.x - n2.x), (a.y + b.y) + w * (x2.y - n2.y));
    double2 rhs = make_double2(dtirk * dv.x, dtirk * dv.y);
    double a3 = w * dvq;
    if (col0) Qx0 = Qx0 + dyv * x1.x;
    yp = make_double2((rhs.x - a3 * yp.x) * d, (rhs.y - a3 * yp.y) * d);
    STC(psi, j + 1, yp);
    x1 = n1;
    x3 = n3;
    x2 = n2;
  };
  j = nyu - 3;  // pressure row ne: a3 = 0
  row2(j, LDC(u1, j), LDC(u3, j), LDC(u2, j), LDR(cUt, j), LDR(cVt, j), LDR(dP, j + 1), du[j + 1], 0.0,
       yv[j + 1] - yv[j]);
  j = nyu - 4;
  for (; j - SB + 1 >= 1; j -= SB) {
    double2 X1[SB], X3[SB], X2[SB];
    double C[SB], CV[SB], Dp[SB];
#pragma unroll
    for (int q = 0; q < SB; q++) {
      X1[q] = LDC(u1, j - q);
      X3[q] = LDC(u3, j - q);
      X2[q] = LDC(u2, j - q);
      C[q] = LDR(cUt, j - q);
      CV[q] = LDR(cVt, j - q);
      Dp[q] = LDR(dP, j - q + 1);
    }
#pragma unroll
    for (int q = 0; q < SB; q++) {
      const int jj = j - q;
      row2(jj, X1[q], X3[q], X2[q], C[q], CV[q], Dp[q], du[jj + 1], dvy[jj + 1], yv[jj + 1] - yv[jj]);
    }
  }
  for (; j >= 1; j--)
    row2(j, LDC(u1, j), LDC(u3, j), LDC(u2, j), LDR(cUt, j), LDR(cVt, j), LDR(dP, j + 1), du[j + 1], dvy[j + 1],
         yv[j + 1] - yv[j]);
  {
    // j = 0: pressure row ns = 1; the mean mode's row is pinned (rhs = 0)
    const double d = LDR(dP, 1);
    const double2 yp_before = yp;
    row2(0, LDC(u1, 0), LDC(u3, 0), LDC(u2, 0), LDR(cUt, 0), LDR(cVt, 0), d, du[1], col0 ? 0.0 : dvy[1],
         yv[1] - yv[0]);
    if (col0) {
      yp = make_double2((0.0 - 0.0 * yp_before.x) * d, (0.0 - 0.0 * yp_before.y) * d);
      STC(psi, 1, yp);
    }
  }
  // flow-rate correction of mode (0,0): flowrate_corr, littleharsh.f90:476-509
  double g = 0.0;
  if (col0 && ch.flag_ctpress == 0) {
    g = (v.scal[SC_QXT] - Qx0) / v.scal[SC_QX1 + kRK - 1];
    v.scal[SC_MPGX] = v.scal[SC_MPGX] + g;
    v.scal[SC_DGX] = g;
  }
  const double *__restrict__ u11 = v.u11 + (size_t)(kRK - 1) * nyu;

  // ---- sweep 3: Poisson back substitution (rows ns..ne), p += psi, v_corr,
  //      divergence of pressure row j-1
  const double dtrk = dt * (rk_a(kRK) + rk_b(kRK));
  const double dkx = dtrk * kxi, dkz = dtrk * kzk;
  double2 pprev, u1p, u3p, vpp = make_double2(0, 0);  // v(0) = 0
  double2 u1first, u3first;
  {
    // row ns: psi(1) = y(1)
    pprev = LDC(psi, 1);
    double2 a = cimul(dkx, pprev), b = cimul(dkz, pprev);
    double2 U1 = LDC(u1, 1), U3 = LDC(u3, 1), P = LDC(pp, 1);
    u1p = make_double2(U1.x - a.x, U1.y - a.y);
    u3p = make_double2(U3.x - b.x, U3.y - b.y);
    if (col0) u1p.x = u1p.x + g * u11[1];
    STC(pp, 1, make_double2(P.x + pprev.x, P.y + pprev.y));
    STC(u1, 1, u1p);
    STC(u3, 1, u3p);
    u1first = u1p;
    u3first = u3p;
  }
  auto row3 = [&](int j, double2 Y, double e, double2 P, double2 U1, double2 U3, double2 V) {
    double2 ps = make_double2(Y.x - e * pprev.x, Y.y - e * pprev.y);
    double2 a = cimul(dkx, ps), b = cimul(dkz, ps);
    double2 n1 = make_double2(U1.x - a.x, U1.y - a.y);
    double2 n3 = make_double2(U3.x - b.x, U3.y - b.y);
    if (col0) n1.x = n1.x + g * u11[j];
    STC(psi, j, ps);
    STC(pp, j, make_double2(P.x + ps.x, P.y + ps.y));
    STC(u1, j, n1);
    STC(u3, j, n3);
    // v row j-1 and the divergence of pressure row j-1
    const double sv = dtrk * dvy[j - 1], w = du[j - 1];
    double2 vv = make_double2(V.x - sv * (ps.x - pprev.x), V.y - sv * (ps.y - pprev.y));
    double2 da = cimul(kxi, u1p), db = cimul(kzk, u3p);
    double2 dd = make_double2((da.x + db.x) + w * (vv.x - vpp.x), (da.y + db.y) + w * (vv.y - vpp.y));
    STC(u2, j - 1, vv);
    STC(dvg, j - 1, dd);
    vpp = vv;
    pprev = ps;
    u1p = n1;
    u3p = n3;
  };
  j = 2;
  for (; j + SB - 1 <= ne; j += SB) {
    double2 Y[SB], P[SB], U1[SB], U3[SB], V[SB];
    double E[SB];
#pragma unroll
    for (int q = 0; q < SB; q++) {
      Y[q] = LDC(psi, j + q);
      E[q] = LDR(eP, j + q);
      P[q] = LDC(pp, j + q);
      U1[q] = LDC(u1, j + q);
      U3[q] = LDC(u3, j + q);
      V[q] = LDC(u2, j + q - 1);
    }
#pragma unroll
    for (int q = 0; q < SB; q++) row3(j + q, Y[q], E[q], P[q], U1[q], U3[q], V[q]);
  }
  for (; j <= ne; j++) row3(j, LDC(psi, j), LDR(eP, j), LDC(pp, j), LDC(u1, j), LDC(u3, j), LDC(u2, j - 1));
  if (live) {
    // divergence of row ne: v(ne) is the upper wall
    double2 da = cimul(kxi, u1p), db = cimul(kzk, u3p);
    double w = du[ne];
    STC(dvg, ne, make_double2((da.x + db.x) + w * (0.0 - vpp.x), (da.y + db.y) + w * (0.0 - vpp.y)));
    // ghost cells and walls (v_corr, littleharsh.f90:992-1044); a band
    // edge keeps its solved ghost value
    if (!(ch.band_edge & 1)) {
      STC(u1, 0, cscale(ch.gwbc1, u1first));
      STC(u3, 0, cscale(ch.gwbc3, u3first));
    }
    if (!(ch.band_edge & 2)) {
      STC(u1, nyu - 1, cscale(ch.gwbc1, u1p));
      STC(u3, nyu - 1, cscale(ch.gwbc3, u3p));
    }
    STC(u2, 0, make_double2(0, 0));
    STC(u2, nyv - 1, make_double2(0, 0));
  }
}
#undef RB16
#undef RB8
#undef LDC
#undef LDR
#undef STC

// ------------------------------------------------------------------------
// K5 (LDS form): column solves with every column of a workgroup staged in
// LDS.  A workgroup of one wave owns C consecutive columns.  The phases:
//   L  all lanes load the Helmholtz right-hand sides r1, r3 (u rows), r2
//      (v rows) and the U, V, Poisson pivots d of this substep into LDS;
//      no dependency between loads, so they are all in flight at once
//   H  6 lanes per column (U1, U3, V) x (re, im) run forward and back
//      substitution of the Helmholtz systems in place (tridLU_3D.f90:55-189
//      with LU_dec's factors, tridLU_3D.f90:437-474)
//   R  all lanes: divergence of the solved velocity and the Poisson right-
//      hand side div/(dt (a+b)) (littleharsh.f90:237-272,581-606)
//   P  2 lanes per column (re, im): Poisson elimination from the top row
//      down and back substitution (LUsolP, tridLU_3D.f90:192-226, with the
//      mean-mode pin, tridLU_3D.f90:384-390); the mean-mode lane also
//      accumulates the flow rate and applies flowrate_corr
//      (littleharsh.f90:476-509)
//   U  all lanes: p += psi and v_corr's projection (littleharsh.f90:956-990)
//   W  all lanes: divergence, ghost rows and wall rows (littleharsh.f90:
//      992-1044), stores of u1, u2, u3, p, psi, div
// Only the recurrences of H and P are serial; the upper diagonals c = a3 d
// and Poisson multipliers e = a1 d are rebuilt from the pivots with the same
// products the factor set-up uses, so they need no LDS.
__host__ __device__ inline size_t solve_lds_col_bytes(int nyu) {
  const int nyv = nyu - 1;
  return 16 * (size_t)(3 * nyu + nyv) + 8 * (size_t)(2 * nyu + nyv) + 16;  // + kx, kz
}
// per-row coefficients shared by the columns of a workgroup
constexpr int SOLVE_ROW_ARRAYS = 9;
__host__ __device__ size_t solve_lds_bytes(int nyu, int C) {
  return (size_t)C * solve_lds_col_bytes(nyu) + 8 * (size_t)SOLVE_ROW_ARRAYS * nyu;
}
#define SB2 8
#define LB 8

// One real first-order recurrence of a Thomas sweep, split over the L lanes
// k = 0 .. L-1 of an aligned lane group (L a power of two):
//   FORM 0: z_j = (r_j - a_j z_prev) d_j      FORM 1: z_j = r_j - (a_j d_j) z_prev
// over rows j = first + dir i, i = 0 .. count-1, in place in X (row j at
// X[j xs]; a_j at A[j], d_j at D[j ds]); zin is the value before the first
// row.  Each lane owns a contiguous chunk of rows: it composes the chunk's
// affine map z -> alpha z + beta (alpha = -a d), the group scans the maps
// (log2 L shuffle steps), and each lane reruns its chunk from its true start
// value, so every stored row comes from the same operations as the serial
// sweep and only the chunk start values carry the scan's rounding.  With
// serial set (or L = 1), lane 0 runs the whole sweep alone.  Returns the
// last row's value on every lane of the group.
template <int FORM>
__device__ __forceinline__ double rec_step(double z, double r, double a, double d) {
  return FORM == 0 ? (r - a * z) * d : r - (a * d) * z;
}
template <int FORM>
__device__ __forceinline__ double rec_sweep(double *X, int xs, const double *A, const double *D, int ds, int first, int dir,
                            int count, double zin, int k, int L, bool serial) {
  serial = serial || L == 1;
  const int cs = serial ? count : (count + L - 1) / L;
  const int b = min(count, k * cs), e = min(count, b + cs);
  double z = zin;
  if (!serial) {
    double Am = 1.0, Bm = 0.0;
    for (int i = b; i < e; i++) {
      const int j = first + dir * i;
      const double a = A[j], d = D[j * ds];
      Bm = rec_step<FORM>(Bm, X[j * xs], a, d);
      Am = FORM == 0 ? (0.0 - a * Am) * d : 0.0 - (a * d) * Am;
    }
    for (int o = 1; o < L; o <<= 1) {
      const double Ap = __shfl_up(Am, o, L), Bp = __shfl_up(Bm, o, L);
      if (k >= o) {
        Bm = Am * Bp + Bm;
        Am = Am * Ap;
      }
    }
    const double Ai = __shfl_up(Am, 1, L), Bi = __shfl_up(Bm, 1, L);
    z = k == 0 ? zin : Ai * zin + Bi;
  }
  for (int i = b; i < e; i++) {
    const int j = first + dir * i;
    z = rec_step<FORM>(z, X[j * xs], A[j], D[j * ds]);
    X[j * xs] = z;
  }
  return __shfl(z, serial ? 0 : L - 1, L);
}

// lanes per sequence: the largest power of two L <= 64 with nseq * L <= nt
// (nt threads in the workgroup; a group never straddles two waves)
__host__ __device__ inline int lanes_per_seq(int nseq, int nt) {
  int L = 1;
  while (nseq * 2 * L <= nt && 2 * L <= 64) L *= 2;
  return L;
}

__global__ void __launch_bounds__(256) k_solve_lds(char *arena, const Chan *chans, int kRK, int C, int scan, int fuse) {
  const Chan ch = chans[blockIdx.z];
  const View v(arena, ch);
  const int nxh = ch.nxh, ldc = ch.ldc, nyu = ch.nyu, nyv = ch.nyv, ncol = ch.ncol;
  // XCD-aware order: workgroups are dispatched round-robin over the 8 XCDs;
  // give each XCD a contiguous run of column groups so that neighbouring
  // groups, which share cache lines, share one L2.  The launch pads
  // gridDim.x to a multiple of 8, so this is a permutation.
  const int per = gridDim.x / 8;
  const int grp = (blockIdx.x % 8) * per + blockIdx.x / 8;
  const int c0 = grp * C;
  if (c0 >= ncol) return;
  const int lane = threadIdx.x, NT = blockDim.x;  // 1, 2 or 4 waves
  const size_t plane = (size_t)(nyu + PADROWS) * ldc;

  extern __shared__ double2 smem_d[];
  double2 *X1 = smem_d, *X3 = X1 + (size_t)nyu * C, *X2 = X3 + (size_t)nyu * C, *Ps = X2 + (size_t)nyv * C;
  double *DU = reinterpret_cast<double *>(Ps + (size_t)nyu * C), *DV = DU + (size_t)nyu * C, *DP = DV + (size_t)nyv * C;
  // row coefficients: Helmholtz sub/super-diagonals a1, a3 of U and V (wall
  // rows included), Poisson a3 = du dthdyv and a1 = du dthdyv(j-1), du, dy
  double *A1U = DP + (size_t)nyu * C, *A3U = A1U + nyu, *A1V = A3U + nyu, *A3V = A1V + nyu;
  double *A3P = A3V + nyu, *A1P = A3P + nyu, *DUY = A1P + nyu, *DYV = DUY + nyu, *DVY = DYV + nyu;
  double *KX = DVY + nyu, *KZ = KX + C;

  const double dt = v.scal[SC_DT];
  const double beta = rk_b(kRK) * dt / ch.Re;
  const double dtirk = (1.0 / dt) / (rk_a(kRK) + rk_b(kRK));
  const double dtrk = dt * (rk_a(kRK) + rk_b(kRK));
  const double *fU = v.fac + (size_t)((kRK - 1) * 4) * plane, *fV = fU + 2 * plane, *fP = v.fac + (size_t)12 * plane;
  const int ne = nyu - 2;
  const int ncl = min(C, ncol - c0);  // live columns of this group

  // p of the projection phase's first PB items per thread, loaded now so
  // that the load is not a round trip after the recurrences
  constexpr int PB = 4;
  double2 Ppre[PB];
#pragma unroll
  for (int u = 0; u < PB; u++) {
    const int t = min(u * NT + lane, ne * C - 1);
    const int q = 1 + t / C, c = min(c0 + t % C, ncol - 1);
    Ppre[u] = v.p[(size_t)q * ldc + c];
  }
  // ---- L: stage right-hand sides, pivots, row coefficients and the
  //      columns' wavenumbers (dead columns read column ncol-1); loads are
  //      issued LB iterations at a time so that they are all in flight.
  //      With `fuse` the right-hand sides are computed here (rhs_point)
  //      rather than read from r1, r2, r3.
  if (fuse) {
    const int nU = nyu * C;
#pragma unroll 2
    for (int t = lane; t < nU; t += NT) {
      const int j = t / C, cl = t - j * C, c = min(c0 + cl, ncol - 1);
      const size_t o = (size_t)j * ldc + c;
      const double du = fU[o], dp = fP[o], dv = j < nyv ? fV[o] : 0.0;
      double2 a, b, e;
      rhs_point(v, ch, kRK, j, c, c0 + cl < ncol, a, b, e);
      X1[t] = a;
      X3[t] = b;
      DU[t] = du;
      DP[t] = dp;
      if (j < nyv) {
        X2[t] = e;
        DV[t] = dv;
      }
    }
  } else {
    const int nU = nyu * C, nV = nyv * C;
    for (int t0 = 0; t0 < nU; t0 += NT * LB) {
      double2 a[LB], b[LB];
      double d[LB], e[LB];
#pragma unroll
      for (int u = 0; u < LB; u++) {
        const int t = min(t0 + u * NT + lane, nU - 1);
        const int j = t / C, c = min(c0 + (t - j * C), ncol - 1);
        const size_t o = (size_t)j * ldc + c;
        a[u] = v.r1[o];
        b[u] = v.r3[o];
        d[u] = fU[o];
        e[u] = fP[o];
      }
#pragma unroll
      for (int u = 0; u < LB; u++) {
        const int t = t0 + u * NT + lane;
        if (t < nU) {
          X1[t] = a[u];
          X3[t] = b[u];
          DU[t] = d[u];
          DP[t] = e[u];
        }
      }
    }
    for (int t0 = 0; t0 < nV; t0 += NT * LB) {
      double2 a[LB];
      double d[LB];
#pragma unroll
      for (int u = 0; u < LB; u++) {
        const int t = min(t0 + u * NT + lane, nV - 1);
        const int j = t / C, c = min(c0 + (t - j * C), ncol - 1);
        const size_t o = (size_t)j * ldc + c;
        a[u] = v.r2[o];
        d[u] = fV[o];
      }
#pragma unroll
      for (int u = 0; u < LB; u++) {
        const int t = t0 + u * NT + lane;
        if (t < nV) {
          X2[t] = a[u];
          DV[t] = d[u];
        }
      }
    }
  }
  for (int j0 = 0; j0 < nyu; j0 += 4 * NT) {
    double c0u[4], c2u[4], c0v[4], c2v[4], hu[4], hv[4], hvm[4], yvj[4], yvm[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int j = min(j0 + u * NT + lane, nyu - 1), jv = min(j, nyv - 1), jm = max(min(j - 1, nyv - 1), 0);
      c0u[u] = v.dyu2i[3 * j];
      c2u[u] = v.dyu2i[3 * j + 2];
      c0v[u] = v.dyv2i[3 * jv];
      c2v[u] = v.dyv2i[3 * jv + 2];
      hu[u] = v.dthdyu[j];
      hv[u] = v.dthdyv[jv];
      hvm[u] = v.dthdyv[jm];
      yvj[u] = v.yv[jv];
      yvm[u] = v.yv[jm];
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int j = j0 + u * NT + lane;
      if (j >= nyu) break;
      const bool iu = j >= 1 && j <= nyu - 2, iv = j >= 1 && j <= nyv - 2;
      A1U[j] = iu ? -beta * c0u[u] : ch.gw2;
      A3U[j] = iu ? -beta * c2u[u] : 0.0;
      A1V[j] = iv ? -beta * c0v[u] : 0.0;
      A3V[j] = iv ? -beta * c2v[u] : 0.0;
      A3P[j] = (j >= 1 && j < ne) ? hu[u] * hv[u] : 0.0;
      A1P[j] = (j >= 2 && j <= ne) ? (-hu[u]) * (-hvm[u]) : 0.0;
      DUY[j] = hu[u];
      DYV[j] = (j >= 1 && j < nyv) ? yvj[u] - yvm[u] : 0.0;
      DVY[j] = j < nyv ? hv[u] : 0.0;
    }
  }
  if (lane < C) {
    const int c = min(c0 + lane, ncol - 1);
    KX[lane] = v.kx[c % nxh];
    KZ[lane] = v.kz[c / nxh];
  }
  __syncthreads();

  // ---- H: Helmholtz forward and back substitution, one real sequence per
  //      lane (rows read SB2 ahead into registers), or, with scan, split over
  //      the lanes of a group (rec_sweep)
  if (scan) {
    const int L = lanes_per_seq(6 * C, NT), sq = lane / L, k = lane % L;
    const int role = sq / C, cl = sq % C;
    if (role < 6 && cl < ncl) {
      const int sys = role >> 1, part = role & 1;
      double *X = reinterpret_cast<double *>(sys == 0 ? X1 : (sys == 1 ? X3 : X2)) + 2 * cl + part;
      const double *D = (sys == 2 ? DV : DU) + cl;
      const double *A1 = sys == 2 ? A1V : A1U, *A3 = sys == 2 ? A3V : A3U;
      const int n = sys == 2 ? nyv : nyu;
      const int C2 = 2 * C;
      // forward: rows 1 .. n-1 from y0 = r0 (row 0: a(2,0) = 1)
      const double y = rec_sweep<0>(X, C2, A1, D, C, 1, 1, n - 1, X[0], k, L, false);
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      // back substitution: rows n-2 .. 1 from x(n-1) = y(n-1); c_j = a3_j d_j
      const double x = rec_sweep<1>(X, C2, A3, D, C, n - 2, -1, n - 2, y, k, L, false);
      if (k == 0) X[0] = X[0] - (sys == 2 ? 0.0 : ch.gw1) * x;
    }
  } else {
    const int cl = lane % C, role = lane / C;
    if (role < 6 && cl < ncl) {
      const int sys = role >> 1, part = role & 1;
      double *X = reinterpret_cast<double *>(sys == 0 ? X1 : (sys == 1 ? X3 : X2)) + 2 * cl + part;
      const double *D = (sys == 2 ? DV : DU) + cl;
      const double *A1 = sys == 2 ? A1V : A1U, *A3 = sys == 2 ? A3V : A3U;
      const int n = sys == 2 ? nyv : nyu;
      const int C2 = 2 * C;
      double y = X[0];  // row 0: a(2,0) = 1
      int j = 1;
      for (; j + SB2 - 1 <= n - 1; j += SB2) {
        double r[SB2], d[SB2], a[SB2];
#pragma unroll
        for (int q = 0; q < SB2; q++) {
          r[q] = X[(j + q) * C2];
          d[q] = D[(j + q) * C];
          a[q] = A1[j + q];
        }
#pragma unroll
        for (int q = 0; q < SB2; q++) {
          y = (r[q] - a[q] * y) * d[q];
          X[(j + q) * C2] = y;
        }
      }
      for (; j <= n - 1; j++) {
        y = (X[j * C2] - A1[j] * y) * D[j * C];
        X[j * C2] = y;
      }
      // back substitution; c_j = a3_j d_j, and c_0 = gw1 (U) or 0 (V)
      double x = y;
      j = n - 2;
      for (; j - SB2 + 1 >= 1; j -= SB2) {
        double r[SB2], d[SB2], a[SB2];
#pragma unroll
        for (int q = 0; q < SB2; q++) {
          r[q] = X[(j - q) * C2];
          d[q] = D[(j - q) * C];
          a[q] = A3[j - q];
        }
#pragma unroll
        for (int q = 0; q < SB2; q++) {
          x = r[q] - (a[q] * d[q]) * x;
          X[(j - q) * C2] = x;
        }
      }
      for (; j >= 1; j--) {
        x = X[j * C2] - (A3[j] * D[j * C]) * x;
        X[j * C2] = x;
      }
      X[0] = X[0] - (sys == 2 ? 0.0 : ch.gw1) * x;
    }
  }
  __syncthreads();

  // ---- R: Poisson right-hand side on pressure rows 1..ne
  for (int t = lane; t < ne * C; t += NT) {
    const int q = 1 + t / C, cl = t % C;
    const double2 x1 = X1[q * C + cl], x3 = X3[q * C + cl], x2 = X2[q * C + cl], n2 = X2[(q - 1) * C + cl];
    const double2 a = cimul(KX[cl], x1), b = cimul(KZ[cl], x3);
    const double w = DUY[q];
    const double2 dv = make_double2((a.x + b.x) + w * (x2.x - n2.x), (a.y + b.y) + w * (x2.y - n2.y));
    Ps[q * C + cl] = make_double2(dtirk * dv.x, dtirk * dv.y);
  }
  __syncthreads();

  // ---- P: flow rate of the mean mode (its workgroup sums over all lanes)
  //      and flowrate_corr's g; Poisson elimination (top down) and back
  //      substitution
  __shared__ double g_sh;
  if (c0 == 0 && lane < 64) {  // one wave
    const double *x1r = reinterpret_cast<const double *>(X1);  // column 0, real part
    double part = 0.0;
    for (int r = 1 + lane; r <= ne; r += 64) part = part + DYV[r] * x1r[r * 2 * C];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) part += __shfl_xor(part, off);
    if (lane == 0) {
      double g = 0.0;
      if (ch.flag_ctpress == 0) {
        g = (v.scal[SC_QXT] - part) / v.scal[SC_QX1 + kRK - 1];
        v.scal[SC_MPGX] = v.scal[SC_MPGX] + g;
        v.scal[SC_DGX] = g;
      }
      g_sh = g;
    }
  }
  if (scan) {
    const int L = lanes_per_seq(2 * C, NT), sq = lane / L, k = lane % L;
    const int role = sq / C, cl = sq % C;
    const bool col0 = c0 + cl == 0;
    if (role < 2 && cl < ncl) {
      double *Y = reinterpret_cast<double *>(Ps) + 2 * cl + role;
      const double *D = DP + cl;
      const int C2 = 2 * C;
      // elimination from the top: rows ne .. 2 from 0 (the mean mode too: its
      // Neumann multipliers are 1, a prefix sum that the scan reassociates)
      const double yp = rec_sweep<0>(Y, C2, A3P, D, C, ne, -1, ne - 1, 0.0, k, L, false);
      // row ns = 1; the mean mode's row is pinned (a3 = 0, rhs = 0)
      const double d1 = D[C];
      const double y1 = col0 ? (0.0 - 0.0 * yp) * d1 : (Y[C2] - A3P[1] * yp) * d1;
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      if (k == 0) Y[C2] = y1;
      // back substitution: rows 2 .. ne from psi(1)
      rec_sweep<1>(Y, C2, A1P, D, C, 2, 1, ne - 1, y1, k, L, false);
    }
  } else {
    const int cl = lane % C, role = lane / C;
    const bool col0 = c0 + cl == 0;
    if (role < 2 && cl < ncl) {
      double *Y = reinterpret_cast<double *>(Ps) + 2 * cl + role;
      const double *D = DP + cl;
      const int C2 = 2 * C;
      double yp = 0.0;
      int q = ne;
      for (; q - SB2 + 1 >= 2; q -= SB2) {
        double r[SB2], d[SB2], a[SB2];
#pragma unroll
        for (int u = 0; u < SB2; u++) {
          r[u] = Y[(q - u) * C2];
          d[u] = D[(q - u) * C];
          a[u] = A3P[q - u];
        }
#pragma unroll
        for (int u = 0; u < SB2; u++) {
          yp = (r[u] - a[u] * yp) * d[u];
          Y[(q - u) * C2] = yp;
        }
      }
      for (; q >= 2; q--) {
        yp = (Y[q * C2] - A3P[q] * yp) * D[q * C];
        Y[q * C2] = yp;
      }
      {  // row ns = 1; the mean mode's row is pinned (a3 = 0, rhs = 0)
        const double d = D[C];
        const double yb = yp;
        yp = col0 ? (0.0 - 0.0 * yb) * d : (Y[C2] - A3P[1] * yp) * d;
        Y[C2] = yp;
      }
      double ps = yp;
      int j = 2;
      for (; j + SB2 - 1 <= ne; j += SB2) {
        double r[SB2], d[SB2], a[SB2];
#pragma unroll
        for (int u = 0; u < SB2; u++) {
          r[u] = Y[(j + u) * C2];
          d[u] = D[(j + u) * C];
          a[u] = A1P[j + u];
        }
#pragma unroll
        for (int u = 0; u < SB2; u++) {
          ps = r[u] - (a[u] * d[u]) * ps;
          Y[(j + u) * C2] = ps;
        }
      }
      for (; j <= ne; j++) {
        ps = Y[j * C2] - (A1P[j] * D[j * C]) * ps;
        Y[j * C2] = ps;
      }
    }
  }
  __syncthreads();

  // ---- U: p += psi; u1, u3 -= dt(a+b) grad_xz psi (+ flow-rate correction);
  //      v -= dt(a+b) D_y psi
  const double *__restrict__ u11 = v.u11 + (size_t)(kRK - 1) * nyu;
  const bool has0 = c0 == 0;
  const double g = has0 ? g_sh : 0.0;
  for (int t0 = 0; t0 < ne * C; t0 += NT * LB) {
    double2 P[LB];
#pragma unroll
    for (int u = 0; u < LB; u++) {
      if (t0 == 0 && u < PB) {
        P[u] = Ppre[u];
        continue;
      }
      const int t = min(t0 + u * NT + lane, ne * C - 1);
      const int q = 1 + t / C, c = min(c0 + t % C, ncol - 1);
      P[u] = v.p[(size_t)q * ldc + c];
    }
#pragma unroll
    for (int u = 0; u < LB; u++) {
      const int t = t0 + u * NT + lane;
      if (t >= ne * C) break;
      const int q = 1 + t / C, cl = t % C, c = c0 + cl;
      const double2 ps = Ps[q * C + cl];
      const double dkx = dtrk * KX[cl], dkz = dtrk * KZ[cl];
      const double2 a = cimul(dkx, ps), b = cimul(dkz, ps);
      const double2 U1 = X1[q * C + cl], U3 = X3[q * C + cl];
      double2 n1 = make_double2(U1.x - a.x, U1.y - a.y);
      const double2 n3 = make_double2(U3.x - b.x, U3.y - b.y);
      if (c == 0) n1.x = n1.x + g * u11[q];
      X1[q * C + cl] = n1;
      X3[q * C + cl] = n3;
      if (q <= ne - 1) {
        const double2 pn = Ps[(q + 1) * C + cl], V = X2[q * C + cl];
        const double sv = dtrk * DVY[q];
        X2[q * C + cl] = make_double2(V.x - sv * (pn.x - ps.x), V.y - sv * (pn.y - ps.y));
      }
      if (c < ncol) {
        const size_t o = (size_t)q * ldc + c;
        v.p[o] = make_double2(P[u].x + ps.x, P[u].y + ps.y);
        v.psi[o] = ps;
      }
    }
  }
  __syncthreads();

  // ---- W: divergence, ghost and wall rows, stores
  for (int t = lane; t < nyu * C; t += NT) {
    const int j = t / C, cl = t % C, c = c0 + cl;
    if (c >= ncol) continue;
    const size_t o = (size_t)j * ldc + c;
    double2 n1, n3;
    if (j == 0 || j == nyu - 1) {  // a band edge keeps its solved ghost value
      const bool edge = (ch.band_edge >> (j == 0 ? 0 : 1)) & 1;
      const int jn = j == 0 ? 1 : ne;
      n1 = edge ? X1[j * C + cl] : cscale(ch.gwbc1, X1[jn * C + cl]);
      n3 = edge ? X3[j * C + cl] : cscale(ch.gwbc3, X3[jn * C + cl]);
    } else {
      n1 = X1[j * C + cl];
      n3 = X3[j * C + cl];
      const double2 da = cimul(KX[cl], n1), db = cimul(KZ[cl], n3);
      const double2 vq = j == ne ? make_double2(0, 0) : X2[j * C + cl];
      const double2 vm = j == 1 ? make_double2(0, 0) : X2[(j - 1) * C + cl];
      const double w = DUY[j];
      v.div[o] = make_double2((da.x + db.x) + w * (vq.x - vm.x), (da.y + db.y) + w * (vq.y - vm.y));
    }
    v.u1[o] = n1;
    v.u3[o] = n3;
    if (j < nyv) v.u2[o] = (j == 0 || j == nyv - 1) ? make_double2(0, 0) : X2[j * C + cl];
  }
}
#undef SB2
#undef LB

// ------------------------------------------------------------------------
// K5 (wave form): one spectral column per wavefront, NW consecutive
// columns per workgroup.  Lane l of a wave owns the contiguous rows
// l K .. l K + K - 1 of its column and keeps every value of them in
// registers between the loads and the stores.  Each recurrence of the solve
// (U/W and V Helmholtz forward and back substitution, Poisson elimination
// from the top and back substitution) runs as a chunked scan: every lane
// composes its chunk's affine map z -> A z + B, the wave scans the maps in
// six shuffle steps, and every lane reruns its chunk from its true start
// value.  Every stored row therefore comes from the serial sweep's
// operations (LUsolU_W / LUsolV / LUsolP, tridLU_3D.f90:55-226, with
// LU_dec's factors); only the chunk start values carry the scan's
// reassociation.  Each input row is read once and each output row written
// once, with no intermediate in HBM.
// HBM is reached through an LDS tile: the workgroup moves one field at a
// time, a row of its NW columns per NW * 16 contiguous bytes (64 / NW rows per
// wave instruction), and each wave reads or writes its own column in the
// tile.  (Gathering a column straight from HBM would touch 64 rows per
// instruction, 528 KB apart at 256 x 256: a separate cache line and page
// translation per lane.)
// Rows outside a recurrence's range are identity maps (A = 1, B = 0), and
// every scan starts from z = 0: the first row of each forward range ignores
// its input (a1 = 0 on row 0; the top or pinned Poisson row), and the last
// row of each backward range has c = 0.
template <int S>
__device__ __forceinline__ void wave_scan(double &A, double (&B)[S], int lane, bool up) {
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const double Ap = up ? __shfl_up(A, o, 64) : __shfl_down(A, o, 64);
    double Bp[S];
#pragma unroll
    for (int s = 0; s < S; s++) Bp[s] = up ? __shfl_up(B[s], o, 64) : __shfl_down(B[s], o, 64);
    if (up ? lane >= o : lane + o < 64) {
#pragma unroll
      for (int s = 0; s < S; s++) B[s] = A * Bp[s] + B[s];
      A = A * Ap;
    }
  }
  // exclusive: the value entering the lane's chunk (0 for the first lane)
#pragma unroll
  for (int s = 0; s < S; s++) {
    const double b = up ? __shfl_up(B[s], 1, 64) : __shfl_down(B[s], 1, 64);
    B[s] = (up ? lane == 0 : lane == 63) ? 0.0 : b;
  }
}

// staging helpers of k_solve_wave: thread tid of a workgroup of NW waves
// moves row (u 64 NW + tid) / NW of column c0 + tid % NW, u < K
template <int K, int NW>
__device__ __forceinline__ void sw_fetch(double2 (&t)[K], const double2 *F, int n, int ldc, int c0, int ncol,
                                         int tid) {
#pragma unroll
  for (int u = 0; u < K; u++) {
    const int idx = u * 64 * NW + tid, r = min(idx / NW, n - 1), cc = min(c0 + idx % NW, ncol - 1);
    t[u] = F[(size_t)r * ldc + cc];
  }
}
// the same elements' right-hand sides computed in place of the loads of r1
// and r3 (rhs_u) or r2 (rhs_v); each live element puts its new explicit
// terms once (clamped copies do not); the wall rows of textured walls hold
// the wall sources, which are read
template <int K, int NW>
__device__ __forceinline__ void sw_rhs_u(double2 (&t)[K], double2 (&t2)[K], const View &v, const Chan &ch, int kRK,
                                         int c0, int tid) {
  const int nyu = ch.nyu, ncol = ch.ncol, ldc = ch.ldc;
  double2 n1[K], n3[K];
  bool put[K];
#pragma unroll
  for (int u = 0; u < K; u++) {
    const int idx = u * 64 * NW + tid, r = min(idx / NW, nyu - 1), cr = c0 + idx % NW, cc = min(cr, ncol - 1);
    put[u] = false;
    if ((r == 0 && (ch.tex & 1)) || (r == nyu - 1 && (ch.tex & 2))) {
      t[u] = v.r1[(size_t)r * ldc + cc];
      t2[u] = v.r3[(size_t)r * ldc + cc];
    } else {
      put[u] = rhs_u(v, ch, kRK, r, cc, false, t[u], t2[u], &n1[u], &n3[u]) && cr < ncol && idx / NW < nyu;
    }
  }
#pragma unroll
  for (int u = 0; u < K; u++) {
    const int idx = u * 64 * NW + tid;
    if (put[u]) {
      const size_t o = (size_t)(idx / NW) * ldc + c0 + idx % NW;
      v.N1o[o] = n1[u];
      v.N3o[o] = n3[u];
    }
  }
}
template <int K, int NW>
__device__ __forceinline__ void sw_rhs_v(double2 (&t)[K], const View &v, const Chan &ch, int kRK, int c0, int tid) {
  const int nyv = ch.nyv, ncol = ch.ncol, ldc = ch.ldc;
  double2 n2[K];
  bool put[K];
#pragma unroll
  for (int u = 0; u < K; u++) {
    const int idx = u * 64 * NW + tid, r = min(idx / NW, nyv - 1), cr = c0 + idx % NW, cc = min(cr, ncol - 1);
    put[u] = rhs_v(v, ch, kRK, r, cc, false, t[u], &n2[u]) && cr < ncol && idx / NW < nyv;
  }
#pragma unroll
  for (int u = 0; u < K; u++) {
    const int idx = u * 64 * NW + tid;
    if (put[u]) v.N2o[(size_t)(idx / NW) * ldc + c0 + idx % NW] = n2[u];
  }
}
template <int K, int NW>
__device__ __forceinline__ void sw_fetch2(double2 (&t)[K], const double *Fa, int na, const double *Fb, int nb, int ldc,
                                          int c0, int ncol, int tid) {
#pragma unroll
  for (int u = 0; u < K; u++) {
    const int idx = u * 64 * NW + tid, r = idx / NW, cc = min(c0 + idx % NW, ncol - 1);
    t[u] = make_double2(Fa[(size_t)min(r, na - 1) * ldc + cc], Fb[(size_t)min(r, nb - 1) * ldc + cc]);
  }
}
template <int K, int NW>
__device__ __forceinline__ void sw_put(double2 *T, const double2 (&t)[K], int S, int nyu, int tid) {
#pragma unroll
  for (int u = 0; u < K; u++) {
    const int idx = u * 64 * NW + tid, r = min(idx / NW, nyu - 1);
    T[(idx % NW) * S + r] = t[u];  // rows past nyu - 1 rewrite row nyu - 1 with its own value
  }
}
template <int K>
__device__ __forceinline__ void sw_get(double2 (&X)[K], const double2 *col, int j0, int nyu) {
#pragma unroll
  for (int q = 0; q < K; q++) X[q] = col[min(j0 + q, nyu - 1)];
}

template <int K>
__device__ __forceinline__ void sw_putcol(double2 *col, const double2 (&X)[K], int j0, int nyu) {
#pragma unroll
  for (int q = 0; q < K; q++)
    if (j0 + q < nyu) col[j0 + q] = X[q];
}
// rows jlo..jhi of the live columns from the tile to F (none when F is
// null); with ADDP, p (its values in tp, fetched by sw_fetch) gets the tile
// added
template <int K, int NW, bool ADDP>
__device__ __forceinline__ void sw_store(const double2 *T, double2 *F, double2 *p, const double2 (&tp)[K], int S,
                                         int jlo, int jhi, int ldc, int c0, int ncol, int tid) {
#pragma unroll
  for (int u = 0; u < K; u++) {
    const int idx = u * 64 * NW + tid, r = idx / NW, cc = c0 + idx % NW;
    if (r >= jlo && r <= jhi && cc < ncol) {
      const double2 x = T[(idx % NW) * S + r];
      const size_t o = (size_t)r * ldc + cc;
      if (ADDP) p[o] = make_double2(tp[u].x + x.x, tp[u].y + x.y);
      if (F) F[o] = x;
    }
  }
}

// FUSE: the right-hand sides are computed in the staging (sw_rhs_u/_v)
// instead of read from r1, r3, r2
// keep = 0: psi and div are not stored (p still gets psi added): the first
// two substeps of a captured step, whose psi and div nothing reads
template <int K, int NW, bool FUSE>
__global__ void __launch_bounds__(64 * NW) k_solve_wave(char *arena, const Chan *chans, int kRK, int keep) {
  const Chan ch = chans[blockIdx.z];
  const int per = gridDim.x / 8;
  const int grp = (blockIdx.x % 8) * per + blockIdx.x / 8;
  const int c0 = grp * NW;
  if (c0 >= ch.ncol) return;  // the whole workgroup, before any barrier
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int c = c0 + w;  // waves past the last column compute a copy of it and store nothing
  const View v(arena, ch);
  const int nxh = ch.nxh, ldc = ch.ldc, nyu = ch.nyu, nyv = ch.nyv, ne = nyu - 2, ncol = ch.ncol;
  const int S = solve_wave_stride(nyu);
  const size_t plane = (size_t)(nyu + PADROWS) * ldc;
  const double dt = v.scal[SC_DT];
  const double beta = rk_b(kRK) * dt / ch.Re;
  const double dtirk = (1.0 / dt) / (rk_a(kRK) + rk_b(kRK));
  const double dtrk = dt * (rk_a(kRK) + rk_b(kRK));
  const double *fU = v.fac + (size_t)((kRK - 1) * 4) * plane, *fV = fU + 2 * plane, *fP = v.fac + (size_t)12 * plane;
  const int cl = min(c, ncol - 1);
  const double kxi = v.kx[cl % nxh], kzk = v.kz[cl / nxh];
  const bool col0 = c == 0;
  const int j0 = lane * K;

  extern __shared__ double2 smem_w[];
  double2 *T = smem_w;  // staging tile [NW][S]
  double *A1U = reinterpret_cast<double *>(T + NW * S), *A3U = A1U + nyu, *A1V = A3U + nyu, *A3V = A1V + nyu;
  double *A3P = A3V + nyu, *A1P = A3P + nyu, *DUY = A1P + nyu, *DVY = DUY + nyu, *DYV = DVY + nyu;

  // ---- staged loads, two fields in flight at a time (issued before the
  //      barrier that frees the tile); p is loaded with the last pair, for
  //      the final p += psi.  (Three in flight make the compiler keep one
  //      staging array in scratch memory.)
  double2 t[K], t2[K], tp[K];
  if constexpr (FUSE) {
    sw_rhs_u<K, NW>(t, t2, v, ch, kRK, c0, tid);
  } else {
    sw_fetch<K, NW>(t, v.r1, nyu, ldc, c0, ncol, tid);
    sw_fetch<K, NW>(t2, v.r3, nyu, ldc, c0, ncol, tid);
  }

  // ---- row coefficients, shared by the workgroup's columns: sub- and
  //      super-diagonals of the Helmholtz operators (wall rows included:
  //      a1 = 0 and a3 = gw1 on row 0, a1 = gw2 on row nyu-1), of the
  //      Poisson operator, and the y metric terms
  for (int j = tid; j < nyu; j += 64 * NW) {
    const int jv = min(j, nyv - 1), jm = max(min(j - 1, nyv - 1), 0);
    const bool iu = j >= 1 && j <= nyu - 2, iv = j >= 1 && j <= nyv - 2;
    const double huj = v.dthdyu[j], hvj = v.dthdyv[jv], hvm = v.dthdyv[jm];
    A1U[j] = iu ? -beta * v.dyu2i[3 * j] : (j == 0 ? 0.0 : ch.gw2);
    A3U[j] = iu ? -beta * v.dyu2i[3 * j + 2] : (j == 0 ? ch.gw1 : 0.0);
    A1V[j] = iv ? -beta * v.dyv2i[3 * jv] : 0.0;
    A3V[j] = iv ? -beta * v.dyv2i[3 * jv + 2] : 0.0;
    A3P[j] = (j >= 1 && j < ne) ? huj * hvj : 0.0;
    A1P[j] = (j >= 2 && j <= ne) ? (-huj) * (-hvm) : 0.0;
    DUY[j] = huj;
    DVY[j] = j < nyv ? hvj : 0.0;
    DYV[j] = (j >= 1 && j < nyv) ? v.yv[jv] - v.yv[jm] : 0.0;
  }

  double2 X1[K], X3[K], X2[K], DUV[K];
  double DP[K];
  sw_put<K, NW>(T, t, S, nyu, tid);
  __syncthreads();
  sw_get<K>(X1, T + w * S, j0, nyu);
  __syncthreads();
  sw_put<K, NW>(T, t2, S, nyu, tid);
  __syncthreads();
  sw_get<K>(X3, T + w * S, j0, nyu);
  if constexpr (FUSE)
    sw_rhs_v<K, NW>(t, v, ch, kRK, c0, tid);
  else
    sw_fetch<K, NW>(t, v.r2, nyv, ldc, c0, ncol, tid);
  sw_fetch2<K, NW>(t2, fU, nyu, fV, nyv, ldc, c0, ncol, tid);
  __syncthreads();
  sw_put<K, NW>(T, t, S, nyu, tid);
  __syncthreads();
  sw_get<K>(X2, T + w * S, j0, nyu);
  __syncthreads();
  sw_put<K, NW>(T, t2, S, nyu, tid);
  __syncthreads();
  sw_get<K>(DUV, T + w * S, j0, nyu);
  sw_fetch2<K, NW>(t, fP, nyu, fP, nyu, ldc, c0, ncol, tid);
  sw_fetch<K, NW>(tp, v.p, nyu, ldc, c0, ncol, tid);
  __syncthreads();
  sw_put<K, NW>(T, t, S, nyu, tid);
  __syncthreads();
  {
    double2 dp[K];
    sw_get<K>(dp, T + w * S, j0, nyu);
#pragma unroll
    for (int q = 0; q < K; q++) DP[q] = dp[q].x;
  }

  // ---- Helmholtz forward substitution, U/W rows 0..nyu-1 and V rows
  //      0..nyv-1: y_j = (r_j - a1_j y_(j-1)) d_j
  {
    double AU = 1.0, AV = 1.0, BU[4] = {0.0, 0.0, 0.0, 0.0}, BV[2] = {0.0, 0.0};
#pragma unroll
    for (int q = 0; q < K; q++) {
      const int j = j0 + q, jc = min(j, nyu - 1);
      if (j < nyu) {
        const double a = A1U[jc], d = DUV[q].x;
        BU[0] = (X1[q].x - a * BU[0]) * d;
        BU[1] = (X1[q].y - a * BU[1]) * d;
        BU[2] = (X3[q].x - a * BU[2]) * d;
        BU[3] = (X3[q].y - a * BU[3]) * d;
        AU = (0.0 - a * AU) * d;
      }
      if (j < nyv) {
        const double a = A1V[jc], d = DUV[q].y;
        BV[0] = (X2[q].x - a * BV[0]) * d;
        BV[1] = (X2[q].y - a * BV[1]) * d;
        AV = (0.0 - a * AV) * d;
      }
    }
    wave_scan<4>(AU, BU, lane, true);
    wave_scan<2>(AV, BV, lane, true);
    asm volatile("" ::: "memory");  // reread the coefficients rather than hold them across the scan
    double2 y1 = make_double2(BU[0], BU[1]), y3 = make_double2(BU[2], BU[3]), y2 = make_double2(BV[0], BV[1]);
#pragma unroll
    for (int q = 0; q < K; q++) {
      const int j = j0 + q, jc = min(j, nyu - 1);
      if (j < nyu) {
        const double a = A1U[jc], d = DUV[q].x;
        y1 = make_double2((X1[q].x - a * y1.x) * d, (X1[q].y - a * y1.y) * d);
        y3 = make_double2((X3[q].x - a * y3.x) * d, (X3[q].y - a * y3.y) * d);
        X1[q] = y1;
        X3[q] = y3;
      }
      if (j < nyv) {
        const double a = A1V[jc], d = DUV[q].y;
        y2 = make_double2((X2[q].x - a * y2.x) * d, (X2[q].y - a * y2.y) * d);
        X2[q] = y2;
      }
    }
  }
  // ---- back substitution: x_j = y_j - (a3_j d_j) x_(j+1); c = 0 on the
  //      last row of each system, c_0 = gw1 (U/W) or 0 (V)
  {
    double AU = 1.0, AV = 1.0, BU[4] = {0.0, 0.0, 0.0, 0.0}, BV[2] = {0.0, 0.0};
#pragma unroll
    for (int q = K - 1; q >= 0; q--) {
      const int j = j0 + q, jc = min(j, nyu - 1);
      if (j < nyu) {
        const double e = A3U[jc] * DUV[q].x;
        BU[0] = X1[q].x - e * BU[0];
        BU[1] = X1[q].y - e * BU[1];
        BU[2] = X3[q].x - e * BU[2];
        BU[3] = X3[q].y - e * BU[3];
        AU = 0.0 - e * AU;
      }
      if (j < nyv) {
        const double e = A3V[jc] * DUV[q].y;
        BV[0] = X2[q].x - e * BV[0];
        BV[1] = X2[q].y - e * BV[1];
        AV = 0.0 - e * AV;
      }
    }
    wave_scan<4>(AU, BU, lane, false);
    wave_scan<2>(AV, BV, lane, false);
    asm volatile("" ::: "memory");
    double2 x1 = make_double2(BU[0], BU[1]), x3 = make_double2(BU[2], BU[3]), x2 = make_double2(BV[0], BV[1]);
#pragma unroll
    for (int q = K - 1; q >= 0; q--) {
      const int j = j0 + q, jc = min(j, nyu - 1);
      if (j < nyu) {
        const double e = A3U[jc] * DUV[q].x;
        x1 = make_double2(X1[q].x - e * x1.x, X1[q].y - e * x1.y);
        x3 = make_double2(X3[q].x - e * x3.x, X3[q].y - e * x3.y);
        X1[q] = x1;
        X3[q] = x3;
      }
      if (j < nyv) {
        const double e = A3V[jc] * DUV[q].y;
        x2 = make_double2(X2[q].x - e * x2.x, X2[q].y - e * x2.y);
        X2[q] = x2;
      }
    }
  }

  // ---- flow rate of the mean mode and flowrate_corr's g
  //      (littleharsh.f90:476-509): Qx = sum (yv_j - yv_(j-1)) Re u1_j
  double g = 0.0;
  if (col0) {
    double part = 0.0;
#pragma unroll
    for (int q = 0; q < K; q++) {
      const int j = j0 + q;
      if (j >= 1 && j <= ne) part = part + DYV[j] * X1[q].x;
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) part += __shfl_xor(part, off, 64);
    if (ch.flag_ctpress == 0) {
      g = (v.scal[SC_QXT] - part) / v.scal[SC_QX1 + kRK - 1];
      if (lane == 0) {
        v.scal[SC_MPGX] = v.scal[SC_MPGX] + g;
        v.scal[SC_DGX] = g;
      }
    }
  }

  // ---- Poisson right-hand side on pressure rows 1..ne: div / (dt (a+b))
  //      (littleharsh.f90:237-272,581-606); the mean mode's row 1 is pinned
  //      (a3 = 0, rhs = 0, tridLU_3D.f90:384-390)
  double2 PS[K];
  {
    const double2 vlast = X2[K - 1];
    const double2 vin = make_double2(__shfl_up(vlast.x, 1, 64), __shfl_up(vlast.y, 1, 64));
#pragma unroll
    for (int q = 0; q < K; q++) {
      const int j = j0 + q, jc = min(j, nyu - 1);
      const double2 vm = q ? X2[q - 1] : vin;
      const double2 a = cimul(kxi, X1[q]), b = cimul(kzk, X3[q]);
      const double wy = DUY[jc];
      const double2 dv = make_double2((a.x + b.x) + wy * (X2[q].x - vm.x), (a.y + b.y) + wy * (X2[q].y - vm.y));
      PS[q] = (j >= 1 && j <= ne && !(col0 && j == 1)) ? make_double2(dtirk * dv.x, dtirk * dv.y)
                                                       : make_double2(0, 0);
    }
  }
  // ---- Poisson elimination from the top, rows ne..1
  {
    double A = 1.0, B[2] = {0.0, 0.0};
#pragma unroll
    for (int q = K - 1; q >= 0; q--) {
      const int j = j0 + q, jc = min(j, nyu - 1);
      if (j >= 1 && j <= ne) {
        const double a = (col0 && j == 1) ? 0.0 : A3P[jc], d = DP[q];
        B[0] = (PS[q].x - a * B[0]) * d;
        B[1] = (PS[q].y - a * B[1]) * d;
        A = (0.0 - a * A) * d;
      }
    }
    wave_scan<2>(A, B, lane, false);
    asm volatile("" ::: "memory");
    double2 z = make_double2(B[0], B[1]);
#pragma unroll
    for (int q = K - 1; q >= 0; q--) {
      const int j = j0 + q, jc = min(j, nyu - 1);
      if (j >= 1 && j <= ne) {
        const double a = (col0 && j == 1) ? 0.0 : A3P[jc], d = DP[q];
        z = make_double2((PS[q].x - a * z.x) * d, (PS[q].y - a * z.y) * d);
        PS[q] = z;
      }
    }
  }
  // ---- back substitution from psi(1): psi_j = y_j - (a1_j d_j) psi_(j-1)
  {
    double A = 1.0, B[2] = {0.0, 0.0};
#pragma unroll
    for (int q = 0; q < K; q++) {
      const int j = j0 + q, jc = min(j, nyu - 1);
      if (j >= 1 && j <= ne) {
        const double e = A1P[jc] * DP[q];
        B[0] = PS[q].x - e * B[0];
        B[1] = PS[q].y - e * B[1];
        A = 0.0 - e * A;
      }
    }
    wave_scan<2>(A, B, lane, true);
    asm volatile("" ::: "memory");
    double2 z = make_double2(B[0], B[1]);
#pragma unroll
    for (int q = 0; q < K; q++) {
      const int j = j0 + q, jc = min(j, nyu - 1);
      if (j >= 1 && j <= ne) {
        const double e = A1P[jc] * DP[q];
        z = make_double2(PS[q].x - e * z.x, PS[q].y - e * z.y);
        PS[q] = z;
      }
    }
  }

  // ---- u1, u3 -= dt(a+b) grad_xz psi (+ the flow-rate correction of the
  //      mean mode); v -= dt(a+b) D_y psi (littleharsh.f90:956-990)
  const double dkx = dtrk * kxi, dkz = dtrk * kzk;
  {
    const double2 pfirst = PS[0];
    const double2 pnext = make_double2(__shfl_down(pfirst.x, 1, 64), __shfl_down(pfirst.y, 1, 64));
#pragma unroll
    for (int q = 0; q < K; q++) {
      const int j = j0 + q, jc = min(j, nyu - 1);
      if (j >= 1 && j <= ne) {
        const double2 ps = PS[q];
        const double2 a = cimul(dkx, ps), b = cimul(dkz, ps);
        X1[q] = make_double2(X1[q].x - a.x, X1[q].y - a.y);
        if (col0) X1[q].x = X1[q].x + g * v.u11[(size_t)(kRK - 1) * nyu + j];
        X3[q] = make_double2(X3[q].x - b.x, X3[q].y - b.y);
        if (j <= ne - 1) {
          const double2 pn = q < K - 1 ? PS[q + 1] : pnext;
          const double sv = dtrk * DVY[jc];
          X2[q] = make_double2(X2[q].x - sv * (pn.x - ps.x), X2[q].y - sv * (pn.y - ps.y));
        }
      }
    }
  }

  // ---- ghost and wall rows (littleharsh.f90:992-1044)
  double2 u1a = make_double2(0, 0), u1b = u1a, u3a = u1a, u3b = u1a;
#pragma unroll
  for (int q = 0; q < K; q++) {
    const int j = j0 + q;
    if (j == 1) {
      u1a = X1[q];
      u3a = X3[q];
    }
    if (j == ne) {
      u1b = X1[q];
      u3b = X3[q];
    }
  }
  const int la = 1 / K, lb = ne / K;
  u1a = make_double2(__shfl(u1a.x, la, 64), __shfl(u1a.y, la, 64));
  u3a = make_double2(__shfl(u3a.x, la, 64), __shfl(u3a.y, la, 64));
  u1b = make_double2(__shfl(u1b.x, lb, 64), __shfl(u1b.y, lb, 64));
  u3b = make_double2(__shfl(u3b.x, lb, 64), __shfl(u3b.y, lb, 64));
#pragma unroll
  for (int q = 0; q < K; q++) {
    const int j = j0 + q;
    if (j == 0 && !(ch.band_edge & 1)) {  // a band edge keeps its solved ghost value
      X1[q] = cscale(ch.gwbc1, u1a);
      X3[q] = cscale(ch.gwbc3, u3a);
    }
    if (j == nyu - 1 && !(ch.band_edge & 2)) {
      X1[q] = cscale(ch.gwbc1, u1b);
      X3[q] = cscale(ch.gwbc3, u3b);
    }
    if (j == 0 || j == nyv - 1) X2[q] = make_double2(0, 0);
  }
  // divergence of the corrected field on pressure rows 1..ne (v(0) and
  // v(ne) are the walls)
  double2 DV[K];
  {
    const double2 vlast = X2[K - 1];
    const double2 vin = make_double2(__shfl_up(vlast.x, 1, 64), __shfl_up(vlast.y, 1, 64));
#pragma unroll
    for (int q = 0; q < K; q++) {
      const int j = j0 + q, jc = min(j, nyu - 1);
      const double2 da = cimul(kxi, X1[q]), db = cimul(kzk, X3[q]);
      const double2 vq = X2[q], vm = q ? X2[q - 1] : vin;
      const double wy = DUY[jc];
      DV[q] = make_double2((da.x + db.x) + wy * (vq.x - vm.x), (da.y + db.y) + wy * (vq.y - vm.y));
    }
  }

  // ---- staged stores: each wave writes its column into the tile, the
  //      workgroup stores rows jlo..jhi of its live columns (p gets psi
  //      added on the way)
  __syncthreads();
  sw_putcol<K>(T + w * S, X1, j0, nyu);
  __syncthreads();
  sw_store<K, NW, false>(T, v.u1, v.p, tp, S, 0, nyu - 1, ldc, c0, ncol, tid);
  __syncthreads();
  sw_putcol<K>(T + w * S, X3, j0, nyu);
  __syncthreads();
  sw_store<K, NW, false>(T, v.u3, v.p, tp, S, 0, nyu - 1, ldc, c0, ncol, tid);
  __syncthreads();
  sw_putcol<K>(T + w * S, X2, j0, nyu);
  __syncthreads();
  sw_store<K, NW, false>(T, v.u2, v.p, tp, S, 0, nyv - 1, ldc, c0, ncol, tid);
  __syncthreads();
  sw_putcol<K>(T + w * S, PS, j0, nyu);
  __syncthreads();
  sw_store<K, NW, true>(T, keep ? v.psi : nullptr, v.p, tp, S, 1, ne, ldc, c0, ncol, tid);
  if (!keep) return;
  __syncthreads();
  sw_putcol<K>(T + w * S, DV, j0, nyu);
  __syncthreads();
  sw_store<K, NW, false>(T, v.div, v.p, tp, S, 1, ne, ldc, c0, ncol, tid);
}

// columns (waves) per workgroup of the wave-form solve: 8 (128-byte rows of
// the staging tile), or 2 for columns of at most 192 rows when
// LH_TUNE_SOLVE_WAVES asks for it (4): small channels then spread over four
// times the workgroups
int solve_wave_cols(int nyu, int waves) { return waves == 4 && solve_wave_rows(nyu) <= 3 ? 2 : 8; }
// rows per lane of the wave-form solve (0: the column is too tall for it)
int solve_wave_rows(int nyu) {
  const int K = (nyu + 63) / 64;
  return K <= SOLVE_WAVE_KMAX ? K : 0;
}

hipError_t launch_factors(char *arena, const Chan *d_chans, const ShapeMax &m, hipStream_t s) {
  dim3 g((m.ncol + 63) / 64, 1, m.nchan);
  hipLaunchKernelGGL(k_factors, g, dim3(64), 0, s, arena, d_chans);
  return hipGetLastError();
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
// Time step at the first substep (FOU3D.f90:288-297): dtc from the CFL
// maxima of this substep's physical planes (dtc_calc, FOU3D.f90:699-736,
// with u2mloc starting at 1e-6), then dt = min(2.5 dtv, CFL dtc) when the
// channel's step is adaptive, and t += dt.  One workgroup per channel.
__global__ void __launch_bounds__(1024) k_dt_update(char *arena, const Chan *chans, int ntiles_max, int e_xprod) {
  __shared__ double red[48];
  dt_update_block(arena, chans[blockIdx.x], ntiles_max, e_xprod, red);
}

// unit-forcing solutions of the mean mode per substep (flowrate_corr's
// u11 = LUsol0(-dRK dt), littleharsh.f90:484-494, with LU_build0/LU_dec0,
// tridLU_3D.f90:409-435,520-537) and their flow rates, for the current dt;
// one thread per (kRK, channel)
__global__ void k_u11(char *arena, const Chan *chans) {
  const Chan ch = chans[blockIdx.x];
  View v(arena, ch);
  const int kRK = threadIdx.x + 1;
  if (kRK > 3) return;
  const int n = ch.nyu;
  const double dt = v.scal[SC_DT];
  const double beta = rk_b(kRK) * dt / ch.Re, bdt = -rk_d(kRK) * dt;
  double *x = v.u11 + (size_t)(kRK - 1) * n;
  // forward elimination with the pivots held in x's storage order
  // (d_0 = 1: row 0 is u0 + gw1 u1 = 0)
  double c = ch.gw1;  // a(3,0) * d_0
  x[0] = 0.0;
  double *cs = v.scr + (size_t)(kRK - 1) * n;  // upper diagonals c_j
  cs[0] = c;
  for (int j = 1; j < n; j++) {
    double a1, a2, a3;
    if (j <= n - 2) {
      a1 = -beta * v.dyu2i[3 * j];
      a2 = 1.0 - beta * v.dyu2i[3 * j + 1];
      a3 = -beta * v.dyu2i[3 * j + 2];
    } else {
      a1 = ch.gw2;
      a2 = 1.0;
      a3 = 0.0;
    }
    const double d = 1.0 / (a2 - a1 * c);
    c = a3 * d;
    cs[j] = c;
    const double r = j <= n - 2 ? bdt : 0.0;
    x[j] = (r - a1 * x[j - 1]) * d;
  }
  for (int j = n - 2; j >= 0; j--) x[j] = x[j] - cs[j] * x[j + 1];
  double Q = 0.0;
  for (int j = 1; j <= n - 2; j++) Q = Q + (v.yv[j] - v.yv[j - 1]) * x[j];
  v.scal[SC_QX1 + kRK - 1] = Q;
}

// History record of record_out (FOU3D.f90:3040-3070): err = max |div| over
// the pressure rows (error, littleharsh.f90:342-375), the flow rate of the
// mean mode (flowrateIm), Umax = max Re u1(0,0) over rows 0..ny (maxvel,
// littleharsh.f90:511-527) and the mean slip velocity interpolated to
// y = -1.  One workgroup per channel; out[ch*8 + 0..3] = err, Qx, Umax, Uslip.
__global__ void __launch_bounds__(1024) k_history(char *arena, const Chan *chans, double *out) {
  const Chan ch = chans[blockIdx.x];
  const View v(arena, ch);
  const int ncol = ch.ncol, L = ch.ldc, nyu = ch.nyu;
  double e = 0.0;
  const int n = (nyu - 2) * ncol;  // p rows 1..ny-1
  for (int q = threadIdx.x; q < n; q += blockDim.x) {
    const int r = q / ncol, c = q - r * ncol;
    const double2 d = v.div[(size_t)(r + 1) * L + c];
    e = fmax(e, sqrt(d.x * d.x + d.y * d.y));
  }
  for (int off = 32; off >= 1; off >>= 1) e = fmax(e, __shfl_xor(e, off));
  __shared__ double red[16];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = e;
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (int q = 1; q < (int)(blockDim.x >> 6); q++) e = fmax(e, red[q]);
  e = fmax(e, red[0]);
  double Q = 0.0, um = 0.0;
  for (int j = 1; j <= nyu - 2; j++) Q = Q + (v.yv[j] - v.yv[j - 1]) * v.u1[(size_t)j * L].x;
  for (int j = 0; j < nyu; j++) um = fmax(um, v.u1[(size_t)j * L].x);
  const double2 a = v.u1[L], b = v.u1[0];
  const double us = (a.x * (-1.0 - v.yu[0]) + b.x * (v.yu[1] + 1.0)) / (v.yu[1] - v.yu[0]);
  double *o = out + 8 * blockIdx.x;
  o[0] = e;
  o[1] = Q;
  o[2] = um;
  o[3] = us;
}

hipError_t launch_history(char *arena, const Chan *d_chans, const ShapeMax &m, double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_history, dim3(m.nchan), dim3(1024), 0, s, arena, d_chans, out);
  return hipGetLastError();
}

hipError_t launch_dt_update(char *arena, const Chan *d_chans, const ShapeMax &m, hipStream_t s) {
  hipLaunchKernelGGL(k_dt_update, dim3(m.nchan), dim3(1024), 0, s, arena, d_chans, m.xprod_tiles, m.e_xprod);
  return hipGetLastError();
}

hipError_t launch_u11(char *arena, const Chan *d_chans, const ShapeMax &m, hipStream_t s) {
  hipLaunchKernelGGL(k_u11, dim3(m.nchan), dim3(64), 0, s, arena, d_chans);
  return hipGetLastError();
}

// ------------------------------------------------------------------------
// launch wrappers
// LDS of the three stages, in complex elements: two ping-pong buffers of
// nb padded sequences (the padded stride of either element size) and the
// twiddle table
static size_t seq_buf(int n, int nb) {
  return (size_t)nb * std::max(seq_stride(n, nb, sizeof(double2)), seq_stride(n, nb, sizeof(float2)));
}
static size_t lds_zinv(int E, int Ng2, int nxh, int ip) {
  const int ti = tile_of8(E, Ng2, nxh);
  return (ip && zinv_inplace(Ng2, ti, ip == 2 ? 512 : 256) ? 1 : 2) * seq_buf(Ng2, ti) + Ng2;
}
static size_t lds_xprod(int E, int Ng1, int Ng2, bool ip, int nt) {
  const int tz = tile_of(E, Ng1, Ng2);
  return (ip && xprod_inplace(Ng1, tz, nt) ? 1 : 2) * seq_buf(Ng1, 2 * tz) + Ng1;
}
static size_t lds_zfwd(int E, int Ng2, int nxh, bool ip) {
  const int ti = tile_of8(E, Ng2, nxh);
  return (ip && zfwd_inplace(Ng2, ti) ? 1 : 2) * seq_buf(Ng2, 4 * ti) + Ng2;
}
// the field-by-field z-forward stage runs when it is asked for and every
// channel's tile fits, all channels sharing one compiled length or all of
// them of length 24, 48 or 96; returns the length to specialise for (0: the
// mixed form), or -1 when it does not run
int zfwd_fs_len(const ShapeMax &m) {
  if (!m.zfwd_fs || m.dims.empty()) return -1;
  const int nt = m.zfwd_fs == 2 ? 512 : 256;
  bool same = true, small = true;
  for (const ChanDims &d : m.dims) {
    if (!zfwd_fs_fits(d.Ng2, d.N2, tile_of8(m.e_zfwd, d.Ng2, d.N1 / 2 + 1), nt)) return -1;
    same = same && d.Ng2 == m.dims[0].Ng2;
    small = small && d.Ng2 <= 96;
  }
  return same ? m.dims[0].Ng2 : small ? 0 : -1;
}
bool zfwd_fs_on(const ShapeMax &m) { return zfwd_fs_len(m) >= 0; }

// tile counts and LDS sizes for the current targets, over every channel
void shape_tiles(ShapeMax &m) {
  m.zinv_tiles = m.xprod_tiles = m.zfwd_tiles = 0;
  m.lds_zinv = m.lds_xprod = m.lds_zfwd = 0;
  for (const ChanDims &d : m.dims) {
    const int nxh = d.N1 / 2 + 1;
    const int ti = tile_of8(m.e_zinv, d.Ng2, nxh), tx = tile_of(m.e_xprod, d.Ng1, d.Ng2),
              tf = tile_of8(m.e_zfwd, d.Ng2, nxh);
    m.zinv_tiles = max(m.zinv_tiles, (nxh + ti - 1) / ti);
    m.xprod_tiles = max(m.xprod_tiles, (d.Ng2 + tx - 1) / tx);
    m.zfwd_tiles = max(m.zfwd_tiles, (nxh + tf - 1) / tf);
    m.lds_zinv = std::max(m.lds_zinv, lds_zinv(m.e_zinv, d.Ng2, nxh, m.zinv_ip));
    m.lds_xprod = std::max(m.lds_xprod, lds_xprod(m.e_xprod, d.Ng1, d.Ng2, m.xprod_ip != 0, m.xprod_threads));
    m.lds_zfwd = std::max(m.lds_zfwd, lds_zfwd(m.e_zfwd, d.Ng2, nxh, m.zfwd_ip != 0));
  }
  if (zfwd_fs_on(m)) {
    m.lds_zfwd = 0;
    for (const ChanDims &d : m.dims)
      m.lds_zfwd = std::max(m.lds_zfwd, seq_buf(d.Ng2, tile_of8(m.e_zfwd, d.Ng2, d.N1 / 2 + 1)) + d.Ng2);
  }
}

// The length the host specialises a stage for: the length every channel of
// the launch shares, if it is one of the specialised lengths and (in-place
// forms) every channel's tile runs in place; else 0 (the multi-length form).
static int nl_length(const ShapeMax &m, int stage) {
  int n = -1;
  for (const ChanDims &d : m.dims) {
    const int nxh = d.N1 / 2 + 1;
    int len = stage == 1 ? d.Ng1 : d.Ng2;
    if (stage == 1 && m.xprod_ip && !xprod_inplace(len, tile_of(m.e_xprod, d.Ng1, d.Ng2), m.xprod_threads)) len = 0;
    if (stage == 0 && m.zinv_ip && !zinv_inplace(len, tile_of8(m.e_zinv, d.Ng2, nxh), m.zinv_ip == 2 ? 512 : 256))
      len = 0;
    if (stage == 2 && m.zfwd_ip && !zfwd_inplace(len, tile_of8(m.e_zfwd, d.Ng2, nxh))) len = 0;
    if (n >= 0 && len != n) return 0;
    n = len;
  }
  return (n == 24 || n == 48 || n == 96 || n == 192 || n == 384) ? n : 0;
}

template <typename T, int NL>
static void zinv_launch(dim3 g, dim3 b, size_t lds, hipStream_t s, char *arena, const Chan *d_chans, const ShapeMax &m,
                        int nz, int extra) {
  if (m.zinv_ip == 2)
    hipLaunchKernelGGL((k_zinv<T, true, NL, 512>), g, dim3(512), lds, s, arena, d_chans, m.e_zinv, m.zinv_tiles,
                       m.nslot, nz, extra);
  else if (m.zinv_ip)
    hipLaunchKernelGGL((k_zinv<T, true, NL>), g, b, lds, s, arena, d_chans, m.e_zinv, m.zinv_tiles, m.nslot, nz, extra);
  else
    hipLaunchKernelGGL((k_zinv<T, false, NL>), g, b, lds, s, arena, d_chans, m.e_zinv, m.zinv_tiles, m.nslot, nz,
                       extra);
}
template <typename T, bool F, bool I, int NL>
static void xprod_launch(dim3 g, dim3 b, size_t lds, hipStream_t s, char *arena, const Chan *d_chans,
                         const ShapeMax &m) {
  hipLaunchKernelGGL((k_xprod<T, F, I, NL>), g, b, lds, s, arena, d_chans, m.e_xprod);
}
template <typename T, int NL>
static void zfwd_fs_launch(dim3 g, dim3 b, size_t lds, hipStream_t s, char *arena, const Chan *d_chans,
                           const ShapeMax &m, int nz, int dtblk) {
  if (m.zfwd_fs == 2)
    hipLaunchKernelGGL((k_zfwd_fs<T, NL, 512>), g, dim3(512), lds, s, arena, d_chans, m.e_zfwd, m.zfwd_tiles, m.nslot,
                       nz, dtblk ? 1 : 0, m.xprod_tiles, m.e_xprod);
  else
    hipLaunchKernelGGL((k_zfwd_fs<T, NL, 256>), g, dim3(256), lds, s, arena, d_chans, m.e_zfwd, m.zfwd_tiles, m.nslot,
                       nz, dtblk ? 1 : 0, m.xprod_tiles, m.e_xprod);
}
template <typename T, bool I, int NL>
static void zfwd_launch(dim3 g, dim3 b, size_t lds, hipStream_t s, char *arena, const Chan *d_chans, const ShapeMax &m,
                        int nz, int dtblk) {
  hipLaunchKernelGGL((k_zfwd<T, I, NL>), g, b, lds, s, arena, d_chans, m.e_zfwd, m.zfwd_tiles, m.nslot, nz,
                     dtblk ? 1 : 0, m.xprod_tiles, m.e_xprod);
}
#define LH_BY_LEN(nl, call) \
  switch (nl) {             \
    case 24: call(24); break;   \
    case 48: call(48); break;   \
    case 192: call(192); break; \
    case 96: call(96); break;   \
    case 384: call(384); break; \
    default: call(0); break;    \
  }

hipError_t launch_nl_stage(char *arena, const Chan *d_chans, const ShapeMax &m, bool mixed, int stage, hipStream_t s,
                           int kRK, int ghost, int dtblk) {
  size_t es = mixed ? sizeof(float2) : sizeof(double2);
  dim3 b(256);
  const int nl = nl_length(m, stage);
  if (stage == 0) {
    const int nz = (int)xcd_grid(m.zinv_tiles, m.nslot);
    int extra = 0, n2 = 0;
    if (ghost)
      for (const ChanDims &d : m.dims) {
        extra = std::max(extra, 2 * (d.N1 / 2 + 1));
        n2 = std::max(n2, d.N2);
      }
    dim3 g(nz + extra, 1, m.nchan * 3);
    const size_t lds = std::max(m.lds_zinv * es, ghost ? tex_1d_lds(n2) : (size_t)0);
    // FP64 in place at a specialised length reads its twiddles from HBM:
    // every channel has Ng2 = nl, and lds_zinv holds nl elements for them
    const size_t lds64 =
        std::max((m.lds_zinv - (m.zinv_ip && nl > 0 ? nl : 0)) * es, ghost ? tex_1d_lds(n2) : (size_t)0);
#define LH_ZI(n)                                                                      \
  if (mixed) zinv_launch<float2, n>(g, b, lds, s, arena, d_chans, m, nz, extra);     \
  else zinv_launch<double2, n>(g, b, lds64, s, arena, d_chans, m, nz, extra)
    bool small = nl == 0;  // a mix of the lengths 24, 48, 96 only
    for (const ChanDims &d : m.dims) small = small && (d.Ng2 == 24 || d.Ng2 == 48 || d.Ng2 == 96);
    if (small) {
      LH_ZI(-1);
    } else {
      LH_BY_LEN(nl, LH_ZI)
    }
#undef LH_ZI
  } else if (stage == 1) {
    dim3 g(m.xprod_tiles, m.nslot, m.nchan);
    const bool cfl = kRK == 1;
    const bool ip = m.xprod_ip != 0;
    const dim3 bx(m.xprod_threads);
    const size_t lds = m.lds_xprod * es;
#define LH_XP(n)                                                                                       \
  if (mixed) {                                                                                         \
    if (cfl) {                                                                                         \
      if (ip) xprod_launch<float2, true, true, n>(g, bx, lds, s, arena, d_chans, m);                   \
      else xprod_launch<float2, true, false, n>(g, bx, lds, s, arena, d_chans, m);                     \
    } else {                                                                                           \
      if (ip) xprod_launch<float2, false, true, n>(g, bx, lds, s, arena, d_chans, m);                  \
      else xprod_launch<float2, false, false, n>(g, bx, lds, s, arena, d_chans, m);                    \
    }                                                                                                  \
  } else {                                                                                             \
    if (cfl) {                                                                                         \
      if (ip) xprod_launch<double2, true, true, n>(g, bx, lds, s, arena, d_chans, m);                  \
      else xprod_launch<double2, true, false, n>(g, bx, lds, s, arena, d_chans, m);                    \
    } else {                                                                                           \
      if (ip) xprod_launch<double2, false, true, n>(g, bx, lds, s, arena, d_chans, m);                 \
      else xprod_launch<double2, false, false, n>(g, bx, lds, s, arena, d_chans, m);                   \
    }                                                                                                  \
  }
    LH_BY_LEN(nl, LH_XP)
#undef LH_XP
  } else {
    const int nz = (int)xcd_grid(m.zfwd_tiles, m.nslot);
    dim3 g(nz + (dtblk ? 1 : 0), 1, m.nchan);
    const size_t lds = std::max(m.lds_zfwd * es, (size_t)48 * sizeof(double));
    if (zfwd_fs_on(m)) {
#define LH_ZFS(n)                                                                               \
  case n:                                                                                       \
    if (mixed) zfwd_fs_launch<float2, n>(g, b, lds, s, arena, d_chans, m, nz, dtblk);           \
    else zfwd_fs_launch<double2, n>(g, b, lds, s, arena, d_chans, m, nz, dtblk);                \
    break;
      switch (zfwd_fs_len(m)) {
        LH_ZFS(0)
        LH_ZFS(24)
        LH_ZFS(48)
        LH_ZFS(96)
        LH_ZFS(192)
        LH_ZFS(384)
        default: return hipErrorInvalidValue;
      }
#undef LH_ZFS
      return hipGetLastError();
    }
#define LH_ZF(n)                                                                                  \
  if (mixed) {                                                                                    \
    if (m.zfwd_ip) zfwd_launch<float2, true, n>(g, b, lds, s, arena, d_chans, m, nz, dtblk);       \
    else zfwd_launch<float2, false, n>(g, b, lds, s, arena, d_chans, m, nz, dtblk);               \
  } else {                                                                                        \
    if (m.zfwd_ip) zfwd_launch<double2, true, n>(g, b, lds, s, arena, d_chans, m, nz, dtblk);      \
    else zfwd_launch<double2, false, n>(g, b, lds, s, arena, d_chans, m, nz, dtblk);              \
  }
    LH_BY_LEN(nl, LH_ZF)
#undef LH_ZF
  }
  return hipGetLastError();
}
#undef LH_BY_LEN

hipError_t launch_nonlinear(char *arena, const Chan *d_chans, const ShapeMax &m, bool mixed, hipStream_t s, int kRK,
                            int ghost, int dtblk) {
  for (int st = 0; st < 3; st++) {
    hipError_t e = launch_nl_stage(arena, d_chans, m, mixed, st, s, kRK, st == 0 ? ghost : 0, st == 2 ? dtblk : 0);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

void shape_add(ShapeMax &m, int N1, int N2, int Ng1, int Ng2, int ny) {
  int nxh = N1 / 2 + 1, ncol = nxh * N2;
  int nslot = (ny - 1) + (ny - 2);
  m.nslot = max(m.nslot, nslot);
  m.ncol = max(m.ncol, ncol);
  m.nyu = max(m.nyu, ny + 1);
  m.solve_cols = solve_cols_default(m.nyu);
  m.nphys = max(m.nphys, N1 * N2);
  m.dims.push_back(ChanDims{N1, N2, Ng1, Ng2, ny});
  shape_tiles(m);
  m.nchan++;
}

hipError_t launch_rhs(char *arena, const Chan *d_chans, const ShapeMax &m, int kRK, hipStream_t s, int band) {
  const int gx = (m.ncol + 255) / 256, gy = band ? 2 * RB_BAND : m.nyu;
  dim3 g((gx * gy + 7) / 8 * 8, 1, m.nchan);
  hipLaunchKernelGGL(k_rhs, g, dim3(256), 0, s, arena, d_chans, kRK, gx, gy, band);
  return hipGetLastError();
}

hipError_t launch_solve(char *arena, const Chan *d_chans, const ShapeMax &m, int kRK, hipStream_t s, int fuse,
                        int keep) {
  const int C = m.solve_cols;
  if (C < 0) {  // one column per wave, 8 (or, for short columns, 2)
                // columns per workgroup; with fuse the right-hand sides are
                // computed in its staging
    const int nw = solve_wave_cols(m.nyu, m.solve_waves);
    const int K = solve_wave_rows(m.nyu);
    const int ngrp = (m.ncol + nw - 1) / nw;
    dim3 g((ngrp + 7) / 8 * 8, 1, m.nchan), b(64 * nw);
    const size_t lds = solve_wave_lds(m.nyu, nw);
#define LH_SW(k, n)                                                                              \
  case k * 16 + n:                                                                               \
    if (fuse)                                                                                    \
      hipLaunchKernelGGL((k_solve_wave<k, n, true>), g, b, lds, s, arena, d_chans, kRK, keep);   \
    else                                                                                         \
      hipLaunchKernelGGL((k_solve_wave<k, n, false>), g, b, lds, s, arena, d_chans, kRK, keep);  \
    break;
    switch (K * 16 + nw) {
      LH_SW(1, 8)
      LH_SW(2, 8)
      LH_SW(3, 8)
      LH_SW(4, 8)
      LH_SW(5, 8)
      LH_SW(6, 8)
      LH_SW(1, 2)
      LH_SW(2, 2)
      LH_SW(3, 2)
      default: return hipErrorInvalidValue;
    }
#undef LH_SW
  } else if (C == 0) {
    if (fuse) {  // the one-column-per-lane solve reads r: compute it first
      hipError_t e = launch_rhs(arena, d_chans, m, kRK, s);
      if (e != hipSuccess) return e;
    }
    dim3 g((m.ncol + 63) / 64, 1, m.nchan);
    hipLaunchKernelGGL(k_solve, g, dim3(64), 0, s, arena, d_chans, kRK);
  } else {
    const int ngrp = (m.ncol + C - 1) / C;
    dim3 g((ngrp + 7) / 8 * 8, 1, m.nchan);
    hipLaunchKernelGGL(k_solve_lds, g, dim3(64 * m.solve_waves), solve_lds_bytes(m.nyu, C), s, arena, d_chans, kRK,
                       C, m.solve_scan, fuse);
  }
  return hipGetLastError();
}

// columns per workgroup of the LDS column solve: the most that keep two
// workgroups' staging within 128 KiB of LDS, at most 10 (6 lanes each)
int solve_cols_default(int nyu) {
  size_t b = solve_lds_col_bytes(nyu);
  int C = (int)((64 * 1024 - 8 * SOLVE_ROW_ARRAYS * (size_t)nyu) / b);
  if (C > 10) C = 10;
  if (C < 1 && 1 * b <= 160 * 1024) C = 1;
  return C;
}

// ------------------------------------------------------------------------
// Multi-resolution bands.  A banded channel is three channels of one
// context: the middle band (the long pencils, its N1 x N2 modes over the
// whole channel, transforming only its own band's planes) and the bottom and
// top wall bands (their finer N1 x N2 modes over their own rows, with a band
// edge on one side).  A wall band's planes hold every mode present at their
// rows (modes_to_planes_UVP with dk, FOU3D.f90:1000-1098): before the
// nonlinear stage the long pencils' rows are copied into the wall bands'
// columns of the same (kx, kz) (out = 0), and after it the nonlinear terms
// and fluxes of those columns' plane rows go back to the middle band
// (out = 1; planes_to_modes, FOU3D.f90:2295-2442,2772-2854).  In a wall band
// those columns are no pencils of the band: they carry the copies, and
// nothing reads the values its column phases leave in them.  p is copied
// too, so that they stay bounded.
__global__ void __launch_bounds__(256) k_band_copy(char *arena, const Chan *chans, const BandLink *links, int out) {
  const BandLink L = links[blockIdx.z >> 1];
  const bool top = blockIdx.z & 1;
  const Chan M = chans[L.cm], W = chans[top ? L.ct : L.cb];
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= M.ncol) return;
  const int r0 = top ? L.r0t : 0;
  const int i = c % M.nxh, km = c / M.nxh, m = km < M.N2 / 2 ? km : km - M.N2;
  const int cw = i + W.nxh * (m < 0 ? m + W.N2 : m);
  const View vm(arena, M), vw(arena, W);
  for (int r = blockIdx.y; r < W.nyu; r += gridDim.y) {
    const size_t om = (size_t)(r0 + r) * M.ldc + c, ow = (size_t)r * W.ldc + cw;
    if (!out) {
      vw.u1[ow] = vm.u1[om];
      vw.u3[ow] = vm.u3[om];
      vw.p[ow] = vm.p[om];
      if (r < W.nyv) vw.u2[ow] = vm.u2[om];
    } else {
      if (r >= 1 && r <= W.nyu - 2) {
        vm.nl1[om] = vw.nl1[ow];
        vm.nl3[om] = vw.nl3[ow];
        vm.fvv[om] = vw.fvv[ow];
      }
      if (r >= 1 && r <= W.nyv - 2) {
        vm.nl2[om] = vw.nl2[ow];
        vm.fuv[om] = vw.fuv[ow];
        vm.fwv[om] = vw.fwv[ow];
      }
    }
  }
}

hipError_t launch_band_copy(char *arena, const Chan *d_chans, const BandLink *d_links, int nlinks, int ncol_mid,
                            int out, hipStream_t s) {
  if (nlinks <= 0) return hipSuccess;
  dim3 g((ncol_mid + 255) / 256, 32, 2 * nlinks);
  hipLaunchKernelGGL(k_band_copy, g, dim3(256), 0, s, arena, d_chans, d_links, out);
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
