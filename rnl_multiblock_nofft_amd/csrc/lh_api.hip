// lh_api.hip -- host side of liblhgpu: setup that start() performs in the
// reference (ygrid, def_k, grid weights), device allocation, field
// transfers, the substep/step drivers (HIP-graph replayed) and the C-ABI
// declared in include/lhgpu.h.
#include <algorithm>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <mutex>
#include <string>
#include <vector>

#include "../../include/lhgpu.h"
#include "lh_closure.h"
#include "lh_common.h"
#include "lh_kernels.h"
#include "lh_tex1d.h"

using namespace lh;

static thread_local std::string g_err;
static int set_err(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
static int set_err(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}
#define HIPCHK(x)                                                                          \
  do {                                                                                     \
    hipError_t e_ = (x);                                                                   \
    if (e_ != hipSuccess) return set_err(LH_EHIP, "%s: %s", #x, hipGetErrorString(e_));    \
  } while (0)

extern "C" const char *lh_last_error(void) { return g_err.c_str(); }

namespace {

// host copy of one channel's setup
struct HostChan {
  lh_config cfg;
  int nyu, nyv, nxh, ncol, ldc, nslot;
  std::vector<double> yu, yv, dthdyu, dthdyv, dyu2i, dyv2i, kx, kz, k2x, k2z;
  std::vector<double2> twz, twx;
  double gw1, gw2, gwbc1, gwbc3, gwi1, gwi2, dtv;
  Chan dev;  // offsets into the context arena
  int iter = 0;
  int xinv = 0;  // every textured wall's mask is independent of x
  int band = -1;  // in a banded channel: 0 bottom wall band, 1 middle band, 2 top wall band
};

double ipow(double x, int n) {
  double r = 1.0;
  for (int i = 0; i < n; i++) r *= x;
  return r;
}

// y(theta), dtheta/dy of the wall-normal mapping (ygrid, start.f90:693-1077;
// 'Wall' grid without immersed region).  grid_kind 1 is a tanh stretching.
void ymap(const lh_config &c, double aaa, double bbb, double qqq, double th, double *y, double *dthdy) {
  if (c.grid_kind == 1) {
    double g = c.gamma, tg = tanh(g);
    double s = g * (th / qqq - 1.0);
    double ts = tanh(s);
    *y = ts / tg;
    *dthdy = 1.0 / ((g / (qqq * tg)) * (1.0 - ts * ts));
  } else {
    int p = c.ppp;
    double x = th - qqq;
    *y = aaa * x + bbb * ipow(x, p);
    *dthdy = 1.0 / (aaa + p * bbb * ipow(x, p - 1));
  }
}

// dyu2i/dyv2i (start.f90:943-987), grid weights (start.f90:537-559), dtv
void derive_grid(HostChan &h) {
  const int nn = h.nyv - 1;
  const std::vector<double> &yu = h.yu, &yv = h.yv;
  std::vector<double> &a = h.dyu2i, &b = h.dyv2i;
  a.assign(3 * h.nyu, 0.0);
  b.assign(3 * h.nyv, 0.0);
  a[1] = -2.0 / ((yu[1] - yu[0]) * (yu[1] - yu[0]));
  a[2] = 1.0 / ((yu[1] - yu[0]) * (yu[1] - yu[0]));
  b[1] = -2.0 / ((yv[1] - yv[0]) * (yv[1] - yv[0]));
  b[2] = 1.0 / ((yv[1] - yv[0]) * (yv[1] - yv[0]));
  for (int j = 1; j <= nn - 1; j++) {
    a[3 * j + 0] = 2.0 / ((yu[j] - yu[j - 1]) * (yu[j + 1] - yu[j - 1]));
    a[3 * j + 1] = -2.0 / ((yu[j + 1] - yu[j]) * (yu[j] - yu[j - 1]));
    a[3 * j + 2] = 2.0 / ((yu[j + 1] - yu[j]) * (yu[j + 1] - yu[j - 1]));
    b[3 * j + 0] = 2.0 / ((yv[j] - yv[j - 1]) * (yv[j + 1] - yv[j - 1]));
    b[3 * j + 1] = -2.0 / ((yv[j + 1] - yv[j]) * (yv[j] - yv[j - 1]));
    b[3 * j + 2] = 2.0 / ((yv[j + 1] - yv[j]) * (yv[j + 1] - yv[j - 1]));
  }
  {
    int j = nn;
    a[3 * j + 0] = 2.0 / ((yu[j] - yu[j - 1]) * (yu[j + 1] - yu[j - 1]));
    a[3 * j + 1] = -2.0 / ((yu[j + 1] - yu[j]) * (yu[j] - yu[j - 1]));
    a[3 * j + 2] = 2.0 / ((yu[j + 1] - yu[j]) * (yu[j + 1] - yu[j - 1]));
    double d = (yv[j] - yv[j - 1]) * (yv[j] - yv[j - 1]);
    b[3 * j + 0] = 1.0 / d;
    b[3 * j + 1] = -2.0 / d;
  }
  {
    int j = nn + 1;
    double d = (yu[j] - yu[j - 1]) * (yu[j] - yu[j - 1]);
    a[3 * j + 0] = 1.0 / d;
    a[3 * j + 1] = -2.0 / d;
  }
  const int ju = h.nyu - 1, jv = h.nyv - 1;
  const double ddv = 1.0;
  h.gw1 = -(yv[0] - yu[0]) / (yv[0] - yu[1]);
  h.gw2 = (yu[ju] - yv[jv]) / (yv[jv] - yu[ju - 1]);
  const double b1 = h.cfg.bslpu1, b3 = h.cfg.bslpu3, dv0 = h.dthdyv[0];
  h.gwbc1 = ((yu[0] - yv[0]) + b1 * ddv * dv0 * (yu[1] - yu[0])) / ((yu[1] - yv[0]) + b1 * ddv * dv0 * (yu[1] - yu[0]));
  h.gwbc3 = ((yu[0] - yv[0]) + b3 * ddv * dv0 * (yu[1] - yu[0])) / ((yu[1] - yv[0]) + b3 * ddv * dv0 * (yu[1] - yu[0]));
  h.gwi1 = (yu[0] - yu[1]) / (yv[0] - yu[1]);
  h.gwi2 = (yu[ju] - yu[ju - 1]) / (yv[jv] - yu[ju - 1]);
  h.dtv = h.cfg.Re / (-h.k2x[h.cfg.N1 / 2] - h.k2z[h.cfg.N2 / 2 - 1] + 4.0 / ((yv[1] - yv[0]) * (yv[1] - yv[0])));
}

void twiddles(std::vector<double2> &tw, int n) {
  tw.resize(n);
  for (int m = 0; m < n; m++) {
    long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)m / (long double)n;
    tw[m] = make_double2((double)cosl(a), (double)sinl(a));
  }
}

int valid_len(int n) {
  if (n < 2) return 0;
  while (n % 3 == 0) n /= 3;
  while (n % 2 == 0) n /= 2;
  return n == 1;
}

int setup_host(HostChan &h, const lh_config &c) {
  if (c.N1 < 2 || c.N2 < 2 || (c.N1 & 1) || (c.N2 & 1))
    return set_err(LH_EINVAL, "N1, N2 must be even and >= 2 (got %d, %d)", c.N1, c.N2);
  if (c.Ngal1 < c.N1 + 2 || c.Ngal2 < c.N2 + 2 || (c.Ngal1 & 1) || (c.Ngal2 & 1))
    return set_err(LH_EINVAL, "Ngal must be even and exceed N (got %d,%d for %d,%d)", c.Ngal1, c.Ngal2, c.N1, c.N2);
  if (2 * c.Ngal1 < 3 * c.N1 || 2 * c.Ngal2 < 3 * c.N2)
    return set_err(LH_EINVAL, "Ngal below the 3/2 rule (%d,%d for %d,%d)", c.Ngal1, c.Ngal2, c.N1, c.N2);
  if (!valid_len(c.Ngal1) || !valid_len(c.Ngal2))
    return set_err(LH_EINVAL, "Ngal must be 2^a 3^b (got %d, %d)", c.Ngal1, c.Ngal2);
  if (c.ny < 5) return set_err(LH_EINVAL, "ny must be >= 5 (got %d)", c.ny);
  if (!(c.dt > 0.0)) return set_err(LH_EINVAL, "dt must be positive");
  if (!(c.Re > 0.0)) return set_err(LH_EINVAL, "Re must be positive");
  if (c.grid_kind == 0 && (c.ppp < 2 || !(c.dyq > 0))) return set_err(LH_EINVAL, "bad polynomial grid");
  if (c.grid_kind == 1 && !(c.gamma > 0)) return set_err(LH_EINVAL, "bad tanh grid");
  if (c.grid_kind != 0 && c.grid_kind != 1) return set_err(LH_EINVAL, "grid_kind must be 0 or 1");
  h.cfg = c;
  h.nyv = c.ny;
  h.nyu = c.ny + 1;
  h.nxh = c.N1 / 2 + 1;
  h.ncol = h.nxh * c.N2;
  h.ldc = (h.ncol + 63) / 64 * 64;
  h.nslot = (h.nyu - 2) + (h.nyv - 2);
  // def_k, start.f90:1079-1136
  h.kx.resize(h.nxh);
  h.k2x.resize(h.nxh);
  h.kz.resize(c.N2);
  h.k2z.resize(c.N2);
  for (int i = 0; i < h.nxh; i++) {
    double t = c.alp * i;
    h.kx[i] = t;
    h.k2x[i] = -(t * t);
  }
  for (int k = 1; k <= c.N2 / 2; k++) {
    double t = c.bet * (k - 1);
    h.kz[k - 1] = t;
    h.k2z[k - 1] = -(t * t);
  }
  for (int k = c.N2 / 2 + 1; k <= c.N2; k++) {
    double t = c.bet * (c.N2 - k + 1);
    h.kz[k - 1] = -t;
    h.k2z[k - 1] = -(t * t);
  }
  // ygrid
  const int nn = c.ny - 1;
  double qqq = nn / 2.0, aaa = 0.0, bbb = 0.0;
  if (c.grid_kind == 0) {
    int p = c.ppp;
    double dyq = c.dyq;
    bbb = (dyq - 1.0) / (qqq * (1.0 - ipow(qqq, p - 1) + dyq * ((1.0 - qqq) * ipow(qqq, p - 1) - ipow(1.0 - qqq, p))));
    aaa = (1.0 - bbb * ipow(qqq, p)) / qqq;
  }
  h.yu.assign(h.nyu, 0.0);
  h.yv.assign(h.nyv, 0.0);
  h.dthdyu.assign(h.nyu, 0.0);
  h.dthdyv.assign(h.nyv, 0.0);
  ymap(c, aaa, bbb, qqq, -0.5, &h.yu[0], &h.dthdyu[0]);
  for (int j = 0; j <= nn; j++) {
    ymap(c, aaa, bbb, qqq, (double)j, &h.yv[j], &h.dthdyv[j]);
    ymap(c, aaa, bbb, qqq, j + 0.5, &h.yu[j + 1], &h.dthdyu[j + 1]);
  }
  derive_grid(h);
  twiddles(h.twz, c.Ngal2);
  twiddles(h.twx, c.Ngal1);
  return LH_OK;
}

template <typename T> size_t al(size_t n) { return (n * sizeof(T) + 255) / 256 * 256; }

}  // namespace

// channels whose nonlinear stages launch together: one plane shape, so
// that every stage runs its length's specialised code
struct NlGroup {
  ShapeMax m;  // the members' dimensions (the context's knobs are applied per launch)
  std::vector<int> members;
  Chan *d = nullptr;  // the members' descriptors
};

struct lh_ctx {
  std::vector<HostChan> ch;
  bool mixed = false;
  char *arena = nullptr;  // every channel's device arrays
  size_t arena_bytes = 0;
  hipStream_t stream = nullptr;
  Chan *d_chans = nullptr;
  ShapeMax shape;
  double2 *staging = nullptr;
  size_t staging_elems = 0;
  hipGraphExec_t graph = nullptr;   // one captured step (the first of an lh_step call)
  hipGraphExec_t graph2 = nullptr;  // the steps after it, when they take over the ghost pass (ghost_deferrable)
  int device = 0;
  std::vector<TexGroup> tex_groups;  // N-grid work planes per (N1, N2) shape
  int ntex = 0;                      // channels with a textured wall
  int tex_nit = 3;                   // wall-source iterations per substep
  int tex_fused = 1;                 // LDS-resident wall transforms when they fit
  int shape_groups = 1;              // launches per plane shape without bands too
  int tex_sep = 1;                   // per-x-mode wall iteration for x-independent masks
  int ntex1d = 0;                    // channels on that path (the rest: ntex - ntex1d)
  bool adaptive = false;             // a channel with an adaptive time step
  char *stats = nullptr;             // statistics accumulators (allocated on first use)
  double *hist = nullptr;            // history-record reductions, 8 doubles per channel
  std::vector<char *> tri_mem;       // per channel: triad-list tables and products (or null)
  int ntri = 0;                      // channels on the triad-list path
  size_t stats_bytes = 0;
  int istat = 0;                     // samples since the last reset
  // learned closure: the packed network (CL_FRAG_BYTES of weight fragments,
  // then CL_NBIAS biases), its normalisation, and per channel the physical
  // planes, correction planes and training-target modes (allocated on use)
  char *cl_net = nullptr;
  ClNorm cl_norm{};
  std::vector<char *> cl_mem;
  std::vector<int> cl_on;
  int ncl = 0;  // channels whose step adds the correction
  // banded channels (lh_set_bands): their links on the host and device,
  // and the largest middle band's column count
  std::vector<BandLink> bands;
  BandLink *d_bands = nullptr;
  int band_ncol = 0;
  // launch groups of the nonlinear stages (empty: one launch covers every
  // channel); a context with banded channels mixes two plane shapes, and
  // its groups keep each launch specialised
  std::vector<NlGroup> nlg;
};

namespace {

int upload_chan_constants(lh_ctx *x, HostChan &h) {
  Chan &dc = h.dev;
  View d(x->arena, dc);
  HIPCHK(hipMemcpy((void *)d.yu, h.yu.data(), h.nyu * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy((void *)d.yv, h.yv.data(), h.nyv * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy((void *)d.dthdyu, h.dthdyu.data(), h.nyu * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy((void *)d.dthdyv, h.dthdyv.data(), h.nyv * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy((void *)d.dyu2i, h.dyu2i.data(), 3 * h.nyu * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy((void *)d.dyv2i, h.dyv2i.data(), 3 * h.nyv * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy((void *)d.kx, h.kx.data(), h.nxh * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy((void *)d.k2x, h.k2x.data(), h.nxh * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy((void *)d.kz, h.kz.data(), h.cfg.N2 * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy((void *)d.k2z, h.k2z.data(), h.cfg.N2 * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy((void *)d.twz, h.twz.data(), h.cfg.Ngal2 * sizeof(double2), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy((void *)d.twx, h.twx.data(), h.cfg.Ngal1 * sizeof(double2), hipMemcpyHostToDevice));
  // unit-forcing solutions of the mean mode for each substep (flowrate_corr's
  // u11 = LUsol0(-dRK dt), littleharsh.f90:484-494) and their flow rates
  {
    const int n = h.nyu;
    std::vector<double> u11(3 * n), a(3 * n);
    double qx1[3];
    for (int kRK = 1; kRK <= 3; kRK++) {
      const double beta = rk_b(kRK) * h.cfg.dt / h.cfg.Re, bdt = -rk_d(kRK) * h.cfg.dt;
      double *x = u11.data() + (kRK - 1) * n;
      a[0] = 0.0;
      a[1] = 1.0;
      a[2] = h.gw1;
      for (int j = 1; j <= n - 2; j++) {
        a[3 * j] = -beta * h.dyu2i[3 * j];
        a[3 * j + 1] = 1.0 - beta * h.dyu2i[3 * j + 1];
        a[3 * j + 2] = -beta * h.dyu2i[3 * j + 2];
      }
      a[3 * (n - 1)] = h.gw2;
      a[3 * (n - 1) + 1] = 1.0;
      a[3 * (n - 1) + 2] = 0.0;
      a[1] = 1.0 / a[1];
      for (int j = 1; j < n; j++) {
        a[3 * (j - 1) + 2] = a[3 * (j - 1) + 2] * a[3 * (j - 1) + 1];
        a[3 * j + 1] = 1.0 / (a[3 * j + 1] - a[3 * j] * a[3 * (j - 1) + 2]);
      }
      for (int j = 0; j < n; j++) x[j] = (j == 0 || j == n - 1) ? 0.0 : bdt;
      x[0] = x[0] * a[1];
      for (int j = 1; j < n; j++) x[j] = (x[j] - a[3 * j] * x[j - 1]) * a[3 * j + 1];
      for (int j = n - 2; j >= 0; j--) x[j] = x[j] - a[3 * j + 2] * x[j + 1];
      double Q = 0.0;
      for (int j = 1; j <= n - 2; j++) Q = Q + (h.yv[j] - h.yv[j - 1]) * x[j];
      qx1[kRK - 1] = Q;
    }
    HIPCHK(hipMemcpy((void *)d.u11, u11.data(), 3 * n * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy((void *)(d.scal + SC_QX1), qx1, sizeof qx1, hipMemcpyHostToDevice));
  }
  dc.gw1 = h.gw1;
  dc.gw2 = h.gw2;
  dc.gwbc1 = h.gwbc1;
  dc.gwbc3 = h.gwbc3;
  dc.gwi1 = h.gwi1;
  dc.gwi2 = h.gwi2;
  return LH_OK;
}

// assigns h.dev's offsets from `base`, returns the end offset
size_t layout_chan(lh_ctx *x, HostChan &h, size_t base) {
  const lh_config &c = h.cfg;
  const size_t fld = (size_t)(h.nyu + PADROWS) * h.ldc;  // complex elements per field (with padding rows)
  const int nxp = (h.nxh + 15) / 16 * 16;
  const size_t ninter = (size_t)h.nslot * c.Ngal2 * nxp;
  Chan &d = h.dev;
  memset(&d, 0, sizeof d);
  size_t off = base;
  auto take = [&](size_t b) -> size_t {
    size_t q = off;
    off += b;
    return q;
  };
  size_t *fl[18] = {&d.o_u1, &d.o_u2, &d.o_u3, &d.o_p, &d.o_N1o, &d.o_N2o, &d.o_N3o, &d.o_r1, &d.o_r2,
                    &d.o_r3, &d.o_nl1, &d.o_nl2, &d.o_nl3, &d.o_fuv, &d.o_fwv, &d.o_fvv, &d.o_psi, &d.o_div};
  for (int q = 0; q < 18; q++) *fl[q] = take(al<double2>(fld));
  d.o_scr = take(al<double>(2 * fld));
  d.o_yu = take(al<double>(h.nyu));
  d.o_dthdyu = take(al<double>(h.nyu));
  d.o_yv = take(al<double>(h.nyv));
  d.o_dthdyv = take(al<double>(h.nyv));
  d.o_dyu2i = take(al<double>(3 * h.nyu));
  d.o_dyv2i = take(al<double>(3 * h.nyv));
  d.o_u11 = take(al<double>(3 * h.nyu));
  d.o_kx = take(al<double>(h.nxh));
  d.o_k2x = take(al<double>(h.nxh));
  d.o_kz = take(al<double>(c.N2));
  d.o_k2z = take(al<double>(c.N2));
  d.o_twz = take(al<double2>(c.Ngal2));
  d.o_twx = take(al<double2>(c.Ngal1));
  d.o_scal = take(al<double>(SC_NSCAL));
  d.o_fac = take(al<double>(14 * fld));
  d.o_mask = take(al<int>(2 * (size_t)c.N1 * c.N2));
  d.o_iscal = take(al<int>(4));
  d.o_wtab = take(al<double>(6 * (size_t)h.nyu * h.ldc));
  d.o_texu = take(al<double2>(4 * (size_t)h.ncol));
  int ng2max = 0;
  for (const ChanDims &q : x->shape.dims) ng2max = std::max(ng2max, q.Ng2);
  d.o_cfl = take(al<double>(2 * (size_t)h.nslot * ng2max));
  d.o_pmean = take(al<double>(2 * (size_t)h.nslot));
  if (x->mixed) {
    d.o_hzf = take(al<float2>(3 * ninter));
    d.o_gxf = take(al<float2>(4 * ninter));
    d.o_hz = d.o_hzf;
    d.o_gx = d.o_gxf;
  } else {
    d.o_hz = take(al<double2>(3 * ninter));
    d.o_gx = take(al<double2>(4 * ninter));
    d.o_hzf = d.o_hz;
    d.o_gxf = d.o_gx;
  }
  d.N1 = c.N1;
  d.N2 = c.N2;
  d.Ng1 = c.Ngal1;
  d.Ng2 = c.Ngal2;
  d.ny = c.ny;
  d.nyu = h.nyu;
  d.nyv = h.nyv;
  d.nxh = h.nxh;
  d.nxp = nxp;
  d.ncol = h.ncol;
  d.ldc = h.ldc;
  d.nplu = h.nyu - 2;
  d.nplv = h.nyv - 2;
  d.nslot = h.nslot;
  d.pu0 = d.pv0 = 1;
  d.band_edge = 0;
  d.flag_ctpress = c.flag_ctpress;
  d.adaptive = c.adaptive_dt != 0;
  d.CFL = c.CFL;
  d.Re = c.Re;
  d.alp = c.alp;
  d.bet = c.bet;
  return off;
}

int init_chan(lh_ctx *x, HostChan &h) {
  View d(x->arena, h.dev);
  std::vector<int> ns(2 * (size_t)h.cfg.N1 * h.cfg.N2, 1);
  HIPCHK(hipMemcpy(d.mask, ns.data(), ns.size() * sizeof(int), hipMemcpyHostToDevice));
  double sc[SC_NSCAL] = {0};
  sc[SC_MPGX] = h.cfg.mpgx;
  sc[SC_DT] = h.cfg.dt;
  sc[SC_DTV] = h.dtv;
  HIPCHK(hipMemcpy(d.scal, sc, sizeof sc, hipMemcpyHostToDevice));
  return upload_chan_constants(x, h);
}

int push_descriptors(lh_ctx *x) {
  std::vector<Chan> v;
  for (auto &h : x->ch) v.push_back(h.dev);
  HIPCHK(hipMemcpy(x->d_chans, v.data(), v.size() * sizeof(Chan), hipMemcpyHostToDevice));
  for (auto &g : x->nlg) {
    std::vector<Chan> gv;
    for (int c : g.members) gv.push_back(x->ch[c].dev);
    HIPCHK(hipMemcpy(g.d, gv.data(), gv.size() * sizeof(Chan), hipMemcpyHostToDevice));
  }
  return LH_OK;
}

// a group's launch shape with the knobs of `k`
ShapeMax group_shape(const NlGroup &g, const ShapeMax &k) {
  ShapeMax m = g.m;
  m.e_zinv = k.e_zinv;
  m.e_xprod = k.e_xprod;
  m.e_zfwd = k.e_zfwd;
  m.xprod_ip = k.xprod_ip;
  m.zfwd_ip = k.zfwd_ip;
  m.zinv_ip = k.zinv_ip;
  m.zfwd_fs = k.zfwd_fs;
  m.xprod_threads = k.xprod_threads;
  m.ghost_zinv = k.ghost_zinv;
  shape_tiles(m);
  return m;
}

// one nonlinear stage over every channel, with the knobs of `k`
hipError_t nl_stage(lh_ctx *x, const ShapeMax &k, int st, int kRK, int ghost = 0, int dtblk = 0) {
  if (x->nlg.empty()) return launch_nl_stage(x->arena, x->d_chans, k, x->mixed, st, x->stream, kRK, ghost, dtblk);
  for (auto &g : x->nlg) {
    hipError_t e = launch_nl_stage(x->arena, g.d, group_shape(g, k), x->mixed, st, x->stream, kRK, ghost, dtblk);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// the right-hand sides and the column solve over every channel, with the
// knobs of `k` (per group where there are groups: their grids then fit
// each group's columns and rows)
ShapeMax col_shape(const NlGroup &g, const ShapeMax &k) {
  ShapeMax m = group_shape(g, k);
  m.solve_cols = k.solve_cols;
  m.solve_scan = k.solve_scan;
  m.solve_waves = k.solve_waves;
  m.rhs_fused = k.rhs_fused;
  return m;
}
hipError_t rhs_stage(lh_ctx *x, const ShapeMax &k, int kRK, int band = 0) {
  if (x->nlg.empty()) return launch_rhs(x->arena, x->d_chans, k, kRK, x->stream, band);
  for (auto &g : x->nlg) {
    hipError_t e = launch_rhs(x->arena, g.d, col_shape(g, k), kRK, x->stream, band);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}
hipError_t solve_stage(lh_ctx *x, const ShapeMax &k, int kRK, int fuse = 0, int keep = 1) {
  if (x->nlg.empty()) return launch_solve(x->arena, x->d_chans, k, kRK, x->stream, fuse, keep);
  for (auto &g : x->nlg) {
    hipError_t e = launch_solve(x->arena, g.d, col_shape(g, k), kRK, x->stream, fuse, keep);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// the field-by-field z-forward form is on for some launch of these knobs
bool nl_fs_on(lh_ctx *x, const ShapeMax &k) {
  if (x->nlg.empty()) return zfwd_fs_on(k);
  for (auto &g : x->nlg)
    if (zfwd_fs_on(group_shape(g, k))) return true;
  return false;
}

// groups channels of equal plane shape (contexts with banded channels)
int build_nl_groups(lh_ctx *x) {
  for (auto &g : x->nlg)
    if (g.d) (void)hipFree(g.d);
  x->nlg.clear();
  if (x->bands.empty()) {
    // without bands: only when asked and the channels have several shapes
    bool several = false;
    for (auto &h : x->ch)
      several = several || h.cfg.Ngal1 != x->ch[0].cfg.Ngal1 || h.cfg.Ngal2 != x->ch[0].cfg.Ngal2;
    if (!x->shape_groups || !several) return LH_OK;
  }
  for (int c = 0; c < (int)x->ch.size(); c++) {
    const lh_config &cf = x->ch[c].cfg;
    NlGroup *g = nullptr;
    for (auto &q : x->nlg)
      if (q.m.dims[0].Ng1 == cf.Ngal1 && q.m.dims[0].Ng2 == cf.Ngal2) g = &q;
    if (!g) {
      x->nlg.emplace_back();
      g = &x->nlg.back();
    }
    g->members.push_back(c);
    shape_add(g->m, cf.N1, cf.N2, cf.Ngal1, cf.Ngal2, cf.ny);
  }
  for (auto &g : x->nlg) {
    hipError_t e = hipMalloc((void **)&g.d, g.members.size() * sizeof(Chan));
    if (e != hipSuccess) {
      g.d = nullptr;
      return set_err(LH_ENOMEM, "hipMalloc(group descriptors): %s", hipGetErrorString(e));
    }
  }
  return push_descriptors(x);
}

// rows of a field in the reference layout, device array, jlo
int field_map(lh_ctx *x, HostChan &h, int field, double2 **arr, int *jlo, int *nrows) {
  View d(x->arena, h.dev);
  switch (field) {
    case LH_U1: *arr = d.u1; *jlo = 0; *nrows = h.nyu; return 0;
    case LH_U2: *arr = d.u2; *jlo = 0; *nrows = h.nyv; return 0;
    case LH_U3: *arr = d.u3; *jlo = 0; *nrows = h.nyu; return 0;
    case LH_P: *arr = d.p; *jlo = 1; *nrows = h.nyu - 2; return 0;
    case LH_NU1: *arr = d.N1o; *jlo = 1; *nrows = h.nyu - 2; return 0;
    case LH_NU2: *arr = d.N2o; *jlo = 1; *nrows = h.nyv - 2; return 0;
    case LH_NU3: *arr = d.N3o; *jlo = 1; *nrows = h.nyu - 2; return 0;
    case LH_R1: *arr = d.r1; *jlo = 0; *nrows = h.nyu; return 0;
    case LH_R2: *arr = d.r2; *jlo = 0; *nrows = h.nyv; return 0;
    case LH_R3: *arr = d.r3; *jlo = 0; *nrows = h.nyu; return 0;
    case LH_PSI: *arr = d.psi; *jlo = 1; *nrows = h.nyu - 2; return 0;
    case LH_DIV: *arr = d.div; *jlo = 1; *nrows = h.nyu - 2; return 0;
  }
  return -1;
}

int check_ctx(lh_ctx *x) {
  if (!x) return set_err(LH_EINVAL, "null context");
  HIPCHK(hipSetDevice(x->device));
  return LH_OK;
}

// textured walls: the wall sources before the column solves (part 0) and
// the masked ghost rows after them (part 1); each field's wall planes stay in
// one workgroup's LDS when they fit
int tex_part(lh_ctx *x, int kRK, int part, bool skip1d = false) {
  if (!x->ntex) return LH_OK;
  const size_t lds = tex_fused_bytes(x->tex_groups);
  const bool fused = x->tex_fused && lds <= 160 * 1024;
  const bool any2d = x->ntex > x->ntex1d;
  if (part == 0) {
    HIPCHK(launch_tex_adj(x->arena, x->d_chans, x->shape, kRK, x->stream));
    if (any2d) {
      if (fused)
        HIPCHK(launch_tex_sources_fused(x->arena, x->d_chans, x->shape, lds, kRK, x->tex_nit, x->stream));
      else
        HIPCHK(launch_tex_sources(x->arena, x->d_chans, x->shape, x->tex_groups, kRK, x->tex_nit, x->stream));
    }
    if (x->ntex1d) HIPCHK(launch_tex_1d(x->arena, x->d_chans, x->shape, kRK, x->tex_nit, 0, x->stream));
  } else {
    if (any2d) {
      if (fused)
        HIPCHK(launch_tex_ghosts_fused(x->arena, x->d_chans, x->shape, lds, x->stream));
      else
        HIPCHK(launch_tex_ghosts(x->arena, x->d_chans, x->shape, x->tex_groups, x->stream));
    }
    if (x->ntex1d && !skip1d) HIPCHK(launch_tex_1d(x->arena, x->d_chans, x->shape, 1, 1, 1, x->stream));
  }
  return LH_OK;
}

// which textured channels take the per-x-mode path (all their textured
// walls x-independent, and the path enabled)
void tex_classify(lh_ctx *x) {
  x->ntex = x->ntex1d = 0;
  for (auto &h : x->ch) {
    h.dev.tex1d = (h.dev.tex && x->tex_sep && h.xinv) ? 1 : 0;
    x->ntex += h.dev.tex != 0;
    x->ntex1d += h.dev.tex1d;
  }
}

// column solves of a substep (with the texture parts around them)
int solve_phase(lh_ctx *x, int kRK, int fuse = 0, bool defer_ghost = false, int keep = 1) {
  int rc = tex_part(x, kRK, 0);
  if (rc) return rc;
  HIPCHK(solve_stage(x, x->shape, kRK, fuse, keep));
  return tex_part(x, kRK, 1, defer_ghost);
}

// whether a captured step may leave the per-x-mode ghost pass of substeps 1
// and 2 to extra workgroups of the next substep's z-inverse launch (the
// nonlinear stage reads no ghost rows; the right-hand side, after it, does)
bool ghost_deferrable(const lh_ctx *x) {
  if (!x->shape.ghost_zinv || !x->ntex1d || x->ntri >= (int)x->ch.size()) return false;
  for (const auto &h : x->ch)
    if (h.dev.tex1d && !tex_ghost_small(h.cfg.N2)) return false;
  return true;
}

// first substep: dtc and (adaptive channels) the new dt, then everything
// that depends on dt -- pivots, mean-mode unit-forcing solutions, texture
// inverse rows -- for the whole step
// the nonlinear products of every channel: the FFT stages for channels
// without triad lists, the list kernels for those with
int raw_nonlinear(lh_ctx *x, int kRK, int ghost = 0, int dtblk = 0) {
  const int nb = (int)x->bands.size();
  HIPCHK(launch_band_copy(x->arena, x->d_chans, x->d_bands, nb, x->band_ncol, 0, x->stream));
  if (x->ntri < (int)x->ch.size())
    for (int st = 0; st < 3; st++) HIPCHK(nl_stage(x, x->shape, st, kRK, st == 0 ? ghost : 0, st == 2 ? dtblk : 0));
  HIPCHK(launch_band_copy(x->arena, x->d_chans, x->d_bands, nb, x->band_ncol, 1, x->stream));
  if (x->ntri) HIPCHK(launch_triads(x->arena, x->d_chans, x->shape, x->stream));
  return LH_OK;
}

// per-channel closure buffers: phys and corr planes (float [3][nyu][N2][N1]),
// training-target modes and the held closure's correction modes (each
// double2 [3][nyu][ldc])
size_t cl_plane_floats(const HostChan &h) { return (size_t)3 * h.nyu * h.cfg.N1 * h.cfg.N2; }
int cl_buffers(lh_ctx *x, int c, float **phys, float **corr, double2 **tau, double2 **held = nullptr) {
  HostChan &h = x->ch[c];
  const size_t pf = al<float>(cl_plane_floats(h));
  const size_t mf = al<double2>((size_t)3 * h.nyu * h.ldc);
  if (!x->cl_mem[c]) {
    const size_t bytes = 2 * pf + 2 * mf;
    hipError_t e = hipMalloc((void **)&x->cl_mem[c], bytes);
    if (e != hipSuccess) {
      x->cl_mem[c] = nullptr;
      return set_err(LH_ENOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    }
    HIPCHK(hipMemset(x->cl_mem[c], 0, bytes));
  }
  *phys = (float *)x->cl_mem[c];
  *corr = (float *)(x->cl_mem[c] + pf);
  *tau = (double2 *)(x->cl_mem[c] + 2 * pf);
  if (held) *held = (double2 *)(x->cl_mem[c] + 2 * pf + mf);
  return LH_OK;
}

// the network's correction of the nonlinear term of every channel that has
// it switched on (after the products, before the right-hand side).  Mode 2
// evaluates the network at the first substep of a step only and adds the
// same correction planes at the other two.
int closure_stage(lh_ctx *x, int kRK) {
  for (size_t c = 0; c < x->ch.size(); c++) {
    if (!x->cl_on[c]) continue;
    HostChan &h = x->ch[c];
    float *phys, *corr;
    double2 *tau, *held;
    int rc = cl_buffers(x, (int)c, &phys, &corr, &tau, &held);
    if (rc) return rc;
    View d(x->arena, h.dev);
    ClSrc src{{d.u1, d.u2, d.u3}, {h.nyu, h.nyv, h.nyu}};
    if (x->cl_on[c] == 2 && kRK != 1) {  // held: the modes stored at kRK = 1
      HIPCHK(launch_cl_add_modes(x->arena, x->d_chans + c, h.dev, held, x->stream));
      continue;
    }
    HIPCHK(launch_cl_phys(x->arena, x->d_chans + c, h.dev, src, phys, x->stream));
    HIPCHK(launch_cl_infer(x->cl_net, (const float *)(x->cl_net + CL_FRAG_BYTES), x->cl_norm, phys, corr, h.cfg.N1,
                           h.cfg.N2, h.nyu, h.nyv, x->stream));
    HIPCHK(launch_cl_to_modes(x->arena, x->d_chans + c, h.dev, corr, x->stream, x->cl_on[c] == 2 ? held : nullptr));
  }
  return LH_OK;
}

int nonlinear_stage(lh_ctx *x, int kRK, int ghost = 0, int dtblk = 0) {
  int rc = raw_nonlinear(x, kRK, ghost, dtblk);
  if (rc || !x->ncl) return rc;
  return closure_stage(x, kRK);
}

// dt_done: the z-forward launch has made the time-step update already
int time_step(lh_ctx *x, int kRK, bool dt_done = false) {
  if (kRK != 1) return LH_OK;
  if (!dt_done) HIPCHK(launch_dt_update(x->arena, x->d_chans, x->shape, x->stream));
  if (x->adaptive) {
    HIPCHK(launch_factors(x->arena, x->d_chans, x->shape, x->stream));
    HIPCHK(launch_u11(x->arena, x->d_chans, x->shape, x->stream));
    if (x->ntex) HIPCHK(launch_tex_tables(x->arena, x->d_chans, x->shape, x->stream));
  }
  return LH_OK;
}

// in_step: one of the three substeps of a captured step, which may hand
// substeps 1 and 2's ghost pass to the next substep (ghost_deferrable);
// with chained steps, the third substep's pass goes to the next step's first
// substep as well (prev_ghost: this step's first substep takes it; keep_last:
// leave this step's third pass to the next step)
// whether the column solve computes the right-hand sides (LH_TUNE_RHS_FUSED):
// any solve form without textured walls, the wave form with them
static int fused_rhs(const lh_ctx *x, const ShapeMax &m) {
  return m.rhs_fused && (!x->ntex || m.solve_cols < 0);
}
int run_substep(lh_ctx *x, int kRK, bool in_step = false, bool prev_ghost = false, bool keep_last = false) {
  const bool defer = in_step && ghost_deferrable(x);
  // the time-step update rides on the z-forward launch when there is one
  // (it reads the x-product stage's CFL maxima; the closure and triad-list
  // kernels between the two read neither dt nor t)
  const bool dt_in_zfwd = kRK == 1 && x->ntri < (int)x->ch.size();
  int rc = nonlinear_stage(x, kRK, defer && (kRK > 1 || prev_ghost), dt_in_zfwd);
  if (rc) return rc;
  rc = time_step(x, kRK, dt_in_zfwd);
  if (rc) return rc;
  // the textured-wall sources read r1 and r3 near the walls before the
  // solve: with textured walls only the one-column-per-wave solve computes
  // the right-hand sides itself, after a pass over the rows the sources read
  const int fuse = fused_rhs(x, x->shape);
  if (!fuse) HIPCHK(rhs_stage(x, x->shape, kRK));
  else if (x->ntex) HIPCHK(rhs_stage(x, x->shape, kRK, 1));
  // psi and div of substeps 1 and 2 of a step are read by nothing (lh_history
  // and downloads see the last substep's)
  return solve_phase(x, kRK, fuse, defer && (kRK < 3 || keep_last), !in_step || kRK == 3);
}

// capture one step into *out
int capture_step(lh_ctx *x, bool prev_ghost, bool keep_last, hipGraphExec_t *out) {
  hipGraph_t g;
  HIPCHK(hipStreamBeginCapture(x->stream, hipStreamCaptureModeThreadLocal));
  int r = LH_OK;
  for (int k = 1; k <= 3 && r == LH_OK; k++) r = run_substep(x, k, true, prev_ghost && k == 1, keep_last && k == 3);
  hipError_t e = hipStreamEndCapture(x->stream, &g);
  if (r) return r;
  if (e != hipSuccess) return set_err(LH_EHIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
  e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    *out = nullptr;
    return set_err(LH_EHIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
  }
  return LH_OK;
}

// Thomas factors and the textured-wall inverse rows (after a grid change)
int refresh_tables(lh_ctx *x) {
  for (auto &h : x->ch) HIPCHK(hipMemsetAsync(x->arena + h.dev.o_iscal, 0, 4 * sizeof(int), x->stream));
  HIPCHK(launch_factors(x->arena, x->d_chans, x->shape, x->stream));
  HIPCHK(launch_tex_tables(x->arena, x->d_chans, x->shape, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  return LH_OK;
}

}  // namespace

extern "C" {

int lh_ctx_create(const lh_config *cfgs, int nchan, lh_ctx **out) {
  if (!cfgs || nchan < 1 || !out) return set_err(LH_EINVAL, "lh_ctx_create: bad arguments");
  *out = nullptr;
  for (int c = 1; c < nchan; c++)
    if (cfgs[c].precision != cfgs[0].precision)
      return set_err(LH_EINVAL, "all channels of a context share one precision mode");
  if (cfgs[0].precision != LH_PREC_FP64 && cfgs[0].precision != LH_PREC_MIXED)
    return set_err(LH_EINVAL, "unknown precision mode %d", cfgs[0].precision);
  lh_ctx *x = new lh_ctx();
  x->mixed = cfgs[0].precision == LH_PREC_MIXED;
  hipError_t e = hipGetDevice(&x->device);
  if (e != hipSuccess) {
    delete x;
    return set_err(LH_EHIP, "hipGetDevice: %s", hipGetErrorString(e));
  }
  x->ch.resize(nchan);
  int rc = LH_OK;
  size_t stage = 0;
  for (int c = 0; c < nchan && rc == LH_OK; c++) {
    rc = setup_host(x->ch[c], cfgs[c]);
    if (rc == LH_OK && cfgs[c].adaptive_dt) {
      if (!(cfgs[c].CFL > 0.0)) rc = set_err(LH_EINVAL, "channel %d: adaptive_dt needs CFL > 0", c);
      x->adaptive = true;
    }
    if (rc == LH_OK) {
      shape_add(x->shape, cfgs[c].N1, cfgs[c].N2, cfgs[c].Ngal1, cfgs[c].Ngal2, cfgs[c].ny);
      size_t s = (size_t)x->ch[c].ncol * x->ch[c].nyu;
      if (s > stage) stage = s;
    }
  }
  if (rc == LH_OK) {
    e = hipStreamCreateWithFlags(&x->stream, hipStreamNonBlocking);
    if (e != hipSuccess) rc = set_err(LH_EHIP, "hipStreamCreate: %s", hipGetErrorString(e));
  }
  if (rc == LH_OK) {
    size_t off = 0;
    for (int c = 0; c < nchan; c++) off = layout_chan(x, x->ch[c], off);
    // texture work planes, channels of one (N1, N2) shape side by side, and
    // the N-grid twiddles of each shape
    for (int c = 0; c < nchan; c++) {
      HostChan &h = x->ch[c];
      TexGroup *g = nullptr;
      for (auto &q : x->tex_groups)
        if (q.N1 == h.cfg.N1 && q.N2 == h.cfg.N2) g = &q;
      if (!g) {
        x->tex_groups.push_back(TexGroup{h.cfg.N1, h.cfg.N2, 0, 0, 0, 0});
        g = &x->tex_groups.back();
      }
      g->nplanes += 4;
    }
    for (auto &g : x->tex_groups) {
      g.o_twz = off;
      off += al<double2>(g.N2);
      g.o_twx = off;
      off += al<double2>(g.N1);
      g.o_texw = off;
      off += al<double2>((size_t)g.nplanes * (g.N1 / 2 + 1) * g.N2);
      int q = 0;
      for (auto &h : x->ch)
        if (h.cfg.N1 == g.N1 && h.cfg.N2 == g.N2) {
          h.dev.o_texw = g.o_texw + (size_t)4 * (q++) * h.ncol * sizeof(double2);
          h.dev.o_twzN = g.o_twz;
          h.dev.o_twxN = g.o_twx;
        }
    }
    e = hipMalloc((void **)&x->arena, off);
    if (e != hipSuccess) rc = set_err(LH_ENOMEM, "hipMalloc(%zu): %s", off, hipGetErrorString(e));
    x->arena_bytes = off;
    // statistics accumulators: ST_N row profiles and SP_N spectra per channel
    size_t so = 0;
    for (auto &h : x->ch) {
      h.dev.o_stat = so;
      so += al<double>((size_t)ST_N * h.nyu);
      h.dev.o_spec = so;
      so += al<double>((size_t)SP_N * h.nyu * h.ldc);
      h.dev.o_slip = so;
      so += al<double2>((size_t)8 * h.ncol);
    }
    x->stats_bytes = so;
  }
  if (rc == LH_OK) {
    e = hipMemset(x->arena, 0, x->arena_bytes);
    if (e != hipSuccess) rc = set_err(LH_EHIP, "hipMemset: %s", hipGetErrorString(e));
  }
  for (int c = 0; c < nchan && rc == LH_OK; c++) rc = init_chan(x, x->ch[c]);
  for (size_t q = 0; q < x->tex_groups.size() && rc == LH_OK; q++) {
    const TexGroup &g = x->tex_groups[q];
    std::vector<double2> tz, tx;
    twiddles(tz, g.N2);
    twiddles(tx, g.N1);
    e = hipMemcpy(x->arena + g.o_twz, tz.data(), tz.size() * sizeof(double2), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(x->arena + g.o_twx, tx.data(), tx.size() * sizeof(double2), hipMemcpyHostToDevice);
    if (e != hipSuccess) rc = set_err(LH_EHIP, "texture twiddles: %s", hipGetErrorString(e));
  }
  if (rc == LH_OK) {
    e = hipMalloc(&x->d_chans, nchan * sizeof(Chan));
    if (e != hipSuccess) rc = set_err(LH_ENOMEM, "hipMalloc(descriptors): %s", hipGetErrorString(e));
  }
  if (rc == LH_OK) {
    e = hipMalloc(&x->staging, stage * sizeof(double2));
    if (e != hipSuccess) rc = set_err(LH_ENOMEM, "hipMalloc(staging): %s", hipGetErrorString(e));
    x->staging_elems = stage;
  }
  if (rc == LH_OK) rc = push_descriptors(x);
  if (rc == LH_OK) rc = build_nl_groups(x);  // channels of several plane shapes
  if (rc == LH_OK) rc = refresh_tables(x);
  if (rc != LH_OK) {
    std::string keep = g_err;
    lh_ctx_destroy(x);
    g_err = keep;
    return rc;
  }
  *out = x;
  return LH_OK;
}

int lh_ctx_destroy(lh_ctx *x) {
  if (!x) return LH_OK;
  (void)hipSetDevice(x->device);
  if (x->stream) (void)hipStreamSynchronize(x->stream);
  if (x->graph) (void)hipGraphExecDestroy(x->graph);
  if (x->graph2) (void)hipGraphExecDestroy(x->graph2);
  if (x->arena) (void)hipFree(x->arena);
  if (x->d_chans) (void)hipFree(x->d_chans);
  if (x->d_bands) (void)hipFree(x->d_bands);
  for (auto &g : x->nlg)
    if (g.d) (void)hipFree(g.d);
  if (x->staging) (void)hipFree(x->staging);
  if (x->stats) (void)hipFree(x->stats);
  if (x->hist) (void)hipFree(x->hist);
  for (char *p : x->tri_mem)
    if (p) (void)hipFree(p);
  for (char *p : x->cl_mem)
    if (p) (void)hipFree(p);
  if (x->cl_net) (void)hipFree(x->cl_net);
  if (x->stream) (void)hipStreamDestroy(x->stream);
  delete x;
  return LH_OK;
}

int lh_set_grid(lh_ctx *x, int ch, const double *yu, const double *yv, const double *dthdyu,
                const double *dthdyv) {
  int rc = check_ctx(x);
  if (rc) return rc;
  if (ch < 0 || ch >= (int)x->ch.size() || !yu || !yv || !dthdyu || !dthdyv)
    return set_err(LH_EINVAL, "lh_set_grid: bad arguments");
  HostChan &h = x->ch[ch];
  h.yu.assign(yu, yu + h.nyu);
  h.yv.assign(yv, yv + h.nyv);
  h.dthdyu.assign(dthdyu, dthdyu + h.nyu);
  h.dthdyv.assign(dthdyv, dthdyv + h.nyv);
  derive_grid(h);
  HIPCHK(hipStreamSynchronize(x->stream));
  rc = upload_chan_constants(x, h);
  if (rc) return rc;
  rc = push_descriptors(x);
  if (rc) return rc;
  return refresh_tables(x);
}

static void drop_graph(lh_ctx *x);

int lh_band_rows(const lh_config *mid, double band_h, int *nb) {
  if (!mid || !nb || !(band_h > 0.0)) return set_err(LH_EINVAL, "lh_band_rows: bad arguments");
  HostChan h;
  int rc = setup_host(h, *mid);
  if (rc) return rc;
  // ygrid's band split (start.f90:999-1011): the first row with
  // yu(j) >= yu(0) + h
  int j = 0;
  while (j < h.nyu - 1 && h.yu[j] < h.yu[0] + band_h) j++;
  *nb = j;
  return LH_OK;
}

int lh_set_bands(lh_ctx *x, int cb, int cm, int ct, int nb) {
  int rc = check_ctx(x);
  if (rc) return rc;
  const int n = (int)x->ch.size();
  if (cb < 0 || cm < 0 || ct < 0 || cb >= n || cm >= n || ct >= n || cb == cm || cb == ct || cm == ct)
    return set_err(LH_EINVAL, "lh_set_bands: bad channels %d, %d, %d", cb, cm, ct);
  HostChan &B = x->ch[cb], &M = x->ch[cm], &T = x->ch[ct];
  if (B.band >= 0 || M.band >= 0 || T.band >= 0)
    return set_err(LH_EINVAL, "lh_set_bands: a channel is a band already");
  const int ny = M.cfg.ny;
  if (nb < 1 || 2 * (nb + 2) + 2 > ny)
    return set_err(LH_EINVAL, "lh_set_bands: the wall bands (v rows 0..%d and %d..%d) leave no middle band", nb + 1,
                   ny - 2 - nb, ny - 1);
  for (HostChan *w : {&B, &T}) {
    const lh_config &c = w->cfg;
    if (c.ny != nb + 2)
      return set_err(LH_EINVAL, "lh_set_bands: a wall band has ny = %d, not nb + 2 = %d", c.ny, nb + 2);
    if (c.N1 < M.cfg.N1 || c.N2 < M.cfg.N2)
      return set_err(LH_EINVAL, "lh_set_bands: a wall band has fewer modes than the middle band");
    if (c.alp != M.cfg.alp || c.bet != M.cfg.bet || c.Re != M.cfg.Re || c.dt != M.cfg.dt ||
        c.bslpu1 != M.cfg.bslpu1 || c.bslpu3 != M.cfg.bslpu3)
      return set_err(LH_EINVAL, "lh_set_bands: the bands differ in alp, bet, Re, dt or slip lengths");
    if (c.flag_ctpress != 1 || c.mpgx != 0.0)
      return set_err(LH_EINVAL, "lh_set_bands: a wall band holds no mean mode (flag_ctpress 1, mpgx 0)");
    if (c.adaptive_dt || M.cfg.adaptive_dt) return set_err(LH_EINVAL, "lh_set_bands: banded channels take a fixed dt");
    if (w->dev.tex || (x->tri_mem.size() > (size_t)(w - x->ch.data()) && x->tri_mem[w - x->ch.data()]))
      return set_err(LH_EINVAL, "lh_set_bands: a wall band has textured walls or triad lists");
  }
  if (M.dev.tex || (x->tri_mem.size() > (size_t)cm && x->tri_mem[cm]))
    return set_err(LH_EINVAL, "lh_set_bands: the middle band has textured walls or triad lists");
  for (int c : {cb, cm, ct})
    if (x->cl_on.size() > (size_t)c && x->cl_on[c])
      return set_err(LH_EINVAL, "lh_set_bands: channel %d runs the closure", c);
  HIPCHK(hipStreamSynchronize(x->stream));
  // the wall bands' grids are the middle band's rows (jlim, start.f90:
  // 1677-1711): v rows 0..nb+1 and ny-2-nb..ny-1, u rows one more; their
  // edge weights then follow from their own end rows (start.f90:539-559)
  const int r0t = ny - 2 - nb;
  for (int q = 0; q < 2; q++) {
    HostChan &w = q == 0 ? B : T;
    const int r0 = q == 0 ? 0 : r0t;
    w.yu.assign(M.yu.begin() + r0, M.yu.begin() + r0 + w.nyu);
    w.yv.assign(M.yv.begin() + r0, M.yv.begin() + r0 + w.nyv);
    w.dthdyu.assign(M.dthdyu.begin() + r0, M.dthdyu.begin() + r0 + w.nyu);
    w.dthdyv.assign(M.dthdyv.begin() + r0, M.dthdyv.begin() + r0 + w.nyv);
    derive_grid(w);
    if ((rc = upload_chan_constants(x, w))) return rc;
    w.dev.band_edge = q == 0 ? 2 : 1;
  }
  // the middle band's planes: u rows nb+2..ny-2-nb, v rows nb+1..ny-2-nb
  // (proc_lims_planes, start.f90:1940-1965)
  M.dev.pu0 = nb + 2;
  M.dev.nplu = ny - 3 - 2 * nb;
  M.dev.pv0 = nb + 1;
  M.dev.nplv = ny - 2 - 2 * nb;
  M.dev.nslot = M.dev.nplu + M.dev.nplv;
  B.band = 0;
  M.band = 1;
  T.band = 2;
  x->bands.push_back(BandLink{cb, cm, ct, r0t});
  x->band_ncol = std::max(x->band_ncol, M.ncol);
  if (x->d_bands) (void)hipFree(x->d_bands);
  x->d_bands = nullptr;
  hipError_t e = hipMalloc((void **)&x->d_bands, x->bands.size() * sizeof(BandLink));
  if (e != hipSuccess) return set_err(LH_ENOMEM, "hipMalloc(band links): %s", hipGetErrorString(e));
  HIPCHK(hipMemcpy(x->d_bands, x->bands.data(), x->bands.size() * sizeof(BandLink), hipMemcpyHostToDevice));
  if ((rc = build_nl_groups(x))) return rc;
  if ((rc = refresh_tables(x))) return rc;
  drop_graph(x);
  return LH_OK;
}

int lh_get_grid(lh_ctx *x, int ch, double *yu, double *yv, double *dthdyu, double *dthdyv) {
  if (!x || ch < 0 || ch >= (int)x->ch.size()) return set_err(LH_EINVAL, "lh_get_grid: bad arguments");
  HostChan &h = x->ch[ch];
  if (yu) memcpy(yu, h.yu.data(), h.nyu * sizeof(double));
  if (yv) memcpy(yv, h.yv.data(), h.nyv * sizeof(double));
  if (dthdyu) memcpy(dthdyu, h.dthdyu.data(), h.nyu * sizeof(double));
  if (dthdyv) memcpy(dthdyv, h.dthdyv.data(), h.nyv * sizeof(double));
  return LH_OK;
}

int lh_field_upload(lh_ctx *x, int ch, int field, const double *reim) {
  int rc = check_ctx(x);
  if (rc) return rc;
  if (ch < 0 || ch >= (int)x->ch.size() || !reim) return set_err(LH_EINVAL, "lh_field_upload: bad arguments");
  HostChan &h = x->ch[ch];
  double2 *arr;
  int jlo, nrows;
  if (field_map(x, h, field, &arr, &jlo, &nrows)) return set_err(LH_EINVAL, "lh_field_upload: bad field %d", field);
  size_t n = (size_t)h.ncol * nrows;
  HIPCHK(hipMemcpyAsync(x->staging, reim, n * sizeof(double2), hipMemcpyHostToDevice, x->stream));
  HIPCHK(launch_to_rows(x->staging, arr, h.ncol, nrows, jlo, h.ldc, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  return LH_OK;
}

int lh_field_download(lh_ctx *x, int ch, int field, double *reim) {
  int rc = check_ctx(x);
  if (rc) return rc;
  if (ch < 0 || ch >= (int)x->ch.size() || !reim) return set_err(LH_EINVAL, "lh_field_download: bad arguments");
  HostChan &h = x->ch[ch];
  double2 *arr;
  int jlo, nrows;
  if (field_map(x, h, field, &arr, &jlo, &nrows)) return set_err(LH_EINVAL, "lh_field_download: bad field %d", field);
  size_t n = (size_t)h.ncol * nrows;
  HIPCHK(launch_from_rows(arr, x->staging, h.ncol, nrows, jlo, h.ldc, x->stream));
  HIPCHK(hipMemcpyAsync(reim, x->staging, n * sizeof(double2), hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  return LH_OK;
}

int lh_get_scalars(lh_ctx *x, int ch, double *o) {
  int rc = check_ctx(x);
  if (rc) return rc;
  if (ch < 0 || ch >= (int)x->ch.size() || !o) return set_err(LH_EINVAL, "lh_get_scalars: bad arguments");
  HostChan &h = x->ch[ch];
  double sc[SC_NSCAL];
  HIPCHK(hipStreamSynchronize(x->stream));
  HIPCHK(hipMemcpy(sc, View(x->arena, h.dev).scal, sizeof sc, hipMemcpyDeviceToHost));
  o[0] = sc[SC_T];
  o[1] = sc[SC_DT];
  o[2] = sc[SC_MPGX];
  o[3] = sc[SC_QXT];
  o[4] = sc[SC_QX];
  o[5] = h.dtv;
  o[6] = sc[SC_DTC];
  o[7] = h.iter;
  return LH_OK;
}

int lh_history(lh_ctx *x, int ch, double *o) {
  int rc = check_ctx(x);
  if (rc) return rc;
  if (ch < 0 || ch >= (int)x->ch.size() || !o) return set_err(LH_EINVAL, "lh_history: bad arguments");
  if (!x->hist) {
    hipError_t e = hipMalloc((void **)&x->hist, 8 * x->ch.size() * sizeof(double));
    if (e != hipSuccess) {
      x->hist = nullptr;
      return set_err(LH_ENOMEM, "lh_history: hipMalloc: %s", hipGetErrorString(e));
    }
  }
  HostChan &h = x->ch[ch];
  HIPCHK(launch_history(x->arena, x->d_chans, x->shape, x->hist, x->stream));
  double r[8], sc[SC_NSCAL];
  HIPCHK(hipMemcpyAsync(r, x->hist + 8 * ch, sizeof r, hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipMemcpyAsync(sc, View(x->arena, h.dev).scal, sizeof sc, hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  o[0] = sc[SC_T];
  o[1] = h.dtv;
  o[2] = sc[SC_DTC];
  o[3] = sc[SC_DT];
  o[4] = r[0];
  o[5] = r[1];
  o[6] = sc[SC_QXT];
  o[7] = sc[SC_MPGX];
  o[8] = sc[SC_DGX];
  o[9] = r[2];
  o[10] = r[3];
  o[11] = h.iter;
  return LH_OK;
}

int lh_set_scalars(lh_ctx *x, int ch, double t, double mpgx, double QxT, int iter) {
  int rc = check_ctx(x);
  if (rc) return rc;
  if (ch < 0 || ch >= (int)x->ch.size()) return set_err(LH_EINVAL, "lh_set_scalars: bad channel");
  HostChan &h = x->ch[ch];
  double sc[SC_NSCAL];
  HIPCHK(hipStreamSynchronize(x->stream));
  HIPCHK(hipMemcpy(sc, View(x->arena, h.dev).scal, sizeof sc, hipMemcpyDeviceToHost));
  sc[SC_MPGX] = mpgx;
  sc[SC_QXT] = QxT;
  sc[SC_T] = t;
  HIPCHK(hipMemcpy(View(x->arena, h.dev).scal, sc, sizeof sc, hipMemcpyHostToDevice));
  h.iter = iter;
  return LH_OK;
}

int lh_init_flowrate(lh_ctx *x) {
  int rc = check_ctx(x);
  if (rc) return rc;
  HIPCHK(launch_flowrate(x->arena, x->d_chans, (int)x->ch.size(), 1, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  return LH_OK;
}

int lh_run_phase(lh_ctx *x, int phase, int kRK) {
  int rc = check_ctx(x);
  if (rc) return rc;
  if (kRK < 1 || kRK > 3) return set_err(LH_EINVAL, "kRK must be 1..3");
  switch (phase) {
    case LH_PHASE_NONLINEAR: {
      int r = nonlinear_stage(x, kRK);
      if (r) return r;
      r = time_step(x, kRK);
      if (r) return r;
      break;
    }
    case LH_PHASE_RHS: HIPCHK(rhs_stage(x, x->shape, kRK)); break;
    case LH_PHASE_SOLVE: {
      int r = solve_phase(x, kRK);
      if (r) return r;
      break;
    }
    default: return set_err(LH_EINVAL, "unknown phase %d", phase);
  }
  HIPCHK(hipStreamSynchronize(x->stream));
  return LH_OK;
}

int lh_substep(lh_ctx *x, int kRK) {
  int rc = check_ctx(x);
  if (rc) return rc;
  if (kRK < 1 || kRK > 3) return set_err(LH_EINVAL, "kRK must be 1..3");
  rc = run_substep(x, kRK);
  if (rc) return rc;
  if (kRK == 3)
    for (auto &h : x->ch) h.iter += 1;
  return LH_OK;
}

int lh_step(lh_ctx *x, int nsteps) {
  int rc = check_ctx(x);
  if (rc) return rc;
  if (nsteps < 0) return set_err(LH_EINVAL, "nsteps must be >= 0");
  if (nsteps == 0) return LH_OK;
  // chained steps: each step after the first takes over the previous
  // step's last ghost pass, and one launch after the last step makes it
  const bool chain = ghost_deferrable(x);
  if (!x->graph && (rc = capture_step(x, false, chain, &x->graph))) return rc;
  if (chain && nsteps > 1 && !x->graph2 && (rc = capture_step(x, true, true, &x->graph2))) return rc;
  HIPCHK(hipGraphLaunch(x->graph, x->stream));
  for (int n = 1; n < nsteps; n++) HIPCHK(hipGraphLaunch(chain ? x->graph2 : x->graph, x->stream));
  if (chain) HIPCHK(launch_tex_1d(x->arena, x->d_chans, x->shape, 1, 1, 1, x->stream));
  HIPCHK(hipGetLastError());
  for (auto &h : x->ch) h.iter += nsteps;
  return LH_OK;
}

int lh_time_kernels(lh_ctx *x, int nsub, double *ms) {
  int rc = check_ctx(x);
  if (rc) return rc;
  if (nsub < 1 || !ms) return set_err(LH_EINVAL, "lh_time_kernels: bad arguments");
  hipEvent_t ev[8];
  for (int q = 0; q < 8; q++) HIPCHK(hipEventCreate(&ev[q]));
  for (int q = 0; q < 6; q++) ms[q] = 0.0;
  int kRK = 1;
  for (int n = 0; n < nsub && rc == LH_OK; n++) {
    HIPCHK(hipEventRecord(ev[0], x->stream));
    for (int st = 0; st < 3; st++) {
      // the band copies are counted with the first and the last stage
      if (st == 0)
        HIPCHK(launch_band_copy(x->arena, x->d_chans, x->d_bands, (int)x->bands.size(), x->band_ncol, 0, x->stream));
      if (x->ntri < (int)x->ch.size()) HIPCHK(nl_stage(x, x->shape, st, kRK));
      if (st == 2)
        HIPCHK(launch_band_copy(x->arena, x->d_chans, x->d_bands, (int)x->bands.size(), x->band_ncol, 1, x->stream));
      // the triad-list kernels are counted with the x-product stage, the
      // closure with the z-forward one
      if (st == 1 && x->ntri) HIPCHK(launch_triads(x->arena, x->d_chans, x->shape, x->stream));
      if (st == 2 && x->ncl && (rc = closure_stage(x, kRK))) break;
      HIPCHK(hipEventRecord(ev[st + 1], x->stream));
    }
    if (rc) break;
    if ((rc = time_step(x, kRK))) break;
    HIPCHK(rhs_stage(x, x->shape, kRK));
    HIPCHK(hipEventRecord(ev[4], x->stream));
    if ((rc = tex_part(x, kRK, 0))) break;
    HIPCHK(hipEventRecord(ev[5], x->stream));
    HIPCHK(solve_stage(x, x->shape, kRK));
    HIPCHK(hipEventRecord(ev[6], x->stream));
    if ((rc = tex_part(x, kRK, 1))) break;
    HIPCHK(hipEventRecord(ev[7], x->stream));
    HIPCHK(hipEventSynchronize(ev[7]));
    float t[7];
    for (int q = 0; q < 7; q++) HIPCHK(hipEventElapsedTime(&t[q], ev[q], ev[q + 1]));
    ms[0] += t[0];
    ms[1] += t[1];
    ms[2] += t[2];
    ms[3] += t[3];
    ms[4] += t[5];
    ms[5] += t[4] + t[6];
    if (kRK == 3)
      for (auto &h : x->ch) h.iter += 1;
    kRK = kRK % 3 + 1;
  }
  for (int q = 0; q < 8; q++) (void)hipEventDestroy(ev[q]);
  return rc;
}

int lh_sync(lh_ctx *x) {
  int rc = check_ctx(x);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(x->stream));
  return LH_OK;
}

// ---------------------------------------------------------------------------
// statistics and spectra (stats.f90:1-107, spectra.f90:1-56)

static int stats_alloc(lh_ctx *x) {
  if (x->stats) return LH_OK;
  hipError_t e = hipMalloc((void **)&x->stats, x->stats_bytes);
  if (e != hipSuccess) {
    x->stats = nullptr;
    return set_err(LH_ENOMEM, "statistics buffer hipMalloc(%zu): %s", x->stats_bytes, hipGetErrorString(e));
  }
  HIPCHK(hipMemsetAsync(x->stats, 0, x->stats_bytes, x->stream));
  x->istat = 0;
  return LH_OK;
}

static_assert(LH_ST_N == ST_N && LH_SP_N == SP_N && LH_ST_WX2M == ST_WX2M && LH_SP_P == SP_P,
              "statistics enums of the C-ABI and the kernels differ");

int lh_stats_accumulate(lh_ctx *x) {
  int rc = check_ctx(x);
  if (rc) return rc;
  if ((rc = stats_alloc(x))) return rc;
  HIPCHK(launch_stats(x->arena, x->d_chans, x->shape, x->stats, x->stream));
  x->istat++;
  return LH_OK;
}

int lh_stats_reset(lh_ctx *x) {
  int rc = check_ctx(x);
  if (rc) return rc;
  if (x->stats) HIPCHK(hipMemsetAsync(x->stats, 0, x->stats_bytes, x->stream));
  x->istat = 0;
  return LH_OK;
}

int lh_stats_get(lh_ctx *x, int ch, double *prof, int *istat) {
  int rc = check_ctx(x);
  if (rc) return rc;
  if (ch < 0 || ch >= (int)x->ch.size() || !prof) return set_err(LH_EINVAL, "lh_stats_get: bad arguments");
  const HostChan &h = x->ch[ch];
  const size_t n = (size_t)LH_ST_N * h.nyu;
  if (!x->stats) {
    std::fill(prof, prof + n, 0.0);
  } else {
    HIPCHK(hipMemcpyAsync(prof, x->stats + h.dev.o_stat, n * sizeof(double), hipMemcpyDeviceToHost, x->stream));
    HIPCHK(hipStreamSynchronize(x->stream));
  }
  if (istat) *istat = x->istat;
  return LH_OK;
}

int lh_slip_stats_get(lh_ctx *x, int ch, int wall, double *maps) {
  int rc = check_ctx(x);
  if (rc) return rc;
  if (ch < 0 || ch >= (int)x->ch.size() || wall < 0 || wall > 1 || !maps)
    return set_err(LH_EINVAL, "lh_slip_stats_get: bad arguments");
  const HostChan &h = x->ch[ch];
  const int N1 = h.cfg.N1, N2 = h.cfg.N2, ig = N1 + 2;
  const size_t msz = (size_t)ig * N2;
  std::fill(maps, maps + 4 * msz, 0.0);
  if (!x->stats) return LH_OK;
  std::vector<double2> acc((size_t)4 * h.ncol);
  HIPCHK(hipMemcpyAsync(acc.data(), x->stats + h.dev.o_slip + (size_t)wall * 4 * h.ncol * sizeof(double2),
                        acc.size() * sizeof(double2), hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  // plane layout of modes_to_planes_phys_lims_2: mode (i, k) at (2i+1, k+1)
  for (int q = 0; q < 4; q++)
    for (int c = 0; c < h.ncol; c++) {
      const int i = c % h.nxh, k = c / h.nxh;
      double *m = maps + q * msz + (size_t)k * ig + 2 * i;
      m[0] = acc[(size_t)q * h.ncol + c].x;
      m[1] = acc[(size_t)q * h.ncol + c].y;
    }
  return LH_OK;
}

int lh_spectrum_get(lh_ctx *x, int ch, int which, double *sp) {
  int rc = check_ctx(x);
  if (rc) return rc;
  if (ch < 0 || ch >= (int)x->ch.size() || which < 0 || which >= LH_SP_N || !sp)
    return set_err(LH_EINVAL, "lh_spectrum_get: bad arguments");
  const HostChan &h = x->ch[ch];
  // rows of the spectrum's grid: u (0..ny), v (0..ny-1), p (1..ny-1)
  const int jlo = which == LH_SP_P ? 1 : 0;
  const int nrows = which == LH_SP_V ? h.nyv : (which == LH_SP_P ? h.nyu - 2 : h.nyu);
  std::vector<double> rows((size_t)nrows * h.ldc);
  if (!x->stats) {
    std::fill(sp, sp + (size_t)nrows * h.ncol, 0.0);
    return LH_OK;
  }
  const char *src = x->stats + h.dev.o_spec + ((size_t)which * h.nyu + jlo) * h.ldc * sizeof(double);
  HIPCHK(hipMemcpyAsync(rows.data(), src, rows.size() * sizeof(double), hipMemcpyDeviceToHost, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  for (int c = 0; c < h.ncol; c++)
    for (int r = 0; r < nrows; r++) sp[(size_t)c * nrows + r] = rows[(size_t)r * h.ldc + c];
  return LH_OK;
}

static std::vector<double2> &tw_cache(int n) {
  static std::vector<std::vector<double2>> store;
  for (auto &v : store)
    if ((int)v.size() == n) return v;
  store.emplace_back();
  twiddles(store.back(), n);
  return store.back();
}

static int plane_args(int N1, int Ngal1, int Ngal2, double *d, int np) {
  if (!d || np < 1 || N1 < 2 || (N1 & 1) || Ngal1 < N1 + 2 || !valid_len(Ngal1) || !valid_len(Ngal2))
    return set_err(LH_EINVAL, "plane transform: bad arguments");
  return LH_OK;
}

static int plane_tw(int Ngal1, int Ngal2, double2 **dz, double2 **dx) {
  static std::vector<std::pair<int, double2 *>> dev;
  int dv = 0;
  (void)hipGetDevice(&dv);
  auto find = [&](int n) -> double2 * {
    int key = n * 64 + dv;
    for (auto &p : dev)
      if (p.first == key) return p.second;
    std::vector<double2> &h = tw_cache(n);
    double2 *d = nullptr;
    if (hipMalloc(&d, n * sizeof(double2)) != hipSuccess) return nullptr;
    if (hipMemcpy(d, h.data(), n * sizeof(double2), hipMemcpyHostToDevice) != hipSuccess) return nullptr;
    dev.push_back({key, d});
    return d;
  };
  *dz = find(Ngal2);
  *dx = find(Ngal1);
  if (!*dz || !*dx) return set_err(LH_ENOMEM, "plane transform: twiddle table allocation failed");
  return LH_OK;
}

int lh_four_to_phys_planes(int N1, int Ngal1, int Ngal2, double *d, int np, void *stream) {
  int rc = plane_args(N1, Ngal1, Ngal2, d, np);
  if (rc) return rc;
  double2 *tz, *tx;
  rc = plane_tw(Ngal1, Ngal2, &tz, &tx);
  if (rc) return rc;
  HIPCHK(launch_plane_f2p(N1, Ngal1, Ngal2, d, np, tz, tx, (hipStream_t)stream));
  return LH_OK;
}

int lh_phys_to_four_planes(int N1, int Ngal1, int Ngal2, double *d, int np, void *stream) {
  int rc = plane_args(N1, Ngal1, Ngal2, d, np);
  if (rc) return rc;
  double2 *tz, *tx;
  rc = plane_tw(Ngal1, Ngal2, &tz, &tx);
  if (rc) return rc;
  HIPCHK(launch_plane_p2f(N1, Ngal1, Ngal2, d, np, tz, tx, (hipStream_t)stream));
  return LH_OK;
}

// FFTPACK's cft / rft (cft_buff.f:1-65, rft_buff.f:24-90) on device data.
// The vectors are transformed concurrently, so they must not share
// elements (the reference runs them one after another).
static int fft_vec_args(const char *who, int n, int ise, int isa, int m, const double *d, int elem) {
  if (!d || m < 1 || n < 2 || (n & 1) || !valid_len(n) || n > 2048 || ise < elem || isa < 1)
    return set_err(LH_EINVAL, "%s: bad arguments (n %d: even, 2^a 3^b, <= 2048)", who, n);
  const size_t span = (size_t)(m - 1) * isa + (size_t)(n - 1) * ise + elem;
  if (span > ((size_t)1 << 32)) return set_err(LH_EINVAL, "%s: %zu doubles is too large", who, span);
  std::vector<unsigned char> used(span, 0);
  for (int k = 0; k < m; k++)
    for (int j = 0; j < n; j++)
      for (int e = 0; e < elem; e++) {
        unsigned char &u = used[(size_t)k * isa + (size_t)j * ise + e];
        if (u) return set_err(LH_EINVAL, "%s: vectors overlap (ise %d, isa %d, m %d)", who, ise, isa, m);
        u = 1;
      }
  return LH_OK;
}

int lh_cft(double *d, int ise, int isa, int m, int iopt, int n, void *stream) {
  int rc = fft_vec_args("lh_cft", n, ise, isa, m, d, 2);
  if (rc) return rc;
  double2 *tw, *tw2;
  rc = plane_tw(n, n, &tw, &tw2);
  if (rc) return rc;
  HIPCHK(launch_cft(n, ise, isa, m, iopt, d, tw, (hipStream_t)stream));
  return LH_OK;
}

int lh_rft(double *d, int isa, int m, int iopt, int n, void *stream) {
  if (isa < n + 2) return set_err(LH_EINVAL, "lh_rft: isa %d below n + 2 = %d", isa, n + 2);
  int rc = fft_vec_args("lh_rft", n, 1, isa, m, d, 1);
  if (rc) return rc;
  double2 *tw, *tw2;
  rc = plane_tw(n, n, &tw, &tw2);
  if (rc) return rc;
  HIPCHK(launch_rft(n, isa, m, iopt, d, tw, (hipStream_t)stream));
  return LH_OK;
}

// ---------------------------------------------------------------------------
// The plane transforms without allocation (no counterpart in the reference;
// for callers that cannot allocate device memory: code under HIP graph
// capture, tight host loops).  lh_four_to_phys_planes / lh_phys_to_four_planes
// allocate their twiddle tables on the first use of a length; here the two
// tables live in `scratch`, a file-scope device array (48 KiB, held by every
// GPU the library runs on until the process ends), and a small kernel writes
// them on the caller's stream at every call, before the transforms that read
// them.  Calls that overlap on one GPU share the array and race.
#define LH_NOALLOC_MAXN 1536
#define LH_NOALLOC_MAXDEV 64
static __device__ __align__(256) double2 scratch[2 * LH_NOALLOC_MAXN];

// tw[m] = exp(-2 pi i m / n), table t of the launch (z: n2 at 0, x: n1 at
// LH_NOALLOC_MAXN); sincospi keeps the quarter turns exact
__global__ void __launch_bounds__(256) k_noalloc_twiddles(double2 *tw, int n2, int n1) {
  const int n = blockIdx.y == 0 ? n2 : n1;
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  double sn, cs;
  sincospi(2.0 * (double)m / (double)n, &sn, &cs);
  tw[blockIdx.y * LH_NOALLOC_MAXN + m] = make_double2(cs, -sn);
}

// the device address of `scratch` on the current device, looked up once per
// device (host code must never pass the array's name to a kernel)
static int noalloc_scratch(double2 **out) {
  static double2 *ptr[LH_NOALLOC_MAXDEV];
  static std::mutex mu;
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= LH_NOALLOC_MAXDEV) return set_err(LH_EINVAL, "lh_plane_transform_noalloc: device %d", dev);
  std::lock_guard<std::mutex> g(mu);
  if (!ptr[dev]) {
    double2 *p = nullptr;
    hipError_t e = hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return set_err(LH_EHIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    ptr[dev] = p;
  }
  *out = ptr[dev];
  return LH_OK;
}

int lh_plane_transform_noalloc(int dir, int N1, int Ngal1, int Ngal2, double *d, int np, void *stream) {
  if (dir != 1 && dir != -1) return set_err(LH_EINVAL, "lh_plane_transform_noalloc: dir must be +1 or -1");
  int rc = plane_args(N1, Ngal1, Ngal2, d, np);
  if (rc) return rc;
  if (Ngal1 > LH_NOALLOC_MAXN || Ngal2 > LH_NOALLOC_MAXN)
    return set_err(LH_EINVAL, "lh_plane_transform_noalloc: Ngal1, Ngal2 at most %d (got %d, %d)", LH_NOALLOC_MAXN,
                   Ngal1, Ngal2);
  double2 *tw = nullptr;
  rc = noalloc_scratch(&tw);
  if (rc) return rc;
  const hipStream_t s = (hipStream_t)stream;
  const int nmax = Ngal1 > Ngal2 ? Ngal1 : Ngal2;
  hipLaunchKernelGGL(k_noalloc_twiddles, dim3((nmax + 255) / 256, 2), dim3(256), 0, s, tw, Ngal2, Ngal1);
  HIPCHK(hipGetLastError());
  if (dir > 0)
    HIPCHK(launch_plane_f2p(N1, Ngal1, Ngal2, d, np, tw, tw + LH_NOALLOC_MAXN, s));
  else
    HIPCHK(launch_plane_p2f(N1, Ngal1, Ngal2, d, np, tw, tw + LH_NOALLOC_MAXN, s));
  return LH_OK;
}

int lh_lusol_uw(lh_ctx *x, int ch, int kRK, const double *ru, const double *rw, double *u, double *w) {
  int rc = check_ctx(x);
  if (rc) return rc;
  if (ch < 0 || ch >= (int)x->ch.size() || kRK < 1 || kRK > 3 || !ru || !rw || !u || !w)
    return set_err(LH_EINVAL, "lh_lusol_uw: bad arguments");
  HostChan &h = x->ch[ch];
  const Chan *dc = x->d_chans + ch;
  HIPCHK(launch_lusol_ref2(x->arena, dc, kRK, 0, (const double2 *)ru, (const double2 *)rw, (double2 *)u, (double2 *)w,
                           View(x->arena, h.dev).scr, h.ncol, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  return LH_OK;
}

int lh_lusol_v(lh_ctx *x, int ch, int kRK, const double *rv, double *v) {
  int rc = check_ctx(x);
  if (rc) return rc;
  if (ch < 0 || ch >= (int)x->ch.size() || kRK < 1 || kRK > 3 || !rv || !v)
    return set_err(LH_EINVAL, "lh_lusol_v: bad arguments");
  HostChan &h = x->ch[ch];
  const Chan *dc = x->d_chans + ch;
  HIPCHK(launch_lusol_ref2(x->arena, dc, kRK, 1, (const double2 *)rv, nullptr, (double2 *)v, nullptr, View(x->arena, h.dev).scr, h.ncol,
                           x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  return LH_OK;
}

int lh_lusol_p(lh_ctx *x, int ch, double *d) {
  int rc = check_ctx(x);
  if (rc) return rc;
  if (ch < 0 || ch >= (int)x->ch.size() || !d) return set_err(LH_EINVAL, "lh_lusol_p: bad arguments");
  HostChan &h = x->ch[ch];
  HIPCHK(launch_lusol_p_ref(x->arena, x->d_chans + ch, (double2 *)d, h.ncol, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  return LH_OK;
}

int lh_lusol0(lh_ctx *x, int ch, int kRK, double *d) {
  int rc = check_ctx(x);
  if (rc) return rc;
  if (ch < 0 || ch >= (int)x->ch.size() || kRK < 1 || kRK > 3 || !d) return set_err(LH_EINVAL, "lh_lusol0: bad arguments");
  HostChan &h = x->ch[ch];
  HIPCHK(launch_lusol0_ref(x->arena, x->d_chans + ch, kRK, d, View(x->arena, h.dev).scr, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  return LH_OK;
}

static int planeN_args(int N1, int N2, double *d, int np) {
  if (!d || np < 1 || N1 < 2 || (N1 & 1) || N2 < 2 || (N2 & 1) || !valid_len(N1) || !valid_len(N2))
    return set_err(LH_EINVAL, "N-grid plane transform: bad arguments");
  return LH_OK;
}

int lh_four_to_phys_N_planes(int N1, int N2, double *d, int np, void *stream) {
  int rc = planeN_args(N1, N2, d, np);
  if (rc) return rc;
  double2 *tz, *tx;
  rc = plane_tw(N1, N2, &tz, &tx);
  if (rc) return rc;
  HIPCHK(launch_plane_f2p(N1, N1, N2, d, np, tz, tx, (hipStream_t)stream, false));
  return LH_OK;
}

int lh_phys_to_four_N_planes(int N1, int N2, double *d, int np, void *stream) {
  int rc = planeN_args(N1, N2, d, np);
  if (rc) return rc;
  double2 *tz, *tx;
  rc = plane_tw(N1, N2, &tz, &tx);
  if (rc) return rc;
  HIPCHK(launch_plane_p2f(N1, N1, N2, d, np, tz, tx, (hipStream_t)stream, false));
  return LH_OK;
}

static int transpose_args(int N1, int N2, int Ngal1, int Ngal2, const void *a, const void *b, int nrows) {
  if (!a || !b || nrows < 1 || N1 < 2 || (N1 & 1) || N2 < 2 || (N2 & 1) || Ngal1 < N1 || Ngal2 < N2 || (Ngal1 & 1))
    return set_err(LH_EINVAL, "transpose: bad arguments");
  return LH_OK;
}

int lh_modes_to_planes(int N1, int N2, int Ngal1, int Ngal2, const double *cols, int nrows, double *planes,
                       void *stream) {
  int rc = transpose_args(N1, N2, Ngal1, Ngal2, cols, planes, nrows);
  if (rc) return rc;
  HIPCHK(launch_modes_to_planes(N1, N2, Ngal1, Ngal2, (const double2 *)cols, nrows, planes, (hipStream_t)stream));
  return LH_OK;
}

int lh_planes_to_modes(int N1, int N2, int Ngal1, int Ngal2, const double *planes, int nrows, double *cols,
                       void *stream) {
  int rc = transpose_args(N1, N2, Ngal1, Ngal2, planes, cols, nrows);
  if (rc) return rc;
  HIPCHK(launch_planes_to_modes(N1, N2, Ngal1, Ngal2, planes, nrows, (double2 *)cols, (hipStream_t)stream));
  return LH_OK;
}

int lh_dev_alloc(size_t bytes, void **out) {
  if (!out || bytes == 0) return set_err(LH_EINVAL, "lh_dev_alloc: bad arguments");
  *out = nullptr;
  if (hipMalloc(out, bytes) != hipSuccess) {
    (void)hipGetLastError();
    *out = nullptr;
    return set_err(LH_ENOMEM, "lh_dev_alloc: hipMalloc of %zu bytes failed", bytes);
  }
  return LH_OK;
}

int lh_dev_free(void *p) {
  if (p) HIPCHK(hipFree(p));
  return LH_OK;
}

int lh_copy_to_device(void *d, const void *h, size_t bytes) {
  if (!d || !h) return set_err(LH_EINVAL, "lh_copy_to_device: null pointer");
  HIPCHK(hipMemcpy(d, h, bytes, hipMemcpyHostToDevice));
  return LH_OK;
}

int lh_copy_to_host(void *h, const void *d, size_t bytes) {
  if (!d || !h) return set_err(LH_EINVAL, "lh_copy_to_host: null pointer");
  HIPCHK(hipMemcpy(h, d, bytes, hipMemcpyDeviceToHost));
  return LH_OK;
}

// ------------------------------------------------------------------------
// tuning
static int solve_cols_ok(const lh_ctx *x, int C) {
  if (C == 0) return 1;
  if (C == -1) return solve_wave_rows(x->shape.nyu) > 0;
  if (C < 1 || C > 10) return 0;
  return solve_lds_bytes(x->shape.nyu, C) <= 160 * 1024;
}

static int *stage_knob(ShapeMax &m, int key) {
  switch (key) {
    case LH_TUNE_ZINV_POINTS: return &m.e_zinv;
    case LH_TUNE_XPROD_POINTS: return &m.e_xprod;
    case LH_TUNE_ZFWD_POINTS: return &m.e_zfwd;
  }
  return nullptr;
}

// LDS bytes of the largest nonlinear stage with the shape's targets
static size_t stage_lds_bytes(const lh_ctx *x, const ShapeMax &m) {
  size_t es = x->mixed ? sizeof(float2) : sizeof(double2);
  return std::max(m.lds_zinv, std::max(m.lds_xprod, m.lds_zfwd)) * es;
}

static void drop_graph(lh_ctx *x) {
  if (x->graph) {
    (void)hipGraphExecDestroy(x->graph);
    x->graph = nullptr;
  }
  if (x->graph2) {
    (void)hipGraphExecDestroy(x->graph2);
    x->graph2 = nullptr;
  }
}

int lh_set_tuning(lh_ctx *x, int key, int value) {
  int rc = check_ctx(x);
  if (rc) return rc;
  if (key == LH_TUNE_TEX_FUSED) {
    if (value != 0 && value != 1) return set_err(LH_EINVAL, "lh_set_tuning: LH_TUNE_TEX_FUSED is 0 or 1");
    HIPCHK(hipStreamSynchronize(x->stream));
    x->tex_fused = value;
    drop_graph(x);
    return LH_OK;
  }
  if (key == LH_TUNE_SHAPE_GROUPS) {
    if (value != 0 && value != 1) return set_err(LH_EINVAL, "lh_set_tuning: LH_TUNE_SHAPE_GROUPS is 0 or 1");
    HIPCHK(hipStreamSynchronize(x->stream));
    x->shape_groups = value;
    if ((rc = build_nl_groups(x))) return rc;
    drop_graph(x);
    return LH_OK;
  }
  if (key == LH_TUNE_TEX_SEPARABLE) {
    if (value != 0 && value != 1) return set_err(LH_EINVAL, "lh_set_tuning: LH_TUNE_TEX_SEPARABLE is 0 or 1");
    HIPCHK(hipStreamSynchronize(x->stream));
    x->tex_sep = value;
    tex_classify(x);
    if ((rc = push_descriptors(x))) return rc;
    drop_graph(x);
    return LH_OK;
  }
  if (key == LH_TUNE_XPROD_THREADS) {
    if (value != 128 && value != 192 && value != 256)
      return set_err(LH_EINVAL, "lh_set_tuning: LH_TUNE_XPROD_THREADS is 128, 192 or 256");
    ShapeMax m = x->shape;
    m.xprod_threads = value;
    shape_tiles(m);
    if (stage_lds_bytes(x, m) > 160 * 1024) return set_err(LH_EINVAL, "lh_set_tuning: LDS exceeds 160 KiB");
    HIPCHK(hipStreamSynchronize(x->stream));
    x->shape = m;
    drop_graph(x);
    return LH_OK;
  }
  if (key == LH_TUNE_XPROD_INPLACE || key == LH_TUNE_ZFWD_INPLACE || key == LH_TUNE_ZINV_INPLACE ||
      key == LH_TUNE_ZFWD_FIELDWISE) {
    if (value != 0 && value != 1 && !(value == 2 && (key == LH_TUNE_ZFWD_FIELDWISE || key == LH_TUNE_ZINV_INPLACE)))
      return set_err(LH_EINVAL, "lh_set_tuning: in-place switches are 0 or 1 (LH_TUNE_ZINV_INPLACE and "
                                "LH_TUNE_ZFWD_FIELDWISE also 2)");
    ShapeMax m = x->shape;
    (key == LH_TUNE_XPROD_INPLACE  ? m.xprod_ip
     : key == LH_TUNE_ZFWD_INPLACE ? m.zfwd_ip
     : key == LH_TUNE_ZINV_INPLACE ? m.zinv_ip
                                   : m.zfwd_fs) = value;
    shape_tiles(m);
    if (stage_lds_bytes(x, m) > 160 * 1024) return set_err(LH_EINVAL, "lh_set_tuning: LDS exceeds 160 KiB");
    HIPCHK(hipStreamSynchronize(x->stream));
    x->shape = m;
    drop_graph(x);
    return LH_OK;
  }
  if (key == LH_TUNE_SOLVE_WAVES) {
    if (value != 1 && value != 2 && value != 4)
      return set_err(LH_EINVAL, "lh_set_tuning: LH_TUNE_SOLVE_WAVES is 1, 2 or 4");
    HIPCHK(hipStreamSynchronize(x->stream));
    x->shape.solve_waves = value;
    drop_graph(x);
    return LH_OK;
  }
  if (key == LH_TUNE_RHS_FUSED) {
    if (value != 0 && value != 1) return set_err(LH_EINVAL, "lh_set_tuning: LH_TUNE_RHS_FUSED is 0 or 1");
    HIPCHK(hipStreamSynchronize(x->stream));
    x->shape.rhs_fused = value;
    drop_graph(x);
    return LH_OK;
  }
  if (key == LH_TUNE_GHOST_IN_ZINV) {
    if (value != 0 && value != 1) return set_err(LH_EINVAL, "lh_set_tuning: LH_TUNE_GHOST_IN_ZINV is 0 or 1");
    HIPCHK(hipStreamSynchronize(x->stream));
    x->shape.ghost_zinv = value;
    drop_graph(x);
    return LH_OK;
  }
  if (key == LH_TUNE_SOLVE_SCAN) {
    if (value != 0 && value != 1) return set_err(LH_EINVAL, "lh_set_tuning: LH_TUNE_SOLVE_SCAN is 0 or 1");
    HIPCHK(hipStreamSynchronize(x->stream));
    x->shape.solve_scan = value;
    drop_graph(x);
    return LH_OK;
  }
  if (key == LH_TUNE_SOLVE_COLS) {
    if (!solve_cols_ok(x, value))
      return set_err(LH_EINVAL,
                     "lh_set_tuning: %d columns per workgroup do not fit (0, -1 for ny < %d, or 1..10 "
                     "within 160 KiB of LDS)",
                     value, 64 * SOLVE_WAVE_KMAX);
    HIPCHK(hipStreamSynchronize(x->stream));
    x->shape.solve_cols = value;
    drop_graph(x);
    return LH_OK;
  }
  ShapeMax m = x->shape;
  int *k = stage_knob(m, key);
  if (!k) return set_err(LH_EINVAL, "lh_set_tuning: unknown key %d", key);
  if (value < 1) return set_err(LH_EINVAL, "lh_set_tuning: points per workgroup must be positive");
  *k = value;
  shape_tiles(m);
  if (stage_lds_bytes(x, m) > 160 * 1024)
    return set_err(LH_EINVAL, "lh_set_tuning: %d points per workgroup need more than 160 KiB of LDS", value);
  HIPCHK(hipStreamSynchronize(x->stream));
  x->shape = m;
  drop_graph(x);
  return LH_OK;
}

int lh_get_tuning(lh_ctx *x, int key, int *value) {
  int rc = check_ctx(x);
  if (rc) return rc;
  if (!value) return set_err(LH_EINVAL, "lh_get_tuning: null value");
  if (key == LH_TUNE_TEX_FUSED) {
    *value = x->tex_fused;
    return LH_OK;
  }
  if (key == LH_TUNE_SHAPE_GROUPS) {
    *value = x->shape_groups;
    return LH_OK;
  }
  if (key == LH_TUNE_TEX_SEPARABLE) {
    *value = x->tex_sep;
    return LH_OK;
  }
  if (key == LH_TUNE_SOLVE_COLS) {
    *value = x->shape.solve_cols;
    return LH_OK;
  }
  if (key == LH_TUNE_SOLVE_SCAN) {
    *value = x->shape.solve_scan;
    return LH_OK;
  }
  if (key == LH_TUNE_SOLVE_WAVES) {
    *value = x->shape.solve_waves;
    return LH_OK;
  }
  if (key == LH_TUNE_XPROD_INPLACE) {
    *value = x->shape.xprod_ip;
    return LH_OK;
  }
  if (key == LH_TUNE_ZFWD_INPLACE) {
    *value = x->shape.zfwd_ip;
    return LH_OK;
  }
  if (key == LH_TUNE_ZINV_INPLACE) {
    *value = x->shape.zinv_ip;
    return LH_OK;
  }
  if (key == LH_TUNE_ZFWD_FIELDWISE) {
    *value = x->shape.zfwd_fs;
    return LH_OK;
  }
  if (key == LH_TUNE_XPROD_THREADS) {
    *value = x->shape.xprod_threads;
    return LH_OK;
  }
  if (key == LH_TUNE_RHS_FUSED) {
    *value = x->shape.rhs_fused;
    return LH_OK;
  }
  if (key == LH_TUNE_GHOST_IN_ZINV) {
    *value = x->shape.ghost_zinv;
    return LH_OK;
  }
  int *k = stage_knob(x->shape, key);
  if (!k) return set_err(LH_EINVAL, "lh_get_tuning: unknown key %d", key);
  *value = *k;
  return LH_OK;
}

}  // extern "C"

// time `reps` launches of what `launch` issues; returns ms or a negative
// value when the launch is rejected
template <typename F> static float time_launches(lh_ctx *x, hipEvent_t e0, hipEvent_t e1, int reps, F launch) {
  if (launch(0) != hipSuccess) {
    (void)hipGetLastError();
    return -1.f;
  }
  if (hipEventRecord(e0, x->stream) != hipSuccess) return -1.f;
  for (int r = 0; r < reps; r++)
    if (launch(r) != hipSuccess) {
      (void)hipGetLastError();
      return -1.f;
    }
  if (hipEventRecord(e1, x->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) return -1.f;
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return -1.f;
  return ms;
}

extern "C" {

int lh_autotune(lh_ctx *x, int reps) {
  int rc = check_ctx(x);
  if (rc) return rc;
  if (reps < 1) return set_err(LH_EINVAL, "lh_autotune: reps must be >= 1");
  HIPCHK(hipStreamSynchronize(x->stream));
  // the trials run real kernels on the current state: keep a copy of the arena
  char *save = nullptr;
  if (hipMalloc(&save, x->arena_bytes) != hipSuccess) {
    (void)hipGetLastError();
    return set_err(LH_ENOMEM, "lh_autotune: no room for a %zu-byte state copy", x->arena_bytes);
  }
  HIPCHK(hipMemcpyAsync(save, x->arena, x->arena_bytes, hipMemcpyDeviceToDevice, x->stream));
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  ShapeMax best = x->shape;
  // nonlinear stages, in pipeline order (each reads what the previous wrote)
  const int pts[] = {256, 384, 512, 768, 1024, 1536, 2048, 3072, 4096};
  const int keys[3] = {LH_TUNE_ZINV_POINTS, LH_TUNE_XPROD_POINTS, LH_TUNE_ZFWD_POINTS};
  for (int st = 0; st < 3; st++) {
    float best_ms = 1e30f;
    for (int E : pts)
      for (int v2 = 0; v2 < (st == 0 ? 3 : st == 1 ? 6 : 4); v2++) {
        // z-inverse: in place in workgroups of 256 or 512 threads; x-product:
        // in place x 256/128/192 threads; z-forward: in place, or field-wise
        // in workgroups of 256 or 512 threads
        const int ip = st == 0 ? v2 : (v2 & 1) || (st == 2 && v2 >= 2);
        const int nt = st == 1 ? (v2 >> 1 == 0 ? 256 : v2 >> 1 == 1 ? 128 : 192) : 256;
        const int fs = st == 2 && v2 >= 2 ? v2 - 1 : 0;
        ShapeMax m = best;
        *stage_knob(m, keys[st]) = E;
        if (st == 1) {
          m.xprod_ip = ip;
          m.xprod_threads = nt;
        }
        if (st == 2) {
          m.zfwd_ip = ip;
          m.zfwd_fs = fs;
        }
        if (st == 0) m.zinv_ip = ip;
        shape_tiles(m);
        if (fs && !nl_fs_on(x, m)) continue;
        if (stage_lds_bytes(x, m) > 160 * 1024) continue;
        float ms = time_launches(
            x, e0, e1, reps, [&](int) { return nl_stage(x, m, st, 1); });
        if (ms >= 0.f && ms < best_ms) {
          best_ms = ms;
          *stage_knob(best, keys[st]) = E;
          if (st == 1) {
            best.xprod_ip = ip;
            best.xprod_threads = nt;
          }
          if (st == 2) {
            best.zfwd_ip = ip;
            best.zfwd_fs = fs;
          }
          if (st == 0) best.zinv_ip = ip;
        }
      }
    shape_tiles(best);
  }
  // column solve: one lane per column, or C columns per workgroup in LDS
  // with serial or lane-split recurrences; timed together with the
  // right-hand side, which smooth-walled contexts may compute inside the LDS
  // solve (a win only where the columns are many: at a few hundred columns
  // the separate kernel hides the loads' latency better)
  {
    const int cand[] = {0, -1, 1, 2, 3, 4, 5, 6, 8, 10};
    float best_ms = 1e30f;
    for (int C : cand) {
      if (!solve_cols_ok(x, C)) continue;
      for (int sc = 0; sc <= (C > 0 ? 1 : 0); sc++)
        for (int nw = 1; nw <= (C > 0 || solve_wave_rows(x->shape.nyu) <= 3 ? 4 : 1); nw *= (C > 0 ? 2 : 4))
          for (int fu = 0; fu <= ((C > 0 && !x->ntex) || C < 0 ? 1 : 0); fu++) {
            ShapeMax m = best;
            m.solve_cols = C;
            m.solve_scan = sc;
            m.solve_waves = nw;
            float ms = time_launches(x, e0, e1, reps, [&](int r) {
              const int kRK = 1 + r % 3;
              if (!fu || x->ntex) {
                hipError_t e = rhs_stage(x, m, kRK, fu);
                if (e != hipSuccess) return e;
              }
              return solve_stage(x, m, kRK, fu);
            });
            if (ms >= 0.f && ms < best_ms) {
              best_ms = ms;
              best.solve_cols = C;
              best.solve_scan = sc;
              best.solve_waves = nw;
              best.rhs_fused = fu;
            }
          }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  HIPCHK(hipMemcpyAsync(x->arena, save, x->arena_bytes, hipMemcpyDeviceToDevice, x->stream));
  HIPCHK(hipStreamSynchronize(x->stream));
  HIPCHK(hipFree(save));
  x->shape = best;
  drop_graph(x);
  return LH_OK;
}

int lh_set_triads(lh_ctx *x, int ch, int nfile, const int *lens, const signed char *rows, const int *map_u,
                  const int *map_v) {
  int rc = check_ctx(x);
  if (rc) return rc;
  if (ch < 0 || ch >= (int)x->ch.size()) return set_err(LH_EINVAL, "lh_set_triads: bad channel %d", ch);
  if (nfile < 0 || (nfile > 0 && (!lens || !rows || !map_u || !map_v)))
    return set_err(LH_EINVAL, "lh_set_triads: bad arguments");
  HostChan &h = x->ch[ch];
  if (nfile > 0 && h.band >= 0)
    return set_err(LH_EINVAL, "lh_set_triads: channel %d is a band of a banded channel", ch);
  if (x->tri_mem.size() != x->ch.size()) x->tri_mem.assign(x->ch.size(), nullptr);
  HIPCHK(hipStreamSynchronize(x->stream));
  if (nfile == 0) {
    if (x->tri_mem[ch]) HIPCHK(hipFree(x->tri_mem[ch]));
    x->tri_mem[ch] = nullptr;
    h.dev.tri = 0;
    h.dev.tri_map = nullptr;
    h.dev.tri_base = nullptr;
    h.dev.tri_off = h.dev.tri_ent = nullptr;
    h.dev.tri_prod = nullptr;
  } else {
    const lh_config &c = h.cfg;
    if (c.adaptive_dt) return set_err(LH_EINVAL, "lh_set_triads: the list path runs with a fixed time step");
    const int ncol = h.ncol, nxh = h.nxh, N2 = c.N2, nx = c.N1 / 2 - 1, nz = c.N2 / 2 - 1;
    if (ncol > 32767) return set_err(LH_EINVAL, "lh_set_triads: %d spectral columns exceed the 15-bit index", ncol);
    const int nslot = h.nslot, nplu = h.nyu - 2;
    std::vector<int> map(nslot);
    for (int s = 0; s < nslot; s++) {
      const int q = s < nplu ? map_u[s] : map_v[s - nplu];
      if (q >= nfile) return set_err(LH_EINVAL, "lh_set_triads: plane %d names file %d of %d", s, q, nfile);
      map[s] = q < 0 ? -1 : q;
    }
    auto kk = [&](int k) { return k >= 0 ? k : k + N2; };
    // (file, type) target-sorted entries: counting sort by target column
    const size_t nlist = (size_t)nfile * 9;
    std::vector<size_t> base(nlist);
    std::vector<unsigned> off(nlist * (ncol + 1));
    std::vector<unsigned> ent;
    size_t roff = 0;
    std::vector<unsigned> tgt, val;
    for (size_t lt = 0; lt < nlist; lt++) {
      const int n = lens[lt];
      if (n < 0) return set_err(LH_EINVAL, "lh_set_triads: negative list length");
      const signed char *d = rows + roff;
      roff += (size_t)4 * n;
      tgt.clear();
      val.clear();
      for (int l = 0; l < n; l++) {
        const int la = d[l], ka = d[n + l], lp = d[2 * n + l], kp = d[3 * n + l];
        if (abs(la) > nx || abs(lp) > nx || abs(ka) > nz || abs(kp) > nz || la + lp < 0 || la + lp > nx ||
            abs(ka + kp) > nz)
          return set_err(LH_EINVAL, "lh_set_triads: row (%d,%d,%d,%d) of list %zu outside the band", la, ka, lp, kp,
                         lt);
        // p = 0: the z mirror into -(ka+kp); p = 1 unless ka + kp = 0
        for (int p = 0; p <= 1; p++) {
          if (p == 1 && ka + kp == 0) continue;
          const int sg = p == 0 ? -1 : 1;
          const unsigned ca = la >= 0 ? (unsigned)(la + nxh * kk(sg * ka)) : (unsigned)(-la + nxh * kk(-sg * ka));
          const unsigned cp = lp >= 0 ? (unsigned)(lp + nxh * kk(sg * kp)) : (unsigned)(-lp + nxh * kk(-sg * kp));
          const unsigned w = ca | (la < 0 ? 0x8000u : 0u) | (cp << 16) | (lp < 0 ? 0x80000000u : 0u);
          tgt.push_back((unsigned)((la + lp) + nxh * kk(sg * (ka + kp))));
          val.push_back(w);
        }
      }
      unsigned *o = off.data() + lt * (ncol + 1);
      for (unsigned t : tgt) o[t + 1]++;
      for (int c2 = 0; c2 < ncol; c2++) o[c2 + 1] += o[c2];
      base[lt] = ent.size();
      ent.resize(ent.size() + val.size());
      std::vector<unsigned> fill(o, o + ncol);
      for (size_t e = 0; e < val.size(); e++) ent[base[lt] + fill[tgt[e]]++] = val[e];
    }
    // one device buffer: map, base, off, entries, products
    auto a8 = [](size_t b) { return (b + 255) / 256 * 256; };
    const size_t b_map = a8(nslot * sizeof(int)), b_base = a8(nlist * sizeof(size_t));
    const size_t b_off = a8(off.size() * sizeof(unsigned)), b_ent = a8(std::max<size_t>(ent.size(), 1) * sizeof(unsigned));
    const size_t b_prod = a8((size_t)nslot * 5 * ncol * sizeof(double2));
    char *mem = nullptr;
    hipError_t e = hipMalloc((void **)&mem, b_map + b_base + b_off + b_ent + b_prod);
    if (e != hipSuccess) return set_err(LH_ENOMEM, "lh_set_triads: hipMalloc: %s", hipGetErrorString(e));
    char *p = mem;
    e = hipMemcpy(p, map.data(), nslot * sizeof(int), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p + b_map, base.data(), nlist * sizeof(size_t), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p + b_map + b_base, off.data(), off.size() * sizeof(unsigned), hipMemcpyHostToDevice);
    if (e == hipSuccess && !ent.empty())
      e = hipMemcpy(p + b_map + b_base + b_off, ent.data(), ent.size() * sizeof(unsigned), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      (void)hipFree(mem);
      return set_err(LH_EHIP, "lh_set_triads: copy: %s", hipGetErrorString(e));
    }
    if (x->tri_mem[ch]) HIPCHK(hipFree(x->tri_mem[ch]));
    x->tri_mem[ch] = mem;
    h.dev.tri = 1;
    h.dev.tri_map = (const int *)p;
    h.dev.tri_base = (const size_t *)(p + b_map);
    h.dev.tri_off = (const unsigned *)(p + b_map + b_base);
    h.dev.tri_ent = (const unsigned *)(p + b_map + b_base + b_off);
    h.dev.tri_prod = (double2 *)(p + b_map + b_base + b_off + b_ent);
  }
  x->ntri = 0;
  for (auto &q : x->ch) x->ntri += q.dev.tri;
  rc = push_descriptors(x);
  if (rc) return rc;
  drop_graph(x);
  return LH_OK;
}

int lh_set_texture(lh_ctx *x, int ch, const int *bottom, const int *top) {
  int rc = check_ctx(x);
  if (rc) return rc;
  if (ch < 0 || ch >= (int)x->ch.size()) return set_err(LH_EINVAL, "lh_set_texture: bad channel %d", ch);
  HostChan &h = x->ch[ch];
  if ((bottom || top) && h.band >= 0)
    return set_err(LH_EINVAL, "lh_set_texture: channel %d is a band of a banded channel", ch);
  // the wall passes transform at lengths N1 and N2 (the N grid of
  // four_to_phys_N), which the LDS transforms support for 2^a 3^b only
  if ((bottom || top) && (!valid_len(h.cfg.N1) || !valid_len(h.cfg.N2)))
    return set_err(LH_EINVAL, "lh_set_texture: textured walls need N1, N2 of the form 2^a 3^b (got %d, %d)",
                   h.cfg.N1, h.cfg.N2);
  const size_t n = (size_t)h.cfg.N1 * h.cfg.N2;
  std::vector<int> m(2 * n, 1);
  const int *src[2] = {bottom, top};
  int bits = 0, xinv = 1;
  const int N1 = h.cfg.N1;
  for (int w = 0; w < 2; w++) {
    if (!src[w]) continue;
    for (size_t q = 0; q < n; q++) {
      if (src[w][q] != 1 && src[w][q] != 2)
        return set_err(LH_EINVAL, "lh_set_texture: mask codes are 1 (no-slip) or 2 (free-shear), got %d", src[w][q]);
      m[w * n + q] = src[w][q];
      if (src[w][q] != src[w][q - q % N1]) xinv = 0;
    }
    bits |= 1 << w;
  }
  HIPCHK(hipStreamSynchronize(x->stream));
  View d(x->arena, h.dev);
  HIPCHK(hipMemcpy(d.mask, m.data(), m.size() * sizeof(int), hipMemcpyHostToDevice));
  h.dev.tex = bits;
  h.xinv = xinv;
  tex_classify(x);
  rc = push_descriptors(x);
  if (rc) return rc;
  drop_graph(x);
  return LH_OK;
}


// ---------------------------------------------------------------------------
// learned closure (no counterpart in the reference)

static unsigned short host_bf16(float f) {
  unsigned u;
  memcpy(&u, &f, 4);
  return (unsigned short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
// compact parameter offsets: W1[128][135] b1 W2[128][128] b2 W3 b3 W4[3][128] b4[3]
static const int CC_W1 = 0, CC_B1 = CC_W1 + CL_NH * CL_NIN, CC_W2 = CC_B1 + CL_NH, CC_B2 = CC_W2 + CL_NH * CL_NH,
                 CC_W3 = CC_B2 + CL_NH, CC_B3 = CC_W3 + CL_NH * CL_NH, CC_W4 = CC_B3 + CL_NH,
                 CC_B4 = CC_W4 + CL_NOUT * CL_NH, CC_N = CC_B4 + CL_NOUT;
static_assert(CC_N == LH_CL_NPARAM, "parameter count");

static ClNorm make_norm(const float *n9) {
  ClNorm nm;
  for (int f = 0; f < 3; f++) {
    nm.mean[f] = n9[f];
    nm.inv_scale[f] = n9[3 + f];
    nm.out_scale[f] = n9[6 + f];
  }
  return nm;
}
static void cl_ensure(lh_ctx *x) {
  if (x->cl_mem.size() != x->ch.size()) {
    x->cl_mem.assign(x->ch.size(), nullptr);
    x->cl_on.assign(x->ch.size(), 0);
  }
}
static int cl_chan_ok(lh_ctx *x, int ch, const char *who) {
  if (ch < 0 || ch >= (int)x->ch.size()) return set_err(LH_EINVAL, "%s: bad channel %d", who, ch);
  const HostChan &h = x->ch[ch];
  if (h.cfg.N1 * h.cfg.N2 > CL_MAXPLANE)
    return set_err(LH_EINVAL, "%s: the closure runs on grids of at most %d modes (N1*N2 = %d)", who, CL_MAXPLANE,
                   h.cfg.N1 * h.cfg.N2);
  cl_ensure(x);
  return LH_OK;
}

int lh_closure_set(lh_ctx *x, const float *params, const float *norm9) {
  int rc = check_ctx(x);
  if (rc) return rc;
  if (!params || !norm9) return set_err(LH_EINVAL, "lh_closure_set: null argument");
  std::vector<unsigned short> fr((size_t)CLF_N * 64 * 8, 0);
  std::vector<float> bias(CL_NBIAS, 0.f);
  for (int lane = 0; lane < 64; lane++) {
    const int q = lane & 15, g = lane >> 4;
    for (int v = 0; v < 8; v++) {
      const int cs = v < 4 ? 4 * g + v : 16 + 4 * g + (v - 4);
      for (int b = 0; b < 8; b++)
        for (int st = 0; st < 5; st++) {
          const int row = 16 * b + q, col = 32 * st + 8 * g + v;
          const float w = col < CL_NIN ? params[CC_W1 + row * CL_NIN + col] : 0.f;
          fr[((size_t)(CLF_L1 + b * 5 + st) * 64 + lane) * 8 + v] = host_bf16(w);
        }
      for (int L = 0; L < 2; L++)
        for (int b = 0; b < 8; b++)
          for (int st = 0; st < 4; st++) {
            const int row = 16 * b + q, col = 32 * st + cs;
            const float w = params[(L == 0 ? CC_W2 : CC_W3) + row * CL_NH + col];
            fr[((size_t)((L == 0 ? CLF_L2 : CLF_L3) + b * 4 + st) * 64 + lane) * 8 + v] = host_bf16(w);
          }
      for (int st = 0; st < 4; st++) {
        const int col = 32 * st + cs;
        const float w = q < CL_NOUT ? params[CC_W4 + q * CL_NH + col] : 0.f;
        fr[((size_t)(CLF_L4 + st) * 64 + lane) * 8 + v] = host_bf16(w);
      }
    }
  }
  for (int r = 0; r < CL_NH; r++) {
    bias[r] = params[CC_B1 + r];
    bias[CL_NH + r] = params[CC_B2 + r];
    bias[2 * CL_NH + r] = params[CC_B3 + r];
  }
  for (int o = 0; o < CL_NOUT; o++) bias[3 * CL_NH + o] = params[CC_B4 + o];
  HIPCHK(hipStreamSynchronize(x->stream));
  if (!x->cl_net) {
    hipError_t e = hipMalloc((void **)&x->cl_net, CL_FRAG_BYTES + CL_NBIAS * sizeof(float));
    if (e != hipSuccess) {
      x->cl_net = nullptr;
      return set_err(LH_ENOMEM, "hipMalloc: %s", hipGetErrorString(e));
    }
  }
  HIPCHK(hipMemcpy(x->cl_net, fr.data(), CL_FRAG_BYTES, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(x->cl_net + CL_FRAG_BYTES, bias.data(), CL_NBIAS * sizeof(float), hipMemcpyHostToDevice));
  x->cl_norm = make_norm(norm9);
  drop_graph(x);  // the normalisation is a kernel argument
  return LH_OK;
}

int lh_closure_enable(lh_ctx *x, int ch, int on) {
  int rc = check_ctx(x);
  if (rc) return rc;
  if ((rc = cl_chan_ok(x, ch, "lh_closure_enable"))) return rc;
  if (on < 0 || on > 2) return set_err(LH_EINVAL, "lh_closure_enable: on is 0, 1 or 2");
  if (on && x->ch[ch].band >= 0)
    return set_err(LH_EINVAL, "lh_closure_enable: channel %d is a band of a banded channel", ch);
  if (on && !x->cl_net) return set_err(LH_ESTATE, "lh_closure_enable: no network set (lh_closure_set)");
  if (on) {
    float *ph, *co;
    double2 *ta;
    if ((rc = cl_buffers(x, ch, &ph, &co, &ta))) return rc;
  }
  x->cl_on[ch] = on;
  x->ncl = 0;
  for (int c : x->cl_on) x->ncl += c != 0;
  drop_graph(x);
  return LH_OK;
}

int lh_closure_eval(lh_ctx *x, int ch, float *d_phys, float *d_corr) {
  int rc = check_ctx(x);
  if (rc) return rc;
  if ((rc = cl_chan_ok(x, ch, "lh_closure_eval"))) return rc;
  if (!d_phys) return set_err(LH_EINVAL, "lh_closure_eval: null d_phys");
  if (d_corr && !x->cl_net) return set_err(LH_ESTATE, "lh_closure_eval: no network set (lh_closure_set)");
  HostChan &h = x->ch[ch];
  View d(x->arena, h.dev);
  ClSrc src{{d.u1, d.u2, d.u3}, {h.nyu, h.nyv, h.nyu}};
  HIPCHK(launch_cl_phys(x->arena, x->d_chans + ch, h.dev, src, d_phys, x->stream));
  if (d_corr) {
    HIPCHK(hipMemsetAsync(d_corr, 0, cl_plane_floats(h) * sizeof(float), x->stream));
    HIPCHK(launch_cl_infer(x->cl_net, (const float *)(x->cl_net + CL_FRAG_BYTES), x->cl_norm, d_phys, d_corr,
                           h.cfg.N1, h.cfg.N2, h.nyu, h.nyv, x->stream));
  }
  HIPCHK(hipStreamSynchronize(x->stream));
  return LH_OK;
}

int lh_closure_sample(lh_ctx *f, int chf, lh_ctx *c, int chc, float *d_phys, float *d_tau) {
  int rc = check_ctx(f);
  if (rc) return rc;
  if (!c || !d_phys || !d_tau) return set_err(LH_EINVAL, "lh_closure_sample: null argument");
  if (c->device != f->device) return set_err(LH_EINVAL, "lh_closure_sample: contexts on different devices");
  if (chf < 0 || chf >= (int)f->ch.size()) return set_err(LH_EINVAL, "lh_closure_sample: bad fine channel %d", chf);
  if ((rc = cl_chan_ok(c, chc, "lh_closure_sample"))) return rc;
  if (f->ntri || c->ntri) return set_err(LH_EINVAL, "lh_closure_sample: FFT-path channels only");
  const HostChan &F = f->ch[chf];
  HostChan &C = c->ch[chc];
  if (F.cfg.ny != C.cfg.ny || F.cfg.N1 < C.cfg.N1 || F.cfg.N2 < C.cfg.N2 || F.cfg.alp != C.cfg.alp ||
      F.cfg.bet != C.cfg.bet)
    return set_err(LH_EINVAL, "lh_closure_sample: the fine channel must have the coarse one's box and ny and at least "
                              "its modes");
  for (int j = 0; j < F.nyu; j++)
    if (F.yu[j] != C.yu[j]) return set_err(LH_EINVAL, "lh_closure_sample: y grids differ (row %d)", j);
  float *ph, *co;
  double2 *tau;
  if ((rc = cl_buffers(c, chc, &ph, &co, &tau))) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  // fine: products, filtered full nonlinear term into tau, filtered fields
  // into the coarse state
  if ((rc = raw_nonlinear(f, 2))) return rc;
  HIPCHK(launch_cl_tau(f->arena, f->d_chans + chf, c->arena, c->d_chans + chc, C.dev, tau, 0, f->stream));
  HIPCHK(launch_cl_restrict_fields(f->arena, f->d_chans + chf, c->arena, c->d_chans + chc, C.dev, f->stream));
  HIPCHK(hipStreamSynchronize(f->stream));
  // coarse: its own nonlinear term of the filtered fields, subtracted
  if ((rc = raw_nonlinear(c, 2))) return rc;
  HIPCHK(launch_cl_tau(c->arena, c->d_chans + chc, c->arena, c->d_chans + chc, C.dev, tau, 1, c->stream));
  View d(c->arena, C.dev);
  const size_t pl = (size_t)C.ldc * C.nyu;
  ClSrc su{{d.u1, d.u2, d.u3}, {C.nyu, C.nyv, C.nyu}};
  ClSrc st{{tau, tau + pl, tau + 2 * pl}, {C.nyu, C.nyu, C.nyu}};
  HIPCHK(launch_cl_phys(c->arena, c->d_chans + chc, C.dev, su, d_phys, c->stream));
  HIPCHK(launch_cl_phys(c->arena, c->d_chans + chc, C.dev, st, d_tau, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return LH_OK;
}

int lh_closure_time(lh_ctx *x, int nrep, double *ms3) {
  int rc = check_ctx(x);
  if (rc) return rc;
  if (nrep < 1 || !ms3) return set_err(LH_EINVAL, "lh_closure_time: bad arguments");
  if (!x->ncl) return set_err(LH_ESTATE, "lh_closure_time: no channel has the closure on");
  hipEvent_t ev[4];
  for (int q = 0; q < 4; q++) HIPCHK(hipEventCreate(&ev[q]));
  for (int q = 0; q < 3; q++) ms3[q] = 0.0;
  for (int n = 0; n < nrep && rc == LH_OK; n++) {
    HIPCHK(hipEventRecord(ev[0], x->stream));
    for (int part = 0; part < 3 && rc == LH_OK; part++) {
      for (size_t c = 0; c < x->ch.size(); c++) {
        if (!x->cl_on[c]) continue;
        HostChan &h = x->ch[c];
        float *phys, *corr;
        double2 *tau;
        if ((rc = cl_buffers(x, (int)c, &phys, &corr, &tau))) break;
        View d(x->arena, h.dev);
        ClSrc src{{d.u1, d.u2, d.u3}, {h.nyu, h.nyv, h.nyu}};
        if (part == 0)
          HIPCHK(launch_cl_phys(x->arena, x->d_chans + c, h.dev, src, phys, x->stream));
        else if (part == 1)
          HIPCHK(launch_cl_infer(x->cl_net, (const float *)(x->cl_net + CL_FRAG_BYTES), x->cl_norm, phys, corr,
                                 h.cfg.N1, h.cfg.N2, h.nyu, h.nyv, x->stream));
        else
          HIPCHK(launch_cl_to_modes(x->arena, x->d_chans + c, h.dev, corr, x->stream));
      }
      HIPCHK(hipEventRecord(ev[part + 1], x->stream));
    }
    HIPCHK(hipEventSynchronize(ev[3]));
    for (int q = 0; q < 3; q++) {
      float t;
      HIPCHK(hipEventElapsedTime(&t, ev[q], ev[q + 1]));
      ms3[q] += t;
    }
  }
  for (int q = 0; q < 4; q++) (void)hipEventDestroy(ev[q]);
  for (int q = 0; q < 3; q++) ms3[q] /= nrep;
  return rc;
}

// ---------------------------------------------------------------------------
// closure training: the network's forward and backward passes as bf16
// matrix-core GEMMs (fp32 accumulation) over a batch of B points

struct lh_mlp {
  int B = 0, device = 0, step = 0;
  hipStream_t stream = nullptr;
  char *mem = nullptr;
  float *P, *G, *Mo, *Vo;  // padded master parameters, gradient, AdamW moments
  uint16_t *Wb;            // bf16 copies of the matrices
  uint16_t *X, *H[3], *dY, *dZ;
  float *T, *Z[3], *Y, *dH;
  double *loss;
  int *pts;
  ClNorm norm{};
};

static int mlp_ok(lh_mlp *m) {
  if (!m) return set_err(LH_EINVAL, "null trainer");
  HIPCHK(hipSetDevice(m->device));
  return LH_OK;
}

static int gemm(lh_mlp *m, const uint16_t *A, long sam, long sak, const uint16_t *B, long sbk, long sbn, float *C,
                long scm, int M, int N, int K, int kchunk, int accum) {
  GemmP p{A, B, C, M, N, K, sam, sak, sbk, sbn, scm, 1, kchunk, accum};
  HIPCHK(launch_gemm(p, m->stream));
  return LH_OK;
}

static int mlp_forward(lh_mlp *m) {
  const int B = m->B;
  int rc;
  const uint16_t *in = m->X;
  int K = CL_KIN;
  const int wo[3] = {CLB_W1, CLB_W2, CLB_W3}, bo[3] = {CLP_B1, CLP_B2, CLP_B3};
  for (int L = 0; L < 3; L++) {
    if ((rc = gemm(m, m->Wb + wo[L], K, 1, in, B, 1, m->Z[L], B, CL_NH, B, K, K, 0))) return rc;
    HIPCHK(launch_bias_gelu(m->Z[L], m->P + bo[L], m->H[L], CL_NH, B, m->stream));
    in = m->H[L];
    K = CL_NH;
  }
  return gemm(m, m->Wb + CLB_W4, CL_NH, 1, in, B, 1, m->Y, B, CL_MOUT, B, CL_NH, CL_NH, 0);
}

int lh_mlp_create(int batch, lh_mlp **out) {
  if (!out || batch < 64 || batch % 64 || batch > (1 << 22))
    return set_err(LH_EINVAL, "lh_mlp_create: batch must be a multiple of 64 in [64, 2^22]");
  *out = nullptr;
  lh_mlp *m = new lh_mlp();
  m->B = batch;
  hipError_t e = hipGetDevice(&m->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete m;
    return set_err(LH_EHIP, "lh_mlp_create: %s", hipGetErrorString(e));
  }
  const size_t B = batch;
  size_t off = 0;
  size_t o_P = off; off += al<float>(CLP_N);
  size_t o_G = off; off += al<float>(CLP_N);
  size_t o_M = off; off += al<float>(CLP_N);
  size_t o_V = off; off += al<float>(CLP_N);
  size_t o_Wb = off; off += al<uint16_t>(CLB_N);
  size_t o_X = off; off += al<uint16_t>(CL_KIN * B);
  size_t o_T = off; off += al<float>(CL_NOUT * B);
  size_t o_Z = off; off += 3 * al<float>(CL_NH * B);
  size_t o_H = off; off += 3 * al<uint16_t>(CL_NH * B);
  size_t o_Y = off; off += al<float>(CL_MOUT * B);
  size_t o_dY = off; off += al<uint16_t>(CL_MOUT * B);
  size_t o_dH = off; off += al<float>(CL_NH * B);
  size_t o_dZ = off; off += al<uint16_t>(CL_NH * B);
  size_t o_loss = off; off += al<double>(1);
  size_t o_pts = off; off += al<int>(B);
  e = hipMalloc((void **)&m->mem, off);
  if (e != hipSuccess) {
    (void)hipStreamDestroy(m->stream);
    delete m;
    return set_err(LH_ENOMEM, "lh_mlp_create: hipMalloc(%zu): %s", off, hipGetErrorString(e));
  }
  (void)hipMemset(m->mem, 0, off);
  m->P = (float *)(m->mem + o_P);
  m->G = (float *)(m->mem + o_G);
  m->Mo = (float *)(m->mem + o_M);
  m->Vo = (float *)(m->mem + o_V);
  m->Wb = (uint16_t *)(m->mem + o_Wb);
  m->X = (uint16_t *)(m->mem + o_X);
  m->T = (float *)(m->mem + o_T);
  for (int L = 0; L < 3; L++) {
    m->Z[L] = (float *)(m->mem + o_Z + L * al<float>(CL_NH * B));
    m->H[L] = (uint16_t *)(m->mem + o_H + L * al<uint16_t>(CL_NH * B));
  }
  m->Y = (float *)(m->mem + o_Y);
  m->dY = (uint16_t *)(m->mem + o_dY);
  m->dH = (float *)(m->mem + o_dH);
  m->dZ = (uint16_t *)(m->mem + o_dZ);
  m->loss = (double *)(m->mem + o_loss);
  m->pts = (int *)(m->mem + o_pts);
  *out = m;
  return LH_OK;
}

int lh_mlp_destroy(lh_mlp *m) {
  if (!m) return LH_OK;
  (void)hipSetDevice(m->device);
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  if (m->mem) (void)hipFree(m->mem);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
  return LH_OK;
}

int lh_mlp_set_params(lh_mlp *m, const float *params, const float *norm9) {
  int rc = mlp_ok(m);
  if (rc) return rc;
  if (!params || !norm9) return set_err(LH_EINVAL, "lh_mlp_set_params: null argument");
  std::vector<float> P(CLP_N, 0.f);
  for (int r = 0; r < CL_NH; r++) {
    for (int i = 0; i < CL_NIN; i++) P[CLP_W1 + r * CL_KIN + i] = params[CC_W1 + r * CL_NIN + i];
    for (int i = 0; i < CL_NH; i++) {
      P[CLP_W2 + r * CL_NH + i] = params[CC_W2 + r * CL_NH + i];
      P[CLP_W3 + r * CL_NH + i] = params[CC_W3 + r * CL_NH + i];
    }
    P[CLP_B1 + r] = params[CC_B1 + r];
    P[CLP_B2 + r] = params[CC_B2 + r];
    P[CLP_B3 + r] = params[CC_B3 + r];
  }
  for (int o = 0; o < CL_NOUT; o++) {
    for (int i = 0; i < CL_NH; i++) P[CLP_W4 + o * CL_NH + i] = params[CC_W4 + o * CL_NH + i];
    P[CLP_B4 + o] = params[CC_B4 + o];
  }
  HIPCHK(hipStreamSynchronize(m->stream));
  HIPCHK(hipMemcpy(m->P, P.data(), CLP_N * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(hipMemset(m->Mo, 0, CLP_N * sizeof(float)));
  HIPCHK(hipMemset(m->Vo, 0, CLP_N * sizeof(float)));
  m->step = 0;
  m->norm = make_norm(norm9);
  HIPCHK(launch_pack_bf16(m->P, m->Wb, m->stream));
  HIPCHK(hipStreamSynchronize(m->stream));
  return LH_OK;
}

static void padded_to_compact(const std::vector<float> &P, float *out) {
  for (int r = 0; r < CL_NH; r++) {
    for (int i = 0; i < CL_NIN; i++) out[CC_W1 + r * CL_NIN + i] = P[CLP_W1 + r * CL_KIN + i];
    for (int i = 0; i < CL_NH; i++) {
      out[CC_W2 + r * CL_NH + i] = P[CLP_W2 + r * CL_NH + i];
      out[CC_W3 + r * CL_NH + i] = P[CLP_W3 + r * CL_NH + i];
    }
    out[CC_B1 + r] = P[CLP_B1 + r];
    out[CC_B2 + r] = P[CLP_B2 + r];
    out[CC_B3 + r] = P[CLP_B3 + r];
  }
  for (int o = 0; o < CL_NOUT; o++) {
    for (int i = 0; i < CL_NH; i++) out[CC_W4 + o * CL_NH + i] = P[CLP_W4 + o * CL_NH + i];
    out[CC_B4 + o] = P[CLP_B4 + o];
  }
}

int lh_mlp_get_params(lh_mlp *m, float *params) {
  int rc = mlp_ok(m);
  if (rc) return rc;
  if (!params) return set_err(LH_EINVAL, "lh_mlp_get_params: null argument");
  std::vector<float> P(CLP_N);
  HIPCHK(hipStreamSynchronize(m->stream));
  HIPCHK(hipMemcpy(P.data(), m->P, CLP_N * sizeof(float), hipMemcpyDeviceToHost));
  padded_to_compact(P, params);
  return LH_OK;
}

int lh_mlp_get_grad(lh_mlp *m, float *grad) {
  int rc = mlp_ok(m);
  if (rc) return rc;
  if (!grad) return set_err(LH_EINVAL, "lh_mlp_get_grad: null argument");
  std::vector<float> P(CLP_N);
  HIPCHK(hipStreamSynchronize(m->stream));
  HIPCHK(hipMemcpy(P.data(), m->G, CLP_N * sizeof(float), hipMemcpyDeviceToHost));
  padded_to_compact(P, grad);
  return LH_OK;
}

int lh_mlp_set_batch(lh_mlp *m, const float *x, const float *t) {
  int rc = mlp_ok(m);
  if (rc) return rc;
  if (!x || !t) return set_err(LH_EINVAL, "lh_mlp_set_batch: null argument");
  const size_t B = m->B;
  std::vector<unsigned short> X(CL_KIN * B, 0);
  std::vector<float> T(CL_NOUT * B);
  for (size_t n = 0; n < B; n++) {
    for (int f = 0; f < CL_NIN; f++) X[f * B + n] = host_bf16(x[n * CL_NIN + f]);
    for (int o = 0; o < CL_NOUT; o++) T[o * B + n] = t[n * CL_NOUT + o];
  }
  HIPCHK(hipStreamSynchronize(m->stream));
  HIPCHK(hipMemcpy(m->X, X.data(), X.size() * 2, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(m->T, T.data(), T.size() * 4, hipMemcpyHostToDevice));
  return LH_OK;
}

int lh_mlp_sample(lh_mlp *m, int N1, int N2, int ny, const float *d_phys, const float *d_tau, int nsnap,
                  uint64_t seed) {
  int rc = mlp_ok(m);
  if (rc) return rc;
  if (!d_phys || !d_tau || nsnap < 1 || N1 < 4 || N2 < 4 || ny < 4)
    return set_err(LH_EINVAL, "lh_mlp_sample: bad arguments");
  HIPCHK(launch_cl_batch(d_phys, d_tau, nsnap, N1, N2, ny + 1, ny, m->norm, seed, m->X, m->T, m->B, m->stream));
  return LH_OK;
}

int lh_mlp_set_points(lh_mlp *m, int N1, int N2, int ny, const float *d_phys, const int *pts) {
  int rc = mlp_ok(m);
  if (rc) return rc;
  if (!d_phys || !pts || N1 < 4 || N2 < 4 || ny < 4) return set_err(LH_EINVAL, "lh_mlp_set_points: bad arguments");
  const int npts = (ny - 1) * N1 * N2;
  for (int n = 0; n < m->B; n++)
    if (pts[n] < 0 || pts[n] >= npts) return set_err(LH_EINVAL, "lh_mlp_set_points: point %d out of range", pts[n]);
  HIPCHK(hipStreamSynchronize(m->stream));
  HIPCHK(hipMemcpy(m->pts, pts, m->B * sizeof(int), hipMemcpyHostToDevice));
  HIPCHK(launch_cl_features(d_phys, N1, N2, ny + 1, ny, m->norm, m->pts, m->X, m->B, m->stream));
  HIPCHK(hipMemsetAsync(m->T, 0, (size_t)CL_NOUT * m->B * sizeof(float), m->stream));
  return LH_OK;
}

int lh_mlp_forward(lh_mlp *m, float *y) {
  int rc = mlp_ok(m);
  if (rc) return rc;
  if (!y) return set_err(LH_EINVAL, "lh_mlp_forward: null argument");
  if ((rc = mlp_forward(m))) return rc;
  const size_t B = m->B;
  std::vector<float> Y(CL_NOUT * B), b4(CL_NOUT);
  HIPCHK(hipStreamSynchronize(m->stream));
  HIPCHK(hipMemcpy(Y.data(), m->Y, Y.size() * 4, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(b4.data(), m->P + CLP_B4, CL_NOUT * 4, hipMemcpyDeviceToHost));
  for (size_t n = 0; n < B; n++)
    for (int o = 0; o < CL_NOUT; o++) y[n * CL_NOUT + o] = Y[o * B + n] + b4[o];
  return LH_OK;
}

int lh_mlp_grad(lh_mlp *m, double *loss) {
  int rc = mlp_ok(m);
  if (rc) return rc;
  const int B = m->B;
  HIPCHK(hipMemsetAsync(m->G, 0, CLP_N * sizeof(float), m->stream));
  HIPCHK(hipMemsetAsync(m->loss, 0, sizeof(double), m->stream));
  if ((rc = mlp_forward(m))) return rc;
  const float scale = 1.f / (3.f * (float)B);
  HIPCHK(launch_out_loss(m->Y, m->P + CLP_B4, m->T, m->dY, m->loss, B, scale, m->stream));
  const int KC = 256;  // points per block of the weight-gradient GEMMs
  // output layer
  if ((rc = gemm(m, m->dY, B, 1, m->H[2], 1, B, m->G + CLP_W4, CL_NH, CL_MOUT, CL_NH, B, KC, 1))) return rc;
  HIPCHK(launch_rowsum(m->dY, m->G + CLP_B4, CL_MOUT, B, m->stream));
  if ((rc = gemm(m, m->Wb + CLB_W4, 1, CL_NH, m->dY, B, 1, m->dH, B, CL_NH, B, CL_MOUT, CL_MOUT, 0))) return rc;
  // hidden layers 3, 2, 1
  const int wo[3] = {CLB_W1, CLB_W2, CLB_W3}, gw[3] = {CLP_W1, CLP_W2, CLP_W3}, gb[3] = {CLP_B1, CLP_B2, CLP_B3};
  for (int L = 2; L >= 0; L--) {
    HIPCHK(launch_dgelu(m->dH, m->Z[L], m->dZ, CL_NH, B, m->stream));
    const uint16_t *in = L ? m->H[L - 1] : m->X;
    const int K = L ? CL_NH : CL_KIN;
    if ((rc = gemm(m, m->dZ, B, 1, in, 1, B, m->G + gw[L], K, CL_NH, K, B, KC, 1))) return rc;
    HIPCHK(launch_rowsum(m->dZ, m->G + gb[L], CL_NH, B, m->stream));
    if (L)
      if ((rc = gemm(m, m->Wb + wo[L], 1, CL_NH, m->dZ, B, 1, m->dH, B, CL_NH, B, CL_NH, CL_NH, 0))) return rc;
  }
  double l = 0.0;
  HIPCHK(hipMemcpyAsync(&l, m->loss, sizeof(double), hipMemcpyDeviceToHost, m->stream));
  HIPCHK(hipStreamSynchronize(m->stream));
  if (loss) *loss = l / (3.0 * B);
  return LH_OK;
}

int lh_mlp_adam(lh_mlp *m, float lr, float beta1, float beta2, float eps, float wd) {
  int rc = mlp_ok(m);
  if (rc) return rc;
  m->step += 1;
  const float bc1 = 1.f - powf(beta1, (float)m->step), bc2 = 1.f - powf(beta2, (float)m->step);
  HIPCHK(launch_adam(m->P, m->G, m->Mo, m->Vo, CLP_N, lr, beta1, beta2, eps, wd, bc1, bc2, m->stream));
  HIPCHK(launch_pack_bf16(m->P, m->Wb, m->stream));
  return LH_OK;
}

int lh_mlp_sync(lh_mlp *m) {
  int rc = mlp_ok(m);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(m->stream));
  return LH_OK;
}

}  // extern "C"
