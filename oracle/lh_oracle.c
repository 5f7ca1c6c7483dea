/*
 * lh_oracle.c -- CPU FP64 restatement of the plain-FFT RK3 step of LittleHarsh
 * (bananapancake12/RNL_MULTIBLOCK_NOFFT), for one channel whose bands share
 * one x/z resolution, run as a single domain (the reference's np=1 view).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Each function cites the reference routine it restates.  The nonlinear term
 * is the dormant plain path `ops_in_planes2` (FOU3D.f90:2090-2251), which the
 * reference would run if FOU3D.f90:245 called it instead of the triad list.
 * Deviations, each deliberate and tested:
 *   - the flux planes uv/wv at the two wall faces and vv at the two ghost
 *     centres are zero (the active path zeroes them, FOU3D.f90:1752,1891-1894;
 *     the plain path never writes them);
 *   - dtc_calc takes the max over the transformed physical planes, without
 *     the one-plane shift the reference's explicit-shape dummy causes
 *     (FOU3D.f90:708-716 vs start.f90:567-569);
 *   - optional fixed dt (the reference only has the adaptive dt).
 */
#include "lh_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

/* ------------------------------------------------------------------------ */
/* Mixed-radix complex FFT (own implementation: recursive decimation in time,
 * radices 4, 2, 3, 5).  Sign convention as FFTPACK: forward e^{-i}, backward
 * e^{+i}, both unscaled (cfftf/cfftb, cft_buff.f:186-279,631-700).          */

#define MAXFAC 32
typedef struct {
  int n, nf;
  int fac[MAXFAC];
  cplx *w; /* w[m] = exp(-2 pi i m / n) */
} fftplan;

static void plan_init(fftplan *p, int n) {
  p->n = n;
  p->nf = 0;
  int r = n;
  const int tryr[4] = {4, 2, 3, 5};
  for (int t = 0; t < 4 && r > 1; t++) {
    while (r % tryr[t] == 0 && !(tryr[t] == 4 && r % 4 != 0)) {
      p->fac[p->nf++] = tryr[t];
      r /= tryr[t];
    }
  }
  if (r != 1) { /* generic leftover factor (not used by this solver) */
    p->fac[p->nf++] = r;
  }
  p->w = (cplx *)malloc(sizeof(cplx) * (size_t)n);
  for (int m = 0; m < n; m++) {
    double a = -2.0 * M_PI * (double)m / (double)n;
    p->w[m] = cos(a) + I * sin(a);
  }
}
static void plan_free(fftplan *p) { free(p->w); }

static void fft_rec(const cplx *in, cplx *out, int n, int sin_, const int *fac,
                    const fftplan *pl, int wstr, int sign) {
  if (n == 1) {
    out[0] = in[0];
    return;
  }
  int p = fac[0], m = n / p;
  for (int q = 0; q < p; q++)
    fft_rec(in + (size_t)q * sin_, out + (size_t)q * m, m, sin_ * p, fac + 1, pl,
            wstr * p, sign);
  const int N = pl->n;
  cplx t[MAXFAC > 64 ? MAXFAC : 64];
  for (int k = 0; k < m; k++) {
    for (int q = 0; q < p; q++) {
      cplx w = pl->w[((size_t)q * k * wstr) % N];
      if (sign > 0) w = conj(w);
      t[q] = out[(size_t)q * m + k] * w;
    }
    for (int s2 = 0; s2 < p; s2++) {
      cplx acc = 0;
      for (int q = 0; q < p; q++) {
        cplx w = pl->w[((size_t)q * s2 * (N / p)) % N];
        if (sign > 0) w = conj(w);
        acc += t[q] * w;
      }
      out[(size_t)s2 * m + k] = acc;
    }
  }
}

/* plan cache (small n only) */
#define PLANCACHE 16
static fftplan g_plans[PLANCACHE];
static int g_nplans = 0;
static const fftplan *get_plan(int n) {
  static __thread const fftplan *last = NULL;
  if (last && last->n == n) return last;
  const fftplan *r = NULL;
#pragma omp critical(lho_plan)
  {
    for (int i = 0; i < g_nplans; i++)
      if (g_plans[i].n == n) r = &g_plans[i];
    if (!r && g_nplans < PLANCACHE) {
      plan_init(&g_plans[g_nplans], n);
      r = &g_plans[g_nplans++];
    }
  }
  last = r;
  return r;
}

/* x (n complex, unit stride) -> transform in place; sign -1 forward, +1 back */
static void cfft(cplx *x, int n, int sign) {
  const fftplan *pl = get_plan(n);
  cplx buf[1024];
  cplx *tmp = n <= 1024 ? buf : (cplx *)malloc(sizeof(cplx) * (size_t)n);
  fft_rec(x, tmp, n, 1, pl->fac, pl, 1, sign);
  memcpy(x, tmp, sizeof(cplx) * (size_t)n);
  if (tmp != buf) free(tmp);
}

/* cft(c,ise,isa,m,iopt,aux), cft_buff.f:1-65: m complex vectors of length n,
 * element stride ise and vector stride isa (both in reals); forward (iopt<0)
 * scaled by 1/n, inverse unscaled. */
void lho_cft(double *c, int n, int ise, int isa, int m, int iopt) {
  cplx buf[1024];
  cplx *v = n <= 1024 ? buf : (cplx *)malloc(sizeof(cplx) * (size_t)n);
  for (int k = 0; k < m; k++) {
    double *base = c + (size_t)k * isa;
    for (int j = 0; j < n; j++) v[j] = base[(size_t)j * ise] + I * base[(size_t)j * ise + 1];
    cfft(v, n, iopt < 0 ? -1 : +1);
    double s = iopt < 0 ? 1.0 / n : 1.0;
    for (int j = 0; j < n; j++) {
      base[(size_t)j * ise] = creal(v[j]) * s;
      base[(size_t)j * ise + 1] = cimag(v[j]) * s;
    }
  }
  if (v != buf) free(v);
}

/* rft(c,isa,m,iopt,aux), rft_buff.f:24-90: m real vectors of length n stored
 * with stride isa >= n+2.  Fourier side: cc(k) = c(2k+1) + i c(2k+2), k=0..n/2.
 * Forward: cc = DFT(x)/n with imag(cc(0)) = imag(cc(n/2)) = 0.  Inverse:
 * x_j = sum over the Hermitian extension of cc, with imag(cc(0)) and
 * imag(cc(n/2)) ignored; c(n+1), c(n+2) are left as they were. */
void lho_rft(double *c, int n, int isa, int m, int iopt) {
  cplx buf[1024];
  cplx *v = n <= 1024 ? buf : (cplx *)malloc(sizeof(cplx) * (size_t)n);
  for (int k = 0; k < m; k++) {
    double *x = c + (size_t)k * isa;
    if (iopt < 0) {
      for (int j = 0; j < n; j++) v[j] = x[j];
      cfft(v, n, -1);
      double s = 1.0 / n;
      x[0] = creal(v[0]) * s;
      x[1] = 0.0;
      for (int q = 1; q < n / 2; q++) {
        x[2 * q] = creal(v[q]) * s;
        x[2 * q + 1] = cimag(v[q]) * s;
      }
      x[n] = creal(v[n / 2]) * s;
      x[n + 1] = 0.0;
    } else {
      v[0] = x[0];
      for (int q = 1; q < n / 2; q++) {
        v[q] = x[2 * q] + I * x[2 * q + 1];
        v[n - q] = conj(v[q]);
      }
      v[n / 2] = x[n];
      cfft(v, n, +1);
      for (int j = 0; j < n; j++) x[j] = creal(v[j]);
    }
  }
  if (v != buf) free(v);
}

/* ------------------------------------------------------------------------ */

struct lho_state {
  lho_cfg c;
  int nyu, nyv;     /* rows of u grid (ny+1) and v grid (ny) */
  int nxh, ncol;    /* N1/2+1, nxh*N2 */
  int igal, kgal;   /* Ngal1+2, Ngal2 */
  double *yu, *yv, *dthdyu, *dthdyv, *dyu2i, *dyv2i; /* dyu2i[3*j+m] */
  double gw1, gw2, gwbc1, gwbc3, gwi1, gwi2;
  double ddthetavi;
  double *kx, *kz, *k2x, *k2z; /* kx[i]: imag part of k1F_x; kz[k] of k1F_z */
  double aRK[3], bRK[3], gRK[3], cRK[3], dRK[3];
  double t, dt, dti, mpgx, QxT, Qx, dtv, dtc;
  int iter;
  /* fields: f[col*rows + j] with rows = nyu (u and p like) or nyv (v like) */
  cplx *u1, *u2, *u3, *p, *Nu1, *Nu2, *Nu3, *du1, *du2, *du3, *psi, *div;
  cplx *u1i, *u2i, *u3i, *uvf, *wvf, *vvc;
  double *dg; /* prefactored Poisson LU, dg[(col*nyu + j)*3 + m] */
  /* wall textures: tex[w][z*N1 + x] = 1 no-slip, 2 free-shear (planeBC
   * codes, start.f90:1776-1829), w = 0 bottom, 1 top; NULL = smooth wall */
  int *tex[2];
  /* triad lists of the restricted nonlinear path (nonlinRead,
   * start.f90:2983-3069): tri_len[q*9 + t-1] rows of type t in file q,
   * tri_rows[q*9 + t-1] their int8 (len, 4) Fortran-order data; tri_u/tri_v
   * the file of each u plane j = 1..ny-1 / v plane j = 1..ny-2 (-1: none) */
  int ntri;
  int *tri_len;
  signed char **tri_rows;
  int *tri_u, *tri_v;
};

static double ipow(double x, int n) {
  double r = 1.0;
  for (int i = 0; i < n; i++) r *= x;
  return r;
}

int lho_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}

/* y(theta) and dy/dtheta of the wall-normal mapping */
static void ymap(const lho_cfg *c, double aaa, double bbb, double qqq, double th,
                 double *y, double *dthdy) {
  if (c->grid_kind == 1) {
    double g = c->gamma, tg = tanh(g);
    double s = g * (th / qqq - 1.0);
    double ts = tanh(s);
    *y = ts / tg;
    *dthdy = 1.0 / ((g / (qqq * tg)) * (1.0 - ts * ts));
  } else {
    int p = c->ppp;
    double x = th - qqq;
    *y = aaa * x + bbb * ipow(x, p);
    *dthdy = 1.0 / (aaa + p * bbb * ipow(x, p - 1));
  }
}

/* derived quantities once yu,yv,dthdyu,dthdyv are known (start.f90:537-559,
 * :943-987 second-derivative coefficients, :456 dtv) */
static void finish_grid(lho_state *s) {
  int nyu = s->nyu, nyv = s->nyv, nn = nyv - 1; /* nn: last v index */
  double *yu = s->yu, *yv = s->yv;
  double *a = s->dyu2i, *b = s->dyv2i;
  /* j = 0 */
  a[0] = 0.0;
  a[1] = -2.0 / ((yu[1] - yu[0]) * (yu[1] - yu[0]));
  a[2] = 1.0 / ((yu[1] - yu[0]) * (yu[1] - yu[0]));
  b[0] = 0.0;
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
    b[3 * j + 2] = 0.0;
  }
  {
    int j = nn + 1;
    double d = (yu[j] - yu[j - 1]) * (yu[j] - yu[j - 1]);
    a[3 * j + 0] = 1.0 / d;
    a[3 * j + 1] = -2.0 / d;
    a[3 * j + 2] = 0.0;
  }
  int ju = nyu - 1, jv = nyv - 1;
  s->gw1 = -(yv[0] - yu[0]) / (yv[0] - yu[1]);
  s->gw2 = (yu[ju] - yv[jv]) / (yv[jv] - yu[ju - 1]);
  double ddv = s->ddthetavi;
  s->gwbc1 = ((yu[0] - yv[0]) + s->c.bslpu1 * ddv * s->dthdyv[0] * (yu[1] - yu[0])) /
             ((yu[1] - yv[0]) + s->c.bslpu1 * ddv * s->dthdyv[0] * (yu[1] - yu[0]));
  s->gwbc3 = ((yu[0] - yv[0]) + s->c.bslpu3 * ddv * s->dthdyv[0] * (yu[1] - yu[0])) /
             ((yu[1] - yv[0]) + s->c.bslpu3 * ddv * s->dthdyv[0] * (yu[1] - yu[0]));
  s->gwi1 = (yu[0] - yu[1]) / (yv[0] - yu[1]);
  s->gwi2 = (yu[ju] - yu[ju - 1]) / (yv[jv] - yu[ju - 1]);
  s->dtv = s->c.Re / (-s->k2x[s->c.N1 / 2] - s->k2z[s->c.N2 / 2 - 1] +
                      4.0 / ((yv[1] - yv[0]) * (yv[1] - yv[0])));
}

static void build_dg(lho_state *s);

lho_state *lho_create(const lho_cfg *cfg) {
  lho_state *s = (lho_state *)calloc(1, sizeof(lho_state));
  s->c = *cfg;
  int ny = cfg->ny;
  s->nyv = ny;
  s->nyu = ny + 1;
  s->nxh = cfg->N1 / 2 + 1;
  s->ncol = s->nxh * cfg->N2;
  s->igal = cfg->Ngal1 + 2;
  s->kgal = cfg->Ngal2;
  s->ddthetavi = 1.0;
  s->yu = calloc(s->nyu, sizeof(double));
  s->yv = calloc(s->nyv, sizeof(double));
  s->dthdyu = calloc(s->nyu, sizeof(double));
  s->dthdyv = calloc(s->nyv, sizeof(double));
  s->dyu2i = calloc(3 * (size_t)s->nyu, sizeof(double));
  s->dyv2i = calloc(3 * (size_t)s->nyv, sizeof(double));
  /* def_k, start.f90:1079-1136 */
  s->kx = calloc(s->nxh, sizeof(double));
  s->k2x = calloc(s->nxh, sizeof(double));
  s->kz = calloc(cfg->N2, sizeof(double));
  s->k2z = calloc(cfg->N2, sizeof(double));
  for (int i = 0; i < s->nxh; i++) {
    s->kx[i] = cfg->alp * i;
    double t = cfg->alp * i;
    s->k2x[i] = -(t * t);
  }
  for (int k = 1; k <= cfg->N2 / 2; k++) {
    double t = cfg->bet * (k - 1);
    s->kz[k - 1] = t;
    s->k2z[k - 1] = -(t * t);
  }
  for (int k = cfg->N2 / 2 + 1; k <= cfg->N2; k++) {
    double t = cfg->bet * (cfg->N2 - k + 1);
    s->kz[k - 1] = -t;
    s->k2z[k - 1] = -(t * t);
  }
  /* RK3 coefficients (Le & Moin), start.f90:419-433 */
  const double a[3] = {4.0 / 15.0, 1.0 / 15.0, 1.0 / 6.0};
  const double g[3] = {8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0};
  const double cc[3] = {0.0, -17.0 / 60.0, -5.0 / 12.0};
  const double d[3] = {8.0 / 15.0, 2.0 / 15.0, 1.0 / 3.0};
  for (int q = 0; q < 3; q++) {
    s->aRK[q] = a[q];
    s->bRK[q] = a[q];
    s->gRK[q] = g[q];
    s->cRK[q] = cc[q];
    s->dRK[q] = d[q];
  }
  /* ygrid, start.f90:693-1077 ('Wall' grid, no immersed region) */
  int nn = ny - 1;
  double qqq = nn / 2.0, aaa = 0, bbb = 0;
  if (cfg->grid_kind == 0) {
    int p = cfg->ppp;
    double dyq = cfg->dyq;
    bbb = (dyq - 1.0) /
          (qqq * (1.0 - ipow(qqq, p - 1) +
                  dyq * ((1.0 - qqq) * ipow(qqq, p - 1) - ipow(1.0 - qqq, p))));
    aaa = (1.0 - bbb * ipow(qqq, p)) / qqq;
  }
  ymap(cfg, aaa, bbb, qqq, -0.5, &s->yu[0], &s->dthdyu[0]);
  for (int j = 0; j <= nn; j++) {
    ymap(cfg, aaa, bbb, qqq, (double)j, &s->yv[j], &s->dthdyv[j]);
    ymap(cfg, aaa, bbb, qqq, j + 0.5, &s->yu[j + 1], &s->dthdyu[j + 1]);
  }
  finish_grid(s);

  size_t nu = (size_t)s->ncol * s->nyu, nv = (size_t)s->ncol * s->nyv;
  s->u1 = calloc(nu, sizeof(cplx));
  s->u3 = calloc(nu, sizeof(cplx));
  s->u2 = calloc(nv, sizeof(cplx));
  s->p = calloc(nu, sizeof(cplx));
  s->Nu1 = calloc(nu, sizeof(cplx));
  s->Nu3 = calloc(nu, sizeof(cplx));
  s->Nu2 = calloc(nv, sizeof(cplx));
  s->du1 = calloc(nu, sizeof(cplx));
  s->du3 = calloc(nu, sizeof(cplx));
  s->du2 = calloc(nv, sizeof(cplx));
  s->psi = calloc(nu, sizeof(cplx));
  s->div = calloc(nu, sizeof(cplx));
  s->u1i = calloc(nv, sizeof(cplx));
  s->u3i = calloc(nv, sizeof(cplx));
  s->u2i = calloc(nu, sizeof(cplx));
  s->uvf = calloc(nv, sizeof(cplx));
  s->wvf = calloc(nv, sizeof(cplx));
  s->vvc = calloc(nu, sizeof(cplx));
  s->dg = calloc(3 * nu, sizeof(double));
  build_dg(s);
  s->mpgx = cfg->mpgx;
  s->dt = cfg->dt;
  s->dti = cfg->dt > 0 ? 1.0 / cfg->dt : 0.0;
  return s;
}

void lho_destroy(lho_state *s) {
  if (!s) return;
  void *ptrs[] = {s->yu, s->yv, s->dthdyu, s->dthdyv, s->dyu2i, s->dyv2i, s->kx, s->kz,
                  s->k2x, s->k2z, s->u1, s->u2, s->u3, s->p, s->Nu1, s->Nu2, s->Nu3,
                  s->du1, s->du2, s->du3, s->psi, s->div, s->u1i, s->u2i, s->u3i,
                  s->uvf, s->wvf, s->vvc, s->dg, s->tex[0], s->tex[1]};
  for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); i++) free(ptrs[i]);
  lho_set_triads(s, 0, NULL, NULL, NULL, NULL);
  free(s);
}

int lho_nrows_u(const lho_state *s) { return s->nyu; }
int lho_ncol(const lho_state *s) { return s->ncol; }

void lho_get_grid(const lho_state *s, double *yu, double *yv, double *dthdyu,
                  double *dthdyv, double *dyu2i, double *dyv2i) {
  if (yu) memcpy(yu, s->yu, sizeof(double) * s->nyu);
  if (yv) memcpy(yv, s->yv, sizeof(double) * s->nyv);
  if (dthdyu) memcpy(dthdyu, s->dthdyu, sizeof(double) * s->nyu);
  if (dthdyv) memcpy(dthdyv, s->dthdyv, sizeof(double) * s->nyv);
  if (dyu2i) memcpy(dyu2i, s->dyu2i, sizeof(double) * 3 * s->nyu);
  if (dyv2i) memcpy(dyv2i, s->dyv2i, sizeof(double) * 3 * s->nyv);
}

void lho_set_grid(lho_state *s, const double *yu, const double *yv, const double *dthdyu,
                  const double *dthdyv) {
  memcpy(s->yu, yu, sizeof(double) * s->nyu);
  memcpy(s->yv, yv, sizeof(double) * s->nyv);
  memcpy(s->dthdyu, dthdyu, sizeof(double) * s->nyu);
  memcpy(s->dthdyv, dthdyv, sizeof(double) * s->nyv);
  finish_grid(s);
  build_dg(s);
}

void lho_get_scalars(const lho_state *s, double *o) {
  o[0] = s->t;
  o[1] = s->dt;
  o[2] = s->mpgx;
  o[3] = s->QxT;
  o[4] = s->Qx;
  o[5] = s->dtv;
  o[6] = s->dtc;
  o[7] = s->iter;
}

void lho_set_scalars(lho_state *s, double t, double mpgx, double QxT, int iter) {
  s->t = t;
  s->mpgx = mpgx;
  s->QxT = QxT;
  s->iter = iter;
}

static cplx *field_ptr(const lho_state *s, int fid, int *rows) {
  switch (fid) {
    case LHO_U1: *rows = s->nyu; return s->u1;
    case LHO_U2: *rows = s->nyv; return s->u2;
    case LHO_U3: *rows = s->nyu; return s->u3;
    case LHO_P: *rows = s->nyu; return s->p;
    case LHO_NU1: *rows = s->nyu; return s->Nu1;
    case LHO_NU2: *rows = s->nyv; return s->Nu2;
    case LHO_NU3: *rows = s->nyu; return s->Nu3;
    case LHO_DU1: *rows = s->nyu; return s->du1;
    case LHO_DU2: *rows = s->nyv; return s->du2;
    case LHO_DU3: *rows = s->nyu; return s->du3;
    case LHO_PSI: *rows = s->nyu; return s->psi;
    case LHO_DIV: *rows = s->nyu; return s->div;
  }
  *rows = 0;
  return NULL;
}

int lho_field_rows(const lho_state *s, int fid, int *jlo, int *jhi) {
  switch (fid) {
    case LHO_U1: case LHO_U3: case LHO_DU1: case LHO_DU3:
      *jlo = 0; *jhi = s->nyu - 1; break;
    case LHO_U2: case LHO_DU2:
      *jlo = 0; *jhi = s->nyv - 1; break;
    case LHO_P: case LHO_PSI: case LHO_DIV: case LHO_NU1: case LHO_NU3:
      *jlo = 1; *jhi = s->nyu - 2; break;
    case LHO_NU2:
      *jlo = 1; *jhi = s->nyv - 2; break;
    default: return -1;
  }
  return *jhi - *jlo + 1;
}

void lho_set_field(lho_state *s, int fid, const double *reim) {
  int rows, jlo, jhi;
  cplx *f = field_ptr(s, fid, &rows);
  int n = lho_field_rows(s, fid, &jlo, &jhi);
  for (int c = 0; c < s->ncol; c++)
    for (int j = jlo; j <= jhi; j++) {
      size_t q = (size_t)c * n + (j - jlo);
      f[(size_t)c * rows + j] = reim[2 * q] + I * reim[2 * q + 1];
    }
}

void lho_get_field(const lho_state *s, int fid, double *reim) {
  int rows, jlo, jhi;
  cplx *f = field_ptr(s, fid, &rows);
  int n = lho_field_rows(s, fid, &jlo, &jhi);
  for (int c = 0; c < s->ncol; c++)
    for (int j = jlo; j <= jhi; j++) {
      size_t q = (size_t)c * n + (j - jlo);
      reim[2 * q] = creal(f[(size_t)c * rows + j]);
      reim[2 * q + 1] = cimag(f[(size_t)c * rows + j]);
    }
}

/* flowrateIm / flowrateRe, littleharsh.f90:435-474 */
static double flowrate_c(const lho_state *s, const cplx *u) {
  double Q = 0.0;
  for (int j = 1; j <= s->nyu - 2; j++) Q = Q + (s->yv[j] - s->yv[j - 1]) * creal(u[j]);
  return Q;
}
static double flowrate_r(const lho_state *s, const double *u) {
  double Q = 0.0;
  for (int j = 1; j <= s->nyu - 2; j++) Q = Q + (s->yv[j] - s->yv[j - 1]) * u[j];
  return Q;
}

void lho_init_flowrate(lho_state *s) {
  s->Qx = flowrate_c(s, s->u1);
  if (s->c.flag_ctpress != 1) s->QxT = s->Qx;
}

/* ------------------------------------------------------------------------ */
/* RHS0_u1/u2/u3 + laplacian_U/V, littleharsh.f90:274-340,635-788          */
void lho_rhs0(lho_state *s, int kRK) {
  int q = kRK - 1;
  double C1 = s->aRK[q] / s->c.Re, C2 = -s->cRK[q];
  int nyu = s->nyu, nyv = s->nyv;
#pragma omp parallel for schedule(static)
  for (int c = 0; c < s->ncol; c++) {
    int i = c % s->nxh, k = c / s->nxh;
    cplx C3x = -s->dRK[q] * (I * s->kx[i]);
    cplx C3z = -s->dRK[q] * (I * s->kz[k]);
    const cplx *u1 = s->u1 + (size_t)c * nyu, *u3 = s->u3 + (size_t)c * nyu;
    const cplx *u2 = s->u2 + (size_t)c * nyv, *p = s->p + (size_t)c * nyu;
    cplx *d1 = s->du1 + (size_t)c * nyu, *d3 = s->du3 + (size_t)c * nyu;
    cplx *d2 = s->du2 + (size_t)c * nyv;
    const cplx *N1 = s->Nu1 + (size_t)c * nyu, *N3 = s->Nu3 + (size_t)c * nyu;
    const cplx *N2 = s->Nu2 + (size_t)c * nyv;
    double k2x = s->k2x[i], k2z = s->k2z[k];
    for (int j = 1; j <= nyu - 2; j++) {
      const double *a = s->dyu2i + 3 * j;
      cplx L1 = a[0] * u1[j - 1] + (k2x + k2z + a[1]) * u1[j] + a[2] * u1[j + 1];
      cplx L3 = a[0] * u3[j - 1] + (k2x + k2z + a[1]) * u3[j] + a[2] * u3[j + 1];
      d1[j] = C1 * L1 + C2 * N1[j] + C3x * p[j];
      d3[j] = C1 * L3 + C2 * N3[j] + C3z * p[j];
    }
    for (int j = 1; j <= nyv - 2; j++) {
      const double *a = s->dyv2i + 3 * j;
      cplx L2 = a[0] * u2[j - 1] + (k2x + k2z + a[1]) * u2[j] + a[2] * u2[j + 1];
      double C3 = -s->dRK[q] * s->ddthetavi * s->dthdyv[j];
      d2[j] = C1 * L2 + C2 * N2[j] + C3 * (p[j + 1] - p[j]);
    }
  }
  /* mean pressure gradient on mode (0,0): littleharsh.f90:670-676 */
  double Cm = -s->dRK[q] * s->mpgx;
  for (int j = 1; j <= nyu - 2; j++) s->du1[j] = s->du1[j] + Cm;
}

/* ------------------------------------------------------------------------ */
/* plane helpers (FOU3D.f90)                                               */

/* modes_to_planes_UVP (FOU3D.f90:1000-1098) for one plane: column (i,k) ->
 * P(2i+1, krow), P(2i+2, krow), krow = k (k<=N2/2) else k+Ngal2-N2 (1-based) */
static void modes_to_plane(const lho_state *s, const cplx *f, int rows, int j, double *P) {
  memset(P, 0, sizeof(double) * (size_t)s->igal * s->kgal);
  int N2 = s->c.N2, dk = N2 - s->c.Ngal2;
  for (int c = 0; c < s->ncol; c++) {
    int i = c % s->nxh, k = c / s->nxh + 1;
    int krow = k <= N2 / 2 ? k : k - dk;
    cplx v = f[(size_t)c * rows + j];
    P[(size_t)(krow - 1) * s->igal + 2 * i] = creal(v);
    P[(size_t)(krow - 1) * s->igal + 2 * i + 1] = cimag(v);
  }
}
static void plane_to_modes(const lho_state *s, const double *P, cplx *f, int rows, int j) {
  int N2 = s->c.N2, dk = N2 - s->c.Ngal2;
  for (int c = 0; c < s->ncol; c++) {
    int i = c % s->nxh, k = c / s->nxh + 1;
    int krow = k <= N2 / 2 ? k : k - dk;
    f[(size_t)c * rows + j] = P[(size_t)(krow - 1) * s->igal + 2 * i] +
                              I * P[(size_t)(krow - 1) * s->igal + 2 * i + 1];
  }
}

/* four_to_phys_du, FOU3D.f90:766-784 */
void lho_four_to_phys_plane(const lho_state *s, double *P) {
  int igal = s->igal, kgal = s->kgal;
  for (int ii = 0; ii < igal; ii++) P[(size_t)(kgal / 2) * igal + ii] = 0.0;
  lho_cft(P, s->c.Ngal2, igal, 2, (s->c.N1 + 2) / 2, +1);
  lho_rft(P, s->c.Ngal1, igal, kgal, +1);
}
/* phys_to_four_du, FOU3D.f90:2253-2272 */
void lho_phys_to_four_plane(const lho_state *s, double *P) {
  int igal = s->igal, kgal = s->kgal;
  lho_rft(P, s->c.Ngal1, igal, kgal, -1);
  lho_cft(P, s->c.Ngal2, igal, 2, (s->c.N1 + 2) / 2, -1);
  for (int ii = 0; ii < igal; ii++) P[(size_t)(kgal / 2) * igal + ii] = 0.0;
}

/* der_x + der_z and their sum (FOU3D.f90:476-527,552-593, 2160-2165):
 * out = ikx*A + ikz*B on the retained modes, zero elsewhere */
static void derxz_sum(const lho_state *s, const double *A, const double *B, double *out) {
  int igal = s->igal, kgal = s->kgal, N1 = s->c.N1, N2 = s->c.N2;
  int Ngal1 = s->c.Ngal1;
  int dk2 = N2 - kgal;
  for (int k = 1; k <= kgal; k++) {
    int live = (k <= N2 / 2) || (k >= kgal - N2 / 2 + 2);
    int kzi = k <= N2 / 2 ? k : k + dk2; /* 1-based index into k1F_z */
    for (int i = 0; i <= Ngal1 / 2; i++) {
      size_t o = (size_t)(k - 1) * igal + 2 * i;
      if (live && i <= N1 / 2) {
        cplx a = A[o] + I * A[o + 1], b = B[o] + I * B[o + 1];
        cplx dx = (I * s->kx[i]) * a;
        cplx dz = (I * s->kz[kzi - 1]) * b;
        cplx r = dx + dz;
        out[o] = creal(r);
        out[o + 1] = cimag(r);
      } else {
        out[o] = 0.0;
        out[o + 1] = 0.0;
      }
    }
  }
}

/* interp_u / interp_v, FOU3D.f90:427-474 */
static void interp_all(lho_state *s) {
  int nyu = s->nyu, nyv = s->nyv;
  const double *yu = s->yu, *yv = s->yv;
#pragma omp parallel for schedule(static)
  for (int c = 0; c < s->ncol; c++) {
    const cplx *u1 = s->u1 + (size_t)c * nyu, *u3 = s->u3 + (size_t)c * nyu;
    const cplx *u2 = s->u2 + (size_t)c * nyv;
    cplx *o1 = s->u1i + (size_t)c * nyv, *o3 = s->u3i + (size_t)c * nyv;
    cplx *o2 = s->u2i + (size_t)c * nyu;
    for (int j = 0; j <= nyv - 1; j++) {
      o1[j] = ((yv[j] - yu[j]) * u1[j + 1] + (yu[j + 1] - yv[j]) * u1[j]) / (yu[j + 1] - yu[j]);
      o3[j] = ((yv[j] - yu[j]) * u3[j + 1] + (yu[j + 1] - yv[j]) * u3[j]) / (yu[j + 1] - yu[j]);
    }
    for (int j = 1; j <= nyu - 2; j++)
      o2[j] = ((yu[j] - yv[j - 1]) * u2[j] + (yv[j] - yu[j]) * u2[j - 1]) / (yv[j] - yv[j - 1]);
    o2[0] = s->gwi1 * (u2[0] - o2[1]) + o2[1];
    o2[nyu - 1] = s->gwi2 * (u2[nyv - 1] - o2[nyu - 2]) + o2[nyu - 2];
  }
}

/* ------------------------------------------------------------------------ */
/* Restricted nonlinear path: ops_in_planes (FOU3D.f90:1724-1987) with
 * nonlinInter (FOU3D.f90:1989-2030) and the index maps iLkup/kLkup/iNeg
 * (start.f90:1139-1163), on (Ngal1+2) x Ngal2 spectral planes. */
int lho_set_triads(lho_state *s, int nfile, const int *lens, const signed char *rows, const int *map_u,
                   const int *map_v) {
  if (s->tri_rows) {
    for (int q = 0; q < 9 * s->ntri; q++) free(s->tri_rows[q]);
  }
  free(s->tri_rows);
  free(s->tri_len);
  free(s->tri_u);
  free(s->tri_v);
  s->tri_rows = NULL;
  s->tri_len = s->tri_u = s->tri_v = NULL;
  s->ntri = 0;
  if (nfile <= 0) return 0;
  int nx = s->c.N1 / 2 - 1, nz = s->c.N2 / 2 - 1;
  s->ntri = nfile;
  s->tri_len = malloc(sizeof(int) * 9 * nfile);
  s->tri_rows = calloc((size_t)9 * nfile, sizeof(signed char *));
  size_t off = 0;
  for (int q = 0; q < 9 * nfile; q++) {
    int n = lens[q];
    s->tri_len[q] = n;
    s->tri_rows[q] = malloc((size_t)4 * (n > 0 ? n : 1));
    memcpy(s->tri_rows[q], rows + off, (size_t)4 * n);
    off += (size_t)4 * n;
    const signed char *d = s->tri_rows[q];
    for (int l = 0; l < n; l++) {
      int la = d[l], ka = d[n + l], lp = d[2 * n + l], kp = d[3 * n + l];
      if (abs(la) > nx || abs(lp) > nx || abs(ka) > nz || abs(kp) > nz || la + lp < 0 || la + lp > nx ||
          abs(ka + kp) > nz) {
        lho_set_triads(s, 0, NULL, NULL, NULL, NULL);
        return -1;
      }
    }
  }
  s->tri_u = malloc(sizeof(int) * (s->nyu - 2));
  s->tri_v = malloc(sizeof(int) * (s->nyv - 2));
  memcpy(s->tri_u, map_u, sizeof(int) * (s->nyu - 2));
  memcpy(s->tri_v, map_v, sizeof(int) * (s->nyv - 2));
  return 0;
}

static int ilk(int l) { return 2 * abs(l); } /* iLkup - 1 */
static int ineg(int l) { return l < 0 ? -1 : 1; }
/* kLkup - 1: k in 0..N2/2-1 -> k, -N2/2..-1 -> k + Ngal2, N2/2 -> Ngal2 - N2/2 */
static int klk(const lho_state *s, int k) {
  int h = s->c.N2 / 2;
  if (k >= 0 && k < h) return k;
  if (k == h) return s->c.Ngal2 - h;
  return k + s->c.Ngal2;
}

/* nonlinInter(jlist, x_cPL, uaPL, upPL), FOU3D.f90:1989-2030 */
static void nonlin_inter(const lho_state *s, int q, int type, double *X, const double *A, const double *P) {
  if (q < 0) return;
  int n = s->tri_len[q * 9 + type - 1], ig = s->igal;
  const signed char *d = s->tri_rows[q * 9 + type - 1];
  for (int l = 0; l < n; l++) {
    int la = d[l], ka0 = d[n + l], lp = d[2 * n + l], kp0 = d[3 * n + l];
    int ia = ilk(la), ip = ilk(lp), it = ilk(la + lp);
    for (int p = 0; p <= 1; p++) {
      if (ka0 + kp0 == 0 && p == 1) continue;
      int pn = p * 2 - 1;
      int ka = klk(s, pn * ineg(la) * ka0), kp = klk(s, pn * ineg(lp) * kp0), kt = klk(s, pn * (ka0 + kp0));
      const double *a = A + (size_t)ka * ig, *b = P + (size_t)kp * ig;
      double *x = X + (size_t)kt * ig;
      x[it] = x[it] + a[ia] * b[ip] - ineg(la) * ineg(lp) * a[ia + 1] * b[ip + 1];
      x[it + 1] = x[it + 1] + ineg(lp) * a[ia] * b[ip + 1] + ineg(la) * a[ia + 1] * b[ip];
    }
  }
}

/* the "0th mode" sums of ops_in_planes (FOU3D.f90:1763-1778, 1892-1908):
 * x(0,0) += sum over i, k of a(i,k) b(-i,-k), k from -kneg to N2/2-1 */
static void zeroth_mode(const lho_state *s, double *X, const double *A, const double *B, int kneg) {
  int h1 = s->c.N1 / 2, h2 = s->c.N2 / 2, ig = s->igal;
  for (int i = -h1; i <= h1; i++)
    for (int k = -kneg; k <= h2 - 1; k++) {
      int la = i, lp = -i;
      int ia = ilk(la), ip = ilk(lp);
      int ka = klk(s, ineg(la) * k), kp = klk(s, ineg(lp) * (-k));
      const double *a = A + (size_t)ka * ig, *b = B + (size_t)kp * ig;
      X[0] = X[0] + a[ia] * b[ip] - ineg(i) * ineg(-i) * a[ia + 1] * b[ip + 1];
      X[1] = X[1] + a[ia] * b[ip + 1] * ineg(-i) + ineg(i) * a[ia + 1] * b[ip];
    }
}

/* u plane j of ops_in_planes (FOU3D.f90:1751-1881): Nu1, Nu3 rows and vv */
static void ops_list_u(lho_state *s, int j, double *P1, double *P2, double *P3, double *W[6]) {
  size_t psz = (size_t)s->igal * s->kgal;
  int nyu = s->nyu;
  double *uu = W[0], *uw = W[1], *vv = W[2], *wu = W[3], *ww = W[4], *R = W[5];
  memset(uu, 0, psz * 8); memset(uw, 0, psz * 8); memset(vv, 0, psz * 8);
  memset(wu, 0, psz * 8); memset(ww, 0, psz * 8);
  modes_to_plane(s, s->u1, nyu, j, P1);
  modes_to_plane(s, s->u2i, nyu, j, P2);
  modes_to_plane(s, s->u3, nyu, j, P3);
  zeroth_mode(s, vv, P2, P2, s->c.N2 / 2);
  /* linear advection by the mean mode (FOU3D.f90:1784-1798) */
  double a1 = P1[0], a2 = P2[0], a3 = P3[0];
  for (size_t o = 0; o < psz; o++) {
    uu[o] = uu[o] + 2 * (a1 * P1[o]);
    ww[o] = ww[o] + 2 * (a3 * P3[o]);
    double b = o < 2 ? 0.0 : 2 * (a2 * P2[o]);
    vv[o] = vv[o] + b;
    double c = a1 * P3[o] + a3 * P1[o];
    wu[o] = wu[o] + c;
    uw[o] = uw[o] + c;
  }
  int q = s->tri_u[j - 1];
  nonlin_inter(s, q, 1, uu, P1, P1);
  nonlin_inter(s, q, 3, uw, P3, P1);
  nonlin_inter(s, q, 5, vv, P2, P2);
  nonlin_inter(s, q, 7, wu, P1, P3);
  nonlin_inter(s, q, 9, ww, P3, P3);
  derxz_sum(s, uu, uw, R);
  plane_to_modes(s, R, s->Nu1, nyu, j);
  derxz_sum(s, wu, ww, R);
  plane_to_modes(s, R, s->Nu3, nyu, j);
  plane_to_modes(s, vv, s->vvc, nyu, j);
}

/* v plane j of ops_in_planes (FOU3D.f90:1889-1984): Nu2 rows, uv and wv */
static void ops_list_v(lho_state *s, int j, double *P1, double *P2, double *P3, double *W[6]) {
  size_t psz = (size_t)s->igal * s->kgal;
  int nyv = s->nyv;
  double *uv = W[0], *vu = W[1], *vw = W[2], *wv = W[3], *R = W[5];
  memset(uv, 0, psz * 8); memset(vu, 0, psz * 8); memset(vw, 0, psz * 8); memset(wv, 0, psz * 8);
  modes_to_plane(s, s->u1i, nyv, j, P1);
  modes_to_plane(s, s->u2, nyv, j, P2);
  modes_to_plane(s, s->u3i, nyv, j, P3);
  zeroth_mode(s, uv, P1, P2, s->c.N2 / 2 - 1);
  zeroth_mode(s, wv, P3, P2, s->c.N2 / 2 - 1);
  double a1 = P1[0], a2 = P2[0], a3 = P3[0];
  for (size_t o = 0; o < psz; o++) {
    double b = a1 * P2[o] + a2 * P1[o];
    vu[o] = vu[o] + b;
    uv[o] = uv[o] + (o < 2 ? 0.0 : b);
    double c = a2 * P3[o] + a3 * P2[o];
    vw[o] = vw[o] + c;
    wv[o] = wv[o] + (o < 2 ? 0.0 : c);
  }
  int q = s->tri_v[j - 1];
  nonlin_inter(s, q, 2, uv, P2, P1);
  nonlin_inter(s, q, 4, vu, P1, P2);
  nonlin_inter(s, q, 6, vw, P3, P2);
  nonlin_inter(s, q, 8, wv, P2, P3);
  derxz_sum(s, vu, vw, R);
  plane_to_modes(s, R, s->Nu2, nyv, j);
  plane_to_modes(s, uv, s->uvf, nyv, j);
  plane_to_modes(s, wv, s->wvf, nyv, j);
}

/* nonlinear (FOU3D.f90:41-410) with ops_in_planes2 (FOU3D.f90:2090-2251) */
void lho_nonlinear(lho_state *s, int kRK) {
  int nyu = s->nyu, nyv = s->nyv, igal = s->igal, kgal = s->kgal;
  int Ngal1 = s->c.Ngal1, Ngal2 = s->c.Ngal2;
  size_t psz = (size_t)igal * kgal;
  interp_all(s);
  double m1 = 0.0, m2 = 1e-6, m3 = 0.0; /* dtc_calc maxima */
  /* u-planes: uu, uw, vv, ww -> Nu1, Nu3 (PL part) and vv */
#pragma omp parallel reduction(max : m1, m3)
  {
    double *P1 = malloc(sizeof(double) * psz), *P2 = malloc(sizeof(double) * psz);
    double *P3 = malloc(sizeof(double) * psz), *Q = malloc(sizeof(double) * psz);
    double *R = malloc(sizeof(double) * psz);
    double *W[6] = {NULL};
    if (s->ntri)
      for (int q = 0; q < 6; q++) W[q] = malloc(sizeof(double) * psz);
#pragma omp for schedule(dynamic)
    for (int j = 1; j <= nyu - 2; j++) {
      if (s->ntri) { /* restricted path; u1PL, u3PL go to physical space after */
        ops_list_u(s, j, P1, P2, P3, W);
        lho_four_to_phys_plane(s, P1);
        lho_four_to_phys_plane(s, P3);
        for (int k = 0; k < Ngal2; k++)
          for (int i = 0; i < Ngal1; i++) {
            size_t o = (size_t)k * igal + i;
            if (fabs(P1[o]) > m1) m1 = fabs(P1[o]);
            if (fabs(P3[o]) > m3) m3 = fabs(P3[o]);
          }
        continue;
      }
      modes_to_plane(s, s->u1, nyu, j, P1);
      modes_to_plane(s, s->u2i, nyu, j, P2);
      modes_to_plane(s, s->u3, nyu, j, P3);
      lho_four_to_phys_plane(s, P1);
      lho_four_to_phys_plane(s, P2);
      lho_four_to_phys_plane(s, P3);
      for (int k = 0; k < Ngal2; k++)
        for (int i = 0; i < Ngal1; i++) {
          size_t o = (size_t)k * igal + i;
          double a = P1[o], b = P2[o], c = P3[o];
          if (fabs(a) > m1) m1 = fabs(a);
          if (fabs(c) > m3) m3 = fabs(c);
          Q[o] = a * c;  /* uw */
          P1[o] = a * a; /* uu */
          P2[o] = b * b; /* vv */
          P3[o] = c * c; /* ww */
        }
      lho_phys_to_four_plane(s, P1);
      lho_phys_to_four_plane(s, Q);
      lho_phys_to_four_plane(s, P2);
      lho_phys_to_four_plane(s, P3);
      derxz_sum(s, P1, Q, R);
      plane_to_modes(s, R, s->Nu1, nyu, j);
      derxz_sum(s, Q, P3, R);
      plane_to_modes(s, R, s->Nu3, nyu, j);
      plane_to_modes(s, P2, s->vvc, nyu, j);
    }
    free(P1); free(P2); free(P3); free(Q); free(R);
    for (int q = 0; q < 6; q++) free(W[q]);
  }
  /* v-planes: uv, wv -> Nu2 (PL part), uv, wv */
#pragma omp parallel reduction(max : m2)
  {
    double *P1 = malloc(sizeof(double) * psz), *P2 = malloc(sizeof(double) * psz);
    double *P3 = malloc(sizeof(double) * psz), *R = malloc(sizeof(double) * psz);
    double *W[6] = {NULL};
    if (s->ntri)
      for (int q = 0; q < 6; q++) W[q] = malloc(sizeof(double) * psz);
#pragma omp for schedule(dynamic)
    for (int j = 1; j <= nyv - 2; j++) {
      if (s->ntri) {
        ops_list_v(s, j, P1, P2, P3, W);
        lho_four_to_phys_plane(s, P2);
        for (int k = 0; k < Ngal2; k++)
          for (int i = 0; i < Ngal1; i++) {
            double av = fabs(P2[(size_t)k * igal + i]) * s->dthdyv[j];
            if (av > m2) m2 = av;
          }
        continue;
      }
      modes_to_plane(s, s->u1i, nyv, j, P1);
      modes_to_plane(s, s->u2, nyv, j, P2);
      modes_to_plane(s, s->u3i, nyv, j, P3);
      lho_four_to_phys_plane(s, P1);
      lho_four_to_phys_plane(s, P2);
      lho_four_to_phys_plane(s, P3);
      for (int k = 0; k < Ngal2; k++)
        for (int i = 0; i < Ngal1; i++) {
          size_t o = (size_t)k * igal + i;
          double v = P2[o];
          double av = fabs(v) * s->dthdyv[j];
          if (av > m2) m2 = av;
          P1[o] = P1[o] * v; /* uv */
          P3[o] = P3[o] * v; /* wv */
        }
      lho_phys_to_four_plane(s, P1);
      lho_phys_to_four_plane(s, P3);
      derxz_sum(s, P1, P3, R);
      plane_to_modes(s, R, s->Nu2, nyv, j);
      plane_to_modes(s, P1, s->uvf, nyv, j);
      plane_to_modes(s, P3, s->wvf, nyv, j);
    }
    free(P1); free(P2); free(P3); free(R);
    for (int q = 0; q < 6; q++) free(W[q]);
  }
  /* wall faces / ghost centres of the flux planes are zero */
  for (int c = 0; c < s->ncol; c++) {
    s->uvf[(size_t)c * nyv] = 0;
    s->uvf[(size_t)c * nyv + nyv - 1] = 0;
    s->wvf[(size_t)c * nyv] = 0;
    s->wvf[(size_t)c * nyv + nyv - 1] = 0;
    s->vvc[(size_t)c * nyu] = 0;
    s->vvc[(size_t)c * nyu + nyu - 1] = 0;
  }
  /* y-derivatives of the fluxes, FOU3D.f90:266-283 (der_yu_h, der_yv_h) */
  double ddv = s->ddthetavi;
#pragma omp parallel for schedule(static)
  for (int c = 0; c < s->ncol; c++) {
    cplx *N1 = s->Nu1 + (size_t)c * nyu, *N3 = s->Nu3 + (size_t)c * nyu;
    cplx *N2 = s->Nu2 + (size_t)c * nyv;
    const cplx *uv = s->uvf + (size_t)c * nyv, *wv = s->wvf + (size_t)c * nyv;
    const cplx *vv = s->vvc + (size_t)c * nyu;
    for (int j = 1; j <= nyu - 2; j++) {
      N1[j] = N1[j] + (uv[j] - uv[j - 1]) * s->dthdyu[j] * ddv;
      N3[j] = N3[j] + (wv[j] - wv[j - 1]) * s->dthdyu[j] * ddv;
    }
    for (int j = 1; j <= nyv - 2; j++) N2[j] = N2[j] + (vv[j + 1] - vv[j]) * s->dthdyv[j] * ddv;
  }
  /* time step, FOU3D.f90:288-297 and dtc_calc FOU3D.f90:699-736 */
  if (kRK == 1) {
    int N1m = s->c.N1, N2m = s->c.N2;
    /* The third candidate of FOU3D.f90:726, 1/(u2mloc*dthetavi), divides by
     * dthetavi, declared (declaration.f90:76) but never assigned (start.f90
     * sets only ddthetavi, :324): it is +Inf and never binds.  The
     * reference's history file shows dtc = the u3 term (test_oracle.py pins
     * it).  m2 is kept for the record only. */
    (void)m2;
    s->dtc = fmin(1.0 / (s->c.alp * (N1m / 2) * m1), 1.0 / (s->c.bet * (N2m / 2) * m3));
    if (!s->c.fixed_dt) s->dt = fmin(2.5 * s->dtv, s->c.CFL * s->dtc);
    s->dti = 1.0 / s->dt;
    s->t = s->t + s->dt;
  }
  /* du += -gRK*N, FOU3D.f90:398-410 */
  double C1 = -s->gRK[kRK - 1];
#pragma omp parallel for schedule(static)
  for (int c = 0; c < s->ncol; c++) {
    for (int j = 1; j <= nyu - 2; j++) {
      s->du1[(size_t)c * nyu + j] += C1 * s->Nu1[(size_t)c * nyu + j];
      s->du3[(size_t)c * nyu + j] += C1 * s->Nu3[(size_t)c * nyu + j];
    }
    for (int j = 1; j <= nyv - 2; j++) s->du2[(size_t)c * nyv + j] += C1 * s->Nu2[(size_t)c * nyv + j];
  }
}

/* ------------------------------------------------------------------------ */
/* LU_dec (tridLU_3D.f90:437-474): a[3*j+0] lower, [1] diag, [2] upper      */
static void lu_dec(double *a, int n) {
  a[1] = 1.0 / a[1];
  for (int j = 1; j < n; j++) {
    a[3 * (j - 1) + 2] = a[3 * (j - 1) + 2] * a[3 * (j - 1) + 1];
    a[3 * j + 1] = 1.0 / (a[3 * j + 1] - a[3 * j + 0] * a[3 * (j - 1) + 2]);
  }
}
static void lu_sol_c(const double *a, const cplx *rhs, cplx *x, int n) {
  x[0] = rhs[0] * a[1];
  for (int j = 1; j < n; j++) x[j] = (rhs[j] - a[3 * j] * x[j - 1]) * a[3 * j + 1];
  for (int j = n - 2; j >= 0; j--) x[j] = x[j] - a[3 * j + 2] * x[j + 1];
}

/* matrices of LU_buildU / LU_buildV (tridLU_3D.f90:257-329) */
static void build_helm(const lho_state *s, int kRK, int c, int vgrid, double *a) {
  int n = vgrid ? s->nyv : s->nyu;
  const double *dy = vgrid ? s->dyv2i : s->dyu2i;
  double beta = s->bRK[kRK - 1] * s->dt / s->c.Re;
  int i = c % s->nxh, k = c / s->nxh;
  double k2x = s->k2x[i], k2z = s->k2z[k];
  a[0] = 0.0;
  a[1] = 1.0;
  a[2] = vgrid ? 0.0 : s->gw1;
  for (int j = 1; j <= n - 2; j++) {
    const double *d = dy + 3 * j;
    a[3 * j] = -beta * d[0];
    a[3 * j + 1] = 1.0 - beta * (d[1] + k2x + k2z);
    a[3 * j + 2] = -beta * d[2];
  }
  a[3 * (n - 1)] = vgrid ? 0.0 : s->gw2;
  a[3 * (n - 1) + 1] = 1.0;
  a[3 * (n - 1) + 2] = 0.0;
}

/* LUsolU_W / LUsolV on arbitrary right-hand sides in the reference layout
 * (ncol x rows, re/im interleaved); vgrid selects LUsolV */
void lho_lusol_helm(const lho_state *s, int kRK, int vgrid, const double *rhs, double *x) {
  int n = vgrid ? s->nyv : s->nyu;
  double *a = malloc(sizeof(double) * 3 * n);
  for (int c = 0; c < s->ncol; c++) {
    build_helm(s, kRK, c, vgrid, a);
    lu_dec(a, n);
    lu_sol_c(a, (const cplx *)rhs + (size_t)c * n, (cplx *)x + (size_t)c * n, n);
  }
  free(a);
}

/* LUsolP with the prefactored LU_buildP operator on p rows 1..ny-1 */
void lho_lusolp(const lho_state *s, double *reim) {
  int nyu = s->nyu, ns = 1, ne = nyu - 2, nr = ne - ns + 1;
  for (int c = 0; c < s->ncol; c++) {
    cplx *ps = (cplx *)reim + (size_t)c * nr - ns;
    const double *a = s->dg + (size_t)c * nyu * 3;
    ps[ns] = ps[ns] * a[3 * ns + 1];
    for (int j = ns + 1; j <= ne; j++) ps[j] = (ps[j] - a[3 * j] * ps[j - 1]) * a[3 * j + 1];
    for (int j = ne - 1; j >= ns; j--) ps[j] = ps[j] - a[3 * j + 2] * ps[j + 1];
  }
}

/* solveU_W + LUsolU_W + LU_buildU, littleharsh.f90:791-828,
 * tridLU_3D.f90:55-132,257-291 */
static void build_u_matrix(const lho_state *s, double beta, int c, double *a) {
  int nyu = s->nyu, i = c % s->nxh, k = c / s->nxh;
  double k2x = s->k2x[i], k2z = s->k2z[k];
  a[0] = 0.0;
  a[1] = 1.0;
  a[2] = s->gw1;
  for (int j = 1; j <= nyu - 2; j++) {
    const double *d = s->dyu2i + 3 * j;
    a[3 * j] = -beta * d[0];
    a[3 * j + 1] = 1.0 - beta * (d[1] + k2x + k2z);
    a[3 * j + 2] = -beta * d[2];
  }
  a[3 * (nyu - 1)] = s->gw2;
  a[3 * (nyu - 1) + 1] = 1.0;
  a[3 * (nyu - 1) + 2] = 0.0;
}

static void tex_solve(lho_state *s, double beta, cplx *rhs[2], cplx *out[2]);

void lho_solve_uw(lho_state *s, int kRK) {
  int nyu = s->nyu;
  double beta = s->bRK[kRK - 1] * s->dt / s->c.Re;
#pragma omp parallel for schedule(static)
  for (int c = 0; c < s->ncol; c++) {
    cplx *u = s->u1 + (size_t)c * nyu, *w = s->u3 + (size_t)c * nyu;
    cplx *du = s->du1 + (size_t)c * nyu, *dw = s->du3 + (size_t)c * nyu;
    du[0] = 0;
    dw[0] = 0;
    for (int j = 1; j <= nyu - 2; j++) {
      du[j] = u[j] + s->dt * du[j];
      dw[j] = w[j] + s->dt * dw[j];
    }
    du[nyu - 1] = 0;
    dw[nyu - 1] = 0;
  }
  cplx *rhs[2] = {s->du1, s->du3}, *out[2] = {s->u1, s->u3};
  tex_solve(s, beta, rhs, out);
}

/* ------------------------------------------------------------------------ */
/* Textured walls.  The reference builds the free-shear / no-slip mask
 * planeBC (start.f90:1776-1829) and, in its commented-out LUsolU
 * (tridLU_3D.f90:753-1078) with LU_build (tridLU_3D.f90:1284-1358), replaces
 * the wall row of the U/W Helmholtz systems point by point in physical space
 * on the N1 x N2 grid (four_to_phys_N / phys_to_four_N): no-slip keeps
 * u0 + gw1 u1 = 0, free-shear has the zero-gradient row u0 - u1 = 0 (top
 * wall: u_ny - u_{ny-1} = 0 against u_ny + gw2 u_{ny-1} = 0).
 *
 * This restatement solves that mode-coupled system exactly, for the
 * unfactored operator: the wall rows differ from the smooth ones by a
 * boundary source, A(k) u = r + e_0 sb + e_ny st, where
 *   sb = T[Pb (T^-1 u_1)],   Pb = 1 + gw1 on free-shear points, 0 else,
 *   st = T[Pt (T^-1 u_{ny-1})], Pt = 1 + gw2 on free-shear points,
 * T the N-grid transform.  With u = A^-1 r + gb sb + .. the sources solve a
 * fixed point, iterated here to convergence; then u = A^-1 (r + e_0 sb +
 * e_ny st).  The ghost rows of v_corr follow the mask point by point:
 * u0 = gwbc u1 on no-slip points, u0 = u1 on free-shear points (tex_ghosts). */

/* four_to_phys_N / phys_to_four_N (FOU3D.f90:786-804, 2274-2293) on an
 * (N1+2) x N2 plane: no oddball row zeroing */
static void four_to_phys_N(const lho_state *s, double *P) {
  int ig = s->c.N1 + 2;
  lho_cft(P, s->c.N2, ig, 2, (s->c.N1 + 2) / 2, +1);
  lho_rft(P, s->c.N1, ig, s->c.N2, +1);
}
static void phys_to_four_N(const lho_state *s, double *P) {
  int ig = s->c.N1 + 2;
  lho_rft(P, s->c.N1, ig, s->c.N2, -1);
  lho_cft(P, s->c.N2, ig, 2, (s->c.N1 + 2) / 2, -1);
}
void lho_four_to_phys_N_plane(const lho_state *s, double *P) { four_to_phys_N(s, P); }
void lho_phys_to_four_N_plane(const lho_state *s, double *P) { phys_to_four_N(s, P); }
/* column (i,k) <-> plane (2i, 2i+1) of row k */
static void col_to_planeN(const lho_state *s, const cplx *v, double *P) {
  int ig = s->c.N1 + 2;
  for (int c = 0; c < s->ncol; c++) {
    int i = c % s->nxh, k = c / s->nxh;
    P[(size_t)k * ig + 2 * i] = creal(v[c]);
    P[(size_t)k * ig + 2 * i + 1] = cimag(v[c]);
  }
}
static void planeN_to_col(const lho_state *s, const double *P, cplx *v) {
  int ig = s->c.N1 + 2;
  for (int c = 0; c < s->ncol; c++) {
    int i = c % s->nxh, k = c / s->nxh;
    v[c] = P[(size_t)k * ig + 2 * i] + I * P[(size_t)k * ig + 2 * i + 1];
  }
}
/* v <- T[m(x) . T^-1 v] with m = fs on free-shear points and ns elsewhere */
static void mask_product(const lho_state *s, const int *mask, double fs, double ns, cplx *v, double *P) {
  int ig = s->c.N1 + 2, N1 = s->c.N1, N2 = s->c.N2;
  memset(P, 0, sizeof(double) * (size_t)ig * N2);
  col_to_planeN(s, v, P);
  four_to_phys_N(s, P);
  for (int z = 0; z < N2; z++)
    for (int x = 0; x < N1; x++) P[(size_t)z * ig + x] *= mask[z * N1 + x] == 2 ? fs : ns;
  phys_to_four_N(s, P);
  planeN_to_col(s, P, v);
}

static void tex_solve(lho_state *s, double beta, cplx *rhs[2], cplx *out[2]) {
  int nyu = s->nyu, ncol = s->ncol;
  int tex = s->tex[0] || s->tex[1];
  /* smooth solve of every right-hand side; for textured walls also the
   * responses g of rows 1 / ny-1 to a unit source in the wall rows */
  double *gbb = NULL, *gbt = NULL, *gtb = NULL, *gtt = NULL;
  if (tex) {
    gbb = calloc(ncol, sizeof(double));
    gbt = calloc(ncol, sizeof(double));
    gtb = calloc(ncol, sizeof(double));
    gtt = calloc(ncol, sizeof(double));
  }
#pragma omp parallel
  {
    double *a = malloc(sizeof(double) * 3 * nyu);
    cplx *e = malloc(sizeof(cplx) * nyu), *x = malloc(sizeof(cplx) * nyu);
#pragma omp for schedule(static)
    for (int c = 0; c < ncol; c++) {
      build_u_matrix(s, beta, c, a);
      lu_dec(a, nyu);
      for (int f = 0; f < 2; f++)
        lu_sol_c(a, rhs[f] + (size_t)c * nyu, out[f] + (size_t)c * nyu, nyu);
      if (tex) {
        memset(e, 0, sizeof(cplx) * nyu);
        e[0] = 1.0;
        lu_sol_c(a, e, x, nyu);
        gbb[c] = creal(x[1]);
        gtb[c] = creal(x[nyu - 2]);
        memset(e, 0, sizeof(cplx) * nyu);
        e[nyu - 1] = 1.0;
        lu_sol_c(a, e, x, nyu);
        gbt[c] = creal(x[1]);
        gtt[c] = creal(x[nyu - 2]);
      }
    }
    free(a);
    free(e);
    free(x);
  }
  if (!tex) return;
  int ig = s->c.N1 + 2;
  double *P = malloc(sizeof(double) * (size_t)ig * s->c.N2);
  cplx *sb = calloc(ncol, sizeof(cplx)), *st = calloc(ncol, sizeof(cplx));
  cplx *tb = malloc(sizeof(cplx) * ncol), *tt = malloc(sizeof(cplx) * ncol);
  const double Pb = 1.0 + s->gw1, Pt = 1.0 + s->gw2;
  for (int f = 0; f < 2; f++) {
    memset(sb, 0, sizeof(cplx) * ncol);
    memset(st, 0, sizeof(cplx) * ncol);
    for (int it = 0; it < 200; it++) {
      double change = 0.0, scale = 0.0;
      for (int c = 0; c < ncol; c++) {
        tb[c] = out[f][(size_t)c * nyu + 1] + gbb[c] * sb[c] + gbt[c] * st[c];
        tt[c] = out[f][(size_t)c * nyu + nyu - 2] + gtb[c] * sb[c] + gtt[c] * st[c];
      }
      if (s->tex[0]) mask_product(s, s->tex[0], Pb, 0.0, tb, P);
      else memset(tb, 0, sizeof(cplx) * ncol);
      if (s->tex[1]) mask_product(s, s->tex[1], Pt, 0.0, tt, P);
      else memset(tt, 0, sizeof(cplx) * ncol);
      for (int c = 0; c < ncol; c++) {
        change = fmax(change, cabs(tb[c] - sb[c]) + cabs(tt[c] - st[c]));
        scale = fmax(scale, cabs(tb[c]) + cabs(tt[c]));
        sb[c] = tb[c];
        st[c] = tt[c];
      }
      if (change <= 1e-17 * scale || scale == 0.0) break;
    }
    /* final solve with the boundary sources in the wall rows */
    for (int c = 0; c < ncol; c++) {
      rhs[f][(size_t)c * nyu] += sb[c];
      rhs[f][(size_t)c * nyu + nyu - 1] += st[c];
    }
#pragma omp parallel
    {
      double *a = malloc(sizeof(double) * 3 * nyu);
#pragma omp for schedule(static)
      for (int c = 0; c < ncol; c++) {
        build_u_matrix(s, beta, c, a);
        lu_dec(a, nyu);
        lu_sol_c(a, rhs[f] + (size_t)c * nyu, out[f] + (size_t)c * nyu, nyu);
      }
      free(a);
    }
  }
  free(P);
  free(sb);
  free(st);
  free(tb);
  free(tt);
  free(gbb);
  free(gbt);
  free(gtb);
  free(gtt);
}

/* ghost rows of v_corr with the wall masks: u0 = gwbc u1 on no-slip points,
 * u0 = u1 on free-shear points, applied as u0 = gwbc u1 + T[(1 - gwbc)
 * chi_fs . T^-1 u1] so that the uniform part stays in Fourier space (the
 * N-grid transform cannot hold the x-Nyquist column i = N1/2, which the
 * 3/2-padded step keeps) */
static void tex_ghosts(lho_state *s) {
  int nyu = s->nyu, ncol = s->ncol, ig = s->c.N1 + 2;
  double *P = malloc(sizeof(double) * (size_t)ig * s->c.N2);
  cplx *v = malloc(sizeof(cplx) * ncol);
  cplx *fld[2] = {s->u1, s->u3};
  double gwbc[2] = {s->gwbc1, s->gwbc3};
  for (int f = 0; f < 2; f++)
    for (int w = 0; w < 2; w++) {
      if (!s->tex[w]) continue;
      int src = w == 0 ? 1 : nyu - 2, dst = w == 0 ? 0 : nyu - 1;
      for (int c = 0; c < ncol; c++) v[c] = fld[f][(size_t)c * nyu + src];
      mask_product(s, s->tex[w], 1.0 - gwbc[f], 0.0, v, P);
      for (int c = 0; c < ncol; c++) fld[f][(size_t)c * nyu + dst] += v[c];
    }
  free(P);
  free(v);
}

int lho_set_texture(lho_state *s, const int *bottom, const int *top) {
  size_t n = (size_t)s->c.N1 * s->c.N2;
  const int *m[2] = {bottom, top};
  for (int w = 0; w < 2; w++) {
    free(s->tex[w]);
    s->tex[w] = NULL;
    if (!m[w]) continue;
    for (size_t q = 0; q < n; q++)
      if (m[w][q] != 1 && m[w][q] != 2) return -1;
    s->tex[w] = malloc(sizeof(int) * n);
    memcpy(s->tex[w], m[w], sizeof(int) * n);
  }
  return 0;
}

/* LUsolU_W on arbitrary right-hand sides with the channel's wall textures
 * (rhs rows 0 and ny are the wall-row right-hand sides) */
void lho_lusol_uw_tex(lho_state *s, int kRK, const double *rhs_u, const double *rhs_w, double *u, double *w) {
  size_t n = (size_t)s->ncol * s->nyu;
  cplx *r[2] = {malloc(sizeof(cplx) * n), malloc(sizeof(cplx) * n)};
  memcpy(r[0], rhs_u, sizeof(cplx) * n);
  memcpy(r[1], rhs_w, sizeof(cplx) * n);
  cplx *o[2] = {(cplx *)u, (cplx *)w};
  tex_solve(s, s->bRK[kRK - 1] * s->dt / s->c.Re, r, o);
  free(r[0]);
  free(r[1]);
}

/* solveV + LUsolV + LU_buildV, littleharsh.f90:877-907, tridLU_3D.f90:135-189,294-329 */
void lho_solve_v(lho_state *s, int kRK) {
  int nyv = s->nyv;
  double beta = s->bRK[kRK - 1] * s->dt / s->c.Re;
#pragma omp parallel
  {
    double *a = malloc(sizeof(double) * 3 * nyv);
#pragma omp for schedule(static)
    for (int c = 0; c < s->ncol; c++) {
      int i = c % s->nxh, k = c / s->nxh;
      cplx *v = s->u2 + (size_t)c * nyv, *dv = s->du2 + (size_t)c * nyv;
      dv[0] = 0;
      for (int j = 1; j <= nyv - 2; j++) dv[j] = v[j] + s->dt * dv[j];
      dv[nyv - 1] = 0;
      double k2x = s->k2x[i], k2z = s->k2z[k];
      a[0] = 0.0;
      a[1] = 1.0;
      a[2] = 0.0;
      for (int j = 1; j <= nyv - 2; j++) {
        const double *d = s->dyv2i + 3 * j;
        a[3 * j] = -beta * d[0];
        a[3 * j + 1] = 1.0 - beta * (d[1] + k2x + k2z);
        a[3 * j + 2] = -beta * d[2];
      }
      a[3 * (nyv - 1)] = 0.0;
      a[3 * (nyv - 1) + 1] = 1.0;
      a[3 * (nyv - 1) + 2] = 0.0;
      lu_dec(a, nyv);
      lu_sol_c(a, dv, v, nyv);
    }
    free(a);
  }
}

/* LUsol0 + LU_build0 + LU_dec0, tridLU_3D.f90:228-254,409-435,520-537 */
void lho_lusol0(const lho_state *s, int kRK, double *x) {
  int n = s->nyu;
  double beta = s->bRK[kRK - 1] * s->dt / s->c.Re;
  double *a = malloc(sizeof(double) * 3 * n);
  a[0] = 0.0;
  a[1] = 1.0;
  a[2] = s->gw1;
  for (int j = 1; j <= n - 2; j++) {
    const double *d = s->dyu2i + 3 * j;
    a[3 * j] = -beta * d[0];
    a[3 * j + 1] = 1.0 - beta * d[1];
    a[3 * j + 2] = -beta * d[2];
  }
  a[3 * (n - 1)] = s->gw2;
  a[3 * (n - 1) + 1] = 1.0;
  a[3 * (n - 1) + 2] = 0.0;
  lu_dec(a, n);
  x[0] = x[0] * a[1];
  for (int j = 1; j < n; j++) x[j] = (x[j] - a[3 * j] * x[j - 1]) * a[3 * j + 1];
  for (int j = n - 2; j >= 0; j--) x[j] = x[j] - a[3 * j + 2] * x[j + 1];
  free(a);
}

/* meanflow_ctU + flowrate_corr, littleharsh.f90:476-509,547-562 */
void lho_meanflow_ctU(lho_state *s, int kRK) {
  int n = s->nyu;
  double *u11 = malloc(sizeof(double) * n);
  double bdt = -s->dRK[kRK - 1] * s->dt;
  u11[0] = 0.0;
  for (int j = 1; j <= n - 2; j++) u11[j] = bdt;
  u11[n - 1] = 0.0;
  lho_lusol0(s, kRK, u11);
  double Qx1 = flowrate_r(s, u11);
  double Qx0 = flowrate_c(s, s->u1);
  double g = (s->QxT - Qx0) / Qx1;
  s->mpgx = s->mpgx + g;
  for (int j = 0; j <= n - 1; j++) s->u1[j] = s->u1[j] + g * u11[j];
  free(u11);
}

/* divergence, littleharsh.f90:237-272: p rows 1..ny-1 */
void lho_divergence(lho_state *s) {
  int nyu = s->nyu, nyv = s->nyv;
#pragma omp parallel for schedule(static)
  for (int c = 0; c < s->ncol; c++) {
    int i = c % s->nxh, k = c / s->nxh;
    cplx kx = I * s->kx[i], kz = I * s->kz[k];
    const cplx *u1 = s->u1 + (size_t)c * nyu, *u3 = s->u3 + (size_t)c * nyu;
    const cplx *u2 = s->u2 + (size_t)c * nyv;
    cplx *d = s->div + (size_t)c * nyu;
    for (int j = 1; j <= nyu - 2; j++)
      d[j] = kx * u1[j] + kz * u3[j] + s->ddthetavi * s->dthdyu[j] * (u2[j] - u2[j - 1]);
  }
}

/* LU_buildP + LU_decP, tridLU_3D.f90:333-407,477-518 (p rows 1..ny-1) */
static void build_dg(lho_state *s) {
  int ns = 1, ne = s->nyu - 2, nyu = s->nyu;
  double ddv = s->ddthetavi;
  double *Db = calloc(nyu + 1, sizeof(double)), *Gb = calloc(nyu + 1, sizeof(double));
  double *Dt = calloc(nyu + 1, sizeof(double)), *Gt = calloc(nyu + 1, sizeof(double));
  for (int j = ns; j <= ne; j++) {
    Db[j - 1] = -ddv * s->dthdyu[j];
    Gb[j] = -ddv * s->dthdyv[j];
    Dt[j] = ddv * s->dthdyu[j];
    Gt[j + 1] = ddv * s->dthdyv[j];
  }
  for (int c = 0; c < s->ncol; c++) {
    int i = c % s->nxh, k = c / s->nxh;
    double k2x = s->k2x[i], k2z = s->k2z[k];
    double *a = s->dg + (size_t)c * nyu * 3; /* rows indexed by j directly */
    memset(a, 0, sizeof(double) * 3 * nyu);
    a[3 * ns + 1] = (Dt[ns] * Gb[ns]) + k2x + k2z;
    a[3 * ns + 2] = (Dt[ns] * Gt[ns + 1]);
    a[3 * ne + 1] = (Db[ne - 1] * Gt[ne]) + k2x + k2z;
    a[3 * ne + 0] = (Db[ne - 1] * Gb[ne - 1]);
    for (int j = ns + 1; j <= ne - 1; j++) {
      a[3 * j + 1] = (Db[j - 1] * Gt[j]) + (Dt[j] * Gb[j]) + k2x + k2z;
      a[3 * j + 0] = (Db[j - 1] * Gb[j - 1]);
      a[3 * j + 2] = (Dt[j] * Gt[j + 1]);
    }
    if (c == 0) { /* pin of the mean-mode pressure, tridLU_3D.f90:384-390 */
      a[3 * ns + 0] = 0.0;
      a[3 * ns + 1] = 1.0 / ((s->yu[2] - s->yu[1]) * (s->yu[2] - s->yu[1]));
      a[3 * ns + 2] = 0.0;
    }
    lu_dec(a + 3 * ns, ne - ns + 1);
  }
  free(Db); free(Gb); free(Dt); free(Gt);
}

/* solveP + LUsolP, littleharsh.f90:581-633, tridLU_3D.f90:192-226 */
void lho_solve_p(lho_state *s, int kRK) {
  int nyu = s->nyu, ns = 1, ne = nyu - 2;
  double dtirk = s->dti / (s->aRK[kRK - 1] + s->bRK[kRK - 1]);
  lho_divergence(s);
#pragma omp parallel for schedule(static)
  for (int c = 0; c < s->ncol; c++) {
    cplx *ps = s->psi + (size_t)c * nyu;
    const cplx *d = s->div + (size_t)c * nyu;
    for (int j = ns; j <= ne; j++) ps[j] = dtirk * d[j];
    if (c == 0) ps[ns] = 0.0;
    const double *a = s->dg + (size_t)c * nyu * 3;
    ps[ns] = ps[ns] * a[3 * ns + 1];
    for (int j = ns + 1; j <= ne; j++) ps[j] = (ps[j] - a[3 * j] * ps[j - 1]) * a[3 * j + 1];
    for (int j = ne - 1; j >= ns; j--) ps[j] = ps[j] - a[3 * j + 2] * ps[j + 1];
    cplx *p = s->p + (size_t)c * nyu;
    for (int j = ns; j <= ne; j++) p[j] = p[j] + ps[j];
  }
}

/* v_corr, littleharsh.f90:956-1052 */
void lho_v_corr(lho_state *s, int kRK) {
  int nyu = s->nyu, nyv = s->nyv;
  double dtrk = s->dt * (s->aRK[kRK - 1] + s->bRK[kRK - 1]);
  double dtrk_u2 = dtrk * s->ddthetavi;
#pragma omp parallel for schedule(static)
  for (int c = 0; c < s->ncol; c++) {
    int i = c % s->nxh, k = c / s->nxh;
    cplx kx = dtrk * (I * s->kx[i]), kz = dtrk * (I * s->kz[k]);
    cplx *u1 = s->u1 + (size_t)c * nyu, *u3 = s->u3 + (size_t)c * nyu;
    cplx *u2 = s->u2 + (size_t)c * nyv;
    const cplx *ps = s->psi + (size_t)c * nyu;
    const int elo = s->c.band_edge & 1, ehi = (s->c.band_edge >> 1) & 1;
    for (int j = 1; j <= nyu - 2; j++) u1[j] = u1[j] - kx * ps[j];
    if (!elo) u1[0] = s->gwbc1 * u1[1];
    if (!ehi) u1[nyu - 1] = s->gwbc1 * u1[nyu - 2];
    for (int j = 1; j <= nyv - 2; j++)
      u2[j] = u2[j] - dtrk_u2 * s->dthdyv[j] * (ps[j + 1] - ps[j]);
    u2[0] = 0;
    u2[nyv - 1] = 0;
    for (int j = 1; j <= nyu - 2; j++) u3[j] = u3[j] - kz * ps[j];
    if (!elo) u3[0] = s->gwbc3 * u3[1];
    if (!ehi) u3[nyu - 1] = s->gwbc3 * u3[nyu - 2];
  }
  if (s->tex[0] || s->tex[1]) tex_ghosts(s);
  lho_divergence(s);
}

/* one RK3 substep, littleharsh.f90:176-215 */
void lho_substep(lho_state *s, int kRK) {
  lho_rhs0(s, kRK);
  lho_nonlinear(s, kRK);
  lho_solve_uw(s, kRK);
  lho_solve_v(s, kRK);
  if (s->c.flag_ctpress == 0) lho_meanflow_ctU(s, kRK);
  lho_solve_p(s, kRK);
  lho_v_corr(s, kRK);
}

/* main loop body, littleharsh.f90:172-228 */
void lho_step(lho_state *s, int nsteps) {
  for (int n = 0; n < nsteps; n++) {
    for (int kRK = 1; kRK <= 3; kRK++) lho_substep(s, kRK);
    if (s->c.flag_ctpress == 1) s->Qx = flowrate_c(s, s->u1);
    s->iter = s->iter + 1;
  }
}

/* ------------------------------------------------------------------------ */
/* stats (stats.f90:1-107) and spectra / buildsp / buildspUV
 * (spectra.f90:1-56): one sample of the current fields, summed in physical
 * space exactly as the reference forms its planes at a statistics step:
 * u1PL, u3PL, u2PL_itp on the u rows and u2PL, u3PL_itp on the v rows are
 * filled by modes_to_planes_UVP (all rows, FOU3D.f90:1000-1098) and only
 * the interior rows are transformed by ops_in_planes2 (FOU3D.f90:2116,
 * 2151); ppPL is the spectral plane of p (FOU3D.f90:156-157); wx is
 * -der_z_N(u2) + der_yv_h_wx(u3) on the N grid, transformed by
 * four_to_phys_N (FOU3D.f90:302-327).  Sums run over i = 1..igal-2,
 * k = 1..kgal (N1 x N2 for wx) as in stats.f90. */
void lho_stats(lho_state *s, double *prof, double *sp) {
  const int nyu = s->nyu, nyv = s->nyv, ny = s->c.ny, igal = s->igal, kgal = s->kgal;
  const int N1 = s->c.N1, N2 = s->c.N2, ig = N1 + 2;
  const size_t psz = (size_t)igal * kgal;
  interp_all(s);
  double *P1 = malloc(sizeof(double) * psz), *P2 = malloc(sizeof(double) * psz);
  double *P3 = malloc(sizeof(double) * psz), *W = malloc(sizeof(double) * (size_t)ig * N2);
  cplx *wc = malloc(sizeof(cplx) * (size_t)s->ncol);
#define PROF(q, j) prof[(size_t)(q) * nyu + (j)]
  for (int j = 0; j < nyu; j++) {
    /* u rows */
    modes_to_plane(s, s->u1, nyu, j, P1);
    modes_to_plane(s, s->u2i, nyu, j, P2);
    modes_to_plane(s, s->u3, nyu, j, P3);
    if (j >= 1 && j <= ny - 1) {
      lho_four_to_phys_plane(s, P1);
      lho_four_to_phys_plane(s, P2);
      lho_four_to_phys_plane(s, P3);
    }
    for (int k = 0; k < kgal; k++)
      for (int i = 0; i < igal - 2; i++) {
        size_t o = (size_t)k * igal + i;
        PROF(0, j) += P1[o];
        PROF(1, j) += P1[o] * P1[o];
        PROF(4, j) += P3[o];
        PROF(5, j) += P3[o] * P3[o];
        PROF(8, j) += P1[o] * P2[o];
        PROF(9, j) += P1[o] * P3[o];
      }
    /* v rows */
    if (j <= nyv - 1) {
      modes_to_plane(s, s->u2, nyv, j, P1);
      modes_to_plane(s, s->u3i, nyv, j, P3);
      if (j >= 1 && j <= ny - 2) {
        lho_four_to_phys_plane(s, P1);
        lho_four_to_phys_plane(s, P3);
      }
      for (int k = 0; k < kgal; k++)
        for (int i = 0; i < igal - 2; i++) {
          size_t o = (size_t)k * igal + i;
          PROF(2, j) += P1[o];
          PROF(3, j) += P1[o] * P1[o];
          PROF(10, j) += P1[o] * P3[o];
        }
      for (int c = 0; c < s->ncol; c++) {
        int k = c / s->nxh;
        cplx u2 = s->u2[(size_t)c * nyv + j];
        cplx dz = (I * s->kz[k]) * u2;
        cplx dy = (s->u3[(size_t)c * nyu + j + 1] - s->u3[(size_t)c * nyu + j]) * s->dthdyv[j] * s->ddthetavi;
        wc[c] = -dz + dy;
      }
      memset(W, 0, sizeof(double) * (size_t)ig * N2);
      col_to_planeN(s, wc, W);
      four_to_phys_N(s, W);
      for (int k = 0; k < N2; k++)
        for (int i = 0; i < N1; i++) {
          double w = W[(size_t)k * ig + i];
          PROF(11, j) += w;
          PROF(12, j) += w * w;
        }
    }
    /* p rows: the spectral plane */
    if (j >= 1 && j <= ny - 1) {
      modes_to_plane(s, s->p, nyu, j, P1);
      for (int k = 0; k < kgal; k++)
        for (int i = 0; i < igal - 2; i++) {
          size_t o = (size_t)k * igal + i;
          PROF(6, j) += P1[o];
          PROF(7, j) += P1[o] * P1[o];
        }
    }
  }
#undef PROF
  /* spectra: sp[w][c*nyu + j], w = U, V, W, UV, P */
  const size_t fs = (size_t)s->ncol * nyu;
  for (int c = 0; c < s->ncol; c++)
    for (int j = 0; j < nyu; j++) {
      size_t o = (size_t)c * nyu + j;
      cplx a = s->u1[o], w = s->u3[o], vi = s->u2i[o];
      sp[0 * fs + o] = sp[0 * fs + o] + creal(a) * creal(a) + cimag(a) * cimag(a);
      sp[2 * fs + o] = sp[2 * fs + o] + creal(w) * creal(w) + cimag(w) * cimag(w);
      sp[3 * fs + o] = sp[3 * fs + o] + creal(a) * creal(vi) + cimag(a) * cimag(vi);
      if (j <= nyv - 1) {
        cplx v = s->u2[(size_t)c * nyv + j];
        sp[1 * fs + o] = sp[1 * fs + o] + creal(v) * creal(v) + cimag(v) * cimag(v);
      }
      if (j >= 1 && j <= ny - 1) {
        cplx p = s->p[o];
        sp[4 * fs + o] = sp[4 * fs + o] + creal(p) * creal(p) + cimag(p) * cimag(p);
      }
    }
  free(P1);
  free(P2);
  free(P3);
  free(W);
  free(wc);
}

/* sl_stats (sl_stats2.f90:33-161): one sample of the wall maps, in the
 * (N1+2) x N2 plane layout of modes_to_planes_phys_lims_2 (spectral
 * coefficients, re/im interleaved): u1 and u3 interpolated to the wall
 * faces (interp_u) and their wall-normal derivatives, the top wall's with
 * the sign flipped.  maps[w][q][(N1+2)*N2], w = 0 bottom (v row 0), 1 top
 * (v row ny-1), q = u1, du1dy, u3, du3dy; added to what maps holds. */
void lho_slstats(lho_state *s, double *maps) {
  const int nyu = s->nyu, ig = s->c.N1 + 2;
  const size_t msz = (size_t)ig * s->c.N2;
  interp_all(s);
  for (int w = 0; w < 2; w++) {
    const int j = w == 0 ? 0 : s->nyv - 1;
    const double sg = w == 0 ? 1.0 : -1.0;
    for (int c = 0; c < s->ncol; c++) {
      int i = c % s->nxh, k = c / s->nxh;
      const cplx *u1 = s->u1 + (size_t)c * nyu, *u3 = s->u3 + (size_t)c * nyu;
      cplx v[4];
      v[0] = s->u1i[(size_t)c * s->nyv + j];
      v[1] = sg * ((u1[j + 1] - u1[j]) * s->dthdyv[j] * s->ddthetavi);
      v[2] = s->u3i[(size_t)c * s->nyv + j];
      v[3] = sg * ((u3[j + 1] - u3[j]) * s->dthdyv[j] * s->ddthetavi);
      for (int q = 0; q < 4; q++) {
        double *m = maps + (size_t)(w * 4 + q) * msz + (size_t)k * ig + 2 * i;
        m[0] = m[0] + creal(v[q]);
        m[1] = m[1] + cimag(v[q]);
      }
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Multi-resolution bands.                                                   */
struct lhob_state {
  lho_state *b[3];  /* bottom (wall resolution), middle, top (wall resolution) */
  int row0[3];      /* global row of each band's local row 0 */
  int nb;           /* N(3,1): last v row of the bottom band */
  int N1m, N2m;     /* middle-band modes */
};

/* signed z wavenumber index of column c of s (k1F_z, def_k start.f90:
 * 1107-1130: indices N2/2..N2-1 are the negative modes) */
static int zmode(const lho_state *s, int c) {
  int k = c / s->nxh;
  return k < s->c.N2 / 2 ? k : k - s->c.N2;
}
/* a mode of the middle band's resolution: the long pencils of
 * proc_lims_columns (start.f90:1624-1648): i <= N1m/2, -N2m/2 <= kz < N2m/2 */
static int is_long(const lhob_state *B, int i, int m) { return i <= B->N1m / 2 && m >= -B->N2m / 2 && m < B->N2m / 2; }

int lhob_owns(const lhob_state *B, int q, int c) {
  const lho_state *s = B->b[q];
  int l = is_long(B, c % s->nxh, zmode(s, c));
  return q == 1 ? 1 : !l;
}
lho_state *lhob_band(lhob_state *B, int q) { return B->b[q]; }
int lhob_row0(const lhob_state *B, int q) { return B->row0[q]; }

lhob_state *lhob_create(const lho_cfg *wall, int N1m, int N2m, int Ngal1m, int Ngal2m, int nb) {
  int ny = wall->ny;
  if (nb < 1 || 2 * (nb + 2) + 2 > ny || N1m > wall->N1 || N2m > wall->N2) return NULL;
  lhob_state *B = calloc(1, sizeof(lhob_state));
  B->nb = nb;
  B->N1m = N1m;
  B->N2m = N2m;
  /* the middle band: the whole channel at the middle resolution */
  lho_cfg cm = *wall;
  cm.N1 = N1m;
  cm.N2 = N2m;
  cm.Ngal1 = Ngal1m;
  cm.Ngal2 = Ngal2m;
  cm.band_edge = 0;
  B->b[1] = lho_create(&cm);
  /* the wall bands: v rows 0..nb+1 and ny-2-nb..ny-1 (jlim, start.f90:
   * 1677-1711), so nb+2 v rows and nb+3 u rows each, at the wall
   * resolution; no mean-flow work (the mean mode is a long pencil) */
  lho_cfg cw = *wall;
  cw.ny = nb + 2;
  cw.flag_ctpress = 1;
  cw.mpgx = 0.0;
  B->row0[0] = 0;
  B->row0[1] = 0;
  B->row0[2] = ny - 2 - nb;
  const lho_state *m = B->b[1];
  for (int q = 0; q < 3; q += 2) {
    cw.band_edge = q == 0 ? 2 : 1;
    lho_state *s = lho_create(&cw);
    int r0 = B->row0[q];
    lho_set_grid(s, m->yu + r0, m->yv + r0, m->dthdyu + r0, m->dthdyv + r0);
    B->b[q] = s;
  }
  return B;
}

void lhob_destroy(lhob_state *B) {
  if (!B) return;
  for (int q = 0; q < 3; q++) lho_destroy(B->b[q]);
  free(B);
}

void lhob_init_flowrate(lhob_state *B) { lho_init_flowrate(B->b[1]); }

/* band of global u row j (planes N(4,q-1)+1..N(4,q), proc_lims_planes
 * start.f90:1940-1965; N(4,1) = nb+1, N(4,2) = ny-2-nb) and v row j
 * (N(3,q-1)+1..N(3,q); N(3,1) = nb, N(3,2) = ny-2-nb) */
static int uband(const lhob_state *B, int j) {
  int ny = B->b[1]->c.ny;
  return j <= B->nb + 1 ? 0 : (j <= ny - 2 - B->nb ? 1 : 2);
}
static int vband(const lhob_state *B, int j) {
  int ny = B->b[1]->c.ny;
  return j <= B->nb ? 0 : (j <= ny - 2 - B->nb ? 1 : 2);
}

/* modes_to_planes_UVP for one plane of band q (FOU3D.f90:1000-1098): the
 * long pencils from the middle band's row j and, in a wall band, that
 * band's short pencils from its local row; column (i, kz) to P(2i+1, row)
 * with negative kz at the top of the plane band's Ngal2 rows (dk,
 * start.f90:1720-1730) */
static void band_gather(const lhob_state *B, int q, int vgrid, const cplx *fm, const cplx *fw, int j, double *P) {
  const lho_state *S = B->b[q], *M = B->b[1];
  int igal = S->igal, kgal = S->kgal;
  memset(P, 0, sizeof(double) * (size_t)igal * kgal);
  int rm = vgrid ? M->nyv : M->nyu;
  for (int c = 0; c < M->ncol; c++) {
    int i = c % M->nxh, m = zmode(M, c), kr = m >= 0 ? m : m + kgal;
    cplx v = fm[(size_t)c * rm + j];
    P[(size_t)kr * igal + 2 * i] = creal(v);
    P[(size_t)kr * igal + 2 * i + 1] = cimag(v);
  }
  if (q == 1) return;
  int rw = vgrid ? S->nyv : S->nyu, jl = j - B->row0[q];
  for (int c = 0; c < S->ncol; c++) {
    int i = c % S->nxh, m = zmode(S, c);
    if (is_long(B, i, m)) continue;
    int kr = m >= 0 ? m : m + kgal;
    cplx v = fw[(size_t)c * rw + jl];
    P[(size_t)kr * igal + 2 * i] = creal(v);
    P[(size_t)kr * igal + 2 * i + 1] = cimag(v);
  }
}
/* planes_to_modes_UVP / _NUVP (FOU3D.f90:2295-2442,2772-2854): the inverse */
static void band_scatter(const lhob_state *B, int q, int vgrid, const double *P, cplx *fm, cplx *fw, int j) {
  const lho_state *S = B->b[q], *M = B->b[1];
  int igal = S->igal, kgal = S->kgal;
  int rm = vgrid ? M->nyv : M->nyu;
  for (int c = 0; c < M->ncol; c++) {
    int i = c % M->nxh, m = zmode(M, c), kr = m >= 0 ? m : m + kgal;
    fm[(size_t)c * rm + j] = P[(size_t)kr * igal + 2 * i] + I * P[(size_t)kr * igal + 2 * i + 1];
  }
  if (q == 1) return;
  int rw = vgrid ? S->nyv : S->nyu, jl = j - B->row0[q];
  for (int c = 0; c < S->ncol; c++) {
    int i = c % S->nxh, m = zmode(S, c);
    if (is_long(B, i, m)) continue;
    int kr = m >= 0 ? m : m + kgal;
    fw[(size_t)c * rw + jl] = P[(size_t)kr * igal + 2 * i] + I * P[(size_t)kr * igal + 2 * i + 1];
  }
}

/* nonlinear (FOU3D.f90:95-283) with ops_in_planes2 (FOU3D.f90:2090-2251)
 * over the bands: every plane at its band's resolution */
void lhob_nonlinear(lhob_state *B, int kRK) {
  lho_state *M = B->b[1];
  int ny = M->c.ny;
  for (int q = 0; q < 3; q++) interp_all(B->b[q]);
  /* u-planes: uu, uw, vv, ww -> Nu1, Nu3 (plane part) and vv */
#pragma omp parallel for schedule(dynamic)
  for (int j = 1; j <= ny - 1; j++) {
    int q = uband(B, j);
    lho_state *S = B->b[q];
    size_t psz = (size_t)S->igal * S->kgal;
    double *P1 = malloc(sizeof(double) * psz), *P2 = malloc(sizeof(double) * psz);
    double *P3 = malloc(sizeof(double) * psz), *Q = malloc(sizeof(double) * psz);
    double *R = malloc(sizeof(double) * psz);
    band_gather(B, q, 0, M->u1, S->u1, j, P1);
    band_gather(B, q, 0, M->u2i, S->u2i, j, P2);
    band_gather(B, q, 0, M->u3, S->u3, j, P3);
    lho_four_to_phys_plane(S, P1);
    lho_four_to_phys_plane(S, P2);
    lho_four_to_phys_plane(S, P3);
    for (int k = 0; k < S->c.Ngal2; k++)
      for (int i = 0; i < S->c.Ngal1; i++) {
        size_t o = (size_t)k * S->igal + i;
        double a = P1[o], b = P2[o], c = P3[o];
        Q[o] = a * c;
        P1[o] = a * a;
        P2[o] = b * b;
        P3[o] = c * c;
      }
    lho_phys_to_four_plane(S, P1);
    lho_phys_to_four_plane(S, Q);
    lho_phys_to_four_plane(S, P2);
    lho_phys_to_four_plane(S, P3);
    derxz_sum(S, P1, Q, R);
    band_scatter(B, q, 0, R, M->Nu1, S->Nu1, j);
    derxz_sum(S, Q, P3, R);
    band_scatter(B, q, 0, R, M->Nu3, S->Nu3, j);
    band_scatter(B, q, 0, P2, M->vvc, S->vvc, j);
    free(P1); free(P2); free(P3); free(Q); free(R);
  }
  /* v-planes: uv, wv -> Nu2 (plane part), uv, wv */
#pragma omp parallel for schedule(dynamic)
  for (int j = 1; j <= ny - 2; j++) {
    int q = vband(B, j);
    lho_state *S = B->b[q];
    size_t psz = (size_t)S->igal * S->kgal;
    double *P1 = malloc(sizeof(double) * psz), *P2 = malloc(sizeof(double) * psz);
    double *P3 = malloc(sizeof(double) * psz), *R = malloc(sizeof(double) * psz);
    band_gather(B, q, 1, M->u1i, S->u1i, j, P1);
    band_gather(B, q, 1, M->u2, S->u2, j, P2);
    band_gather(B, q, 1, M->u3i, S->u3i, j, P3);
    lho_four_to_phys_plane(S, P1);
    lho_four_to_phys_plane(S, P2);
    lho_four_to_phys_plane(S, P3);
    for (int k = 0; k < S->c.Ngal2; k++)
      for (int i = 0; i < S->c.Ngal1; i++) {
        size_t o = (size_t)k * S->igal + i;
        double v = P2[o];
        P1[o] = P1[o] * v;
        P3[o] = P3[o] * v;
      }
    lho_phys_to_four_plane(S, P1);
    lho_phys_to_four_plane(S, P3);
    derxz_sum(S, P1, P3, R);
    band_scatter(B, q, 1, R, M->Nu2, S->Nu2, j);
    band_scatter(B, q, 1, P1, M->uvf, S->uvf, j);
    band_scatter(B, q, 1, P3, M->wvf, S->wvf, j);
    free(P1); free(P2); free(P3); free(R);
  }
  /* per band: the flux rows no plane fills are zero (the wall faces and
   * ghost centres, and a wall band's edge rows, which planes_to_modes skips,
   * FOU3D.f90:2295-2442), then der_yu_h / der_yv_h over the band's rows
   * (FOU3D.f90:266-283) and du += -gRK N (FOU3D.f90:398-410) */
  for (int q = 0; q < 3; q++) {
    lho_state *s = B->b[q];
    int nyu = s->nyu, nyv = s->nyv;
    double ddv = s->ddthetavi, C1 = -s->gRK[kRK - 1];
#pragma omp parallel for schedule(static)
    for (int c = 0; c < s->ncol; c++) {
      cplx *uv = s->uvf + (size_t)c * nyv, *wv = s->wvf + (size_t)c * nyv, *vv = s->vvc + (size_t)c * nyu;
      uv[0] = uv[nyv - 1] = wv[0] = wv[nyv - 1] = 0;
      vv[0] = vv[nyu - 1] = 0;
      if (!lhob_owns(B, q, c)) continue;
      cplx *N1 = s->Nu1 + (size_t)c * nyu, *N3 = s->Nu3 + (size_t)c * nyu, *N2 = s->Nu2 + (size_t)c * nyv;
      for (int j = 1; j <= nyu - 2; j++) {
        N1[j] = N1[j] + (uv[j] - uv[j - 1]) * s->dthdyu[j] * ddv;
        N3[j] = N3[j] + (wv[j] - wv[j - 1]) * s->dthdyu[j] * ddv;
      }
      for (int j = 1; j <= nyv - 2; j++) N2[j] = N2[j] + (vv[j + 1] - vv[j]) * s->dthdyv[j] * ddv;
      for (int j = 1; j <= nyu - 2; j++) {
        s->du1[(size_t)c * nyu + j] += C1 * N1[j];
        s->du3[(size_t)c * nyu + j] += C1 * N3[j];
      }
      for (int j = 1; j <= nyv - 2; j++) s->du2[(size_t)c * nyv + j] += C1 * N2[j];
    }
  }
  /* fixed dt: the time advances at the first substep */
  if (kRK == 1)
    for (int q = 0; q < 3; q++) B->b[q]->t = B->b[q]->t + B->b[q]->dt;
}

/* one RK3 substep over the bands (littleharsh.f90:176-215): each band's
 * column phases run over its own rows and edge rows (RHS0, solveU_W,
 * solveV, solveP, v_corr loop iband = sband..eband); the flow-rate
 * correction acts on the mean mode, a long pencil */
void lhob_substep(lhob_state *B, int kRK) {
  for (int q = 0; q < 3; q++) lho_rhs0(B->b[q], kRK);
  lhob_nonlinear(B, kRK);
  for (int q = 0; q < 3; q++) {
    lho_solve_uw(B->b[q], kRK);
    lho_solve_v(B->b[q], kRK);
  }
  if (B->b[1]->c.flag_ctpress == 0) lho_meanflow_ctU(B->b[1], kRK);
  for (int q = 0; q < 3; q++) {
    lho_solve_p(B->b[q], kRK);
    lho_v_corr(B->b[q], kRK);
  }
}

void lhob_step(lhob_state *B, int nsteps) {
  for (int n = 0; n < nsteps; n++) {
    for (int kRK = 1; kRK <= 3; kRK++) lhob_substep(B, kRK);
    lho_state *M = B->b[1];
    if (M->c.flag_ctpress == 1) M->Qx = flowrate_c(M, M->u1);
    for (int q = 0; q < 3; q++) B->b[q]->iter = B->b[q]->iter + 1;
  }
}
