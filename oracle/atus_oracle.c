/*
 * atus_oracle.c -- CPU restatement of the ATUS-2 split-step Fourier
 * propagator and Sobolev-gradient minimiser.
 *
 * TEST INFRASTRUCTURE ONLY: this is the checker used by tests/, the smoke
 * check and the CPU-baseline leg of bench.py.  The product never links it.
 *
 * Each function names the reference file:line whose semantics it restates.
 * The FFT is this file's own (iterative radix-2, powers of two; a direct
 * O(n^2) DFT for other lengths); it stands in for FFTW 3.3.8, which the
 * reference links (cft_3d.cpp:41-42) and which is not installed here.
 */
#define _GNU_SOURCE
#include "atus_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------ */
/* grid helpers                                                        */
/* ------------------------------------------------------------------ */

long long orc_npts(const orc_grid *g) { return g->n[0] * g->n[1] * g->n[2]; }

int orc_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* CRT_shared::Read_header m_ar / m_ar_k (CRT_shared.h:142-169) */
double orc_ar(const orc_grid *g) {
  double a = g->dx[0];
  if (g->dim > 1) a *= g->dx[1];
  if (g->dim > 2) a *= g->dx[2];
  return a;
}
double orc_ar_k(const orc_grid *g) {
  double a = g->dk[0];
  if (g->dim > 1) a *= g->dk[1];
  if (g->dim > 2) a *= g->dk[2];
  return a;
}

static void idx3(const orc_grid *g, long long l, long long ijk[3]) {
  long long nyz = g->n[1] * g->n[2];
  ijk[0] = l / nyz;
  ijk[1] = (l - ijk[0] * nyz) / g->n[2];
  ijk[2] = l - ijk[0] * nyz - ijk[1] * g->n[2];
}

/* cft_3d::Get_x (cft_3d.cpp:128-138); cft_1d.cpp:85-90 for 1D */
void orc_get_x(const orc_grid *g, long long l, double x[3]) {
  long long ijk[3];
  idx3(g, l, ijk);
  for (int a = 0; a < 3; a++) {
    long long shift = g->n[a] / 2;
    x[a] = (a < g->dim) ? (double)(ijk[a] - shift) * g->dx[a] : 0.0;
  }
}

/* cft_3d::Get_k with m_bfix == false (cft_3d.cpp:145-166) */
void orc_get_k(const orc_grid *g, long long l, double k[3]) {
  long long ijk[3];
  idx3(g, l, ijk);
  for (int a = 0; a < 3; a++) {
    long long n = g->n[a], shift = n / 2;
    long long t = (ijk[a] + shift) % n;
    k[a] = (a < g->dim) ? g->dk[a] * (double)(t - shift) : 0.0;
  }
}

/* ------------------------------------------------------------------ */
/* FFT (stands in for fftw_execute, cft_3d.cpp:64,70)                   */
/* ------------------------------------------------------------------ */

static int is_pow2(long long n) { return n > 0 && (n & (n - 1)) == 0; }

/* twiddle table w[j] = exp(-2 pi i j / n), j < n/2, rounded from long
 * double so each entry is within half an ulp */
static double *make_twiddles(long long n) {
  long long h = n / 2 > 0 ? n / 2 : 1;
  double *w = (double *)malloc(sizeof(double) * 2 * h);
  for (long long j = 0; j < h; j++) {
    long double a = -2.0L * 3.141592653589793238462643383279502884L *
                    (long double)j / (long double)n;
    w[2 * j] = (double)cosl(a);
    w[2 * j + 1] = (double)sinl(a);
  }
  return w;
}

/* direct DFT for lengths that are not powers of two (small n only) */
static void dft_direct(double *x, long long n, int sign, double *tmp) {
  for (long long k = 0; k < n; k++) {
    long double sr = 0, si = 0;
    for (long long j = 0; j < n; j++) {
      long double a = (long double)sign * 2.0L *
                      3.141592653589793238462643383279502884L *
                      (long double)((j * k) % n) / (long double)n;
      long double c = cosl(a), s = sinl(a);
      sr += x[2 * j] * c - x[2 * j + 1] * s;
      si += x[2 * j] * s + x[2 * j + 1] * c;
    }
    tmp[2 * k] = (double)sr;
    tmp[2 * k + 1] = (double)si;
  }
  memcpy(x, tmp, sizeof(double) * 2 * n);
}

#define ORC_BLK 16

/* the radix-2 DIT (bit reversal, then butterflies of growing length with
 * w[j * n/len]) applied to up to ORC_BLK pencils at once, held split as
 * re[i * ORC_BLK + c], im[i * ORC_BLK + c] (element i of pencil c); the
 * loops over c vectorise */
static void fft_pow2_blk(double *re, double *im, long long n, int sign,
                         const double *w, int cw) {
  for (long long i = 1, j = 0; i < n; i++) {
    long long bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j)
      for (int c = 0; c < cw; c++) {
        double tr = re[i * ORC_BLK + c], ti = im[i * ORC_BLK + c];
        re[i * ORC_BLK + c] = re[j * ORC_BLK + c];
        im[i * ORC_BLK + c] = im[j * ORC_BLK + c];
        re[j * ORC_BLK + c] = tr;
        im[j * ORC_BLK + c] = ti;
      }
  }
  for (long long len = 2; len <= n; len <<= 1) {
    long long half = len >> 1, step = n / len;
    for (long long i = 0; i < n; i += len) {
      for (long long j = 0; j < half; j++) {
        const double wr = w[2 * j * step];
        const double wi = (sign < 0) ? w[2 * j * step + 1] : -w[2 * j * step + 1];
        double *ar = re + (i + j) * ORC_BLK, *ai = im + (i + j) * ORC_BLK;
        double *br = re + (i + j + half) * ORC_BLK, *bi = im + (i + j + half) * ORC_BLK;
        for (int c = 0; c < ORC_BLK; c++) {
          const double tr = br[c] * wr - bi[c] * wi;
          const double ti = br[c] * wi + bi[c] * wr;
          br[c] = ar[c] - tr;
          bi[c] = ai[c] - ti;
          ar[c] += tr;
          ai[c] += ti;
        }
      }
    }
  }
}

/* what the last pass does to each element at its store (the cft scale,
 * then optionally a pointwise complex or real multiplier) */
typedef struct {
  int on;
  double fak;
  const double *ctab; /* complex multiplier, interleaved, or NULL */
  const double *rmul; /* real multiplier, or NULL */
} orc_post;

/* pencils p0 .. p0+cw-1 of a power-of-two axis of length n whose elements
 * lie `inner` apart (pencil p starts at (p / inner) * n * inner + p % inner
 * of `data`); `post` applies at the store, indexed by the element's offset
 * from `whole` (the start of the full array) */
static void axis_block(double *data, const double *whole, long long n,
                       long long inner, long long p0, int cw, int sign,
                       const double *w, double *re, double *im,
                       const orc_post *post) {
  long long start[ORC_BLK];
  for (int c = 0; c < cw; c++) {
    const long long p = p0 + c;
    start[c] = (p / inner) * n * inner + p % inner;
  }
  for (long long i = 0; i < n; i++) {
    for (int c = 0; c < cw; c++) {
      const double *s = data + 2 * (start[c] + i * inner);
      re[i * ORC_BLK + c] = s[0];
      im[i * ORC_BLK + c] = s[1];
    }
    for (int c = cw; c < ORC_BLK; c++) {
      re[i * ORC_BLK + c] = 0.0;
      im[i * ORC_BLK + c] = 0.0;
    }
  }
  fft_pow2_blk(re, im, n, sign, w, cw);
  for (long long i = 0; i < n; i++)
    for (int c = 0; c < cw; c++) {
      double *d = data + 2 * (start[c] + i * inner);
      double vr = re[i * ORC_BLK + c], vi = im[i * ORC_BLK + c];
      if (post && post->on) {
        /* scale_all, then the step's multiply, in that order */
        vr *= post->fak;
        vi *= post->fak;
        const long long l = (d - whole) / 2;
        if (post->ctab) {
          const double tr = post->ctab[2 * l], ti = post->ctab[2 * l + 1];
          const double xr = vr;
          vr = xr * tr - vi * ti;
          vi = vi * tr + xr * ti;
        } else if (post->rmul) {
          vr *= post->rmul[l];
          vi *= post->rmul[l];
        }
      }
      d[0] = vr;
      d[1] = vi;
    }
}

/* direct transform along axis `a` for a length that is not a power of two
 * (small test grids only), pencil by pencil */
static void dft_axis(const orc_grid *g, double *data, int a, int sign) {
  long long n = g->n[a];
  long long inner = 1, outer = 1;
  for (int b = a + 1; b < 3; b++) inner *= g->n[b];
  for (int b = 0; b < a; b++) outer *= g->n[b];
#pragma omp parallel
  {
    double *line = (double *)malloc(sizeof(double) * 2 * n);
    double *tmp = (double *)malloc(sizeof(double) * 2 * n);
#pragma omp for schedule(static)
    for (long long p = 0; p < outer * inner; p++) {
      double *base = data + 2 * ((p / inner) * n * inner + p % inner);
      for (long long i = 0; i < n; i++) {
        line[2 * i] = base[2 * i * inner];
        line[2 * i + 1] = base[2 * i * inner + 1];
      }
      dft_direct(line, n, sign, tmp);
      for (long long i = 0; i < n; i++) {
        base[2 * i * inner] = line[2 * i];
        base[2 * i * inner + 1] = line[2 * i + 1];
      }
    }
    free(line);
    free(tmp);
  }
}

/* every pencil of axis `a` over the whole array (ORC_BLK adjacent columns
 * at a time for a strided axis, so every read touches whole cache lines) */
static void pow2_axis(const orc_grid *g, double *data, int a, int sign,
                      const double *w, const orc_post *post) {
  long long n = g->n[a];
  long long inner = 1, outer = 1;
  for (int b = a + 1; b < 3; b++) inner *= g->n[b];
  for (int b = 0; b < a; b++) outer *= g->n[b];
  const long long npen = outer * inner;
  const long long nblk = (npen + ORC_BLK - 1) / ORC_BLK;
#pragma omp parallel
  {
    double *re = (double *)malloc(sizeof(double) * n * ORC_BLK);
    double *im = (double *)malloc(sizeof(double) * n * ORC_BLK);
#pragma omp for schedule(static)
    for (long long b = 0; b < nblk; b++) {
      const long long p0 = b * ORC_BLK;
      const int cw = (int)(npen - p0 < ORC_BLK ? npen - p0 : ORC_BLK);
      axis_block(data, data, n, inner, p0, cw, sign, w, re, im, post);
    }
    free(re);
    free(im);
  }
}

static void scale_all(double *data, long long npts, double fak) {
#pragma omp parallel for schedule(static)
  for (long long i = 0; i < 2 * npts; i++) data[i] *= fak;
}

static void apply_post(double *data, long long np, const orc_post *post) {
  scale_all(data, np, post->fak);
  if (post->ctab) {
#pragma omp parallel for schedule(static)
    for (long long l = 0; l < np; l++) {
      double re = data[2 * l], im = data[2 * l + 1];
      data[2 * l] = re * post->ctab[2 * l] - im * post->ctab[2 * l + 1];
      data[2 * l + 1] = im * post->ctab[2 * l] + re * post->ctab[2 * l + 1];
    }
  } else if (post->rmul) {
#pragma omp parallel for schedule(static)
    for (long long l = 0; l < np; l++) {
      data[2 * l] *= post->rmul[l];
      data[2 * l + 1] *= post->rmul[l];
    }
  }
}

/* the unscaled transform along axes 0, 1, 2 in that order, then `post`.
 * With power-of-two axes the last axis applies `post` at its store, and
 * on 3D grids axes 1 and 2 run plane by plane while the plane is in cache;
 * each element sees the same operations in the same order either way */
static void fft_post(const orc_grid *g, double *data, int sign, const orc_post *post) {
  int last = -1, all_pow2 = 1;
  for (int a = 0; a < 3; a++) {
    if (g->n[a] == 1) continue;
    last = a;
    if (!is_pow2(g->n[a])) all_pow2 = 0;
  }
  if (last < 0 || !all_pow2) {
    for (int a = 0; a < 3; a++)
      if (g->n[a] > 1) {
        if (is_pow2(g->n[a])) {
          double *w = make_twiddles(g->n[a]);
          pow2_axis(g, data, a, sign, w, NULL);
          free(w);
        } else {
          dft_axis(g, data, a, sign);
        }
      }
    if (post && post->on) apply_post(data, orc_npts(g), post);
    return;
  }
  double *w[3] = {NULL, NULL, NULL};
  for (int a = 0; a < 3; a++)
    if (g->n[a] > 1) w[a] = make_twiddles(g->n[a]);
  const long long n0 = g->n[0], n1 = g->n[1], n2 = g->n[2];
  if (n0 > 1 && n1 > 1 && n2 > 1) {
    pow2_axis(g, data, 0, sign, w[0], NULL);
    const long long plane = n1 * n2;
#pragma omp parallel
    {
      const long long nm = n1 > n2 ? n1 : n2;
      double *re = (double *)malloc(sizeof(double) * nm * ORC_BLK);
      double *im = (double *)malloc(sizeof(double) * nm * ORC_BLK);
#pragma omp for schedule(static)
      for (long long i0 = 0; i0 < n0; i0++) {
        double *pl = data + 2 * i0 * plane;
        for (long long p0 = 0; p0 < n2; p0 += ORC_BLK) {
          const int cw = (int)(n2 - p0 < ORC_BLK ? n2 - p0 : ORC_BLK);
          axis_block(pl, data, n1, n2, p0, cw, sign, w[1], re, im, NULL);
        }
        for (long long p0 = 0; p0 < n1; p0 += ORC_BLK) {
          const int cw = (int)(n1 - p0 < ORC_BLK ? n1 - p0 : ORC_BLK);
          axis_block(pl, data, n2, 1, p0, cw, sign, w[2], re, im, post);
        }
      }
      free(re);
      free(im);
    }
  } else {
    for (int a = 0; a < 3; a++)
      if (g->n[a] > 1) pow2_axis(g, data, a, sign, w[a], a == last ? post : NULL);
  }
  for (int a = 0; a < 3; a++) free(w[a]);
}

void orc_fft(const orc_grid *g, double *data, int sign) {
  fft_post(g, data, sign, NULL);
}

/* the scale() factor of cft_Nd (cft_1d.cpp:193, cft_2d.cpp:451,
 * cft_3d.cpp:629), same floating-point expression per dimension */
static double ft_factor(const orc_grid *g, const double *s) {
  switch (g->dim) {
  case 1: return s[0] / sqrt(2.0 * M_PI);
  case 2: return 0.5 * s[0] * s[1] / M_PI;
  default: return s[0] * s[1] * s[2] / pow(2 * M_PI, 1.5);
  }
}

/* cft_Nd::ft (cft_3d.cpp:58-74, m_bfix == false path): FFTW's unscaled
 * transform, then scale(); ctab / rmul: the pointwise multiply a caller
 * applies right after the transform (NULL: none) */
static void ft_mul(const orc_grid *g, double *data, int isign, const double *ctab, const double *rmul) {
  if (isign != 1 && isign != -1) return;
  orc_post post = {1, ft_factor(g, isign == -1 ? g->dx : g->dk), ctab, rmul};
  fft_post(g, data, isign, &post);
}

void orc_ft(const orc_grid *g, double *data, int isign) {
  ft_mul(g, data, isign, NULL, NULL);
}

/* ------------------------------------------------------------------ */
/* real-time propagator                                                */
/* ------------------------------------------------------------------ */

/* CRT_Base::Init (CRT_Base.h:253-274): phi = -dt * k.scale(alpha).k, the
 * table holds cos/sin(s * phi) with s = 0.5 (half) or 1 (full) */
void orc_kinetic_table(const orc_grid *g, const double alpha[3], double dt,
                       double s, double *tab) {
  long long np = orc_npts(g);
  const double mdt = -dt;
#pragma omp parallel for schedule(static)
  for (long long l = 0; l < np; l++) {
    double k[3];
    orc_get_k(g, l, k);
    double ks = 0;
    for (int a = 0; a < g->dim; a++) ks += (k[a] * alpha[a]) * k[a];
    double phi = mdt * ks;
    double ph = (s == 1.0) ? phi : s * phi;
    tab[2 * l] = cos(ph);
    tab[2 * l + 1] = sin(ph);
  }
}

/* CRT_Base::Do_FT_Step_full / _half (CRT_Base.h:337-392) */
void orc_ft_step(const orc_grid *g, int ncomp, double **psi,
                 const double *tab) {
  /* ft(-1), psi *= tab, ft(+1) per component (the components are
   * independent, so each runs through all three before the next) */
  for (int c = 0; c < ncomp; c++) {
    ft_mul(g, psi[c], -1, tab, NULL);
    orc_ft(g, psi[c], 1);
  }
}

/* CRT_Base::Do_NL_Step (CRT_Base.h:397-441): all phases from the
 * pre-update densities, then psi_i *= exp(i phi_i) */
void orc_nl_step(const orc_grid *g, int ncomp, double **psi, const double *gs,
                 double dt, double **pot) {
  long long np = orc_npts(g);
  const double mdt = -dt;
#pragma omp parallel for schedule(static)
  for (long long l = 0; l < np; l++) {
    double phi[4];
    for (int i = 0; i < ncomp; i++) {
      phi[i] = (pot && pot[i]) ? pot[i][l] : 0.0;
      for (int j = 0; j < ncomp; j++)
        phi[i] += gs[j + ncomp * i] *
                  (psi[j][2 * l] * psi[j][2 * l] +
                   psi[j][2 * l + 1] * psi[j][2 * l + 1]);
      phi[i] *= mdt;
    }
    for (int i = 0; i < ncomp; i++) {
      double s, c;
      sincos(phi[i], &s, &c);
      double re = psi[i][2 * l], im = psi[i][2 * l + 1];
      psi[i][2 * l] = re * c - im * s;
      psi[i][2 * l + 1] = im * c + re * s;
    }
  }
}

/* CRT_Base_IF::Do_NL_Step (CRT_Base_IF.h:381-419): gravity along the first
 * axis and per-component detuning; the reference loop is serial, the result
 * does not depend on the order */
void orc_nl_if_step(const orc_grid *g, int ncomp, double **psi,
                    const double *gs, double dt, double beta0,
                    const double *deltaL) {
  long long np = orc_npts(g);
  const double mdt = -dt;
#pragma omp parallel for schedule(static)
  for (long long l = 0; l < np; l++) {
    double phi[4], x[3];
    orc_get_x(g, l, x);
    for (int i = 0; i < ncomp; i++) {
      phi[i] = 0;
      for (int j = 0; j < ncomp; j++)
        phi[i] += gs[j + ncomp * i] *
                  (psi[j][2 * l] * psi[j][2 * l] +
                   psi[j][2 * l + 1] * psi[j][2 * l + 1]);
      phi[i] += beta0 * x[0] - deltaL[i];
      phi[i] *= mdt;
    }
    for (int i = 0; i < ncomp; i++) {
      double s, c;
      sincos(phi[i], &s, &c);
      double re = psi[i][2 * l], im = psi[i][2 * l + 1];
      psi[i][2 * l] = re * c - im * s;
      psi[i][2 * l + 1] = im * c + re * s;
    }
  }
}

/* Bragg_single::Do_Bragg_ad (bragg.cpp:186-304): per point, the analytic
 * exponential of the 2x2 Hermitian H = [[V11, Om*eta],[Om*conj(eta), V22]]
 * applied to (psi1, psi2).  Operation order follows the reference so that
 * rounding matches closely. */
void orc_bragg_ad(const orc_grid *g, double *psi1, double *psi2,
                  const orc_bragg_params *p) {
  long long np = orc_npts(g);
  const double dt = p->dt;
  const double t1 = p->t + 0.5 * dt;
  const double F = p->F;
#pragma omp parallel for schedule(static)
  for (long long l = 0; l < np; l++) {
    double x[3];
    orc_get_x(g, l, x);
    double *a = psi1 + 2 * l, *b = psi2 + 2 * l;
    double n1 = a[0] * a[0] + a[1] * a[1];
    double n2 = b[0] * b[0] + b[1] * b[1];
    double V11 = p->gs[0] * n1 + p->gs[1] * n2 + p->beta0 * x[0];
    double V22 = p->gs[2] * n1 + p->gs[3] * n2 - p->deltaL1 + p->beta0 * x[0];
    double Om = F * p->amp0 *
                cos(p->laser_k * x[0] -
                    (p->laser_domh + p->chirp * t1 + p->chirp_rate0 * t1) * t1 +
                    p->phase0 / 2);
    double s, c;
    if (Om == 0.0) {
      sincos(-dt * V11, &s, &c);
      double re = a[0];
      a[0] = re * c - a[1] * s;
      a[1] = re * s + a[1] * c;
      sincos(-dt * V22, &s, &c);
      re = b[0];
      b[0] = re * c - b[1] * s;
      b[1] = re * s + b[1] * c;
      continue;
    }
    double etr, eti;
    sincos(-0.5 * (p->laser_dk * x[0]), &eti, &etr);
    double r = sqrt((V11 - V22) * (V11 - V22) + 4.0 * Om * Om);
    double Ep = 0.5 * (V11 + V22 + r);
    double Em = 0.5 * (V11 + V22 - r);
    double gpr, gpi, gmr, gmi, nrm;
    sincos(-dt * Ep, &s, &c);
    nrm = 1.0 / fabs((V11 - Ep) * (V11 - Ep) + Om * Om);
    gpr = nrm * c;
    gpi = nrm * s;
    sincos(-dt * Em, &s, &c);
    nrm = 1.0 / fabs((V11 - Em) * (V11 - Em) + Om * Om);
    gmr = nrm * c;
    gmi = nrm * s;
    double o2 = Om * Om;
    double O11r = o2 * (gpr + gmr), O11i = o2 * (gpi + gmi);
    double dp = V11 - Ep, dm = V11 - Em;
    double O22r = dp * dp * gpr + dm * dm * gmr;
    double O22i = dp * dp * gpi + dm * dm * gmi;
    double Xr = -Om * (dp * gpr + dm * gmr);
    double Xi = -Om * (dp * gpi + dm * gmi);
    /* O12 = X * eta, O21 = X * conj(eta) */
    double O12r = etr * Xr - eti * Xi, O12i = eti * Xr + etr * Xi;
    double O21r = etr * Xr + eti * Xi, O21i = etr * Xi - eti * Xr;
    double ar = a[0], ai = a[1], br = b[0], bi = b[1];
    a[0] = (O11r * ar - O11i * ai) + (O12r * br - O12i * bi);
    a[1] = (O11r * ai + O11i * ar) + (O12r * bi + O12i * br);
    b[0] = (O21r * ar - O21i * ai) + (O22r * br - O22i * bi);
    b[1] = (O21r * ai + O21i * ar) + (O22r * bi + O22i * br);
  }
}

/* ------------------------------------------------------------------ */
/* reductions                                                          */
/* ------------------------------------------------------------------ */

/* CRT_Base::Get_Particle_Number (CRT_Base.h:587-599) */
double orc_particle_number(const orc_grid *g, const double *psi) {
  long long np = orc_npts(g);
  double s = 0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (long long l = 0; l < np; l++)
    s += psi[2 * l] * psi[2 * l] + psi[2 * l + 1] * psi[2 * l + 1];
  return orc_ar(g) * s;
}

/* CRT_Base::Expval_Position (CRT_Base.h:481-525) */
void orc_expval_x(const orc_grid *g, const double *psi, double out[3]) {
  long long np = orc_npts(g);
  double s0 = 0, s1 = 0, s2 = 0;
#pragma omp parallel for reduction(+ : s0, s1, s2) schedule(static)
  for (long long l = 0; l < np; l++) {
    double x[3];
    orc_get_x(g, l, x);
    double d = psi[2 * l] * psi[2 * l] + psi[2 * l + 1] * psi[2 * l + 1];
    s0 += x[0] * d;
    s1 += x[1] * d;
    s2 += x[2] * d;
  }
  double ar = orc_ar(g);
  out[0] = ar * s0;
  out[1] = g->dim > 1 ? ar * s1 : 0.0;
  out[2] = g->dim > 2 ? ar * s2 : 0.0;
}

/* CRT_Base::Expval_Momentum (CRT_Base.h:533-580) */
void orc_expval_k(const orc_grid *g, double *psi, double out[3]) {
  long long np = orc_npts(g);
  orc_ft(g, psi, -1);
  double s0 = 0, s1 = 0, s2 = 0;
#pragma omp parallel for reduction(+ : s0, s1, s2) schedule(static)
  for (long long l = 0; l < np; l++) {
    double k[3];
    orc_get_k(g, l, k);
    double d = psi[2 * l] * psi[2 * l] + psi[2 * l + 1] * psi[2 * l + 1];
    s0 += k[0] * d;
    s1 += k[1] * d;
    s2 += k[2] * d;
  }
  orc_ft(g, psi, 1);
  double ark = orc_ar_k(g);
  out[0] = ark * s0;
  out[1] = g->dim > 1 ? ark * s1 : 0.0;
  out[2] = g->dim > 2 ? ark * s2 : 0.0;
}

/* CRT_Base_IF::compute_rabi_integrals (CRT_Base_IF.h:208-281), including
 * the reference's slot counter, which advances only on a miss */
void orc_rabi_integrals(const orc_grid *g, double *psi, const double *k0,
                        int nk0, double threshold, double *out) {
  long long np = orc_npts(g);
  orc_ft(g, psi, -1);
  double *acc = (double *)calloc((size_t)nk0, sizeof(double));
#pragma omp parallel
  {
    double *loc = (double *)calloc((size_t)nk0, sizeof(double));
#pragma omp for schedule(static)
    for (long long l = 0; l < np; l++) {
      double k[3];
      orc_get_k(g, l, k);
      int j = 0;
      for (int q = 0; q < nk0; q++) {
        double dd = 0;
        for (int a = 0; a < g->dim; a++) {
          double d = k0[3 * q + a] - k[a];
          dd += d * d;
        }
        if (sqrt(dd) < threshold) {
          loc[j] += psi[2 * l] * psi[2 * l] + psi[2 * l + 1] * psi[2 * l + 1];
          continue;
        }
        j++;
      }
    }
#pragma omp critical
    for (int q = 0; q < nk0; q++) acc[q] += loc[q];
    free(loc);
  }
  double ark = orc_ar_k(g);
  for (int q = 0; q < nk0; q++) out[q] = ark * acc[q];
  free(acc);
  orc_ft(g, psi, 1);
}

/* Energy functional.  The reference defines no energy; this is
 *   E = ar_k sum_i sum_k (sum_a alpha_a k_a^2) |psi_i^(k)|^2
 *     + ar sum_i sum_x [ V_i |psi_i|^2 + 1/2 sum_j g_ij |psi_i|^2 |psi_j|^2 ]
 * with psi^ = ft(-1) psi, i.e. the Hamiltonian whose kinetic propagator is
 * the reference's exp(-i dt alpha k^2) (CRT_Base.h:253-274). */
double orc_energy(const orc_grid *g, int ncomp, double **psi, const double *gs,
                  const double alpha[3], double **pot) {
  long long np = orc_npts(g);
  double ekin = 0, epot = 0;
  double *tmp = (double *)malloc(sizeof(double) * 2 * np);
  for (int i = 0; i < ncomp; i++) {
    memcpy(tmp, psi[i], sizeof(double) * 2 * np);
    orc_ft(g, tmp, -1);
    double s = 0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (long long l = 0; l < np; l++) {
      double k[3];
      orc_get_k(g, l, k);
      double ks = 0;
      for (int a = 0; a < g->dim; a++) ks += alpha[a] * k[a] * k[a];
      s += ks * (tmp[2 * l] * tmp[2 * l] + tmp[2 * l + 1] * tmp[2 * l + 1]);
    }
    ekin += s;
  }
  free(tmp);
#pragma omp parallel for reduction(+ : epot) schedule(static)
  for (long long l = 0; l < np; l++) {
    double n[4];
    for (int i = 0; i < ncomp; i++)
      n[i] = psi[i][2 * l] * psi[i][2 * l] + psi[i][2 * l + 1] * psi[i][2 * l + 1];
    for (int i = 0; i < ncomp; i++) {
      double v = (pot && pot[i]) ? pot[i][l] : 0.0;
      double nl = 0;
      for (int j = 0; j < ncomp; j++) nl += gs[j + ncomp * i] * n[j];
      epot += (v + 0.5 * nl) * n[i];
    }
  }
  return orc_ar_k(g) * ekin + orc_ar(g) * epot;
}

/* ------------------------------------------------------------------ */
/* sequence block                                                      */
/* ------------------------------------------------------------------ */

void orc_rt_set_dt(orc_rt *s, double dt) {
  s->dt = dt;
  orc_kinetic_table(&s->g, s->alpha, dt, 0.5, s->half_tab);
  orc_kinetic_table(&s->g, s->alpha, dt, 1.0, s->full_tab);
  if (s->two_species) {
    /* CRT_Base_2::Init (CRT_Base_2.h:195-221), second pair of tables from Alpha_2 */
    orc_kinetic_table(&s->g, s->alpha2, dt, 0.5, s->half_tab2);
    orc_kinetic_table(&s->g, s->alpha2, dt, 1.0, s->full_tab2);
  }
  s->tab_dt = dt;
}

void orc_ft_step2(const orc_grid *g, int ncomp, double **psi, const double *tab1, const double *tab2) {
  for (int c = 0; c < ncomp; c++) {
    ft_mul(g, psi[c], -1, c < ncomp / 2 ? tab1 : tab2, NULL);
    orc_ft(g, psi[c], 1);
  }
}

void orc_nl_if2_step(const orc_grid *g, int ncomp, double **psi, const double *gs, double dt, double beta,
                     double beta2, const double *deltaL) {
  long long np = orc_npts(g);
  const double mdt = -dt;
#pragma omp parallel for schedule(static)
  for (long long l = 0; l < np; l++) {
    double x[3], phi[4];
    orc_get_x(g, l, x);
    for (int i = 0; i < ncomp; i++) {
      phi[i] = 0;
      for (int j = 0; j < ncomp; j++) {
        const double re = psi[j][2 * l], im = psi[j][2 * l + 1];
        phi[i] += gs[j + ncomp * i] * (re * re + im * im);
      }
      phi[i] += (i < ncomp / 2 ? beta : beta2) * x[0] - deltaL[i];
      phi[i] *= mdt;
    }
    for (int i = 0; i < ncomp; i++) {
      double si, co;
      sincos(phi[i], &si, &co);
      const double re = psi[i][2 * l];
      psi[i][2 * l] = psi[i][2 * l] * co - psi[i][2 * l + 1] * si;
      psi[i][2 * l + 1] = psi[i][2 * l + 1] * co + re * si;
    }
  }
}

/* psi_i *= exp(-i dt phi_i) as in the Omega == 0 branches of bragg_ds.cpp:152-160 */
static void ds_diag(double **psi, long long l, int i, double phi, double dt) {
  double im1, re1;
  sincos(-phi * dt, &im1, &re1);
  const double tmp1 = psi[i][2 * l];
  psi[i][2 * l] = tmp1 * re1 - psi[i][2 * l + 1] * im1;
  psi[i][2 * l + 1] = tmp1 * im1 + psi[i][2 * l + 1] * re1;
}

/* the 2x2 exponential of one species (bragg_ds.cpp:174-206), std::complex
 * arithmetic written out (products with a zero imaginary part are exact) */
static void ds_pair(double **psi, long long l, int i0, double V11, double V22, double Om, double theta,
                    double dt) {
  const double er = cos(theta), ei = sin(theta);
  const double tmp1 = sqrt((V11 - V22) * (V11 - V22) + 4.0 * Om * Om);
  const double Ep = 0.5 * (V11 + V22 + tmp1);
  const double Em = 0.5 * (V11 + V22 - tmp1);
  const double d0 = V11 - Ep, d1 = V11 - Em;
  const double den0 = fabs(d0 * d0 + Om * Om), den1 = fabs(d1 * d1 + Om * Om);
  const double g0r = cos(-dt * Ep) / den0, g0i = sin(-dt * Ep) / den0;
  const double g1r = cos(-dt * Em) / den1, g1i = sin(-dt * Em) / den1;
  const double tr = d0 * g0r + d1 * g1r, ti = d0 * g0i + d1 * g1i;
  const double o2 = Om * Om;
  const double M00r = o2 * (g0r + g1r), M00i = o2 * (g0i + g1i);
  const double ar = -Om * er, ai = -Om * ei;          /* -Omega * eta */
  const double M01r = ar * tr - ai * ti, M01i = ar * ti + ai * tr;
  const double cr = -Om * er, ci = -Om * (-ei);       /* -Omega * conj(eta) */
  const double M10r = cr * tr - ci * ti, M10i = cr * ti + ci * tr;
  const double dd0 = d0 * d0, dd1 = d1 * d1;
  const double M11r = dd0 * g0r + dd1 * g1r, M11i = dd0 * g0i + dd1 * g1i;
  const double p0r = psi[i0][2 * l], p0i = psi[i0][2 * l + 1];
  const double p1r = psi[i0 + 1][2 * l], p1i = psi[i0 + 1][2 * l + 1];
  psi[i0][2 * l] = (M00r * p0r - M00i * p0i) + (M01r * p1r - M01i * p1i);
  psi[i0][2 * l + 1] = (M00r * p0i + M00i * p0r) + (M01r * p1i + M01i * p1r);
  psi[i0 + 1][2 * l] = (M10r * p0r - M10i * p0i) + (M11r * p1r - M11i * p1i);
  psi[i0 + 1][2 * l + 1] = (M10r * p0i + M10i * p0r) + (M11r * p1i + M11i * p1r);
}

void orc_bragg_ds(const orc_grid *g, double **psi, const orc_rt *s, double t1, int mode1, int mode2) {
  long long np = orc_npts(g);
  const double dt = s->dt;
#pragma omp parallel for schedule(static)
  for (long long l = 0; l < np; l++) {
    double x[3], phi[4];
    orc_get_x(g, l, x);
    for (int i = 0; i < 4; i++) {
      phi[i] = 0;
      for (int j = 0; j < 4; j++) {
        const double re = psi[j][2 * l], im = psi[j][2 * l + 1];
        phi[i] += s->gs[j + 4 * i] * (re * re + im * im);
      }
      phi[i] += (i < 2 ? s->beta0 : s->beta2) * x[0] - s->deltaL[i];
    }
    const double Om = mode1 * s->amp0 * cos(s->laser_k * x[0] - (s->laser_domh + s->chirp_rate0 * t1) * t1 + s->phase0 / 2);
    const double Om2 =
        mode2 * s->amp2 * cos(s->laser_k2 * x[0] - (s->laser_domh2 + s->chirp_rate2 * t1) * t1 + s->phase2 / 2);
    const double th1 = -0.5 * (s->laser_dk * x[0] + s->phase0);
    const double th2 = -0.5 * (s->laser_dk2 * x[0] + s->phase2);
    if (Om == 0 && Om2 == 0) {
      for (int i = 0; i < 4; i++) ds_diag(psi, l, i, phi[i], dt);
    } else if (Om == 0.0) {
      for (int i = 0; i < 2; i++) ds_diag(psi, l, i, phi[i], dt);
      ds_pair(psi, l, 2, phi[2], phi[3], Om2, th2, dt);
    } else if (Om2 == 0.0) {
      for (int i = 2; i < 4; i++) ds_diag(psi, l, i, phi[i], dt);
      ds_pair(psi, l, 0, phi[0], phi[1], Om, th1, dt);
    } else {
      ds_pair(psi, l, 0, phi[0], phi[1], Om, th1, dt);
      ds_pair(psi, l, 2, phi[2], phi[3], Om2, th2, dt);
    }
  }
}

static void v_step(orc_rt *s, int kind) {
  if (kind == ORC_FREEPROP_IF2) {
    orc_nl_if2_step(&s->g, s->ncomp, s->psi, s->gs, s->dt, s->beta0, s->beta2, s->deltaL);
    return;
  }
  if (kind == ORC_BRAGG_DS) {
    /* pulse gates (bragg_ds.cpp:120-128): a pulse is off once t1 - seq.time
     * exceeds its duration */
    const double t1 = s->t + 0.5 * s->dt;
    const double time = t1 - s->seq_time;
    orc_bragg_ds(&s->g, s->psi, s, t1, time > s->duration[0] ? 0 : 1, time > s->duration[1] ? 0 : 1);
    return;
  }
  if (kind == ORC_DOUBLE_BRAGG) {
    const double t1 = s->t + 0.5 * s->dt;
    orc_double_bragg(&s->g, s->psi, s, t1, s->amp ? s->amp(t1, s->amp_user) : 1.0);
    return;
  }
  if (kind == ORC_NUM_BRAGG) {
    orc_num_bragg(&s->g, s->ncomp, s->psi, s, s->t);
    return;
  }
  if (kind == ORC_RAMAN) {
    orc_raman(&s->g, s->ncomp, s->psi, s);
    return;
  }
  if (kind == ORC_FREEPROP) {
    orc_nl_step(&s->g, s->ncomp, s->psi, s->gs, s->dt, s->pot);
  } else if (kind == ORC_FREEPROP_IF) {
    orc_nl_if_step(&s->g, s->ncomp, s->psi, s->gs, s->dt, s->beta0, s->deltaL);
  } else {
    orc_bragg_params p;
    p.dt = s->dt;
    p.t = s->t;
    /* Amplitude_at_time (CRT_Base_IF.h:171-198) at Get_t() + 0.5*Get_dt() */
    p.F = s->amp ? s->amp(s->t + 0.5 * s->dt, s->amp_user) : 1.0;
    p.amp0 = s->amp0;
    p.laser_k = s->laser_k;
    p.laser_domh = s->laser_domh;
    p.laser_dk = s->laser_dk;
    p.chirp = s->chirp;
    p.chirp_rate0 = s->chirp_rate0;
    p.phase0 = s->phase0;
    p.beta0 = s->beta0;
    p.deltaL1 = s->deltaL[1];
    for (int q = 0; q < 4; q++) p.gs[q] = s->gs[q];
    orc_bragg_ad(&s->g, s->psi[0], s->psi[1], &p);
  }
}

/* ------------------------------------------------------------------ */
/* three-state steps                                                   */
/* ------------------------------------------------------------------ */

void orc_herm_expm(int n, const double *H, double dt, double *U) {
  long double ar[4][4], ai[4][4], vr[4][4], vi[4][4];
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) {
      ar[i][j] = H[2 * (i * n + j)];
      ai[i][j] = H[2 * (i * n + j) + 1];
      vr[i][j] = i == j;
      vi[i][j] = 0;
    }
  for (int it = 0; it < 200; it++) {
    /* the largest off-diagonal element */
    int p = 0, q = 1;
    long double big = -1, dg = 0;
    for (int i = 0; i < n; i++) {
      dg += ar[i][i] * ar[i][i];
      for (int j = i + 1; j < n; j++) {
        long double m = ar[i][j] * ar[i][j] + ai[i][j] * ai[i][j];
        if (m > big) {
          big = m;
          p = i;
          q = j;
        }
      }
    }
    if (n < 2 || big <= 1e-44L * dg || big == 0) break;
    /* H_pq = r e^{i phi}; with D = diag(1, e^{-i phi}) on (p, q) the pair
     * is real symmetric, R = [[c, s], [-s, c]] zeroes it; J = D R */
    long double r = sqrtl(big);
    long double cr = ar[p][q] / r, ci = -ai[p][q] / r; /* e^{-i phi} */
    long double th = (ar[q][q] - ar[p][p]) / (2 * r);
    long double t = (th >= 0 ? 1.0L : -1.0L) / (fabsl(th) + sqrtl(1 + th * th));
    long double c = 1 / sqrtl(1 + t * t), sn = t * c;
    for (int k = 0; k < n; k++) { /* columns: H J */
      long double hpr = ar[k][p], hpi = ai[k][p];
      long double hqr = cr * ar[k][q] - ci * ai[k][q], hqi = cr * ai[k][q] + ci * ar[k][q];
      ar[k][p] = c * hpr - sn * hqr;
      ai[k][p] = c * hpi - sn * hqi;
      ar[k][q] = sn * hpr + c * hqr;
      ai[k][q] = sn * hpi + c * hqi;
    }
    for (int k = 0; k < n; k++) { /* rows: J^H (H J), conj(e^{-i phi}) = e^{i phi} */
      long double bpr = ar[p][k], bpi = ai[p][k];
      long double bqr = cr * ar[q][k] + ci * ai[q][k], bqi = cr * ai[q][k] - ci * ar[q][k];
      ar[p][k] = c * bpr - sn * bqr;
      ai[p][k] = c * bpi - sn * bqi;
      ar[q][k] = sn * bpr + c * bqr;
      ai[q][k] = sn * bpi + c * bqi;
    }
    ar[p][q] = ai[p][q] = ar[q][p] = ai[q][p] = 0;
    ai[p][p] = ai[q][q] = 0;
    for (int k = 0; k < n; k++) { /* V J */
      long double vpr = vr[k][p], vpi = vi[k][p];
      long double vqr = cr * vr[k][q] - ci * vi[k][q], vqi = cr * vi[k][q] + ci * vr[k][q];
      vr[k][p] = c * vpr - sn * vqr;
      vi[k][p] = c * vpi - sn * vqi;
      vr[k][q] = sn * vpr + c * vqr;
      vi[k][q] = sn * vpi + c * vqi;
    }
  }
  /* U = V diag(exp(-i dt lambda)) V^H */
  long double er[4], ei[4];
  for (int i = 0; i < n; i++) {
    er[i] = cosl(-(long double)dt * ar[i][i]);
    ei[i] = sinl(-(long double)dt * ar[i][i]);
  }
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) {
      long double sr = 0, si = 0;
      for (int k = 0; k < n; k++) {
        /* V_ik e_k conj(V_jk) */
        long double xr = vr[i][k] * er[k] - vi[i][k] * ei[k], xi = vr[i][k] * ei[k] + vi[i][k] * er[k];
        sr += xr * vr[j][k] + xi * vi[j][k];
        si += xi * vr[j][k] - xr * vi[j][k];
      }
      U[2 * (i * n + j)] = (double)sr;
      U[2 * (i * n + j) + 1] = (double)si;
    }
}

/* beta.x as CPoint::operator* sums it (CPoint.h:145-151) */
static double beta_dot(const orc_grid *g, const double *beta, const double *x) {
  double r = 0;
  for (int a = 0; a < g->dim; a++) r += beta[a] * x[a];
  return r;
}

static void apply_u(int n, const double *U, double **psi, long long l) {
  double in[8], out[8];
  for (int i = 0; i < n; i++) {
    in[2 * i] = psi[i][2 * l];
    in[2 * i + 1] = psi[i][2 * l + 1];
  }
  for (int i = 0; i < n; i++) {
    double sr = 0, si = 0;
    for (int k = 0; k < n; k++) {
      const double ur = U[2 * (i * n + k)], ui = U[2 * (i * n + k) + 1];
      sr += ur * in[2 * k] - ui * in[2 * k + 1];
      si += ur * in[2 * k + 1] + ui * in[2 * k];
    }
    out[2 * i] = sr;
    out[2 * i + 1] = si;
  }
  for (int i = 0; i < n; i++) {
    psi[i][2 * l] = out[2 * i];
    psi[i][2 * l + 1] = out[2 * i + 1];
  }
}

/* diagonal: phi_i = sum_j g_ij |psi_j|^2, then += beta.x - DeltaL_i
 * (CRT_Base_IF.h:455-466) */
static void three_state_diag(int n, double *H, double **psi, long long l, const orc_rt *s, double bx) {
  for (int i = 0; i < 2 * n * n; i++) H[i] = 0;
  for (int i = 0; i < n; i++) {
    double phi = 0;
    for (int j = 0; j < n; j++) {
      const double re = psi[j][2 * l], im = psi[j][2 * l + 1];
      phi += s->gs[j + n * i] * (re * re + im * im);
    }
    phi += bx - s->deltaL[i];
    H[2 * (i * n + i)] = phi;
  }
}

void orc_num_bragg(const orc_grid *g, int ncomp, double **psi, const orc_rt *s, double t) {
  long long np = orc_npts(g);
  const int n = ncomp;
  /* laser_k[1] = -laser_k[0], chirp_rate[1] = -chirp_rate[0] (:447-448) */
  const double lk[2] = {s->laser_k, -s->laser_k};
  const double cr[2] = {s->chirp_rate0, -s->chirp_rate0};
#pragma omp parallel for schedule(static)
  for (long long l = 0; l < np; l++) {
    double x[3], H[32], U[32];
    orc_get_x(g, l, x);
    three_state_diag(n, H, psi, l, s, beta_dot(g, s->beta, x));
    for (int i = 0; i < n - 1; i++) {
      const double th = (-s->laser_domh + cr[i] * t) * t + lk[i] * x[0] - 0.5 * s->phase0;
      const double re1 = cos(th), im1 = sin(th);
      const double e0 = s->amp0 * re1 / 2 + s->amp1 * re1 / 2;
      const double e1 = s->amp0 * im1 / 2 - s->amp1 * im1 / 2;
      H[2 * ((i + 1) * n + 0)] = e0;
      H[2 * ((i + 1) * n + 0) + 1] = e1;
      H[2 * (0 * n + i + 1)] = e0;
      H[2 * (0 * n + i + 1) + 1] = -e1;
    }
    orc_herm_expm(n, H, s->dt, U);
    apply_u(n, U, psi, l);
  }
}

void orc_raman(const orc_grid *g, int ncomp, double **psi, const orc_rt *s) {
  long long np = orc_npts(g);
  const int n = ncomp;
#pragma omp parallel for schedule(static)
  for (long long l = 0; l < np; l++) {
    double x[3], H[32], U[32];
    orc_get_x(g, l, x);
    three_state_diag(n, H, psi, l, s, beta_dot(g, s->beta, x));
    H[2 * (2 * n + 2)] += s->laser_domh;
    const double re1 = cos(s->laser_k * x[0]), im1 = sin(s->laser_k * x[0]);
    H[2 * (0 * n + 2)] = s->amp0 / 2 * re1;
    H[2 * (0 * n + 2) + 1] = s->amp0 / 2 * im1;
    H[2 * (2 * n + 0)] = s->amp0 / 2 * re1;
    H[2 * (2 * n + 0) + 1] = -(s->amp0 / 2 * im1);
    H[2 * (1 * n + 2)] = s->amp1 / 2 * re1;
    H[2 * (1 * n + 2) + 1] = -(s->amp1 / 2 * im1);
    H[2 * (2 * n + 1)] = s->amp1 / 2 * re1;
    H[2 * (2 * n + 1) + 1] = s->amp1 / 2 * im1;
    orc_herm_expm(n, H, s->dt, U);
    apply_u(n, U, psi, l);
  }
}

/* the reference's closed form restated term by term; O23 and O32 take
 * their eta factors as written there (double_bragg.cpp:216-222) */
void orc_double_bragg(const orc_grid *g, double **psi, const orc_rt *s, double t1, double F) {
  long long np = orc_npts(g);
  const double dt = s->dt;
  if (F < 0) F = 0;
#pragma omp parallel for schedule(static)
  for (long long l = 0; l < np; l++) {
    double x[3];
    orc_get_x(g, l, x);
    double *P1 = psi[0] + 2 * l, *P2 = psi[1] + 2 * l, *P3 = psi[2] + 2 * l;
    double tmp1 = P1[0] * P1[0] + P1[1] * P1[1];
    double tmp2 = P2[0] * P2[0] + P2[1] * P2[1];
    double tmp3 = P3[0] * P3[0] + P3[1] * P3[1];
    const double bx = beta_dot(g, s->beta, x);
    const double V11 = s->gs[0] * tmp1 + s->gs[1] * tmp2 + s->gs[2] * tmp3 + bx;
    const double V22 = s->gs[3] * tmp1 + s->gs[4] * tmp2 + s->gs[5] * tmp3 - s->deltaL[1] + bx;
    double dw = s->laser_domh + s->chirp_rate0 * t1;
    const double Op = F * s->amp0 * cos(s->laser_k * x[0] - dw * t1 - s->phase0 / 2);
    dw = s->laser_domh - s->chirp_rate0 * t1;
    const double Om = F * s->amp0 * cos(-s->laser_k * x[0] - dw * t1 - s->phase0 / 2);
    double re1, im1;
    if (Om == 0.0 && Op == 0.0) {
      re1 = cos(-dt * V11);
      im1 = sin(-dt * V11);
      tmp1 = P1[0];
      P1[0] = tmp1 * re1 - P1[1] * im1;
      P1[1] = tmp1 * im1 + P1[1] * re1;
      re1 = cos(-dt * V22);
      im1 = sin(-dt * V22);
      tmp1 = P2[0];
      P2[0] = tmp1 * re1 - P2[1] * im1;
      P2[1] = tmp1 * im1 + P2[1] * re1;
      tmp1 = P3[0];
      P3[0] = tmp1 * re1 - P3[1] * im1;
      P3[1] = tmp1 * im1 + P3[1] * re1;
      continue;
    }
    const double eta0 = cos(-0.5 * s->laser_dk * x[0]), eta1 = sin(-0.5 * s->laser_dk * x[0]);
    tmp1 = sqrt((V11 - V22) * (V11 - V22) + 4.0 * (Op * Op + Om * Om));
    const double E1 = 0.5 * (V11 + V22 + tmp1), E2 = 0.5 * (V11 + V22 - tmp1);
    double G1[2], G2[2], G3[2];
    tmp1 = 1.0 / fabs((V22 - E1) * (V22 - E1) + Op * Op + Om * Om);
    G1[0] = tmp1 * cos(-dt * E1);
    G1[1] = tmp1 * sin(-dt * E1);
    tmp1 = 1.0 / fabs((V22 - E2) * (V22 - E2) + Op * Op + Om * Om);
    G2[0] = tmp1 * cos(-dt * E2);
    G2[1] = tmp1 * sin(-dt * E2);
    tmp1 = 1.0 / fabs(Op * Op + Om * Om);
    G3[0] = tmp1 * cos(-dt * V22);
    G3[1] = tmp1 * sin(-dt * V22);
    double O11[2], O12[2], O13[2], O21[2], O22[2], O23[2], O31[2], O32[2], O33[2];
    tmp1 = (V22 - E2) * (V22 - E2);
    tmp2 = (V22 - E1) * (V22 - E1);
    for (int c = 0; c < 2; c++) O11[c] = tmp1 * G2[c] + tmp2 * G1[c];
    const double w[2] = {(V22 - E2) * G2[0] + (V22 - E1) * G1[0], (V22 - E2) * G2[1] + (V22 - E1) * G1[1]};
    for (int c = 0; c < 2; c++) {
      O12[c] = -Op * w[c];
      O13[c] = -Om * w[c];
      O21[c] = O12[c];
      O31[c] = O13[c];
    }
    /* eta * z for O12 and O31, conj(eta) * z for O21 and O13 */
    double z0 = O12[0], z1 = O12[1];
    O12[0] = eta0 * z0 - eta1 * z1;
    O12[1] = eta1 * z0 + eta0 * z1;
    z0 = O21[0];
    z1 = O21[1];
    O21[0] = eta0 * z0 + eta1 * z1;
    O21[1] = eta0 * z1 - eta1 * z0;
    z0 = O31[0];
    z1 = O31[1];
    O31[0] = eta0 * z0 - eta1 * z1;
    O31[1] = eta1 * z0 + eta0 * z1;
    z0 = O13[0];
    z1 = O13[1];
    O13[0] = eta0 * z0 + eta1 * z1;
    O13[1] = eta0 * z1 - eta1 * z0;
    const double pp = Op * Op, mm = Om * Om;
    for (int c = 0; c < 2; c++) {
      O22[c] = pp * (G2[c] + G1[c]) + mm * G3[c];
      O33[c] = mm * (G2[c] + G1[c]) + pp * G3[c];
      O23[c] = (Om * Op) * (G1[c] + G2[c] - G3[c]);
      O32[c] = O23[c];
    }
    const double sq = eta0 * eta0 + eta1 * eta1, df = eta0 * eta0 - eta1 * eta1, cr2 = 2 * eta0 * eta1;
    z0 = O23[0];
    O23[0] = sq * z0 + cr2 * O23[1];
    O23[1] = sq * O23[1] - cr2 * z0;
    z0 = O32[0];
    O32[0] = df * z0 - cr2 * O32[1];
    O32[1] = df * O23[1] + cr2 * z0; /* O23[1] after its update, as written */
    const double a1[2] = {P1[0], P1[1]}, a2[2] = {P2[0], P2[1]}, a3[2] = {P3[0], P3[1]};
#define ORC_ROW(O1, O2, O3, D)                                                                              \
  D[0] = (O1[0] * a1[0] - O1[1] * a1[1]) + (O2[0] * a2[0] - O2[1] * a2[1]) + (O3[0] * a3[0] - O3[1] * a3[1]); \
  D[1] = (O1[0] * a1[1] + O1[1] * a1[0]) + (O2[0] * a2[1] + O2[1] * a2[0]) + (O3[0] * a3[1] + O3[1] * a3[0]);
    ORC_ROW(O11, O12, O13, P1)
    ORC_ROW(O21, O22, O23, P2)
    ORC_ROW(O31, O32, O33, P3)
#undef ORC_ROW
  }
}

/* CRT_Base::run_sequence hot loop (CRT_Base.h:776-786); the time
 * bookkeeping is Do_FT_Step_half: t += 0.5*dt (:391), full: t += dt (:361) */
static void kin_step(orc_rt *s, int full) {
  if (s->two_species)
    orc_ft_step2(&s->g, s->ncomp, s->psi, full ? s->full_tab : s->half_tab, full ? s->full_tab2 : s->half_tab2);
  else
    orc_ft_step(&s->g, s->ncomp, s->psi, full ? s->full_tab : s->half_tab);
}

/* SplitMix64 (Steele, Lea & Flood 2014) of one 64-bit counter */
static unsigned long long splitmix64(unsigned long long z) {
  z += 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

/* the rounding model of orc_rt (jitter > 0) */
static void jitter_all(orc_rt *s) {
  if (!(s->jitter > 0.0)) return;
  const long long np = orc_npts(&s->g);
  const unsigned long long key = splitmix64(s->jitter_seed ^ splitmix64((unsigned long long)s->jitter_count));
  s->jitter_count++;
  for (int c = 0; c < s->ncomp; c++) {
    double *p = s->psi[c];
    const unsigned long long kc = splitmix64(key + (unsigned long long)c);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < 2 * np; i++) {
      const double u = (double)(splitmix64(kc + (unsigned long long)i) >> 11) * 0x1.0p-52 - 1.0;
      p[i] *= 1.0 + s->jitter * u;
    }
  }
}

void orc_rt_block(orc_rt *s, int kind, int Nk) {
  kin_step(s, 0);
  jitter_all(s);
  s->t += 0.5 * s->dt;
  for (int j = 2; j <= Nk; j++) {
    v_step(s, kind);
    jitter_all(s);
    kin_step(s, 1);
    jitter_all(s);
    s->t += s->dt;
  }
  v_step(s, kind);
  jitter_all(s);
  kin_step(s, 0);
  jitter_all(s);
  s->t += 0.5 * s->dt;
}

/* ------------------------------------------------------------------ */
/* Sobolev gradient minimiser (CSOB_Base.h)                            */
/* ------------------------------------------------------------------ */

/* CSOB_Base::Init (CSOB_Base.h:229-248): op = 1/(1+alpha|k|^2),
 * Lap = -alpha|k|^2 with isotropic alpha = Alpha_i[0] */
void orc_sob_init_ops(orc_sob *s) {
  long long np = orc_npts(&s->g);
  for (int j = 0; j < s->nwf; j++) {
    double al = s->alpha[j];
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < np; i++) {
      double k[3];
      orc_get_k(&s->g, i, k);
      double phi = 0;
      for (int a = 0; a < s->g.dim; a++) phi += (k[a] * al) * k[a];
      s->op[j][i] = 1.0 / (1.0 + phi);
      s->lapop[j][i] = -phi;
    }
  }
}

static double sob_pn(const orc_sob *s, int i) {
  return orc_particle_number(&s->g, s->psi[i]);
}

/* CSOB_Base::Renormalize_All_Psi (CSOB_Base.h:523-543) */
void orc_sob_renormalize(orc_sob *s) {
  long long np = orc_npts(&s->g);
  for (int i = 0; i < s->nwf; i++) s->N[i] = sob_pn(s, i);
  for (int i = 0; i < s->nwf; i++) {
    double fak = 1.0 / sqrt(s->N[i]);
    scale_all(s->psi[i], np, fak);
  }
  for (int i = 0; i < s->nwf; i++) s->N[i] = 1;
}

/* out = ft(+1)[ mult * ft(-1)[in] ]; the reference copies into the field
 * buffer, transforms in place and copies out (CSOB_Base.h:251-301) */
static void sob_filter(const orc_sob *s, const double *in, const double *mult,
                       double *out) {
  long long np = orc_npts(&s->g);
  if (out != in) memcpy(out, in, sizeof(double) * 2 * np);
  ft_mul(&s->g, out, -1, NULL, mult);
  orc_ft(&s->g, out, 1);
}

/* one body of the do-while loop of CSOB_Base::run (CSOB_Base.h:431-503) */
void orc_sob_iter(orc_sob *s) {
  long long np = orc_npts(&s->g);
  int nw = s->nwf;
  /* Compute_Sobolev_Psi (:277-301) */
  for (int i = 0; i < nw; i++) sob_filter(s, s->psi[i], s->op[i], s->sob[i]);
  /* Compute_Laplace (:251-274) */
  for (int i = 0; i < nw; i++) sob_filter(s, s->psi[i], s->lapop[i], s->lap[i]);
  /* Compute_Sobolev_Gradient (:304-337) */
  for (int i = 0; i < nw; i++) {
    double *in = s->grad[i];
#pragma omp parallel for schedule(static)
    for (long long l = 0; l < np; l++) {
      double nl = 0;
      for (int j = 0; j < nw; j++)
        nl += s->gs[j + i * nw] * (s->psi[j][2 * l] * s->psi[j][2 * l] +
                                   s->psi[j][2 * l + 1] * s->psi[j][2 * l + 1]);
      double v = s->pot[i][l] + nl;
      in[2 * l] = -s->lap[i][2 * l] + v * s->psi[i][2 * l];
      in[2 * l + 1] = -s->lap[i][2 * l + 1] + v * s->psi[i][2 * l + 1];
    }
    sob_filter(s, in, s->op[i], in);
  }
  /* Project_Sobolev_Gradient (:340-367) */
  for (int i = 0; i < nw; i++) {
    double t1 = 0, t2 = 0;
    const double *ps = s->psi[i], *sb = s->sob[i];
    double *gr = s->grad[i];
#pragma omp parallel for reduction(+ : t1, t2) schedule(static)
    for (long long l = 0; l < np; l++) {
      t1 += ps[2 * l] * gr[2 * l] + ps[2 * l + 1] * gr[2 * l + 1];
      t2 += ps[2 * l] * sb[2 * l] + ps[2 * l + 1] * sb[2 * l + 1];
    }
    double fak = t1 / t2;
#pragma omp parallel for schedule(static)
    for (long long l = 0; l < np; l++) {
      gr[2 * l] -= fak * sb[2 * l];
      gr[2 * l + 1] -= fak * sb[2 * l + 1];
    }
  }
  /* Compute_res (:370-397); the reference's m_res array has `dim`
   * entries of which only the first no_wf are written: the unwritten ones
   * are taken as zero here */
  s->res_tot = 0;
  for (int i = 0; i < nw; i++) {
    const double *gr = s->grad[i];
    double t = 0;
#pragma omp parallel for reduction(+ : t) schedule(static)
    for (long long l = 0; l < np; l++)
      t += gr[2 * l] * gr[2 * l] + gr[2 * l + 1] * gr[2 * l + 1];
    s->res[i] = t * orc_ar(&s->g);
    s->res_tot += s->res[i] * s->res[i];
  }
  s->res_tot = sqrt(s->res_tot);
  /* norm check and renormalisation after the gradient (:470-481) */
  int brenorm = 0;
  for (int i = 0; i < nw; i++) {
    s->N[i] = sob_pn(s, i);
    if (fabs(s->N[i] - 1) > 1e-5) brenorm = 1;
  }
  s->renormalised = brenorm;
  if (brenorm) orc_sob_renormalize(s);
  /* update (:483-492) */
  for (int i = 0; i < nw; i++) {
    double *ps = s->psi[i];
    const double *gr = s->grad[i];
    const double st = s->stepsize;
#pragma omp parallel for schedule(static)
    for (long long l = 0; l < np; l++) {
      ps[2 * l] -= st * gr[2 * l];
      ps[2 * l + 1] -= st * gr[2 * l + 1];
    }
  }
}

/* CSOB_Base::Compute_mu (CSOB_Base.h:400-424; never called by the
 * reference, Get_Ar() taken as m_ar):
 *   mu_i = ar * sum [ -Re(conj(psi_i) * Lap_i) + (V_i + sum_j g_ij |psi_j|^2) |psi_i|^2 ]
 * with Lap_i = ft(+1)[-alpha |k|^2 ft(-1) psi_i] */
void orc_sob_mu(orc_sob *s, double *mu) {
  long long np = orc_npts(&s->g);
  int nw = s->nwf;
  for (int i = 0; i < nw; i++) sob_filter(s, s->psi[i], s->lapop[i], s->lap[i]);
  for (int i = 0; i < nw; i++) {
    double m = 0;
    const double *ps = s->psi[i], *lp = s->lap[i];
#pragma omp parallel for reduction(+ : m) schedule(static)
    for (long long l = 0; l < np; l++) {
      double nl = 0;
      for (int j = 0; j < nw; j++)
        nl += s->gs[j + i * nw] * (s->psi[j][2 * l] * s->psi[j][2 * l] +
                                   s->psi[j][2 * l + 1] * s->psi[j][2 * l + 1]);
      m += -(ps[2 * l] * lp[2 * l] + ps[2 * l + 1] * lp[2 * l + 1]) +
           (s->pot[i][l] + nl) * (ps[2 * l] * ps[2 * l] + ps[2 * l + 1] * ps[2 * l + 1]);
    }
    mu[i] = orc_ar(&s->g) * m;
  }
}
