/*
 * lmw_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C, single-threaded FP64 restatement of lmweather's dynamical core (reference files cpp/sphere.cpp and
 * cpp/plane.cpp), kept operation-for-operation in the reference's order so that it agrees with the compiled
 * reference bit for bit on x86-64.  It is the checker of the HIP library and the CPU baseline timed beside it:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  The product path never calls it.
 *
 * Every function cites the reference line it follows.  Pinning (tests/test_oracle_*.py):
 *   - against oracle/_ref (the reference's own sphere.cpp/plane.cpp compiled by oracle/Makefile) where that builds,
 *     bitwise on transforms, tendencies, steps and diagnostics;
 *   - against the committed golden vectors in tests/golden/ (written from oracle/_ref);
 *   - against closed forms from the reference's own tests: naive DFT, Gauss-Legendre exactness, Legendre
 *     orthonormality, spectral round trip, solid-body rotation, rest state, TC2 steadiness, gravity-mode frequency,
 *     Parseval, plane gravity-wave speed (tests/core.cpp and the .lms programs under tests/).
 * Complex numbers are (re, im) pairs of doubles; products are written out as (ac - bd, ad + bc), which is what
 * std::complex<double> computes for finite operands, so the rounding matches the reference.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define KPI 3.14159265358979323846
enum { ORC_OK = 0, ORC_EINVAL = 1, ORC_ERUNTIME = 2 };

static __thread char g_err[512];
const char *orc_last_error(void) { return g_err; }
static int fail(int code, const char *msg) { snprintf(g_err, sizeof g_err, "%s", msg); return code; }

typedef struct { double re, im; } cx;
static inline cx cmk(double r, double i) { cx z = {r, i}; return z; }
static inline cx cadd(cx a, cx b) { return cmk(a.re + b.re, a.im + b.im); }
static inline cx csub(cx a, cx b) { return cmk(a.re - b.re, a.im - b.im); }
static inline cx cmul(cx a, cx b) { return cmk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
static inline cx cscale(cx a, double s) { return cmk(a.re * s, a.im * s); }      /* cplx * double */
static inline cx cdiv(cx a, double s) { return cmk(a.re / s, a.im / s); }        /* cplx / double */
static inline cx cimul(double m, cx a) { return cmul(cmk(0.0 * m, 1.0 * m), a); } /* (I * double(m)) * a */

/* eps(n, m): sphere.cpp:11 */
static inline double eps_nm(int n, int m) { return sqrt((double)(n * n - m * m) / (double)(4 * n * n - 1)); }

/* splitmix64: sphere.cpp:13-18, plane.cpp:12-17 */
static uint64_t splitmix(uint64_t *x) {
  uint64_t z = (*x += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
/* Irwin-Hall normal: sphere.cpp:316-320 */
static double irwin_hall(uint64_t *st) {
  double s = 0;
  for (int i = 0; i < 12; i++) s += (double)(splitmix(st) >> 11) * (1.0 / 9007199254740992.0);
  return s - 6.0;
}

/* ------------------------------------------------------------------ free functions (sphere.cpp:21-73) */

/* radix-2 DIT FFT, unnormalized: sphere.cpp:21-45 */
static void fft_cx(cx *a, size_t n, int sign) {
  for (size_t i = 1, j = 0; i < n; i++) {
    size_t bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) { cx t = a[i]; a[i] = a[j]; a[j] = t; }
  }
  for (size_t len = 2; len <= n; len <<= 1) {
    double ang = 2 * KPI / (double)len * sign;
    cx wl = cmk(cos(ang), sin(ang));
    for (size_t i = 0; i < n; i += len) {
      cx w = cmk(1.0, 0.0);
      for (size_t k = 0; k < len / 2; k++) {
        if ((k & 15) == 0) w = cmk(cos(ang * (double)k), sin(ang * (double)k));
        cx u = a[i + k], v = cmul(a[i + k + len / 2], w);
        a[i + k] = cadd(u, v);
        a[i + k + len / 2] = csub(u, v);
        w = cmul(w, wl);
      }
    }
  }
}
int orc_fft(double *re_im, size_t n, int sign) {
  if (n < 2 || (n & (n - 1))) return fail(ORC_EINVAL, "fft length must be a power of two");
  fft_cx((cx *)re_im, n, sign);
  return ORC_OK;
}

/* Gauss-Legendre nodes (ascending) and weights: sphere.cpp:47-64 */
void orc_gauss_legendre(int n, double *x, double *w) {
  memset(x, 0, sizeof(double) * (size_t)n);
  memset(w, 0, sizeof(double) * (size_t)n);
  for (int i = 0; i < (n + 1) / 2; i++) {
    double z = cos(KPI * (i + 0.75) / (n + 0.5)), pp = 1;
    for (int it = 0; it < 100; it++) {
      double p1 = 1, p2 = 0;
      for (int j = 1; j <= n; j++) { double p3 = p2; p2 = p1; p1 = ((2.0 * j - 1) * z * p2 - (j - 1.0) * p3) / j; }
      pp = n * (z * p1 - p2) / (z * z - 1);
      double dz = p1 / pp;
      z -= dz;
      if (fabs(dz) < 1e-15) break;
    }
    x[i] = -z;
    x[n - 1 - i] = z;
    w[i] = w[n - 1 - i] = 2 / ((1 - z * z) * pp * pp);
  }
}

/* orthonormal associated Legendre column n = m..nmax+1: sphere.cpp:66-73 (out has nmax + 2 - m entries) */
void orc_legendre_column(int m, int nmax, double mu, double *out) {
  int len = nmax + 2 - m;
  for (int i = 0; i < len; i++) out[i] = 0.0;
  double c = sqrt(fmax(0.0, 1 - mu * mu)), p = 1 / sqrt(2.0);
  for (int k = 1; k <= m; k++) p *= sqrt((2.0 * k + 1) / (2.0 * k)) * c;
  out[0] = p;
  if (nmax + 1 >= m + 1) out[1] = sqrt(2.0 * m + 3) * mu * p;
  for (int n = m + 2; n <= nmax + 1; n++) out[n - m] = (mu * out[n - 1 - m] - eps_nm(n - 1, m) * out[n - 2 - m]) / eps_nm(n, m);
}

/* ------------------------------------------------------------------ sphere (sphere.hpp:27-121, sphere.cpp:75-652) */

typedef struct orc_sphere {
  int N, nlat, nlon;
  double radius, omega, gravity, diffusion;
  double t;
  double *mu, *w;
  double *P, *H; /* [j][m][n-m], row stride (N+1) per m: P[(j*(N+1) + m)*(N+1) + (n-m)] */
  cx *zeta, *delta, *phi, *q;
  int has_tracer, advect_only, has_orography;
  double *ref_phi, *ref_q, *f, *hs;
  size_t nspec, ngrid;
} orc_sphere;

static inline size_t sidx(const orc_sphere *s, int m, int n) { return (size_t)m * (s->N + 1) + n; }
static inline double Ptab(const orc_sphere *s, int j, int m, int k) { return s->P[((size_t)j * (s->N + 1) + m) * (s->N + 1) + k]; }
static inline double Htab(const orc_sphere *s, int j, int m, int k) { return s->H[((size_t)j * (s->N + 1) + m) * (s->N + 1) + k]; }

static double lon_of(const orc_sphere *s, int k) { return 2 * KPI * k / s->nlon; } /* sphere.cpp:121 */

/* set_axis_tilt: sphere.cpp:111-118 */
void orc_sphere_set_axis_tilt(orc_sphere *s, double alpha) {
  for (int j = 0; j < s->nlat; j++)
    for (int k = 0; k < s->nlon; k++) {
      double lam = lon_of(s, k), c = sqrt(1 - s->mu[j] * s->mu[j]);
      s->f[(size_t)j * s->nlon + k] = 2 * s->omega * (-cos(lam) * c * sin(alpha) + s->mu[j] * cos(alpha));
    }
}

/* ctor: sphere.cpp:75-103 */
orc_sphere *orc_sphere_create(int N, double radius, double omega, double gravity, double diffusion) {
  if (N < 4 || N > 170) { fail(ORC_EINVAL, "truncation must be between 4 and 170"); return NULL; }
  orc_sphere *s = (orc_sphere *)calloc(1, sizeof *s);
  s->N = N; s->radius = radius; s->omega = omega; s->gravity = gravity; s->diffusion = diffusion;
  s->nlon = 4;
  while (s->nlon < 3 * N + 1) s->nlon *= 2;
  s->nlat = s->nlon / 2;
  s->mu = (double *)malloc(sizeof(double) * s->nlat);
  s->w = (double *)malloc(sizeof(double) * s->nlat);
  orc_gauss_legendre(s->nlat, s->mu, s->w);
  size_t tab = (size_t)s->nlat * (N + 1) * (N + 1);
  s->P = (double *)calloc(tab, sizeof(double));
  s->H = (double *)calloc(tab, sizeof(double));
  double *col = (double *)malloc(sizeof(double) * (N + 2));
  for (int j = 0; j < s->nlat; j++)
    for (int m = 0; m <= N; m++) {
      orc_legendre_column(m, N, s->mu[j], col);
      double *Pr = s->P + ((size_t)j * (N + 1) + m) * (N + 1), *Hr = s->H + ((size_t)j * (N + 1) + m) * (N + 1);
      for (int n = m; n <= N; n++) Pr[n - m] = col[n - m];
      for (int n = m; n <= N; n++) {
        double plus = col[n + 1 - m], minus = (n > m) ? col[n - 1 - m] : 0.0;
        Hr[n - m] = -n * eps_nm(n + 1, m) * plus + (n + 1) * eps_nm(n, m) * minus;
      }
    }
  free(col);
  s->nspec = (size_t)(N + 1) * (N + 1);
  s->ngrid = (size_t)s->nlat * s->nlon;
  s->zeta = (cx *)calloc(s->nspec, sizeof(cx));
  s->delta = (cx *)calloc(s->nspec, sizeof(cx));
  s->phi = (cx *)calloc(s->nspec, sizeof(cx));
  s->q = (cx *)calloc(s->nspec, sizeof(cx));
  s->ref_phi = (double *)calloc(s->ngrid, sizeof(double));
  s->ref_q = (double *)calloc(s->ngrid, sizeof(double));
  s->f = (double *)calloc(s->ngrid, sizeof(double));
  s->hs = (double *)calloc(s->ngrid, sizeof(double));
  orc_sphere_set_axis_tilt(s, 0);
  return s;
}
void orc_sphere_destroy(orc_sphere *s) {
  if (!s) return;
  free(s->mu); free(s->w); free(s->P); free(s->H);
  free(s->zeta); free(s->delta); free(s->phi); free(s->q);
  free(s->ref_phi); free(s->ref_q); free(s->f); free(s->hs);
  free(s);
}
void orc_sphere_dims(const orc_sphere *s, int *nlat, int *nlon) { *nlat = s->nlat; *nlon = s->nlon; }
double orc_sphere_time(const orc_sphere *s) { return s->t; }
void orc_sphere_gauss(const orc_sphere *s, double *mu, double *w) {
  memcpy(mu, s->mu, sizeof(double) * s->nlat);
  memcpy(w, s->w, sizeof(double) * s->nlat);
}
void orc_sphere_set_diffusion(orc_sphere *s, double nu4) { s->diffusion = nu4; }
void orc_sphere_set_advect_only(orc_sphere *s, int on) { s->advect_only = on; }

/* set_orography_grid: sphere.cpp:105-109 */
int orc_sphere_set_orography_grid(orc_sphere *s, const double *hs, size_t n) {
  if (n != s->ngrid) return fail(ORC_EINVAL, "grid size mismatch");
  memcpy(s->hs, hs, sizeof(double) * n);
  s->has_orography = 1;
  return ORC_OK;
}

/* fourier_forward / fourier_backward: sphere.cpp:123-136 */
static void fourier_forward(const orc_sphere *s, const double *row, cx *fm, cx *work) {
  for (int k = 0; k < s->nlon; k++) work[k] = cmk(row[k], 0.0);
  fft_cx(work, (size_t)s->nlon, -1);
  for (int m = 0; m <= s->N; m++) fm[m] = cdiv(work[m], (double)s->nlon);
}
static void fourier_backward(const orc_sphere *s, const cx *fm, double *row, cx *work) {
  for (int k = 0; k < s->nlon; k++) work[k] = cmk(0, 0);
  work[0] = cmk(fm[0].re, 0);
  for (int m = 1; m <= s->N; m++) { work[m] = fm[m]; work[s->nlon - m] = cmk(fm[m].re, -fm[m].im); }
  fft_cx(work, (size_t)s->nlon, +1);
  for (int k = 0; k < s->nlon; k++) row[k] = work[k].re;
}

/* analyze_scalar: sphere.cpp:138-149 */
static void analyze_scalar(const orc_sphere *s, const double *grid, cx *spec) {
  const int N = s->N;
  cx *fm = (cx *)malloc(sizeof(cx) * (N + 1)), *work = (cx *)malloc(sizeof(cx) * s->nlon);
  for (size_t i = 0; i < s->nspec; i++) spec[i] = cmk(0, 0);
  for (int j = 0; j < s->nlat; j++) {
    fourier_forward(s, grid + (size_t)j * s->nlon, fm, work);
    for (int m = 0; m <= N; m++) {
      cx f = cscale(fm[m], s->w[j]);
      for (int n = m; n <= N; n++) spec[sidx(s, m, n)] = cadd(spec[sidx(s, m, n)], cscale(f, Ptab(s, j, m, n - m)));
    }
  }
  free(fm); free(work);
}
/* synthesize_scalar: sphere.cpp:150-162 */
static void synthesize_scalar(const orc_sphere *s, const cx *spec, double *grid) {
  const int N = s->N;
  cx *fm = (cx *)malloc(sizeof(cx) * (N + 1)), *work = (cx *)malloc(sizeof(cx) * s->nlon);
  for (int j = 0; j < s->nlat; j++) {
    for (int m = 0; m <= N; m++) {
      cx acc = cmk(0, 0);
      for (int n = m; n <= N; n++) acc = cadd(acc, cscale(spec[sidx(s, m, n)], Ptab(s, j, m, n - m)));
      fm[m] = acc;
    }
    fourier_backward(s, fm, grid + (size_t)j * s->nlon, work);
  }
  free(fm); free(work);
}
/* analyze_divergence: sphere.cpp:164-177 */
static void analyze_divergence(const orc_sphere *s, const double *A, const double *B, cx *spec) {
  const int N = s->N;
  cx *fa = (cx *)malloc(sizeof(cx) * (N + 1)), *fb = (cx *)malloc(sizeof(cx) * (N + 1)), *work = (cx *)malloc(sizeof(cx) * s->nlon);
  for (size_t i = 0; i < s->nspec; i++) spec[i] = cmk(0, 0);
  for (int j = 0; j < s->nlat; j++) {
    fourier_forward(s, A + (size_t)j * s->nlon, fa, work);
    fourier_forward(s, B + (size_t)j * s->nlon, fb, work);
    double wj = s->w[j] / (1 - s->mu[j] * s->mu[j]) / s->radius;
    for (int m = 0; m <= N; m++) {
      cx ca = cscale(cimul((double)m, fa[m]), wj), cb = cscale(fb[m], wj);
      for (int n = m; n <= N; n++) {
        size_t q = sidx(s, m, n);
        spec[q] = cadd(spec[q], csub(cscale(ca, Ptab(s, j, m, n - m)), cscale(cb, Htab(s, j, m, n - m))));
      }
    }
  }
  free(fa); free(fb); free(work);
}
/* winds_spectral_to_grid: sphere.cpp:179-206 */
static void winds_spectral_to_grid(const orc_sphere *s, const cx *z, const cx *d, double *U, double *V) {
  const int N = s->N;
  const double a = s->radius;
  cx *psi = (cx *)calloc(s->nspec, sizeof(cx)), *chi = (cx *)calloc(s->nspec, sizeof(cx));
  for (int m = 0; m <= N; m++)
    for (int n = (m > 1 ? m : 1); n <= N; n++) {
      double k = -a * a / (n * (n + 1.0));
      psi[sidx(s, m, n)] = cscale(z[sidx(s, m, n)], k);
      chi[sidx(s, m, n)] = cscale(d[sidx(s, m, n)], k);
    }
  cx *fu = (cx *)malloc(sizeof(cx) * (N + 1)), *fv = (cx *)malloc(sizeof(cx) * (N + 1)), *work = (cx *)malloc(sizeof(cx) * s->nlon);
  for (int j = 0; j < s->nlat; j++) {
    for (int m = 0; m <= N; m++) {
      cx sp = cmk(0, 0), sc = cmk(0, 0), hp = cmk(0, 0), hc = cmk(0, 0);
      for (int n = m; n <= N; n++) {
        double P = Ptab(s, j, m, n - m), H = Htab(s, j, m, n - m);
        size_t q = sidx(s, m, n);
        sp = cadd(sp, cscale(psi[q], P)); sc = cadd(sc, cscale(chi[q], P));
        hp = cadd(hp, cscale(psi[q], H)); hc = cadd(hc, cscale(chi[q], H));
      }
      fu[m] = cdiv(csub(cimul((double)m, sc), hp), a);
      fv[m] = cdiv(cadd(cimul((double)m, sp), hc), a);
    }
    fourier_backward(s, fu, U + (size_t)j * s->nlon, work);
    fourier_backward(s, fv, V + (size_t)j * s->nlon, work);
  }
  free(psi); free(chi); free(fu); free(fv); free(work);
}
/* analyze_vort_div: sphere.cpp:208-215 */
static void analyze_vort_div(const orc_sphere *s, const double *U, const double *V, cx *z, cx *d) {
  double *mU = (double *)malloc(sizeof(double) * s->ngrid);
  for (size_t i = 0; i < s->ngrid; i++) mU[i] = -U[i];
  analyze_divergence(s, V, mU, z);
  analyze_divergence(s, U, V, d);
  z[0] = cmk(0, 0);
  d[0] = cmk(0, 0);
  free(mU);
}

/* mark_reference: sphere.cpp:236-239 */
void orc_sphere_mark_reference(orc_sphere *s) {
  synthesize_scalar(s, s->phi, s->ref_phi);
  if (s->has_tracer) synthesize_scalar(s, s->q, s->ref_q);
}
/* set_state_grid: sphere.cpp:217-234 */
int orc_sphere_set_state_grid(orc_sphere *s, const double *h, const double *u, const double *v, size_t n) {
  if (n != s->ngrid) return fail(ORC_EINVAL, "grid size mismatch");
  double *U = (double *)malloc(sizeof(double) * n), *V = (double *)malloc(sizeof(double) * n), *ph = (double *)malloc(sizeof(double) * n);
  for (int j = 0; j < s->nlat; j++) {
    double c = sqrt(1 - s->mu[j] * s->mu[j]);
    for (int k = 0; k < s->nlon; k++) {
      size_t i = (size_t)j * s->nlon + k;
      U[i] = u[i] * c;
      V[i] = v[i] * c;
      ph[i] = s->gravity * h[i];
    }
  }
  analyze_vort_div(s, U, V, s->zeta, s->delta);
  analyze_scalar(s, ph, s->phi);
  s->t = 0;
  orc_sphere_mark_reference(s);
  free(U); free(V); free(ph);
  return ORC_OK;
}
/* set_tracer_grid: sphere.cpp:241-246 */
int orc_sphere_set_tracer_grid(orc_sphere *s, const double *q, size_t n) {
  if (n != s->ngrid) return fail(ORC_EINVAL, "grid size mismatch");
  analyze_scalar(s, q, s->q);
  s->has_tracer = 1;
  synthesize_scalar(s, s->q, s->ref_q);
  return ORC_OK;
}
/* tracer_tendency: sphere.cpp:250-263 */
static void tracer_tendency(const orc_sphere *s, const cx *z, const cx *d, const cx *q, cx *dq) {
  const size_t ng = s->ngrid;
  double *U = malloc(sizeof(double) * ng), *V = malloc(sizeof(double) * ng), *qg = malloc(sizeof(double) * ng), *dg = malloc(sizeof(double) * ng);
  winds_spectral_to_grid(s, z, d, U, V);
  synthesize_scalar(s, q, qg);
  synthesize_scalar(s, d, dg);
  double *A = malloc(sizeof(double) * ng), *B = malloc(sizeof(double) * ng), *C = malloc(sizeof(double) * ng);
  for (size_t i = 0; i < ng; i++) { A[i] = U[i] * qg[i]; B[i] = V[i] * qg[i]; C[i] = qg[i] * dg[i]; }
  cx *flux = malloc(sizeof(cx) * s->nspec), *corr = malloc(sizeof(cx) * s->nspec);
  analyze_divergence(s, A, B, flux);
  analyze_scalar(s, C, corr);
  for (size_t i = 0; i < s->nspec; i++) dq[i] = csub(corr[i], flux[i]);
  free(U); free(V); free(qg); free(dg); free(A); free(B); free(C); free(flux); free(corr);
}

/* add_mode: sphere.cpp:301-311 */
int orc_sphere_add_mode(orc_sphere *s, int n, int m, double amplitude, double mean_depth) {
  if (m < 0 || n < (m > 1 ? m : 1) || n > s->N) return fail(ORC_EINVAL, "mode (n, m) needs 0 <= m <= n <= truncation and n >= 1");
  for (size_t i = 0; i < s->nspec; i++) { s->zeta[i] = cmk(0, 0); s->delta[i] = cmk(0, 0); s->phi[i] = cmk(0, 0); }
  s->phi[0] = cmk(s->gravity * mean_depth * sqrt(2.0), 0);
  size_t q = sidx(s, m, n);
  s->phi[q] = cadd(s->phi[q], cmk(s->gravity * amplitude * (m == 0 ? 1.0 : 0.5), 0));
  s->t = 0;
  orc_sphere_mark_reference(s);
  return ORC_OK;
}

/* perturb: sphere.cpp:313-336 */
int orc_sphere_perturb(orc_sphere *s, double amplitude, uint64_t seed, int max_wave) {
  if (max_wave < 1 || max_wave > s->N) return fail(ORC_EINVAL, "max_wave must be between 1 and the truncation");
  uint64_t st = seed * 0x9E3779B97F4A7C15ULL + 12345;
  cx *add = (cx *)calloc(s->nspec, sizeof(cx));
  for (int m = 0; m <= max_wave; m++)
    for (int n = (m > 1 ? m : 1); n <= max_wave; n++) {
      double re = irwin_hall(&st), im = m == 0 ? 0.0 : irwin_hall(&st);
      add[sidx(s, m, n)] = cmk(re, im);
    }
  double *g = (double *)malloc(sizeof(double) * s->ngrid);
  synthesize_scalar(s, add, g);
  double ms = 0, wsum = 0;
  for (int j = 0; j < s->nlat; j++)
    for (int k = 0; k < s->nlon; k++) { double x = g[(size_t)j * s->nlon + k]; ms += s->w[j] * x * x; wsum += s->w[j]; }
  double rms = sqrt(ms / wsum);
  double scale = rms > 0 ? amplitude / rms * s->gravity : 0;
  for (size_t i = 0; i < s->nspec; i++) s->phi[i] = cadd(s->phi[i], cscale(add[i], scale));
  s->t = 0;
  orc_sphere_mark_reference(s);
  free(add); free(g);
  return ORC_OK;
}

/* tendencies: sphere.cpp:338-390 */
static void tendencies(const orc_sphere *s, const cx *z, const cx *d, const cx *f, cx *dz, cx *dd, cx *df) {
  const size_t ng = s->ngrid;
  const int N = s->N;
  const double a = s->radius;
  double *U = malloc(sizeof(double) * ng), *V = malloc(sizeof(double) * ng), *zg = malloc(sizeof(double) * ng), *pg = malloc(sizeof(double) * ng);
  winds_spectral_to_grid(s, z, d, U, V);
  synthesize_scalar(s, z, zg);
  synthesize_scalar(s, f, pg);
  double *A = malloc(sizeof(double) * ng), *B = malloc(sizeof(double) * ng), *UP = malloc(sizeof(double) * ng), *VP = malloc(sizeof(double) * ng),
         *PE = malloc(sizeof(double) * ng);
  for (int j = 0; j < s->nlat; j++) {
    double om = 1 - s->mu[j] * s->mu[j];
    for (int k = 0; k < s->nlon; k++) {
      size_t i = (size_t)j * s->nlon + k;
      double eta = zg[i] + s->f[i];
      A[i] = U[i] * eta;
      B[i] = V[i] * eta;
      UP[i] = U[i] * pg[i];
      VP[i] = V[i] * pg[i];
      PE[i] = pg[i] + s->gravity * s->hs[i] + (U[i] * U[i] + V[i] * V[i]) / (2 * om);
    }
  }
  for (size_t i = 0; i < s->nspec; i++) { dz[i] = cmk(0, 0); dd[i] = cmk(0, 0); df[i] = cmk(0, 0); }
  cx *fa = malloc(sizeof(cx) * (N + 1)), *fb = malloc(sizeof(cx) * (N + 1)), *fu = malloc(sizeof(cx) * (N + 1)), *fv = malloc(sizeof(cx) * (N + 1)),
     *fe = malloc(sizeof(cx) * (N + 1)), *work = malloc(sizeof(cx) * s->nlon);
  for (int j = 0; j < s->nlat; j++) {
    size_t o = (size_t)j * s->nlon;
    fourier_forward(s, A + o, fa, work);
    fourier_forward(s, B + o, fb, work);
    fourier_forward(s, UP + o, fu, work);
    fourier_forward(s, VP + o, fv, work);
    fourier_forward(s, PE + o, fe, work);
    double wj = s->w[j], wd = s->w[j] / (1 - s->mu[j] * s->mu[j]) / a;
    for (int m = 0; m <= N; m++) {
      cx ca = cscale(cimul((double)m, fa[m]), wd), cb = cscale(fb[m], wd), cu = cscale(cimul((double)m, fu[m]), wd), cv = cscale(fv[m], wd);
      cx cb2 = cscale(cimul((double)m, fb[m]), wd), ca2 = cscale(fa[m], wd);
      cx ce = cscale(fe[m], wj);
      for (int n = m; n <= N; n++) {
        double P = Ptab(s, j, m, n - m), H = Htab(s, j, m, n - m);
        size_t q = sidx(s, m, n);
        dz[q] = csub(dz[q], csub(cscale(ca, P), cscale(cb, H)));
        dd[q] = cadd(dd[q], cadd(cscale(cb2, P), cscale(ca2, H)));
        dd[q] = cadd(dd[q], cscale(cscale(ce, P), (n * (n + 1.0) / (a * a))));
        df[q] = csub(df[q], csub(cscale(cu, P), cscale(cv, H)));
      }
    }
  }
  dz[0] = cmk(0, 0);
  dd[0] = cmk(0, 0);
  df[0] = cmk(0, 0);
  free(U); free(V); free(zg); free(pg); free(A); free(B); free(UP); free(VP); free(PE);
  free(fa); free(fb); free(fu); free(fv); free(fe); free(work);
}

/* filter: sphere.cpp:392-405 */
static void sphere_filter(orc_sphere *s, double dt) {
  if (s->diffusion <= 0) return;
  const double a2 = s->radius * s->radius;
  for (int m = 0; m <= s->N; m++)
    for (int n = (m > 1 ? m : 1); n <= s->N; n++) {
      double l = n * (n + 1.0);
      double kv = (l - 2) / a2, kp = l / a2;
      double fv = exp(-s->diffusion * kv * kv * dt), fp = exp(-s->diffusion * kp * kp * dt);
      size_t q = sidx(s, m, n);
      s->zeta[q] = cscale(s->zeta[q], fv);
      s->delta[q] = cscale(s->delta[q], fv);
      s->phi[q] = cscale(s->phi[q], fp);
      if (s->has_tracer) s->q[q] = cscale(s->q[q], fp);
    }
}

/* step: sphere.cpp:407-436 */
typedef struct { cx *z, *d, *f, *q; } sstate;
static void rhs(const orc_sphere *s, const sstate *x, sstate *k) {
  if (s->advect_only) {
    for (size_t i = 0; i < s->nspec; i++) { k->z[i] = cmk(0, 0); k->d[i] = cmk(0, 0); k->f[i] = cmk(0, 0); }
  } else {
    tendencies(s, x->z, x->d, x->f, k->z, k->d, k->f);
  }
  if (s->has_tracer) tracer_tendency(s, x->z, x->d, x->q, k->q);
}
static void axpy(const cx *x, const cx *k, double c, cx *r, size_t n) {
  for (size_t i = 0; i < n; i++) r[i] = cadd(x[i], cscale(k[i], c));
}
void orc_sphere_step(orc_sphere *s, double dt) {
  const size_t ns = s->nspec;
  cx *buf = (cx *)calloc(ns * 4 * 5, sizeof(cx));
  sstate s0 = {s->zeta, s->delta, s->phi, s->q}, kk[4], t;
  for (int i = 0; i < 4; i++) { kk[i].z = buf + ns * (4 * i); kk[i].d = buf + ns * (4 * i + 1); kk[i].f = buf + ns * (4 * i + 2); kk[i].q = buf + ns * (4 * i + 3); }
  t.z = buf + ns * 16; t.d = buf + ns * 17; t.f = buf + ns * 18; t.q = buf + ns * 19;
  rhs(s, &s0, &kk[0]);
  const double cs[3] = {0.5 * dt, 0.5 * dt, dt};
  for (int st = 1; st < 4; st++) {
    axpy(s0.z, kk[st - 1].z, cs[st - 1], t.z, ns);
    axpy(s0.d, kk[st - 1].d, cs[st - 1], t.d, ns);
    axpy(s0.f, kk[st - 1].f, cs[st - 1], t.f, ns);
    if (s->has_tracer) axpy(s0.q, kk[st - 1].q, cs[st - 1], t.q, ns);
    rhs(s, &t, &kk[st]);
  }
  for (size_t i = 0; i < ns; i++) {
#define RK4(FIELD, K)                                                                                                         \
  s->FIELD[i] = cadd(s->FIELD[i], cscale(cadd(cadd(cadd(kk[0].K[i], cscale(kk[1].K[i], 2.0)), cscale(kk[2].K[i], 2.0)), kk[3].K[i]), dt / 6))
    RK4(zeta, z);
    RK4(delta, d);
    RK4(phi, f);
    if (s->has_tracer) RK4(q, q);
#undef RK4
  }
  free(buf);
  sphere_filter(s, dt);
  s->t += dt;
}

/* grid diagnostics: sphere.cpp:486-508 */
void orc_sphere_grid_height(const orc_sphere *s, double *g) {
  synthesize_scalar(s, s->phi, g);
  for (size_t i = 0; i < s->ngrid; i++) g[i] /= s->gravity;
}
void orc_sphere_grid_vorticity(const orc_sphere *s, double *g) { synthesize_scalar(s, s->zeta, g); }
static void grid_uv(const orc_sphere *s, double *u, double *v) {
  winds_spectral_to_grid(s, s->zeta, s->delta, u, v);
  for (int j = 0; j < s->nlat; j++) {
    double c = sqrt(1 - s->mu[j] * s->mu[j]);
    for (int k = 0; k < s->nlon; k++) { u[(size_t)j * s->nlon + k] /= c; v[(size_t)j * s->nlon + k] /= c; }
  }
}
void orc_sphere_grid_u(const orc_sphere *s, double *u) {
  double *v = malloc(sizeof(double) * s->ngrid);
  grid_uv(s, u, v);
  free(v);
}
void orc_sphere_grid_v(const orc_sphere *s, double *v) {
  double *u = malloc(sizeof(double) * s->ngrid);
  grid_uv(s, u, v);
  free(u);
}
int orc_sphere_grid_tracer(const orc_sphere *s, double *g) {
  if (!s->has_tracer) return fail(ORC_ERUNTIME, "no tracer: set one with tracer() first");
  synthesize_scalar(s, s->q, g);
  return ORC_OK;
}

/* stable_dt: sphere.cpp:438-447 */
int orc_sphere_stable_dt(const orc_sphere *s, double safety, double *out) {
  double *h = malloc(sizeof(double) * s->ngrid), *u = malloc(sizeof(double) * s->ngrid), *v = malloc(sizeof(double) * s->ngrid);
  orc_sphere_grid_height(s, h);
  grid_uv(s, u, v);
  double hmax = 0, wmax = 0;
  for (size_t i = 0; i < s->ngrid; i++) { hmax = fmax(hmax, h[i]); wmax = fmax(wmax, hypot(u[i], v[i])); }
  free(h); free(u); free(v);
  if (!(hmax > 0)) return fail(ORC_ERUNTIME, "the height field must be positive to step the model");
  double omega = s->advect_only ? (s->N / s->radius) * wmax : (s->N / s->radius) * (sqrt(s->gravity * hmax) + wmax) + 2 * s->omega;
  *out = !(omega > 0) ? 1e9 : safety / omega;
  return ORC_OK;
}
/* check_finite: sphere.cpp:449-455 */
static int check_finite(const orc_sphere *s) {
  double sum = 0;
  for (size_t i = 0; i < s->nspec; i++) sum += hypot(s->zeta[i].re, s->zeta[i].im) + hypot(s->delta[i].re, s->delta[i].im) + hypot(s->phi[i].re, s->phi[i].im);
  if (s->has_tracer)
    for (size_t i = 0; i < s->nspec; i++) sum += hypot(s->q[i].re, s->q[i].im);
  if (!isfinite(sum)) return fail(ORC_ERUNTIME, "the integration blew up (not finite): use a smaller dt");
  return ORC_OK;
}
/* run: sphere.cpp:457-484 (messages identical, std::to_string(double) is "%f") */
int orc_sphere_run(orc_sphere *s, double seconds, double dt) {
  if (!(seconds >= 0)) return fail(ORC_EINVAL, "run: seconds must not be negative");
  if (s->advect_only && !s->has_tracer) return fail(ORC_EINVAL, "advect_only: set a tracer first (Sphere.tracer)");
  if (seconds == 0) return ORC_OK;
  int rc;
  if (!(dt > 0)) { if ((rc = orc_sphere_stable_dt(s, 1.5, &dt))) return rc; }
  double limit;
  if ((rc = orc_sphere_stable_dt(s, 2.0 * sqrt(2.0), &limit))) return rc;
  if (!(dt <= limit)) {
    snprintf(g_err, sizeof g_err, "run: dt = %f s exceeds the RK4 stability limit of about %f s for the present state: use a smaller dt", dt, limit);
    return ORC_EINVAL;
  }
  int64_t n = (int64_t)ceil(seconds / dt - 1e-9);
  if (n < 1) n = 1;
  if (n > 50000000) return fail(ORC_EINVAL, "run: more than 5e7 steps; use a larger dt");
  double h = seconds / (double)n;
  for (int64_t i = 0; i < n; i++) {
    orc_sphere_step(s, h);
    if ((i & 63) == 63) {
      if ((rc = check_finite(s))) return rc;
      double now;
      if ((rc = orc_sphere_stable_dt(s, 2.0 * sqrt(2.0), &now))) return rc;
      if (!(h <= now)) {
        snprintf(g_err, sizeof g_err,
                 "run: dt = %f s exceeds the RK4 stability limit of about %f s reached at t = %f s (the flow accelerated): use a smaller dt", h, now, s->t);
        return ORC_ERUNTIME;
      }
    }
  }
  return check_finite(s);
}

/* scalar diagnostics: sphere.cpp:510-569 */
double orc_sphere_mean_height(const orc_sphere *s) { return s->phi[0].re / sqrt(2.0) / s->gravity; }
double orc_sphere_kinetic_energy(const orc_sphere *s) {
  double *h = malloc(sizeof(double) * s->ngrid), *u = malloc(sizeof(double) * s->ngrid), *v = malloc(sizeof(double) * s->ngrid);
  orc_sphere_grid_height(s, h);
  grid_uv(s, u, v);
  double sum = 0;
  for (int j = 0; j < s->nlat; j++) {
    double row = 0;
    for (int k = 0; k < s->nlon; k++) { size_t i = (size_t)j * s->nlon + k; row += 0.5 * h[i] * (u[i] * u[i] + v[i] * v[i]); }
    sum += s->w[j] * row / s->nlon;
  }
  free(h); free(u); free(v);
  return 0.5 * sum;
}
double orc_sphere_potential_energy(const orc_sphere *s) {
  double *h = malloc(sizeof(double) * s->ngrid);
  orc_sphere_grid_height(s, h);
  double sum = 0;
  for (int j = 0; j < s->nlat; j++) {
    double row = 0;
    for (int k = 0; k < s->nlon; k++) { size_t i = (size_t)j * s->nlon + k; row += 0.5 * s->gravity * h[i] * h[i] + s->gravity * h[i] * s->hs[i]; }
    sum += s->w[j] * row / s->nlon;
  }
  free(h);
  return 0.5 * sum;
}
double orc_sphere_energy(const orc_sphere *s) { return orc_sphere_kinetic_energy(s) + orc_sphere_potential_energy(s); }
/* spectrum: sphere.cpp:536-551 (ke, pe have N+1 entries) */
void orc_sphere_spectrum(const orc_sphere *s, double *ke, double *pe) {
  const double a2 = s->radius * s->radius, g2 = s->gravity * s->gravity;
  for (int n = 0; n <= s->N; n++) { ke[n] = 0; pe[n] = 0; }
  for (int m = 0; m <= s->N; m++) {
    double cm = m == 0 ? 1.0 : 2.0;
    for (int n = (m > 1 ? m : 1); n <= s->N; n++) {
      size_t q = sidx(s, m, n);
      double nz = s->zeta[q].re * s->zeta[q].re + s->zeta[q].im * s->zeta[q].im, nd = s->delta[q].re * s->delta[q].re + s->delta[q].im * s->delta[q].im;
      double np = s->phi[q].re * s->phi[q].re + s->phi[q].im * s->phi[q].im;
      ke[n] += 0.25 * cm * a2 * (nz + nd) / (n * (n + 1.0));
      pe[n] += 0.25 * cm * s->gravity * np / g2;
    }
  }
}
double orc_sphere_potential_enstrophy(const orc_sphere *s) {
  double *h = malloc(sizeof(double) * s->ngrid), *z = malloc(sizeof(double) * s->ngrid);
  orc_sphere_grid_height(s, h);
  synthesize_scalar(s, s->zeta, z);
  double sum = 0;
  for (int j = 0; j < s->nlat; j++) {
    double row = 0;
    for (int k = 0; k < s->nlon; k++) { size_t i = (size_t)j * s->nlon + k; double f = s->f[i]; row += 0.5 * (z[i] + f) * (z[i] + f) / h[i]; }
    sum += s->w[j] * row / s->nlon;
  }
  free(h); free(z);
  return 0.5 * sum;
}
double orc_sphere_max_wind(const orc_sphere *s) {
  double *u = malloc(sizeof(double) * s->ngrid), *v = malloc(sizeof(double) * s->ngrid);
  grid_uv(s, u, v);
  double m = 0;
  for (size_t i = 0; i < s->ngrid; i++) m = fmax(m, hypot(u[i], v[i]));
  free(u); free(v);
  return m;
}
/* errors: sphere.cpp:571-590 */
static void norms(const orc_sphere *s, const double *ref, const double *h, double g, double *l2, double *linf) {
  double num = 0, den = 0, mn = 0, mr = 0;
  for (int j = 0; j < s->nlat; j++)
    for (int k = 0; k < s->nlon; k++) {
      size_t i = (size_t)j * s->nlon + k;
      double r = ref[i] / g, d = h[i] - r;
      num += s->w[j] * d * d;
      den += s->w[j] * r * r;
      mn = fmax(mn, fabs(d));
      mr = fmax(mr, fabs(r));
    }
  *l2 = sqrt(num / den);
  *linf = mn / mr;
}
void orc_sphere_height_errors(const orc_sphere *s, double *l2, double *linf) {
  double *h = malloc(sizeof(double) * s->ngrid);
  orc_sphere_grid_height(s, h);
  norms(s, s->ref_phi, h, s->gravity, l2, linf);
  free(h);
}
int orc_sphere_tracer_errors(const orc_sphere *s, double *l2, double *linf) {
  if (!s->has_tracer) return fail(ORC_ERUNTIME, "no tracer: set one with tracer() first");
  double *q = malloc(sizeof(double) * s->ngrid);
  synthesize_scalar(s, s->q, q);
  norms(s, s->ref_q, q, 1.0, l2, linf);
  free(q);
  return ORC_OK;
}
/* tracer scalars: sphere.cpp:271-286 */
int orc_sphere_tracer_mean(const orc_sphere *s, double *out) {
  if (!s->has_tracer) return fail(ORC_ERUNTIME, "no tracer: set one with tracer() first");
  *out = s->q[0].re / sqrt(2.0);
  return ORC_OK;
}
int orc_sphere_tracer_mass(const orc_sphere *s, double *out) {
  if (!s->has_tracer) return fail(ORC_ERUNTIME, "no tracer: set one with tracer() first");
  double *h = malloc(sizeof(double) * s->ngrid), *q = malloc(sizeof(double) * s->ngrid);
  orc_sphere_grid_height(s, h);
  synthesize_scalar(s, s->q, q);
  double sum = 0;
  for (int j = 0; j < s->nlat; j++) {
    double row = 0;
    for (int k = 0; k < s->nlon; k++) { size_t i = (size_t)j * s->nlon + k; row += h[i] * q[i]; }
    sum += s->w[j] * row / s->nlon;
  }
  free(h); free(q);
  *out = 0.5 * sum;
  return ORC_OK;
}

/* point evaluators: sphere.cpp:287-299, 592-637 */
static double point_scalar(const orc_sphere *s, const cx *spec, double lat, double lon) {
  double mu = sin(lat);
  double *col = malloc(sizeof(double) * (s->N + 2));
  double sum = 0;
  for (int m = 0; m <= s->N; m++) {
    orc_legendre_column(m, s->N, mu, col);
    cx g = cmk(0, 0);
    for (int n = m; n <= s->N; n++) g = cadd(g, cscale(spec[sidx(s, m, n)], col[n - m]));
    cx e = cmk(cos(m * lon), sin(m * lon));
    sum += (m == 0 ? 1.0 : 2.0) * cmul(g, e).re;
  }
  free(col);
  return sum;
}
double orc_sphere_height_at(const orc_sphere *s, double lat, double lon) { return point_scalar(s, s->phi, lat, lon) / s->gravity; }
double orc_sphere_vorticity_at(const orc_sphere *s, double lat, double lon) { return point_scalar(s, s->zeta, lat, lon); }
int orc_sphere_tracer_at(const orc_sphere *s, double lat, double lon, double *out) {
  if (!s->has_tracer) return fail(ORC_ERUNTIME, "no tracer: set one with tracer() first");
  *out = point_scalar(s, s->q, lat, lon);
  return ORC_OK;
}
double orc_sphere_wind_at(const orc_sphere *s, double lat, double lon, int zonal) {
  double mu = sin(lat), a = s->radius, c = cos(lat);
  double *col = malloc(sizeof(double) * (s->N + 2));
  double sum = 0;
  for (int m = 0; m <= s->N; m++) {
    orc_legendre_column(m, s->N, mu, col);
    cx sp = cmk(0, 0), sc = cmk(0, 0), hp = cmk(0, 0), hc = cmk(0, 0);
    for (int n = (m > 1 ? m : 1); n <= s->N; n++) {
      double k = -a * a / (n * (n + 1.0));
      double P = col[n - m];
      double H = -n * eps_nm(n + 1, m) * col[n + 1 - m] + (n > m ? (n + 1) * eps_nm(n, m) * col[n - 1 - m] : 0.0);
      cx ps = cscale(s->zeta[sidx(s, m, n)], k), ch = cscale(s->delta[sidx(s, m, n)], k);
      sp = cadd(sp, cscale(ps, P)); sc = cadd(sc, cscale(ch, P)); hp = cadd(hp, cscale(ps, H)); hc = cadd(hc, cscale(ch, H));
    }
    cx im = cmk(0, (double)m);
    cx g = zonal ? cdiv(csub(cmul(im, sc), hp), a) : cdiv(cadd(cmul(im, sp), hc), a);
    cx e = cmk(cos(m * lon), sin(m * lon));
    sum += (m == 0 ? 1.0 : 2.0) * cmul(g, e).re;
  }
  free(col);
  return sum / c;
}
int orc_sphere_mode_coefficient(const orc_sphere *s, int n, int m, int imag, double *out) {
  if (m < 0 || n < m || n > s->N) return fail(ORC_EINVAL, "mode (n, m) outside the truncation");
  cx c = cdiv(s->phi[sidx(s, m, n)], s->gravity);
  *out = imag ? c.im : c.re;
  return ORC_OK;
}
int orc_sphere_wave_crest(const orc_sphere *s, int m, double lat, double *out) {
  if (m < 1 || m > s->N) return fail(ORC_EINVAL, "zonal wavenumber must be between 1 and the truncation");
  double *col = malloc(sizeof(double) * (s->N + 2));
  orc_legendre_column(m, s->N, sin(lat), col);
  cx g = cmk(0, 0);
  for (int n = m; n <= s->N; n++) g = cadd(g, cscale(s->phi[sidx(s, m, n)], col[n - m]));
  free(col);
  *out = -atan2(g.im, g.re) / m;
  return ORC_OK;
}

/* exposed transforms (sphere.hpp:97-100): spectral arrays are (N+1)^2 complex = 2 (N+1)^2 doubles, index m*(N+1)+n */
void orc_sphere_synthesize(const orc_sphere *s, const double *spec, double *grid) { synthesize_scalar(s, (const cx *)spec, grid); }
void orc_sphere_analyze(const orc_sphere *s, const double *grid, double *spec) { analyze_scalar(s, grid, (cx *)spec); }
void orc_sphere_analyze_divergence(const orc_sphere *s, const double *A, const double *B, double *spec) { analyze_divergence(s, A, B, (cx *)spec); }
void orc_sphere_winds(const orc_sphere *s, const double *z, const double *d, double *U, double *V) {
  winds_spectral_to_grid(s, (const cx *)z, (const cx *)d, U, V);
}
/* one evaluation of tendencies() on the current state */
void orc_sphere_tendencies(const orc_sphere *s, double *dz, double *dd, double *df) {
  tendencies(s, s->zeta, s->delta, s->phi, (cx *)dz, (cx *)dd, (cx *)df);
}
/* raw spectral state access: field 0 zeta, 1 delta, 2 phi, 3 q */
static cx *field_ptr(orc_sphere *s, int field) { return field == 0 ? s->zeta : field == 1 ? s->delta : field == 2 ? s->phi : s->q; }
void orc_sphere_get_spectral(orc_sphere *s, int field, double *out) { memcpy(out, field_ptr(s, field), sizeof(cx) * s->nspec); }
void orc_sphere_set_spectral(orc_sphere *s, int field, const double *in) {
  memcpy(field_ptr(s, field), in, sizeof(cx) * s->nspec);
  if (field == 3) s->has_tracer = 1;
}
void orc_sphere_get_tables(const orc_sphere *s, double *P, double *H) {
  size_t tab = (size_t)s->nlat * (s->N + 1) * (s->N + 1);
  memcpy(P, s->P, sizeof(double) * tab);
  memcpy(H, s->H, sizeof(double) * tab);
}
void orc_sphere_get_coriolis(const orc_sphere *s, double *f) { memcpy(f, s->f, sizeof(double) * s->ngrid); }

/* ------------------------------------------------------------------ plane (plane.hpp:9-68, plane.cpp:20-318) */

typedef struct orc_plane {
  int nx, ny;
  double lx, ly, gravity, depth, f0, beta, viscosity, drag;
  int linear;
  double t;
  double *h, *u, *v; /* h, u: ny*nx; v: (ny+1)*nx */
} orc_plane;

/* reset: plane.cpp:27-32 */
void orc_plane_reset(orc_plane *p) {
  for (size_t i = 0; i < (size_t)p->nx * p->ny; i++) { p->h[i] = p->depth; p->u[i] = 0; }
  for (size_t i = 0; i < (size_t)p->nx * (p->ny + 1); i++) p->v[i] = 0;
  p->t = 0;
}
/* ctor: plane.cpp:20-25 */
orc_plane *orc_plane_create(int nx, int ny, double lx, double ly, double gravity, double depth, double f0, double beta, double viscosity,
                            double drag, int linear) {
  if (nx < 4 || ny < 4 || nx > 4000 || ny > 4000 || (int64_t)nx * ny > 8000000) {
    fail(ORC_EINVAL, "grid must be at least 4 x 4 and at most 8 million cells");
    return NULL;
  }
  if (!(lx > 0 && ly > 0)) { fail(ORC_EINVAL, "domain sizes must be positive"); return NULL; }
  orc_plane *p = (orc_plane *)calloc(1, sizeof *p);
  p->nx = nx; p->ny = ny; p->lx = lx; p->ly = ly; p->gravity = gravity; p->depth = depth; p->f0 = f0; p->beta = beta;
  p->viscosity = viscosity; p->drag = drag; p->linear = linear;
  p->h = (double *)malloc(sizeof(double) * nx * ny);
  p->u = (double *)malloc(sizeof(double) * nx * ny);
  p->v = (double *)malloc(sizeof(double) * nx * (ny + 1));
  orc_plane_reset(p);
  return p;
}
void orc_plane_destroy(orc_plane *p) {
  if (!p) return;
  free(p->h); free(p->u); free(p->v); free(p);
}
void orc_plane_set_params(orc_plane *p, double gravity, double depth, double f0, double beta, double viscosity, double drag, int linear) {
  p->gravity = gravity; p->depth = depth; p->f0 = f0; p->beta = beta; p->viscosity = viscosity; p->drag = drag; p->linear = linear;
}
double orc_plane_time(const orc_plane *p) { return p->t; }
static inline double pdx(const orc_plane *p) { return p->lx / p->nx; }
static inline double pdy(const orc_plane *p) { return p->ly / p->ny; }
static inline double coriolis(const orc_plane *p, double y) { return p->f0 + p->beta * (y - 0.5 * p->ly); } /* plane.hpp:67 */

/* perturb: plane.cpp:34-57 */
void orc_plane_perturb(orc_plane *p, double amplitude, uint64_t seed) {
  uint64_t st = seed * 0x9E3779B97F4A7C15ULL + 777;
  const int nx = p->nx, ny = p->ny;
  size_t nh = (size_t)nx * ny;
  double *n = malloc(sizeof(double) * nh), *t = malloc(sizeof(double) * nh);
  for (size_t i = 0; i < nh; i++) n[i] = irwin_hall(&st);
  for (int pass = 0; pass < 3; pass++) {
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++)
        t[(size_t)j * nx + i] = 0.25 * n[(size_t)j * nx + (i + nx - 1) % nx] + 0.5 * n[(size_t)j * nx + i] + 0.25 * n[(size_t)j * nx + (i + 1) % nx];
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++)
        n[(size_t)j * nx + i] = 0.25 * t[(size_t)(j - 1 > 0 ? j - 1 : 0) * nx + i] + 0.5 * t[(size_t)j * nx + i] + 0.25 * t[(size_t)(j + 1 < ny - 1 ? j + 1 : ny - 1) * nx + i];
  }
  double mean = 0, ms = 0;
  for (size_t i = 0; i < nh; i++) mean += n[i];
  mean /= nh;
  for (size_t i = 0; i < nh; i++) { n[i] -= mean; ms += n[i] * n[i]; }
  double rms = sqrt(ms / nh);
  double scale = rms > 0 ? amplitude / rms : 0;
  for (size_t i = 0; i < nh; i++) p->h[i] += n[i] * scale;
  free(n); free(t);
}

/* tendency: plane.cpp:59-129 */
static void plane_tendency(const orc_plane *p, const double *h, const double *u, const double *v, double *dh, double *du, double *dv) {
  const int nx = p->nx, ny = p->ny;
  const double dxx = pdx(p), dyy = pdy(p), g = p->gravity;
#define HH(i, j) h[(size_t)(j) * nx + (i)]
#define UU(i, j) u[(size_t)(j) * nx + ((((i) % nx) + nx) % nx)]
#define VV(i, j) v[(size_t)(j) * nx + ((((i) % nx) + nx) % nx)]
  for (size_t i = 0; i < (size_t)nx * ny; i++) { dh[i] = 0; du[i] = 0; }
  for (size_t i = 0; i < (size_t)nx * (ny + 1); i++) dv[i] = 0;
  const int lin = p->linear;
  double *B = malloc(sizeof(double) * nx * ny), *eta = malloc(sizeof(double) * nx * (ny + 1));
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      double k = 0;
      if (!lin) k = 0.5 * (0.5 * (UU(i, j) * UU(i, j) + UU(i + 1, j) * UU(i + 1, j)) + 0.5 * (VV(i, j) * VV(i, j) + VV(i, j + 1) * VV(i, j + 1)));
      B[(size_t)j * nx + i] = g * HH(i, j) + k;
    }
#define BF(i, j) B[(size_t)(j) * nx + ((((i) % nx) + nx) % nx)]
  for (int j = 0; j <= ny; j++) {
    double f = coriolis(p, j * dyy);
    for (int i = 0; i < nx; i++) {
      double zeta = 0;
      if (!lin) {
        double dvdx = (VV(i, j) - VV(i - 1, j)) / dxx;
        double uj = j < ny ? UU(i, j) : UU(i, ny - 1), ujm = j > 0 ? UU(i, j - 1) : UU(i, 0);
        zeta = dvdx - (uj - ujm) / dyy;
      }
      eta[(size_t)j * nx + i] = f + zeta;
    }
  }
#define EE(i, j) eta[(size_t)(j) * nx + ((((i) % nx) + nx) % nx)]
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      double q = 0.5 * (EE(i, j) + EE(i, j + 1));
      double vbar = 0.25 * (VV(i - 1, j) + VV(i, j) + VV(i - 1, j + 1) + VV(i, j + 1));
      double r = q * vbar - (BF(i, j) - BF(i - 1, j)) / dxx - p->drag * UU(i, j);
      if (p->viscosity != 0) {
        double uy1 = j + 1 < ny ? UU(i, j + 1) : UU(i, j), uy0 = j > 0 ? UU(i, j - 1) : UU(i, j);
        r += p->viscosity * ((UU(i + 1, j) - 2 * UU(i, j) + UU(i - 1, j)) / (dxx * dxx) + (uy1 - 2 * UU(i, j) + uy0) / (dyy * dyy));
      }
      du[(size_t)j * nx + i] = r;
    }
  for (int j = 1; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      double q = 0.5 * (EE(i, j) + EE(i + 1, j));
      double ubar = 0.25 * (UU(i, j - 1) + UU(i + 1, j - 1) + UU(i, j) + UU(i + 1, j));
      double r = -q * ubar - (BF(i, j) - BF(i, j - 1)) / dyy - p->drag * VV(i, j);
      if (p->viscosity != 0) {
        double vy1 = j + 1 < ny ? VV(i, j + 1) : 0.0, vy0 = j - 1 > 0 ? VV(i, j - 1) : 0.0;
        r += p->viscosity * ((VV(i + 1, j) - 2 * VV(i, j) + VV(i - 1, j)) / (dxx * dxx) + (vy1 - 2 * VV(i, j) + vy0) / (dyy * dyy));
      }
      dv[(size_t)j * nx + i] = r;
    }
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      double hw = lin ? p->depth : 0.5 * (HH(i - 1 < 0 ? nx - 1 : i - 1, j) + HH(i % nx, j));
      double fe = (lin ? p->depth : 0.5 * (HH(i, j) + HH((i + 1) % nx, j))) * UU(i + 1, j);
      double fw = hw * UU(i, j);
      double fn = 0, fs = 0;
      if (j + 1 < ny) fn = (lin ? p->depth : 0.5 * (HH(i, j) + HH(i, j + 1))) * VV(i, j + 1);
      if (j > 0) fs = (lin ? p->depth : 0.5 * (HH(i, j - 1) + HH(i, j))) * VV(i, j);
      dh[(size_t)j * nx + i] = -((fe - fw) / dxx + (fn - fs) / dyy);
    }
#undef HH
#undef UU
#undef VV
#undef BF
#undef EE
  free(B); free(eta);
}
void orc_plane_tendency(const orc_plane *p, double *dh, double *du, double *dv) { plane_tendency(p, p->h, p->u, p->v, dh, du, dv); }

/* max_speed: plane.cpp:209-218 */
double orc_plane_max_speed(const orc_plane *p) {
  const int nx = p->nx, ny = p->ny;
  double m = 0;
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      double uc = 0.5 * (p->u[(size_t)j * nx + i] + p->u[(size_t)j * nx + (i + 1) % nx]), vc = 0.5 * (p->v[(size_t)j * nx + i] + p->v[(size_t)(j + 1) * nx + i]);
      m = fmax(m, hypot(uc, vc));
    }
  return m;
}
/* stable_dt: plane.cpp:131-141 */
int orc_plane_stable_dt(const orc_plane *p, double safety, double *out) {
  double hmax = p->linear ? p->depth : 0, umax = p->linear ? 0 : orc_plane_max_speed(p);
  if (!p->linear)
    for (size_t i = 0; i < (size_t)p->nx * p->ny; i++) hmax = fmax(hmax, p->h[i]);
  if (!(hmax > 0)) return fail(ORC_ERUNTIME, "the depth must be positive to step the model");
  double c = sqrt(p->gravity * hmax) + umax;
  double inv = c * sqrt(1.0 / (pdx(p) * pdx(p)) + 1.0 / (pdy(p) * pdy(p)));
  double visc = p->viscosity * (4.0 / (pdx(p) * pdx(p)) + 4.0 / (pdy(p) * pdy(p)));
  *out = safety * 2.0 / (inv + visc + fabs(p->f0) + fabs(p->beta) * p->ly);
  return ORC_OK;
}
/* step: plane.cpp:143-158 */
void orc_plane_step(orc_plane *p, double dt) {
  const size_t nh = (size_t)p->nx * p->ny, nv = (size_t)p->nx * (p->ny + 1);
  double *kb = calloc((nh * 2 + nv) * 4, sizeof(double));
  double *kh[4], *ku[4], *kv[4];
  for (int s = 0; s < 4; s++) { kh[s] = kb + (nh * 2 + nv) * s; ku[s] = kh[s] + nh; kv[s] = ku[s] + nh; }
  double *ht = malloc(sizeof(double) * nh), *ut = malloc(sizeof(double) * nh), *vt = malloc(sizeof(double) * nv);
  plane_tendency(p, p->h, p->u, p->v, kh[0], ku[0], kv[0]);
  const double cs[3] = {0.5 * dt, 0.5 * dt, dt};
  for (int s = 1; s < 4; s++) {
    for (size_t i = 0; i < nh; i++) ht[i] = p->h[i] + cs[s - 1] * kh[s - 1][i];
    for (size_t i = 0; i < nh; i++) ut[i] = p->u[i] + cs[s - 1] * ku[s - 1][i];
    for (size_t i = 0; i < nv; i++) vt[i] = p->v[i] + cs[s - 1] * kv[s - 1][i];
    plane_tendency(p, ht, ut, vt, kh[s], ku[s], kv[s]);
  }
  for (size_t i = 0; i < nh; i++) p->h[i] += dt / 6 * (kh[0][i] + 2 * kh[1][i] + 2 * kh[2][i] + kh[3][i]);
  for (size_t i = 0; i < nh; i++) p->u[i] += dt / 6 * (ku[0][i] + 2 * ku[1][i] + 2 * ku[2][i] + ku[3][i]);
  for (size_t i = 0; i < nv; i++) p->v[i] += dt / 6 * (kv[0][i] + 2 * kv[1][i] + 2 * kv[2][i] + kv[3][i]);
  p->t += dt;
  free(kb); free(ht); free(ut); free(vt);
}
/* run: plane.cpp:160-189 */
int orc_plane_run(orc_plane *p, double seconds, double dt) {
  if (!(seconds >= 0)) return fail(ORC_EINVAL, "run: seconds must not be negative");
  if (seconds == 0) return ORC_OK;
  int rc;
  if (!(dt > 0)) { if ((rc = orc_plane_stable_dt(p, 0.5, &dt))) return rc; }
  double limit;
  if ((rc = orc_plane_stable_dt(p, sqrt(0.5), &limit))) return rc;
  if (!(dt <= limit)) {
    snprintf(g_err, sizeof g_err, "run: dt = %f s exceeds the RK4 stability limit of about %f s for the present state: use a smaller dt", dt, limit);
    return ORC_EINVAL;
  }
  int64_t n = (int64_t)ceil(seconds / dt - 1e-9);
  if (n < 1) n = 1;
  if (n > 200000000) return fail(ORC_EINVAL, "run: too many steps; use a larger dt");
  double hstep = seconds / (double)n;
  const size_t nh = (size_t)p->nx * p->ny, nv = (size_t)p->nx * (p->ny + 1);
  for (int64_t i = 0; i < n; i++) {
    orc_plane_step(p, hstep);
    if ((i & 63) == 63 || i + 1 == n) {
      double s = 0;
      for (size_t k = 0; k < nh; k++) s += p->h[k];
      for (size_t k = 0; k < nh; k++) s += p->u[k];
      for (size_t k = 0; k < nv; k++) s += p->v[k];
      if (!isfinite(s)) return fail(ORC_ERUNTIME, "the integration blew up (not finite): use a smaller dt");
      double now;
      if ((rc = orc_plane_stable_dt(p, sqrt(0.5), &now))) return rc;
      if (!(hstep <= now)) {
        snprintf(g_err, sizeof g_err,
                 "run: dt = %f s exceeds the RK4 stability limit of about %f s reached at t = %f s (the flow accelerated): use a smaller dt", hstep, now, p->t);
        return ORC_ERUNTIME;
      }
    }
  }
  return ORC_OK;
}
/* diagnostics: plane.cpp:191-243 */
double orc_plane_mean_depth(const orc_plane *p) {
  double s = 0;
  size_t nh = (size_t)p->nx * p->ny;
  for (size_t i = 0; i < nh; i++) s += p->h[i];
  return s / nh;
}
double orc_plane_energy(const orc_plane *p) {
  const int nx = p->nx, ny = p->ny;
  double s = 0;
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      double uc = 0.5 * (p->u[(size_t)j * nx + i] + p->u[(size_t)j * nx + (i + 1) % nx]), vc = 0.5 * (p->v[(size_t)j * nx + i] + p->v[(size_t)(j + 1) * nx + i]);
      double hh = p->h[(size_t)j * nx + i];
      s += 0.5 * hh * (uc * uc + vc * vc) + 0.5 * p->gravity * hh * hh;
    }
  return s / ((size_t)nx * ny);
}
double orc_plane_zeta_at_corner(const orc_plane *p, int i, int j) {
  const int nx = p->nx, ny = p->ny;
  int im = (i + nx - 1) % nx;
  double dvdx = (p->v[(size_t)j * nx + i % nx] - p->v[(size_t)j * nx + im]) / pdx(p);
  double uj = j < ny ? p->u[(size_t)j * nx + i % nx] : p->u[(size_t)(ny - 1) * nx + i % nx];
  double ujm = j > 0 ? p->u[(size_t)(j - 1) * nx + i % nx] : p->u[(size_t)0 * nx + i % nx];
  return dvdx - (uj - ujm) / pdy(p);
}
double orc_plane_pv_anomaly_rms(const orc_plane *p) {
  const int nx = p->nx, ny = p->ny;
  double s = 0;
  size_t cnt = 0;
  for (int j = 1; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      int im = (i + nx - 1) % nx;
      double hav = 0.25 * (p->h[(size_t)(j - 1) * nx + im] + p->h[(size_t)(j - 1) * nx + i] + p->h[(size_t)j * nx + im] + p->h[(size_t)j * nx + i]);
      double q = orc_plane_zeta_at_corner(p, i, j) - p->f0 * (hav - p->depth) / p->depth;
      s += q * q;
      cnt++;
    }
  return sqrt(s / (double)cnt);
}
/* interp: plane.cpp:245-260 (kind 0 h, 1 u, 2 v) */
double orc_plane_at(const orc_plane *p, double x, double y, int kind) {
  double dxx = p->lx / p->nx, dyy = p->ly / p->ny;
  double xs = kind == 1 ? 0.0 : 0.5, ys = kind == 2 ? 0.0 : 0.5;
  double fx = x / dxx - xs, fy = y / dyy - ys;
  int rows = kind == 2 ? p->ny + 1 : p->ny;
  double fl = floor(fx);
  int i0 = (int)fl, j0 = (int)floor(fy);
  double ax = fx - fl, ay = fy - floor(fy);
  double g[2][2];
  for (int di = 0; di < 2; di++)
    for (int dj = 0; dj < 2; dj++) {
      int i = (((i0 + di) % p->nx) + p->nx) % p->nx;
      int j = j0 + dj;
      j = j < 0 ? 0 : (j > rows - 1 ? rows - 1 : j);
      const double *f = kind == 0 ? p->h : kind == 1 ? p->u : p->v;
      g[di][dj] = f[(size_t)j * p->nx + i];
    }
  return (1 - ax) * ((1 - ay) * g[0][0] + ay * g[0][1]) + ax * ((1 - ay) * g[1][0] + ay * g[1][1]);
}
/* harmonic: plane.cpp:265-269 */
static cx harmonic(int nx, int m, const double *row, double shift) {
  cx c = cmk(0, 0);
  for (int i = 0; i < nx; i++) {
    double th = -2 * KPI * m * (i + shift) / nx;
    c = cadd(c, cscale(cmk(cos(th), sin(th)), row[i]));
  }
  return cdiv(c, (double)nx);
}
int orc_plane_zonal_amplitude(const orc_plane *p, int m, int j, double *out) {
  if (m < 1 || m >= p->nx / 2 || j < 0 || j >= p->ny) return fail(ORC_EINVAL, "zonal harmonic or row out of range");
  cx c = harmonic(p->nx, m, p->h + (size_t)j * p->nx, 0.5);
  *out = 2 * hypot(c.re, c.im);
  return ORC_OK;
}
static double crest(const orc_plane *p, cx c, int m) {
  double ph = atan2(c.im, c.re);
  double x = -ph / (2 * KPI * m) * p->lx;
  double per = p->lx / m;
  x = fmod(x, per);
  return x < 0 ? x + per : x;
}
int orc_plane_zonal_crest(const orc_plane *p, int m, int j, double *out) {
  if (m < 1 || m >= p->nx / 2 || j < 0 || j >= p->ny) return fail(ORC_EINVAL, "zonal harmonic or row out of range");
  *out = crest(p, harmonic(p->nx, m, p->h + (size_t)j * p->nx, 0.5), m);
  return ORC_OK;
}
int orc_plane_harmonic_energy(const orc_plane *p, int m, double *out) {
  if (m < 1 || m >= p->nx / 2) return fail(ORC_EINVAL, "zonal harmonic out of range");
  double s = 0;
  for (int j = 0; j < p->ny; j++) { cx c = harmonic(p->nx, m, p->u + (size_t)j * p->nx, 0.0); s += c.re * c.re + c.im * c.im; }
  for (int j = 1; j < p->ny; j++) { cx c = harmonic(p->nx, m, p->v + (size_t)j * p->nx, 0.5); s += c.re * c.re + c.im * c.im; }
  *out = s / p->ny;
  return ORC_OK;
}
int orc_plane_v_zonal_crest(const orc_plane *p, int m, int j, double *out) {
  if (m < 1 || m >= p->nx / 2 || j < 1 || j >= p->ny) return fail(ORC_EINVAL, "zonal harmonic or v row out of range");
  *out = crest(p, harmonic(p->nx, m, p->v + (size_t)j * p->nx, 0.5), m);
  return ORC_OK;
}
/* setters and getters: plane.cpp:312-318 (field 0 h, 1 u, 2 v) */
int orc_plane_set_field(orc_plane *p, int field, const double *g, size_t n) {
  size_t want = field == 2 ? (size_t)p->nx * (p->ny + 1) : (size_t)p->nx * p->ny;
  if (n != want) return fail(ORC_EINVAL, "grid size mismatch");
  double *dst = field == 0 ? p->h : field == 1 ? p->u : p->v;
  memcpy(dst, g, sizeof(double) * n);
  if (field == 2)
    for (int i = 0; i < p->nx; i++) { p->v[i] = 0; p->v[(size_t)p->ny * p->nx + i] = 0; }
  return ORC_OK;
}
void orc_plane_get_field(const orc_plane *p, int field, double *out) {
  size_t n = field == 2 ? (size_t)p->nx * (p->ny + 1) : (size_t)p->nx * p->ny;
  memcpy(out, field == 0 ? p->h : field == 1 ? p->u : p->v, sizeof(double) * n);
}
