/*
 * smx_oracle.c -- FP64 CPU oracle (TEST INFRASTRUCTURE ONLY; see smx_oracle.h).
 *
 * A restatement of the reference time loops in plain C.  Every block cites the
 * reference file:line it follows (paths relative to the reference root).  The
 * FFT is this file's own FP64 Stockham autosort (radices 4, 2, 3, 5 and a
 * direct DFT for other primes), following the DFT definitions FFTW documents:
 * forward e^{-2 pi i jk/n}, backward e^{+2 pi i jk/n}, all unnormalised.
 */
#include "smx_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

/* ======================================================================
 * 1-D complex FFT (Stockham autosort)
 * ==================================================================== */
typedef struct {
  int n, nf;
  int rad[64];
  cplx *tw; /* tw[m] = exp(-2 pi i m / n) */
} plan1d;

static plan1d *plan_create(int n) {
  plan1d *p = (plan1d *)calloc(1, sizeof(plan1d));
  p->n = n;
  int m = n;
  while (m % 4 == 0) { p->rad[p->nf++] = 4; m /= 4; }
  while (m % 2 == 0) { p->rad[p->nf++] = 2; m /= 2; }
  while (m % 3 == 0) { p->rad[p->nf++] = 3; m /= 3; }
  while (m % 5 == 0) { p->rad[p->nf++] = 5; m /= 5; }
  for (int f = 7; m > 1; f += 2)
    while (m % f == 0) { p->rad[p->nf++] = f; m /= f; }
  p->tw = (cplx *)malloc(sizeof(cplx) * (size_t)(n > 0 ? n : 1));
  for (int i = 0; i < n; ++i) {
    double a = -2.0 * M_PI * (double)i / (double)n;
    p->tw[i] = cos(a) + I * sin(a);
  }
  return p;
}

static void plan_destroy(plan1d *p) {
  if (!p) return;
  free(p->tw);
  free(p);
}

/* x: input/output (length n), w: scratch (length n). sign -1 forward, +1 backward. */
static void fft_exec(const plan1d *p, cplx *x, cplx *w, int sign) {
  const int n = p->n;
  cplx *X = x, *Y = w;
  int Ns = 1;
  for (int f = 0; f < p->nf; ++f) {
    const int R = p->rad[f], m = n / R, tws = n / (Ns * R);
    for (int j = 0; j < m; ++j) {
      const int k = j % Ns, base = (j - k) * R + k;
      cplx a[64], b[64];
      for (int r = 0; r < R; ++r) {
        cplx t = p->tw[r * k * tws];
        if (sign > 0) t = conj(t);
        a[r] = X[j + r * m] * t;
      }
      if (R == 2) {
        b[0] = a[0] + a[1];
        b[1] = a[0] - a[1];
      } else if (R == 4) {
        const cplx is = (sign > 0) ? I : -I; /* e^{sign*2 pi i/4} */
        cplx t0 = a[0] + a[2], t1 = a[0] - a[2], t2 = a[1] + a[3], t3 = (a[1] - a[3]) * is;
        b[0] = t0 + t2;
        b[2] = t0 - t2;
        b[1] = t1 + t3;
        b[3] = t1 - t3;
      } else if (R == 3) {
        const double s3 = 0.86602540378443864676 * (double)sign;
        cplx t = a[1] + a[2], d = (a[1] - a[2]) * (I * s3);
        b[0] = a[0] + t;
        b[1] = a[0] - 0.5 * t + d;
        b[2] = a[0] - 0.5 * t - d;
      } else {
        const int st = n / R;
        for (int q = 0; q < R; ++q) {
          cplx s = 0;
          for (int r = 0; r < R; ++r) {
            cplx t = p->tw[((q * r) % R) * st];
            if (sign > 0) t = conj(t);
            s += a[r] * t;
          }
          b[q] = s;
        }
      }
      for (int r = 0; r < R; ++r) Y[base + r * Ns] = b[r];
    }
    Ns *= R;
    cplx *t = X;
    X = Y;
    Y = t;
  }
  if (X != x) memcpy(x, X, sizeof(cplx) * (size_t)n);
}

void or_fft1d(int n, double *re_im, int sign) {
  plan1d *p = plan_create(n);
  cplx *w = (cplx *)malloc(sizeof(cplx) * (size_t)n);
  fft_exec(p, (cplx *)re_im, w, sign);
  free(w);
  plan_destroy(p);
}

/* Apply a complex FFT along one axis of a complex array: lines of length n
 * with element stride `st`, line starts given by (outer o, inner q):
 * base = o*ostride + q, o < nouter, q < ninner. */
static void fft_axis(cplx *a, int n, int64_t st, int64_t nouter, int64_t ostride, int64_t ninner, int sign) {
  plan1d *p = plan_create(n);
#pragma omp parallel
  {
    cplx *line = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    cplx *w = (cplx *)malloc(sizeof(cplx) * (size_t)n);
#pragma omp for collapse(2) schedule(static)
    for (int64_t o = 0; o < nouter; ++o)
      for (int64_t q = 0; q < ninner; ++q) {
        cplx *b = a + o * ostride + q;
        for (int i = 0; i < n; ++i) line[i] = b[i * st];
        fft_exec(p, line, w, sign);
        for (int i = 0; i < n; ++i) b[i * st] = line[i];
      }
    free(line);
    free(w);
  }
  plan_destroy(p);
}

/* ======================================================================
 * 3-D real transforms with FFTW r2c / c2r semantics
 * ==================================================================== */

/* r2c: z (half spectrum), then y, then x.  Two real z-lines share one complex
 * FFT: Z = FFT(a + i b), A = (Z[k] + conj Z[-k])/2, B = (Z[k] - conj Z[-k])/2i. */
int or_rfft3(int nx, int ny, int nz, const double *in, double *out_ri) {
  const int nzh = nz / 2 + 1;
  const int64_t nlines = (int64_t)nx * ny;
  cplx *out = (cplx *)out_ri;
  plan1d *p = plan_create(nz);
#pragma omp parallel
  {
    cplx *z = (cplx *)malloc(sizeof(cplx) * (size_t)nz);
    cplx *w = (cplx *)malloc(sizeof(cplx) * (size_t)nz);
#pragma omp for schedule(static)
    for (int64_t l = 0; l < nlines; l += 2) {
      const int two = (l + 1 < nlines);
      const double *a = in + l * nz, *b = two ? in + (l + 1) * nz : NULL;
      for (int i = 0; i < nz; ++i) z[i] = a[i] + I * (two ? b[i] : 0.0);
      fft_exec(p, z, w, -1);
      for (int k = 0; k < nzh; ++k) {
        cplx zk = z[k], zm = conj(z[(nz - k) % nz]);
        out[l * nzh + k] = 0.5 * (zk + zm);
        if (two) out[(l + 1) * nzh + k] = -0.5 * I * (zk - zm);
      }
    }
    free(z);
    free(w);
  }
  plan_destroy(p);
  fft_axis(out, ny, nzh, nx, (int64_t)ny * nzh, nzh, -1);  /* y */
  fft_axis(out, nx, (int64_t)ny * nzh, 1, 0, (int64_t)ny * nzh, -1); /* x */
  return 0;
}

/* c2r: complex backward DFTs along x and y of the half spectrum, then the 1-D
 * c2r along z, which uses only the real parts of the kz = 0 and kz = nz/2
 * entries (the half-complex stage of a multi-dimensional FFTW c2r). */
int or_irfft3(int nx, int ny, int nz, const double *in_ri, double *out) {
  const int nzh = nz / 2 + 1;
  const int64_t nlines = (int64_t)nx * ny;
  cplx *t = (cplx *)malloc(sizeof(cplx) * (size_t)(nlines * nzh));
  memcpy(t, in_ri, sizeof(cplx) * (size_t)(nlines * nzh));
  fft_axis(t, nx, (int64_t)ny * nzh, 1, 0, (int64_t)ny * nzh, +1);
  fft_axis(t, ny, nzh, nx, (int64_t)ny * nzh, nzh, +1);
  plan1d *p = plan_create(nz);
#pragma omp parallel
  {
    cplx *z = (cplx *)malloc(sizeof(cplx) * (size_t)nz);
    cplx *w = (cplx *)malloc(sizeof(cplx) * (size_t)nz);
#pragma omp for schedule(static)
    for (int64_t l = 0; l < nlines; l += 2) {
      const int two = (l + 1 < nlines);
      const cplx *A = t + l * nzh, *B = two ? t + (l + 1) * nzh : NULL;
      for (int k = 0; k < nzh; ++k) {
        cplx a = A[k], b = two ? B[k] : 0.0;
        int self = (k == 0) || (2 * k == nz);
        if (self) { a = creal(a); b = creal(b); }
        z[k] = a + I * b;
        if (!self && k > 0) z[nz - k] = conj(a) + I * conj(b);
      }
      fft_exec(p, z, w, +1);
      for (int i = 0; i < nz; ++i) {
        out[l * nz + i] = creal(z[i]);
        if (two) out[(l + 1) * nz + i] = cimag(z[i]);
      }
    }
    free(z);
    free(w);
  }
  plan_destroy(p);
  free(t);
  return 0;
}

/* ======================================================================
 * DCT-II (REDFT10) / DCT-III (REDFT01) via an n-point complex FFT
 * (Makhoul's even/odd reordering and quarter-wave twiddle).
 * REDFT10: y_k = 2 sum_j x_j cos(pi (j+1/2) k / n)
 * REDFT01: y_k = x_0 + 2 sum_{j>=1} x_j cos(pi j (k+1/2) / n)
 * ==================================================================== */
static void redft10_line(const plan1d *p, const double *x, int64_t xs, double *y, int64_t ys, cplx *v,
                         cplx *w) {
  const int n = p->n;
  for (int m = 0; 2 * m < n; ++m) v[m] = x[(2 * m) * xs];
  for (int m = 0; 2 * m + 1 < n; ++m) v[n - 1 - m] = x[(2 * m + 1) * xs];
  fft_exec(p, v, w, -1);
  for (int k = 0; k < n; ++k) {
    double a = -M_PI * (double)k / (2.0 * n);
    y[k * ys] = 2.0 * creal((cos(a) + I * sin(a)) * v[k]);
  }
}

static void redft01_line(const plan1d *p, const double *x, int64_t xs, double *y, int64_t ys, cplx *v,
                         cplx *w) {
  const int n = p->n;
  for (int k = 0; k < n; ++k) {
    double a = M_PI * (double)k / (2.0 * n);
    double xnk = (k == 0) ? 0.0 : x[(n - k) * xs];
    v[k] = (cos(a) + I * sin(a)) * (x[k * xs] - I * xnk);
  }
  fft_exec(p, v, w, +1);
  for (int m = 0; 2 * m < n; ++m) y[(2 * m) * ys] = creal(v[m]);
  for (int m = 0; 2 * m + 1 < n; ++m) y[(2 * m + 1) * ys] = creal(v[n - 1 - m]);
}

void or_redft10_1d(int n, const double *in, double *out) {
  plan1d *p = plan_create(n);
  cplx *v = (cplx *)malloc(sizeof(cplx) * n), *w = (cplx *)malloc(sizeof(cplx) * n);
  redft10_line(p, in, 1, out, 1, v, w);
  free(v); free(w); plan_destroy(p);
}

void or_redft01_1d(int n, const double *in, double *out) {
  plan1d *p = plan_create(n);
  cplx *v = (cplx *)malloc(sizeof(cplx) * n), *w = (cplx *)malloc(sizeof(cplx) * n);
  redft01_line(p, in, 1, out, 1, v, w);
  free(v); free(w); plan_destroy(p);
}

/* DST-II (RODFT10) / DST-III (RODFT01), FFTW's definitions:
 *   RODFT10: y_k = 2 sum_j x_j sin(pi (j+1/2)(k+1)/n)
 *   RODFT01: y_k = (-1)^k x_{n-1} + 2 sum_{j<n-1} x_j sin(pi (j+1)(k+1/2)/n)
 * through the cosine lines: sin(pi (j+1/2)(n-k)/n) = (-1)^j cos(pi (j+1/2) k/n), so
 * RODFT10[x]_k = REDFT10[(-1)^j x_j]_{n-1-k} and RODFT01[X]_j = (-1)^j REDFT01[X_{n-1-k}]_j.
 * t: scratch of 2n. */
static void rodft10_line(const plan1d *p, const double *x, int64_t xs, double *y, int64_t ys, cplx *v, cplx *w,
                         double *t) {
  const int n = p->n;
  for (int j = 0; j < n; ++j) t[j] = (j & 1) ? -x[j * xs] : x[j * xs];
  redft10_line(p, t, 1, t + n, 1, v, w);
  for (int k = 0; k < n; ++k) y[k * ys] = t[n + (n - 1 - k)];
}

static void rodft01_line(const plan1d *p, const double *x, int64_t xs, double *y, int64_t ys, cplx *v, cplx *w,
                         double *t) {
  const int n = p->n;
  for (int k = 0; k < n; ++k) t[k] = x[(n - 1 - k) * xs];
  redft01_line(p, t, 1, t + n, 1, v, w);
  for (int j = 0; j < n; ++j) y[j * ys] = (j & 1) ? -t[n + j] : t[n + j];
}

void or_rodft10_1d(int n, const double *in, double *out) {
  plan1d *p = plan_create(n);
  cplx *v = (cplx *)malloc(sizeof(cplx) * n), *w = (cplx *)malloc(sizeof(cplx) * n);
  double *t = (double *)malloc(sizeof(double) * 2 * n);
  rodft10_line(p, in, 1, out, 1, v, w, t);
  free(v); free(w); free(t); plan_destroy(p);
}

void or_rodft01_1d(int n, const double *in, double *out) {
  plan1d *p = plan_create(n);
  cplx *v = (cplx *)malloc(sizeof(cplx) * n), *w = (cplx *)malloc(sizeof(cplx) * n);
  double *t = (double *)malloc(sizeof(double) * 2 * n);
  rodft01_line(p, in, 1, out, 1, v, w, t);
  free(v); free(w); free(t); plan_destroy(p);
}

/* kind: 10 REDFT10, 1 REDFT01, 110 RODFT10, 101 RODFT01 */
static void r2r_axis(double *a, int n, int64_t st, int64_t nouter, int64_t ostride, int64_t ninner, int kind) {
  plan1d *p = plan_create(n);
#pragma omp parallel
  {
    cplx *v = (cplx *)malloc(sizeof(cplx) * (size_t)n), *w = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    double *line = (double *)malloc(sizeof(double) * (size_t)n);
    double *t = (double *)malloc(sizeof(double) * 2 * (size_t)n);
#pragma omp for collapse(2) schedule(static)
    for (int64_t o = 0; o < nouter; ++o)
      for (int64_t q = 0; q < ninner; ++q) {
        double *b = a + o * ostride + q;
        if (kind == 10) redft10_line(p, b, st, line, 1, v, w);
        else if (kind == 1) redft01_line(p, b, st, line, 1, v, w);
        else if (kind == 110) rodft10_line(p, b, st, line, 1, v, w, t);
        else rodft01_line(p, b, st, line, 1, v, w, t);
        for (int i = 0; i < n; ++i) b[i * st] = line[i];
      }
    free(v); free(w); free(line); free(t);
  }
  plan_destroy(p);
}

/* separable 3-D r2r transform of a dense [x][y][z] array, one kind per axis */
static void r2r3k(int nx, int ny, int nz, const double *in, double *out, int kx, int ky, int kz) {
  if (out != in) memcpy(out, in, sizeof(double) * (size_t)nx * ny * nz);
  r2r_axis(out, nz, 1, (int64_t)nx * ny, nz, 1, kz);
  r2r_axis(out, ny, nz, nx, (int64_t)ny * nz, nz, ky);
  r2r_axis(out, nx, (int64_t)ny * nz, 1, 0, (int64_t)ny * nz, kx);
}

static void r2r3(int nx, int ny, int nz, const double *in, double *out, int kind) {
  r2r3k(nx, ny, nz, in, out, kind, kind, kind);
}

int or_redft10_3d(int nx, int ny, int nz, const double *in, double *out) { r2r3(nx, ny, nz, in, out, 10); return 0; }
int or_redft01_3d(int nx, int ny, int nz, const double *in, double *out) { r2r3(nx, ny, nz, in, out, 1); return 0; }

/* ======================================================================
 * Solver state
 * ==================================================================== */
struct or_state {
  smx_config c;
  int nx, ny, nz, nzh;
  int64_t N, Nh; /* real points; spectral entries (complex for periodic, real for Neumann) */
  double dx, dy, dz, scale;
  double *vqx, *vqy, *vqz;
  double *psi, *psi_prev, *nqpast; /* nqpast: complex Nh (periodic) or real N (Neumann) */
  int64_t steps;
  int has_state, has_prev;
};

/* Periodic wavenumbers in FFT order, Nyquist kept at +pi/dx: pf3d.cpp:162-190. */
static void wavenumbers_periodic(double *v, int n, double d) {
  for (int i = 0; i < n; ++i) v[i] = (i < n / 2 + 1) ? 2.0 * M_PI * i / (d * n) : -(double)(n - i) * 2.0 * M_PI / (d * n);
}
/* DCT wavenumbers pi k / L: cosNoAdvMorph.cpp:216-231. */
static void wavenumbers_dct(double *v, int n, double d) {
  for (int i = 0; i < n; ++i) v[i] = M_PI * i / (d * n);
}

or_state *or_create(const smx_config *cfg) {
  if (!cfg || cfg->nx <= 0 || cfg->ny <= 0 || cfg->nz <= 1) return NULL;
  if (cfg->bc == SMX_BC_PERIODIC && (cfg->nz % 2)) return NULL;
  or_state *s = (or_state *)calloc(1, sizeof(or_state));
  s->c = *cfg;
  s->nx = cfg->nx; s->ny = cfg->ny; s->nz = cfg->nz; s->nzh = cfg->nz / 2 + 1;
  s->N = (int64_t)s->nx * s->ny * s->nz;
  s->dx = cfg->lx / s->nx; s->dy = cfg->ly / s->ny; s->dz = cfg->lz / s->nz;
  s->vqx = (double *)malloc(sizeof(double) * s->nx);
  s->vqy = (double *)malloc(sizeof(double) * s->ny);
  s->vqz = (double *)malloc(sizeof(double) * s->nz);
  if (cfg->bc == SMX_BC_PERIODIC) {
    s->Nh = (int64_t)s->nx * s->ny * s->nzh;
    s->scale = 1.0 / (double)s->N; /* pf3d.cpp:62 */
    wavenumbers_periodic(s->vqx, s->nx, s->dx);
    wavenumbers_periodic(s->vqy, s->ny, s->dy);
    wavenumbers_periodic(s->vqz, s->nz, s->dz);
    s->nqpast = (double *)calloc((size_t)(2 * s->Nh), sizeof(double));
  } else {
    s->Nh = s->N;
    s->scale = 0.125 / (double)s->N; /* cosNoAdvMorph.cpp:112 */
    wavenumbers_dct(s->vqx, s->nx, s->dx);
    wavenumbers_dct(s->vqy, s->ny, s->dy);
    wavenumbers_dct(s->vqz, s->nz, s->dz);
    s->nqpast = (double *)calloc((size_t)s->N, sizeof(double));
  }
  s->psi = (double *)calloc((size_t)s->N, sizeof(double));
  s->psi_prev = (double *)calloc((size_t)s->N, sizeof(double));
  return s;
}

void or_destroy(or_state *s) {
  if (!s) return;
  free(s->vqx); free(s->vqy); free(s->vqz);
  free(s->psi); free(s->psi_prev); free(s->nqpast);
  free(s);
}

/* opSH = alpha (|k|^2 - q0^2)^2, aLin = eps - opSH: pf3d.cpp:205-206, pf3dMorph.cpp:236-238 */
static inline double op_sh(const or_state *s, double k2) {
  double d = k2 - s->c.q0 * s->c.q0;
  return s->c.alpha * d * d;
}

/* Local nonlinearity: pf3d.cpp:368-369 (quintic) or the cubic model. */
static inline double nl(const or_state *s, double p) {
  if (s->c.nl_model == SMX_NL_CUBIC) return s->c.beta * p * p - s->c.gamma * p * p * p;
  double p2 = p * p;
  return s->c.beta * p2 * p - s->c.gamma * p2 * p2 * p;
}

/* N(psi) without advection, with the pf3dMorph mid-plane penalty
 * 2 wMid (psi_{k-1} - 2 psi_k + psi_{k+1})/dz^2 at k == mid_plane
 * (pf3dMorph.cpp:456-462, 572-581). */
static void nl_field(const or_state *s, const double *psi, double *out) {
  const int nz = s->nz;
  const int mid = (s->c.w_mid != 0.0) ? s->c.mid_plane : -1;
#pragma omp parallel for schedule(static)
  for (int64_t l = 0; l < (int64_t)s->nx * s->ny; ++l) {
    const double *p = psi + l * nz;
    double *o = out + l * nz;
    for (int k = 0; k < nz; ++k) o[k] = nl(s, p[k]);
    if (mid > 0 && mid < nz - 1)
      o[mid] += 2.0 * s->c.w_mid * (p[mid - 1] - 2.0 * p[mid] + p[mid + 1]) / (s->dz * s->dz);
  }
}

int or_set_psi(or_state *s, const double *psi) {
  memcpy(s->psi, psi, sizeof(double) * (size_t)s->N);
  memcpy(s->psi_prev, psi, sizeof(double) * (size_t)s->N);
  double *nr = (double *)malloc(sizeof(double) * (size_t)s->N);
  nl_field(s, s->psi, nr);
  if (s->c.bc == SMX_BC_PERIODIC)
    or_rfft3(s->nx, s->ny, s->nz, nr, s->nqpast); /* "First NqPast", pf3d.cpp:363-374 */
  else
    or_redft10_3d(s->nx, s->ny, s->nz, nr, s->nqpast); /* cosNoAdvMorph.cpp:436-442 */
  free(nr);
  s->steps = 0;
  s->has_state = 1;
  s->has_prev = 0;
  return 0;
}

int or_get_psi(const or_state *s, double *psi) {
  memcpy(psi, s->psi, sizeof(double) * (size_t)s->N);
  return 0;
}

int or_get_nqpast(const or_state *s, double *out) {
  memcpy(out, s->nqpast, sizeof(double) * (size_t)(s->c.bc == SMX_BC_PERIODIC ? 2 * s->Nh : s->N));
  return 0;
}

int64_t or_step_count(const or_state *s) { return s->steps; }

#define KIDX(i, j, k) (((int64_t)(i) * s->ny + (j)) * s->nzh + (k))

/* Advective term -U.grad(psi) of pf3d.cpp:399-451 added into nr. */
static void periodic_advection(or_state *s, const cplx *psiq, double *nr) {
  const int64_t N = s->N, Nh = s->Nh;
  const double sc = s->scale;
  cplx *gq[3], *shq = (cplx *)malloc(sizeof(cplx) * Nh);
  double *g[3], *sh = (double *)malloc(sizeof(double) * N);
  for (int a = 0; a < 3; ++a) {
    gq[a] = (cplx *)malloc(sizeof(cplx) * Nh);
    g[a] = (double *)malloc(sizeof(double) * N);
  }
  /* spectral gradient and psiSH: pf3d.cpp:399-408 */
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = 0; i < s->nx; ++i)
    for (int j = 0; j < s->ny; ++j)
      for (int k = 0; k < s->nzh; ++k) {
        int64_t id = KIDX(i, j, k);
        double k2 = s->vqx[i] * s->vqx[i] + s->vqy[j] * s->vqy[j] + s->vqz[k] * s->vqz[k];
        gq[0][id] = sc * I * s->vqx[i] * psiq[id];
        gq[1][id] = sc * I * s->vqy[j] * psiq[id];
        gq[2][id] = sc * I * s->vqz[k] * psiq[id];
        shq[id] = sc * op_sh(s, k2) * psiq[id];
      }
  for (int a = 0; a < 3; ++a) or_irfft3(s->nx, s->ny, s->nz, (double *)gq[a], g[a]); /* :410-412 */
  or_irfft3(s->nx, s->ny, s->nz, (double *)shq, sh);                                 /* :413 */
  /* force S = psiSH grad psi (pf3d.cpp:417-425) or mu grad psi with
   * mu = psiSH - eps psi - N(psi) (the variational derivative of F). */
  double *S[3];
  for (int a = 0; a < 3; ++a) S[a] = (double *)malloc(sizeof(double) * N);
#pragma omp parallel for schedule(static)
  for (int64_t x = 0; x < N; ++x) {
    double f = sh[x];
    if (s->c.force_form == SMX_FORCE_MU) f = sh[x] - s->c.eps * s->psi[x] - nl(s, s->psi[x]);
    for (int a = 0; a < 3; ++a) S[a][x] = f * g[a][x];
  }
  cplx *Sq[3];
  for (int a = 0; a < 3; ++a) {
    Sq[a] = (cplx *)malloc(sizeof(cplx) * Nh);
    or_rfft3(s->nx, s->ny, s->nz, S[a], (double *)Sq[a]); /* :427-429 */
  }
  /* Stokes: U = s CM1 (S - CM2 (k.S)), CM1 = 1/(rho eta |k|^2), CM2 = k/|k|^2, zero at k = 0
   * (pf3d.cpp:209-222, 433-447). */
  const double reta = s->c.rho * s->c.eta;
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = 0; i < s->nx; ++i)
    for (int j = 0; j < s->ny; ++j)
      for (int k = 0; k < s->nzh; ++k) {
        int64_t id = KIDX(i, j, k);
        double kv[3] = {s->vqx[i], s->vqy[j], s->vqz[k]};
        double k2 = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
        if (id == 0 || k2 == 0.0) {
          for (int a = 0; a < 3; ++a) gq[a][id] = 0.0;
          continue;
        }
        cplx dot = kv[0] * Sq[0][id] + kv[1] * Sq[1][id] + kv[2] * Sq[2][id];
        double cm1 = 1.0 / (reta * k2);
        for (int a = 0; a < 3; ++a) gq[a][id] = sc * cm1 * (Sq[a][id] - (kv[a] / k2) * dot);
      }
  double *U[3];
  for (int a = 0; a < 3; ++a) {
    U[a] = S[a];
    or_irfft3(s->nx, s->ny, s->nz, (double *)gq[a], U[a]); /* :449-451 */
  }
  /* N -= U.grad psi: pf3d.cpp:455-465 */
#pragma omp parallel for schedule(static)
  for (int64_t x = 0; x < N; ++x) nr[x] -= U[0][x] * g[0][x] + U[1][x] * g[1][x] + U[2][x] * g[2][x];
  for (int a = 0; a < 3; ++a) {
    free(gq[a]); free(g[a]); free(S[a]); free(Sq[a]);
  }
  free(shq); free(sh);
}

/* One periodic step, pf3d.cpp:395-485,549. */
static void periodic_step(or_state *s) {
  const int64_t N = s->N, Nh = s->Nh;
  const double sc = s->scale, dtd2 = 0.5 * s->c.dt;
  cplx *psiq = (cplx *)malloc(sizeof(cplx) * Nh);
  cplx *nq = (cplx *)malloc(sizeof(cplx) * Nh);
  double *nr = (double *)malloc(sizeof(double) * N);
  or_rfft3(s->nx, s->ny, s->nz, s->psi, (double *)psiq); /* :395 */
  nl_field(s, s->psi, nr);
  if (s->c.advect) periodic_advection(s, psiq, nr);
  or_rfft3(s->nx, s->ny, s->nz, nr, (double *)nq); /* :467 */
  cplx *nqp = (cplx *)s->nqpast;
  /* CNAB2: psiqNew = s (C1 psiq + dt/2 (3 Nq - NqPast)) / C2, pf3d.cpp:471-479 */
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = 0; i < s->nx; ++i)
    for (int j = 0; j < s->ny; ++j)
      for (int k = 0; k < s->nzh; ++k) {
        int64_t id = KIDX(i, j, k);
        double k2 = s->vqx[i] * s->vqx[i] + s->vqy[j] * s->vqy[j] + s->vqz[k] * s->vqz[k];
        double aLin = s->c.eps - op_sh(s, k2);
        double C1 = 1.0 + dtd2 * aLin, C2 = 1.0 - dtd2 * aLin;
        psiq[id] = sc * (C1 * psiq[id] + dtd2 * (3.0 * nq[id] - nqp[id])) / C2;
        nqp[id] = nq[id]; /* :485 */
      }
  double *t = s->psi_prev;
  s->psi_prev = s->psi;
  s->psi = t;
  or_irfft3(s->nx, s->ny, s->nz, (double *)psiq, s->psi); /* :481 */
  free(psiq); free(nq); free(nr);
}

#define NIDX(i, j, k) (((int64_t)(i) * s->ny + (j)) * s->nz + (k))

/* Advective term -U.grad(psi) of the Neumann step, added into nr; psiq = REDFT10(psi),
 * unnormalised.  The reference has the sine/cosine Stokes machinery for its Neumann box
 * (morphSurfK.cpp:331-348 wavenumbers, :392-418 plans, :864-882 projection) but never applies
 * it (:896-898) and projects every component in one all-sine basis; this is the formulation
 * with per-component parity.  psi is a cosine series along every axis (REDFT10 grid,
 * wavenumbers pi k/L).  d_a psi, the force component S_a and the velocity U_a are sine series
 * along axis a (RODFT10/01: entry m <-> wavenumber pi (m+1)/L_a, so m = n_a - 1 is pi/d_a)
 * and cosine series along the other two axes.  The Stokes projection
 * U = s (S - K (K.S)/|K|^2) / (rho eta |K|^2) acts per wavevector K = (Kx, Ky, Kz),
 * 0 <= K_a <= n_a, on the components present there (a sine component needs K_a >= 1, a
 * cosine one K_b <= n_b - 1).  With the force psiSH grad psi (pf3d.cpp:417-425) or
 * mu grad psi, this is the periodic step (pf3d.cpp:399-465) of the box mirrored evenly to
 * [0, 2L) along every axis, restricted to [0, L). */
static void neumann_advection(or_state *s, const double *psiq, double *nr, double *uout) {
  const int64_t N = s->N;
  const int n[3] = {s->nx, s->ny, s->nz};
  const double d[3] = {s->dx, s->dy, s->dz};
  const double sc = s->scale;
  double *g[3], *S[3], *Sq[3], *Uq[3];
  double *shq = (double *)malloc(sizeof(double) * N), *sh = (double *)malloc(sizeof(double) * N);
  for (int a = 0; a < 3; ++a) {
    g[a] = (double *)malloc(sizeof(double) * N);
    S[a] = (double *)malloc(sizeof(double) * N);
    Sq[a] = (double *)malloc(sizeof(double) * N);
    Uq[a] = (double *)calloc((size_t)N, sizeof(double));
  }
  /* d/dx cos(q_m x) = -q_m sin(q_m x): sine entry m holds -q_{m+1} psiq_{m+1}; entry n-1 is 0 */
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = 0; i < s->nx; ++i)
    for (int j = 0; j < s->ny; ++j)
      for (int k = 0; k < s->nz; ++k) {
        const int64_t id = NIDX(i, j, k);
        g[0][id] = (i + 1 < s->nx) ? -sc * s->vqx[i + 1] * psiq[NIDX(i + 1, j, k)] : 0.0;
        g[1][id] = (j + 1 < s->ny) ? -sc * s->vqy[j + 1] * psiq[NIDX(i, j + 1, k)] : 0.0;
        g[2][id] = (k + 1 < s->nz) ? -sc * s->vqz[k + 1] * psiq[NIDX(i, j, k + 1)] : 0.0;
        const double k2 = s->vqx[i] * s->vqx[i] + s->vqy[j] * s->vqy[j] + s->vqz[k] * s->vqz[k];
        shq[id] = sc * op_sh(s, k2) * psiq[id];
      }
  r2r3k(s->nx, s->ny, s->nz, g[0], g[0], 101, 1, 1);
  r2r3k(s->nx, s->ny, s->nz, g[1], g[1], 1, 101, 1);
  r2r3k(s->nx, s->ny, s->nz, g[2], g[2], 1, 1, 101);
  or_redft01_3d(s->nx, s->ny, s->nz, shq, sh);
#pragma omp parallel for schedule(static)
  for (int64_t x = 0; x < N; ++x) {
    double f = sh[x];
    if (s->c.force_form == SMX_FORCE_MU) f = sh[x] - s->c.eps * s->psi[x] - nl(s, s->psi[x]);
    for (int a = 0; a < 3; ++a) S[a][x] = f * g[a][x];
  }
  r2r3k(s->nx, s->ny, s->nz, S[0], Sq[0], 110, 10, 10);
  r2r3k(s->nx, s->ny, s->nz, S[1], Sq[1], 10, 110, 10);
  r2r3k(s->nx, s->ny, s->nz, S[2], Sq[2], 10, 10, 110);
  const double reta = s->c.rho * s->c.eta;
#pragma omp parallel for collapse(2) schedule(static)
  for (int K0 = 0; K0 <= n[0]; ++K0)
    for (int K1 = 0; K1 <= n[1]; ++K1)
      for (int K2 = 0; K2 <= n[2]; ++K2) {
        const int K[3] = {K0, K1, K2};
        double q[3], k2 = 0.0;
        for (int a = 0; a < 3; ++a) {
          q[a] = M_PI * K[a] / (d[a] * n[a]);
          k2 += q[a] * q[a];
        }
        int64_t idx[3];
        int present[3];
        double dot = 0.0;
        for (int a = 0; a < 3; ++a) {
          int e[3];
          present[a] = K[a] >= 1;
          for (int b = 0; b < 3; ++b) {
            e[b] = (b == a) ? K[b] - 1 : K[b];
            if (b != a && K[b] > n[b] - 1) present[a] = 0;
          }
          idx[a] = present[a] ? NIDX(e[0], e[1], e[2]) : -1;
          if (present[a]) dot += q[a] * Sq[a][idx[a]];
        }
        if (k2 == 0.0) continue;
        const double cm1 = 1.0 / (reta * k2);
        for (int a = 0; a < 3; ++a)
          if (present[a]) Uq[a][idx[a]] = sc * cm1 * (Sq[a][idx[a]] - (q[a] / k2) * dot);
      }
  r2r3k(s->nx, s->ny, s->nz, Uq[0], Uq[0], 101, 1, 1);
  r2r3k(s->nx, s->ny, s->nz, Uq[1], Uq[1], 1, 101, 1);
  r2r3k(s->nx, s->ny, s->nz, Uq[2], Uq[2], 1, 1, 101);
  if (uout)
    for (int a = 0; a < 3; ++a) memcpy(uout + a * N, Uq[a], sizeof(double) * (size_t)N);
  if (nr) {
#pragma omp parallel for schedule(static)
    for (int64_t x = 0; x < N; ++x) nr[x] -= Uq[0][x] * g[0][x] + Uq[1][x] * g[1][x] + Uq[2][x] * g[2][x];
  }
  for (int a = 0; a < 3; ++a) {
    free(g[a]); free(S[a]); free(Sq[a]); free(Uq[a]);
  }
  free(shq); free(sh);
}

/* One Neumann step, cosFFT/mesabi/cosNoAdvMorph.cpp:490-530,619 (plus the advective term
 * above when advection is on). */
static void neumann_step(or_state *s) {
  const int64_t N = s->N;
  const double sc = s->scale, dtd2 = 0.5 * s->c.dt;
  double *nr = (double *)malloc(sizeof(double) * N);
  double *nq = (double *)malloc(sizeof(double) * N);
  double *psiq = (double *)malloc(sizeof(double) * N);
  or_redft10_3d(s->nx, s->ny, s->nz, s->psi, psiq);      /* :512 */
  nl_field(s, s->psi, nr);                               /* :494-504 */
  if (s->c.advect) neumann_advection(s, psiq, nr, NULL);
  or_redft10_3d(s->nx, s->ny, s->nz, nr, nq);            /* :508 */
#pragma omp parallel for collapse(2) schedule(static)
  for (int i = 0; i < s->nx; ++i)
    for (int j = 0; j < s->ny; ++j)
      for (int k = 0; k < s->nz; ++k) {
        int64_t id = NIDX(i, j, k);
        double k2 = s->vqx[i] * s->vqx[i] + s->vqy[j] * s->vqy[j] + s->vqz[k] * s->vqz[k];
        double aLin = s->c.eps - op_sh(s, k2); /* :431-433 */
        double C1 = 1.0 + dtd2 * aLin, C2 = 1.0 - dtd2 * aLin;
        psiq[id] = sc * (C1 * psiq[id] + dtd2 * (3.0 * nq[id] - s->nqpast[id])) / C2; /* :524-525 */
        s->nqpast[id] = nq[id]; /* :490 of the next step */
      }
  double *t = s->psi_prev;
  s->psi_prev = s->psi;
  s->psi = t;
  or_redft01_3d(s->nx, s->ny, s->nz, psiq, s->psi); /* :530 */
  free(nr); free(nq); free(psiq);
}

/* Stokes velocity (U_x, U_y, U_z), each dense [x][y][z], of the current Neumann state, as the
 * advected step forms it. */
int or_velocity(or_state *s, double *u) {
  if (!s->has_state) return SMX_ERR_STATE;
  if (s->c.bc != SMX_BC_NEUMANN) return SMX_ERR_UNSUPPORTED;
  double *psiq = (double *)malloc(sizeof(double) * (size_t)s->N);
  or_redft10_3d(s->nx, s->ny, s->nz, s->psi, psiq);
  neumann_advection(s, psiq, NULL, u);
  free(psiq);
  return 0;
}

int or_step(or_state *s, int64_t nsteps) {
  if (!s->has_state) return SMX_ERR_STATE;
  for (int64_t n = 0; n < nsteps; ++n) {
    if (s->c.bc == SMX_BC_PERIODIC) periodic_step(s);
    else neumann_step(s);
    s->steps++;
    s->has_prev = 1;
  }
  return 0;
}

/* L1 = sum|psiNew - psi| / (dt sum|psiNew|), pf3d.cpp:494-509 */
double or_l1(const or_state *s) {
  if (!s->has_prev) return NAN;
  double a = 0.0, b = 0.0;
  for (int64_t x = 0; x < s->N; ++x) {
    a += fabs(s->psi[x] - s->psi_prev[x]);
    b += fabs(s->psi[x]);
  }
  return a / (s->c.dt * b);
}

/* Local energy density terms of the model: energyPsi.m:225-228 (quintic:
 * -beta/4 psi^4 + gamma/6 psi^6) or the cubic model (-beta/3 psi^3 + gamma/4 psi^4). */
static inline double f_local(const or_state *s, double p) {
  double p2 = p * p;
  if (s->c.nl_model == SMX_NL_CUBIC) return -s->c.beta / 3.0 * p2 * p + s->c.gamma / 4.0 * p2 * p2;
  return -s->c.beta / 4.0 * p2 * p2 + s->c.gamma / 6.0 * p2 * p2 * p2;
}

/* Spectral energy: quadratic part by Parseval, local part by the rectangle rule. */
double or_energy(const or_state *s) {
  const double dV = s->dx * s->dy * s->dz, q02 = s->c.q0 * s->c.q0;
  double quad = 0.0, loc = 0.0;
  if (s->c.bc == SMX_BC_PERIODIC) {
    cplx *q = (cplx *)malloc(sizeof(cplx) * s->Nh);
    or_rfft3(s->nx, s->ny, s->nz, s->psi, (double *)q);
    for (int i = 0; i < s->nx; ++i)
      for (int j = 0; j < s->ny; ++j)
        for (int k = 0; k < s->nzh; ++k) {
          double k2 = s->vqx[i] * s->vqx[i] + s->vqy[j] * s->vqy[j] + s->vqz[k] * s->vqz[k];
          double w = (k == 0 || 2 * k == s->nz) ? 1.0 : 2.0;
          double m = 0.5 * s->c.alpha * (q02 - k2) * (q02 - k2) - 0.5 * s->c.eps;
          cplx v = q[KIDX(i, j, k)];
          quad += w * m * (creal(v) * creal(v) + cimag(v) * cimag(v));
        }
    quad *= dV / (double)s->N; /* sum_x psi^2 = (1/N) sum_k |psiq_k|^2 */
    free(q);
  } else {
    double *q = (double *)malloc(sizeof(double) * s->N);
    or_redft10_3d(s->nx, s->ny, s->nz, s->psi, q);
    for (int i = 0; i < s->nx; ++i)
      for (int j = 0; j < s->ny; ++j)
        for (int k = 0; k < s->nz; ++k) {
          double k2 = s->vqx[i] * s->vqx[i] + s->vqy[j] * s->vqy[j] + s->vqz[k] * s->vqz[k];
          double w = (i == 0 ? 0.5 : 1.0) * (j == 0 ? 0.5 : 1.0) * (k == 0 ? 0.5 : 1.0);
          double m = 0.5 * s->c.alpha * (q02 - k2) * (q02 - k2) - 0.5 * s->c.eps;
          double v = q[((int64_t)i * s->ny + j) * s->nz + k];
          quad += w * m * v * v;
        }
    quad *= dV * s->scale; /* sum_x psi^2 = s sum_k w_k y_k^2, s = 1/(8N) */
    free(q);
  }
  for (int64_t x = 0; x < s->N; ++x) loc += f_local(s, s->psi[x]);
  return quad + loc * dV;
}

/* energyPsi.m:131-234: periodic-wrap second differences, trapz over
 * 0..L-dx on each axis (end points weighted 1/2). */
double or_energy_fd_trapz(const or_state *s) {
  const int nx = s->nx, ny = s->ny, nz = s->nz;
  const double q02 = s->c.q0 * s->c.q0;
  const double idx2 = 1.0 / (s->dx * s->dx), idy2 = 1.0 / (s->dy * s->dy), idz2 = 1.0 / (s->dz * s->dz);
  double tot = 0.0;
  for (int i = 0; i < nx; ++i) {
    int im = (i + nx - 1) % nx, ip = (i + 1) % nx;
    double wi = (i == 0 || i == nx - 1) ? 0.5 : 1.0;
    for (int j = 0; j < ny; ++j) {
      int jm = (j + ny - 1) % ny, jp = (j + 1) % ny;
      double wj = (j == 0 || j == ny - 1) ? 0.5 : 1.0;
      for (int k = 0; k < nz; ++k) {
        int km = (k + nz - 1) % nz, kp = (k + 1) % nz;
        double wk = (k == 0 || k == nz - 1) ? 0.5 : 1.0;
#define P(a, b, c) s->psi[((int64_t)(a) * ny + (b)) * nz + (c)]
        double p = P(i, j, k);
        double lap = idz2 * (P(i, j, km) - 2 * p + P(i, j, kp)) + idy2 * (P(i, jm, k) - 2 * p + P(i, jp, k)) +
                     idx2 * (P(im, j, k) - 2 * p + P(ip, j, k));
#undef P
        double t = q02 * p + lap;
        double e = -0.5 * s->c.eps * p * p + 0.5 * s->c.alpha * t * t + f_local(s, p);
        tot += wi * wj * wk * e;
      }
    }
  }
  return tot * s->dx * s->dy * s->dz;
}

int or_eckhaus_scan(const smx_config *cfg, const double *psi0, int nq, const double *q, const double *amp,
                    double *energy) {
  or_state *s = or_create(cfg);
  if (!s) return -1;
  const int nx = s->nx, ny = s->ny, nz = s->nz;
  double *dpz = (double *)malloc(sizeof(double) * (size_t)s->N);
  if (!dpz) {
    or_destroy(s);
    return -2;
  }
  /* energyPsi.m:68-101, periodic neighbours of :86-98 */
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k) {
        const int kp = (k + 1) % nz, kpp = (k + 2) % nz, km = (k + nz - 1) % nz, kmm = (k + nz - 2) % nz;
        const double *c = psi0 + ((int64_t)i * ny + j) * nz;
        dpz[((int64_t)i * ny + j) * nz + k] =
            -(1.0 / 12) * c[kpp] + (2.0 / 3) * c[kp] - (2.0 / 3) * c[kmm] + (1.0 / 12) * c[km];
      }
  for (int m = 0; m < nq; ++m) { /* :108-124 */
    for (int64_t x = 0; x < s->N; ++x) {
      const int k = (int)(x % nz);
      s->psi[x] = psi0[x] + dpz[x] * amp[m] * cos((1.0 + q[m]) * (k + 1) * s->dz);
    }
    energy[m] = or_energy_fd_trapz(s); /* :131-234 */
  }
  free(dpz);
  or_destroy(s);
  return 0;
}

double or_dominant_wavenumber(const or_state *s) {
  double num = 0.0, den = 0.0;
  if (s->c.bc == SMX_BC_PERIODIC) {
    cplx *q = (cplx *)malloc(sizeof(cplx) * s->Nh);
    or_rfft3(s->nx, s->ny, s->nz, s->psi, (double *)q);
    for (int i = 0; i < s->nx; ++i)
      for (int j = 0; j < s->ny; ++j)
        for (int k = 0; k < s->nzh; ++k) {
          if (i == 0 && j == 0 && k == 0) continue;
          double k2 = s->vqx[i] * s->vqx[i] + s->vqy[j] * s->vqy[j] + s->vqz[k] * s->vqz[k];
          double w = (k == 0 || 2 * k == s->nz) ? 1.0 : 2.0;
          cplx v = q[KIDX(i, j, k)];
          double pw = w * (creal(v) * creal(v) + cimag(v) * cimag(v));
          num += pw * sqrt(k2);
          den += pw;
        }
    free(q);
  } else {
    double *q = (double *)malloc(sizeof(double) * s->N);
    or_redft10_3d(s->nx, s->ny, s->nz, s->psi, q);
    for (int i = 0; i < s->nx; ++i)
      for (int j = 0; j < s->ny; ++j)
        for (int k = 0; k < s->nz; ++k) {
          if (i == 0 && j == 0 && k == 0) continue;
          double k2 = s->vqx[i] * s->vqx[i] + s->vqy[j] * s->vqy[j] + s->vqz[k] * s->vqz[k];
          double w = (i == 0 ? 0.5 : 1.0) * (j == 0 ? 0.5 : 1.0) * (k == 0 ? 0.5 : 1.0);
          double v = q[((int64_t)i * s->ny + j) * s->nz + k];
          num += w * v * v * sqrt(k2);
          den += w * v * v;
        }
    free(q);
  }
  return den > 0 ? num / den : 0.0;
}

/* dpsi/dt = F^-1[s (aLin psiq + Nq)] of the current state:
 * surfTrack.cpp:606,622 (periodic; Nq includes the advective term, :585-594),
 * cosNoAdvMorphCurv.cpp:744,764 (Neumann; with advection on, the Neumann advective term). */
int or_dtpsi(or_state *s, double *out) {
  const int64_t N = s->N;
  double *nr = (double *)malloc(sizeof(double) * N);
  nl_field(s, s->psi, nr);
  if (s->c.bc == SMX_BC_PERIODIC) {
    cplx *psiq = (cplx *)malloc(sizeof(cplx) * s->Nh), *nq = (cplx *)malloc(sizeof(cplx) * s->Nh);
    or_rfft3(s->nx, s->ny, s->nz, s->psi, (double *)psiq);
    if (s->c.advect) periodic_advection(s, psiq, nr);
    or_rfft3(s->nx, s->ny, s->nz, nr, (double *)nq);
    for (int i = 0; i < s->nx; ++i)
      for (int j = 0; j < s->ny; ++j)
        for (int k = 0; k < s->nzh; ++k) {
          int64_t id = KIDX(i, j, k);
          double k2 = s->vqx[i] * s->vqx[i] + s->vqy[j] * s->vqy[j] + s->vqz[k] * s->vqz[k];
          nq[id] = s->scale * ((s->c.eps - op_sh(s, k2)) * psiq[id] + nq[id]);
        }
    or_irfft3(s->nx, s->ny, s->nz, (double *)nq, out);
    free(psiq); free(nq);
  } else {
    double *psiq = (double *)malloc(sizeof(double) * N), *nq = (double *)malloc(sizeof(double) * N);
    or_redft10_3d(s->nx, s->ny, s->nz, s->psi, psiq);
    if (s->c.advect) neumann_advection(s, psiq, nr, NULL);
    or_redft10_3d(s->nx, s->ny, s->nz, nr, nq);
    for (int i = 0; i < s->nx; ++i)
      for (int j = 0; j < s->ny; ++j)
        for (int k = 0; k < s->nz; ++k) {
          int64_t id = ((int64_t)i * s->ny + j) * s->nz + k;
          double k2 = s->vqx[i] * s->vqx[i] + s->vqy[j] * s->vqy[j] + s->vqz[k] * s->vqz[k];
          nq[id] = s->scale * ((s->c.eps - op_sh(s, k2)) * psiq[id] + nq[id]);
        }
    or_redft01_3d(s->nx, s->ny, s->nz, nq, out);
    free(psiq); free(nq);
  }
  free(nr);
  return 0;
}

/* Periodic surfTrack on the current state (surfTrack.cpp:611-626 spectral
 * derivatives, :767-806 column scan). */
static int surftrack_periodic(or_state *s, double arm, double vsgn, double *surfz, double *vel, double *H,
                              double *Kc, int64_t *found) {
  const int64_t N = s->N, Nh = s->Nh;
  const double sc = s->scale;
  cplx *psiq = (cplx *)malloc(sizeof(cplx) * Nh), *t = (cplx *)malloc(sizeof(cplx) * Nh);
  double *f[8];
  for (int a = 0; a < 8; ++a) f[a] = (double *)malloc(sizeof(double) * N);
  /* f: 0 gx, 1 gy, 2 gz, 3 curvH numerator, 4 dxx, 5 dyy, 6 dxy, 7 dtpsi */
  or_rfft3(s->nx, s->ny, s->nz, s->psi, (double *)psiq);
  for (int a = 0; a < 7; ++a) {
    for (int i = 0; i < s->nx; ++i)
      for (int j = 0; j < s->ny; ++j)
        for (int k = 0; k < s->nzh; ++k) {
          int64_t id = KIDX(i, j, k);
          double kx = s->vqx[i], ky = s->vqy[j], kz = s->vqz[k];
          cplx m = 0;
          switch (a) {
          case 0: m = I * kx; break;
          case 1: m = I * ky; break;
          case 2: m = I * kz; break;
          case 3: m = (I * kx) * (I * kx) + (I * ky) * (I * ky); break; /* :612-613 */
          case 4: m = (I * kx) * (I * kx); break;                       /* :615 */
          case 5: m = (I * ky) * (I * ky); break;                       /* :616 */
          case 6: m = (I * ky) * (I * kx); break;                       /* :617 */
          }
          t[id] = sc * m * psiq[id];
        }
    or_irfft3(s->nx, s->ny, s->nz, (double *)t, f[a]);
  }
  or_dtpsi(s, f[7]);
  int64_t nf = 0;
  for (int i = 0; i < s->nx; ++i)
    for (int j = 0; j < s->ny; ++j) {
      int track = 0;
      const int64_t col = ((int64_t)i * s->ny + j) * s->nz;
      for (int k = 0; k < s->nz; ++k) {
        int64_t id = col + k;
        if (s->psi[id] > arm && track == 0) track = 1;
        if (s->psi[id] < 0 && track == 1) {
          if (fabs(s->psi[id]) > fabs(s->psi[id - 1])) id = id - 1;
          double g = sqrt(f[0][id] * f[0][id] + f[1][id] * f[1][id] + f[2][id] * f[2][id]);
          int64_t c2 = (int64_t)i * s->ny + j;
          surfz[c2] = k; /* the reference writes k, not the shifted index (:785) */
          vel[c2] = vsgn * f[7][id] / g;
          H[c2] = f[3][id] / g;
          Kc[c2] = (f[4][id] * f[5][id] - f[6][id] * f[6][id]) / (g * g);
          track = 2;
          nf++;
        }
      }
    }
  *found = nf;
  for (int a = 0; a < 8; ++a) free(f[a]);
  free(psiq); free(t);
  return 0;
}

/* Neumann surfTrack on the current state (cosNoAdvMorphCurv.cpp:744-762
 * spectral second derivatives, :811-873 central differences, one-sided x2 at
 * the walls, :877-961 column scan with Megrabov H and K). */
static int surftrack_neumann(or_state *s, double arm, double vsgn, double *surfz, double *vel, double *H,
                             double *Kc, int64_t *found) {
  const int nx = s->nx, ny = s->ny, nz = s->nz;
  const int64_t N = s->N;
  double *psiq = (double *)malloc(sizeof(double) * N), *t = (double *)malloc(sizeof(double) * N);
  double *dxx = (double *)malloc(sizeof(double) * N), *dyy = (double *)malloc(sizeof(double) * N),
         *dzz = (double *)malloc(sizeof(double) * N), *dt = (double *)malloc(sizeof(double) * N);
  double *gx = (double *)malloc(sizeof(double) * N), *gy = (double *)malloc(sizeof(double) * N),
         *gz = (double *)malloc(sizeof(double) * N);
  or_redft10_3d(nx, ny, nz, s->psi, psiq);
  for (int a = 0; a < 3; ++a) {
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < ny; ++j)
        for (int k = 0; k < nz; ++k) {
          int64_t id = ((int64_t)i * ny + j) * nz + k;
          double q = a == 0 ? s->vqx[i] : (a == 1 ? s->vqy[j] : s->vqz[k]);
          t[id] = -s->scale * q * q * psiq[id]; /* :746-750 */
        }
    or_redft01_3d(nx, ny, nz, t, a == 0 ? dxx : (a == 1 ? dyy : dzz));
  }
  or_dtpsi(s, dt);
  const double tdx = 2 * s->dx, tdy = 2 * s->dy, tdz = 2 * s->dz; /* :167-169 */
#define P(a, b, c) s->psi[((int64_t)(a) * ny + (b)) * nz + (c)]
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k) {
        int64_t id = ((int64_t)i * ny + j) * nz + k;
        /* x from the current state (the reference uses a stale gathered copy, :811-841) */
        if (i > 0 && i < nx - 1) gx[id] = (P(i + 1, j, k) - P(i - 1, j, k)) / tdx;
        else if (i == 0) gx[id] = 2 * (P(1, j, k) - P(0, j, k)) / tdx;
        else gx[id] = 2 * (P(i, j, k) - P(i - 1, j, k)) / tdx;
        if (j != 0 && j != ny - 1) gy[id] = (P(i, j + 1, k) - P(i, j - 1, k)) / tdy; /* :859-865 */
        else if (j == 0) gy[id] = 2 * (P(i, 1, k) - P(i, 0, k)) / tdy;
        else gy[id] = 2 * (P(i, j, k) - P(i, j - 1, k)) / tdy;
        if (k != 0 && k != nz - 1) gz[id] = (P(i, j, k + 1) - P(i, j, k - 1)) / tdz; /* :866-872 */
        else if (k == 0) gz[id] = 2 * (P(i, j, 1) - P(i, j, 0)) / tdz;
        else gz[id] = 2 * (P(i, j, k) - P(i, j, k - 1)) / tdz;
      }
  int64_t nf = 0;
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j) {
      int track = 0;
      const int64_t col = ((int64_t)i * ny + j) * nz, c2 = (int64_t)i * ny + j;
      for (int k = nz - 1; k > -1; --k) { /* :884 */
        int64_t id = col + k;
        if (s->psi[id] > arm && track == 0) track = 1;
        if (s->psi[id] < 0.0 && track == 1) {
          int k2 = k;
          if (fabs(s->psi[id]) > fabs(s->psi[id + 1])) { id = id + 1; k2 = k + 1; } /* :898-902 */
          surfz[c2] = k2;
          double g = sqrt(gx[id] * gx[id] + gy[id] * gy[id] + gz[id] * gz[id]);
          vel[c2] = vsgn * dt[id] / g; /* :909 */
          double dxy;
#define GX(jj, kk) gx[((int64_t)i * ny + (jj)) * nz + (kk)]
          if (j > 0 && j < ny - 1) dxy = (GX(j + 1, k2) - GX(j - 1, k2)) / tdy; /* :913-924 */
          else if (j == 0) dxy = 2 * (GX(1, k2) - GX(0, k2)) / tdy;
          else dxy = 2 * (GX(j, k2) - GX(j - 1, k2)) / tdy;
#undef GX
          /* :926-940 chooses the stencil by k while centring it on index k2; a
           * centred stencil that would leave the column falls back to one-sided. */
          double dxz, dyz;
          int lo = (int)(id - col) - 1, hi = (int)(id - col) + 1;
          if (k > 0 && k < nz - 1 && lo >= 0 && hi <= nz - 1) {
            dxz = (gx[id + 1] - gx[id - 1]) / tdz;
            dyz = (gy[id + 1] - gy[id - 1]) / tdz;
          } else if (k == 0 || lo < 0) {
            dxz = 2 * (gx[id + 1] - gx[id]) / tdz;
            dyz = 2 * (gy[id + 1] - gy[id]) / tdz;
          } else {
            dxz = 2 * (gx[id] - gx[id - 1]) / tdz;
            dyz = 2 * (gy[id] - gy[id - 1]) / tdz;
          }
          double px = gx[id], py = gy[id], pz = gz[id];
          double fxx = dxx[id], fyy = dyy[id], fzz = dzz[id];
          H[c2] = ((py * py + pz * pz) * fxx + (px * px + pz * pz) * fyy + (px * px + py * py) * fzz -
                   2 * (px * py * dxy + px * pz * dxz + py * pz * dyz)) /
                  (g * g * g); /* :945-952 */
          Kc[c2] = (pz * pz * (fxx * fyy - dxy * dxy) + px * px * (fyy * fzz - dyz * dyz) +
                    py * py * (fxx * fzz - dxz * dxz) +
                    2 * (py * dxy * (pz * dxz - px * fzz) + px * dxz * (py * dyz - pz * fyy) +
                         pz * dyz * (px * dxy - py * fxx))) /
                   (g * g * g * g); /* :954-961 */
          track = 2;
          nf++;
        }
      }
    }
#undef P
  *found = nf;
  free(psiq); free(t); free(dxx); free(dyy); free(dzz); free(dt); free(gx); free(gy); free(gz);
  return 0;
}

int or_surftrack(or_state *s, double arm, double vel_sign, double *surfz, double *vel, double *H, double *K,
                 int64_t *found) {
  if (!s->has_state) return SMX_ERR_STATE;
  if (s->c.bc == SMX_BC_PERIODIC) return surftrack_periodic(s, arm, vel_sign, surfz, vel, H, K, found);
  return surftrack_neumann(s, arm, vel_sign, surfz, vel, H, K, found);
}

/* ======================================================================
 * Initial conditions
 * ==================================================================== */

/* Counter-based SplitMix64: the idx-th output of the stream seeded with
 * seed ^ box; u = (x >> 11) * 2^-53 in [0,1). */
double or_splitmix_uniform(uint64_t seed, uint64_t box, uint64_t idx) {
  uint64_t z = (seed ^ box) + (idx + 1) * 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  z ^= z >> 31;
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

void or_ic_noise(int nx, int ny, int nz, uint64_t seed, uint64_t box, double amp, double *psi) {
  const int64_t N = (int64_t)nx * ny * nz;
  for (int64_t i = 0; i < N; ++i) psi[i] = amp * (2.0 * or_splitmix_uniform(seed, box, (uint64_t)i) - 1.0);
}

/* psi0 = a cos(q0 sqrt(max(x^2+y^2-z^2, 0) + 4)) about the box centre plus
 * uniform noise of amplitude `noise`. */
void or_ic_hyperboloid(int nx, int ny, int nz, double dx, double dy, double dz, double q0, double a,
                       double noise, uint64_t seed, double *psi) {
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k) {
        int64_t id = ((int64_t)i * ny + j) * nz + k;
        double x = (i - nx / 2) * dx, y = (j - ny / 2) * dy, z = (k - nz / 2) * dz;
        double r2 = x * x + y * y - z * z;
        if (r2 < 0) r2 = 0;
        psi[id] = a * cos(q0 * sqrt(r2 + 4.0)) + noise * (2.0 * or_splitmix_uniform(seed, 0, (uint64_t)id) - 1.0);
      }
}

/* pf3dVal.cpp:214-326 */
/* trapz of morphSurfAmp.cpp:76-86: 0.5 h sum_{n < nend-nstart} (f[n] + f[n+1]) */
static double ref_trapz(int nstart, int nend, double h, const double *f) {
  double t = 0.0;
  for (int n = 0; n < nend - nstart; n++) t = t + f[n] + f[n + 1];
  return 0.5 * h * t;
}

/* trapz over x (innermost), then y, then z of psi^2 on [i0,i1]x[j0,j1]x[k0,k1]
 * (morphSurfAmp.cpp:1018-1049 and 1085-1114) */
static double ref_box(int ny, int nz, double dx, double dy, double dz, const double *psi, int i0, int i1, int j0,
                      int j1, int k0, int k1, double *fx, double *fy, double *fz) {
  int c3 = 0;
  for (int k2 = k0; k2 < k1 + 1; k2++) {
    int c2 = 0;
    for (int j2 = j0; j2 < j1 + 1; j2++) {
      int c1 = 0;
      for (int i2 = i0; i2 < i1 + 1; i2++) {
        const double v = psi[((int64_t)i2 * ny + j2) * nz + k2];
        fx[c1++] = v * v;
      }
      fy[c2++] = ref_trapz(i0, i1, dx, fx);
    }
    fz[c3++] = ref_trapz(j0, j1, dy, fy);
  }
  return ref_trapz(k0, k1, dz, fz);
}

int or_psi_avg(int nx, int ny, int nz, double dx, double dy, double dz, int nw, double q0, const double *psi,
               double *out) {
  const int nl = nw / 2; /* nlambda = Nw/2, :1011 */
  if (nl < 1 || 3 * nl > nx - 1 || 3 * nl > ny - 1 || 3 * nl > nz - 1) return -1;
  const int wmax = 2 * nl + 1;
  double *fx = (double *)malloc(sizeof(double) * 3 * wmax);
  if (!fx) return -2;
  double *fy = fx + wmax, *fz = fy + wmax;
  /* psiRef over [nl, 3 nl]^3, :1013-1049 */
  const double ref = ref_box(ny, nz, dx, dy, dz, psi, nl, 3 * nl, nl, 3 * nl, nl, 3 * nl, fx, fy, fz);
  const double lam = 2.0 * M_PI / q0, lam3 = lam * lam * lam; /* 8 pi^3 at q0 = 1, :1053 */
  for (int k = 0; k < nz; k++) {
    const int k0 = k - nl < 0 ? 0 : k - nl, k1 = k + nl > nz - 1 ? nz - 1 : k + nl;
    for (int j = 0; j < ny; j++) {
      const int j0 = j - nl < 0 ? 0 : j - nl, j1 = j + nl > ny - 1 ? ny - 1 : j + nl;
      for (int i = 0; i < nx; i++) {
        const int i0 = i - nl < 0 ? 0 : i - nl, i1 = i + nl > nx - 1 ? nx - 1 : i + nl;
        const double l1 = (i1 - i0) * dx, l2 = (j1 - j0) * dy, l3 = (k1 - k0) * dz; /* :1078-1080 */
        const double t = ref_box(ny, nz, dx, dy, dz, psi, i0, i1, j0, j1, k0, k1, fx, fy, fz);
        out[((int64_t)i * ny + j) * nz + k] = t / (ref * l1 * l2 * l3 / lam3); /* :1118 */
      }
    }
  }
  free(fx);
  return 0;
}

int or_growth_sweep(const smx_config *cfg, double mq0, double dmq, int nq, double *growth) {
  or_state *s = or_create(cfg);
  if (!s) return SMX_ERR_INVALID;
  double *psi = (double *)malloc(sizeof(double) * s->N);
  for (int q = 0; q < nq; ++q) {
    double mq = mq0 + q * dmq;
    for (int i = 0; i < s->nx; ++i)
      for (int j = 0; j < s->ny; ++j)
        for (int k = 0; k < s->nz; ++k) psi[((int64_t)i * s->ny + j) * s->nz + k] = 1e-4 * cos(mq * k * s->dz);
    or_set_psi(s, psi);
    or_step(s, 3);
    int64_t c = ((int64_t)(s->nx / 2) * s->ny + s->ny / 2) * s->nz + s->nz / 2;
    growth[q] = (s->psi[c] - s->psi_prev[c]) / (s->c.dt * s->psi_prev[c]); /* :320 */
  }
  free(psi);
  or_destroy(s);
  return 0;
}

/* ======================================================================
 * Discrete adjoint of the periodic CNAB2 step (no advection, no penalty).
 *
 * Forward map in real space (pf3d.cpp:471-481 with the bootstrap :363-374):
 *   psi_{n+1} = A psi_n + B (c_n N(psi_n) - N(psi_{n-1})),  c_0 = 2 (N_{-1} := N_0 drops out), c_n = 3,
 *   A = F^-1 diag(C1/C2) F,  B = F^-1 diag(dt/2 / C2) F  (symmetric: real even multipliers).
 * Loss J = 1/2 sum_x (psi_T - g)^2, g = amp cos(q0 z), z = k dz.
 * Adjoint a_n = dJ/dpsi_n:
 *   a_T = psi_T - g,  a_n = A a_{n+1} + N'(psi_n) B (c_n a_{n+1} - a_{n+2}),  a_{T+1} = 0.
 * Parameter gradients (theta = eps or q0, dL/deps = 1, dL/dq0 = 4 alpha q0 (k^2 - q0^2)):
 *   dJ/dtheta = sum_n < a_{n+1}, F^-1[ dt/2 dL/dtheta (Psi_n + Psi_{n+1}) / C2 ] >  (+ target term for q0).
 * ==================================================================== */
static inline double nl_prime(const or_state *s, double p) {
  if (s->c.nl_model == SMX_NL_CUBIC) return 2.0 * s->c.beta * p - 3.0 * s->c.gamma * p * p;
  return 3.0 * s->c.beta * p * p - 5.0 * s->c.gamma * p * p * p * p;
}

/* out = F^-1[m(k) F(a)] with F normalised, m chosen by `which`: 0 -> C1/C2, 1 -> dt/2/C2. */
static void apply_mult(or_state *s, const double *a, double *out, int which) {
  cplx *q = (cplx *)malloc(sizeof(cplx) * s->Nh);
  or_rfft3(s->nx, s->ny, s->nz, a, (double *)q);
  const double dtd2 = 0.5 * s->c.dt;
  for (int i = 0; i < s->nx; ++i)
    for (int j = 0; j < s->ny; ++j)
      for (int k = 0; k < s->nzh; ++k) {
        double k2 = s->vqx[i] * s->vqx[i] + s->vqy[j] * s->vqy[j] + s->vqz[k] * s->vqz[k];
        double L = s->c.eps - op_sh(s, k2), C1 = 1.0 + dtd2 * L, C2 = 1.0 - dtd2 * L;
        q[KIDX(i, j, k)] *= s->scale * (which == 0 ? C1 / C2 : dtd2 / C2);
      }
  or_irfft3(s->nx, s->ny, s->nz, (double *)q, out);
  free(q);
}

int or_adjoint_grad(const smx_config *cfg, const double *psi0, int64_t T, double amp, double *loss, double *dpsi0,
                    double *deps, double *dq0) {
  if (cfg->bc != SMX_BC_PERIODIC || cfg->advect || cfg->w_mid != 0.0 || T < 1) return SMX_ERR_UNSUPPORTED;
  or_state *s = or_create(cfg);
  if (!s) return SMX_ERR_INVALID;
  const int64_t N = s->N;
  double *traj = (double *)malloc(sizeof(double) * N * (size_t)(T + 1));
  or_set_psi(s, psi0);
  memcpy(traj, s->psi, sizeof(double) * N);
  for (int64_t n = 0; n < T; ++n) {
    or_step(s, 1);
    memcpy(traj + (n + 1) * N, s->psi, sizeof(double) * N);
  }
  double *a1 = (double *)calloc(N, sizeof(double)), *a2 = (double *)calloc(N, sizeof(double));
  double *w = (double *)malloc(sizeof(double) * N), *t = (double *)malloc(sizeof(double) * N);
  /* a_T and the loss */
  double J = 0.0, tq = 0.0;
  const double *pT = traj + T * N;
  for (int i = 0; i < s->nx; ++i)
    for (int j = 0; j < s->ny; ++j)
      for (int k = 0; k < s->nz; ++k) {
        int64_t id = ((int64_t)i * s->ny + j) * s->nz + k;
        double z = k * s->dz, g = amp * cos(s->c.q0 * z);
        a1[id] = pT[id] - g;
        J += 0.5 * a1[id] * a1[id];
        tq += a1[id] * amp * z * sin(s->c.q0 * z); /* -dg/dq0 */
      }
  /* parameter terms: < a_{n+1}, F^-1[dt/2 dL (Psi_n + Psi_{n+1})/C2] > = N sum_k w Re(conj alpha_{n+1} ...) */
  double ge = 0.0, gq = 0.0;
  cplx *qa = (cplx *)malloc(sizeof(cplx) * s->Nh), *qp = (cplx *)malloc(sizeof(cplx) * s->Nh);
  double *sum = (double *)malloc(sizeof(double) * N);
  const double dtd2 = 0.5 * s->c.dt, q02 = s->c.q0 * s->c.q0;
  for (int64_t n = T - 1; n >= 0; --n) {
    /* a1 = a_{n+1}, a2 = a_{n+2} */
    for (int64_t x = 0; x < N; ++x) sum[x] = traj[n * N + x] + traj[(n + 1) * N + x];
    or_rfft3(s->nx, s->ny, s->nz, a1, (double *)qa);
    or_rfft3(s->nx, s->ny, s->nz, sum, (double *)qp);
    for (int i = 0; i < s->nx; ++i)
      for (int j = 0; j < s->ny; ++j)
        for (int k = 0; k < s->nzh; ++k) {
          int64_t id = KIDX(i, j, k);
          double k2 = s->vqx[i] * s->vqx[i] + s->vqy[j] * s->vqy[j] + s->vqz[k] * s->vqz[k];
          double L = s->c.eps - op_sh(s, k2), C2 = 1.0 - dtd2 * L;
          double wk = (k == 0 || 2 * k == s->nz) ? 1.0 : 2.0;
          /* normalised spectra: alpha = s qa, Psi = s qp; inner product N sum w Re(conj alpha Psi) */
          double re = creal(conj(qa[id]) * qp[id]) * s->scale * s->scale * (double)N;
          double base = wk * re * dtd2 / C2;
          ge += base;
          gq += base * 4.0 * s->c.alpha * s->c.q0 * (k2 - q02);
        }
    const double c = (n >= 1) ? 3.0 : 2.0;
    for (int64_t x = 0; x < N; ++x) t[x] = c * a1[x] - a2[x];
    apply_mult(s, t, w, 1);
    for (int64_t x = 0; x < N; ++x) w[x] *= nl_prime(s, traj[n * N + x]);
    apply_mult(s, a1, t, 0);
    for (int64_t x = 0; x < N; ++x) t[x] += w[x];
    /* shift: a2 <- a1, a1 <- a_n */
    double *tmp = a2;
    a2 = a1;
    a1 = tmp;
    memcpy(a1, t, sizeof(double) * N);
  }
  memcpy(dpsi0, a1, sizeof(double) * N);
  *loss = J;
  *deps = ge;
  *dq0 = gq + tq;
  free(traj); free(a1); free(a2); free(w); free(t); free(qa); free(qp); free(sum);
  or_destroy(s);
  return 0;
}
