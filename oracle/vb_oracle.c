/*
 * vb_oracle.c -- CPU FP64 restatement of cudaBubble's pseudo-spectral step.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Only the tests, the smoke check
 * and the benchmark's CPU-baseline leg load this library.  The product never
 * links or calls it.
 *
 * Compiled with -ffp-contract=off so that every expression is evaluated in
 * the order written, which is the order of the reference expression it
 * restates (cited per function).  OpenMP parallelises the batched 1-D FFTs
 * and the pointwise loops; reductions use pairwise summation so that the
 * checker's integrals are accurate to a few ulps.
 */
#include "vb_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* tool.h:7 */
#define VB_PI 3.1415926535897932384626433832795

typedef struct { double re, im; } cpx;

static inline cpx cmul(cpx a, cpx b) {
  cpx r; r.re = a.re * b.re - a.im * b.im; r.im = a.re * b.im + a.im * b.re; return r;
}
static inline cpx cadd(cpx a, cpx b) { cpx r; r.re = a.re + b.re; r.im = a.im + b.im; return r; }

/* ------------------------------------------------------------------------ */
/* 1-D mixed-radix Stockham FFT (own implementation).                       */
/* ------------------------------------------------------------------------ */
typedef struct {
  int n, nf;
  int fac[40];
  cpx* tw;   /* tw[m] = exp(-2 pi i m / n), m in [0, n) */
} fftplan;

static void plan_init(fftplan* p, int n) {
  p->n = n; p->nf = 0;
  int m = n;
  while (m % 4 == 0) { p->fac[p->nf++] = 4; m /= 4; }
  while (m % 2 == 0) { p->fac[p->nf++] = 2; m /= 2; }
  for (int q = 3; m > 1; q += 2)
    while (m % q == 0) { p->fac[p->nf++] = q; m /= q; }
  p->tw = (cpx*)malloc(sizeof(cpx) * (size_t)n);
  for (int k = 0; k < n; k++) {
    /* long double angle, rounded once to double */
    long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)k / (long double)n;
    p->tw[k].re = (double)cosl(a);
    p->tw[k].im = (double)sinl(a);
  }
}

static void plan_free(fftplan* p) { free(p->tw); p->tw = NULL; }

static inline cpx twid(const fftplan* p, int m, int sign) {
  cpx w = p->tw[m];
  if (sign > 0) w.im = -w.im;
  return w;
}

/* x: length n, work: length n.  sign -1 forward, +1 inverse, unscaled. */
static void fft_exec(const fftplan* p, cpx* x, cpx* work, int sign) {
  const int n = p->n;
  if (n == 1) return;
  cpx* src = x;
  cpx* dst = work;
  int Ns = 1;
  cpx v[64], y[64];
  for (int f = 0; f < p->nf; f++) {
    const int R = p->fac[f];
    const int m = n / R;
    const int tstep = n / (Ns * R);
    for (int j = 0; j < m; j++) {
      const int k = j % Ns;
      for (int r = 0; r < R; r++) {
        v[r] = src[j + r * m];
        if (k) v[r] = cmul(v[r], twid(p, (r * k * tstep) % n, sign));
      }
      if (R == 2) {
        y[0] = cadd(v[0], v[1]);
        y[1].re = v[0].re - v[1].re; y[1].im = v[0].im - v[1].im;
      } else if (R == 4) {
        cpx a = cadd(v[0], v[2]);
        cpx b; b.re = v[0].re - v[2].re; b.im = v[0].im - v[2].im;
        cpx c = cadd(v[1], v[3]);
        cpx d; d.re = v[1].re - v[3].re; d.im = v[1].im - v[3].im;
        /* rotate d by sign*i */
        cpx di; if (sign < 0) { di.re = d.im; di.im = -d.re; } else { di.re = -d.im; di.im = d.re; }
        y[0] = cadd(a, c);
        y[2].re = a.re - c.re; y[2].im = a.im - c.im;
        y[1] = cadd(b, di);
        y[3].re = b.re - di.re; y[3].im = b.im - di.im;
      } else {
        /* generic radix-R DFT with roots exp(sign 2 pi i r q / R) = tw[(r q mod R) n/R] */
        for (int q = 0; q < R; q++) {
          cpx acc = v[0];
          for (int r = 1; r < R; r++)
            acc = cadd(acc, cmul(v[r], twid(p, ((r * q) % R) * m, sign)));
          y[q] = acc;
        }
      }
      const int base = (j / Ns) * Ns * R + k;
      for (int q = 0; q < R; q++) dst[base + q * Ns] = y[q];
    }
    cpx* t = src; src = dst; dst = t;
    Ns *= R;
  }
  if (src != x) memcpy(x, src, sizeof(cpx) * (size_t)n);
}

void vbo_fft1d(int n, double* x, int sign) {
  fftplan p; plan_init(&p, n);
  cpx* w = (cpx*)malloc(sizeof(cpx) * (size_t)n);
  fft_exec(&p, (cpx*)x, w, sign);
  free(w);
  plan_free(&p);
}

/* ------------------------------------------------------------------------ */
/* context                                                                  */
/* ------------------------------------------------------------------------ */
struct vbo_ctx {
  vbo_params p;
  int nh;
  size_t nr, nc;          /* real points, half-spectrum points */
  fftplan px, py, pz;
  /* scratch: real fields and spectra */
  double *r1, *r2, *r3, *r4, *r5, *r6, *r7;
  double *c1, *c2, *c3, *c4;
};

vbo_ctx* vbo_new(const vbo_params* p) {
  if (p->nx <= 0 || p->nz <= 0 || (p->nx % 2) != 0) return NULL;
  vbo_ctx* c = (vbo_ctx*)calloc(1, sizeof(vbo_ctx));
  c->p = *p;
  c->p.ny = p->nx;  /* Calc.cu:46 forces ny = nx */
  c->nh = p->nx / 2 + 1;
  c->nr = (size_t)p->nx * p->nx * p->nz;
  c->nc = (size_t)c->nh * p->nx * p->nz;
  plan_init(&c->px, p->nx);
  plan_init(&c->py, p->nx);
  plan_init(&c->pz, p->nz);
  double** rr[] = {&c->r1, &c->r2, &c->r3, &c->r4, &c->r5, &c->r6, &c->r7};
  for (int i = 0; i < 7; i++) *rr[i] = (double*)calloc(c->nr, sizeof(double));
  double** cc[] = {&c->c1, &c->c2, &c->c3, &c->c4};
  for (int i = 0; i < 4; i++) *cc[i] = (double*)calloc(2 * c->nc, sizeof(double));
  return c;
}

void vbo_free(vbo_ctx* c) {
  if (!c) return;
  plan_free(&c->px); plan_free(&c->py); plan_free(&c->pz);
  free(c->r1); free(c->r2); free(c->r3); free(c->r4); free(c->r5); free(c->r6); free(c->r7);
  free(c->c1); free(c->c2); free(c->c3); free(c->c4);
  free(c);
}

void vbo_set_targets(vbo_ctx* c, double alpha, double beta) { c->p.alpha = alpha; c->p.beta = beta; }
void vbo_get_params(const vbo_ctx* c, vbo_params* out) { *out = c->p; }

int vbo_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* threads of the following parallel regions (n >= 1); the results do not
 * depend on it (fixed reduction splits) */
void vbo_set_threads(int n) {
#ifdef _OPENMP
  if (n >= 1) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/* ------------------------------------------------------------------------ */
/* 3-D transforms.  Forward: D2Z then scale by 1/(nx ny nz) (Calc.cu:88-96). */
/* Backward: Z2D, unnormalised (Calc.cu:99-104).                            */
/* ------------------------------------------------------------------------ */
static void fft_y_z(vbo_ctx* c, cpx* s, int sign) {
  const int nx = c->p.nx, ny = c->p.ny, nz = c->p.nz, nh = c->nh;
  /* y axis: columns (z, kx), stride nh */
#pragma omp parallel
  {
    cpx* col = (cpx*)malloc(sizeof(cpx) * (size_t)(ny > nz ? ny : nz));
    cpx* wk = (cpx*)malloc(sizeof(cpx) * (size_t)(ny > nz ? ny : nz));
#pragma omp for schedule(static)
    for (int zk = 0; zk < nz * nh; zk++) {
      int z = zk / nh, k = zk % nh;
      cpx* base = s + (size_t)z * ny * nh + k;
      for (int y = 0; y < ny; y++) col[y] = base[(size_t)y * nh];
      fft_exec(&c->py, col, wk, sign);
      for (int y = 0; y < ny; y++) base[(size_t)y * nh] = col[y];
    }
    /* z axis: columns (y, kx), stride ny*nh */
#pragma omp for schedule(static)
    for (int yk = 0; yk < ny * nh; yk++) {
      cpx* base = s + yk;
      for (int z = 0; z < nz; z++) col[z] = base[(size_t)z * ny * nh];
      fft_exec(&c->pz, col, wk, sign);
      for (int z = 0; z < nz; z++) base[(size_t)z * ny * nh] = col[z];
    }
    free(col); free(wk);
  }
  (void)nx;
}

void vbo_r2c(vbo_ctx* c, const double* u, double* sd) {
  const int nx = c->p.nx, ny = c->p.ny, nz = c->p.nz, nh = c->nh;
  cpx* s = (cpx*)sd;
#pragma omp parallel
  {
    cpx* row = (cpx*)malloc(sizeof(cpx) * (size_t)nx);
    cpx* wk = (cpx*)malloc(sizeof(cpx) * (size_t)nx);
#pragma omp for schedule(static)
    for (int zy = 0; zy < nz * ny; zy++) {
      const double* in = u + (size_t)zy * nx;
      for (int x = 0; x < nx; x++) { row[x].re = in[x]; row[x].im = 0.0; }
      fft_exec(&c->px, row, wk, -1);
      cpx* out = s + (size_t)zy * nh;
      for (int k = 0; k < nh; k++) out[k] = row[k];
    }
    free(row); free(wk);
  }
  fft_y_z(c, s, -1);
  /* normalize_kernel (Calc.cu:77-85) with factor 1.0/(nx*ny*nz) (Calc.cu:94) */
  const double factor = 1.0 / (nx * ny * nz);
#pragma omp parallel for schedule(static)
  for (long i = 0; i < (long)c->nc; i++) { s[i].re *= factor; s[i].im *= factor; }
}

void vbo_c2r(vbo_ctx* c, const double* sd, double* u) {
  const int nx = c->p.nx, ny = c->p.ny, nz = c->p.nz, nh = c->nh;
  cpx* s = (cpx*)c->c4;
  memcpy(s, sd, sizeof(cpx) * c->nc);
  fft_y_z(c, s, +1);
#pragma omp parallel
  {
    cpx* row = (cpx*)malloc(sizeof(cpx) * (size_t)nx);
    cpx* wk = (cpx*)malloc(sizeof(cpx) * (size_t)nx);
#pragma omp for schedule(static)
    for (int zy = 0; zy < nz * ny; zy++) {
      const cpx* in = s + (size_t)zy * nh;
      for (int k = 0; k < nh; k++) row[k] = in[k];
      /* Hermitian completion along x; the imaginary parts of kx = 0 and
       * kx = nx/2 only reach the imaginary part of the output, which a
       * real-output inverse discards. */
      for (int k = nh; k < nx; k++) { row[k].re = in[nx - k].re; row[k].im = -in[nx - k].im; }
      fft_exec(&c->px, row, wk, +1);
      double* out = u + (size_t)zy * nx;
      for (int x = 0; x < nx; x++) out[x] = row[x].re;
    }
    free(row); free(wk);
  }
}

/* ------------------------------------------------------------------------ */
/* Calc kernels                                                             */
/* ------------------------------------------------------------------------ */

/* laplace_kernel (Calc.cu:107-124), coefficients from Calc::laplace (Calc.cu:127-137) */
void vbo_laplace(vbo_ctx* c, double* sd) {
  const int nx = c->p.nx, ny = c->p.ny, nz = c->p.nz, nh = c->nh;
  double cx = 2.0 * VB_PI / c->p.xlen, cy = 2.0 * VB_PI / c->p.xlen, cz = 2.0 * VB_PI / c->p.xlen;
  cx *= cx; cy *= cy; cz *= cz;
  cpx* u = (cpx*)sd;
#pragma omp parallel for schedule(static)
  for (long tid = 0; tid < (long)c->nc; tid++) {
    int z = (int)(tid / (ny * nh));
    int y = (int)((tid % (ny * nh)) / nh);
    int x = (int)((tid % (ny * nh)) % nh);
    z = z < nz - z ? z : nz - z;
    y = y < ny - y ? y : ny - y;
    x = x < nx - x ? x : nx - x;
    double tmp = -(x * x * cx + y * y * cy + z * z * cz);
    u[tid].re *= tmp;
    u[tid].im *= tmp;
  }
}

/* gradient_kernel (Calc.cu:300-341) with fac = 2 pi / len (Calc.cu:346) */
void vbo_gradient(vbo_ctx* c, const double* sd, double* s1, double* s2, double* s3) {
  const int nx = c->p.nx, ny = c->p.ny, nz = c->p.nz, nh = c->nh;
  const double fac1 = 2.0 * VB_PI / c->p.xlen, fac2 = fac1, fac3 = fac1;
  const cpx* u = (const cpx*)sd;
  cpx* u1 = (cpx*)s1; cpx* u2 = (cpx*)s2; cpx* u3 = (cpx*)s3;
#pragma omp parallel for schedule(static)
  for (long tid = 0; tid < (long)c->nc; tid++) {
    int x = (int)(tid % nh);
    if (x == nx / 2) { u1[tid].re = u1[tid].im = 0; }
    else {
      double cx = u[tid].re * fac1, cy = u[tid].im * fac1;
      u1[tid].re = -cy * x; u1[tid].im = cx * x;
    }
    int y = (int)((tid % (ny * nh)) / nh);
    if (y == ny / 2) { u2[tid].re = u2[tid].im = 0; }
    else {
      y = y < ny - y ? y : y - ny;
      double cx = u[tid].re * fac2, cy = u[tid].im * fac2;
      u2[tid].re = -cy * y; u2[tid].im = cx * y;
    }
    int z = (int)(tid / (ny * nh));
    if (z == nz / 2) { u3[tid].re = u3[tid].im = 0; }
    else {
      z = z < nz - z ? z : z - nz;
      double cx = u[tid].re * fac3, cy = u[tid].im * fac3;
      u3[tid].re = -cy * z; u3[tid].im = cx * z;
    }
  }
}

/* Calc::lap(fd, out) (Calc.cu:139-144): copy, laplace, C2R */
void vbo_lap_spec(vbo_ctx* c, const double* fd, double* out) {
  memcpy(c->c1, fd, sizeof(cpx) * c->nc);
  vbo_laplace(c, c->c1);
  vbo_c2r(c, c->c1, out);
}

/* Calc::lap(d, out) (Calc.cu:146-151): R2C, laplace, C2R */
void vbo_lap_real(vbo_ctx* c, const double* d, double* out) {
  vbo_r2c(c, d, c->c1);
  vbo_laplace(c, c->c1);
  vbo_c2r(c, c->c1, out);
}

/* Calc::fftgradient (Calc.cu:344-356) */
void vbo_fftgradient(vbo_ctx* c, const double* fd, double* dx, double* dy, double* dz) {
  vbo_gradient(c, fd, c->c1, c->c2, c->c3);
  vbo_c2r(c, c->c1, dx);
  vbo_c2r(c, c->c2, dy);
  vbo_c2r(c, c->c3, dz);
}

/* f_kernel (Calc.cu:153-162) */
void vbo_f(vbo_ctx* c, double* f, const double* lap, const double* d) {
  const double eps = c->p.eps;
#pragma omp parallel for schedule(static)
  for (long i = 0; i < (long)c->nr; i++)
    f[i] = -lap[i] * eps + (d[i] * d[i] - 1.0) * d[i] / eps;
}

/* g0_kernel (Calc.cu:171-180) */
void vbo_g0k(vbo_ctx* c, double* g0, const double* d, const double* f) {
  const double eps = c->p.eps;
#pragma omp parallel for schedule(static)
  for (long i = 0; i < (long)c->nr; i++)
    g0[i] = -g0[i] + (3.0 * d[i] * d[i] - 1.0) * f[i] / (eps * eps);
}

/* pairwise sum: the checker's accurate stand-in for the reference's
 * grid-stride + tree reduction (Calc.cu:191-252) */
static double psum(const double* u, long n) {
  if (n <= 64) {
    double s = 0.0;
    for (long i = 0; i < n; i++) s += u[i];
    return s;
  }
  long h = n / 2;
  return psum(u, h) + psum(u + h, n - h);
}

static double psum_par(const double* u, long n) {
  /* fixed split into 64 chunks so the result does not depend on threads */
  double part[64];
  const long chunk = (n + 63) / 64;
#pragma omp parallel for schedule(static)
  for (int b = 0; b < 64; b++) {
    long lo = b * chunk, hi = lo + chunk;
    if (hi > n) hi = n;
    part[b] = lo < hi ? psum(u + lo, hi - lo) : 0.0;
  }
  return psum(part, 64);
}

/* Calc::intf (Calc.cu:255-263): (sum(u)/N + addition) * multiply * L^3 */
double vbo_intf(vbo_ctx* c, const double* u, double addition, double multiply) {
  const double N = (double)(long)c->nr;
  double S = psum_par(u, (long)c->nr) / N;
  double L = c->p.xlen;
  return (S + addition) * (multiply * L * L * L);
}

/* nld_kernel1 / Calc::lap / nld_kernel2 (Calc.cu:265-297), kappa and D from
 * phasefield::nld (phasefield.cpp:259-260) */
void vbo_nld(vbo_ctx* c, double* nldp, const double* d, const double* f) {
  const double kappa = c->p.kappa, D = c->p.D;
  double* tmp1 = c->r7;
#pragma omp parallel for schedule(static)
  for (long i = 0; i < (long)c->nr; i++) tmp1[i] = d[i] * (d[i] * d[i] - 1.0 - kappa);
  vbo_lap_real(c, tmp1, nldp);
  double eps = 1 / c->p.eps;
  double eps2 = eps * eps;
  double eps3 = eps * eps2;
#pragma omp parallel for schedule(static)
  for (long i = 0; i < (long)c->nr; i++)
    nldp[i] = -nldp[i] * eps + D * tmp1[i] * eps3 + (3 * d[i] * d[i] - D - 1) * eps2 * f[i];
}

/* s_term_kernel (Calc.cu:358-368) */
void vbo_s_term(vbo_ctx* c, double* s, const double* d, const double* dx,
                const double* dy, const double* dz) {
  const double eps = c->p.eps;
#pragma omp parallel for schedule(static)
  for (long i = 0; i < (long)c->nr; i++)
    s[i] = 0.5 * eps * (dx[i] * dx[i] + dy[i] * dy[i] + dz[i] * dz[i])
         + 0.25 / eps * (d[i] * d[i] - 1) * (d[i] * d[i] - 1);
}

/* nld1_kernel (Calc.cu:410-419) */
void vbo_nld1(vbo_ctx* c, double* nld, const double* f, const double* g, double l1, double l2) {
  const double eta = c->p.eta;
#pragma omp parallel for schedule(static)
  for (long i = 0; i < (long)c->nr; i++) nld[i] = l1 * 0.5 + l2 * f[i] + g[i] * eta;
}

/* nld2_kernel (Calc.cu:428-437) */
void vbo_nld2(vbo_ctx* c, double* nld, const double* f, const double* g, double l1, double l2) {
  const double eta = c->p.eta;
#pragma omp parallel for schedule(static)
  for (long i = 0; i < (long)c->nr; i++)
    nld[i] = -l1 * 0.5 - l2 * f[i] * 3.0 / (2.0 * sqrt(2.0)) - g[i] * eta;
}

/* Calc::init_fEIF (Calc.cu:579-606), kappa = 2, D = 1.5 hard-coded there */
void vbo_init_fEIF(vbo_ctx* c, double* Het) {
  const int nx = c->p.nx, ny = c->p.ny, nz = c->p.nz;
  const double eps = c->p.eps, kappa = c->p.kappa, D = c->p.D;
  double cx = 2.0 * VB_PI / c->p.xlen, cy = cx, cz = cx;
  cx *= cx; cy *= cy; cz *= cz;
  const int nh = nx / 2 + 1;
  for (int i = 0; i < nz; i++) {
    int i2 = i < nz - i ? i : nz - i;
    for (int j = 0; j < ny; j++) {
      int j2 = j < ny - j ? j : ny - j;
      for (int k = 0; k < nh; k++) {
        int k2 = k < nx - k ? k : nx - k;
        double cc = -(i2 * i2 * cz + j2 * j2 * cy + k2 * k2 * cx);
        Het[((size_t)i * ny + j) * nh + k] = -eps * cc * cc + kappa * cc / eps + D * (cc - kappa / (eps * eps)) / eps;
      }
    }
  }
}

/* etd1_kernel (Calc.cu:446-457) */
void vbo_etd1(vbo_ctx* c, double* ad, const double* fdd, const double* nldd, const double* Het, double h) {
  cpx* a = (cpx*)ad; const cpx* fd = (const cpx*)fdd; const cpx* nld = (const cpx*)nldd;
#pragma omp parallel for schedule(static)
  for (long t = 0; t < (long)c->nc; t++) {
    double cc = exp(Het[t] * h);
    a[t].re = cc * fd[t].re + (cc - 1.0) * nld[t].re / Het[t];
    a[t].im = cc * fd[t].im + (cc - 1.0) * nld[t].im / Het[t];
  }
}

/* etd2_kernel (Calc.cu:466-479) */
void vbo_etd2(vbo_ctx* c, double* ad, const double* fdd, const double* nnldd, const double* nldd,
              const double* Het, double h) {
  cpx* a = (cpx*)ad; const cpx* fd = (const cpx*)fdd;
  const cpx* nnld = (const cpx*)nnldd; const cpx* nld = (const cpx*)nldd;
#pragma omp parallel for schedule(static)
  for (long t = 0; t < (long)c->nc; t++) {
    double cc = exp(Het[t] * h);
    if (h * Het[t] == 0) cc = h / 2;
    else cc = (cc - 1.0 - h * Het[t]) / (h * Het[t] * Het[t]);
    a[t].re = fd[t].re + cc * (nnld[t].re - nld[t].re);
    a[t].im = fd[t].im + cc * (nnld[t].im - nld[t].im);
  }
}

/* etd41_kernel (Calc.cu:488-499) */
void vbo_etd41(vbo_ctx* c, double* ad, const double* fdd, const double* nldd, const double* Het, double h) {
  cpx* a = (cpx*)ad; const cpx* fd = (const cpx*)fdd; const cpx* nld = (const cpx*)nldd;
#pragma omp parallel for schedule(static)
  for (long t = 0; t < (long)c->nc; t++) {
    double e1 = exp(Het[t] * h / 2);
    a[t].re = e1 * fd[t].re + (e1 - 1.0) * nld[t].re / Het[t];
    a[t].im = e1 * fd[t].im + (e1 - 1.0) * nld[t].im / Het[t];
  }
}

/* etd42_kernel (Calc.cu:509-520) */
void vbo_etd42(vbo_ctx* c, double* bd, const double* fdd, const double* nldad, const double* Het, double h) {
  cpx* b = (cpx*)bd; const cpx* fd = (const cpx*)fdd; const cpx* nlda = (const cpx*)nldad;
#pragma omp parallel for schedule(static)
  for (long t = 0; t < (long)c->nc; t++) {
    double e1 = exp(Het[t] * h / 2);
    b[t].re = e1 * fd[t].re + (e1 - 1.0) * nlda[t].re / Het[t];
    b[t].im = e1 * fd[t].im + (e1 - 1.0) * nlda[t].im / Het[t];
  }
}

/* etd43_kernel (Calc.cu:529-540); cc is updated in place */
void vbo_etd43(vbo_ctx* c, double* ccd, const double* nldd, const double* nldbd, const double* Het, double h) {
  cpx* cc = (cpx*)ccd; const cpx* nld = (const cpx*)nldd; const cpx* nldb = (const cpx*)nldbd;
#pragma omp parallel for schedule(static)
  for (long t = 0; t < (long)c->nc; t++) {
    double e1 = exp(Het[t] * h / 2);
    cc[t].re = cc[t].re * e1 + (e1 - 1.0) * (2 * nldb[t].re - nld[t].re) / Het[t];
    cc[t].im = cc[t].im * e1 + (e1 - 1.0) * (2 * nldb[t].im - nld[t].im) / Het[t];
  }
}

/* etd44_kernel (Calc.cu:549-569): the first bracket term uses N(a) */
void vbo_etd44(vbo_ctx* c, double* dd, const double* fdd, const double* nldad, const double* nldbd,
               const double* nldcd, const double* Het, double h) {
  cpx* d = (cpx*)dd; const cpx* fd = (const cpx*)fdd;
  const cpx* nlda = (const cpx*)nldad; const cpx* nldb = (const cpx*)nldbd; const cpx* nldc = (const cpx*)nldcd;
#pragma omp parallel for schedule(static)
  for (long t = 0; t < (long)c->nc; t++) {
    const double H = Het[t];
    double e1 = exp(H * h / 2);
    double e2 = e1 * e1;
    d[t].re = e2 * fd[t].re + 1.0 / (h * h * H * H * H) * (
        nlda[t].re * (-4.0 - h * H + e2 * (4.0 - 3.0 * h * H + h * h * H * H)) +
        2 * (nlda[t].re + nldb[t].re) * (2.0 + h * H + e2 * (-2 + h * H)) +
        nldc[t].re * (-4 - 3 * h * H - h * h * H * H + e2 * (4 - h * H)));
    d[t].im = e2 * fd[t].im + 1.0 / (h * h * H * H * H) * (
        nlda[t].im * (-4.0 - h * H + e2 * (4.0 - 3.0 * h * H + h * h * H * H)) +
        2 * (nlda[t].im + nldb[t].im) * (2.0 + h * H + e2 * (-2 + h * H)) +
        nldc[t].im * (-4 - 3 * h * H - h * h * H * H + e2 * (4 - h * H)));
  }
}

void vbo_etd_coeffs(double H, double h, double* out) {
  double e1 = exp(H * h / 2);
  double e2 = e1 * e1;
  out[0] = e1;
  out[1] = (e1 - 1.0) / H;
  out[2] = e2;
  out[3] = -4.0 - h * H + e2 * (4.0 - 3.0 * h * H + h * h * H * H);
  out[4] = 2.0 + h * H + e2 * (-2 + h * H);
  out[5] = -4 - 3 * h * H - h * h * H * H + e2 * (4 - h * H);
  double cc = exp(H * h);
  if (h * H == 0) cc = h / 2;
  else cc = (cc - 1.0 - h * H) / (h * H * H);
  out[6] = cc;
}

/* ------------------------------------------------------------------------ */
/* phasefield compositions                                                  */
/* ------------------------------------------------------------------------ */

/* phasefield::vol (phasefield.cpp:283-289) */
double vbo_vol(vbo_ctx* c, const double* d) { return vbo_intf(c, d, 1.0, 0.5); }

/* phasefield::surface -> s -> gradient -> fd (phasefield.cpp:270-316) */
static double surface_with_fd(vbo_ctx* c, const double* d, double* fd_out) {
  double *dx = c->r1, *dy = c->r2, *dz = c->r3, *s = c->r4;
  vbo_r2c(c, d, fd_out);
  vbo_fftgradient(c, fd_out, dx, dy, dz);
  vbo_s_term(c, s, d, dx, dy, dz);
  return vbo_intf(c, s, 0.0, 3.0 / (2.0 * sqrt(2.0)));
}

double vbo_surface(vbo_ctx* c, const double* d) {
  double* fd = (double*)malloc(sizeof(cpx) * c->nc);
  double A = surface_with_fd(c, d, fd);
  free(fd);
  return A;
}

/* phasefield::f via lap (phasefield.cpp:140-181); fd given */
static void f_from_fd(vbo_ctx* c, const double* d, const double* fd, double* f) {
  double* lap = c->r5;
  memcpy(c->c2, fd, sizeof(cpx) * c->nc);  /* copy_array(cptmp1, m_fd) (phasefield.cpp:146) */
  vbo_lap_spec(c, c->c2, lap);
  vbo_f(c, f, lap, d);
}

void vbo_fpf(vbo_ctx* c, const double* d, double* f) {
  double* fd = (double*)malloc(sizeof(cpx) * c->nc);
  vbo_r2c(c, d, fd);
  f_from_fd(c, d, fd, f);
  free(fd);
}

/* phasefield::elastic_eng (phasefield.cpp:318-325) */
double vbo_elastic_eng(vbo_ctx* c, const double* d) {
  double* f = (double*)malloc(sizeof(double) * c->nr);
  vbo_fpf(c, d, f);
  double* t = c->r6;
#pragma omp parallel for schedule(static)
  for (long i = 0; i < (long)c->nr; i++) t[i] = f[i] * f[i];   /* multi_kernel */
  double e = vbo_intf(c, t, 0.0, 1.0 / (c->p.eps * 2));
  free(f);
  return e;
}

/* phasefield::g0 (phasefield.cpp:205-221) */
void vbo_g0(vbo_ctx* c, const double* d, double* g0) {
  double* f = (double*)malloc(sizeof(double) * c->nr);
  vbo_fpf(c, d, f);
  vbo_lap_real(c, f, g0);
  vbo_g0k(c, g0, d, f);
  free(f);
}

/* belastic::getstaticenergy (belastic.cpp:68-86) */
double vbo_energy(vbo_ctx* c, const double* d, int constrained, double* Vo, double* Ao, double* Ebo) {
  double V = vbo_vol(c, d);
  double A = vbo_surface(c, d);
  double eng1 = vbo_elastic_eng(c, d);
  if (Vo) *Vo = V;
  if (Ao) *Ao = A;
  if (Ebo) *Ebo = eng1;
  eng1 *= c->p.eta;
  if (constrained) {
    double volerr = V - c->p.alpha, surerr = A - c->p.beta;
    return eng1 + 0.5 * c->p.M1 * volerr * volerr + 0.5 * c->p.M2 * surerr * surerr;
  }
  return eng1;
}

/* belastic::get_nld_term (belastic.cpp:473-488) on a field d; writes N^ and,
 * optionally, F(d) as the phasefield cache would hold it. */
static void nld_term_fd(vbo_ctx* c, const double* d, double* nhat, double* fd, double* Vo, double* Ao) {
  double V = vbo_vol(c, d);
  double A = surface_with_fd(c, d, fd);
  double lamb1 = c->p.M1 * (V - c->p.alpha);
  double lamb2 = c->p.M2 * (A - c->p.beta);
  double* f = c->r6;
  f_from_fd(c, d, fd, f);
  double* nl = c->r5;
  vbo_nld(c, nl, d, f);
  double* tmp = c->r4;
  vbo_nld2(c, tmp, f, nl, lamb1, lamb2);
  vbo_r2c(c, tmp, nhat);
  if (Vo) *Vo = V;
  if (Ao) *Ao = A;
}

void vbo_nld_term(vbo_ctx* c, const double* d, double* nhat, double* V, double* A) {
  double* fd = (double*)malloc(sizeof(cpx) * c->nc);
  nld_term_fd(c, d, nhat, fd, V, A);
  free(fd);
}

/* belastic::get_sum_fgqterm (belastic.cpp:52-64) with g = g0 (no C field) */
void vbo_sum_fgq(vbo_ctx* c, const double* d, double* rhs) {
  double V = vbo_vol(c, d);
  double A = vbo_surface(c, d);
  double lamb1 = c->p.M1 * (V - c->p.alpha);
  double lamb2 = c->p.M2 * (A - c->p.beta);
  double* f = (double*)malloc(sizeof(double) * c->nr);
  double* g = (double*)malloc(sizeof(double) * c->nr);
  vbo_fpf(c, d, f);
  vbo_g0(c, d, g);
  vbo_nld1(c, rhs, f, g, lamb1, lamb2);
  free(f); free(g);
}

/* ------------------------------------------------------------------------ */
/* steppers                                                                 */
/* ------------------------------------------------------------------------ */

/* belastic::ETD4RK (belastic.cpp:519-552) */
void vbo_etd4rk(vbo_ctx* c, double* d, double h) {
  const size_t cb = sizeof(cpx) * c->nc;
  double* Het = (double*)malloc(sizeof(double) * c->nc);
  vbo_init_fEIF(c, Het);
  double* fu = (double*)malloc(cb);
  double* fa = (double*)malloc(cb);
  double* nld = (double*)malloc(cb);
  double* nlda = (double*)malloc(cb);
  double* nldb = (double*)malloc(cb);
  double* nldc = (double*)malloc(cb);
  double* sp = (double*)malloc(cb);
  double* a = (double*)malloc(sizeof(double) * c->nr);

  nld_term_fd(c, d, nld, fu, NULL, NULL);         /* get_nld_term(m_pd, nld) */
  vbo_etd41(c, sp, fu, nld, Het, h);              /* etd41 -> ctmp1          */
  vbo_c2r(c, sp, a);                              /* -> m_pa->d()            */
  nld_term_fd(c, a, nlda, fa, NULL, NULL);        /* get_nld_term(m_pa,nlda) */
  vbo_etd42(c, sp, fu, nlda, Het, h);             /* b                       */
  double* cc = (double*)malloc(cb);
  memcpy(cc, fa, cb);                             /* c <- m_pa->fd()         */
  vbo_c2r(c, sp, a);                              /* m_pa->update_fd(b)      */
  nld_term_fd(c, a, nldb, fa, NULL, NULL);        /* get_nld_term(m_pa,nldb) */
  vbo_etd43(c, cc, nld, nldb, Het, h);            /* etd43(c, .., nld, nldb) */
  vbo_c2r(c, cc, a);                              /* m_pa->update_fd(c)      */
  nld_term_fd(c, a, nldc, fa, NULL, NULL);        /* get_nld_term(m_pa,nldc) */
  vbo_etd44(c, sp, fu, nlda, nldb, nldc, Het, h); /* etd44 -> ctmp1          */
  vbo_c2r(c, sp, d);                              /* -> m_pd->d()            */

  free(Het); free(fu); free(fa); free(nld); free(nlda); free(nldb); free(nldc);
  free(sp); free(a); free(cc);
}

/* belastic::ETD2RK (belastic.cpp:491-517) */
void vbo_etd2rk(vbo_ctx* c, double* d, double h) {
  const size_t cb = sizeof(cpx) * c->nc;
  double* Het = (double*)malloc(sizeof(double) * c->nc);
  vbo_init_fEIF(c, Het);
  double* fu = (double*)malloc(cb);
  double* fa = (double*)malloc(cb);
  double* nld = (double*)malloc(cb);
  double* nnld = (double*)malloc(cb);
  double* sp = (double*)malloc(cb);
  double* a = (double*)malloc(sizeof(double) * c->nr);
  nld_term_fd(c, d, nld, fu, NULL, NULL);
  vbo_etd1(c, sp, fu, nld, Het, h);
  vbo_c2r(c, sp, a);                              /* m_pa->update_fd(a)      */
  nld_term_fd(c, a, nnld, fa, NULL, NULL);        /* fd = m_pa->fd()         */
  vbo_etd2(c, sp, fa, nnld, nld, Het, h);
  vbo_c2r(c, sp, d);                              /* m_pd->update_fd(tmp)    */
  free(Het); free(fu); free(fa); free(nld); free(nnld); free(sp); free(a);
}

/* ------------------------------------------------------------------------ */
/* drivers                                                                  */
/* ------------------------------------------------------------------------ */

static int all_finite(const double* d, size_t n) {
  int ok = 1;
#pragma omp parallel for reduction(&& : ok) schedule(static)
  for (long i = 0; i < (long)n; i++) ok = ok && isfinite(d[i]);
  return ok;
}

void vbo_fEIF_begin(vbo_ctx* c, const double* d, vbo_drive* s) {
  /* belastic.cpp:588-593 */
  if (c->p.beta <= 0.0) {
    c->p.alpha = vbo_vol(c, d);
    c->p.beta = vbo_surface(c, d);
  }
  s->alpha0 = c->p.alpha;
  s->beta0 = c->p.beta;
  /* belastic.cpp:599-601 */
  if (s->fixed_time_step > 0)
    s->time_step = s->fixed_time_step < s->time_step ? s->fixed_time_step : s->time_step;
  if (s->dt_double_every <= 0) s->dt_double_every = 4;
  if (s->dt_cap <= 0) s->dt_cap = 0.001;
  /* belastic.cpp:617 */
  s->eng0 = vbo_energy(c, d, 1, NULL, NULL, NULL);
  s->eng1 = s->eng0;
  s->steps = 0;
  s->status = 0;
}

static void one_eval(vbo_ctx* c, double* d, double h, int scheme) {
  if (scheme == 2) vbo_etd2rk(c, d, h);
  else vbo_etd4rk(c, d, h);
}

/* loop body of belastic::run_static_fEIF (belastic.cpp:635-746) */
int vbo_fEIF_steps(vbo_ctx* c, double* d, vbo_drive* s, int n, int scheme, double* trace) {
  double* od = (double*)malloc(sizeof(double) * c->nr);
  int done = 0;
  for (int it = 0; it < n && s->status == 0; it++) {
    if (s->fixed_time_step <= 0) {
      /* adaptive (belastic.cpp:639-669) */
      s->eng0 = vbo_energy(c, d, 1, NULL, NULL, NULL);
      memcpy(od, d, sizeof(double) * c->nr);
      if (s->steps % s->dt_double_every == 0) {
        if (s->time_step * 1.4 < s->dt_cap) s->time_step = s->time_step * 2;
        else s->time_step = s->dt_cap;
      }
      for (;;) {
        if (s->time_step < 1e-23) break;
        one_eval(c, d, s->time_step, scheme);
        s->evaluations++;
        s->eng1 = vbo_energy(c, d, 1, NULL, NULL, NULL);
        if (s->eng1 > s->eng0) {
          memcpy(d, od, sizeof(double) * c->nr);
          s->time_step /= 2;
          s->rejections++;
          continue;
        } else break;
      }
    } else if (c->p.M1 != 0 || c->p.M2 != 0) {
      /* fixed dt with penalty drift correction (belastic.cpp:670-691) */
      memcpy(od, d, sizeof(double) * c->nr);
      one_eval(c, d, s->time_step, scheme);
      s->evaluations++;
      double v = vbo_vol(c, d), sf = vbo_surface(c, d);
      c->p.alpha -= (v - s->alpha0);
      c->p.beta -= (sf - s->beta0);
      memcpy(d, od, sizeof(double) * c->nr);
      one_eval(c, d, s->time_step, scheme);
      s->evaluations++;
      v = vbo_vol(c, d); sf = vbo_surface(c, d);
      c->p.alpha -= (v - s->alpha0);
      c->p.beta -= (sf - s->beta0);
    } else {
      one_eval(c, d, s->time_step, scheme);     /* belastic.cpp:692-693 */
      s->evaluations++;
    }
    /* exits (belastic.cpp:713-740); the blow-up test reads the whole field
     * instead of the reference's stale host copy of the centre point */
    if (s->time_step < 1e-23) { s->status = 1; }
    else if (s->eng0 < 1e-3) { s->status = 2; }
    else if (!all_finite(d, c->nr)) { s->status = 3; }
    if (trace) {
      double V, A;
      double E = vbo_energy(c, d, 1, &V, &A, NULL);
      trace[4 * it + 0] = s->time_step;
      trace[4 * it + 1] = E;
      trace[4 * it + 2] = V;
      trace[4 * it + 3] = A;
    }
    if (s->status == 1 || s->status == 3) break;
    if (s->status == 2) break;
    s->time_current += s->time_step;   /* belastic.cpp:742-746 */
    s->steps++;
    done++;
  }
  free(od);
  return done;
}

/* belastic::shape_adjust (belastic.cpp:174-283).  Returns iterations. */
int vbo_shape_adjust(vbo_ctx* c, double* d, int max_iter) {
  double* od = (double*)malloc(sizeof(double) * c->nr);
  double* g0 = (double*)malloc(sizeof(double) * c->nr);
  double tcur = 0.0;
  int steps = 1;
  double time_step = 1e-7;
  const double eta = c->p.eta;
  double eng0 = vbo_energy(c, d, 0, NULL, NULL, NULL);
  double eng1 = eng0, eng = eng0;
  int it;
  for (it = 0; it < max_iter; it++) {
    tcur += time_step;
    steps++;
    if (time_step > 1e-4 || (int)(tcur / time_step) % (int)(1e-4 / time_step) == 0) {
      eng1 = vbo_energy(c, d, 0, NULL, NULL, NULL);
      if (eng1 < 100 * eta || fabs((eng - eng1) / eng) < 0.02) break;
      eng = eng1;
    }
    if (steps % 10 == 0) time_step *= 2;
    memcpy(od, d, sizeof(double) * c->nr);
    vbo_g0(c, d, g0);
    for (;;) {
      const double t = -time_step * 3;
#pragma omp parallel for schedule(static)
      for (long i = 0; i < (long)c->nr; i++) d[i] += t * g0[i];   /* plus_kernel */
      eng1 = vbo_energy(c, d, 0, NULL, NULL, NULL);
      if (eng1 < eng0) { eng0 = eng1; break; }
      memcpy(d, od, sizeof(double) * c->nr);
      time_step /= 2;
      if (time_step < 1e-13) break;
    }
    if (time_step < 1e-13) break;
    if (!all_finite(d, c->nr)) break;
  }
  free(od); free(g0);
  return it;
}

/* loop body of belastic::run_static (belastic.cpp:367-437), Forward Euler */
int vbo_static_steps(vbo_ctx* c, double* d, vbo_drive* s, int n) {
  double* od = (double*)malloc(sizeof(double) * c->nr);
  double* nld1 = (double*)malloc(sizeof(double) * c->nr);
  const double ga = c->p.ga;
  int done = 0;
  vbo_sum_fgq(c, d, nld1);   /* belastic.cpp:347 */
  for (int it = 0; it < n && s->status == 0; it++) {
    if (s->fixed_time_step <= 0) {
      s->eng0 = vbo_energy(c, d, 1, NULL, NULL, NULL);
      memcpy(od, d, sizeof(double) * c->nr);
      if (s->steps % 4 == 0) s->time_step = s->time_step * 1.4;
      for (;;) {
        if (s->time_step < 1e-23) break;
        const double t = -s->time_step * ga;
#pragma omp parallel for schedule(static)
        for (long i = 0; i < (long)c->nr; i++) d[i] += t * nld1[i];
        s->evaluations++;
        s->eng1 = vbo_energy(c, d, 1, NULL, NULL, NULL);
        if (s->eng1 > s->eng0) {
          memcpy(d, od, sizeof(double) * c->nr);
          s->time_step /= 2;
          s->rejections++;
          continue;
        } else break;
      }
    } else {
      const double t = -s->time_step * ga;
#pragma omp parallel for schedule(static)
      for (long i = 0; i < (long)c->nr; i++) d[i] += t * nld1[i];
      s->evaluations++;
    }
    if (s->time_step < 1e-23) { s->status = 1; break; }
    if (!all_finite(d, c->nr)) { s->status = 3; break; }
    s->time_current += s->time_step;
    s->steps++;
    done++;
    vbo_sum_fgq(c, d, nld1);   /* belastic.cpp:436 */
  }
  free(od); free(nld1);
  return done;
}
