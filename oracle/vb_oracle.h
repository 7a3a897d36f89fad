/*
 * vb_oracle.h -- CPU FP64 restatement of cudaBubble's pseudo-spectral
 * vesicle step (Allen-Cahn phase field, elastic bending energy with volume
 * and area penalties, Forward-Euler / ETDRK2 / ETDRK4 time stepping).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: the tests, the
 * smoke check and the CPU-baseline leg of the benchmark load it to compute
 * expected values.  The product (libvesicle.so) never links, loads or calls
 * it, and nothing here is measured as the product.
 *
 * Every function restates one reference routine formula-for-formula and in
 * the same dataflow order (including the reference's forward/backward FFT
 * round trips).  Citations are to the reference sources:
 *   Calc.cu       -- FFT wrappers, spectral/pointwise kernels, ETD combines
 *   phasefield.cpp-- composition of derived fields and integrals
 *   belastic.cpp  -- nonlinear term, steppers, energy, drivers
 *   fluidbase.cpp -- grid, epsilon table
 *
 * Layouts (reference Calc.h:4, Calc.cu:57-58): real fields are [nz][ny][nx]
 * doubles with x fastest; half spectra are [nz][ny][nx/2+1] complex numbers
 * stored as interleaved (re, im) doubles.
 *
 * The FFT is this file's own mixed-radix Stockham transform (the reference
 * uses cuFFT, which is closed source and not available here).
 */
#ifndef VB_ORACLE_H
#define VB_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vbo_params {
  int nx, ny, nz, pad0;   /* nx == ny, even (Calc.cu:45-55)                  */
  double xlen;            /* box length L (already 2*pi*len), all axes      */
  double eps, eta, ga;    /* interface width, bending rigidity, gamma       */
  double M1, M2;          /* volume / area penalty weights                  */
  double alpha, beta;     /* volume / area targets                          */
  double kappa, D;        /* splitting constants, 2 and 1.5 in the reference*/
} vbo_params;

typedef struct vbo_ctx vbo_ctx;

vbo_ctx* vbo_new(const vbo_params* p);
void vbo_free(vbo_ctx* c);
void vbo_set_targets(vbo_ctx* c, double alpha, double beta);
void vbo_get_params(const vbo_ctx* c, vbo_params* out);
int vbo_threads(void);
void vbo_set_threads(int n);

/* ---- transforms (Calc.cu:77-104) ------------------------------------- */
void vbo_r2c(vbo_ctx* c, const double* u, double* s);   /* normalised   */
void vbo_c2r(vbo_ctx* c, const double* s, double* u);   /* unnormalised */
/* plain 1-D complex DFT of length n (sign -1 forward, +1 inverse), for tests */
void vbo_fft1d(int n, double* x, int sign);

/* ---- Calc kernels ----------------------------------------------------- */
void vbo_laplace(vbo_ctx* c, double* s);                          /* Calc.cu:107-137 */
void vbo_gradient(vbo_ctx* c, const double* s, double* s1, double* s2, double* s3); /* Calc.cu:300-341 */
void vbo_lap_spec(vbo_ctx* c, const double* fd, double* out);     /* Calc.cu:139-144 */
void vbo_lap_real(vbo_ctx* c, const double* d, double* out);      /* Calc.cu:146-151 */
void vbo_fftgradient(vbo_ctx* c, const double* fd, double* dx, double* dy, double* dz); /* Calc.cu:344-356 */
void vbo_f(vbo_ctx* c, double* f, const double* lap, const double* d);        /* Calc.cu:153-169 */
void vbo_g0k(vbo_ctx* c, double* g0, const double* d, const double* f);       /* Calc.cu:171-188 */
double vbo_intf(vbo_ctx* c, const double* u, double addition, double multiply); /* Calc.cu:191-263 */
void vbo_nld(vbo_ctx* c, double* nldp, const double* d, const double* f);     /* Calc.cu:265-297 */
void vbo_s_term(vbo_ctx* c, double* s, const double* d, const double* dx,
                const double* dy, const double* dz);                          /* Calc.cu:358-374 */
void vbo_nld1(vbo_ctx* c, double* nld, const double* f, const double* g,
              double l1, double l2);                                          /* Calc.cu:410-426 */
void vbo_nld2(vbo_ctx* c, double* nld, const double* f, const double* g,
              double l1, double l2);                                          /* Calc.cu:428-444 */
void vbo_init_fEIF(vbo_ctx* c, double* Het);                                  /* Calc.cu:579-606 */
void vbo_etd1(vbo_ctx* c, double* a, const double* fd, const double* nld, const double* Het, double h);
void vbo_etd2(vbo_ctx* c, double* a, const double* fd, const double* nnld, const double* nld, const double* Het, double h);
void vbo_etd41(vbo_ctx* c, double* a, const double* fd, const double* nld, const double* Het, double h);
void vbo_etd42(vbo_ctx* c, double* b, const double* fd, const double* nlda, const double* Het, double h);
void vbo_etd43(vbo_ctx* c, double* cc, const double* nld, const double* nldb, const double* Het, double h);
void vbo_etd44(vbo_ctx* c, double* d, const double* fd, const double* nlda, const double* nldb,
               const double* nldc, const double* Het, double h);

/* ---- phasefield / belastic compositions -------------------------------- */
double vbo_vol(vbo_ctx* c, const double* d);                  /* phasefield.cpp:283-289 */
double vbo_surface(vbo_ctx* c, const double* d);              /* phasefield.cpp:293-316 */
double vbo_elastic_eng(vbo_ctx* c, const double* d);          /* phasefield.cpp:318-325 */
void vbo_fpf(vbo_ctx* c, const double* d, double* f);         /* phasefield.cpp:140-181 */
void vbo_g0(vbo_ctx* c, const double* d, double* g0);         /* phasefield.cpp:205-221 */
/* E = eta*Eb + 1/2 M1 (V-alpha)^2 + 1/2 M2 (A-beta)^2  (belastic.cpp:68-86) */
double vbo_energy(vbo_ctx* c, const double* d, int constrained, double* V, double* A, double* Eb);
/* N^ = F(nld2(f, nld; lambda1, lambda2)) (belastic.cpp:473-488); also returns V, A */
void vbo_nld_term(vbo_ctx* c, const double* d, double* nhat, double* V, double* A);
/* Euler RHS = lambda1/2 + lambda2 f + eta g0 (belastic.cpp:52-64) */
void vbo_sum_fgq(vbo_ctx* c, const double* d, double* rhs);

/* ---- steppers (state d updated in place) -------------------------------- */
void vbo_etd4rk(vbo_ctx* c, double* d, double h);   /* belastic.cpp:519-552 */
void vbo_etd2rk(vbo_ctx* c, double* d, double h);   /* belastic.cpp:491-517 */

/* ---- drivers ------------------------------------------------------------ */
typedef struct vbo_drive {
  double time_step, fixed_time_step, time_current;
  double alpha0, beta0;       /* targets at t=0 (belastic.cpp:592-593)     */
  double eng0, eng1;          /* energies of the last attempt              */
  double dt_cap;              /* 1e-3 (belastic.cpp:645)                   */
  int steps, dt_double_every; /* 4 in the reference (belastic.cpp:644)     */
  int rejections, evaluations;
  int status;                 /* 0 running, 1 dt underflow, 2 energy small, 3 non-finite */
  int pad;
} vbo_drive;

/* Start of run_static_fEIF (belastic.cpp:588-618): alpha/beta auto-init when
 * beta <= 0, fixed dt = min(fixed, dt) when fixed > 0, eng0 = E(phi0). */
void vbo_fEIF_begin(vbo_ctx* c, const double* d, vbo_drive* s);
/* n iterations of the run_static_fEIF loop body (belastic.cpp:635-746),
 * scheme 4 = ETDRK4 (the reference), 2 = ETDRK2 in the same loop.           */
int vbo_fEIF_steps(vbo_ctx* c, double* d, vbo_drive* s, int n, int scheme,
                   double* trace /* optional [n][4]: dt, E, V, A */);
/* Forward-Euler shape relaxation (belastic.cpp:174-283); returns iterations */
int vbo_shape_adjust(vbo_ctx* c, double* d, int max_iter);
/* n iterations of the Forward-Euler driver run_static (belastic.cpp:286-469) */
int vbo_static_steps(vbo_ctx* c, double* d, vbo_drive* s, int n);

/* ---- ETD coefficient kernels evaluated for a single mode value H ---------- */
/* out[0..5] = e1, (e1-1)/H, e2, A, B, C of etd44 (Calc.cu:556-567), and
 * out[6] = the ETDRK2 c2 coefficient (Calc.cu:473-475).  Reference order.  */
void vbo_etd_coeffs(double H, double h, double* out);

#ifdef __cplusplus
}
#endif
#endif
