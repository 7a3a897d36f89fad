/*
 * vesicle.h -- C ABI of the MI355X-native pseudo-spectral vesicle solver
 * (libvesicle.so).  It is a drop-in for cudaBubble's solver path: the
 * Allen-Cahn phase-field elastic-bending vesicle model on a periodic
 * [-pi, pi]^3 grid, with volume and area penalties, spectral derivatives and
 * Forward-Euler / ETDRK2 / ETDRK4 time stepping, on one AMD Instinct GPU.
 *
 * The reference has no C ABI: the path sits behind three C++ classes,
 * belastic (model and drivers), phasefield (one field with cached derived
 * quantities) and Calc (FFT plans and kernels).  Each entry point below names
 * the reference routine it replaces.  Everything is plain C: host pointers are
 * caller-owned copies, device pointers (the vb_calc_* level) are plain
 * addresses of device memory, and every call returns a vb_status code.
 *
 * Layouts follow the reference (Calc.h:4, Calc.cu:57-58, tool.cpp:28-42):
 * real fields are [nz][ny][nx] with x fastest; half spectra are
 * [nz][ny][nx/2+1] complex values, interleaved (re, im).
 */
#ifndef VESICLE_H
#define VESICLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum vb_status {
  VB_OK = 0,
  VB_ERR_PARAM = 1,        /* invalid parameter or unsupported grid size        */
  VB_ERR_ODD_NX = 2,       /* nx must be even (Calc.cu:52-55)                   */
  VB_ERR_OOM = 3,          /* device allocation failed                          */
  VB_ERR_HIP = 4,          /* a HIP runtime call failed                          */
  VB_ERR_NONFINITE = 5,    /* the field blew up (belastic.cpp:733-740)          */
  VB_ERR_DT_UNDERFLOW = 6, /* time step < 1e-23 (belastic.cpp:713-718)          */
  VB_ERR_IO = 7,           /* file open / read / write failed                    */
  VB_ERR_NO_DEVICE = 8,    /* no usable GPU                                      */
  VB_FINISHED = 9          /* a normal driver exit: E < 1e-3 (belastic.cpp:727-732),
                              shape-adjust convergence or t > tmax (belastic.cpp:407) */
} vb_status;

enum { VB_EULER = 1, VB_ETDRK2 = 2, VB_ETDRK4 = 4 };
enum { VB_FP64 = 0, VB_FP32_MIXED = 1 };
enum { VB_LUT_REFERENCE = 0, VB_LUT_ACCURATE = 1 };

/* Solver parameters.  Field names follow fluidbase/belastic members.
 * vb_params_default() fills the reference defaults. */
typedef struct vb_params {
  int nx, ny, nz;          /* grid; ny == nx, nx even (Calc.cu:45-55)           */
  int scheme;              /* VB_ETDRK4 (run_static_fEIF), VB_ETDRK2, VB_EULER  */
  double len;              /* box length in units of 2 pi (fluidbase.cpp:52-57)*/
  double eps;              /* interface width; <= 0: from epssize             */
  int epssize;             /* ini code for eps (fluidbase.cpp:68-108)           */
  int precision;           /* VB_FP64 or VB_FP32_MIXED                          */
  double eta, ga;          /* bending rigidity, Euler mobility gamma            */
  double M1, M2;           /* volume / area penalty weights                     */
  double alpha, beta;      /* targets; beta <= 0: taken from phi0 at t = 0      */
  double kappa, D;         /* 2 and 1.5 (Calc.cu:581-582)                       */
  double time_step;        /* initial dt (fluidbase ini field time_step)        */
  double fixed_time_step;  /* > 0: fixed dt = min(fixed, time_step)            */
  double dt_cap;           /* adaptive dt cap, 1e-3 (belastic.cpp:645)          */
  int dt_double_every;     /* adaptive doubling period, 4 (belastic.cpp:644)    */
  int penalty_drift;       /* 1: fixed-dt alpha/beta drift step (belastic.cpp:670-691) */
  int phi_lut_mode;        /* VB_LUT_REFERENCE or VB_LUT_ACCURATE               */
  int shape_adjust;        /* b_shape: Euler pre-relaxation at t = 0            */
  double tmax;             /* Euler driver stop time (belastic.cpp:407)         */
  int device;              /* HIP device ordinal                                */
  int graphs;              /* 1: fixed-dt iterations and adaptive ETD attempts
                              replay captured HIP graphs (same kernels and
                              arguments as eager launches); 2: also the Euler
                              attempts of shape_adjust and of the adaptive
                              VB_EULER driver, one graph for every dt;
                              0: eager launches only                            */
  int accept_chain;        /* with graphs, > 0: adaptive ETD iterations run up
                              to this many attempts (at most 16) per launch,
                              accepted or rejected on the device with the host
                              driver's arithmetic, one read-back per launch,
                              and the shape_adjust pre-pass runs in chains of 8
                              attempts the same way; 0: the host decides after
                              each attempt (belastic.cpp:174-283, 639-669)     */
} vb_params;

/* Counters of one vb_step call. */
typedef struct vb_step_report {
  int steps;               /* accepted driver iterations                        */
  int evaluations;         /* stepper evaluations (ETDRK4 calls)                */
  int rejections;          /* adaptive rejections                               */
  int status;              /* vb_status of the driver exit condition, VB_OK if none */
  double time_step;        /* dt after the call                                 */
  double time_current;     /* simulated time after the call                     */
  double energy;           /* E of the last accepted state (adaptive) or NaN    */
} vb_step_report;

typedef struct vb_solver vb_solver;
typedef struct vb_calc vb_calc;

/* ---- parameters --------------------------------------------------------- */
void vb_params_default(vb_params* p);
/* Reads line 1 of an ini<i>.dat file: 34 whitespace fields in the order of
 * fluidbase::fluidbase (fluidbase.cpp:33-35) and applies its defaults and the
 * eps table (fluidbase.cpp:52-148). */
int vb_params_from_ini(const char* path, vb_params* out);
/* eps for an ini epssize code (fluidbase.cpp:68-108); L is the box length. */
double vb_eps_from_code(int epssize, int nx, double L);

/* ---- solver (replaces belastic + phasefield) ------------------------------ */
/* belastic::belastic + phasefield ctor (belastic.cpp:15-23, phasefield.cpp:14-58):
 * phi0_host is [nz][ny][nx] doubles, copied. */
int vb_create(const vb_params* p, const double* phi0_host, vb_solver** out);
void vb_destroy(vb_solver* s);
/* n iterations of the driver loop of belastic::run_static_fEIF
 * (belastic.cpp:635-746), or of run_static (belastic.cpp:367-437) for
 * VB_EULER.  The first call performs the start of the driver
 * (belastic.cpp:558-618): shape adjustment, alpha/beta auto-init, fixed dt. */
int vb_step(vb_solver* s, int n_steps, vb_step_report* rep);
/* belastic::getstaticenergy + phasefield::vol/surface/elastic_eng
 * (belastic.cpp:68-86, phasefield.cpp:283-325) of the current field. */
int vb_query(vb_solver* s, double* E, double* V, double* A, double* Eb);
/* current targets alpha, beta, dt and time */
int vb_state(vb_solver* s, double* alpha, double* beta, double* time_step, double* time_current);
int vb_get_phi(vb_solver* s, double* host_out);        /* cudaMemcpy D2H, belastic.cpp:163 */
int vb_set_phi(vb_solver* s, const double* host_in);   /* phasefield ctor H2D copy        */
/* belastic::geterror (belastic.cpp:136-143): LSerror of the Euler right-hand
 * side lambda1/2 + lambda2 f + eta g0 of the current field (belastic.cpp:52-64),
 * the "error" column of the reference's output */
int vb_residual(vb_solver* s, double* err);
/* one stepper evaluation without the driver: ETD4RK / ETD2RK
 * (belastic.cpp:519-552, 491-517) with step h */
int vb_evaluate(vb_solver* s, double h);
/* checkpoint: <dir>/<nx>x<ny>x<nz>.info (7 doubles, fluidbase.cpp:2776-2795)
 * and the raw field <dir>/d<nx>x<ny>x<nz>.out, or d<nx>x<ny>x<nz>_<stamp>.out
 * for fields of at most 100000 bytes, stamp = round(t / 1e-6) as %08d
 * (fluidbase.cpp:259-275).  vb_save also writes the state spectrum beside
 * the field (same name, .spec: [nz][ny][nx/2+1] complex doubles); vb_load
 * uses it when its inverse transform is the field file bit for bit, so a
 * resumed run continues exactly, and otherwise loads the field. */
int vb_save(vb_solver* s, const char* dir);
int vb_load(vb_solver* s, const char* dir);
/* device stream the solver launches on (a hipStream_t) */
void* vb_stream(vb_solver* s);
/* per-kernel timing: when on, every launch is bracketed by events on the
 * solver stream; vb_kernel_times returns up to cap entries of
 * (name, launches, total ms). */
int vb_profile(vb_solver* s, int on);
int vb_kernel_times(vb_solver* s, char* names, int name_len, int* counts, double* ms, int cap);
/* bytes of device memory the solver holds (admission sizing) */
size_t vb_footprint(const vb_params* p);
const char* vb_strerror(int status);

/* ---- Calc level (replaces class Calc, Calc.h:8-119) ------------------------
 * Device pointers: real fields double* / float* [nz][ny][nx]; half spectra
 * interleaved complex [nz][ny][nx/2+1].  Precision VB_FP64 unless noted. */
int vb_calc_create(int nx, int nz, double lenx, int precision, vb_calc** out);  /* Calc::Calc, Calc.cu:43-64 */
void vb_calc_destroy(vb_calc* c);                                             /* Calc::~Calc, Calc.cu:66-75 */
int vb_calc_r2c(vb_calc* c, const void* u, void* s);              /* r2c_3d_f, Calc.cu:88-96 (FP64/FP32) */
int vb_calc_c2r(vb_calc* c, const void* s, void* u);              /* c2r_3d_b, Calc.cu:99-104 (FP64/FP32); s is preserved */
int vb_calc_laplace(vb_calc* c, void* s);                         /* laplace, Calc.cu:127-137 */
int vb_calc_lap_spec(vb_calc* c, const void* fd, void* out);      /* lap(fd,out), Calc.cu:139-144 */
int vb_calc_lap_real(vb_calc* c, const void* d, void* out);       /* lap(d,out), Calc.cu:146-151 */
int vb_calc_fftgradient(vb_calc* c, const void* fd, void* dx, void* dy, void* dz); /* Calc.cu:344-356 */
int vb_calc_f(vb_calc* c, void* f, const void* lap, const void* d, double eps);    /* Calc.cu:165-169 */
int vb_calc_g0(vb_calc* c, void* g0, const void* d, const void* f, double eps);    /* Calc.cu:184-188 */
int vb_calc_intf(vb_calc* c, const void* u, double addition, double multiply, double* host_val); /* Calc.cu:255-263 */
int vb_calc_nld(vb_calc* c, void* nldp, const void* d, const void* f, void* tmp1,
                double eps, double kappa, double D);                              /* Calc.cu:286-297 */
int vb_calc_s_term(vb_calc* c, void* s, const void* d, const void* dx, const void* dy,
                   const void* dz, double eps);                                   /* Calc.cu:370-374 */
int vb_calc_multi(vb_calc* c, const void* a, const void* b, void* out);          /* Calc.cu:387-391 */
int vb_calc_plus(vb_calc* c, void* u, const void* nld1, double t);               /* Calc.cu:404-408 */
int vb_calc_nld1(vb_calc* c, void* nld, const void* f, const void* g, double l1, double l2, double eta); /* Calc.cu:422-426 */
int vb_calc_nld2(vb_calc* c, void* nld, const void* f, const void* g, double l1, double l2, double eta); /* Calc.cu:440-444 */
/* L^ table (Calc.cu:579-606), written to device memory [nz][ny][nx/2+1] doubles */
int vb_calc_init_fEIF(vb_calc* c, void* Het, double eps);
/* ETD combines (Calc.cu:446-577).  The coefficient functions of L^ come from
 * host-built FP64 tables indexed by |k|^2 (lut_mode VB_LUT_REFERENCE keeps the
 * reference expressions), so no device exp enters the result. */
int vb_calc_etd1(vb_calc* c, void* a, const void* fd, const void* nld, double eps, double h);
int vb_calc_etd2(vb_calc* c, void* a, const void* fd, const void* nnld, const void* nld, double eps, double h);
int vb_calc_etd41(vb_calc* c, void* d_real, const void* fd, const void* nld, double eps, double h);
int vb_calc_etd42(vb_calc* c, void* b, const void* fd, const void* nlda, double eps, double h);
int vb_calc_etd43(vb_calc* c, void* cc, const void* nld, const void* nldb, double eps, double h);
int vb_calc_etd44(vb_calc* c, void* d_real, const void* fd, const void* nlda, const void* nldb,
                  const void* nldc, double eps, double h);
int vb_calc_sync(vb_calc* c);

/* The integral of vb_calc_intf, (sum(u)/n + addition) * multiply * lenx^3 over
 * n FP64 points of device memory, written to the device double *val on
 * `stream` (a hipStream_t, NULL: the null stream) without waiting, and without
 * allocating anything: no counterpart in the reference (its Calc::intf takes a
 * caller scratch array, Calc.cu:255-263).  It is for callers that cannot
 * allocate device memory: code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc.  The block partial sums live in a static
 * device array of 16384 doubles per GPU (held until the process ends), so
 * n <= 2^30 (VB_ERR_PARAM beyond); every call on a GPU shares that array, and
 * calls that overlap on one GPU (two streams, two threads) race.  The result
 * equals vb_calc_intf's bit for bit (the same kernels). */
int vb_intf_static(const double* u, long long n, double lenx, double addition, double multiply, double* val,
                   void* stream);

/* ---- initial shapes (fluidbase init_* generators, seeded) ----------------- */
/* phi = tanh(d / (sqrt2 eps)) on the reference grid x_i = i L/n - L/2
 * (fluidbase.cpp:152-161).  kind: 0 sphere, 1 ellipsoid, 2 two ellipsoids,
 * 3 torus, 4 biconcave disk.  Shape parameters are drawn from splitmix64(seed)
 * in the ranges of the benchmark configurations; noise > 0 adds N(0, noise). */
int vb_shape(int kind, unsigned long long seed, int nx, int nz, double len, double eps,
             double noise, double* phi_out, double* shape_params /* optional [8] */);

/* The reference's deterministic initial shapes by ini field ninit
 * (fluidbase.cpp:539-699): 0 sphere (:1080-1099), 1 two spheres (:1197-1215),
 * 2 spheroid (:1862-1877), 4 torus of varying tube radius (:2068-2090),
 * 5, 6, 11 oblate spheroids 2:1, 5:1, 3:1 (:1880-1932), 8 torus about the y
 * axis (:2121-2143), 50 the "tian" rod frame (:1825-1858) and 51 the gyroid
 * (:2255-2270), both sharp +-1 fields.  All axes share the box length
 * 2 pi len.  The file-based shapes (7, 35, 43) are read with vb_read_field. */
int vb_shape_ninit(int ninit, int nx, int ny, int nz, double len, double eps, double* phi_out);
/* A raw field file as write_array stores it (tool.cpp:28-42): nx*ny*nz
 * doubles, or (nx+1)*(ny+1)*nz doubles whose last x and y planes repeat the
 * first (the 65x65x64 fields of the reference's results folder), trimmed. */
int vb_read_field(const char* path, int nx, int ny, int nz, double* out);

/* ---- batched solves ---------------------------------------------------------
 * Many independent vesicles of any supported sizes on one GPU, each with its
 * own parameters (no reference counterpart: the reference runs one belastic
 * per process).  The solves are spread over a few streams; fixed-dt solves
 * replay captured HIP graphs, two driver iterations per replay, issued
 * round-robin.  Each solve computes exactly what vb_step computes for it. */
typedef struct vb_batch vb_batch;
int vb_batch_create(int n, const vb_params* params, const double* const* phi0, vb_batch** out);
/* n driver iterations of every solve (vb_step semantics per solve) */
int vb_batch_step(vb_batch* b, int n_steps);
/* the vb_step_report of solve i from the last vb_batch_step */
int vb_batch_report(vb_batch* b, int i, vb_step_report* rep);
int vb_batch_query(vb_batch* b, int i, double* E, double* V, double* A, double* Eb);
int vb_batch_get_phi(vb_batch* b, int i, double* host_out);
/* Shared stage launches across grid sizes (no reference counterpart): with
 * on != 0, the fixed-dt FP32-mixed ETDRK4 size groups of lengths 32, 64, 128
 * and 256 (default kernel schedule) run each ETD stage combine as one launch
 * over the solves of all of them, their other kernels staying one launch per
 * group; *members (may be NULL) receives the number of size groups that share
 * (0: fewer than two qualify, and nothing changes).  Results are unchanged bit
 * for bit.  Call between vb_batch_step calls. */
int vb_batch_set_shared_stages(vb_batch* b, int on, int* members);
void vb_batch_destroy(vb_batch* b);

/* ---- request loop ----------------------------------------------------------
 * Solves arrive at any time (no reference counterpart: the reference runs one
 * belastic per process, main.cpp:17-18).  One worker thread owns the device:
 * it admits queued requests in arrival order while their footprint fits the
 * memory budget and a slot is free, keeps each admitted solve resident in HBM,
 * and keeps one adaptive attempt of every running solve in flight on one of a
 * few streams, deciding each on the host as soon as it completes.  A solve
 * leaves when E dropped by at most converge_tol (relative) over the last
 * converge_window accepted steps, when t >= tmax, at a driver exit
 * (vb_step_report.status) or after max_steps.  Requests must use adaptive dt
 * (fixed_time_step <= 0) and an ETD scheme.  Every result equals what vb_step
 * computes for that request alone.  Client calls are thread-safe. */
typedef struct vb_server_config {
  int device;              /* HIP device ordinal                                  */
  int max_active;          /* solves resident and stepping at once (16)          */
  int streams;             /* device streams the running solves are spread over (4) */
  int keep_phi;            /* 1: keep each final phi for vb_server_take_phi       */
  size_t memory_budget;    /* bytes of HBM for resident solves; 0: 90% of free    */
  int converge_window;     /* accepted steps of the energy-drop test (50)         */
  int max_steps;           /* accepted-step cap per request, 0: none              */
  double converge_tol;     /* relative energy drop (1e-6)                         */
} vb_server_config;

typedef struct vb_result {
  long long id;
  int status;              /* VB_OK when it left by convergence / tmax / cap      */
  int steps, evaluations, rejections;
  double E, V, A, Eb;      /* FP64 integrals of the final state                   */
  double time_step, time_current;
  double t_submit, t_admit, t_done;   /* seconds on the server clock           */
} vb_result;

typedef struct vb_server vb_server;
void vb_server_config_default(vb_server_config* c);
int vb_server_create(const vb_server_config* cfg, vb_server** out);
/* queues a copy of phi0 ([nz][ny][nx] doubles) under the caller's id */
int vb_server_submit(vb_server* s, long long id, const vb_params* p, const double* phi0);
/* moves up to cap finished results to out; *n_out = how many */
int vb_server_poll(vb_server* s, vb_result* out, int cap, int* n_out);
int vb_server_pending(vb_server* s, int* queued, int* running, size_t* bytes_in_use);
/* the final phi of a finished request (keep_phi = 1), once */
int vb_server_take_phi(vb_server* s, long long id, double* host_out);
double vb_server_clock(vb_server* s);
/* finishes every queued and running request, then frees everything */
void vb_server_destroy(vb_server* s);

/* ---- kernel schedules and the autotuner -------------------------------------
 * No reference counterpart: cuFFT picks its own plans inside the reference's
 * Calc (Calc.cu:56-58).  Each grid length is compiled with a few kernel
 * schedules: the column tile of the z kernels and the radix plan of the ETD
 * x kernel (FP32-mixed only; FP64 has one).  A process-wide selection per
 * (length, precision), schedule 0 until set, applies to every solver, batch
 * and Calc created afterwards; solves created under different selections
 * differ in rounding only.  The schedule functions need no GPU. */
#define VB_MAX_SCHEDULES 4
typedef struct vb_tune_report {
  int n, precision, steps;
  int candidates;                          /* schedules compiled for (n, precision) */
  int chosen;                              /* the fastest, now selected             */
  double ms_per_step[VB_MAX_SCHEDULES];    /* best of 3 timings of `steps` driver steps */
} vb_tune_report;
/* schedules compiled for grid length n (0: length not supported) */
int vb_schedule_count(int n, int precision);
/* the selected schedule (-1: length not supported) */
int vb_schedule_get(int n, int precision);
int vb_schedule_set(int n, int precision, int sched);
/* z_tile: columns per workgroup of the z kernels; x_passes, x_radix: passes
 * of the ETD x kernel's FFT plan and the radix of its first pass */
int vb_schedule_describe(int n, int precision, int sched, int* z_tile, int* x_passes, int* x_radix);
/* Times every schedule of (n, precision) on the benchmark workload (n^3
 * torus, eps = 2h, fixed-dt ETDRK4 with penalty drift, dt = 5e-7) on the
 * device and selects the fastest.  Makes the device current on the calling
 * thread; do not call it while other solves of that length are created. */
int vb_autotune(int n, int precision, int steps, int device, vb_tune_report* rep);

/* The ETD coefficient table the solver uses for step h (host only, no GPU):
 * out[m][8] for m = kx^2 + ky'^2 + kz'^2 in [0, mmax] with folded indices.
 * ETDRK4 records {e1, (e1-1)/H, e2, A s, 2 B s, C s, 0, H} (Calc.cu:495-567,
 * s = 1/(h^2 H^3)); ETDRK2 records {e, (e-1)/H, c2, 0, 0, 0, 0, H}
 * (Calc.cu:453-475); H = L^ of Calc::init_fEIF (Calc.cu:599-600). */
int vb_lut(int scheme, int lut_mode, double h, double eps, double len, double kappa, double D, int mmax,
           double* out);

/* ---- measurement -------------------------------------------------------------
 * No reference counterpart.  The streaming ceiling of the device: copies of a
 * `bytes` buffer into a second one with 16-byte loads and stores (one or four
 * per thread, plain or nontemporal), `repeats` times each, timed with HIP
 * events; read + write bytes per second of the fastest variant in *best_GBps,
 * and of each variant in variant_GBps[0..cap) (order: 1x16B, 4x16B,
 * 1x16B nontemporal, 4x16B nontemporal).  Allocates and frees its buffers. */
int vb_copy_bandwidth(int device, size_t bytes, int repeats, double* best_GBps, double* variant_GBps, int cap);

const char* vb_version(void);

#ifdef __cplusplus
}
#endif
#endif
