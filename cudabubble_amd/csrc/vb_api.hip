// vb_api.hip -- C ABI of libvesicle: parameters and ini loading, solver
// lifetime, stepping, queries, checkpoints, seeded initial shapes, batches.
#include <sys/stat.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "vb_host.hpp"
#include "vb_ops.hpp"
#include "vb_solver.hpp"

namespace vb {
static std::mutex g_err_mu;
static std::string g_err;
void set_last_error(const std::string& msg) {
  std::lock_guard<std::mutex> l(g_err_mu);
  g_err = msg;
}
}  // namespace vb

struct vb_solver {
  std::unique_ptr<vb::SolverBase> impl;
};

namespace {

template <class F>
int guard(F&& f) {
  try {
    return f();
  } catch (const vb::HipError& e) {
    vb::set_last_error(std::string(hipGetErrorString(e.err)) + " in " + e.where);
    return e.err == hipErrorOutOfMemory ? VB_ERR_OOM : VB_ERR_HIP;
  } catch (const std::bad_alloc&) {
    return VB_ERR_OOM;
  }
}

const double kPI = 3.1415926535897932384626433832795;   // tool.h:7

}  // namespace

int vb::validate_params(const vb_params* p) {
  if (!p) return VB_ERR_PARAM;
  if (p->nx <= 0 || p->nz <= 0) return VB_ERR_PARAM;
  if (p->nx % 2) return VB_ERR_ODD_NX;
  if (p->nz % 2) return VB_ERR_PARAM;
  if (!vb::size_supported(p->nx) || !vb::size_supported(p->nz)) return VB_ERR_PARAM;
  if (!(p->len > 0) || !(p->eps > 0)) return VB_ERR_PARAM;
  if (p->scheme != VB_ETDRK4 && p->scheme != VB_ETDRK2 && p->scheme != VB_EULER) return VB_ERR_PARAM;
  if (p->precision != VB_FP64 && p->precision != VB_FP32_MIXED) return VB_ERR_PARAM;
  if (p->phi_lut_mode != VB_LUT_REFERENCE && p->phi_lut_mode != VB_LUT_ACCURATE) return VB_ERR_PARAM;
  if (!(p->time_step > 0) || p->dt_double_every <= 0 || !(p->dt_cap > 0)) return VB_ERR_PARAM;
  if (p->accept_chain < 0 || p->accept_chain > 16) return VB_ERR_PARAM;
  if (p->graphs < 0 || p->graphs > 2) return VB_ERR_PARAM;
  return VB_OK;
}

namespace {
int validate(const vb_params* p) { return vb::validate_params(p); }

// splitmix64 (Steele, Lea, Flood 2014)
struct Rng {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uni() { return double(next() >> 11) * (1.0 / 9007199254740992.0); }
  double uni(double a, double b) { return a + (b - a) * uni(); }
  double normal() {   // Box-Muller
    double u1 = uni(), u2 = uni();
    if (u1 < 1e-300) u1 = 1e-300;
    return sqrt(-2.0 * log(u1)) * cos(2.0 * kPI * u2);
  }
};

}  // namespace

extern "C" {

const char* vb_version(void) { return "vesicle-mi355x 0.1 (gfx950)"; }

const char* vb_strerror(int s) {
  switch (s) {
    case VB_OK: return "ok";
    case VB_ERR_PARAM: return "invalid parameter or unsupported grid size";
    case VB_ERR_ODD_NX: return "the grid nx should be an even number";
    case VB_ERR_OOM: return "device out of memory";
    case VB_ERR_HIP: {
      static thread_local std::string m;
      std::lock_guard<std::mutex> l(vb::g_err_mu);
      m = "HIP error: " + vb::g_err;
      return m.c_str();
    }
    case VB_ERR_NONFINITE: return "the field blew up (non-finite values)";
    case VB_ERR_DT_UNDERFLOW: return "time step < 1e-23";
    case VB_ERR_IO: return "file input/output failed";
    case VB_ERR_NO_DEVICE: {
      static thread_local std::string m;
      std::lock_guard<std::mutex> l(vb::g_err_mu);
      m = "no usable GPU (" + vb::g_err + ")";
      return m.c_str();
    }
    case VB_FINISHED: return "driver finished";
    default: return "unknown status";
  }
}

double vb_eps_from_code(int epssize, int nx, double xlen) {
  // fluidbase.cpp:68-108 (xlen is the box length, already 2 pi * len)
  if (epssize == 1) return 2.5 * (sqrt(2.) * xlen / 2.0 / nx);
  if (epssize == 2) return 2.5 * (2 * xlen / 2.0 / nx);
  if (epssize == 3) return 1.5 * (2 * xlen / 2.0 / nx);
  if (epssize == 4) return 1.25 * (2 * xlen / 2.0 / nx);
  if (epssize == 5) return 1.0 * (2 * xlen / 2.0 / nx);
  if (epssize == 6) return 0.7 * (2 * xlen / 2.0 / nx);
  if (epssize == 7) return 4.0 * (2 * xlen / 2.0 / nx);
  if (epssize == 8) return 2.5 * (sqrt(2.) * xlen / 2.0 / nx) * 1.22;
  if (epssize > 10) return epssize / 100.0 * (2 * xlen / 2.0 / nx);
  return 2.0 * (2 * xlen / 2.0 / nx);
}

void vb_params_default(vb_params* p) {
  if (!p) return;
  memset(p, 0, sizeof(*p));
  p->nx = p->ny = p->nz = 64;
  p->scheme = VB_ETDRK4;
  p->len = 1.0;
  p->epssize = 200;
  p->eps = vb_eps_from_code(200, 64, 2.0 * kPI);
  p->precision = VB_FP64;
  p->eta = 1.0;
  p->ga = 1.0;
  p->M1 = 1e4;
  p->M2 = 1e4;
  p->alpha = 0.0;
  p->beta = 0.0;
  p->kappa = 2.0;
  p->D = 1.5;
  p->time_step = 1e-7;
  p->fixed_time_step = 0.0;
  p->dt_cap = 1e-3;
  p->dt_double_every = 4;
  p->penalty_drift = 1;
  p->phi_lut_mode = VB_LUT_REFERENCE;
  p->shape_adjust = 0;
  p->tmax = 4.0;
  p->device = 0;
  p->graphs = 1;
  p->accept_chain = 0;
}

int vb_params_from_ini(const char* path, vb_params* out) {
  if (!path || !out) return VB_ERR_PARAM;
  FILE* fp = fopen(path, "r");
  if (!fp) return VB_ERR_IO;
  char line[4096];
  if (!fgets(line, sizeof line, fp)) {
    fclose(fp);
    return VB_ERR_IO;
  }
  fclose(fp);
  int ifold = 101, nx = 0, ny = 0, nz = 0, b_shape = 0, ninit = 0, nflush = 0, epssize = 2, ncur = 0;
  int output_interval = 0, einit = 0;
  double xlen = 0, ylen = 0, zlen = 0, time_step = 0, fluid_time_step = 0, fixed_time_step = 0, surctrl = 0;
  double m_angle = 0, re = 0, eta = 0, ga = 0, alpha = 0, beta = 0, M1 = 0, M2 = 0, M3 = 0, M4 = 0, tmax = 0;
  double gweight = 0, m_k = 1.0, m_c = 0.0, m_linet = 0, m_alpha3 = 0.0;
  // fluidbase.cpp:33-35: the same 34 fields in the same order
  const int got = sscanf(line,
                         "%d %d %d %d %lf %lf %lf %d %lf %lf %lf %d %d %lf %d %d %lf %lf %lf %lf %lf %lf %lf %lf "
                         "%lf %lf %lf %d %lf %d %lf %lf %lf %lf",
                         &ifold, &nx, &ny, &nz, &xlen, &ylen, &zlen, &b_shape, &time_step, &fluid_time_step,
                         &fixed_time_step, &ninit, &nflush, &surctrl, &epssize, &ncur, &m_angle, &re, &eta, &ga,
                         &alpha, &beta, &M1, &M2, &M3, &M4, &tmax, &output_interval, &gweight, &einit, &m_k, &m_c,
                         &m_linet, &m_alpha3);
  if (got < 4) return VB_ERR_IO;
  (void)ifold; (void)fluid_time_step; (void)ninit; (void)nflush; (void)surctrl; (void)ncur; (void)m_angle;
  (void)re; (void)M3; (void)M4; (void)output_interval; (void)gweight; (void)einit; (void)m_k; (void)m_c;
  (void)m_linet; (void)m_alpha3; (void)ylen; (void)zlen; (void)ny;
  if (nx <= 0 || ny <= 0 || nz <= 0) return VB_ERR_PARAM;   // fluidbase.cpp:39-42
  vb_params_default(out);
  if (xlen <= 0) xlen = 1;                                     // fluidbase.cpp:52
  out->nx = nx;
  out->ny = nx;                                                // Calc.cu:46
  out->nz = nz;
  out->len = xlen;
  out->epssize = epssize;
  out->eps = vb_eps_from_code(epssize, nx, xlen * (2.0 * kPI));
  out->time_step = time_step <= 0 ? 1e-6 : time_step;         // fluidbase.cpp:112
  out->fixed_time_step = fixed_time_step;
  out->eta = eta < 0 ? 10.0 : eta;                             // fluidbase.cpp:135
  out->ga = ga <= 0 ? 1.0 : ga;                                // fluidbase.cpp:136
  out->tmax = tmax < 0 ? 4.0 : tmax;                           // fluidbase.cpp:137
  if (beta < 0) alpha = beta = 0;                              // fluidbase.cpp:143
  out->alpha = alpha;
  out->beta = beta;
  out->M1 = M1;
  out->M2 = M2;
  out->shape_adjust = b_shape;
  return VB_OK;
}

size_t vb_footprint(const vb_params* p) {
  if (validate(p) != VB_OK) return 0;
  return vb::solver_footprint(*p);
}

int vb_create(const vb_params* p, const double* phi0, vb_solver** out) {
  if (!out) return VB_ERR_PARAM;
  *out = nullptr;
  int rc = validate(p);
  if (rc != VB_OK) return rc;
  if (!phi0) return VB_ERR_PARAM;
  int ndev = 0;
  const hipError_t de = hipGetDeviceCount(&ndev);
  if (de != hipSuccess || ndev <= 0 || p->device < 0 || p->device >= ndev) {
    vb::set_last_error(std::string("hipGetDeviceCount: ") + hipGetErrorString(de) + ", devices " +
                       std::to_string(ndev));
    return VB_ERR_NO_DEVICE;
  }
  return guard([&] {
    VB_HIP(hipSetDevice(p->device));
    auto s = std::make_unique<vb_solver>();
    s->impl = vb::make_solver(*p, nullptr);
    s->impl->set_phi(phi0);
    *out = s.release();
    return (int)VB_OK;
  });
}

void vb_destroy(vb_solver* s) { delete s; }

int vb_step(vb_solver* s, int n, vb_step_report* rep) {
  if (!s || !s->impl || n < 0) return VB_ERR_PARAM;
  return guard([&] {
    vb_step_report r{};
    s->impl->step(n, &r);
    if (rep) *rep = r;
    return (int)VB_OK;
  });
}

int vb_query(vb_solver* s, double* E, double* V, double* A, double* Eb) {
  if (!s || !s->impl) return VB_ERR_PARAM;
  return guard([&] {
    s->impl->query(E, V, A, Eb);
    return (int)VB_OK;
  });
}

int vb_state(vb_solver* s, double* a, double* b, double* dt, double* t) {
  if (!s || !s->impl) return VB_ERR_PARAM;
  return guard([&] {
    s->impl->state(a, b, dt, t);
    return (int)VB_OK;
  });
}

int vb_get_phi(vb_solver* s, double* host) {
  if (!s || !s->impl || !host) return VB_ERR_PARAM;
  return guard([&] {
    s->impl->get_phi(host);
    return (int)VB_OK;
  });
}

int vb_set_phi(vb_solver* s, const double* host) {
  if (!s || !s->impl || !host) return VB_ERR_PARAM;
  return guard([&] {
    s->impl->set_phi(host);
    return (int)VB_OK;
  });
}

int vb_evaluate(vb_solver* s, double h) {
  if (!s || !s->impl || !(h > 0)) return VB_ERR_PARAM;
  return guard([&] {
    s->impl->evaluate(h);
    return (int)VB_OK;
  });
}

int vb_residual(vb_solver* s, double* err) {
  if (!s || !s->impl || !err) return VB_ERR_PARAM;
  return guard([&] {
    *err = s->impl->residual();
    return (int)VB_OK;
  });
}

void* vb_stream(vb_solver* s) { return (s && s->impl) ? (void*)s->impl->stream() : nullptr; }

int vb_profile(vb_solver* s, int on) {
  if (!s || !s->impl) return VB_ERR_PARAM;
  return guard([&] {
    s->impl->profile(on != 0);
    return (int)VB_OK;
  });
}

int vb_kernel_times(vb_solver* s, char* names, int name_len, int* counts, double* ms, int cap) {
  if (!s || !s->impl) return -VB_ERR_PARAM;
  int n = 0;
  int rc = guard([&] {
    n = s->impl->kernel_times(names, name_len, counts, ms, cap);
    return (int)VB_OK;
  });
  return rc == VB_OK ? n : -rc;
}

}  // extern "C"

namespace {
// Name of the checkpoint field (fluidbase::output / input, fluidbase.cpp:259-275):
// d<nx>x<ny>x<nz>.out, or for fields of at most 100000 bytes
// d<nx>x<ny>x<nz>_<stamp>.out with stamp = roundx(t / TIME_STEP_OUTPUT)
// (tool.cpp:23-26, TIME_STEP_OUTPUT = 1e-6 in stdafx.h:41).
std::string field_file(const char* dir, const vb_params& p, double t) {
  char fn[1024];
  if ((long long)p.nx * p.nx * p.nz * 8 > 100000)
    snprintf(fn, sizeof fn, "%s/d%dx%dx%d.out", dir, p.nx, p.nx, p.nz);
  else
    snprintf(fn, sizeof fn, "%s/d%dx%dx%d_%08d.out", dir, p.nx, p.nx, p.nz, int(floor(t / 1e-6 + 0.5)));
  return fn;
}

// Sidecar of the field file: the solver's state spectrum, [nz][ny][nx/2+1]
// complex doubles.  The reference format stores only the real field, and this
// solver's state is the spectrum, which a transform round trip reproduces to
// rounding only; with the sidecar a resumed run continues bit for bit.
std::string spec_file(const std::string& field) { return field.substr(0, field.size() - 4) + ".spec"; }

bool read_all(const std::string& fn, std::vector<double>& out) {
  FILE* f = fopen(fn.c_str(), "rb");
  if (!f) return false;
  const size_t n = fread(out.data(), sizeof(double), out.size(), f);
  const bool at_end = fgetc(f) == EOF;
  fclose(f);
  return n == out.size() && at_end;
}
}  // namespace

extern "C" {

// checkpoint files of fluidbase::output_info / output("d", ...) (fluidbase.cpp:224-265, 2776-2795)
int vb_save(vb_solver* s, const char* dir) {
  if (!s || !s->impl || !dir) return VB_ERR_PARAM;
  const vb_params& p = s->impl->p;
  std::vector<double> phi((size_t)p.nx * p.nx * p.nz);
  int rc = vb_get_phi(s, phi.data());
  if (rc != VB_OK) return rc;
  double a, b, dt, t;
  rc = vb_state(s, &a, &b, &dt, &t);
  if (rc != VB_OK) return rc;
  char fn[1024];
  snprintf(fn, sizeof fn, "%s/%dx%dx%d.info", dir, p.nx, p.nx, p.nz);
  FILE* f = fopen(fn, "wb");
  if (!f) return VB_ERR_IO;
  const double info[7] = {t, dt, a, b, s->impl->info_extra[0], s->impl->info_extra[1], s->impl->info_extra[2]};
  const bool ok1 = fwrite(info, sizeof(double), 7, f) == 7;
  fclose(f);
  const std::string ff = field_file(dir, p, t);
  f = fopen(ff.c_str(), "wb");
  if (!f) return VB_ERR_IO;
  const bool ok2 = fwrite(phi.data(), sizeof(double), phi.size(), f) == phi.size();
  fclose(f);
  std::vector<double> spec((size_t)(p.nx / 2 + 1) * p.nx * p.nz * 2);
  rc = guard([&] {
    s->impl->get_spec(spec.data());
    return (int)VB_OK;
  });
  if (rc != VB_OK) return rc;
  f = fopen(spec_file(ff).c_str(), "wb");
  if (!f) return VB_ERR_IO;
  const bool ok3 = fwrite(spec.data(), sizeof(double), spec.size(), f) == spec.size();
  fclose(f);
  return ok1 && ok2 && ok3 ? VB_OK : VB_ERR_IO;
}

// fluidbase::load_info + reload (fluidbase.cpp:2797-2825, 2677-2686, 385-388)
int vb_load(vb_solver* s, const char* dir) {
  if (!s || !s->impl || !dir) return VB_ERR_PARAM;
  const vb_params& p = s->impl->p;
  char fn[1024];
  snprintf(fn, sizeof fn, "%s/%dx%dx%d.info", dir, p.nx, p.nx, p.nz);
  FILE* f = fopen(fn, "rb");
  if (!f) return VB_ERR_IO;
  double info[7] = {0, 0, 0, 0, 0, 0, 0};
  const size_t got = fread(info, sizeof(double), 7, f);
  fclose(f);
  if (got != 7) return VB_ERR_IO;
  double t = info[0], dt = info[1];
  if (dt < 1e-13) dt = 1e-7;          // fluidbase.cpp:2818
  if (t == 0.0) t = 1e-13;            // fluidbase.cpp:2819
  std::vector<double> phi((size_t)p.nx * p.nx * p.nz);
  const std::string ff = field_file(dir, p, t);
  f = fopen(ff.c_str(), "rb");
  if (!f) return VB_ERR_IO;
  const size_t n = fread(phi.data(), sizeof(double), phi.size(), f);
  fclose(f);
  if (n != phi.size()) return VB_ERR_IO;
  // the spectrum sidecar is used when it is the field's own: its inverse
  // transform reproduces the field file bit for bit (a field written by
  // another program, or edited, is taken as it is)
  std::vector<double> spec((size_t)(p.nx / 2 + 1) * p.nx * p.nz * 2);
  bool exact = false;
  int rc = VB_OK;
  if (read_all(spec_file(ff), spec)) {
    std::vector<double> back(phi.size());
    rc = guard([&] {
      s->impl->set_spec(spec.data());
      s->impl->get_phi(back.data());
      return (int)VB_OK;
    });
    if (rc != VB_OK) return rc;
    exact = memcmp(back.data(), phi.data(), sizeof(double) * phi.size()) == 0;
  }
  if (!exact) {
    rc = vb_set_phi(s, phi.data());
    if (rc != VB_OK) return rc;
  }
  s->impl->info_extra[0] = info[4];
  s->impl->info_extra[1] = info[5];
  s->impl->info_extra[2] = info[6];
  // on resume the drift reference targets are the saved targets
  return guard([&] {
    s->impl->set_state(info[2], info[3], dt, t, info[2], info[3]);
    return (int)VB_OK;
  });
}

int vb_lut(int scheme, int lut_mode, double h, double eps, double len, double kappa, double D, int mmax,
           double* out) {
  if (!out || mmax < 0 || !(h > 0) || !(eps > 0) || !(len > 0)) return VB_ERR_PARAM;
  if (scheme != VB_ETDRK4 && scheme != VB_ETDRK2) return VB_ERR_PARAM;
  if (lut_mode != VB_LUT_REFERENCE && lut_mode != VB_LUT_ACCURATE) return VB_ERR_PARAM;
  std::vector<vb::Coef> t;
  const vb::Geo g = vb::make_geo(2, 2, len * (2.0 * kPI));
  vb::build_lut(t, mmax, scheme == VB_ETDRK2 ? 2 : 4, lut_mode, h, g.kfac2, eps, kappa, D);
  memcpy(out, t.data(), sizeof(vb::Coef) * t.size());
  return VB_OK;
}

int vb_shape(int kind, unsigned long long seed, int nx, int nz, double len, double eps, double noise,
             double* phi, double* sp) {
  if (!phi || nx <= 0 || nz <= 0 || !(len > 0) || !(eps > 0) || kind < 0 || kind > 4) return VB_ERR_PARAM;
  Rng rng{seed};
  const int ny = nx;
  const double xlen = len * (2.0 * kPI), ylen = xlen, zlen = xlen;
  std::vector<double> xx(nx), yy(ny), zz(nz);
  for (int i = 0; i < nx; i++) xx[i] = i * xlen / nx - xlen / 2.0;   // fluidbase.cpp:156-161
  for (int i = 0; i < ny; i++) yy[i] = i * ylen / ny - ylen / 2.0;
  for (int i = 0; i < nz; i++) zz[i] = i * zlen / nz - zlen / 2.0;
  double q[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  const double w = sqrt(2.0) * eps;
  switch (kind) {
    case 0: q[0] = rng.uni(0.8, 1.8); break;                                          // sphere radius
    case 1: q[0] = rng.uni(1.0, 2.0); q[1] = rng.uni(1.0, 2.0); q[2] = rng.uni(1.0, 2.0); break;   // semi-axes
    case 2:
      for (int e = 0; e < 2; ++e) {
        for (int c = 0; c < 3; ++c) q[3 * e + c] = rng.uni(-1.0, 1.0);
      }
      q[6] = rng.uni(0.6, 1.4);
      q[7] = rng.uni(0.6, 1.4);
      break;
    case 3: q[0] = rng.uni(1.2, 1.8); q[1] = rng.uni(0.4, 0.7); break;                // torus R, r
    case 4: q[0] = rng.uni(0.8, 1.8); break;                                          // disk radius
  }
  for (int i = 0; i < nz; i++) {
    const double z = zz[i];
    for (int j = 0; j < ny; j++) {
      const double y = yy[j];
      for (int k = 0; k < nx; k++) {
        const double x = xx[k];
        double val;
        if (kind == 0) {
          val = tanh((q[0] - sqrt(x * x + y * y + z * z)) / w);
        } else if (kind == 1) {
          const double rho = sqrt(x * x / (q[0] * q[0]) + y * y / (q[1] * q[1]) + z * z / (q[2] * q[2]));
          const double m = fmin(q[0], fmin(q[1], q[2]));
          val = tanh((1.0 - rho) * m / w);
        } else if (kind == 2) {
          // union of two spheroids with equal x, y semi-axis a and z semi-axis 0.7 a
          double t = 1.0;
          for (int e = 0; e < 2; ++e) {
            const double a = q[6 + e];
            const double dx = x - q[3 * e], dy = y - q[3 * e + 1], dz = z - q[3 * e + 2];
            const double rho = sqrt((dx * dx + dy * dy) / (a * a) + dz * dz / (0.49 * a * a));
            t += tanh((1.0 - rho) * 0.7 * a / w);
          }
          val = t;
        } else if (kind == 3) {
          const double rr = sqrt(x * x + y * y) - q[0];
          val = tanh((q[1] - sqrt(rr * rr + z * z)) / w);
        } else {
          // Evans-Fung biconcave profile: half thickness h(r) = R sqrt(1-s^2)(c0 + c2 s^2 + c4 s^4), s = r/R
          const double R = q[0], r = sqrt(x * x + y * y), s = r / R;
          double d;
          if (s < 1.0) {
            const double hh = 0.5 * R * sqrt(1.0 - s * s) * (0.207 + 2.003 * s * s - 1.123 * s * s * s * s);
            d = fmin(hh - fabs(z), R - r);
          } else {
            d = -sqrt((r - R) * (r - R) + z * z);
          }
          val = tanh(d / w);
        }
        phi[((size_t)i * ny + j) * nx + k] = val;
      }
    }
  }
  if (noise > 0) {
    const size_t n = (size_t)nx * ny * nz;
    for (size_t i = 0; i < n; ++i) phi[i] += noise * rng.normal();
  }
  if (sp) memcpy(sp, q, sizeof q);
  return VB_OK;
}

}  // extern "C"
