// vesicle_app.cpp -- the belastic-shaped driver over libvesicle's C ABI.
#include "vesicle_app.hpp"

#include <sys/stat.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <vector>

namespace vbhost {

namespace {

constexpr double kTimeStepOutput = 1e-6;   // stdafx.h:41

int roundx(double x) { return int(floor(x + 0.5)); }   // tool.cpp:23-26

bool exists(const std::string& fn) {
  struct stat st;
  return stat(fn.c_str(), &st) == 0;
}

void check(int rc, const char* what) {
  if (rc != VB_OK) throw std::runtime_error(std::string(what) + ": " + vb_strerror(rc));
}

}  // namespace

belastic::belastic(int folderindex, const std::string& result_path) {
  folderindex_ = folderindex < 0 ? 0 : folderindex;
  result_path_ = result_path;
  if (result_path_.empty() || result_path_.back() != '/') result_path_ += '/';
  inifile_ = result_path_ + "ini" + std::to_string(folderindex_) + ".dat";
  FILE* f = fopen(inifile_.c_str(), "r");
  if (!f) throw std::runtime_error("Can not open the ini file [" + inifile_ + "]");
  char line[4096] = {0};
  const bool got = fgets(line, sizeof line, f) != nullptr;
  fclose(f);
  int nx = 0, ny = 0, nz = 0, b_shape = 0;
  double xl, yl, zl, ts, fts, fixed;
  if (!got || sscanf(line, "%d %d %d %d %lf %lf %lf %d %lf %lf %lf %d", &ifold_, &nx, &ny, &nz, &xl, &yl, &zl,
                     &b_shape, &ts, &fts, &fixed, &ninit_) != 12)
    throw std::runtime_error("malformed ini file [" + inifile_ + "]");
  check(vb_params_from_ini(inifile_.c_str(), &p_), "vb_params_from_ini");
  if (p_.nx <= 0 || p_.nz <= 0) throw std::runtime_error("Illegal ngrid number!");
  folder_ = result_path_ + std::to_string(ifold_) + "/";
  mkdir(folder_.c_str(), 0755);
  output(folder_.c_str());
  output("\n");
  // keep a copy of the ini with the results (fluidbase.cpp:45-47)
  const std::string bak = folder_ + "ini" + std::to_string(folderindex_) + ".dat";
  if (FILE* o = fopen(bak.c_str(), "w")) {
    fputs(line, o);
    fclose(o);
  }
}

belastic::~belastic() {
  if (s_) vb_destroy(s_);
}

// ---- output sinks ------------------------------------------------------------
void belastic::output(const char* s) {
  char fn[1024];
  snprintf(fn, sizeof fn, "%soutput%dx%dx%d.txt", folder_.c_str(), p_.nx, p_.nx, p_.nz);
  if (FILE* f = fopen(fn, "a+")) {
    fputs(s, f);
    fclose(f);
  }
  fputs(s, stdout);
  fflush(stdout);
}

static double current_time(vb_solver* s) {
  double t = 0.0;
  if (s) vb_state(s, nullptr, nullptr, nullptr, &t);
  return t;
}

void belastic::output(const char* name, double x) {
  char fn[1024];
  snprintf(fn, sizeof fn, "%s%s%dx%dx%d.txt", folder_.c_str(), name, p_.nx, p_.nx, p_.nz);
  if (FILE* f = fopen(fn, "a+")) {
    fprintf(f, "%08d, %15.12f\n", roundx(current_time(s_) / kTimeStepOutput), x);
    fclose(f);
  }
}

void belastic::output(const char* name, double x1, double x2) {
  char fn[1024];
  snprintf(fn, sizeof fn, "%s%s%dx%dx%d.txt", folder_.c_str(), name, p_.nx, p_.nx, p_.nz);
  if (FILE* f = fopen(fn, "a+")) {
    fprintf(f, "%08d, %15.12f, %15.12f\n", roundx(current_time(s_) / kTimeStepOutput), x1, x2);
    fclose(f);
  }
}

void belastic::output9(const char* name, const double* x) {
  char fn[1024];
  snprintf(fn, sizeof fn, "%s%s%dx%dx%d.txt", folder_.c_str(), name, p_.nx, p_.nx, p_.nz);
  if (FILE* f = fopen(fn, "a+")) {
    fprintf(f, "%08d", roundx(current_time(s_) / kTimeStepOutput));
    for (int i = 0; i < 9; ++i) fprintf(f, ", %8.5f", x[i]);
    fprintf(f, "\n");
    fclose(f);
  }
}

// x-z and x-y mid-plane slices as text, and the raw field (fluidbase.cpp:224-265)
void belastic::output_field(const double* phi) {
  const int nx = p_.nx, ny = p_.nx, nz = p_.nz;
  const int stamp = roundx(current_time(s_) / kTimeStepOutput);
  char fn[1024];
  snprintf(fn, sizeof fn, "%sd%dx%dx%d_%08d_xz.txt", folder_.c_str(), nx, ny, nz, stamp);
  if (FILE* f = fopen(fn, "w+")) {
    for (int i = 0; i < nz; ++i) {
      for (int k = 0; k < nx; ++k) fprintf(f, "%7.5f ", phi[((size_t)i * ny + ny / 2) * nx + k]);
      fprintf(f, "\n");
    }
    fclose(f);
  }
  snprintf(fn, sizeof fn, "%sd%dx%dx%d_%08d_xy.txt", folder_.c_str(), nx, ny, nz, stamp);
  if (FILE* f = fopen(fn, "w+")) {
    for (int j = 0; j < ny; ++j) {
      for (int k = 0; k < nx; ++k) fprintf(f, "%6.3f ", phi[((size_t)(nz / 2) * ny + j) * nx + k]);
      fprintf(f, "\n");
    }
    fclose(f);
  }
  // the raw field itself is the checkpoint's (vb_save: d<dims>.out, or
  // d<dims>_<stamp>.out for fields of at most 100000 bytes)
}

double belastic::energy(double* V, double* A, double* Eb) {
  double E = NAN;
  check(vb_query(s_, &E, V, A, Eb), "vb_query");
  return E;
}

void belastic::outputall() {
  double E, V, A, Eb, a, b, dt, t, err = NAN;
  check(vb_query(s_, &E, &V, &A, &Eb), "vb_query");
  check(vb_state(s_, &a, &b, &dt, &t), "vb_state");
  (void)vb_residual(s_, &err);
  // outputenergy (belastic.cpp:88-113): eta E_b, volume and area penalties, total
  const double e1 = p_.eta * Eb, e4 = 0.5 * p_.M1 * (V - a) * (V - a), e5 = 0.5 * p_.M2 * (A - b) * (A - b);
  const double eng[9] = {e1, 0.0, 0.0, e4, e5, 0.0, e1 + e4 + e5, 0.0, e1 + e4 + e5};
  output9("energy", eng);
  output("vol_sur", V, A);
  output("lamb", p_.M1 * (V - a), p_.M2 * (A - b));
  output("error", err);
  output("time_step", dt);
  output("alpha_beta", a, b);
  check(vb_save(s_, folder_.c_str()), "vb_save");   // .info + d<dims>.out
  std::vector<double> phi((size_t)p_.nx * p_.nx * p_.nz);
  check(vb_get_phi(s_, phi.data()), "vb_get_phi");
  output_field(phi.data());
  char line[512];
  snprintf(line, sizeof line, "current time is %12.8f\n", t);
  output(line);
  snprintf(line, sizeof line, "vol: %12.7f surf: %12.7f eng: %12.10f Error: %12.10f Time_step: %e\n", V, A,
           e1 + e4 + e5, err, dt);
  output(line);
}

bool belastic::check_folder() {
  FILE* f = fopen(inifile_.c_str(), "r");
  if (!f) return false;
  int ifold = 0;
  const int n = fscanf(f, "%d", &ifold);
  fclose(f);
  if (n == 1 && ifold != ifold_) {
    output("User terminate!\n");
    return true;
  }
  return false;
}

// Initial field: continue a saved experiment, else the shape of ninit.
bool belastic::load_initial() {
  const size_t n = (size_t)p_.nx * p_.nx * p_.nz;
  std::vector<double> phi(n, 0.0);
  char info[1024];
  snprintf(info, sizeof info, "%s%dx%dx%d.info", folder_.c_str(), p_.nx, p_.nx, p_.nz);
  const bool resume = exists(info);
  if (!resume) {
    output("Starting a new experiment.... start time is 0.\n");
    char fn[1024];
    // a field file that cannot be read leaves zeros, as read_array does (tool.cpp:44-53)
    auto load_or_zero = [&](const char* path) {
      if (vb_read_field(path, p_.nx, p_.nx, p_.nz, phi.data()) != VB_OK) {
        char msg[1200];
        snprintf(msg, sizeof msg, "Can not load file %s! set as zeros\n", path);
        output(msg);
        std::fill(phi.begin(), phi.end(), 0.0);
      }
    };
    const int analytic[] = {0, 1, 2, 4, 5, 6, 8, 11, 50, 51};
    if (std::find(std::begin(analytic), std::end(analytic), ninit_) != std::end(analytic)) {
      check(vb_shape_ninit(ninit_, p_.nx, p_.nx, p_.nz, p_.len, p_.eps, phi.data()), "vb_shape_ninit");
    } else if (ninit_ == 7) {   // the field saved in the folder (fluidbase.cpp:560-562, 268-275), t = 0
      if ((size_t)p_.nx * p_.nx * p_.nz * 8 > 100000)
        snprintf(fn, sizeof fn, "%sd%dx%dx%d.out", folder_.c_str(), p_.nx, p_.nx, p_.nz);
      else
        snprintf(fn, sizeof fn, "%sd%dx%dx%d_%08d.out", folder_.c_str(), p_.nx, p_.nx, p_.nz, 0);
      load_or_zero(fn);
    } else if (ninit_ == 35) {  // the dimpled disk of the results folder (fluidbase.cpp:2232-2238)
      snprintf(fn, sizeof fn, "%sdiskflat%dx%dx%d.out", result_path_.c_str(), p_.nx, p_.nx, p_.nz);
      if (!exists(fn)) snprintf(fn, sizeof fn, "%sdisk%dx%dx%d.out", result_path_.c_str(), p_.nx + 1, p_.nx + 1, p_.nz);
      load_or_zero(fn);
    } else if (ninit_ == 43) {  // the drop of the results folder (fluidbase.cpp:2217-2228)
      snprintf(fn, sizeof fn, "%sdrop%dx%dx%d.out", result_path_.c_str(), p_.nx, p_.nx, p_.nz);
      if (!exists(fn)) snprintf(fn, sizeof fn, "%sdrop%dx%dx%d.out", result_path_.c_str(), p_.nx + 1, p_.nx + 1, p_.nz);
      load_or_zero(fn);
    } else {
      throw std::runtime_error("initial shape ninit = " + std::to_string(ninit_) + " is not supported");
    }
  } else {
    output("Continue the existing experiment....");
  }
  check(vb_create(&p_, phi.data(), &s_), "vb_create");
  if (resume) {
    check(vb_load(s_, folder_.c_str()), "vb_load");
    char line[128];
    snprintf(line, sizeof line, "Set start time to be %12.8f.\n", current_time(s_));
    output(line);
  }
  return resume;
}

void belastic::header(bool fixed) {
  double a, b, dt, t;
  check(vb_state(s_, &a, &b, &dt, &t), "vb_state");
  char line[512];
  snprintf(line, sizeof line, "grid(xyz): %dx%dx%d, timestep = %e\n", p_.nx, p_.nx, p_.nz, p_.time_step);
  output(line);
  if (fixed) {
    snprintf(line, sizeof line, "FIXED timestep = %e    ", dt);
    output(line);
  } else {
    output("Adjustment time_steps; ");
  }
  snprintf(line, sizeof line, "eps = %9.6f, ga = %9.6f, eta = %9.6f, tmax = %9.6f\n", p_.eps, p_.ga, p_.eta, p_.tmax);
  output(line);
  snprintf(line, sizeof line, "alpha = %9.6f, beta = %9.6f, M1 = %8.2f, M2 = %8.2f\n", a, b, p_.M1, p_.M2);
  output(line);
  output("********* Start STATIC Running ***********\n");
  if (t == 0.0) {
    snprintf(line, sizeof line, "Preserving volume and surface. (Initial volume = %12.8f, surface = %12.8f\n", a, b);
    output(line);
    outputall();
  }
}

void belastic::run(int scheme) {
  p_.scheme = scheme;
  if (!s_) load_initial();
  vb_step_report r{};
  check(vb_step(s_, 0, &r), "vb_step");   // start of the driver: shape adjustment, alpha/beta, dt
  const bool fixed = p_.fixed_time_step > 0;
  header(fixed);
  auto last = std::chrono::steady_clock::now();
  for (;;) {
    int chunk = (fixed && scheme != VB_EULER) ? 16 - int(steps_ % 16) : 1;
    if (max_steps > 0) chunk = (int)std::min<long long>(chunk, max_steps - steps_);
    double err_before = NAN;
    if (scheme == VB_EULER) (void)vb_residual(s_, &err_before);
    const int rc = vb_step(s_, chunk, &r);
    steps_ += r.steps;
    status_ = rc != VB_OK ? rc : r.status;
    char line[256];
    if (status_ == VB_ERR_DT_UNDERFLOW) {
      output("time_step < 1e-23, get exit.\n");
      outputall();
      break;
    }
    if (status_ == VB_FINISHED) {
      output(scheme == VB_EULER ? "Too long time running, get exit.\n" : "Energy is too small < 1e-3, get exit.\n");
      outputall();
      break;
    }
    if (status_ == VB_ERR_NONFINITE) {
      snprintf(line, sizeof line, "Static Director Blow up! Blow up at time %12.8f, %lld steps\n", current_time(s_),
               steps_);
      output(line);
      break;
    }
    if (status_ != VB_OK) {
      snprintf(line, sizeof line, "solver error: %s\n", vb_strerror(status_));
      output(line);
      break;
    }
    if (scheme == VB_EULER && err_before < 1e-25) {   // belastic.cpp:415-421
      output("Successfully reached a local minimum! Exit.\n");
      outputall();
      break;
    }
    const auto now = std::chrono::steady_clock::now();
    if (std::chrono::duration<double>(now - last).count() >= output_every_s) {
      outputall();
      last = now;
    }
    if (max_steps > 0 && steps_ >= max_steps) {
      outputall();
      break;
    }
    if (steps_ % 16 == 0 && check_folder()) break;
  }
  output("********* End Running ***********\n\n\n");
}

void belastic::run_static_fEIF() { run(p_.scheme == VB_ETDRK2 ? VB_ETDRK2 : VB_ETDRK4); }

void belastic::run_static() { run(VB_EULER); }

}  // namespace vbhost
