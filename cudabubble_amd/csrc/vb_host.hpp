// vb_host.hpp -- host-side internals of libvesicle (not part of the ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "../../include/vesicle.h"
#include "vb_device.hpp"

namespace vb {

// Every HIP call is checked; a failure becomes VB_ERR_HIP with a message.
struct HipError {
  hipError_t err;
  std::string where;
};
#define VB_HIP(call)                                                              \
  do {                                                                            \
    hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess) throw ::vb::HipError{e_, std::string(#call) + " @ " +   \
                                                  std::to_string(__LINE__)};      \
  } while (0)

void set_last_error(const std::string& msg);

// VB_OK, or the status vb_create returns for these parameters (grid, lengths,
// scheme, precision, table mode, time-step settings); every entry point that
// takes a vb_params checks it
int validate_params(const vb_params* p);

// ---- host tables (vb_lut.cpp, compiled without FMA contraction) -----------
// L^ of Calc::init_fEIF (Calc.cu:599-600) for c = -m (2 pi / L)^2.
double lhat(double m, double kfac2, double eps, double kappa, double D);
// Coefficient records for every m in [0, mmax] (see Coef).
void build_lut(std::vector<Coef>& out, int mmax, int scheme, int mode, double h, double kfac2, double eps,
               double kappa, double D);
// exp(-2 pi i m / n), m in [0, n), as interleaved doubles
void build_twiddles(std::vector<double>& out, int n);
int mmax_of(int nx, int ny, int nz);

// ---- device buffers ---------------------------------------------------------
template <typename X>
struct DBuf {
  X* p = nullptr;
  size_t n = 0;
  void alloc(size_t count) {
    free();
    if (count == 0) return;
    hipError_t e = hipMalloc(&p, count * sizeof(X));
    if (e != hipSuccess) {
      p = nullptr;
      throw HipError{e, "hipMalloc"};
    }
    n = count;
  }
  void free() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  ~DBuf() { free(); }
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
};

Geo make_geo(int nx, int nz, double L);

// A stream with a hardware queue of its own: created with a CU mask that
// enables every CU of the current device.  Plain streams are mapped onto the
// process's few hardware queues in launch order, so streams created after
// others have launched work can end up sharing a queue and serialise; the
// concurrent streams of a batch or a request loop must not.
hipStream_t make_queue_stream();

// arguments of the integrals of one prefix field, and the standalone
// finalize kernel's launcher (vb_solver.hip)
FinArgs fin_args(const double* xpart, int nxb, const double* zpart, const double* dc, int nzb, Scal* sc,
                 FinScratch* fs, const Geo& g, const Phys& ph, int mode);
void launch_finalize(const double* xpart, int nxb, const double* zpart, const double* dc, int nzb, Scal* sc,
                     const Geo& g, const Phys& ph, int mode, hipStream_t s);
void launch_fill_scal(Scal* sc, double alpha, double beta, double alpha0, double beta0, hipStream_t s);

// device-side accept of adaptive attempts (vb_chain.hip): the control kernels
// before and after an attempt, and the copies of an accepted candidate
struct ChainCopy {
  struct Job {
    const void* src;
    void* dst;
    size_t bytes;   // a multiple of 8
  } job[3];
  int n;
};
void launch_chain_begin(ChainCtl* c, const Scal* sc, hipStream_t s);
void launch_chain_table(const ChainCtl* c, double2* c1, double4* c2, int nrec, hipStream_t s);
void launch_chain_end(ChainCtl* c, Scal* sc, const Scal* test, hipStream_t s);
void launch_chain_copy(const ChainCtl* c, const ChainCopy& cp, hipStream_t s);
// the copies of cp when *flag == want (device flag, read by the kernel)
void launch_cond_copy(const int* flag, int want, const ChainCopy& cp, hipStream_t s);
void launch_shape_begin(ShapeCtl* c, const Scal* sc, hipStream_t s);
void launch_shape_end(ShapeCtl* c, Scal* sc, hipStream_t s);
// *d = v on stream s (a scalar a captured graph reads)
void launch_set_double(double* d, double v, hipStream_t s);

}  // namespace vb
