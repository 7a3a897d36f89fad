// vb_solver.hpp -- one vesicle solve on one GPU: buffers, fused stage
// pipeline and the reference's drivers (belastic.cpp).
#pragma once
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "vb_host.hpp"
#include "vb_kernels.hpp"
#include "vb_ops.hpp"

namespace vb {

struct SolverBase {
  vb_params p{};
  virtual ~SolverBase() = default;
  virtual void step(int n, vb_step_report* rep) = 0;
  virtual void query(double* E, double* V, double* A, double* Eb) = 0;
  virtual void state(double* alpha, double* beta, double* dt, double* t) = 0;
  virtual void set_state(double alpha, double beta, double dt, double t, double alpha0, double beta0) = 0;
  virtual void get_phi(double* host) = 0;
  virtual void set_phi(const double* host) = 0;
  virtual void evaluate(double h) = 0;
  virtual hipStream_t stream() = 0;
  virtual void profile(bool on) = 0;
  virtual int kernel_times(char* names, int name_len, int* counts, double* ms, int cap) = 0;
  // extra info stored in the checkpoint (.info): alpha3, alpha_1, beta_1
  double info_extra[3] = {0, 0, 0};
};

std::unique_ptr<SolverBase> make_solver(const vb_params& p, hipStream_t shared_stream);
size_t solver_footprint(const vb_params& p);

}  // namespace vb
