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

// Solves whose pipelines issue the same launches (grid and kernel) and can
// therefore share batched launches.
struct GroupKey {
  int ok;   // 0: never grouped (adaptive dt, Euler, graphs off)
  int nx, nz, precision, scheme, drift;
  bool operator<(const GroupKey& o) const {
    const int a[6] = {ok, nx, nz, precision, scheme, drift}, b[6] = {o.ok, o.nx, o.nz, o.precision, o.scheme, o.drift};
    for (int i = 0; i < 6; ++i)
      if (a[i] != b[i]) return a[i] < b[i];
    return false;
  }
};

struct SolverBase {
  vb_params p{};
  virtual ~SolverBase() = default;
  virtual void step(int n, vb_step_report* rep) = 0;
  virtual void query(double* E, double* V, double* A, double* Eb) = 0;
  virtual void state(double* alpha, double* beta, double* dt, double* t) = 0;
  virtual void set_state(double alpha, double beta, double dt, double t, double alpha0, double beta0) = 0;
  virtual void get_phi(double* host) = 0;
  virtual void set_phi(const double* host) = 0;
  // the state spectrum u^ as [nz][ny][nx/2+1] complex doubles (the real
  // field's forward transform, normalised as Calc::r2c_3d_f)
  virtual void get_spec(double* host) = 0;
  virtual void set_spec(const double* host) = 0;
  virtual void evaluate(double h) = 0;
  virtual double residual() = 0;
  virtual hipStream_t stream() = 0;
  virtual void profile(bool on) = 0;
  virtual int kernel_times(char* names, int name_len, int* counts, double* ms, int cap) = 0;
  // Fixed-dt drivers only: launch_steps() issues n driver iterations without
  // host synchronisation (graph replays once the pipeline recurs), and
  // finish_steps() does what step() does after its loop (target read-back,
  // blow-up test).  step(n) is launch_steps(n) + finish_steps() for them.
  virtual bool fixed_dt() const = 0;
  virtual void launch_steps(int n, vb_step_report* rep) = 0;
  virtual void finish_steps(vb_step_report* rep) = 0;
  // Adaptive ETD drivers: one attempt of the current iteration is issued
  // without waiting (evaluation, candidate energy, scalars copied to the host);
  // attempt_ready() polls it; adaptive_complete() accepts it (returns 1, time
  // and steps advanced, exits in rep) or rejects it (returns 0, dt halved).
  // step(n) on an adaptive solver is issue/complete until n acceptances.
  // Batched launches over the solves of a size group (fixed dt): record the
  // launches of two driver iterations from the current state instead of
  // issuing them (returns their evaluations, 0 if the state does not recur
  // yet); at_recorded_state() tells whether a replay of that recording is
  // valid now; advance_recorded() moves the host state as the two iterations
  // would.
  virtual GroupKey group_key() const = 0;
  // the column tile of this solve's z kernels (its group-major layout)
  virtual int z_tile() const = 0;
  virtual int record_pair(Recorder& rec) = 0;
  virtual bool at_recorded_state() const = 0;
  virtual void advance_recorded(vb_step_report* rep) = 0;
  virtual void adaptive_issue(vb_step_report& rep) = 0;
  virtual bool attempt_ready() = 0;
  virtual int adaptive_complete(vb_step_report& rep) = 0;
  // adaptive_issue() in its parts, so that a scheduler can issue the attempts
  // of several solves as shared launches: attempt_prepare() is the host half
  // (the iteration's dt, the current energy; false after a driver exit in
  // rep.status), then either issue_prepared() launches the attempt itself, or
  // record_attempt() records its launches instead (its tables uploaded on
  // stream() first; the host state advances as if they ran), the caller
  // issues them on stream(), and attempt_issued() adds the copy of the
  // candidate's scalars and the completion event.
  virtual bool attempt_prepare(vb_step_report& rep) = 0;
  virtual void issue_prepared() = 0;
  virtual void record_attempt(Recorder& rec) = 0;
  virtual void attempt_issued() = 0;
  // The request loop's slot pool: reuse this solver's memory for another solve
  // of the same grid, precision and buffer set (false if p needs others).
  // Captured graphs and tables of the old parameters are dropped; phi0 is
  // uploaded on the solver's stream.
  virtual bool reset(const vb_params& p, const double* phi0) = 0;
  // E, V, A and E_b of the current state as the last accepted adaptive attempt
  // left them in host memory (FP64 in both precisions), without a device
  // round trip; false if no attempt was accepted or the state changed since.
  virtual bool accepted_values(double* E, double* V, double* A, double* Eb) const = 0;
  // extra info stored in the checkpoint (.info): alpha3, alpha_1, beta_1
  double info_extra[3] = {0, 0, 0};
};

std::unique_ptr<SolverBase> make_solver(const vb_params& p, hipStream_t shared_stream);
size_t solver_footprint(const vb_params& p);

}  // namespace vb
