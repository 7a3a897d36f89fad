// vb_tune.hip -- kernel schedules per transform length and the autotuner
// that picks one by timing them on the device.
//
// Each length is compiled with a few schedules (vb_size_impl.hpp): the
// column tile of the z kernels (stage combines, z-prep, z passes) and the
// radix plan of the ETD x kernel.  The fastest one depends on the length and
// on the GPU (at 256^3 in FP32: 8-column z tiles and the 16 x 16 plan; at
// 128^3 the plan with fewer passes measured no faster).  vb_autotune builds
// a solve of the benchmark's fixed-dt ETDRK4 penalty-drift workload under
// every schedule, times driver steps on the device and selects the fastest
// for every solver, batch and Calc created afterwards.  The schedules differ
// in rounding only: the x plan changes the FFT's operation order and the z
// tile the grouping of block partial sums.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <mutex>
#include <vector>

#include "vb_ops.hpp"
#include "vesicle.h"

namespace vb {

thread_local int g_sched_override = -1;

namespace {
std::mutex g_sel_mu;
// selected schedule per (FP32?, length); lengths index VB_SIZE_LIST
int g_sel[2][16] = {};

int size_index(int N) {
  int i = 0;
#define VB_IDX(n) \
  if (N == n) return i; \
  ++i;
  VB_SIZE_LIST(VB_IDX)
#undef VB_IDX
  return -1;
}

int count_for(int N, int precision) {
  if (!size_supported(N)) return 0;
  return precision == VB_FP32_MIXED ? sched_count_of<float>(N) : sched_count_of<double>(N);
}
}  // namespace

int sched_selected(bool fp32, int N) {
  const int i = size_index(N);
  if (i < 0) return 0;
  std::lock_guard<std::mutex> l(g_sel_mu);
  return g_sel[fp32 ? 1 : 0][i];
}

}  // namespace vb

extern "C" {

int vb_schedule_count(int n, int precision) {
  if (precision != VB_FP64 && precision != VB_FP32_MIXED) return 0;
  return vb::count_for(n, precision);
}

int vb_schedule_get(int n, int precision) {
  if (vb::count_for(n, precision) == 0) return -1;
  return vb::sched_selected(precision == VB_FP32_MIXED, n);
}

int vb_schedule_set(int n, int precision, int sched) {
  if (precision != VB_FP64 && precision != VB_FP32_MIXED) return VB_ERR_PARAM;
  const int c = vb::count_for(n, precision);
  if (sched < 0 || sched >= c) return VB_ERR_PARAM;
  std::lock_guard<std::mutex> l(vb::g_sel_mu);
  vb::g_sel[precision == VB_FP32_MIXED ? 1 : 0][vb::size_index(n)] = sched;
  return VB_OK;
}

int vb_schedule_describe(int n, int precision, int sched, int* z_tile, int* x_passes, int* x_radix) {
  if (precision != VB_FP64 && precision != VB_FP32_MIXED) return VB_ERR_PARAM;
  if (sched < 0 || sched >= vb::count_for(n, precision)) return VB_ERR_PARAM;
  int zt, xp, xr;
  if (precision == VB_FP32_MIXED) {
    const vb::SizeOps<float>* o = vb::get_ops<float>(n, sched);
    zt = o->ytile, xp = o->xr_passes, xr = o->xr_radix;
  } else {
    const vb::SizeOps<double>* o = vb::get_ops<double>(n, sched);
    zt = o->ytile, xp = o->xr_passes, xr = o->xr_radix;
  }
  if (z_tile) *z_tile = zt;
  if (x_passes) *x_passes = xp;
  if (x_radix) *x_radix = xr;
  return VB_OK;
}

int vb_autotune(int n, int precision, int steps, int device, vb_tune_report* rep) {
  const int count = vb_schedule_count(n, precision);
  if (count == 0 || steps < 1 || !rep) return VB_ERR_PARAM;
  const double L = 2.0 * 3.1415926535897932384626433832795;
  vb_params p;
  vb_params_default(&p);
  p.nx = p.ny = p.nz = n;
  p.precision = precision;
  p.eps = 2.0 * L / n;
  p.M1 = p.M2 = 1e4;
  p.alpha = p.beta = 0.0;
  p.phi_lut_mode = VB_LUT_ACCURATE;
  p.time_step = p.fixed_time_step = 5e-7;
  p.device = device;
  std::vector<double> phi((size_t)n * n * n);
  int rc = vb_shape(3, 20260003ULL, n, n, 1.0, p.eps, 1e-3, phi.data(), nullptr);   // torus
  if (rc != VB_OK) return rc;
  if (hipSetDevice(device) != hipSuccess) return VB_ERR_NO_DEVICE;
  *rep = vb_tune_report{};
  rep->n = n;
  rep->precision = precision;
  rep->steps = steps;
  rep->candidates = count;
  int best = 0;
  for (int sc = 0; sc < count; ++sc) {
    vb_solver* s = nullptr;
    vb::g_sched_override = sc;
    rc = vb_create(&p, phi.data(), &s);
    vb::g_sched_override = -1;
    if (rc != VB_OK) return rc;
    vb_step_report r;
    rc = vb_step(s, 2, &r);   // driver start, graph capture
    double t = INFINITY;
    for (int rep_i = 0; rc == VB_OK && rep_i < 3; ++rep_i) {
      if (hipDeviceSynchronize() != hipSuccess) { rc = VB_ERR_HIP; break; }
      const auto t0 = std::chrono::steady_clock::now();
      rc = vb_step(s, steps, &r);
      if (rc == VB_OK && hipDeviceSynchronize() != hipSuccess) rc = VB_ERR_HIP;
      const auto t1 = std::chrono::steady_clock::now();
      t = std::fmin(t, std::chrono::duration<double, std::milli>(t1 - t0).count() / steps);
    }
    vb_destroy(s);
    if (rc != VB_OK) return rc;
    rep->ms_per_step[sc] = t;
    if (t < rep->ms_per_step[best]) best = sc;
  }
  rep->chosen = best;
  return vb_schedule_set(n, precision, best);
}

}  // extern "C"
