// vb_batch.hip -- many independent vesicle solves on one GPU.
//
// Every solve keeps its own resident state and work buffers in HBM; all solves
// of a batch launch on one shared stream, in size-sorted order, so consecutive
// launches of small grids run back to back on the device without host
// synchronisation.
#include <algorithm>
#include <memory>
#include <numeric>
#include <vector>

#include "vb_host.hpp"
#include "vb_solver.hpp"

struct vb_batch {
  std::vector<std::unique_ptr<vb::SolverBase>> solves;
  std::vector<int> order;
  hipStream_t st = nullptr;
  ~vb_batch() {
    solves.clear();
    if (st) {
      (void)hipStreamSynchronize(st);
      (void)hipStreamDestroy(st);
    }
  }
};

namespace {
template <class F>
int bguard(F&& f) {
  try {
    return f();
  } catch (const vb::HipError& e) {
    vb::set_last_error(std::string(hipGetErrorString(e.err)) + " in " + e.where);
    return e.err == hipErrorOutOfMemory ? VB_ERR_OOM : VB_ERR_HIP;
  } catch (const std::bad_alloc&) {
    return VB_ERR_OOM;
  }
}
}  // namespace

extern "C" {

int vb_batch_create(int n, const vb_params* params, const double* const* phi0, vb_batch** out) {
  if (!out || n <= 0 || !params || !phi0) return VB_ERR_PARAM;
  *out = nullptr;
  for (int i = 0; i < n; ++i) {
    const vb_params& p = params[i];
    if (p.nx <= 0 || p.nz <= 0 || (p.nx % 2) || !(p.eps > 0) || !(p.len > 0) || !phi0[i]) return VB_ERR_PARAM;
    if (!vb::size_supported(p.nx) || !vb::size_supported(p.nz)) return VB_ERR_PARAM;
  }
  return bguard([&] {
    auto b = std::make_unique<vb_batch>();
    VB_HIP(hipSetDevice(params[0].device));
    VB_HIP(hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking));
    for (int i = 0; i < n; ++i) {
      b->solves.push_back(vb::make_solver(params[i], b->st));
      b->solves.back()->set_phi(phi0[i]);
    }
    b->order.resize(n);
    std::iota(b->order.begin(), b->order.end(), 0);
    std::stable_sort(b->order.begin(), b->order.end(), [&](int a, int c) {
      return (long long)params[a].nx * params[a].nx * params[a].nz >
             (long long)params[c].nx * params[c].nx * params[c].nz;
    });
    *out = b.release();
    return (int)VB_OK;
  });
}

int vb_batch_step(vb_batch* b, int n_steps) {
  if (!b || n_steps < 0) return VB_ERR_PARAM;
  return bguard([&] {
    for (int i : b->order) {
      vb_step_report r{};
      b->solves[(size_t)i]->step(n_steps, &r);
    }
    VB_HIP(hipStreamSynchronize(b->st));
    return (int)VB_OK;
  });
}

int vb_batch_query(vb_batch* b, int i, double* E, double* V, double* A, double* Eb) {
  if (!b || i < 0 || i >= (int)b->solves.size()) return VB_ERR_PARAM;
  return bguard([&] {
    b->solves[(size_t)i]->query(E, V, A, Eb);
    return (int)VB_OK;
  });
}

int vb_batch_get_phi(vb_batch* b, int i, double* host_out) {
  if (!b || i < 0 || i >= (int)b->solves.size() || !host_out) return VB_ERR_PARAM;
  return bguard([&] {
    b->solves[(size_t)i]->get_phi(host_out);
    return (int)VB_OK;
  });
}

void vb_batch_destroy(vb_batch* b) { delete b; }

}  // extern "C"
