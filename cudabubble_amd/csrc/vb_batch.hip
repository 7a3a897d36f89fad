// vb_batch.hip -- many independent vesicle solves on one GPU.
//
// Every solve keeps its state, work buffers and coefficient tables resident in
// HBM for the life of the batch.  Solves are spread over a few streams by
// longest-processing-time assignment (grid points per stream balanced).  A
// batch step runs each stream's solves one after the other, all steps of one
// solve before the next (its working set stays cached between steps), the
// streams concurrently; fixed-dt iterations replay the solve's captured graph,
// two iterations per replay, so the host issues one graph launch where the
// eager driver would issue about sixty kernel launches per iteration.  Solves with adaptive dt need
// a host decision every iteration; they step after the fixed-dt work has been
// issued and overlap with it.
#include <algorithm>
#include <memory>
#include <numeric>
#include <vector>

#include "vb_host.hpp"
#include "vb_solver.hpp"

namespace {
// Four streams, each with a hardware queue of its own: measured best for
// mixed sizes (more concurrent solves thrash the caches).
constexpr int kBatchStreams = 4;
}

struct vb_batch {
  std::vector<std::unique_ptr<vb::SolverBase>> solves;
  std::vector<int> order;                 // decreasing grid size
  std::vector<hipStream_t> streams;
  std::vector<int> stream_of;             // stream index of each solve
  std::vector<vb_step_report> reports;    // of the last vb_batch_step
  ~vb_batch() {
    solves.clear();
    for (auto s : streams) {
      (void)hipStreamSynchronize(s);
      (void)hipStreamDestroy(s);
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

long long points(const vb_params& p) { return (long long)p.nx * p.nx * p.nz; }
}  // namespace

extern "C" {

int vb_batch_create(int n, const vb_params* params, const double* const* phi0, vb_batch** out) {
  if (!out || n <= 0 || !params || !phi0) return VB_ERR_PARAM;
  *out = nullptr;
  for (int i = 0; i < n; ++i) {
    const vb_params& p = params[i];
    if (p.nx <= 0 || p.nz <= 0 || !(p.eps > 0) || !(p.len > 0) || !phi0[i]) return VB_ERR_PARAM;
    if (p.nx % 2) return VB_ERR_ODD_NX;
    if (!vb::size_supported(p.nx) || !vb::size_supported(p.nz)) return VB_ERR_PARAM;
    if (p.device != params[0].device) return VB_ERR_PARAM;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= params[0].device) return VB_ERR_NO_DEVICE;
  return bguard([&] {
    auto b = std::make_unique<vb_batch>();
    VB_HIP(hipSetDevice(params[0].device));
    const int ns = std::min(n, kBatchStreams);
    for (int k = 0; k < ns; ++k) {
      b->streams.push_back(vb::make_queue_stream());
    }
    b->order.resize(n);
    std::iota(b->order.begin(), b->order.end(), 0);
    std::stable_sort(b->order.begin(), b->order.end(),
                     [&](int a, int c) { return points(params[a]) > points(params[c]); });
    b->stream_of.assign(n, 0);
    std::vector<long long> load(ns, 0);
    for (int i : b->order) {
      const int k = int(std::min_element(load.begin(), load.end()) - load.begin());
      b->stream_of[i] = k;
      load[k] += points(params[i]);
    }
    for (int i = 0; i < n; ++i) {
      b->solves.push_back(vb::make_solver(params[i], b->streams[b->stream_of[i]]));
      b->solves.back()->set_phi(phi0[i]);
    }
    b->reports.assign(n, vb_step_report{});
    *out = b.release();
    return (int)VB_OK;
  });
}

int vb_batch_step(vb_batch* b, int n_steps) {
  if (!b || n_steps < 0) return VB_ERR_PARAM;
  return bguard([&] {
    for (auto& r : b->reports) {
      r = vb_step_report{};
      r.energy = NAN;
    }
    // Fixed-dt solves, solve-major on each stream: a stream runs all n steps
    // of one solve before the next, so that solve's working set stays in the
    // caches from one step to the next.  The host issues a chunk of steps per
    // stream in turn, which keeps every stream fed without filling one
    // stream's queue while the others wait.
    constexpr int chunk = 16;
    const size_t ns = b->streams.size();
    std::vector<std::vector<int>> lists(ns);
    for (int i : b->order)
      if (b->solves[(size_t)i]->fixed_dt()) lists[(size_t)b->stream_of[(size_t)i]].push_back(i);
    std::vector<size_t> cur(ns, 0);
    std::vector<int> done(ns, 0);
    for (bool any = true; any;) {
      any = false;
      for (size_t k = 0; k < ns; ++k) {
        if (cur[k] >= lists[k].size()) continue;
        any = true;
        const int i = lists[k][cur[k]];
        const int c = std::min(chunk, n_steps - done[k]);
        if (c > 0 && b->reports[(size_t)i].status == VB_OK) b->solves[(size_t)i]->launch_steps(c, &b->reports[(size_t)i]);
        done[k] += std::max(c, 0);
        if (done[k] >= n_steps || b->reports[(size_t)i].status != VB_OK) {
          ++cur[k];
          done[k] = 0;
        }
      }
    }
    // adaptive solves decide on the host every iteration
    for (int i : b->order) {
      auto& s = b->solves[(size_t)i];
      if (!s->fixed_dt()) s->step(n_steps, &b->reports[(size_t)i]);
    }
    for (int i : b->order) {
      auto& s = b->solves[(size_t)i];
      if (s->fixed_dt()) s->finish_steps(&b->reports[(size_t)i]);
    }
    for (auto s : b->streams) VB_HIP(hipStreamSynchronize(s));
    return (int)VB_OK;
  });
}

int vb_batch_report(vb_batch* b, int i, vb_step_report* rep) {
  if (!b || i < 0 || i >= (int)b->solves.size() || !rep) return VB_ERR_PARAM;
  *rep = b->reports[(size_t)i];
  return VB_OK;
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
