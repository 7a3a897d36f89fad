// vb_batch.hip -- many independent vesicle solves on one GPU.
//
// Every solve keeps its state, work buffers and coefficient tables resident in
// HBM for the life of the batch.  Fixed-dt solves of the same grid, precision,
// scheme and drift mode form a size group: their pipelines issue the same
// launches, so each launch is issued once for the whole group (grid.z runs
// over the solves, each solve's argument struct read from a device table),
// and the launches of two driver iterations are captured as one HIP graph per
// group.  A 32^3 solve alone cannot fill the GPU with one kernel; sixteen of
// them in one launch can.  Groups (and solves that cannot be grouped) are
// spread over four streams, each with a hardware queue of its own, by
// longest-processing-time assignment; a stream runs its groups one after the
// other, all steps of one before the next, so the group's working set stays
// cached between steps.  Solves with adaptive dt need a host decision every
// iteration; they step after the fixed-dt work has been issued and overlap
// with it.
#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <numeric>
#include <vector>

#include "vb_host.hpp"
#include "vb_multi.hpp"
#include "vb_solver.hpp"

namespace {
// Four streams: measured best for mixed sizes (more concurrent hardware
// queues slow every size down).
constexpr int kBatchStreams = 4;
// Every size is grouped.  Small grids gain most (a 64^3 kernel cannot fill
// the GPU), and for large ones one group per stream keeps fewer large working
// sets competing for the caches than solves running side by side on four
// streams (measured on the 48-solve mix: 8.1e9 grid-point-steps/s grouped,
// 7.7e9 with 128^3 and 192^3 ungrouped).

// one batched launch sequence over the solves of a size group
struct Group {
  std::vector<int> lanes;
  int stream = 0;
  hipGraphExec_t gx = nullptr;
  std::vector<std::unique_ptr<vb::DBuf<unsigned char>>> tabs;
  ~Group() {
    if (gx) (void)hipGraphExecDestroy(gx);
  }
};

// Size groups of several grid lengths whose ETD stage combines share
// launches (vb_batch_set_shared_stages): each group's other kernels stay one
// launch per group on the group's stream, and each stage combine is one launch
// over the solves of every member (vb_multi.hip), the member streams joined
// before it and forked after it, all in one graph of two iterations.
struct Shared {
  std::vector<Group*> members;
  hipGraphExec_t gx = nullptr;
  std::vector<std::unique_ptr<vb::DBuf<unsigned char>>> tabs;
  std::vector<hipEvent_t> events;           // the captured joins and forks
  std::map<hipStream_t, hipEvent_t> order;  // per member stream: ordering around each graph launch
  void clear() {
    if (gx) (void)hipGraphExecDestroy(gx);
    gx = nullptr;
    tabs.clear();
    for (auto e : events) (void)hipEventDestroy(e);
    events.clear();
  }
  ~Shared() {
    clear();
    for (auto& kv : order) (void)hipEventDestroy(kv.second);
  }
};
}  // namespace

struct vb_batch {
  std::vector<std::unique_ptr<vb::SolverBase>> solves;
  std::vector<std::unique_ptr<Group>> groups;   // a solve that cannot be grouped is a group of one
  std::vector<hipStream_t> streams;
  std::vector<vb_step_report> reports;    // of the last vb_batch_step
  Shared shared;                          // empty unless shared stage launches are on
  ~vb_batch() {
    for (auto s : streams) (void)hipStreamSynchronize(s);
    shared.clear();
    groups.clear();
    solves.clear();
    for (auto s : streams) {
      (void)hipStreamSynchronize(s);
      (void)hipStreamDestroy(s);
    }
  }

  // Records two iterations of every lane and captures their batched launches
  // as one graph.  False if a lane is not in its recurring state yet or the
  // lanes' launch sequences differ.
  bool build(Group& g) {
    if (g.gx) {
      (void)hipGraphExecDestroy(g.gx);
      g.gx = nullptr;
    }
    g.tabs.clear();
    const size_t L = g.lanes.size();
    std::vector<vb::Recorder> rec(L);
    for (size_t l = 0; l < L; ++l)
      if (solves[(size_t)g.lanes[l]]->record_pair(rec[l]) == 0) return false;
    const auto& s0 = rec[0].seq;
    for (size_t l = 1; l < L; ++l) {
      const auto& sl = rec[l].seq;
      if (sl.size() != s0.size()) return false;
      for (size_t k = 0; k < s0.size(); ++k) {
        if (sl[k].op != s0[k].op || sl[k].launch_tab != s0[k].launch_tab || sl[k].args.size() != s0[k].args.size() ||
            sl[k].grid.x != s0[k].grid.x || sl[k].grid.y != s0[k].grid.y || sl[k].block.x != s0[k].block.x)
          return false;
      }
    }
    // one device table per recorded launch: the lanes' argument structs
    for (size_t k = 0; k < s0.size(); ++k) {
      const size_t a = s0[k].args.size();
      std::vector<unsigned char> h(a * L);
      for (size_t l = 0; l < L; ++l) std::copy(rec[l].seq[k].args.begin(), rec[l].seq[k].args.end(), h.begin() + l * a);
      auto t = std::make_unique<vb::DBuf<unsigned char>>();
      t->alloc(h.size());
      VB_HIP(hipMemcpy(t->p, h.data(), h.size(), hipMemcpyHostToDevice));
      g.tabs.push_back(std::move(t));
    }
    hipStream_t st = streams[(size_t)g.stream];
    VB_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    for (size_t k = 0; k < s0.size(); ++k)
      s0[k].launch_tab(s0[k].op, g.tabs[k]->p, s0[k].grid, s0[k].block, (int)L, st);
    hipGraph_t gr = nullptr;
    VB_HIP(hipStreamEndCapture(st, &gr));
    const hipError_t e = hipGraphInstantiateWithFlags(&g.gx, gr, 0);
    (void)hipGraphDestroy(gr);
    if (e != hipSuccess) {
      g.gx = nullptr;
      throw vb::HipError{e, "hipGraphInstantiateWithFlags"};
    }
    return true;
  }

  // The member groups of the shared stage launches: fixed-dt FP32-mixed
  // ETDRK4 groups whose length a shared launch carries, with the default
  // column tile, of the drift mode with the most points; none unless there
  // are at least two such groups.
  void choose_shared() {
    shared.clear();
    shared.members.clear();
    std::map<int, std::vector<Group*>> by_drift;
    std::map<int, long long> pts;
    for (auto& g : groups) {
      const vb::SolverBase& s0 = *solves[(size_t)g->lanes[0]];
      const vb::GroupKey k = s0.group_key();
      if (!k.ok || k.precision != VB_FP32_MIXED || k.scheme != VB_ETDRK4 || k.nx != k.nz ||
          !vb::zstage_multi_size(k.nx, s0.z_tile()))
        continue;
      by_drift[k.drift].push_back(g.get());
      pts[k.drift] += (long long)g->lanes.size() * k.nx * k.nx * k.nz;
    }
    int best = -1;
    for (auto& kv : by_drift)
      if (kv.second.size() >= 2 && (best < 0 || pts[kv.first] > pts[best])) best = kv.first;
    if (best >= 0) shared.members = by_drift[best];
  }

  // the distinct streams of the shared members, each with its ordering event
  std::vector<hipStream_t> member_streams() {
    std::vector<hipStream_t> v;
    for (auto* g : shared.members) {
      hipStream_t s = streams[(size_t)g->stream];
      if (std::find(v.begin(), v.end(), s) == v.end()) v.push_back(s);
      if (!shared.order.count(s)) {
        hipEvent_t e = nullptr;
        VB_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        shared.order[s] = e;
      }
    }
    return v;
  }

  bool is_member(const Group* g) const {
    return std::find(shared.members.begin(), shared.members.end(), g) != shared.members.end();
  }

  hipEvent_t new_event() {
    hipEvent_t e = nullptr;
    VB_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    shared.events.push_back(e);
    return e;
  }

  // Records two iterations of every lane of every member and captures them:
  // per-group table launches on the groups' streams, one shared launch per
  // stage combine.  False if a lane is not in its recurring state yet or the
  // launch sequences do not line up.
  bool build_shared() {
    shared.clear();
    auto& M = shared.members;
    std::vector<std::vector<vb::Recorder>> rec(M.size());
    for (size_t m = 0; m < M.size(); ++m) {
      rec[m].resize(M[m]->lanes.size());
      for (size_t l = 0; l < rec[m].size(); ++l)
        if (solves[(size_t)M[m]->lanes[l]]->record_pair(rec[m][l]) == 0) return false;
    }
    const auto& s0 = rec[0][0].seq;
    for (size_t m = 0; m < M.size(); ++m) {
      const auto& g0 = rec[m][0].seq;
      for (size_t l = 0; l < rec[m].size(); ++l) {
        const auto& sl = rec[m][l].seq;
        if (sl.size() != s0.size()) return false;
        for (size_t k = 0; k < s0.size(); ++k) {
          if (sl[k].op != s0[k].op) return false;   // every lane of every member: the same ops in order
          if (sl[k].launch_tab != g0[k].launch_tab || sl[k].args.size() != g0[k].args.size() ||
              sl[k].grid.x != g0[k].grid.x || sl[k].grid.y != g0[k].grid.y || sl[k].block.x != g0[k].block.x)
            return false;   // within a group: the same launches
        }
      }
    }
    auto upload = [&](const std::vector<unsigned char>& h) {
      auto t = std::make_unique<vb::DBuf<unsigned char>>();
      t->alloc(h.size());
      VB_HIP(hipMemcpy(t->p, h.data(), h.size(), hipMemcpyHostToDevice));
      shared.tabs.push_back(std::move(t));
      return shared.tabs.back()->p;
    };
    // the tables: per member and launch for the group launches; per stage
    // combine, the argument structs, first column groups and lengths of all
    // solves of all members
    std::vector<std::vector<unsigned char*>> gtab(M.size(), std::vector<unsigned char*>(s0.size(), nullptr));
    std::vector<vb::ZsMultiArgs> margs(s0.size());
    std::vector<unsigned> mblocks(s0.size(), 0);
    for (size_t k = 0; k < s0.size(); ++k) {
      if (vb::zstage_multi_op(s0[k].op)) {
        std::vector<unsigned char> args;
        std::vector<int> first, len;
        int acc = 0;
        for (size_t m = 0; m < M.size(); ++m)
          for (size_t l = 0; l < rec[m].size(); ++l) {
            const auto& x = rec[m][l].seq[k];
            if (x.block.x != 256 || x.grid.y != 1) return false;
            args.insert(args.end(), x.args.begin(), x.args.end());
            first.push_back(acc);
            len.push_back(solves[(size_t)M[m]->lanes[l]]->p.nx);
            acc += (int)x.grid.x;
          }
        std::vector<unsigned char> hf(first.size() * sizeof(int)), hn(len.size() * sizeof(int));
        std::memcpy(hf.data(), first.data(), hf.size());
        std::memcpy(hn.data(), len.data(), hn.size());
        margs[k].args = upload(args);
        margs[k].first = reinterpret_cast<const int*>(upload(hf));
        margs[k].n = reinterpret_cast<const int*>(upload(hn));
        margs[k].nsolve = (int)first.size();
        mblocks[k] = (unsigned)acc;
      } else {
        for (size_t m = 0; m < M.size(); ++m) {
          const size_t a = rec[m][0].seq[k].args.size();
          std::vector<unsigned char> h(a * rec[m].size());
          for (size_t l = 0; l < rec[m].size(); ++l)
            std::copy(rec[m][l].seq[k].args.begin(), rec[m][l].seq[k].args.end(), h.begin() + l * a);
          gtab[m][k] = upload(h);
        }
      }
    }
    hipStream_t main = streams[(size_t)M[0]->stream];
    std::vector<hipStream_t> others;
    for (auto* g : M) {
      hipStream_t s = streams[(size_t)g->stream];
      if (s != main && std::find(others.begin(), others.end(), s) == others.end()) others.push_back(s);
    }
    auto join = [&] {
      for (hipStream_t s : others) {
        hipEvent_t e = new_event();
        VB_HIP(hipEventRecord(e, s));
        VB_HIP(hipStreamWaitEvent(main, e, 0));
      }
    };
    auto fork = [&] {
      hipEvent_t e = new_event();
      VB_HIP(hipEventRecord(e, main));
      for (hipStream_t s : others) VB_HIP(hipStreamWaitEvent(s, e, 0));
    };
    VB_HIP(hipStreamBeginCapture(main, hipStreamCaptureModeThreadLocal));
    hipGraph_t gr = nullptr;
    try {
      fork();
      for (size_t k = 0; k < s0.size(); ++k) {
        if (vb::zstage_multi_op(s0[k].op)) {
          join();
          vb::launch_zstage_multi(s0[k].op, margs[k], mblocks[k], main);
          fork();
        } else {
          for (size_t m = 0; m < M.size(); ++m) {
            const auto& x = rec[m][0].seq[k];
            x.launch_tab(x.op, gtab[m][k], x.grid, x.block, (int)rec[m].size(), streams[(size_t)M[m]->stream]);
          }
        }
      }
      join();
    } catch (...) {
      (void)hipStreamEndCapture(main, &gr);
      if (gr) (void)hipGraphDestroy(gr);
      throw;
    }
    VB_HIP(hipStreamEndCapture(main, &gr));
    const hipError_t e = hipGraphInstantiateWithFlags(&shared.gx, gr, 0);
    (void)hipGraphDestroy(gr);
    if (e != hipSuccess) {
      shared.gx = nullptr;
      throw vb::HipError{e, "hipGraphInstantiateWithFlags"};
    }
    return true;
  }

  // up to `c` more iterations of every lane of the shared members
  int advance_shared(int c) {
    bool ok = true, rec = shared.gx != nullptr;
    for (auto* g : shared.members)
      for (int i : g->lanes) {
        ok = ok && reports[(size_t)i].status == VB_OK;
        rec = rec && solves[(size_t)i]->at_recorded_state();
      }
    if (ok && c >= 2 && (rec || build_shared())) {
      // The graph is launched on the first member's stream, but the members'
      // own launches (the per-solve steps before the pipelines recur, an odd
      // last step) run on their groups' streams: the graph waits for every
      // member stream, and every member stream waits for the graph.
      hipStream_t main = streams[(size_t)shared.members[0]->stream];
      for (hipStream_t s : member_streams())
        if (s != main) {
          VB_HIP(hipEventRecord(shared.order[s], s));
          VB_HIP(hipStreamWaitEvent(main, shared.order[s], 0));
        }
      VB_HIP(hipGraphLaunch(shared.gx, main));
      VB_HIP(hipEventRecord(shared.order[main], main));
      for (hipStream_t s : member_streams())
        if (s != main) VB_HIP(hipStreamWaitEvent(s, shared.order[main], 0));
      for (auto* g : shared.members)
        for (int i : g->lanes) solves[(size_t)i]->advance_recorded(&reports[(size_t)i]);
      return 2;
    }
    // lanes not yet in their recurring state: per-solve launches
    for (auto* g : shared.members)
      for (int i : g->lanes)
        if (reports[(size_t)i].status == VB_OK) solves[(size_t)i]->launch_steps(1, &reports[(size_t)i]);
    return 1;
  }

  bool ready(const Group& g) const {
    if (!g.gx) return false;
    for (int i : g.lanes)
      if (!solves[(size_t)i]->at_recorded_state()) return false;
    return true;
  }

  // up to `c` more iterations of every lane of g; returns how many were issued
  int advance(Group& g, int c) {
    bool ok = true;
    for (int i : g.lanes) ok = ok && reports[(size_t)i].status == VB_OK;
    if (ok && g.lanes.size() > 1 && c >= 2 && (ready(g) || build(g))) {
      VB_HIP(hipGraphLaunch(g.gx, streams[(size_t)g.stream]));
      for (int i : g.lanes) solves[(size_t)i]->advance_recorded(&reports[(size_t)i]);
      return 2;
    }
    // a lone solve, or lanes not yet in their recurring state: per-solve launches
    const int c1 = g.lanes.size() > 1 ? 1 : c;
    for (int i : g.lanes)
      if (reports[(size_t)i].status == VB_OK) solves[(size_t)i]->launch_steps(c1, &reports[(size_t)i]);
    return c1;
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
    const int rc = vb::validate_params(&p);
    if (rc != VB_OK) return rc;
    if (!phi0[i] || p.device != params[0].device) return VB_ERR_PARAM;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= params[0].device) return VB_ERR_NO_DEVICE;
  return bguard([&] {
    auto b = std::make_unique<vb_batch>();
    VB_HIP(hipSetDevice(params[0].device));
    const int ns = std::min(n, kBatchStreams);
    for (int k = 0; k < ns; ++k) b->streams.push_back(vb::make_queue_stream());
    // groups: fixed-dt solves with the same key; everything else alone
    std::map<vb::GroupKey, int> gi;
    std::vector<int> group_of(n);
    for (int i = 0; i < n; ++i) {
      const vb_params& p = params[i];
      vb::GroupKey k{};
      k.ok = (p.fixed_time_step > 0 && p.scheme != VB_EULER && p.graphs) ? 1 : 0;
      k.nx = p.nx;
      k.nz = p.nz;
      k.precision = p.precision;
      k.scheme = p.scheme;
      k.drift = ((p.M1 != 0 || p.M2 != 0) && p.penalty_drift) ? 1 : 0;
      if (k.ok) {
        auto it = gi.find(k);
        if (it == gi.end()) {
          it = gi.emplace(k, (int)b->groups.size()).first;
          b->groups.push_back(std::make_unique<Group>());
        }
        group_of[i] = it->second;
      } else {
        group_of[i] = (int)b->groups.size();
        b->groups.push_back(std::make_unique<Group>());
      }
      b->groups[(size_t)group_of[i]]->lanes.push_back(i);
    }
    // longest-processing-time assignment of groups to streams
    std::vector<int> gorder(b->groups.size());
    std::iota(gorder.begin(), gorder.end(), 0);
    auto gpts = [&](int g) {
      long long s = 0;
      for (int i : b->groups[(size_t)g]->lanes) s += points(params[i]);
      return s;
    };
    std::stable_sort(gorder.begin(), gorder.end(), [&](int a, int c) { return gpts(a) > gpts(c); });
    std::vector<long long> load((size_t)ns, 0);
    for (int g : gorder) {
      const int k = int(std::min_element(load.begin(), load.end()) - load.begin());
      b->groups[(size_t)g]->stream = k;
      load[(size_t)k] += gpts(g);
    }
    for (int i = 0; i < n; ++i) {
      b->solves.push_back(vb::make_solver(params[i], b->streams[(size_t)b->groups[(size_t)group_of[i]]->stream]));
      b->solves.back()->set_phi(phi0[i]);
    }
    // issue order: largest groups first
    std::vector<std::unique_ptr<Group>> sorted;
    for (int g : gorder) sorted.push_back(std::move(b->groups[(size_t)g]));
    b->groups = std::move(sorted);
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
    // Fixed-dt groups, group-major on each stream; the host issues a chunk
    // of steps per stream in turn, which keeps every stream fed without
    // filling one stream's queue while the others wait.
    constexpr int chunk = 16;
    const size_t ns = b->streams.size();
    // the members of the shared stage launches go first, all their steps
    if (!b->shared.members.empty())
      for (int done = 0; done < n_steps;) done += b->advance_shared(n_steps - done);
    std::vector<std::vector<Group*>> lists(ns);
    for (auto& g : b->groups)
      if (b->solves[(size_t)g->lanes[0]]->fixed_dt() && !b->is_member(g.get()))
        lists[(size_t)g->stream].push_back(g.get());
    std::vector<size_t> cur(ns, 0);
    std::vector<int> done(ns, 0);
    for (bool any = true; any;) {
      any = false;
      for (size_t k = 0; k < ns; ++k) {
        if (cur[k] >= lists[k].size()) continue;
        any = true;
        Group& g = *lists[k][cur[k]];
        const int target = std::min(n_steps, done[k] + chunk);
        while (done[k] < target) done[k] += b->advance(g, target - done[k]);
        if (done[k] >= n_steps) {
          ++cur[k];
          done[k] = 0;
        }
      }
    }
    // adaptive solves decide on the host every iteration
    for (size_t i = 0; i < b->solves.size(); ++i)
      if (!b->solves[i]->fixed_dt()) b->solves[i]->step(n_steps, &b->reports[i]);
    for (size_t i = 0; i < b->solves.size(); ++i)
      if (b->solves[i]->fixed_dt()) b->solves[i]->finish_steps(&b->reports[i]);
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

int vb_batch_set_shared_stages(vb_batch* b, int on, int* members) {
  if (!b) return VB_ERR_PARAM;
  return bguard([&] {
    for (auto s : b->streams) VB_HIP(hipStreamSynchronize(s));
    if (on) {
      b->choose_shared();
    } else {
      b->shared.clear();
      b->shared.members.clear();
    }
    if (members) *members = (int)b->shared.members.size();
    return (int)VB_OK;
  });
}

void vb_batch_destroy(vb_batch* b) { delete b; }

}  // extern "C"
