// vb_server.hip -- a request loop for vesicle solves on one GPU.
//
// Clients submit solves (parameters + initial field) at any time; one worker
// thread owns the device.  It admits pending requests in arrival order while
// the resident footprint fits the memory budget and a slot is free, keeps
// every admitted solve's state, work buffers and coefficient tables resident
// in HBM, and interleaves the solves' adaptive iterations: each attempt is
// decided on the host when its completion event has fired.  Larger solves
// keep one attempt each in flight on one of a few streams and are re-issued
// as soon as they are decided.  Small solves of one kind (grid, precision,
// scheme) run on one stream and advance in rounds: once none of a kind's
// attempts is in flight, the next attempt of every one of them is issued as
// shared launches, one per pipeline kernel, over all of them (grid.z runs
// over the solves, each solve's argument struct read from a device table),
// so that many small grids fill the GPU together.  A solve leaves when its energy
// has dropped by less than a relative tolerance over a window of accepted
// steps, when t reaches tmax, at a driver exit, or at a step cap.  Requests
// are independent: each computes exactly what vb_step computes for it alone,
// whatever else is running.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "vb_host.hpp"
#include "vb_solver.hpp"

namespace {

struct Request {
  long long id;
  vb_params p;
  std::vector<double> phi0;
  double t_submit;
};

struct Running {
  Request req;
  std::unique_ptr<vb::SolverBase> s;
  int stream;
  size_t bytes;
  double t_admit;
  bool inflight = false;     // an attempt was issued and is not decided yet
  vb_step_report rep{};
  std::deque<double> hist;   // energies of the last window+1 accepted states
};

// the kind of a solve: solves of one kind share a stream and launches
vb::GroupKey kind_of(const vb_params& p) {
  vb::GroupKey k{};
  k.ok = 1;
  k.nx = p.nx;
  k.nz = p.nz;
  k.precision = p.precision;
  k.scheme = p.scheme;
  return k;
}

// Kinds whose solves share launches: grids of at most 64^3 points, which one
// solve's launches cannot fill the GPU with.  Larger solves fill it alone,
// and one launch over several of them streams several working sets through
// the caches at once (measured in the batched form: 4 x 192^3 in shared
// launches ran 14 % slower than one launch per solve); they interleave
// attempt by attempt over the streams instead.
constexpr long long kSharedMaxPoints = 64ll * 64 * 64;
bool shared_kind(const vb_params& p) { return (long long)p.nx * p.nx * p.nz <= kSharedMaxPoints; }

// a reused solver keeps its stream, which is its kind's
bool same_kind(const vb_params& a, const vb_params& b) {
  const vb::GroupKey x = kind_of(a), y = kind_of(b);
  return !(x < y) && !(y < x);
}

// The argument tables of one launch sequence's shared launches (pinned
// staging and its device copy, each launch's table at a fixed offset, room for
// max_active solves) and its captured launches, one graph per number of
// solves.  A kind's next round is issued only after every attempt of its
// previous round has completed, so the tables are free when rewritten.
struct SeqShared {
  vb::DBuf<unsigned char> dev;
  unsigned char* host = nullptr;
  size_t bytes = 0;
  std::vector<size_t> off;
  std::map<int, hipGraphExec_t> graphs;
  ~SeqShared() {
    for (auto& g : graphs)
      if (g.second) (void)hipGraphExecDestroy(g.second);
    if (host) (void)hipHostFree(host);
  }
};

// A finished request's solver, kept with its memory for the next request of
// the same kind and buffer set (vb::SolverBase::reset).
struct Idle {
  std::unique_ptr<vb::SolverBase> s;
  int stream;
  size_t bytes;
  vb_params p;
};

bool same_slot(const vb_params& a, const vb_params& b) {
  const bool wa = a.scheme == VB_EULER || a.shape_adjust, wb = b.scheme == VB_EULER || b.shape_adjust;
  return a.nx == b.nx && a.nz == b.nz && a.precision == b.precision && (wa || !wb);
}

double now_s(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

struct vb_server {
  vb_server_config cfg{};
  std::chrono::steady_clock::time_point t0;
  std::vector<hipStream_t> streams;
  std::vector<long long> stream_load;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<Request> pending;
  std::vector<vb_result> done;
  std::map<long long, std::vector<double>> phis;
  std::vector<Running> running;   // worker thread only
  std::deque<Idle> idle;          // worker thread only: the slot pool, oldest first
  std::map<vb::GroupKey, int> kind_stream;                           // worker thread only
  std::map<std::vector<long long>, std::unique_ptr<SeqShared>> seqs;  // worker thread only
  size_t used = 0;                // bytes of every solver held, running or idle
  int nrun = 0;
  bool stop = false;
  int worker_error = VB_OK;
  std::thread worker;

  ~vb_server() {
    {
      std::lock_guard<std::mutex> g(mu);
      stop = true;
    }
    cv.notify_all();
    if (worker.joinable()) worker.join();
    running.clear();
    idle.clear();
    for (auto s : streams) (void)hipStreamSynchronize(s);
    seqs.clear();
    for (auto s : streams) {
      (void)hipStreamSynchronize(s);
      (void)hipStreamDestroy(s);
    }
  }

  void finish(Running& r) {
    vb_result res{};
    res.id = r.req.id;
    res.status = r.rep.status;
    res.steps = r.rep.steps;
    res.evaluations = r.rep.evaluations;
    res.rejections = r.rep.rejections;
    // the last accepted attempt left the final state's integrals in host
    // memory; a query (a device round trip) only after a driver exit
    if (!r.s->accepted_values(&res.E, &res.V, &res.A, &res.Eb)) r.s->query(&res.E, &res.V, &res.A, &res.Eb);
    r.s->state(nullptr, nullptr, &res.time_step, &res.time_current);
    res.t_submit = r.req.t_submit;
    res.t_admit = r.t_admit;
    std::vector<double> phi;
    if (cfg.keep_phi) {
      phi.resize((size_t)r.req.p.nx * r.req.p.nx * r.req.p.nz);
      r.s->get_phi(phi.data());
    }
    res.t_done = now_s(t0);
    stream_load[(size_t)r.stream] -= (long long)r.req.p.nx * r.req.p.nx * r.req.p.nz;
    idle.push_back(Idle{std::move(r.s), r.stream, r.bytes, r.req.p});   // memory kept for reuse
    std::lock_guard<std::mutex> g(mu);
    nrun--;
    if (cfg.keep_phi) phis[res.id] = std::move(phi);
    done.push_back(res);
  }

  // true when the solve leaves the server after this accepted iteration
  bool converged(Running& r) {
    if (r.rep.status != VB_OK) return true;
    r.hist.push_back(r.rep.energy);
    if ((int)r.hist.size() > cfg.converge_window + 1) r.hist.pop_front();
    if ((int)r.hist.size() == cfg.converge_window + 1) {
      const double e0 = r.hist.front(), e1 = r.hist.back();
      if (fabs(e0 - e1) <= cfg.converge_tol * fabs(e0)) return true;
    }
    if (r.req.p.tmax > 0 && r.rep.time_current >= r.req.p.tmax) return true;
    if (cfg.max_steps > 0 && r.rep.steps >= cfg.max_steps) return true;
    return false;
  }

  // Admission in arrival order.  A request takes an idle solver of its slot
  // kind if there is one (no allocation); otherwise a new one, after idle
  // solvers are released, oldest first, to make room in the memory budget.
  void admit() {
    for (;;) {
      Request q;
      size_t bytes = 0;
      int reuse = -1;
      {
        std::lock_guard<std::mutex> g(mu);
        if (pending.empty() || (int)running.size() >= cfg.max_active) return;
        const vb_params& fp = pending.front().p;
        for (size_t i = 0; i < idle.size() && reuse < 0; ++i)
          if (same_slot(idle[i].p, fp) && same_kind(idle[i].p, fp)) reuse = (int)i;
        if (reuse < 0) {
          bytes = vb::solver_footprint(fp);
          while (used + bytes > cfg.memory_budget && !idle.empty()) {
            used -= idle.front().bytes;
            idle.pop_front();     // releases its device memory
          }
          if (used + bytes > cfg.memory_budget && !running.empty()) return;   // wait for memory
          used += bytes;
        }
        q = std::move(pending.front());
        pending.pop_front();
        nrun++;
      }
      Running r;
      r.t_admit = now_s(t0);
      r.rep.energy = NAN;
      int fail = VB_OK;
      try {
        if (reuse >= 0) {
          Idle it = std::move(idle[(size_t)reuse]);
          idle.erase(idle.begin() + reuse);
          r.stream = it.stream;
          r.bytes = it.bytes;
          r.s = std::move(it.s);
          if (!r.s->reset(q.p, q.phi0.data())) throw std::bad_alloc();   // same_slot guarantees a fit
        } else {
          // small kinds: one stream per kind, the least loaded one when the kind
          // first appears; larger solves: the least loaded stream
          const int least = int(std::min_element(stream_load.begin(), stream_load.end()) - stream_load.begin());
          if (shared_kind(q.p)) {
            auto ks = kind_stream.find(kind_of(q.p));
            if (ks == kind_stream.end()) ks = kind_stream.emplace(kind_of(q.p), least).first;
            r.stream = ks->second;
          } else {
            r.stream = least;
          }
          r.bytes = bytes;
          r.s = vb::make_solver(q.p, streams[(size_t)r.stream]);
          r.s->set_phi(q.phi0.data());
        }
      } catch (const vb::HipError& e) {
        fail = e.err == hipErrorOutOfMemory ? VB_ERR_OOM : VB_ERR_HIP;
      } catch (const std::bad_alloc&) {
        fail = VB_ERR_OOM;
      }
      if (fail != VB_OK) {
        r.s.reset();
        vb_result res{};
        res.id = q.id;
        res.status = fail;
        res.t_submit = q.t_submit;
        res.t_admit = r.t_admit;
        res.t_done = now_s(t0);
        std::lock_guard<std::mutex> g(mu);
        used -= reuse >= 0 ? r.bytes : bytes;
        nrun--;
        done.push_back(res);
        continue;
      }
      stream_load[(size_t)r.stream] += (long long)q.p.nx * q.p.nx * q.p.nz;
      q.phi0.clear();
      q.phi0.shrink_to_fit();
      r.req = std::move(q);
      running.push_back(std::move(r));   // its first attempt goes out with its kind's next round
    }
  }

  // Issues the next attempt of the solves running[idx] (one kind, none in
  // flight): alone through the solver's own graphs, or together as shared
  // launches.  Solves whose pipelines are in different states (an accepted
  // attempt leaves the next state's prefix done, a rejected one does not)
  // record different launch sequences; each sequence goes out as its own
  // shared launches.  Returns the indices of solves that left at a driver
  // exit before their attempt.
  std::vector<size_t> issue_round(const vb::GroupKey& k, const std::vector<size_t>& idx) {
    std::vector<size_t> go, out;
    for (size_t i : idx) {
      if (running[i].s->attempt_prepare(running[i].rep)) go.push_back(i);
      else out.push_back(i);
    }
    if (go.size() == 1) {
      running[go[0]].s->issue_prepared();
      running[go[0]].inflight = true;
    }
    if (go.size() < 2) return out;
    std::vector<vb::Recorder> rec(go.size());
    std::map<std::vector<long long>, std::vector<size_t>> parts;   // sequence -> lanes (into go)
    for (size_t l = 0; l < go.size(); ++l) {
      running[go[l]].s->record_attempt(rec[l]);
      parts[signature(k, running[go[l]].s->stream(), rec[l])].push_back(l);
    }
    for (auto& pv : parts) {
      const auto& lanes = pv.second;
      const auto& s0 = rec[lanes[0]].seq;
      hipStream_t st = running[go[lanes[0]]].s->stream();
      SeqShared& e = seq_shared(pv.first, s0);
      const int L = int(lanes.size());
      for (size_t j = 0; j < s0.size(); ++j) {
        const size_t a = s0[j].args.size();
        for (int m = 0; m < L; ++m) {
          const auto& src = rec[lanes[(size_t)m]].seq[j].args;
          std::copy(src.begin(), src.end(), e.host + e.off[j] + (size_t)m * a);
        }
      }
      hipGraphExec_t& gx = e.graphs[L];
      if (!gx) {
        // the shared launches of this sequence over L solves, captured once:
        // the kernels read their arguments from the sequence's fixed tables
        VB_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        for (size_t j = 0; j < s0.size(); ++j)
          s0[j].launch_tab(s0[j].op, e.dev.p + e.off[j], s0[j].grid, s0[j].block, L, st);
        hipGraph_t gr = nullptr;
        VB_HIP(hipStreamEndCapture(st, &gr));
        const hipError_t err = hipGraphInstantiateWithFlags(&gx, gr, 0);
        (void)hipGraphDestroy(gr);
        if (err != hipSuccess) {
          gx = nullptr;
          throw vb::HipError{err, "hipGraphInstantiateWithFlags"};
        }
      }
      VB_HIP(hipMemcpyAsync(e.dev.p, e.host, e.bytes, hipMemcpyHostToDevice, st));
      VB_HIP(hipGraphLaunch(gx, st));
    }
    for (size_t l = 0; l < go.size(); ++l) {
      running[go[l]].s->attempt_issued();
      running[go[l]].inflight = true;
    }
    return out;
  }

  // the launch sequence of a recorded attempt, with its kind and stream
  static std::vector<long long> signature(const vb::GroupKey& k, hipStream_t st, const vb::Recorder& r) {
    std::vector<long long> v{k.nx, k.nz, k.precision, k.scheme, (long long)(intptr_t)st};
    for (const auto& x : r.seq) {
      v.push_back(x.op);
      v.push_back((long long)(intptr_t)x.launch_tab);
      v.push_back((long long)x.args.size());
      v.push_back(x.grid.x);
      v.push_back(x.grid.y);
      v.push_back(x.block.x);
    }
    return v;
  }

  // the tables and captured graphs of one launch sequence, sized for
  // max_active solves when the sequence first appears
  SeqShared& seq_shared(const std::vector<long long>& sig, const std::vector<vb::RecLaunch>& seq) {
    auto& e = seqs[sig];
    if (!e) {
      e = std::make_unique<SeqShared>();
      e->off.resize(seq.size());
      size_t bytes = 0;
      for (size_t j = 0; j < seq.size(); ++j) {
        e->off[j] = bytes;
        bytes += (seq[j].args.size() * (size_t)cfg.max_active + 255) / 256 * 256;
      }
      e->bytes = bytes;
      VB_HIP(hipHostMalloc((void**)&e->host, bytes, hipHostMallocDefault));
      e->dev.alloc(bytes);
    }
    return *e;
  }

  // the next attempt of every solve that has none in flight: larger solves at
  // once, each alone; small kinds in rounds, once none of the kind's solves
  // has an attempt in flight.  True if anything was issued or left.
  bool issue_rounds() {
    std::map<vb::GroupKey, std::vector<size_t>> ready;
    std::map<vb::GroupKey, bool> busy;
    std::vector<size_t> gone;
    bool any = false;
    for (size_t i = 0; i < running.size(); ++i) {
      const vb::GroupKey k = kind_of(running[i].req.p);
      if (!shared_kind(running[i].req.p)) {
        if (running[i].inflight) continue;
        any = true;
        for (size_t j : issue_round(k, {i})) gone.push_back(j);
      } else if (running[i].inflight) {
        busy[k] = true;
      } else {
        ready[k].push_back(i);
      }
    }
    for (auto& kv : ready) {
      if (busy.count(kv.first)) continue;
      any = true;
      for (size_t i : issue_round(kv.first, kv.second)) gone.push_back(i);
    }
    std::sort(gone.begin(), gone.end());
    for (size_t m = gone.size(); m-- > 0;) {
      finish(running[gone[m]]);
      running.erase(running.begin() + (long)gone[m]);
    }
    return any;
  }

  void loop() {
    try {
      for (;;) {
        admit();
        if (running.empty()) {
          std::unique_lock<std::mutex> g(mu);
          if (stop && pending.empty()) return;
          cv.wait(g, [&] { return stop || !pending.empty(); });
          continue;
        }
        bool progressed = false;
        for (size_t i = 0; i < running.size();) {
          Running& r = running[i];
          if (!r.inflight || !r.s->attempt_ready()) {
            ++i;
            continue;
          }
          progressed = true;
          r.inflight = false;
          const int ended = r.s->adaptive_complete(r.rep);
          if (ended && converged(r)) {
            finish(r);
            running.erase(running.begin() + (long)i);
            continue;
          }
          ++i;
        }
        if (issue_rounds()) progressed = true;
        if (!progressed) std::this_thread::sleep_for(std::chrono::microseconds(20));
      }
    } catch (const vb::HipError& e) {
      vb::set_last_error(std::string(hipGetErrorString(e.err)) + " in " + e.where);
      std::lock_guard<std::mutex> g(mu);
      worker_error = e.err == hipErrorOutOfMemory ? VB_ERR_OOM : VB_ERR_HIP;
    } catch (const std::bad_alloc&) {
      std::lock_guard<std::mutex> g(mu);
      worker_error = VB_ERR_OOM;
    }
  }
};

extern "C" {

void vb_server_config_default(vb_server_config* c) {
  if (!c) return;
  *c = vb_server_config{};
  c->device = 0;
  c->max_active = 16;
  c->streams = 4;
  c->keep_phi = 0;
  c->memory_budget = 0;
  c->converge_window = 50;
  c->converge_tol = 1e-6;
  c->max_steps = 0;
}

int vb_server_create(const vb_server_config* cfg, vb_server** out) {
  if (!cfg || !out || cfg->max_active <= 0 || cfg->streams <= 0 || cfg->converge_window <= 0) return VB_ERR_PARAM;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= cfg->device) return VB_ERR_NO_DEVICE;
  try {
    auto s = std::make_unique<vb_server>();
    s->cfg = *cfg;
    s->t0 = std::chrono::steady_clock::now();
    VB_HIP(hipSetDevice(cfg->device));
    if (s->cfg.memory_budget == 0) {
      size_t fr = 0, tot = 0;
      VB_HIP(hipMemGetInfo(&fr, &tot));
      s->cfg.memory_budget = fr / 10 * 9;
    }
    for (int k = 0; k < cfg->streams; ++k) {
      s->streams.push_back(vb::make_queue_stream());
      s->stream_load.push_back(0);
    }
    vb_server* raw = s.get();
    s->worker = std::thread([raw] {
      (void)hipSetDevice(raw->cfg.device);
      raw->loop();
    });
    *out = s.release();
    return VB_OK;
  } catch (const vb::HipError& e) {
    vb::set_last_error(std::string(hipGetErrorString(e.err)) + " in " + e.where);
    return VB_ERR_HIP;
  } catch (const std::bad_alloc&) {
    return VB_ERR_OOM;
  }
}

int vb_server_submit(vb_server* s, long long id, const vb_params* p, const double* phi0) {
  if (!s || !p || !phi0) return VB_ERR_PARAM;
  const int rc = vb::validate_params(p);
  if (rc != VB_OK) return rc;
  if (p->fixed_time_step > 0 || p->scheme == VB_EULER) return VB_ERR_PARAM;   // adaptive ETD only
  if (p->device != s->cfg.device) return VB_ERR_PARAM;
  Request q;
  q.id = id;
  q.p = *p;
  q.phi0.assign(phi0, phi0 + (size_t)p->nx * p->nx * p->nz);
  q.t_submit = now_s(s->t0);
  {
    std::lock_guard<std::mutex> g(s->mu);
    if (s->worker_error) return s->worker_error;
    if (vb::solver_footprint(*p) > s->cfg.memory_budget) return VB_ERR_OOM;
    s->pending.push_back(std::move(q));
  }
  s->cv.notify_all();
  return VB_OK;
}

int vb_server_poll(vb_server* s, vb_result* out, int cap, int* n_out) {
  if (!s || !n_out || (cap > 0 && !out)) return VB_ERR_PARAM;
  std::lock_guard<std::mutex> g(s->mu);
  const int n = std::min(cap, (int)s->done.size());
  for (int i = 0; i < n; ++i) out[i] = s->done[(size_t)i];
  s->done.erase(s->done.begin(), s->done.begin() + n);
  *n_out = n;
  return s->worker_error;
}

int vb_server_pending(vb_server* s, int* queued, int* running, size_t* bytes_in_use) {
  if (!s) return VB_ERR_PARAM;
  std::lock_guard<std::mutex> g(s->mu);
  if (queued) *queued = (int)s->pending.size();
  if (running) *running = s->nrun;
  if (bytes_in_use) *bytes_in_use = s->used;
  return s->worker_error;
}

int vb_server_take_phi(vb_server* s, long long id, double* host_out) {
  if (!s || !host_out) return VB_ERR_PARAM;
  std::lock_guard<std::mutex> g(s->mu);
  auto it = s->phis.find(id);
  if (it == s->phis.end()) return VB_ERR_PARAM;
  std::copy(it->second.begin(), it->second.end(), host_out);
  s->phis.erase(it);
  return VB_OK;
}

double vb_server_clock(vb_server* s) { return s ? now_s(s->t0) : 0.0; }

void vb_server_destroy(vb_server* s) { delete s; }

}  // extern "C"
