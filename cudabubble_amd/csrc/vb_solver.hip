// vb_solver.hip -- fused stage pipeline and drivers of one vesicle solve.
//
// State: the half spectrum u^ of the phase field (two buffers, current and
// candidate, so a rejected adaptive step needs no copy-back).  Work buffers:
// P, Q (spectra of the stage field and of its Laplacian during the inverse
// transform), Tt (tmp1 / N pipeline), W (f pipeline, Euler only), pre, Na and
// Nb (ETD stage spectra, group-major), v and f (real stage field and f, Euler
// only), block partial sums, and the device scalars (V, A, E_b, E, lambda1,
// lambda2, alpha, beta).
//
// A "prefix" turns the prepared stage field into v, f, lap tmp1 and the
// integrals; a "suffix" forms N, its spectrum and the ETD combine of one
// stage, and prepares the next stage field.  One ETDRK4 evaluation is
// prefix(1) suffix(1) ... prefix(4) suffix(4), where prefix(1) of the next
// evaluation is also the energy evaluation of the new state.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <map>

#include "vb_solver.hpp"

namespace vb {

// ---------------------------------------------------------------------------
// finalize: block partials -> V, A, E_b, E, lambda (Calc.cu:191-263,
// phasefield.cpp:283-325, belastic.cpp:68-86 and 473-479)
// ---------------------------------------------------------------------------
thread_local Recorder* g_recorder = nullptr;

constexpr int kFinThreads = 1024;   // one block; wide so the partial sums load in few rounds

template <bool TAB>
__global__ void __launch_bounds__(kFinThreads) k_finalize(FinArgs a0, const FinArgs* tab) {
  const FinArgs a = TAB ? tab[blockIdx.z] : a0;
  __shared__ double red[6 * (kFinThreads / 64)];
  __shared__ double tot[7];
  finalize_body<kFinThreads>(a, red, tot, 0, 1);
}

// Across the penalty-drift double step: KEEP saves the scalars of the
// current state (its integrals do not depend on the targets); RETARGET puts
// them back under the current (drifted) targets, which is what the finalize
// kernel would write for that state's partial sums.
struct KeepArgs {
  Scal* sc;
  Scal* keep;
  Phys ph;
};

template <bool TAB>
__global__ void k_keep_scal(KeepArgs a0, const KeepArgs* tab) {
  const KeepArgs a = TAB ? tab[blockIdx.z] : a0;
  if (threadIdx.x == 0 && blockIdx.x == 0) *a.keep = *a.sc;
}

template <bool TAB>
__global__ void k_retarget(KeepArgs a0, const KeepArgs* tab) {
  const KeepArgs a = TAB ? tab[blockIdx.z] : a0;
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    Scal c = *a.keep;
    c.alpha = a.sc->alpha;
    c.beta = a.sc->beta;
    c.alpha0 = a.sc->alpha0;
    c.beta0 = a.sc->beta0;
    apply_targets(c, a.ph);
    *a.sc = c;
  }
}

static void keep_tab(int op, const void* tab, dim3 gr, dim3 bl, int lanes, hipStream_t s) {
  gr.z = unsigned(lanes);
  if (op == OP_KEEP)
    hipLaunchKernelGGL(k_keep_scal<true>, gr, bl, 0, s, KeepArgs{}, (const KeepArgs*)tab);
  else
    hipLaunchKernelGGL(k_retarget<true>, gr, bl, 0, s, KeepArgs{}, (const KeepArgs*)tab);
}

// op: OP_KEEP or OP_RETARGET
static void launch_keep(int op, Scal* sc, Scal* keep, const Phys& ph, hipStream_t s) {
  const KeepArgs a{sc, keep, ph};
  if (g_recorder) return record_launch(&keep_tab, op, a, dim3(1), dim3(64));
  if (op == OP_KEEP)
    hipLaunchKernelGGL(k_keep_scal<false>, dim3(1), dim3(64), 0, s, a, nullptr);
  else
    hipLaunchKernelGGL(k_retarget<false>, dim3(1), dim3(64), 0, s, a, nullptr);
}

__global__ void k_fill_scal(Scal* sc, double alpha, double beta, double alpha0, double beta0) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    sc->alpha = alpha;
    sc->beta = beta;
    sc->alpha0 = alpha0;
    sc->beta0 = beta0;
  }
}

template <typename TI>
static void relayout_tab(int, const void* tab, dim3 gr, dim3 bl, int lanes, hipStream_t s) {
  gr.z = unsigned(lanes);
  hipLaunchKernelGGL((k_relayout<TI, true>), gr, bl, 0, s, RelArgs<TI>{}, (const RelArgs<TI>*)tab);
}

static void finalize_tab(int, const void* tab, dim3 gr, dim3 bl, int lanes, hipStream_t s) {
  gr.z = unsigned(lanes);
  hipLaunchKernelGGL(k_finalize<true>, gr, bl, 0, s, FinArgs{}, (const FinArgs*)tab);
}

FinArgs fin_args(const double* xpart, int nxb, const double* zpart, const double* dc, int nzb, Scal* sc,
                 FinScratch* fs, const Geo& g, const Phys& ph, int mode) {
  FinArgs a;
  a.xpart = xpart;
  a.zpart = zpart;
  a.dc = dc;
  a.dcpart = nullptr;
  a.sc = sc;
  a.fs = fs;
  a.nxb = nxb;
  a.nzb = nzb;
  a.mode = mode;
  a.pad = 0;
  a.g = g;
  a.ph = ph;
  return a;
}

void launch_finalize(const double* xpart, int nxb, const double* zpart, const double* dc, int nzb, Scal* sc,
                     const Geo& g, const Phys& ph, int mode, hipStream_t s) {
  const FinArgs a = fin_args(xpart, nxb, zpart, dc, nzb, sc, nullptr, g, ph, mode);
  if (g_recorder) return record_launch(&finalize_tab, OP_FIN, a, dim3(1), dim3(kFinThreads));
  hipLaunchKernelGGL(k_finalize<false>, dim3(1), dim3(kFinThreads), 0, s, a, nullptr);
}

hipStream_t make_queue_stream() {
  int dev = 0;
  VB_HIP(hipGetDevice(&dev));
  int ncu = 0;
  VB_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  std::vector<uint32_t> mask((size_t)(ncu + 31) / 32, 0u);
  for (int c = 0; c < ncu; ++c) mask[(size_t)c / 32] |= 1u << (c % 32);
  hipStream_t s = nullptr;
  VB_HIP(hipExtStreamCreateWithCUMask(&s, (uint32_t)mask.size(), mask.data()));
  return s;
}

void launch_fill_scal(Scal* sc, double alpha, double beta, double alpha0, double beta0, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_scal, dim3(1), dim3(64), 0, s, sc, alpha, beta, alpha0, beta0);
}

// ---------------------------------------------------------------------------
// Solver
// ---------------------------------------------------------------------------
template <typename T>
struct Solver final : SolverBase {
  Geo g{};
  Phys ph{};
  hipStream_t st = nullptr;
  bool own_stream = false;
  const SizeOps<T>* ox = nullptr;
  const SizeOps<T>* oz = nullptr;
  // FP64 operations (state transforms, FP64 queries in mixed mode)
  const SizeOps<double>* oxd = nullptr;
  const SizeOps<double>* ozd = nullptr;
  DBuf<cpx<double>> uh[2], S;            // state (FP64, group-major) and a transform scratch
  DBuf<cpx<T>> P, Q, W;                  // transform work buffers
  DBuf<cpx<T>> Tt, Ts;                   // the state's prefix (group-major)
  DBuf<cpx<T>> Tt1, Ts1;                 // the prefixes of stages 2..4 (group-major)
  DBuf<cpx<T>> pre, Na, Nb, Vh;          // stage spectra kept for later stages (group-major)
  DBuf<T> v, f;
  DBuf<double> xpart, zpart, phid, dc;
  DBuf<double> xdc;                      // FP32-mixed: the x kernel's FP64 sums of tmp1 and eps^3 s
  DBuf<Scal> scal;
  DBuf<Scal> scal_keep;                  // the state's integrals across the drift double step
  DBuf<FinScratch> fin_scratch;
  DBuf<cpx<T>> twx, twz;
  DBuf<cpx<double>> twxd, twzd;
  // FP32-mixed only: FP64 integrals of the state (FP64 transforms of the
  // FP64 state), for the energy-decrease test of the adaptive driver and for
  // vb_query
  DBuf<cpx<double>> P64, Q64;
  DBuf<double> xpart64, zpart64, dc64;
  DBuf<Scal> scal64;
  int nxb64 = 0, nzb64 = 0;
  Scal* hs = nullptr;   // pinned host copy
  DBuf<double> rpart;   // block sums of the Euler residual
  int nxb = 0, nzb = 0;
  long long npairs = 0;
  int mmax = 0;
  // ETD coefficient tables: a fixed pool of slots allocated with the solver.
  // A (scheme, h) table is built on the host at its first use, into pinned
  // staging, and copied on the solver's stream; when the pool is full the
  // least recently used slot is refilled and the captured graphs that read it
  // are dropped.  Stepping never allocates, frees or waits for the device.
  static constexpr int kLutSlots = 24;
  struct Lut {
    double2* c1;
    double4* c2;
    int sch;        // 0: empty
    double h;
    unsigned long long used;
  };
  Lut slots[kLutSlots] = {};
  unsigned long long lut_clock = 0;
  DBuf<double2> lut_c1;
  DBuf<double4> lut_c2;
  double* lut_stage = nullptr;      // pinned, (mmax + 1) * 6 doubles
  hipEvent_t stage_ev = nullptr;    // the last copy out of lut_stage

  // pipeline state
  int cur = 0;
  bool prepped = false;     // P, Q hold zprep of uh[cur]
  bool prefixed = false;    // ETD prefix of uh[cur] done (scal; Tt, Ts: x,y-forward of tmp1, s)
  bool prefixed_f = false;  // Euler prefix of uh[cur] done (scal; v, f; W: x-spectra of lap f)
  // driver state
  bool started = false;
  double time_step = 0, time_current = 0;
  int steps = 0;
  double alpha = 0, beta = 0, alpha0 = 0, beta0 = 0;

  // profiling
  bool prof = false;
  struct Rec { const char* name; hipEvent_t a, b; };
  std::vector<Rec> pending;
  std::vector<hipEvent_t> evpool;
  std::map<std::string, std::pair<int, double>> ktimes;

  // Driver work captured as HIP graphs: two fixed-dt iterations, one
  // adaptive attempt, a chain of attempts decided on the device, or one Euler
  // attempt.  A graph records the launches that work issues from one
  // host-side pipeline state (the state index cur, the prepared/prefixed
  // flags and dt) and is replayed only from that same state; afterwards the
  // host state is set to what the work leaves behind, exactly as if it had
  // run eagerly.  Chains and Euler attempts read their dt from device memory,
  // so theirs are replayed whatever dt is.
  struct PipeState {
    int cur;
    bool prepped, prefixed, prefixed_f;
    double dt;
    bool operator==(const PipeState& o) const {
      return cur == o.cur && prepped == o.prepped && prefixed == o.prefixed && prefixed_f == o.prefixed_f &&
             dt == o.dt;
    }
  };
  enum { kFixedPair = 0, kAttempt = 1, kChain = 2, kShapeAttempt = 3, kEulerAttempt = 4, kShapeChain = 5 };
  struct StepGraph {
    hipGraphExec_t x = nullptr;
    int kind = kFixedPair;
    PipeState pre{}, post{};
    int evals = 0;
  };
  static constexpr int kGraphIters = 2;
  static constexpr int kGraphCache = 16;   // adaptive dt visits a ladder of values
  StepGraph graphs[kGraphCache];
  int graph_next = 0;
  PipeState fixed_post{};        // state after the last fixed iteration
  bool have_fixed_post = false;

  // geometry and physical constants of p (the buffer layout depends on nx,
  // nz and the kernel schedules only)
  void setup_geo() {
    g = make_geo(p.nx, p.nz, p.len * (2.0 * 3.1415926535897932384626433832795));
    ph.eps = p.eps;
    ph.eta = p.eta;
    ph.ga = p.ga;
    ph.kappa = p.kappa;
    ph.D = p.D;
    ph.M1 = p.M1;
    ph.M2 = p.M2;
    // padded, 128-byte aligned spectrum rows when every column tile width in
    // use divides nx/2 (see Geo)
    bool tiles = true;
    for (int t : {ox->ytile, oz->ytile, oxd->ytile, ozd->ytile, ox->ctile, oz->ctile, oxd->ctile, ozd->ctile}) {
      // the group-major index (gm_index) divides by a column tile with shifts
      if (t <= 0 || (t & (t - 1)) != 0) throw HipError{hipErrorInvalidValue, "column tile not a power of two"};
      tiles = tiles && (p.nx / 2) % t == 0;
    }
    if ((p.nx / 2) % 16 == 0 && tiles) {
      // rows of nx/2 modes (whole 128-byte lines), then the plane's Nyquist
      // column; one more aligned unit of 16 keeps the z stride away from large
      // powers of two (which concentrate z columns on few HBM channels)
      g.pitch = p.nx / 2;
      g.aligned = 1;
      g.nq = (long long)g.ny * g.pitch;
      g.pp = g.nq + g.ny + 16;
      g.ncp = (long long)g.nz * g.pp;
    }
  }

  Solver(const vb_params& pp, hipStream_t shared) {
    p = pp;
    ox = get_ops<T>(p.nx);
    oz = get_ops<T>(p.nz);
    oxd = get_ops<double>(p.nx);
    ozd = get_ops<double>(p.nz);
    setup_geo();
    if (shared) {
      st = shared;
    } else {
      VB_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
      own_stream = true;
    }
    const size_t nc = (size_t)g.ncp, nr = (size_t)g.nr;
    for (auto* b : {&uh[0], &uh[1]}) b->alloc((size_t)zs_private(g, oz->ytile));
    // S holds the state in the row layout (phi in and out) or regrouped for
    // the FP64 z kernels of the mixed-mode integrals
    S.alloc((size_t)std::max(zs_private(g, oz->ytile), zs_private(g, ozd->ytile)));
    for (auto* b : {&P, &Q}) b->alloc(nc);
    for (auto* b : {&Tt, &Ts, &Tt1, &Ts1, &pre, &Na, &Nb, &Vh}) b->alloc((size_t)zs_private(g, oz->ytile));
    // W, v and f carry the Euler right-hand side only (run_static, shape
    // adjust, the residual); the ETD pipeline needs no real fields stored
    if (p.scheme == VB_EULER || p.shape_adjust) {
      W.alloc(nc);
      v.alloc(nr);
      f.alloc(nr);
    }
    dc.alloc(1);
    phid.alloc(nr);
    npairs = (long long)g.nz * g.ny / 2;
    nxb = int((npairs + ox->xlanes - 1) / ox->xlanes);
    nzb = int(zgroups(g, oz->ytile));
    xpart.alloc((size_t)nxb * 3);
    if (kMixed) xdc.alloc((size_t)nxb * 2);
    zpart.alloc((size_t)nzb);
    scal.alloc(1);
    scal_keep.alloc(1);
    fin_scratch.alloc(1);
    VB_HIP(hipMemsetAsync(fin_scratch.p, 0, sizeof(FinScratch), st));
    VB_HIP(hipHostMalloc((void**)&hs, sizeof(Scal), hipHostMallocDefault));
    memset(hs, 0, sizeof(Scal));
    VB_HIP(hipMemsetAsync(scal.p, 0, sizeof(Scal), st));
    upload_twiddles(twx, p.nx);
    upload_twiddles(twz, p.nz);
    upload_twiddles(twxd, p.nx);
    upload_twiddles(twzd, p.nz);
    mmax = mmax_of(p.nx, p.ny, p.nz);
    lut_c1.alloc((size_t)kLutSlots * (mmax + 1));
    lut_c2.alloc((size_t)kLutSlots * (mmax + 1));
    for (int k = 0; k < kLutSlots; ++k) {
      slots[k].c1 = lut_c1.p + (size_t)k * (mmax + 1);
      slots[k].c2 = lut_c2.p + (size_t)k * (mmax + 1);
    }
    VB_HIP(hipHostMalloc((void**)&lut_stage, sizeof(double) * 6 * (size_t)(mmax + 1), hipHostMallocDefault));
    VB_HIP(hipEventCreateWithFlags(&stage_ev, hipEventDisableTiming));
    time_step = p.time_step;
    alpha = p.alpha;
    beta = p.beta;
    alpha0 = alpha;
    beta0 = beta;
    set_targets_device();
    VB_HIP(hipStreamSynchronize(st));
  }

  ~Solver() override {
    if (st) (void)hipStreamSynchronize(st);
    for (auto& G : graphs)
      if (G.x) (void)hipGraphExecDestroy(G.x);
    if (att_ev) (void)hipEventDestroy(att_ev);
    if (stage_ev) (void)hipEventDestroy(stage_ev);
    if (lut_stage) (void)hipHostFree(lut_stage);
    if (hchain) (void)hipHostFree(hchain);
    if (hshape) (void)hipHostFree(hshape);
    flush_prof();
    for (auto e : evpool) (void)hipEventDestroy(e);
    if (hs) (void)hipHostFree(hs);
    if (own_stream && st) (void)hipStreamDestroy(st);
  }

  template <typename X>
  void upload_twiddles(DBuf<cpx<X>>& b, int n) {
    std::vector<double> t;
    build_twiddles(t, n);
    std::vector<cpx<X>> h((size_t)n);
    for (int m = 0; m < n; ++m) {
      h[(size_t)m].x = X(t[2 * (size_t)m]);
      h[(size_t)m].y = X(t[2 * (size_t)m + 1]);
    }
    b.alloc((size_t)n);
    VB_HIP(hipMemcpy(b.p, h.data(), sizeof(cpx<X>) * (size_t)n, hipMemcpyHostToDevice));
  }

  const Lut* lut(int sch, double h) {
    Lut* hit = nullptr;
    for (auto& sl : slots)
      if (sl.sch == sch && sl.h == h) hit = &sl;
    if (hit) {
      hit->used = ++lut_clock;
      return hit;
    }
    Lut* sl = &slots[0];
    for (auto& c : slots)
      if (c.sch == 0 || (sl->sch != 0 && c.used < sl->used)) sl = &c;
    if (sl->sch != 0) {
      // captured graphs read this slot's table: drop those of its step
      for (auto& G : graphs)
        if (G.x && G.pre.dt == sl->h) {
          VB_HIP(hipGraphExecDestroy(G.x));
          G.x = nullptr;
        }
    }
    std::vector<Coef> tab;
    build_lut(tab, mmax, sch, p.phi_lut_mode, h, g.kfac2, p.eps, p.kappa, p.D);
    // compact per-stage arrays: {e1, q} and {e2, A s, 2 B s, C s} (ETDRK2: {e, q1}, {c2})
    VB_HIP(hipEventSynchronize(stage_ev));   // the staging buffer's previous copy is done
    double2* h1 = reinterpret_cast<double2*>(lut_stage);
    double4* h2 = reinterpret_cast<double4*>(lut_stage + 2 * (size_t)(mmax + 1));
    for (size_t m = 0; m < tab.size(); ++m) {
      h1[m].x = tab[m].c[0];
      h1[m].y = tab[m].c[1];
      if (sch == 2) {
        h2[m].x = tab[m].c[2];
        h2[m].y = h2[m].z = h2[m].w = 0.0;
      } else {
        h2[m].x = tab[m].c[2];
        h2[m].y = tab[m].c[3];
        h2[m].z = tab[m].c[4];
        h2[m].w = tab[m].c[5];
      }
    }
    VB_HIP(hipMemcpyAsync(sl->c1, h1, sizeof(double2) * tab.size(), hipMemcpyHostToDevice, st));
    VB_HIP(hipMemcpyAsync(sl->c2, h2, sizeof(double4) * tab.size(), hipMemcpyHostToDevice, st));
    VB_HIP(hipEventRecord(stage_ev, st));
    sl->sch = sch;
    sl->h = h;
    sl->used = ++lut_clock;
    return sl;
  }

  // ---- profiling ---------------------------------------------------------
  hipEvent_t ev() {
    if (!evpool.empty()) {
      hipEvent_t e = evpool.back();
      evpool.pop_back();
      return e;
    }
    hipEvent_t e;
    VB_HIP(hipEventCreate(&e));
    return e;
  }
  void flush_prof() {
    if (pending.empty()) return;
    (void)hipStreamSynchronize(st);
    for (auto& r : pending) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
        auto& k = ktimes[r.name];
        k.first += 1;
        k.second += ms;
      }
      evpool.push_back(r.a);
      evpool.push_back(r.b);
    }
    pending.clear();
  }
  template <class F>
  void L(const char* name, F&& fn) {
    if (!prof) {
      fn();
      return;
    }
    Rec r{name, ev(), ev()};
    VB_HIP(hipEventRecord(r.a, st));
    fn();
    VB_HIP(hipEventRecord(r.b, st));
    pending.push_back(r);
    if (pending.size() > 2048) flush_prof();
  }
  void profile(bool on) override {
    flush_prof();
    if (on) ktimes.clear();
    prof = on;
  }
  int kernel_times(char* names, int name_len, int* counts, double* ms, int cap) override {
    flush_prof();
    int i = 0;
    for (auto& kv : ktimes) {
      if (i >= cap) break;
      if (names && name_len > 0) {
        strncpy(names + (size_t)i * name_len, kv.first.c_str(), (size_t)name_len - 1);
        names[(size_t)i * name_len + name_len - 1] = 0;
      }
      if (counts) counts[i] = kv.second.first;
      if (ms) ms[i] = kv.second.second;
      ++i;
    }
    return (int)ktimes.size();
  }
  hipStream_t stream() override { return st; }

  // ---- pipeline pieces ------------------------------------------------------
  // copy of a spectrum between the row layout (tile 0) and a group-major one
  template <typename TI>
  void relayout(const cpx<TI>* in, int tb_in, cpx<double>* out, int tb_out) {
    const long long nm = (long long)g.nz * g.ny * (g.nx / 2 + 1);
    const RelArgs<TI> a{in, out, tb_in, tb_out, g};
    const dim3 gr(unsigned((nm + 255) / 256)), bl(256);
    if (g_recorder) return record_launch(&relayout_tab<TI>, OP_RELAYOUT, a, gr, bl);
    hipLaunchKernelGGL((k_relayout<TI, false>), gr, bl, 0, st, a, nullptr);
  }
  void check_launch() { VB_HIP(hipGetLastError()); }

  void prep(const cpx<double>* vh) {
    L("zprep", [&] { oz->zprep(vh, P.p, Q.p, zpart.p, dc.p, g, twz.p, st); });
    check_launch();
    prepped = true;
  }

  // fin: integrals to evaluate in the same launch as the z pass (extra
  // workgroups of it), or nullptr
  void spectral_lap(cpx<T>* b, const FinArgs* fin) {
    L("y_fwd", [&] { ox->ypass(b, nullptr, 1, -1, g, twx.p, nullptr, st); });
    L(fin ? "z_lap_fin" : "z_lap", [&] { oz->zlap(b, g, twz.p, fin, st); });
    L("y_inv", [&] { ox->ypass(b, nullptr, 1, +1, g, twx.p, nullptr, st); });
  }

  // The prefix of the prepared stage field (P, Q: its z-inverse spectra and
  // those of its Laplacian): the integrals in scal, and
  //  - ETD (withf = false): Tt, Ts = x,y-forward of tmp1 and s, from which the
  //    stage combine forms N^ (k_xa, k_zstage);
  //  - Euler (withf = true): v, f and W = x-spectra of lap f (k_xe).
  // set 0: the prefix of the state (stage 1), kept in Tt, Ts; set 1: of a
  // later stage, in Tt1, Ts1 (so the state's survives an evaluation)
  void prefix(int mode, bool withf, int set = 0) {
    cpx<T>* tt = set ? Tt1.p : Tt.p;
    cpx<T>* ts = set ? Ts1.p : Ts.p;
    L("y_inv2", [&] { ox->ypass(P.p, Q.p, 2, +1, g, twx.p, nullptr, st); });
    XaArgs<T> a{};
    a.P = P.p;
    a.Q = Q.p;
    a.v = v.p;
    a.f = f.p;
    // FP32-mixed ETD prefix: FP64 sums of tmp1 and eps^3 s for the k = 0 modes
    a.dcpart = (kMixed && !withf) ? xdc.p : nullptr;
    // ETD: tmp1 and s go to P, Q in place (each lane reads its two rows of P
    // and Q before it writes them); Euler: tmp1 to a scratch, f to W
    a.Tt = withf ? tt : P.p;
    a.W = W.p;
    a.Ts = Q.p;
    a.part = xpart.p;
    a.tw = twx.p;
    a.npairs = npairs;
    a.eps = p.eps;
    a.kappa = p.kappa;
    a.D = p.D;
    a.inv_n = 1.0 / double(g.nr);
    a.rl = Rows{g.ny, g.pitch, g.pp, g.nq};
    L(withf ? "xa_f" : "xa", [&] { ox->xa(a, withf ? XA_WITHF : XA_WITHS, st); });
    // the integrals of this field (xa's and the z kernel's partials) are
    // evaluated by extra workgroups of the next column pass
    FinArgs fin = fin_args(xpart.p, nxb, zpart.p, dc.p, nzb, scal.p, fin_scratch.p, g, ph, mode);
    fin.dcpart = a.dcpart;
    if (withf) {
      spectral_lap(W.p, &fin);
    } else {
      L("y_fwd2", [&] { ox->ypass_gm(P.p, Q.p, tt, ts, oz->ytile, g, twx.p, &fin, st); });
    }
    check_launch();
    prepped = false;
    prefixed = !withf;
    prefixed_f = withf;
  }

  // V and A of the prepared stage field alone (no stage data, and no E_b):
  // y-inverse of P, the partial sums of v and (v^2-1)^2, and the finalize
  // kernel.  Used where only the targets' drift is needed.
  void prefix_integrals(int mode) {
    L("y_inv", [&] { ox->ypass(P.p, nullptr, 1, +1, g, twx.p, nullptr, st); });
    XaArgs<T> a{};
    a.P = P.p;
    a.Q = Q.p;
    a.part = xpart.p;
    a.tw = twx.p;
    a.npairs = npairs;
    a.eps = p.eps;
    a.kappa = p.kappa;
    a.D = p.D;
    a.inv_n = 1.0 / double(g.nr);
    a.rl = Rows{g.ny, g.pitch, g.pp, g.nq};
    L("xa_sums", [&] { ox->xa(a, XA_SUMS, st); });
    refinalize(mode, (nxb + kXsLoop - 1) / kXsLoop);
    prepped = false;
    prefixed = false;
    prefixed_f = false;
  }

  // nx_parts: the x partials written by the last x kernel (k_xsums writes fewer)
  void refinalize(int mode, int nx_parts = -1) {
    const int nxp = nx_parts < 0 ? nxb : nx_parts;
    L("finalize", [&] { launch_finalize(xpart.p, nxp, zpart.p, dc.p, nzb, scal.p, g, ph, mode, st); });
    check_launch();
  }

  // The combine of stage stg: z-forward of tmp1 and s, N^, the ETD combine,
  // and the z-inverse of the next stage field into P, Q.  probe (last ETDRK4
  // stage): the new state is only needed for its V and A, so neither it nor Q
  // is stored.  lt null: the chain's table buffers (the dt its control block chose).
  void suffix(int sch, int stg, const Lut* lt, int out, bool probe = false) {
    ZsArgs<T> z;
    z.U = stg == 1 ? Tt.p : Tt1.p;
    z.U2 = stg == 1 ? Ts.p : Ts1.p;
    z.Vh = Vh.p;
    z.sc = scal.p;
    z.eps = p.eps;
    z.kappa = p.kappa;
    z.eta = p.eta;
    z.uh = uh[cur].p;
    z.pre = pre.p;
    z.Na = Na.p;
    z.Nb = Nb.p;
    z.uout = uh[out].p;
    z.P = P.p;
    z.Q = Q.p;
    z.part = zpart.p;
    z.dc = dc.p;
    z.c1 = lt ? lt->c1 : chain_c1.p;
    z.c2 = lt ? lt->c2 : chain_c2.p;
    z.tw = twz.p;
    z.g = g;
    static const char* names[5] = {"zstage", "zstage1", "zstage2", "zstage3", "zstage4"};
    if (probe && sch == 4 && stg == 4)
      L("zstage4p", [&] { oz->zstage(z, sch, 5, st); });
    else
      L(names[stg], [&] { oz->zstage(z, sch, stg, st); });
    check_launch();
    prefixed = false;
    prefixed_f = false;
    prepped = !(probe && sch == 4 && stg == 4);   // P, Q hold the next stage field (the probe: P alone)
  }

  void ensure_prefixed(bool withf = false) {
    if (withf ? prefixed_f : prefixed) return;
    if (!prepped) prep(uh[cur].p);
    prefix(FIN_PLAIN, withf);
  }
  // the integrals of uh[cur] in scal: either prefix provides them
  void ensure_integrals() {
    if (prefixed || prefixed_f) return;
    ensure_prefixed(p.scheme == VB_EULER);
  }

  // one ETDRK4 / ETDRK2 evaluation from uh[cur] into uh[1-cur]; on return P, Q
  // hold the prepared new state.  probe: the new state is needed only for V
  // and A (prefix_integrals); for ETDRK4 uh[1-cur] and Q are then not written.
  int eval(double h, bool probe = false) {
    const int sch = p.scheme == VB_ETDRK2 ? 2 : 4;
    const Lut* lt = lut(sch, h);
    const int out = 1 - cur;
    ensure_prefixed();
    for (int stg = 1; stg <= sch; ++stg) {
      if (stg > 1) prefix(FIN_PLAIN, false, 1);
      suffix(sch, stg, lt, out, probe);
    }
    // after the last stage P, Q belong to uh[out]
    return out;
  }

  // Forward-Euler update of uh[cur] into uh[1-cur] with rhs g0 (shape) or nld1;
  // tdev non-null: the step is read from there on the device instead of t
  int euler(double t, bool shape, const double* tdev = nullptr) {
    ensure_prefixed(true);
    const int out = 1 - cur;
    XeArgs<T> e;
    e.W = W.p;
    e.v = v.p;
    e.f = f.p;
    e.out = W.p;   // in place over the rows it reads
    e.sc = scal.p;
    e.tw = twx.p;
    e.npairs = npairs;
    e.eps = p.eps;
    e.eta = p.eta;
    e.t = t;
    e.inv_n = 1.0 / double(g.nr);
    e.shape = shape ? 1 : 0;
    e.rl = Rows{g.ny, g.pitch, g.pp, g.nq};
    e.part = nullptr;
    e.tp = tdev;
    L("xe", [&] { ox->xe(e, st); });
    L("y_fwd", [&] { ox->ypass(W.p, nullptr, 1, -1, g, twx.p, nullptr, st); });
    L("z_fwd", [&] { oz->zpass(W.p, -1, g, twz.p, st); });
    L("widen", [&] {
      relayout(W.p, 0, uh[out].p, oz->ytile);
    });
    check_launch();
    prefixed_f = false;   // W was consumed
    return out;
  }

  bool drift_mode() const { return (p.M1 != 0 || p.M2 != 0) && p.penalty_drift; }

  // One fixed-dt iteration of the run_static_fEIF loop, launches only: the
  // penalty-drift double step (belastic.cpp:670-691) or a plain step
  // (belastic.cpp:692-693).  Returns the number of evaluations.
  int fixed_body() {
    if (drift_mode()) {
      ensure_prefixed();
      launch_keep(OP_KEEP, scal.p, scal_keep.p, ph, st);
      int out = eval(time_step, true);
      const int keep = cur;
      cur = out;
      prefix_integrals(FIN_DRIFT);   // only the drift of the targets is used
      cur = keep;
      // The second evaluation starts again from the kept state with the new
      // targets.  Its prefix data (Tt, Ts) does not depend on the targets and
      // is still there; only lambda1, lambda2 and E change.
      launch_keep(OP_RETARGET, scal.p, scal_keep.p, ph, st);
      prepped = false;
      prefixed = true;
      prefixed_f = false;
      out = eval(time_step);
      cur = out;
      prefix(FIN_DRIFT, false);
      return 2;
    }
    ensure_prefixed();
    cur = eval(time_step);
    return 1;
  }

  PipeState pipe_state() const { return PipeState{cur, prepped, prefixed, prefixed_f, time_step}; }
  void set_pipe_state(const PipeState& q) {
    cur = q.cur;
    prepped = q.prepped;
    prefixed = q.prefixed;
    prefixed_f = q.prefixed_f;
  }

  // Records the launches of `body` (returning its evaluation count) from the
  // current pipeline state.  Nothing runs during the capture; the host state
  // is put back afterwards.
  template <class F>
  StepGraph& capture(int kind, F&& body) {
    StepGraph& G = graphs[graph_next];
    graph_next = (graph_next + 1) % kGraphCache;
    if (G.x) {
      VB_HIP(hipGraphExecDestroy(G.x));
      G.x = nullptr;
    }
    if (kind != kShapeAttempt && kind != kEulerAttempt && kind != kShapeChain)
      (void)lut(p.scheme == VB_ETDRK2 ? 2 : 4, time_step);   // table upload happens outside the capture
    G.kind = kind;
    G.pre = pipe_state();
    VB_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    hipGraph_t gr = nullptr;
    try {
      G.evals = body();
    } catch (...) {
      (void)hipStreamEndCapture(st, &gr);
      if (gr) (void)hipGraphDestroy(gr);
      set_pipe_state(G.pre);
      throw;
    }
    G.post = pipe_state();
    set_pipe_state(G.pre);
    VB_HIP(hipStreamEndCapture(st, &gr));
    const hipError_t e = hipGraphInstantiateWithFlags(&G.x, gr, 0);
    (void)hipGraphDestroy(gr);
    if (e != hipSuccess) {
      G.x = nullptr;
      throw HipError{e, "hipGraphInstantiateWithFlags"};
    }
    return G;
  }

  StepGraph& capture_iters() {
    return capture(kFixedPair, [&] {
      int ev = 0;
      for (int i = 0; i < kGraphIters; ++i) ev += fixed_body();
      return ev;
    });
  }

  StepGraph* find_graph(int kind = kFixedPair) {
    const PipeState now = pipe_state();
    for (auto& G : graphs)
      if (G.x && G.kind == kind && G.pre == now) return &G;
    return nullptr;
  }

  // n fixed-dt iterations launched without host synchronisation.  A graph
  // is captured once the pipeline is back in the state the previous fixed
  // iteration left (the state that recurs from one iteration to the next)
  // and replayed whenever the pipeline is in a captured graph's start state.
  void launch_fixed(int n, vb_step_report& r) {
    const bool use_graphs = p.graphs && !prof;
    while (n > 0) {
      if (use_graphs && n >= kGraphIters) {
        StepGraph* G = find_graph();
        if (!G && have_fixed_post && fixed_post == pipe_state()) G = &capture_iters();
        if (G) {
          VB_HIP(hipGraphLaunch(G->x, st));
          set_pipe_state(G->post);
          fixed_post = G->post;
          r.evaluations += G->evals;
          for (int i = 0; i < kGraphIters; ++i) {
            time_current += time_step;
            steps++;
            r.steps++;
          }
          n -= kGraphIters;
          continue;
        }
      }
      r.evaluations += fixed_body();
      fixed_post = pipe_state();
      have_fixed_post = true;
      time_current += time_step;
      steps++;
      r.steps++;
      --n;
    }
  }

  bool fixed_dt() const override { return p.fixed_time_step > 0 && p.scheme != VB_EULER; }

  // ---- batched launches (see SolverBase) -------------------------------------
  PipeState rec_pre{}, rec_post{};
  int rec_evals = 0;

  int z_tile() const override { return oz->ytile; }

  GroupKey group_key() const override {
    GroupKey k{};
    k.ok = (fixed_dt() && p.graphs) ? 1 : 0;
    k.nx = p.nx;
    k.nz = p.nz;
    k.precision = p.precision;
    k.scheme = p.scheme;
    k.drift = drift_mode() ? 1 : 0;
    return k;
  }

  int record_pair(Recorder& rec) override {
    rec_evals = 0;
    if (!started || prof || !have_fixed_post || !(fixed_post == pipe_state())) return 0;
    (void)lut(p.scheme == VB_ETDRK2 ? 2 : 4, time_step);   // tables exist before recording
    rec_pre = pipe_state();
    g_recorder = &rec;
    int ev = 0;
    try {
      for (int i = 0; i < kGraphIters; ++i) ev += fixed_body();
    } catch (...) {
      g_recorder = nullptr;
      set_pipe_state(rec_pre);
      throw;
    }
    g_recorder = nullptr;
    rec_post = pipe_state();
    set_pipe_state(rec_pre);
    rec_evals = ev;
    return ev;
  }

  bool at_recorded_state() const override { return rec_evals > 0 && pipe_state() == rec_pre; }

  void advance_recorded(vb_step_report* r) override {
    cur_E_known = false;
    set_pipe_state(rec_post);
    fixed_post = rec_post;
    have_fixed_post = true;
    r->evaluations += rec_evals;
    for (int i = 0; i < kGraphIters; ++i) {
      time_current += time_step;
      steps++;
      r->steps++;
    }
  }

  // ---- adaptive ETD iterations in two halves (belastic.cpp:639-669, 713-746)
  // adaptive_issue() launches one attempt: the evaluation at the current dt,
  // the energy of the candidate and the copy of its scalars to pinned host
  // memory; adaptive_complete() waits for that copy and accepts or rejects.
  // A scheduler can issue attempts of many solves before completing any.
  hipEvent_t att_ev = nullptr;
  bool att_pending = false;
  bool iter_open = false;      // dt chosen for the current iteration
  bool cur_E_known = false;    // cur_E is the energy of uh[cur] with the current targets
  bool hs_state = false;       // hs holds the integrals of uh[cur] (an accepted attempt's)
  double cur_E = 0.0, att_eng0 = 0.0;
  int att_keep = 0;

  // the energy-decrease test compares FP64 energies in both precisions: in
  // FP32-mixed mode they come from FP64 transforms of the FP64 state
  static constexpr bool kMixed = sizeof(T) == 4;

  void adaptive_issue(vb_step_report& r) override {
    if (attempt_prepare(r)) issue_prepared();
  }

  bool attempt_prepare(vb_step_report& r) override {
    hs_state = false;
    if (!started) start();
    if (kMixed) need64();   // allocations stay outside graph capture
    if (!iter_open) {
      if (!cur_E_known) {
        if (kMixed) {
          energy64();
          read_scal64();
        } else {
          ensure_prefixed();
          read_scal();
        }
        cur_E = hs->E;
        cur_E_known = true;
      }
      att_eng0 = cur_E;
      if (steps % p.dt_double_every == 0) {
        if (time_step * 1.4 < p.dt_cap) time_step = time_step * 2;
        else time_step = p.dt_cap;
      }
      iter_open = true;
    }
    if (time_step < 1e-23) {
      r.status = VB_ERR_DT_UNDERFLOW;
      return false;
    }
    if (!att_ev) VB_HIP(hipEventCreateWithFlags(&att_ev, hipEventDisableTiming));
    att_keep = cur;
    return true;
  }

  void issue_prepared() override {
    if (p.graphs && !prof) {
      // one graph per pipeline state and dt; the ladder of dt values recurs
      StepGraph* G = find_graph(kAttempt);
      if (!G) G = &capture(kAttempt, [&] { attempt_body(); return 1; });
      VB_HIP(hipGraphLaunch(G->x, st));
      set_pipe_state(G->post);
    } else {
      attempt_body();
    }
    attempt_issued();
  }

  void record_attempt(Recorder& rec) override {
    (void)lut(p.scheme == VB_ETDRK2 ? 2 : 4, time_step);   // the table exists before recording
    const PipeState pre = pipe_state();
    g_recorder = &rec;
    try {
      attempt_body();
    } catch (...) {
      g_recorder = nullptr;
      set_pipe_state(pre);
      throw;
    }
    g_recorder = nullptr;
  }

  void attempt_issued() override {
    VB_HIP(hipMemcpyAsync(hs, kMixed ? scal64.p : scal.p, sizeof(Scal), hipMemcpyDeviceToHost, st));
    VB_HIP(hipEventRecord(att_ev, st));
    att_pending = true;
  }

  // the launches of one adaptive attempt: the evaluation, the prefix of the
  // candidate (its integrals; the first stage of the next attempt if it is
  // accepted) and, in FP32-mixed mode, its FP64 energy
  void attempt_body() {
    const int out = eval(time_step);
    cur = out;
    prefix(FIN_PLAIN, false);
    if (kMixed) energy64();
  }

  bool attempt_ready() override {
    if (!att_pending) return true;
    const hipError_t e = hipEventQuery(att_ev);
    if (e == hipErrorNotReady) return false;
    VB_HIP(e);
    return true;
  }

  // 1: the iteration ended (accepted, or a driver exit in r.status); 0: the
  // attempt was rejected and dt halved, issue again
  int adaptive_complete(vb_step_report& r) override {
    if (!att_pending) {
      iter_open = false;
      return 1;
    }
    VB_HIP(hipEventSynchronize(att_ev));
    att_pending = false;
    r.evaluations++;
    if (hs->E > att_eng0) {
      hs_state = false;
      cur = att_keep;   // the state is untouched; only its prefix is redone
      prefixed = prefixed_f = prepped = false;
      time_step /= 2;
      r.rejections++;
      return 0;
    }
    iter_open = false;
    cur_E = hs->E;
    r.energy = hs->E;
    hs_state = true;
    if (att_eng0 < 1e-3) r.status = VB_FINISHED;          // belastic.cpp:727-732
    else if (hs->nonfinite) r.status = VB_ERR_NONFINITE;  // belastic.cpp:733-740
    if (r.status != VB_OK) return 1;
    time_current += time_step;   // belastic.cpp:742-746
    steps++;
    r.steps++;
    r.time_step = time_step;
    r.time_current = time_current;
    return 1;
  }

  // ---- adaptive iterations accepted on the device ----------------------------
  // With p.accept_chain > 0 the adaptive ETD driver issues chains of up to that
  // many attempts as one graph launch and reads the driver state back once per
  // chain: k_chain_begin / k_chain_end (vb_chain.hip) take the decisions of
  // attempt_prepare() and adaptive_complete() on the device, from the same
  // doubles.  An attempt of a chain evaluates from uh[cur] into uh[out] with the
  // tables the control block selects (copied to chain_c1, chain_c2 by
  // k_chain_table), puts the candidate's prefix into set 1
  // and, if the candidate is accepted, copies it and that prefix over uh[cur]
  // and set 0; a rejection leaves both untouched, and its integrals are put
  // back from the control block.  So the state is always uh[cur], with its
  // prefix in set 0, and one captured chain serves every dt.
  // It is not the default: the host's round trip per attempt costs little
  // beside the attempt's launches.  Measured per attempt (FP32-mixed, 64
  // iterations from a torus, 20% rejected), host decisions vs chains of 4 and
  // 16: 118 / 115 / 120 us at 32^3, 181 / 178 / 182 at 64^3, 353 / 353 / 354 at
  // 128^3, and 1457 / 1493 / 1507 at 256^3, where the accepted candidate's
  // copies outweigh the round trip.
  DBuf<ChainCtl> chain;
  ChainCtl* hchain = nullptr;   // pinned
  DBuf<double2> chain_c1;       // the tables of the current attempt's dt
  DBuf<double4> chain_c2;
  static constexpr int kChainMax = 16;

  bool chain_ok() const {
    return p.accept_chain > 0 && p.graphs && !prof && p.fixed_time_step <= 0 && p.scheme != VB_EULER;
  }

  int chain_body(int K) {
    const int sch = p.scheme == VB_ETDRK2 ? 2 : 4;
    ChainCopy cp{};
    cp.n = 3;
    for (int k = 0; k < K; ++k) {
      launch_chain_begin(chain.p, scal.p, st);
      launch_chain_table(chain.p, chain_c1.p, chain_c2.p, mmax + 1, st);
      const int keep = cur, out = 1 - cur;
      for (int stg = 1; stg <= sch; ++stg) {
        if (stg > 1) prefix(FIN_PLAIN, false, 1);
        suffix(sch, stg, nullptr, out);
      }
      cur = out;
      prefix(FIN_PLAIN, false, 1);
      if (kMixed) energy64();
      cur = keep;
      launch_chain_end(chain.p, scal.p, kMixed ? scal64.p : scal.p, st);
      cp.job[0] = {uh[out].p, uh[keep].p, uh[keep].n * sizeof(cpx<double>)};
      cp.job[1] = {Tt1.p, Tt.p, Tt.n * sizeof(cpx<T>)};
      cp.job[2] = {Ts1.p, Ts.p, Ts.n * sizeof(cpx<T>)};
      launch_chain_copy(chain.p, cp, st);
      check_launch();
      prepped = false;
      prefixed = true;
      prefixed_f = false;
    }
    return K;
  }

  // a captured chain of kind (kChain, kShapeChain) and K attempts for the
  // current pipeline state, whatever dt
  StepGraph* find_chain(int K, int kind = kChain) {
    const PipeState now = pipe_state();
    for (auto& G : graphs)
      if (G.x && G.kind == kind && G.evals == K && G.pre.cur == now.cur && G.pre.prepped == now.prepped &&
          G.pre.prefixed == now.prefixed && G.pre.prefixed_f == now.prefixed_f)
        return &G;
    return nullptr;
  }

  // n iterations (or fewer at a driver exit) of the adaptive ETD driver in
  // chains; the same report as n rounds of adaptive_issue/adaptive_complete
  void chain_steps(int n, vb_step_report& r) {
    const int sch = p.scheme == VB_ETDRK2 ? 2 : 4;
    if (kMixed) need64();
    if (!chain.p) {
      chain.alloc(1);
      chain_c1.alloc((size_t)mmax + 1);
      chain_c2.alloc((size_t)mmax + 1);
      VB_HIP(hipHostMalloc((void**)&hchain, sizeof(ChainCtl), hipHostMallocDefault));
    }
    hs_state = false;
    int done = 0;
    while (done < n) {
      if (!iter_open && !cur_E_known) {
        if (kMixed) {
          energy64();
          read_scal64();
        } else {
          ensure_prefixed();
          read_scal();
        }
        cur_E = hs->E;
        cur_E_known = true;
        hs_state = false;
      }
      ensure_prefixed();
      // tables for the first attempt's dt and its half; the chain asks for
      // any other it reaches
      // (every table is in place before the control block lists them: a new
      // one can take the slot of an old)
      double d = time_step;
      if (!iter_open && steps % p.dt_double_every == 0) d = d * 1.4 < p.dt_cap ? d * 2 : p.dt_cap;
      if (d < 1e-23) d = time_step;
      (void)lut(sch, time_step);
      if (d / 2 >= 1e-23) (void)lut(sch, d / 2);
      const Lut* first = lut(sch, d);
      const int K = std::min(kChainMax, std::min(p.accept_chain, n - done));
      ChainCtl& c = *hchain;
      memset(&c, 0, sizeof(c));
      c.dt = time_step;
      c.dt_cap = p.dt_cap;
      c.t = time_current;
      c.E_cur = cur_E;
      c.E0 = att_eng0;
      c.energy = NAN;
      c.steps = steps;
      c.dt_double_every = p.dt_double_every;
      c.iter_open = iter_open ? 1 : 0;
      c.budget = K;
      c.status = CHAIN_RUN;
      c.mixed = kMixed ? 1 : 0;
      for (auto& sl : slots)
        if (sl.sch == sch && c.nlev < kChainLevels) {
          c.lev_dt[c.nlev] = sl.h;
          c.lev[c.nlev] = LutSel{sl.c1, sl.c2};
          c.nlev++;
        }
      c.sel = LutSel{first->c1, first->c2};   // read by attempts that do not run
      VB_HIP(hipMemcpyAsync(chain.p, hchain, sizeof(ChainCtl), hipMemcpyHostToDevice, st));
      StepGraph* G = find_chain(K);
      if (!G) G = &capture(kChain, [&] { return chain_body(K); });
      VB_HIP(hipGraphLaunch(G->x, st));
      set_pipe_state(G->post);
      VB_HIP(hipMemcpyAsync(hchain, chain.p, sizeof(ChainCtl), hipMemcpyDeviceToHost, st));
      VB_HIP(hipStreamSynchronize(st));
      time_step = c.dt;
      time_current = c.t;
      steps = c.steps;
      iter_open = c.iter_open != 0;
      att_eng0 = c.E0;
      cur_E = c.E_cur;
      cur_E_known = true;
      r.evaluations += c.evals;
      r.rejections += c.rej;
      r.steps += c.acc;
      done += c.acc;
      if (!std::isnan(c.energy)) r.energy = c.energy;
      r.time_step = time_step;
      r.time_current = time_current;
      if (c.status == CHAIN_FINISHED) r.status = VB_FINISHED;
      else if (c.status == CHAIN_NONFINITE) r.status = VB_ERR_NONFINITE;
      else if (c.status == CHAIN_UNDERFLOW) {
        r.status = VB_ERR_DT_UNDERFLOW;
        iter_open = false;   // as adaptive_complete() after an attempt that was not issued
      }
      if (r.status != VB_OK) break;
    }
  }

  void launch_steps(int n, vb_step_report* rep) override {
    hs_state = false;
    if (!started) start();
    cur_E_known = false;
    if (time_step < 1e-23) {
      rep->status = VB_ERR_DT_UNDERFLOW;
      return;
    }
    launch_fixed(n, *rep);
  }

  void finish_steps(vb_step_report* rep) override {
    if (drift_mode()) sync_targets_from_device();
    if (rep->status == VB_OK) {
      ensure_prefixed();   // the next step's first stage needs it anyway
      read_scal();
      if (hs->nonfinite) rep->status = VB_ERR_NONFINITE;
    }
    rep->time_step = time_step;
    rep->time_current = time_current;
  }

  void read_scal() {
    hs_state = false;
    VB_HIP(hipMemcpyAsync(hs, scal.p, sizeof(Scal), hipMemcpyDeviceToHost, st));
    VB_HIP(hipStreamSynchronize(st));
  }
  void read_scal64() {
    hs_state = false;
    VB_HIP(hipMemcpyAsync(hs, scal64.p, sizeof(Scal), hipMemcpyDeviceToHost, st));
    VB_HIP(hipStreamSynchronize(st));
  }

  // make candidate `out` current (P, Q already prepared for it by eval)
  void accept(int out) {
    cur = out;
  }

  // ---- drivers ----------------------------------------------------------------
  void set_targets_device() { launch_fill_scal(scal.p, alpha, beta, alpha0, beta0, st); }

  // One Euler attempt of the shape-adjust and run_static loops: the update of
  // uh[cur] with step t into uh[out], and the prefix of the candidate (its
  // integrals in scal; v, f and W for the next update if it is kept), which
  // becomes cur.  With graphs = 2 the launches replay one captured graph per
  // pipeline state and kind, whatever t: the update reads t from euler_t.
  // Eager launches are the default: measured per attempt (FP32-mixed, the host
  // reading E after each), the graph form took 72 / 97 / 152 us at 32^3 /
  // 64^3 / 128^3 against 76 / 91 / 142 us eager.
  DBuf<double> euler_t;
  void euler_attempt(double t, bool shape) {
    auto body = [&](bool tdev) {
      const int out = euler(t, shape, tdev ? euler_t.p : nullptr);
      prep(uh[out].p);
      cur = out;
      prefix(FIN_PLAIN, true);
    };
    if (p.graphs < 2 || prof) return body(false);
    if (!euler_t.p) euler_t.alloc(1);
    launch_set_double(euler_t.p, t, st);
    const int kind = shape ? kShapeAttempt : kEulerAttempt;
    StepGraph* G = nullptr;
    const PipeState now = pipe_state();
    for (auto& H : graphs)
      if (H.x && H.kind == kind && H.pre.cur == now.cur && H.pre.prepped == now.prepped &&
          H.pre.prefixed == now.prefixed && H.pre.prefixed_f == now.prefixed_f)
        G = &H;
    if (!G) G = &capture(kind, [&] {
      body(true);
      return 1;
    });
    VB_HIP(hipGraphLaunch(G->x, st));
    set_pipe_state(G->post);
  }

  // The shape-adjust loop with its decisions on the device (accept_chain > 0
  // with graphs): chains of kShapeSlots attempts, k_shape_begin / k_shape_end
  // (vb_chain.hip) doing what shape_adjust() does on the host with the same
  // doubles.  Before each update the state's Euler prefix (v, f, W) is copied
  // aside; an accepted candidate is copied over the state (its prefix is then
  // the state's), a rejected one, or an attempt after the loop ended, gets the
  // prefix copied back.  The host reads the control block once per chain.
  static constexpr int kShapeSlots = 8;
  DBuf<ShapeCtl> shape_ctl;
  ShapeCtl* hshape = nullptr;   // pinned
  DBuf<T> v_keep, f_keep;
  DBuf<cpx<T>> W_keep;

  int shape_chain_body(int K) {
    for (int k = 0; k < K; ++k) {
      launch_shape_begin(shape_ctl.p, scal.p, st);
      ChainCopy b{};
      b.n = 3;
      b.job[0] = {v.p, v_keep.p, v.n * sizeof(T)};
      b.job[1] = {f.p, f_keep.p, f.n * sizeof(T)};
      b.job[2] = {W.p, W_keep.p, W.n * sizeof(cpx<T>)};
      launch_cond_copy(&shape_ctl.p->one, 1, b, st);
      const int keep = cur;
      const int out = euler(0.0, true, &shape_ctl.p->tstep);
      prep(uh[out].p);
      cur = out;
      prefix(FIN_PLAIN, true);
      cur = keep;
      launch_shape_end(shape_ctl.p, scal.p, st);
      ChainCopy a{};
      a.n = 1;
      a.job[0] = {uh[out].p, uh[keep].p, uh[keep].n * sizeof(cpx<double>)};
      launch_cond_copy(&shape_ctl.p->accept, 1, a, st);
      ChainCopy r{};
      r.n = 3;
      r.job[0] = {v_keep.p, v.p, v.n * sizeof(T)};
      r.job[1] = {f_keep.p, f.p, f.n * sizeof(T)};
      r.job[2] = {W_keep.p, W.p, W.n * sizeof(cpx<T>)};
      launch_cond_copy(&shape_ctl.p->accept, 0, r, st);
      check_launch();
      prepped = false;
      prefixed = false;
      prefixed_f = true;
    }
    return K;
  }

  void shape_adjust_chain() {
    ensure_prefixed(true);
    read_scal();
    if (!shape_ctl.p) {
      shape_ctl.alloc(1);
      v_keep.alloc(v.n);
      f_keep.alloc(f.n);
      W_keep.alloc(W.n);
      VB_HIP(hipHostMalloc((void**)&hshape, sizeof(ShapeCtl), hipHostMallocDefault));
    }
    ShapeCtl& c = *hshape;
    memset(&c, 0, sizeof(c));
    c.tcur = 0.0;
    c.dt = 1e-7;
    c.eng = c.eng0 = hs->Eu;
    c.eta = p.eta;
    c.nst = 1;
    c.status = SHAPE_RUN;
    c.one = 1;
    VB_HIP(hipMemcpyAsync(shape_ctl.p, hshape, sizeof(ShapeCtl), hipMemcpyHostToDevice, st));
    StepGraph* G = find_chain(kShapeSlots, kShapeChain);
    if (!G) G = &capture(kShapeChain, [&] { return shape_chain_body(kShapeSlots); });
    for (;;) {
      VB_HIP(hipGraphLaunch(G->x, st));
      set_pipe_state(G->post);
      VB_HIP(hipMemcpyAsync(hshape, shape_ctl.p, sizeof(ShapeCtl), hipMemcpyDeviceToHost, st));
      VB_HIP(hipStreamSynchronize(st));
      if (c.status != SHAPE_RUN) break;
    }
    hs_state = false;
  }

  // double to int as the x86 conversion does it (truncation; INT_MIN out of
  // range or NaN), which k_shape_begin repeats on the device
  static int trunc_int(double x) { return (x < 2147483648.0 && x > -2147483649.0) ? int(x) : INT_MIN; }

  // Forward-Euler shape relaxation, unconstrained (belastic.cpp:174-283)
  void shape_adjust() {
    double tcur = 0.0;
    int nst = 1;
    double dt = 1e-7;
    ensure_prefixed(true);
    read_scal();
    double eng0 = hs->Eu, eng1 = eng0, eng = eng0;
    for (int it = 0; it < 1000000; ++it) {
      tcur += dt;
      nst++;
      if (dt > 1e-4 || trunc_int(tcur / dt) % trunc_int(1e-4 / dt) == 0) {
        ensure_prefixed(true);
        read_scal();
        eng1 = hs->Eu;
        if (eng1 < 100 * p.eta || fabs((eng - eng1) / eng) < 0.02) break;
        eng = eng1;
      }
      if (nst % 10 == 0) dt *= 2;
      for (;;) {
        const int keep = cur;
        euler_attempt(-dt * 3, true);
        read_scal();
        eng1 = hs->Eu;
        if (eng1 < eng0) {
          eng0 = eng1;
          break;
        }
        cur = keep;   // restore (belastic.cpp:249-250)
        prefixed = prefixed_f = prepped = false;
        dt /= 2;
        if (dt < 1e-13) break;
      }
      if (dt < 1e-13) break;
      if (hs->nonfinite) break;
    }
  }

  void start() {
    started = true;
    cur_E_known = false;
    if (p.shape_adjust && time_current == 0.0) {
      if (p.accept_chain > 0 && p.graphs && !prof) shape_adjust_chain();
      else shape_adjust();
    }
    ensure_prefixed(p.scheme == VB_EULER);
    if (beta <= 0.0) {   // belastic.cpp:588-591
      read_scal();
      alpha = hs->V;
      beta = hs->A;
    }
    alpha0 = alpha;   // belastic.cpp:592-593
    beta0 = beta;
    set_targets_device();
    refinalize(FIN_PLAIN);
    if (p.fixed_time_step > 0) time_step = std::min(p.fixed_time_step, time_step);   // belastic.cpp:599-601
  }

  void sync_targets_from_device() {
    read_scal();
    alpha = hs->alpha;
    beta = hs->beta;
  }

  void step(int n, vb_step_report* rep) override {
    vb_step_report r{};
    r.status = VB_OK;
    r.energy = NAN;
    if (!started) start();
    const bool adaptive = p.fixed_time_step <= 0;
    const bool drift = !adaptive && (p.M1 != 0 || p.M2 != 0) && p.penalty_drift;
    if (!adaptive || p.scheme == VB_EULER) cur_E_known = false;
    for (int it = 0; it < n; ++it) {
      double eng0 = NAN;
      if (p.scheme == VB_EULER) {
        // run_static loop body (belastic.cpp:367-437)
        ensure_prefixed(true);
        if (adaptive) {
          read_scal();
          eng0 = hs->E;
          if (steps % 4 == 0) time_step = time_step * 1.4;
          for (;;) {
            if (time_step < 1e-23) break;
            const int keep = cur;
            euler_attempt(-time_step * p.ga, false);
            r.evaluations++;
            read_scal();
            if (hs->E > eng0) {
              cur = keep;
              prefixed = prefixed_f = prepped = false;
              time_step /= 2;
              r.rejections++;
              ensure_prefixed(true);
              continue;
            }
            r.energy = hs->E;
            break;
          }
        } else {
          const int out = euler(-time_step * p.ga, false);
          r.evaluations++;
          cur = out;
          prepped = prefixed = prefixed_f = false;
        }
        if (time_step < 1e-23) { r.status = VB_ERR_DT_UNDERFLOW; break; }
        if (time_current > p.tmax) { r.status = VB_FINISHED; break; }
      } else if (adaptive) {
        if (chain_ok()) {   // the decisions on the device, a read-back per chain
          chain_steps(n - it, r);
          break;
        }
        // belastic.cpp:639-669: attempts until one is accepted
        int acc = 0;
        do {
          adaptive_issue(r);
          acc = adaptive_complete(r);
        } while (!acc);
        if (r.status != VB_OK) break;
        continue;   // time, steps and the exits were handled on acceptance
      } else {
        // fixed dt (belastic.cpp:670-693): no host decision inside the loop
        if (time_step < 1e-23) { r.status = VB_ERR_DT_UNDERFLOW; break; }
        launch_fixed(n - it, r);
        break;
      }
      // exits (belastic.cpp:713-740)
      if (time_step < 1e-23) { r.status = VB_ERR_DT_UNDERFLOW; break; }
      if (adaptive && p.scheme != VB_EULER && eng0 < 1e-3) { r.status = VB_FINISHED; break; }
      if (adaptive) {
        if (hs->nonfinite) { r.status = VB_ERR_NONFINITE; break; }
      }
      time_current += time_step;
      steps++;
      r.steps++;
    }
    if (drift) sync_targets_from_device();
    if (r.status == VB_OK && !adaptive) {
      // the blow-up test of the driver (belastic.cpp:733-740), on the integrals
      ensure_prefixed(p.scheme == VB_EULER);
      read_scal();
      if (hs->nonfinite) r.status = VB_ERR_NONFINITE;
    }
    r.time_step = time_step;
    r.time_current = time_current;
    if (rep) *rep = r;
  }

  void need64() {
    if (P64.p) return;
    P64.alloc((size_t)g.ncp);
    Q64.alloc((size_t)g.ncp);
    dc64.alloc(1);
    scal64.alloc(1);
    nxb64 = int((npairs + oxd->xlanes - 1) / oxd->xlanes);
    nzb64 = int(zgroups(g, ozd->ytile));
    xpart64.alloc((size_t)nxb64 * 3);
    zpart64.alloc((size_t)nzb64);
  }

  // FP32-mixed: V, A, E_b and E of uh[cur] evaluated entirely in FP64 into
  // scal64 (launches only): FP64 z and y inverse transforms of the FP64 state
  // and of its Laplacian, the FP64 x inverse with f and the partial sums
  // (k_xrow without its forward transform), the finalize kernel under the
  // targets of scal.  need64() first.
  void energy64() {
    const cpx<double>* src = uh[cur].p;
    if (ozd->ytile != oz->ytile) {   // the FP64 z kernels have their own column tile
      L("regroup", [&] { relayout(uh[cur].p, oz->ytile, S.p, ozd->ytile); });
      src = S.p;
    }
    L("zprep64", [&] { ozd->zprep(src, P64.p, Q64.p, zpart64.p, dc64.p, g, twzd.p, st); });
    L("y_inv2_64", [&] { oxd->ypass(P64.p, Q64.p, 2, +1, g, twxd.p, nullptr, st); });
    XaArgs<double> a{};
    a.P = P64.p;
    a.Q = Q64.p;
    a.part = xpart64.p;
    a.tw = twxd.p;
    a.npairs = npairs;
    a.eps = p.eps;
    a.kappa = p.kappa;
    a.D = p.D;
    a.inv_n = 1.0 / double(g.nr);
    a.rl = Rows{g.ny, g.pitch, g.pp, g.nq};
    L("xsums64", [&] { oxd->xa(a, XA_ROWSUMS, st); });
    launch_keep(OP_KEEP, scal.p, scal64.p, ph, st);   // the targets
    L("finalize64", [&] { launch_finalize(xpart64.p, nxb64, zpart64.p, dc64.p, nzb64, scal64.p, g, ph, FIN_PLAIN, st); });
    check_launch();
  }

  void query(double* E, double* V, double* A, double* Eb) override {
    if (sizeof(T) == sizeof(double)) {
      ensure_integrals();
      read_scal();
    } else {
      if (targets_drift()) sync_targets_from_device();
      need64();
      energy64();
      read_scal64();
    }
    if (E) *E = hs->E;
    if (V) *V = hs->V;
    if (A) *A = hs->A;
    if (Eb) *Eb = hs->Eb;
  }

  // the targets move on the device only under the fixed-dt penalty drift
  bool targets_drift() const { return started && fixed_dt() && drift_mode(); }

  void state(double* a, double* b, double* dt, double* t) override {
    if (targets_drift()) sync_targets_from_device();
    if (a) *a = alpha;
    if (b) *b = beta;
    if (dt) *dt = time_step;
    if (t) *t = time_current;
  }

  void set_state(double a, double b, double dt, double t, double a0, double b0) override {
    cur_E_known = false;
    iter_open = false;
    alpha = a;
    beta = b;
    alpha0 = a0;
    beta0 = b0;
    time_step = dt;
    time_current = t;
    set_targets_device();
    if (prefixed || prefixed_f) refinalize(FIN_PLAIN);
    started = t > 0.0 || started;
  }

  void get_phi(double* host) override {
    L("to_rows", [&] { relayout(uh[cur].p, oz->ytile, S.p, 0); });
    L("z_inv", [&] { ozd->zpass(S.p, +1, g, twzd.p, st); });
    L("y_inv", [&] { oxd->ypass(S.p, nullptr, 1, +1, g, twxd.p, nullptr, st); });
    L("x_c2r", [&] { oxd->xc2r_d(S.p, phid.p, npairs, twxd.p, g, st); });
    check_launch();
    VB_HIP(hipMemcpyAsync(host, phid.p, sizeof(double) * (size_t)g.nr, hipMemcpyDeviceToHost, st));
    VB_HIP(hipStreamSynchronize(st));
  }

  void set_phi(const double* host) override {
    cur_E_known = false;
    hs_state = false;
    iter_open = false;
    VB_HIP(hipMemcpyAsync(phid.p, host, sizeof(double) * (size_t)g.nr, hipMemcpyHostToDevice, st));
    L("x_r2c", [&] { oxd->xr2c_d(phid.p, S.p, npairs, twxd.p, 1.0 / double(g.nr), g, st); });
    L("y_fwd", [&] { oxd->ypass(S.p, nullptr, 1, -1, g, twxd.p, nullptr, st); });
    L("z_fwd", [&] { ozd->zpass(S.p, -1, g, twzd.p, st); });
    L("to_groups", [&] { relayout(S.p, 0, uh[cur].p, oz->ytile); });
    check_launch();
    VB_HIP(hipStreamSynchronize(st));
    prepped = prefixed = prefixed_f = false;
  }

  // the state spectrum in the reference layout [nz][ny][nx/2+1] (the
  // solver's row layout: spec_pos)
  void get_spec(double* host) override {
    L("to_rows", [&] { relayout(uh[cur].p, oz->ytile, S.p, 0); });
    check_launch();
    std::vector<cpx<double>> h((size_t)g.ncp);
    VB_HIP(hipMemcpyAsync(h.data(), S.p, sizeof(cpx<double>) * h.size(), hipMemcpyDeviceToHost, st));
    VB_HIP(hipStreamSynchronize(st));
    const int nh = g.nx / 2 + 1;
    for (int z = 0; z < g.nz; ++z)
      for (int y = 0; y < g.ny; ++y)
        for (int k = 0; k < nh; ++k) {
          const cpx<double>& v = h[(size_t)spec_pos(g, z, y, k)];
          double* o = host + 2 * (((size_t)z * g.ny + y) * nh + k);
          o[0] = v.x;
          o[1] = v.y;
        }
  }

  void set_spec(const double* host) override {
    cur_E_known = false;
    hs_state = false;
    iter_open = false;
    const int nh = g.nx / 2 + 1;
    std::vector<cpx<double>> h((size_t)g.ncp, cpx<double>{0.0, 0.0});
    for (int z = 0; z < g.nz; ++z)
      for (int y = 0; y < g.ny; ++y)
        for (int k = 0; k < nh; ++k) {
          const double* i = host + 2 * (((size_t)z * g.ny + y) * nh + k);
          h[(size_t)spec_pos(g, z, y, k)] = cpx<double>{i[0], i[1]};
        }
    VB_HIP(hipMemcpyAsync(S.p, h.data(), sizeof(cpx<double>) * h.size(), hipMemcpyHostToDevice, st));
    L("to_groups", [&] { relayout(S.p, 0, uh[cur].p, oz->ytile); });
    check_launch();
    VB_HIP(hipStreamSynchronize(st));
    prepped = prefixed = prefixed_f = false;
  }

  // belastic::geterror (belastic.cpp:136-143): I[r^2; 0, 1] of the Euler
  // right-hand side r = lambda1/2 + lambda2 f + eta g0 (belastic.cpp:52-64)
  double residual() override {
    if (!started) start();
    if (!W.p) W.alloc((size_t)g.ncp);
    if (!v.p) v.alloc((size_t)g.nr);
    if (!f.p) f.alloc((size_t)g.nr);
    if (!rpart.p) rpart.alloc((size_t)nxb);
    ensure_prefixed(true);
    XeArgs<T> e;
    e.W = W.p;
    e.v = v.p;
    e.f = f.p;
    e.out = nullptr;
    e.sc = scal.p;
    e.tw = twx.p;
    e.npairs = npairs;
    e.eps = p.eps;
    e.eta = p.eta;
    e.t = 0.0;
    e.inv_n = 1.0 / double(g.nr);
    e.shape = 0;
    e.rl = Rows{g.ny, g.pitch, g.pp, g.nq};
    e.part = rpart.p;
    e.tp = nullptr;
    L("xe_resid", [&] { ox->xe(e, st); });
    check_launch();
    std::vector<double> h((size_t)nxb);
    VB_HIP(hipMemcpyAsync(h.data(), rpart.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, st));
    VB_HIP(hipStreamSynchronize(st));
    double sum = 0.0;
    for (double x : h) sum += x;
    return sum / double(g.nr) * g.L3;   // Calc.cu:250,260 with a = 0, m = 1
  }

  bool accepted_values(double* E, double* V, double* A, double* Eb) const override {
    if (!hs_state) return false;
    if (E) *E = hs->E;
    if (V) *V = hs->V;
    if (A) *A = hs->A;
    if (Eb) *Eb = hs->Eb;
    return true;
  }

  bool reset(const vb_params& pp, const double* phi0) override {
    if (pp.nx != p.nx || pp.nz != p.nz || pp.precision != p.precision) return false;
    if ((pp.scheme == VB_EULER || pp.shape_adjust) && !W.p) return false;
    // the last work of the previous solve has completed (its attempt's event
    // or a synchronising call); graphs and tables hold the old parameters
    for (auto& G : graphs)
      if (G.x) {
        VB_HIP(hipGraphExecDestroy(G.x));
        G.x = nullptr;
      }
    graph_next = 0;
    for (auto& sl : slots) sl.sch = 0;
    flush_prof();
    prof = false;
    p = pp;
    setup_geo();
    cur = 0;
    prepped = prefixed = prefixed_f = false;
    started = false;
    time_step = p.time_step;
    time_current = 0.0;
    steps = 0;
    alpha = alpha0 = p.alpha;
    beta = beta0 = p.beta;
    have_fixed_post = false;
    rec_evals = 0;
    att_pending = iter_open = cur_E_known = hs_state = false;
    info_extra[0] = info_extra[1] = info_extra[2] = 0.0;
    set_targets_device();
    set_phi(phi0);
    return true;
  }

  void evaluate(double h) override {
    cur_E_known = false;
    hs_state = false;
    const int out = eval(h);
    cur = out;
  }
};

std::unique_ptr<SolverBase> make_solver(const vb_params& p, hipStream_t shared) {
  if (p.precision == VB_FP32_MIXED) return std::make_unique<Solver<float>>(p, shared);
  return std::make_unique<Solver<double>>(p, shared);
}

size_t solver_footprint(const vb_params& p) {
  const size_t nh = (size_t)p.nx / 2 + 1;
  const size_t nc = nh * (size_t)p.nx * (size_t)p.nz, nr = (size_t)p.nx * p.nx * p.nz;
  const bool mixed = p.precision == VB_FP32_MIXED;
  const size_t cs = mixed ? 8 : 16, rs = cs / 2;
  const bool euler = p.scheme == VB_EULER || p.shape_adjust;
  const size_t work = 10 + (euler ? 1 : 0);   // P, Q, Tt, Ts, Tt1, Ts1, pre, Na, Nb, Vh (+ W)
  size_t bytes = 3 * nc * 16 + work * nc * cs + (euler ? 2 * nr * rs : 0) + nr * 8;
  if (mixed) bytes += 2 * nc * 16;   // FP64 integrals of the state (P64, Q64)
  const size_t mm = (size_t)mmax_of(p.nx, p.nx, p.nz) + 1;
  return bytes + (size_t)Solver<double>::kLutSlots * mm * 48 + ((size_t)4 << 20);
}

}  // namespace vb
