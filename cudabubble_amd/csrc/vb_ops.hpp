// vb_ops.hpp -- per-transform-length launch table.  Each supported length N
// is compiled in its own translation unit (vb_n<N>.hip) so the build runs in
// parallel; the host pipeline picks the table of nx (x and y kernels) and the
// table of nz (z kernels) at solver creation.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "vb_device.hpp"

namespace vb {

// Launch recording for batched launches.  While a Recorder is installed on
// the calling thread, the pipeline launchers append each launch (its kernel,
// its argument struct and its grid) instead of issuing it; the batched form
// then issues one launch per recorded position for all solves of a size
// group, with the solves' argument structs in a device table.
enum PipeOp {
  OP_XA_RS, OP_XA_F, OP_XA_S, OP_XA_SUMS, OP_Y_FWD, OP_Y_INV, OP_Y_FWD_GM, OP_ZLAP, OP_ZPREP,
  OP_ZS41, OP_ZS42, OP_ZS43, OP_ZS44, OP_ZS21, OP_ZS22, OP_FIN, OP_KEEP, OP_RETARGET, OP_ZS44P, OP_RELAYOUT
};
// x-kernel outputs: the partial sums of v, (v^2-1)^2 and f^2 only (k_xrow
// without its forward transform: the integrals of a field), tmp1 and f
// (Euler), tmp1 and s (spectral N), the partial sums of v and (v^2-1)^2 only
// (from P alone: k_xsums)
enum XaMode { XA_ROWSUMS = 0, XA_WITHF = 1, XA_WITHS = 2, XA_SUMS = 3 };
struct RecLaunch {
  // issues the table form of the recorded kernel over `lanes` solves
  void (*launch_tab)(int op, const void* tab, dim3 grid, dim3 block, int lanes, hipStream_t s);
  int op;
  std::vector<unsigned char> args;
  dim3 grid, block;
};
struct Recorder {
  std::vector<RecLaunch> seq;
};
extern thread_local Recorder* g_recorder;

template <class A>
inline void record_launch(void (*tabfn)(int, const void*, dim3, dim3, int, hipStream_t), int op, const A& a,
                          dim3 grid, dim3 block) {
  RecLaunch r;
  r.launch_tab = tabfn;
  r.op = op;
  r.args.assign(reinterpret_cast<const unsigned char*>(&a), reinterpret_cast<const unsigned char*>(&a) + sizeof(A));
  r.grid = grid;
  r.block = block;
  g_recorder->seq.push_back(std::move(r));
}

template <typename T> struct XaArgs;
template <typename T> struct XeArgs;
template <typename T> struct ZsArgs;

template <typename T>
struct SizeOps {
  int N;
  int sched;       // kernel schedule (z tile, x plan) this table launches
  int xr_passes;   // passes of the ETD x kernel's plan
  int xr_radix;    // radix of its first pass
  int xlanes;      // row pairs per x workgroup
  int ytile;       // columns per workgroup of the z stage kernels
  int ctile;       // columns per workgroup of the y passes and the z Laplacian
  int ythreads;    // threads per y / z workgroup
  void (*xr2c_d)(const double* in, cpx<T>* out, long long npairs, const cpx<T>* tw, double scl, const Geo& g,
                 hipStream_t s);
  void (*xr2c)(const T* in, cpx<T>* out, long long npairs, const cpx<T>* tw, double scl, const Geo& g, hipStream_t s);
  void (*xc2r_d)(const cpx<T>* in, double* out, long long npairs, const cpx<T>* tw, const Geo& g, hipStream_t s);
  void (*xc2r)(const cpx<T>* in, T* out, long long npairs, const cpx<T>* tw, const Geo& g, hipStream_t s);
  void (*xa)(const XaArgs<T>& a, int mode, hipStream_t s);   // mode: XaMode
  void (*xe)(const XeArgs<T>& a, hipStream_t s);
  // fin (or nullptr): integrals evaluated by extra workgroups of the launch
  void (*ypass)(cpx<T>* b0, cpx<T>* b1, int narr, int sign, const Geo& g, const cpx<T>* tw, const FinArgs* fin,
                hipStream_t s);
  // forward y pass of two arrays into the z kernels' group-major layout
  void (*ypass_gm)(const cpx<T>* b0, const cpx<T>* b1, cpx<T>* o0, cpx<T>* o1, int ztb, const Geo& g,
                   const cpx<T>* tw, const FinArgs* fin, hipStream_t s);
  void (*zpass)(cpx<T>* b, int sign, const Geo& g, const cpx<T>* tw, hipStream_t s);
  // fin (or nullptr): integrals evaluated by one extra workgroup of the launch
  void (*zlap)(cpx<T>* b, const Geo& g, const cpx<T>* tw, const FinArgs* fin, hipStream_t s);
  void (*zprep)(const cpx<double>* vh, cpx<T>* P, cpx<T>* Q, double* part, double* dc, const Geo& g, const cpx<T>* tw,
                hipStream_t s);
  // stg 1..sch; 5: the probe form of ETDRK4 stage 4 (V and A of the result
  // only: no state store, P alone)
  void (*zstage)(const ZsArgs<T>& a, int sch, int stg, hipStream_t s);
};

// launch table of kernel schedule sc of length N (nullptr if not compiled),
// and the number of schedules compiled for it
template <typename T, int N> const SizeOps<T>* ops_for(int sc);
template <typename T, int N> int sched_count();

#define VB_SIZE_LIST(X) X(16) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512)

#define VB_DECL(n)                                              \
  template <> const SizeOps<double>* ops_for<double, n>(int);   \
  template <> const SizeOps<float>* ops_for<float, n>(int);     \
  template <> int sched_count<double, n>();                     \
  template <> int sched_count<float, n>();
VB_SIZE_LIST(VB_DECL)
#undef VB_DECL

template <typename T>
inline int sched_count_of(int N) {
  switch (N) {
#define VB_CASE(n) case n: return sched_count<T, n>();
    VB_SIZE_LIST(VB_CASE)
#undef VB_CASE
    default: return 0;
  }
}

// The schedule in use for (precision, N): the autotuner's or the caller's
// choice (vb_schedule_set), else 0; a thread's override (>= 0) wins while
// set (the autotuner builds its candidate solvers under one).
int sched_selected(bool fp32, int N);
extern thread_local int g_sched_override;

template <typename T>
inline const SizeOps<T>* get_ops(int N, int sc) {
  switch (N) {
#define VB_CASE(n) case n: return ops_for<T, n>(sc);
    VB_SIZE_LIST(VB_CASE)
#undef VB_CASE
    default: return nullptr;
  }
}
template <typename T>
inline const SizeOps<T>* get_ops(int N) {
  int sc = g_sched_override >= 0 ? g_sched_override : sched_selected(sizeof(T) == 4, N);
  if (sc >= sched_count_of<T>(N)) sc = 0;
  return get_ops<T>(N, sc);
}

inline bool size_supported(int N) {
  switch (N) {
#define VB_CASE(n) case n: return true;
    VB_SIZE_LIST(VB_CASE)
#undef VB_CASE
    default: return false;
  }
}

}  // namespace vb
