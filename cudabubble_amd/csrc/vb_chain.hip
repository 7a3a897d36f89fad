// vb_chain.hip -- accept/reject of adaptive ETD attempts on the device.
//
// A chain is a run of attempts issued back to back without the host: before
// each attempt k_chain_begin opens the iteration (E0 and the dt doubling of
// belastic.cpp:639-651) and picks the coefficient tables of the dt, which
// k_chain_table copies to where the stage kernels read them; after it
// k_chain_end applies the energy test (belastic.cpp:653-668) and the exits of
// belastic.cpp:713-746; k_chain_copy then makes an accepted candidate the
// state.  The arithmetic is the host driver's (the same doubles, compared and
// halved the same way), so a chain takes the decisions the host would.
#include <algorithm>

#include "vb_host.hpp"

namespace vb {

namespace {

typedef unsigned int v4u __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(64) k_chain_begin(ChainCtl* c, const Scal* sc) {
  if (threadIdx.x != 0) return;
  // at the start of every attempt scal holds the integrals of the state
  c->saved = *sc;
  c->active = 0;
  c->accept = 0;
  if (c->status != CHAIN_RUN) return;
  if (c->acc >= c->budget) {
    c->status = CHAIN_BUDGET;
    return;
  }
  if (!c->iter_open) {
    c->E0 = c->E_cur;
    if (c->steps % c->dt_double_every == 0) c->dt = c->dt * 1.4 < c->dt_cap ? c->dt * 2 : c->dt_cap;
    c->iter_open = 1;
  }
  if (c->dt < 1e-23) {
    c->status = CHAIN_UNDERFLOW;
    return;
  }
  int l = -1;
  for (int i = 0; i < c->nlev; ++i)
    if (c->lev_dt[i] == c->dt) l = i;
  if (l < 0) {
    c->status = CHAIN_NEED_LUT;
    return;
  }
  c->sel = c->lev[l];
  c->active = 1;
}

// test: the scalars of the candidate the energy test reads (FP64 ones in
// FP32-mixed mode)
__global__ void __launch_bounds__(64) k_chain_end(ChainCtl* c, Scal* sc, const Scal* test) {
  if (threadIdx.x != 0) return;
  if (!c->active) {
    *sc = c->saved;   // the launches of an attempt that did not run changed only work buffers
    return;
  }
  const double E = test->E;
  c->evals++;
  if (E > c->E0) {   // rejected: the state stays, dt halves
    c->dt = c->dt / 2;
    c->rej++;
    *sc = c->saved;
    return;
  }
  c->accept = 1;
  c->iter_open = 0;
  c->E_cur = E;
  c->energy = E;
  if (c->E0 < 1e-3) c->status = CHAIN_FINISHED;
  else if (test->nonfinite) c->status = CHAIN_NONFINITE;
  if (c->status != CHAIN_RUN) return;
  c->t += c->dt;
  c->steps++;
  c->acc++;
}

// the copies that make an accepted candidate the state (nothing otherwise)
__global__ void __launch_bounds__(256) k_chain_copy(const ChainCtl* c, ChainCopy cp) {
  if (!c->accept) return;
  const long long stride = (long long)gridDim.x * 256;
  for (int j = 0; j < cp.n; ++j) {
    const v4u* src = static_cast<const v4u*>(cp.job[j].src);
    v4u* dst = static_cast<v4u*>(cp.job[j].dst);
    const long long n = (long long)(cp.job[j].bytes / 16);
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += stride)
      __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
    if (blockIdx.x == 0 && threadIdx.x == 0 && (cp.job[j].bytes & 15)) {   // an 8-byte tail
      const size_t t = cp.job[j].bytes / 8 - 1;
      static_cast<unsigned long long*>(cp.job[j].dst)[t] = static_cast<const unsigned long long*>(cp.job[j].src)[t];
    }
  }
}

// the tables of the current attempt's dt into the chain's own table buffers,
// which the stage kernels read (nrec records each)
__global__ void __launch_bounds__(256) k_chain_table(const ChainCtl* c, double2* c1, double4* c2, int nrec) {
  if (!c->active) return;
  const double2* s1 = c->sel.c1;
  const double4* s2 = c->sel.c2;
  for (int i = int(blockIdx.x * 256 + threadIdx.x); i < nrec; i += int(gridDim.x * 256)) {
    c1[i] = s1[i];
    c2[i] = s2[i];
  }
}

__global__ void __launch_bounds__(64) k_set_double(double* d, double v) {
  if (threadIdx.x == 0) *d = v;
}

}  // namespace

void launch_set_double(double* d, double v, hipStream_t s) {
  hipLaunchKernelGGL(k_set_double, dim3(1), dim3(64), 0, s, d, v);
}

void launch_chain_begin(ChainCtl* c, const Scal* sc, hipStream_t s) {
  hipLaunchKernelGGL(k_chain_begin, dim3(1), dim3(64), 0, s, c, sc);
}

void launch_chain_table(const ChainCtl* c, double2* c1, double4* c2, int nrec, hipStream_t s) {
  const unsigned blocks = unsigned(std::min(1024, std::max(1, (nrec + 255) / 256)));
  hipLaunchKernelGGL(k_chain_table, dim3(blocks), dim3(256), 0, s, c, c1, c2, nrec);
}

void launch_chain_end(ChainCtl* c, Scal* sc, const Scal* test, hipStream_t s) {
  hipLaunchKernelGGL(k_chain_end, dim3(1), dim3(64), 0, s, c, sc, test);
}

void launch_chain_copy(const ChainCtl* c, const ChainCopy& cp, hipStream_t s) {
  size_t most = 0;
  for (int j = 0; j < cp.n; ++j) {
    if (cp.job[j].bytes % 8 != 0) throw HipError{hipErrorInvalidValue, "chain copy: size not a multiple of 8 bytes"};
    most = std::max(most, cp.job[j].bytes / 16);
  }
  const unsigned blocks = unsigned(std::min<size_t>(4096, std::max<size_t>(1, (most + 1023) / 1024)));
  hipLaunchKernelGGL(k_chain_copy, dim3(blocks), dim3(256), 0, s, c, cp);
}

}  // namespace vb
