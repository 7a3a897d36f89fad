// vb_chain.hip -- accept/reject of adaptive ETD attempts on the device.
//
// The shape-adjust loop (belastic.cpp:174-283) runs the same way:
// k_shape_begin / k_shape_end carry its time, counters, step and energies, an
// accepted candidate is copied over the state, and a rejection puts the
// state's Euler prefix back from the copy taken before the update.
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

// x86 conversion of a double to int (truncation; INT_MIN out of range or NaN),
// the host's, so that the check of the shape-adjust loop agrees with it
__device__ __forceinline__ int trunc_int(double x) {
  return (x < 2147483648.0 && x > -2147483649.0) ? int(x) : int(0x80000000u);
}

// Head of an iteration of the shape-adjust loop (belastic.cpp:196-236), when
// none is open, and the attempt's step: the energy check every 1e-4 of time
// (or every step once dt > 1e-4) ends the loop below 100 eta or at a relative
// change under 2 %; dt doubles every tenth iteration.
__global__ void __launch_bounds__(64) k_shape_begin(ShapeCtl* c, const Scal* sc) {
  if (threadIdx.x != 0) return;
  c->saved = *sc;
  c->active = 0;
  c->accept = 0;
  if (c->status != SHAPE_RUN) return;
  if (!c->inner_open) {
    if (c->it >= 1000000) {
      c->status = SHAPE_DONE;
      return;
    }
    c->tcur += c->dt;
    c->nst++;
    const double dt = c->dt;
    if (dt > 1e-4 || trunc_int(c->tcur / dt) % trunc_int(1e-4 / dt) == 0) {
      const double eng1 = c->eng0;   // the state's energy (its prefix is the last accepted candidate's)
      if (eng1 < 100 * c->eta || fabs((c->eng - eng1) / c->eng) < 0.02) {
        c->status = SHAPE_DONE;
        return;
      }
      c->eng = eng1;
    }
    if (c->nst % 10 == 0) c->dt = c->dt * 2;
    c->inner_open = 1;
  }
  c->tstep = -c->dt * 3;
  c->active = 1;
}

// The attempt's energy test (belastic.cpp:238-262): accepted below the
// state's energy; otherwise dt halves, and the loop ends below 1e-13.  An
// accepted non-finite state ends it too.
__global__ void __launch_bounds__(64) k_shape_end(ShapeCtl* c, Scal* sc) {
  if (threadIdx.x != 0) return;
  if (!c->active) {
    *sc = c->saved;
    return;
  }
  c->evals++;
  const double eng1 = sc->Eu;
  if (eng1 < c->eng0) {
    c->eng0 = eng1;
    c->accept = 1;
    c->inner_open = 0;
    c->it++;
    if (sc->nonfinite) c->status = SHAPE_DONE;
    return;
  }
  *sc = c->saved;
  c->rej++;
  c->dt = c->dt / 2;
  if (c->dt < 1e-13) c->status = SHAPE_DONE;
}

// copies of up to three buffers when *flag == want
__global__ void __launch_bounds__(256) k_cond_copy(const int* flag, int want, ChainCopy cp) {
  if (*flag != want) return;
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

void launch_cond_copy(const int* flag, int want, const ChainCopy& cp, hipStream_t s) {
  size_t most = 0;
  for (int j = 0; j < cp.n; ++j) {
    if (cp.job[j].bytes % 8 != 0) throw HipError{hipErrorInvalidValue, "chain copy: size not a multiple of 8 bytes"};
    most = std::max(most, cp.job[j].bytes / 16);
  }
  const unsigned blocks = unsigned(std::min<size_t>(4096, std::max<size_t>(1, (most + 1023) / 1024)));
  hipLaunchKernelGGL(k_cond_copy, dim3(blocks), dim3(256), 0, s, flag, want, cp);
}

void launch_chain_copy(const ChainCtl* c, const ChainCopy& cp, hipStream_t s) {
  launch_cond_copy(&c->accept, 1, cp, s);
}

void launch_shape_begin(ShapeCtl* c, const Scal* sc, hipStream_t s) {
  hipLaunchKernelGGL(k_shape_begin, dim3(1), dim3(64), 0, s, c, sc);
}

void launch_shape_end(ShapeCtl* c, Scal* sc, hipStream_t s) {
  hipLaunchKernelGGL(k_shape_end, dim3(1), dim3(64), 0, s, c, sc);
}

}  // namespace vb
