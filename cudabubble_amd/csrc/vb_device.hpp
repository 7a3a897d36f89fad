// vb_device.hpp -- shared device/host definitions of the vesicle step kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "vb_fft.hpp"

namespace vb {

// Grid geometry.  Real fields are [nz][ny][nx] (x fastest); half spectra are
// [nz][ny][nh] with nh = nx/2 + 1 (reference layout, Calc.cu:57-58).
struct Geo {
  int nx, ny, nz, nh;
  long long nr, nc;     // real points, half-spectrum points
  double kfac;          // 2 pi / L
  double kfac2;         // (2 pi / L)^2
  double L3;            // L^3
  // Row pitch of half spectra in elements.  pitch = nh is the reference
  // layout; the solver pads rows to nx/2 + 16 so every row starts on a
  // 128-byte boundary, and then (aligned = 1) the y and z kernels tile
  // kx in [0, nx/2) with aligned column groups and treat the Nyquist column
  // kx = nx/2 as a separate group.
  int pitch, aligned;
  long long pp;         // plane pitch (z stride) of spectra in elements
  long long ncp;        // allocated spectrum elements, nz * pp
};

// Physical constants of one solve (reference names).
struct Phys {
  double eps, eta, ga, kappa, D, M1, M2;
};

// Device-resident scalars of one solve.  Written by the finalize kernel from
// the block partial sums of the last prefix; read by the stage kernels.
struct Scal {
  double V, A, Eb, E, Eu;      // integrals of the last prefix field; E with
                               // penalties, Eu = eta * Eb (unconstrained)
  double l1, l2;               // penalty multipliers in use
  double alpha, beta;          // current targets
  double alpha0, beta0;        // targets at t = 0 (drift reference)
  double sum_v, sum_w, sum_f2, grad2;
  int nonfinite, pad;
};

// ETD coefficient record, indexed by the integer m = kx^2 + ky'^2 + kz'^2
// (folded indices).  ETDRK4: {e1, (e1-1)/H, e2, A s, 2 B s, C s, -, H} with
// s = 1/(h^2 H^3).  ETDRK2: {e, (e-1)/H, c2, -, -, -, -, H}.
struct alignas(64) Coef {
  double c[8];
};

// finalize modes: PLAIN computes the integrals and multipliers; DRIFT first
// moves the targets, alpha -= V - alpha0, beta -= A - beta0
// (belastic.cpp:680-688), then computes the multipliers with them.
enum { FIN_PLAIN = 0, FIN_DRIFT = 1 };

__device__ __forceinline__ int fold(int k, int n) { return k < n - k ? k : n - k; }
__device__ __forceinline__ int sgnk(int k, int n) { return k < n - k ? k : k - n; }

// block-wide deterministic sum of K doubles; thread 0..K-1 write out[q].
// Uses LDS scratch of K * (NT / 64) doubles (declared by the caller).
template <int K, int NT>
__device__ __forceinline__ void block_sum_store(double (&acc)[K], double* red, double* out) {
#pragma unroll
  for (int q = 0; q < K; ++q) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc[q] += __shfl_xor(acc[q], off, 64);
  }
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) {
#pragma unroll
    for (int q = 0; q < K; ++q) red[q * (NT / 64) + w] = acc[q];
  }
  __syncthreads();
  if (threadIdx.x < K) {
    double s = 0.0;
    for (int i = 0; i < NT / 64; ++i) s += red[threadIdx.x * (NT / 64) + i];
    out[threadIdx.x] = s;
  }
}

}  // namespace vb
