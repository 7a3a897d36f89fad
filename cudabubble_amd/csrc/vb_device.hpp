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
  // layout (nq = -1: the Nyquist column kx = nx/2 ends each row).  The solver
  // (aligned = 1) keeps rows of kx in [0, nx/2) only, pitch = nx/2, so every
  // row is whole 128-byte lines, and stores the Nyquist column of a plane as
  // ny contiguous elements after its rows, at plane offset nq: the row kernels
  // then read no line for one element, and the y and z kernels tile kx in
  // [0, nx/2) with aligned column groups and read the Nyquist column as a
  // separate, contiguous group.
  int pitch, aligned;
  long long pp;         // plane pitch (z stride) of spectra in elements
  long long ncp;        // allocated spectrum elements, nz * pp
  long long nq;         // plane offset of the Nyquist column, or -1 (in the rows)
};

// Position of mode (z, ky, kx) of a half spectrum in the row layout.
__host__ __device__ inline long long spec_pos(const Geo& g, int z, int ky, int kx) {
  return (long long)z * g.pp + ((g.nq >= 0 && kx == g.nx / 2) ? g.nq + ky : (long long)ky * g.pitch + kx);
}

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
  // FP32-mixed stage fields: the means of tmp1 and of eps^3 s, summed in FP64
  // from the x kernel's values (the k = 0 coefficients of their spectra)
  double dc_t1, dc_s;
  int nonfinite, pad;
};

// The ETD tables of one dt.
struct LutSel {
  const double2* c1;
  const double4* c2;
};

// Control block of a chain of adaptive ETD attempts whose accept/reject
// decisions are taken on the device (belastic.cpp:639-669, 713-746).  The host
// writes it before the chain's launches and reads it back after them; between
// the two, k_chain_begin / k_chain_end (one per attempt) carry the driver's
// state: dt, time, the accepted steps, the open iteration's E0.  A chain stops
// making attempts at the first driver exit, after `budget` accepted
// iterations, or when dt leaves the tables it was given (CHAIN_NEED_LUT: the
// host adds the table and goes on).
constexpr int kChainLevels = 24;
enum { CHAIN_RUN = 0, CHAIN_FINISHED = 1, CHAIN_NONFINITE = 2, CHAIN_UNDERFLOW = 3, CHAIN_NEED_LUT = 4, CHAIN_BUDGET = 5 };
struct ChainCtl {
  double dt, dt_cap;      // the current step; the adaptive cap
  double t;               // simulated time
  double E_cur;           // E of the state
  double E0;              // E of the state when the open iteration began
  double energy;          // E of the last accepted candidate (NaN until one is)
  int steps;              // accepted iterations since the start of the run
  int dt_double_every;
  int iter_open;          // dt of the current iteration chosen
  int budget;             // accepted iterations this chain may make
  int status;             // CHAIN_*
  int mixed;              // the energy test reads the FP64 scalars
  int acc, evals, rej;    // this chain's accepted iterations, attempts, rejections
  int active, accept;     // the current attempt runs; it was accepted
  int nlev;
  double lev_dt[kChainLevels];
  LutSel lev[kChainLevels];
  LutSel sel;             // tables of the current attempt
  Scal saved;             // integrals of the state, put back after a rejection
};

// Control block of a chain of shape-adjust attempts decided on the device
// (belastic.cpp:174-283): the loop's time, iteration count, step and energies,
// carried by k_shape_begin / k_shape_end between the host's writes and reads.
enum { SHAPE_RUN = 0, SHAPE_DONE = 1 };
struct ShapeCtl {
  double tcur, dt;        // the loop's time and step
  double eng, eng0;       // energy at the last check; energy of the state
  double eta;             // the stop level is 100 eta
  double tstep;           // the update's factor of this attempt, -3 dt (read by k_xe)
  int nst, it;            // the loop's counters
  int inner_open;         // the current iteration's check and doubling are done
  int status;             // SHAPE_*
  int active, accept;     // the current attempt runs; it was accepted
  int one, pad;           // one = 1: the condition of unconditional copies
  int evals, rej;
  Scal saved;             // integrals of the state, put back after a rejection
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

// E, Eu and the multipliers from the integrals V, A, E_b and the targets
// alpha, beta of c (belastic.cpp:68-86 and 473-479).
__device__ __forceinline__ void apply_targets(Scal& c, const Phys& ph) {
  const double ve = c.V - c.alpha, se = c.A - c.beta;
  c.Eu = c.Eb * ph.eta;
  c.E = c.Eb * ph.eta + 0.5 * ph.M1 * ve * ve + 0.5 * ph.M2 * se * se;
  c.l1 = ph.M1 * ve;
  c.l2 = ph.M2 * se;
}

// Workgroup shares of a split evaluation of the integrals and the ticket that
// picks the workgroup which adds them; one per solve, zeroed at creation and
// left zeroed by every evaluation.
constexpr int kFinShares = 16;
struct FinScratch {
  double part[6 * kFinShares];
  unsigned ticket, pad;
};

// Inputs of the integrals of one prefix field: the block partial sums of the
// x kernel (sum v, sum w, sum f^2 per workgroup) and of the z kernel (Parseval
// |grad|^2), and the field's k = 0 coefficient.
struct FinArgs {
  const double* xpart;
  const double* zpart;
  const double* dc;   // k = 0 coefficient of the field (its mean)
  const double* dcpart;   // FP32-mixed: the x kernel's sums of tmp1 and eps^3 s (2 per block), or nullptr
  Scal* sc;
  FinScratch* fs;
  int nxb, nzb, mode, pad;
  Geo g;
  Phys ph;
};

// The integrals, evaluated by nb workgroups of NT threads (this is workgroup
// b): each sums a contiguous share of the partials in a fixed order; with
// nb > 1 the shares go to fs and the workgroup that takes the last ticket adds
// them in workgroup order, so the result does not depend on which one
// finishes last.  Thread 0 of the final workgroup writes V, A, E_b, E and the
// multipliers (Calc.cu:191-263, phasefield.cpp:283-325, belastic.cpp:68-86
// and 473-479), and with dcpart the means of tmp1 and eps^3 s.  red holds
// 6 * (NT / 64) doubles and tot 7 doubles of LDS.
// UN accumulator sets per thread keep UN loads in flight (fewer registers in
// kernels that carry this as a side branch).
template <int NT, int UN = 4>
__device__ __forceinline__ void finalize_body(const FinArgs& a, double* red, double* tot, int b, int nb) {
  static_assert(NT % 64 == 0, "whole waves");
  const int xs = (a.nxb + nb - 1) / nb, zs = (a.nzb + nb - 1) / nb;
  const int x0 = b * xs, x1 = min(a.nxb, x0 + xs);
  const int z0 = b * zs, z1 = min(a.nzb, z0 + zs);
  double u[UN][4] = {};
  for (int i0 = x0 + threadIdx.x; i0 < x1; i0 += UN * NT) {
#pragma unroll
    for (int k = 0; k < UN; ++k) {
      const int i = i0 + k * NT;
      if (i < x1) {
        u[k][0] += a.xpart[3 * i];
        u[k][1] += a.xpart[3 * i + 1];
        u[k][2] += a.xpart[3 * i + 2];
      }
    }
  }
  for (int i = z0 + threadIdx.x; i < z1; i += NT) u[0][3] += a.zpart[i];
  double s[6];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    s[q] = u[0][q];
#pragma unroll
    for (int k = 1; k < UN; ++k) s[q] += u[k][q];
  }
  s[4] = s[5] = 0.0;
  if (a.dcpart) {
    for (int i = x0 + threadIdx.x; i < x1; i += NT) {
      s[4] += a.dcpart[2 * i];
      s[5] += a.dcpart[2 * i + 1];
    }
  }
  if (nb == 1) {
    block_sum_store<6, NT>(s, red, tot);
    __syncthreads();
  } else {
    unsigned* last = reinterpret_cast<unsigned*>(tot + 6);
    block_sum_store<6, NT>(s, red, a.fs->part + 6 * b);
    __threadfence();   // publish the share before taking a ticket
    __syncthreads();
    if (threadIdx.x == 0) *last = atomicAdd(&a.fs->ticket, 1u) == unsigned(nb - 1) ? 1u : 0u;
    __syncthreads();
    if (!*last) return;
    __threadfence();
    if (threadIdx.x < 6) {
      const volatile double* sh = a.fs->part;
      double t = 0.0;
      for (int k = 0; k < nb; ++k) t += sh[6 * k + threadIdx.x];
      tot[threadIdx.x] = t;
    }
    if (threadIdx.x == 0) a.fs->ticket = 0u;
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    Scal c = *a.sc;
    const double N = double(a.g.nr);
    const double L3 = a.g.L3;
    const double eps = a.ph.eps;
    // vol: intf(d, 1, 1/2); surface: intf(s, 0, 3/(2 sqrt2)) with
    // s = eps/2 |grad d|^2 + (d^2-1)^2/(4 eps); elastic: intf(f^2, 0, 1/(2 eps)).
    // mean(d) is the k = 0 coefficient of the FP64 spectrum (Parseval gives
    // mean |grad d|^2 the same way); the spatial sum is kept as sum_v.
    const double V = (a.dc[0] + 1.0) * (0.5 * L3);
    const double ms = 0.5 * eps * tot[3] + 0.25 / eps * (tot[1] / N);
    const double A = ms * (3.0 / (2.0 * sqrt(2.0)) * L3);
    const double Eb = (tot[2] / N) * (1.0 / (eps * 2) * L3);
    if (a.mode == FIN_DRIFT) {
      c.alpha -= (V - c.alpha0);
      c.beta -= (A - c.beta0);
    }
    c.V = V;
    c.A = A;
    c.Eb = Eb;
    c.sum_v = tot[0];
    c.sum_w = tot[1];
    c.sum_f2 = tot[2];
    c.grad2 = tot[3];
    if (a.dcpart) {
      c.dc_t1 = tot[4] / N;
      c.dc_s = tot[5] / N;
    }
    c.nonfinite = !(isfinite(V) && isfinite(A) && isfinite(Eb)) ? 1 : 0;
    apply_targets(c, a.ph);
    *a.sc = c;
  }
}

}  // namespace vb
