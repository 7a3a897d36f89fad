// vb_kernels.hpp -- kernel templates of the fused pseudo-spectral step.
//
// Pass structure of one nonlinear evaluation N^(v) (reference:
// belastic::get_nld_term, belastic.cpp:473-488), with v^ the stage spectrum:
//
//   z-kernel  : v^ -> z-inverse of v^ and of -|k|^2 v^   (+ Parseval |grad v|^2)
//   y-kernel  : y-inverse of both
//   xrow      : x-inverse of v^ + i lap^ v -> v, lap v; f, tmp1, s, sums of
//               v, (v^2-1)^2, f^2; x-forward of tmp1 + i s
//   y-kernel  : y-forward of both; extra workgroups of the launch evaluate
//               V, A, E_b, lambda1, lambda2, E from the block partials
//               (finalize_body)
//   zstage    : z-forward of both -> N^ in spectral space (k_xrow explains
//               why that is exact), the ETD combine with the state, and the
//               first kernel of the next stage (z-inverse of the next field)
//
// The Euler paths keep a real-space right-hand side: their x kernel (k_xa)
// stores v and f and transforms f, and y, z(lap), y kernels take lap f back to x-spectra.
#pragma once
#include <type_traits>

#include "vb_device.hpp"

namespace vb {

constexpr int XWG = 64;   // threads per x-kernel workgroup: one wave (measured best at 256^3: 64 < 128 < 256)

// Plan variant of the x kernels (row transforms).
template <typename T, int N> constexpr int xplan() { return 0; }
template <typename T, int N> constexpr int xlanes() { return XWG / TFof<N, xplan<T, N>()>(); }

// Minimum waves per SIMD requested for the fused x kernels: in FP32 four
// (the register budget that allows it costs at most a few spilled words and
// gains more in latency hiding than it loses); FP64 keeps the full budget.
template <typename T> constexpr int xwaves() { return sizeof(T) == 4 ? 4 : 1; }

// Plan variant of the FP32 column kernels (y passes, z Laplacian) at 256, 128
// and 192: 16 x 16 instead of 8 x 8 x 4 (see Plan<256, 1>), 16 x 8 instead of
// 8 x 4 x 4 (with the x kernel's, 12 x 128^3 batched 3.74 -> 3.67 ms per step),
// 24 x 8 instead of 4 x 4 x 4 x 3 (forward y pass 32.2 -> 26.7 us at 192^3).
template <typename T, int N> constexpr int colplan() {
  return (sizeof(T) == 4 && (N == 256 || N == 128 || N == 192)) ? 1 : 0;
}

// Column tile width of the y passes and the z Laplacian (columns that are
// contiguous in memory): enough columns for >= 128-byte segments and >= 256
// threads, except FP32 at 64, 16 columns (128 threads): a 64^3 solve's y pass
// is then 1056 workgroups instead of 528 (9.6 -> 7.7 us, a step 0.236 -> 0.209
// ms; 32 columns at 32 changed nothing).
template <typename T, int N> constexpr int colw() {
  if (sizeof(T) == 4 && N == 64) return 16;
  constexpr int seg = 128 / int(sizeof(cpx<T>));
  constexpr int occ = 256 / TFof<N>();
  return seg > occ ? seg : occ;
}

// Column tile of the z stage kernels (stage combines, z-prep, plain z
// passes): the column width, except FP32 at 256 and 512, 8 columns.  At 256
// (256-thread workgroups, two or three per CU) 8 measured 6-14 % faster per
// stage than 16: their inputs are group-major, so only their P, Q outputs are
// written in 64-byte pieces.  At 512, 16 columns make 1024-thread workgroups,
// whose 128-register budget spilled 160-412 bytes per lane in the stage
// combines; 8 took the combines from 1410-2358 to 1118-1440 us (the forward y
// pass, which writes the 8-column groups, from 483 to 590 us) and a step from
// 25.7 to 23.2 ms.  FP32 at 64 keeps 32 (256 threads, the shared stage
// launches' workgroup), though its y passes use 16 columns.
template <typename T, int N> constexpr int tbw() {
  return (sizeof(T) == 4 && (N == 256 || N == 512)) ? 8 : (sizeof(T) == 4 && N == 64) ? 32 : colw<T, N>();
}
// The other z tile a kernel schedule can select: the column width where tbw
// is narrower, else half of it.
template <typename T, int N> constexpr int tbw_alt() { return tbw<T, N>() < colw<T, N>() ? colw<T, N>() : tbw<T, N>() / 2; }

// Z_p = A_ext(p) + i * B_ext(p): Hermitian extension along x of two half
// spectra (rows A and B, whose Nyquist elements are at An and Bn).  The
// imaginary parts at kx = 0 and kx = N/2 are dropped, which is what a
// real-output inverse does with them.
template <typename T, int N>
__device__ __forceinline__ cpx<T> pack2(const cpx<T>* A, const cpx<T>* An, const cpx<T>* B, const cpx<T>* Bn, int p) {
  const bool mir = p > N / 2;
  const int k = mir ? N - p : p;
  cpx<T> a = *(k == N / 2 ? An : A + k), b = *(k == N / 2 ? Bn : B + k);
  if (k == 0 || k == N / 2) { a.y = T(0); b.y = T(0); }
  if (mir) { a.y = -a.y; b.y = -b.y; }
  return mk<T>(a.x - b.y, a.y + b.x);
}

// Layout of half-spectrum rows: row r = z * ny + y starts at z * pp + y * pitch;
// its Nyquist element is at z * pp + nq + y (nq >= 0) or ends the row.
// Rows number at most 512 * 512, so the division is a 32-bit one.
struct Rows {
  int ny, pitch;
  long long pp;
  long long nq;
  __host__ __device__ long long off(long long r) const {
    const unsigned q = unsigned(r) / unsigned(ny), m = unsigned(r) - q * unsigned(ny);
    return (long long)q * pp + (long long)m * pitch;
  }
  __host__ __device__ long long nyq(long long r, int half) const {
    const unsigned q = unsigned(r) / unsigned(ny), m = unsigned(r) - q * unsigned(ny);
    return (long long)q * pp + (nq >= 0 ? nq + m : (long long)m * pitch + half);
  }
};

// After a forward transform of Z = a + i b (a, b real rows), write the half
// spectra of a and b, times s (and b's times sb as well); SCALE = false writes
// them unscaled and doubled, Z_k + conj Z_{N-k} and -i (Z_k - conj Z_{N-k}).
// Uses the lane's LDS; ends with a barrier.  outAn, outBn: where the rows'
// Nyquist elements go.
template <typename T, int N, int V = 0, bool SCALE = true, class Lds>
__device__ __forceinline__ void unpack_store(const cpx<T>* x, int j, Lds& lds, cpx<T>* outA, cpx<T>* outAn,
                                             cpx<T>* outB, cpx<T>* outBn, T s, bool valid, T sb = T(1)) {
  constexpr int E = Plan<N, V>::E, TF = N / E, NH = N / 2 + 1;
#pragma unroll
  for (int e = 0; e < E; ++e) lds(j + TF * e) = x[e];
  __syncthreads();
#pragma unroll
  for (int e = 0; e < (NH + TF - 1) / TF; ++e) {
    const int k = j + TF * e;
    if (k < NH && valid) {
      const cpx<T> zk = lds(k);
      const cpx<T> zm = lds(k == 0 ? 0 : N - k);
      const T h = T(0.5) * s, hb = h * sb;
      // a_k = (Z_k + conj Z_{N-k}) / 2,  b_k = (Z_k - conj Z_{N-k}) / (2i)
      cpx<T>* const oa = k == N / 2 ? outAn : outA + k;
      cpx<T>* const ob = k == N / 2 ? outBn : outB + k;
      if constexpr (SCALE) {
        *oa = mk<T>((zk.x + zm.x) * h, (zk.y - zm.y) * h);
        *ob = mk<T>((zk.y + zm.y) * hb, (zm.x - zk.x) * hb);
      } else {
        *oa = mk<T>(zk.x + zm.x, zk.y - zm.y);
        *ob = mk<T>(zk.y + zm.y, zm.x - zk.x);
      }
    }
  }
  __syncthreads();
}

// Loads and stores of streams whose next reader is more than one kernel away
// (the z kernels' stage-private spectra and state, the last read of a stage
// input) carry the nontemporal hint: they then do not displace, in L2 and the
// memory-side cache, the spectra that the next kernel reads.  At 256^3 FP32
// the hint took the stage combines from 138-198 to 122-157 us and the y
// passes from 58 to 55 us.  Values are unchanged; it is a cache policy only.
template <typename T>
__device__ __forceinline__ cpx<T> ld_nt(const cpx<T>* p) {
  typedef T v2 __attribute__((ext_vector_type(2)));
  const v2 w = __builtin_nontemporal_load(reinterpret_cast<const v2*>(p));
  return mk<T>(w.x, w.y);
}
template <typename T>
__device__ __forceinline__ void st_nt(cpx<T>* p, cpx<T> v) {
  typedef T v2 __attribute__((ext_vector_type(2)));
  v2 w;
  w.x = v.x;
  w.y = v.y;
  __builtin_nontemporal_store(w, reinterpret_cast<v2*>(p));
}

// ---------------------------------------------------------------------------
// plain x transforms (pairs of rows)
// ---------------------------------------------------------------------------
template <typename TI, typename T, int N>
__global__ void __launch_bounds__(XWG) k_xr2c(const TI* __restrict__ in, cpx<T>* __restrict__ out,
                                               long long npairs, const cpx<T>* __restrict__ tw, double scl,
                                               Rows rl) {
  constexpr int XP = xplan<T, N>(), E = Plan<N, XP>::E, TF = N / E, G = xlanes<T, N>(), LL = row_lds_len<N>(), NH = N / 2 + 1;
  __shared__ cpx<T> sm[G * LL];
  const int lane = threadIdx.x / TF, j = threadIdx.x % TF;
  RowLds<T> lds{sm + lane * LL};
  long long pair = (long long)blockIdx.x * G + lane;
  const bool valid = pair < npairs;
  if (!valid) pair = npairs - 1;
  const long long r0 = 2 * pair;
  cpx<T> x[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int p = j + TF * e;
    x[e] = mk<T>(T(in[r0 * N + p]), T(in[(r0 + 1) * N + p]));
  }
  fft<T, N, -1, XP>(x, j, tw, lds);
  unpack_store<T, N, XP>(x, j, lds, out + rl.off(r0), out + rl.nyq(r0, N / 2), out + rl.off(r0 + 1),
                         out + rl.nyq(r0 + 1, N / 2), T(scl), valid);
}

template <typename T, typename TO, int N>
__global__ void __launch_bounds__(XWG) k_xc2r(const cpx<T>* __restrict__ in, TO* __restrict__ out,
                                               long long npairs, const cpx<T>* __restrict__ tw, Rows rl) {
  constexpr int XP = xplan<T, N>(), E = Plan<N, XP>::E, TF = N / E, G = xlanes<T, N>(), LL = row_lds_len<N>(), NH = N / 2 + 1;
  __shared__ cpx<T> sm[G * LL];
  const int lane = threadIdx.x / TF, j = threadIdx.x % TF;
  RowLds<T> lds{sm + lane * LL};
  long long pair = (long long)blockIdx.x * G + lane;
  const bool valid = pair < npairs;
  if (!valid) pair = npairs - 1;
  const long long r0 = 2 * pair;
  cpx<T> x[E];
  const cpx<T>* const i0 = in + rl.off(r0);
  const cpx<T>* const i1 = in + rl.off(r0 + 1);
  const cpx<T>* const n0 = in + rl.nyq(r0, N / 2);
  const cpx<T>* const n1 = in + rl.nyq(r0 + 1, N / 2);
#pragma unroll
  for (int e = 0; e < E; ++e) x[e] = pack2<T, N>(i0, n0, i1, n1, j + TF * e);
  fft<T, N, +1, XP>(x, j, tw, lds);
  if (valid) {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int p = j + TF * e;
      out[r0 * N + p] = TO(x[e].x);
      out[(r0 + 1) * N + p] = TO(x[e].y);
    }
  }
}

// ---------------------------------------------------------------------------
// xa: stage prefix, P1 fused (see header comment)
// ---------------------------------------------------------------------------
template <typename T>
struct XaArgs {
  const cpx<T>* P;   // x-half-spectra of v   (y, z already inverted)
  const cpx<T>* Q;   // x-half-spectra of lap v
  T* v;
  T* f;
  cpx<T>* Tt;        // out: x-forward of tmp1, times inv_n (k_xrow: doubled, unnormalised)
  cpx<T>* W;         // out (Euler only): x-forward of f, times inv_n
  cpx<T>* Ts;        // out (k_xrow): x-forward of eps^3 s, doubled, unnormalised
  double* part;      // [gridDim.x][3]
  double* dcpart;    // k_xrow, FP32: [gridDim.x][2] FP64 sums of tmp1 and eps^3 s, or nullptr
  const cpx<T>* tw;
  long long npairs;
  double eps, kappa, D, inv_n;
  Rows rl;
};

// Every pipeline kernel takes its argument struct by value, or (TAB, a
// batched launch over the solves of a size group) from a per-solve table in
// device memory indexed by blockIdx.z.
// k_xa: the x kernel of the Euler paths (WITHF: v, f and the x-forward of f
// as well); partial sums and the x-forward of tmp1, two rows per lane.
template <typename T, int N, bool WITHF, bool TAB = false>
__global__ void __launch_bounds__(XWG, (xwaves<T>())) k_xa(XaArgs<T> a0, const XaArgs<T>* tab) {
  const XaArgs<T> a = TAB ? tab[blockIdx.z] : a0;
  constexpr int XP = xplan<T, N>(), E = Plan<N, XP>::E, TF = N / E, G = xlanes<T, N>(), LL = row_lds_len<N>(), NH = N / 2 + 1;
  __shared__ cpx<T> sm[2 * G * LL];
  __shared__ double red[3 * (XWG / 64)];
  const int lane = threadIdx.x / TF, j = threadIdx.x % TF;
  RowLds<T> lds{sm + lane * LL};
  RowLds<T> lds2{sm + (G + lane) * LL};
  long long pair = (long long)blockIdx.x * G + lane;
  const bool valid = pair < a.npairs;
  if (!valid) pair = a.npairs - 1;
  const long long r0 = 2 * pair;
  // the pointwise terms multiply by 1/eps instead of dividing (within an ulp
  // of the reference's quotients; an FP64 division per point is a long
  // instruction sequence that these kernels cannot hide)
  const double eps = a.eps, kappa = a.kappa, ieps = 1.0 / eps;
  const double D = a.D, ie2 = ieps * ieps, ie3 = ie2 * ieps;
  double acc[3] = {0.0, 0.0, 0.0};
  T tm[2][E];
  T ff[WITHF ? 2 : 1][E];
  {
    cpx<T> x[2][E];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
#pragma unroll
      for (int e = 0; e < E; ++e)
        x[h][e] = pack2<T, N>(a.P + a.rl.off(r0 + h), a.P + a.rl.nyq(r0 + h, N / 2), a.Q + a.rl.off(r0 + h),
                              a.Q + a.rl.nyq(r0 + h, N / 2), j + TF * e);
    }
    fft2<T, N, +1, XP>(x[0], x[1], j, a.tw, lds, lds2);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const long long r = r0 + h;
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int p = j + TF * e;
        const double v = double(x[h][e].x), lv = double(x[h][e].y);
        const double w = v * v - 1.0;
        const double f = -lv * eps + w * v * ieps;         // f_kernel, Calc.cu:160
        const double t1 = v * (v * v - 1.0 - kappa);       // nld_kernel1, Calc.cu:270
        if (valid) {
          acc[0] += v; acc[1] += w * w; acc[2] += f * f;
          if constexpr (WITHF) {   // the Euler update reads v and f back
            a.v[r * N + p] = T(v);
            a.f[r * N + p] = T(f);
          }
        }
        tm[h][e] = T(t1);
        if constexpr (WITHF) ff[h][e] = T(f);
      }
    }
  }
  {
    cpx<T> z[E];
#pragma unroll
    for (int e = 0; e < E; ++e) z[e] = mk<T>(tm[0][e], tm[1][e]);
    fft<T, N, -1, XP>(z, j, a.tw, lds);
    unpack_store<T, N, XP>(z, j, lds, a.Tt + a.rl.off(r0), a.Tt + a.rl.nyq(r0, N / 2), a.Tt + a.rl.off(r0 + 1),
                           a.Tt + a.rl.nyq(r0 + 1, N / 2), T(a.inv_n), valid);
  }
  if constexpr (WITHF) {
    cpx<T> z[E];
#pragma unroll
    for (int e = 0; e < E; ++e) z[e] = mk<T>(ff[0][e], ff[1][e]);
    fft<T, N, -1, XP>(z, j, a.tw, lds);
    unpack_store<T, N, XP>(z, j, lds, a.W + a.rl.off(r0), a.W + a.rl.nyq(r0, N / 2), a.W + a.rl.off(r0 + 1),
                           a.W + a.rl.nyq(r0 + 1, N / 2), T(a.inv_n), valid);
  }
  block_sum_store<3, XWG>(acc, red, a.part + 3 * (long long)blockIdx.x);
}

// K sums over a one-wave workgroup in a fixed order, for K <= 8: every lane
// stores its K values to LDS (sm: K * 64 doubles), lanes 8q..8q+7 each add
// eight stored values of sum q, and three exchanges within those eight lanes
// finish it.  Sums 0-2 go to o3[0..2], sums 3.. to o2[0..] (skipped when o2
// is null).  Fewer exchanges than one butterfly per sum (block_sum_store: six
// per sum, each two LDS permutes and an FP64 add).
template <int K>
__device__ __forceinline__ void wave_sum_store(const double (&v)[K], double* sm, double* o3, double* o2) {
  static_assert(K <= 8, "eight lanes per sum");
  const int l = threadIdx.x;
#pragma unroll
  for (int q = 0; q < K; ++q) sm[q * 64 + l] = v[q];
  __syncthreads();
  const int q = l >> 3, r = l & 7;
  double s = 0.0;
  if (q < K) {
#pragma unroll
    for (int i = 0; i < 8; ++i) s += sm[q * 64 + r * 8 + i];
  }
  s += __shfl_xor(s, 4, 64);
  s += __shfl_xor(s, 2, 64);
  s += __shfl_xor(s, 1, 64);
  if (q < K && r == 0) {
    if (q < 3) o3[q] = s;
    else if (o2) o2[q - 3] = s;
  }
}

// Plan of k_xrow: at 256 the 16 x 16 plan (one LDS exchange per transform
// instead of two) ran 71 -> 65 us in FP32 (then 204 registers, two waves per
// SIMD, against 116 and four for 8 x 8 x 4) and raised the FP64 rate 3 % (250
// registers, two waves); at 128 in FP32 the 16 x 8 plan (one exchange instead
// of two); at 512 in FP32 16 x 8 x 4 (half the threads per row of 8 x 8 x 8,
// 760 -> 688 us); at 192 in FP32 24 x 8 (one exchange instead of three; 191
// registers and two waves per SIMD against 132 and three, 38.5 -> 33.8 us at
// 192^3).  The FP64 x kernel at 128 keeps 8 x 4 x 4.
template <typename T, int N> constexpr int xrplan() {
  return (N == 256 || ((N == 128 || N == 192 || N == 512) && sizeof(T) == 4)) ? 1 : xplan<T, N>();
}
// the other plan a kernel schedule can select (the same where no two-pass
// plan exists)
template <typename T, int N> constexpr int xrplan_alt() {
  return xrplan<T, N>() == 1 ? 0 : ((N == 256 || (N == 128 && sizeof(T) == 4)) ? 1 : xrplan<T, N>());
}
// threads of a k_xrow workgroup: as many rows as the other x kernels'
// workgroups (two per row pair), so the partial sums line up
template <typename T, int N, int XP = xrplan<T, N>()> constexpr int xrthreads() { return 2 * xlanes<T, N>() * TFof<N, XP>(); }
// Minimum waves per SIMD of k_xrow: three in FP32 at 256 (164 registers
// instead of 191, no spills; 56.1 -> 54.5 us at 256^3), else no minimum (the
// same cap at 192 spilled: 4 x 192^3 batched 3.56 -> 4.03 ms per step).
template <typename T, int N> constexpr int xrwaves() { return (sizeof(T) == 4 && N == 256) ? 3 : 1; }

// The x kernel of an ETD stage, one row per lane: the x-inverse of P + i Q of
// the row gives v and lap v; f, tmp1 and s follow pointwise; the x-forward of
// tmp1 + i s gives both half spectra, written over the row's P and Q.  One
// transform each way per lane: the two-row form needs about 200 registers
// with the 8 x 8 x 4 plan and ran 9 % slower at 256^3.
//
// The spectral form of the nonlinear term: nld = -lap(tmp1)/eps + s with
// s = D tmp1/eps^3 + (3v^2 - D - 1) f/eps^2 (Calc.cu:278-282), and the
// forward transform is linear, so N^ = -lambda1/2 delta_k0 - lambda2 c F(f)
// - eta F(s) - (eta/eps) |k|^2 F(tmp1), with F(f) = eps |k|^2 v^ +
// (F(tmp1) + kappa v^)/eps.  The stage combine forms N^ from the spectra of
// tmp1 and s, so lap(tmp1) never goes back to real space.  s is O(1/eps^3)
// where tmp1 is O(1), and the rounding of a shared transform scales with the
// larger part: s goes in as eps^3 s, of tmp1's size.  The two spectra are
// stored unnormalised and doubled (unpack_store without SCALE): 2 x the row
// transforms of tmp1 and eps^3 s.  The stage combine, which multiplies them by
// per-mode factors anyway, takes the normalisation 1/(2 N) and the 1/eps^3 into
// those factors (zstage_body), which saves this kernel four multiplies per mode.
// SUMS: the partial sums only (no forward transform, nothing stored): the
// integrals V, A, E_b of a field, as in FP32-mixed mode's FP64 energy of the
// state (T = double).
template <typename T, int N, bool TAB = false, int XP = xrplan<T, N>(), bool SUMS = false>
__global__ void __launch_bounds__((xrthreads<T, N, XP>()), (xrwaves<T, N>())) k_xrow(XaArgs<T> a0, const XaArgs<T>* tab) {
  const XaArgs<T> a = TAB ? tab[blockIdx.z] : a0;
  constexpr int E = Plan<N, XP>::E, TF = N / E, RL = 2 * xlanes<T, N>(), LL = row_lds_len<N>(),
                NT = xrthreads<T, N, XP>();
  static_assert(NT % 64 == 0, "whole waves");
  __shared__ cpx<T> sm[RL * LL];
  __shared__ double red[3 * (NT / 64)];
  const int lane = threadIdx.x / TF, j = threadIdx.x % TF;
  RowLds<T> lds{sm + lane * LL};
  const long long nrows = 2 * a.npairs;
  long long row = (long long)blockIdx.x * RL + lane;
  const bool valid = row < nrows;
  if (!valid) row = nrows - 1;
  const double eps = a.eps, kappa = a.kappa, ieps = 1.0 / eps;
  const double D = a.D, e3 = eps * eps * eps;
  cpx<T> x[E];
#pragma unroll
  for (int e = 0; e < E; ++e)
    x[e] = pack2<T, N>(a.P + a.rl.off(row), a.P + a.rl.nyq(row, N / 2), a.Q + a.rl.off(row),
                       a.Q + a.rl.nyq(row, N / 2), j + TF * e);
  fft<T, N, +1, XP>(x, j, a.tw, lds);   // v + i lap v
  double acc[3] = {0.0, 0.0, 0.0};
  // FP32: the k = 0 coefficients of tmp1 and eps^3 s are sums over the whole
  // grid of O(1) values that largely cancel, which an FP32 transform rounds
  // at ~1e-7 of the sum of magnitudes; summed here in FP64 from the same
  // values, they replace the transform's in the stage combine
  constexpr bool DCS = sizeof(T) == 4 && !SUMS;
  double dcs[2] = {0.0, 0.0};
  // a thread's E values are added in T first and then once in FP64: in FP32
  // the E-term share is rounded at ~1e-7 of its size, and the shares' errors
  // are independent, so the grid sums carry ~1e-10 relative (the per-point
  // FP64 conversions and adds were an eighth of this kernel's instructions)
  T ts[5] = {T(0), T(0), T(0), T(0), T(0)};
#pragma unroll
  for (int e = 0; e < E; ++e) {
    // in the work precision: v and lap v come out of a T transform, so an
    // FP64 evaluation would add nothing to their rounding (measured: the same
    // errors against the FP64 run, 48 fewer registers)
    const T v = x[e].x, lv = x[e].y;
    const T w = v * v - T(1);
    const T f = -lv * T(eps) + w * v * T(ieps);                         // f_kernel, Calc.cu:160
    const T t1 = v * (w - T(kappa));                                    // nld_kernel1, Calc.cu:270
    // eps^3 s = D tmp1 + (3 v^2 - D - 1) eps f (Calc.cu:282), as D tmp1 + (3 eps v^2 - eps (D + 1)) f
    const T sv = T(D) * t1 + (T(3 * eps) * v * v - T(eps * (D + 1))) * f;
    if constexpr (sizeof(T) == 4) {
      ts[0] += v; ts[1] += w * w; ts[2] += f * f;
      if constexpr (DCS) { ts[3] += t1; ts[4] += sv; }
    } else if (valid) {
      acc[0] += double(v); acc[1] += double(w) * double(w); acc[2] += double(f) * double(f);
    }
    x[e] = mk<T>(t1, sv);
  }
  if constexpr (sizeof(T) == 4) {
    if (valid) {
      acc[0] = double(ts[0]); acc[1] = double(ts[1]); acc[2] = double(ts[2]);
      dcs[0] = double(ts[3]); dcs[1] = double(ts[4]);
    }
  }
  if constexpr (SUMS) {
    block_sum_store<3, NT>(acc, red, a.part + 3 * (long long)blockIdx.x);
    return;
  }
  if constexpr (DCS && NT == 64) {
    // one wave: the five sums in one exchange (wave_sum_store) after the transform
    fft<T, N, -1, XP>(x, j, a.tw, lds);   // tmp1 + i s
    unpack_store<T, N, XP, false>(x, j, lds, a.Tt + a.rl.off(row), a.Tt + a.rl.nyq(row, N / 2), a.Ts + a.rl.off(row),
                                  a.Ts + a.rl.nyq(row, N / 2), T(1), valid);
    __shared__ double wred[5 * 64];
    const double v5[5] = {acc[0], acc[1], acc[2], dcs[0], dcs[1]};
    double* const o3 = a.part + 3 * (long long)blockIdx.x;
    double* const o2 = a.dcpart ? a.dcpart + 2 * (long long)blockIdx.x : nullptr;
    wave_sum_store<5>(v5, wred, o3, o2);
    return;
  }
  if constexpr (DCS) {
    if (a.dcpart) {
      __shared__ double red2[2 * (NT / 64)];
      block_sum_store<2, NT>(dcs, red2, a.dcpart + 2 * (long long)blockIdx.x);
    }
  }
  fft<T, N, -1, XP>(x, j, a.tw, lds);   // tmp1 + i s
  unpack_store<T, N, XP, false>(x, j, lds, a.Tt + a.rl.off(row), a.Tt + a.rl.nyq(row, N / 2), a.Ts + a.rl.off(row),
                                a.Ts + a.rl.nyq(row, N / 2), T(1), valid);
  block_sum_store<3, NT>(acc, red, a.part + 3 * (long long)blockIdx.x);
}

// The partial sums of v and (v^2-1)^2 from P alone, for the targets' drift in
// the first half of the penalty-drift double step (it needs V and A only, not
// E_b; the sum of f^2 is left zero).  Two rows of P travel as one transform.
// A workgroup covers kXsLoop row groups of the other x kernels in turn (group
// blockIdx.x + k gridDim.x), so the finalize reads kXsLoop times fewer
// partials; v^2 - 1 is formed in the work precision, like v, the sums in FP64.
constexpr int kXsLoop = 4;   // 1, 2, 4, 8: sums + finalize 37, 34, 32, 37 us at 256^3
template <typename T, int N, bool TAB = false>
__global__ void __launch_bounds__(XWG) k_xsums(XaArgs<T> a0, const XaArgs<T>* tab) {
  const XaArgs<T> a = TAB ? tab[blockIdx.z] : a0;
  constexpr int XP = xplan<T, N>(), E = Plan<N, XP>::E, TF = N / E, G = xlanes<T, N>(), LL = row_lds_len<N>();
  __shared__ cpx<T> sm[G * LL];
  __shared__ double red[3 * (XWG / 64)];
  const int lane = threadIdx.x / TF, j = threadIdx.x % TF;
  RowLds<T> lds{sm + lane * LL};
  const long long ngroups = (a.npairs + G - 1) / G;
  double acc[3] = {0.0, 0.0, 0.0};
  // the bound is the same for every thread of the workgroup (the transform
  // has barriers)
  for (long long grp = blockIdx.x; grp < ngroups; grp += gridDim.x) {
    long long pair = grp * G + lane;
    const bool valid = pair < a.npairs;
    if (!valid) pair = a.npairs - 1;
    const long long r0 = 2 * pair;
    cpx<T> x[E];
#pragma unroll
    for (int e = 0; e < E; ++e)
      x[e] = pack2<T, N>(a.P + a.rl.off(r0), a.P + a.rl.nyq(r0, N / 2), a.P + a.rl.off(r0 + 1),
                         a.P + a.rl.nyq(r0 + 1, N / 2), j + TF * e);
    fft<T, N, +1, XP>(x, j, a.tw, lds);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const T v = h ? x[e].y : x[e].x;
        const double w = double(v * v - T(1));
        if (valid) { acc[0] += double(v); acc[1] += w * w; }
      }
    }
  }
  block_sum_store<3, XWG>(acc, red, a.part + 3 * (long long)blockIdx.x);
}

// ---------------------------------------------------------------------------
// xe: Forward-Euler update (run_static / shape_adjust), fused C2R + R2C
// ---------------------------------------------------------------------------
template <typename T>
struct XeArgs {
  const cpx<T>* W;   // x-half-spectra of lap f
  const T* v;
  const T* f;
  cpx<T>* out;       // x-forward of the updated field, times inv_n
  const Scal* sc;
  const cpx<T>* tw;
  long long npairs;
  double eps, eta, t, inv_n;
  int shape;         // 1: shape_adjust (rhs = g0), 0: run_static (rhs = nld1)
  Rows rl;
  double* part;      // non-null: only sum rhs^2 per block here (no update)
  const double* tp;  // non-null: t is read from here (one captured update serves every step)
};

template <typename T, int N>
__global__ void __launch_bounds__(XWG) k_xe(XeArgs<T> a) {
  constexpr int XP = xplan<T, N>(), E = Plan<N, XP>::E, TF = N / E, G = xlanes<T, N>(), LL = row_lds_len<N>(), NH = N / 2 + 1;
  __shared__ cpx<T> sm[G * LL];
  const int lane = threadIdx.x / TF, j = threadIdx.x % TF;
  RowLds<T> lds{sm + lane * LL};
  long long pair = (long long)blockIdx.x * G + lane;
  const bool valid = pair < a.npairs;
  if (!valid) pair = a.npairs - 1;
  const long long r0 = 2 * pair;
  cpx<T> x[E];
#pragma unroll
  for (int e = 0; e < E; ++e)
    x[e] = pack2<T, N>(a.W + a.rl.off(r0), a.W + a.rl.nyq(r0, N / 2), a.W + a.rl.off(r0 + 1),
                       a.W + a.rl.nyq(r0 + 1, N / 2), j + TF * e);
  fft<T, N, +1, XP>(x, j, a.tw, lds);
  const double l1 = a.sc->l1, l2 = a.sc->l2, eta = a.eta, eps = a.eps, t = a.tp ? *a.tp : a.t, ie2 = 1.0 / (eps * eps);
  double acc[1] = {0.0};
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int p = j + TF * e;
    double nv[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const long long r = r0 + h;
      const double lf = double(h ? x[e].y : x[e].x);
      const double v = double(a.v[r * N + p]), f = double(a.f[r * N + p]);
      const double g0 = -lf + (3.0 * v * v - 1.0) * f * ie2;            // Calc.cu:178
      const double rhs = a.shape ? g0 : l1 * 0.5 + l2 * f + g0 * eta;    // Calc.cu:417
      nv[h] = v + t * rhs;                                               // Calc.cu:400
      acc[0] += valid ? rhs * rhs : 0.0;
    }
    x[e] = mk<T>(T(nv[0]), T(nv[1]));
  }
  if (a.part) {   // uniform over the grid: the residual sum only
    __shared__ double red[XWG / 64];
    block_sum_store<1, XWG>(acc, red, a.part + blockIdx.x);
    return;
  }
  fft<T, N, -1, XP>(x, j, a.tw, lds);
  unpack_store<T, N, XP>(x, j, lds, a.out + a.rl.off(r0), a.out + a.rl.nyq(r0, N / 2), a.out + a.rl.off(r0 + 1),
                         a.out + a.rl.nyq(r0 + 1, N / 2), T(a.inv_n), valid);
}

// ---------------------------------------------------------------------------
// y kernels: complex transforms along y, in place, batch over (z, kx).
// blockIdx.y selects one of two arrays.
// ---------------------------------------------------------------------------
// Column of one y transform: (z, kx) -> base address and the stride between
// its ky (pitch; 1 in a plane's Nyquist column, see Geo).
struct YIdx {
  long long base;
  long long stride;
  int z, kx;
  bool valid;
};

// t: the workgroup's column group (blockIdx.x unless the launch has extra
// workgroups in front)
template <int TB>
__device__ __forceinline__ YIdx yidx(const Geo& g, long long t) {
  const int c = threadIdx.x % TB;
  YIdx r;
  int z, kx;
  if (g.aligned) {
    // the Nyquist groups (uncoalesced, slower) come first so they overlap
    // with the main groups instead of trailing them
    const int tpr = (g.nx / 2) / TB;
    const long long nnyq = (g.nz + TB - 1) / TB;
    if (t >= nnyq) {
      z = int((t - nnyq) / tpr);
      kx = int((t - nnyq) % tpr) * TB + c;
      r.valid = true;
    } else {   // Nyquist column of TB planes
      z = int(t * TB + c);
      kx = g.nx / 2;
      r.valid = z < g.nz;
      if (!r.valid) z = g.nz - 1;
    }
  } else {
    const long long batch = (long long)g.nz * g.nh;
    long long b = t * TB + c;
    r.valid = b < batch;
    if (!r.valid) b = batch - 1;
    z = int(b / g.nh);
    kx = int(b % g.nh);
  }
  const bool seg = g.nq >= 0 && kx == g.nx / 2;
  r.base = (long long)z * g.pp + (seg ? g.nq : kx);
  r.stride = seg ? 1 : g.pitch;
  r.z = z;
  r.kx = kx;
  return r;
}
template <int TB>
__device__ __forceinline__ YIdx yidx(const Geo& g) {
  return yidx<TB>(g, blockIdx.x);
}

template <typename T>
struct YArgs {
  cpx<T>* b0;
  cpx<T>* b1;   // blockIdx.y = 1 transforms this one
  const cpx<T>* tw;
  Geo g;
  FinArgs fin;  // workgroups [0, fin_on) of blockIdx.y = 0 evaluate these integrals
  int fin_on;
  cpx<T>* o0;   // GM: out-of-place outputs in the z kernels' group-major layout
  cpx<T>* o1;
  int ztb;      // GM: the z kernels' column tile
};

// Position of spectral element (z, ky, kx) in the group-major layout of the z
// stage kernels: [z workgroup][kz][ztb], the workgroup and lane that zidx
// gives to column (ky, kx).
__device__ __forceinline__ long long gm_index(const Geo& g, int ztb, int z, int ky, int kx) {
  // ztb is a power of two (every column tile is; the solver checks it)
  const int lz = __builtin_ctz((unsigned)ztb), mz = ztb - 1;
  long long grp;
  int c;
  if (g.aligned) {
    if (kx == g.nx / 2) {
      grp = ky >> lz;
      c = ky & mz;
    } else {
      grp = (long long)((g.ny + mz) >> lz) + (long long)ky * ((g.nx / 2) >> lz) + (kx >> lz);
      c = kx & mz;
    }
  } else {
    const long long q = (long long)ky * g.nh + kx;
    grp = q >> lz;
    c = int(q & mz);
  }
  return (grp * g.nz + z) * ztb + c;
}

// GM: the transform is written to o0 / o1 in the z kernels' group-major
// layout instead of in place (the y-forward before a stage combine, so that
// the combine reads its two inputs as contiguous runs).
template <typename T, int N, int S, int TB = tbw<T, N>(), bool TAB = false, int V = 0, bool GM = false>
__global__ void __launch_bounds__((TB * TFof<N, V>())) k_ypass(YArgs<T> a0, const YArgs<T>* tab) {
  constexpr int E = Plan<N, V>::E, TF = N / E;
  __shared__ cpx<T> sm[N * TB];
  const YArgs<T> a = TAB ? tab[blockIdx.z] : a0;
  if ((int)blockIdx.x < a.fin_on) {
    // the integrals read only partials of kernels that ran before this one;
    // their workgroups are dispatched first and overlap the column work
    if (blockIdx.y == 0) {
      double* red = reinterpret_cast<double*>(sm);   // the column LDS, unused here
      finalize_body<TB * TF, 2>(a.fin, red, red + 6 * ((TB * TF) / 64), blockIdx.x, a.fin_on);
    }
    return;
  }
  const Geo& g = a.g;
  const cpx<T>* __restrict__ tw = a.tw;
  cpx<T>* buf = blockIdx.y ? a.b1 : a.b0;
  const int j = threadIdx.x / TB;
  ColLds<T, TB> lds{sm + threadIdx.x % TB};
  const YIdx id = yidx<TB>(g, (long long)blockIdx.x - a.fin_on);
  cpx<T>* col = buf + id.base;
  cpx<T> x[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    // the forward pass into the group-major stage inputs is the last reader of P, Q
    const cpx<T>* src = col + (long long)(j + TF * e) * id.stride;
    x[e] = GM ? ld_nt(src) : *src;
  }
  fft<T, N, S, V>(x, j, tw, lds);
  if (id.valid) {
    if constexpr (GM) {
      cpx<T>* out = blockIdx.y ? a.o1 : a.o0;
      if (g.aligned && id.kx != g.nx / 2) {
        // ky = j + TF e: one ky step is one row of column groups of all z
        cpx<T>* o = out + gm_index(g, a.ztb, id.z, j, id.kx);
        const long long s = (long long)(((g.nx / 2) >> __builtin_ctz((unsigned)a.ztb)) * TF) * g.nz * a.ztb;
#pragma unroll
        for (int e = 0; e < E; ++e) o[e * s] = x[e];
      } else {
#pragma unroll
        for (int e = 0; e < E; ++e) out[gm_index(g, a.ztb, id.z, j + TF * e, id.kx)] = x[e];
      }
    } else {
#pragma unroll
      for (int e = 0; e < E; ++e) col[(long long)(j + TF * e) * id.stride] = x[e];
    }
  }
}

// ---------------------------------------------------------------------------
// z kernels: complex transforms along z, batch over the (ky, kx) plane.
// ---------------------------------------------------------------------------
struct ZIdx {
  long long q;     // column in the (ky, kx) plane
  int ky, kx;
  bool valid;
};

// t: the workgroup's column group, blockIdx.x unless the launch has extra
// workgroups in front
template <int TB>
__device__ __forceinline__ ZIdx zidx(const Geo& g, long long t, int c) {
  ZIdx r;
  if (g.aligned) {
    const int tpr = (g.nx / 2) / TB;
    const long long nnyq = (g.ny + TB - 1) / TB;
    if (t >= nnyq) {
      r.ky = int((t - nnyq) / tpr);
      r.kx = int((t - nnyq) % tpr) * TB + c;
      r.valid = true;
    } else {   // Nyquist column of TB rows (scheduled first, see yidx)
      int ky = int(t * TB + c);
      r.valid = ky < g.ny;
      if (!r.valid) ky = g.ny - 1;
      r.ky = ky;
      r.kx = g.nx / 2;
    }
  } else {
    const long long ncol = (long long)g.ny * g.nh;
    long long q = t * TB + c;
    r.valid = q < ncol;
    if (!r.valid) q = ncol - 1;
    r.ky = int(q / g.nh);
    r.kx = int(q % g.nh);
  }
  r.q = (g.nq >= 0 && r.kx == g.nx / 2) ? g.nq + r.ky : (long long)r.ky * g.pitch + r.kx;
  return r;
}
template <int TB>
__device__ __forceinline__ ZIdx zidx(const Geo& g, long long t) {
  return zidx<TB>(g, t, int(threadIdx.x % TB));
}
template <int TB>
__device__ __forceinline__ ZIdx zidx(const Geo& g) {
  return zidx<TB>(g, blockIdx.x);
}

// number of y / z workgroups for a tile width
__host__ __device__ inline long long ygroups(const Geo& g, int TB) {
  if (g.aligned) return (long long)g.nz * ((g.nx / 2) / TB) + (g.nz + TB - 1) / TB;
  return ((long long)g.nz * g.nh + TB - 1) / TB;
}
__host__ __device__ inline long long zgroups(const Geo& g, int TB) {
  if (g.aligned) return (long long)g.ny * ((g.nx / 2) / TB) + (g.ny + TB - 1) / TB;
  return ((long long)g.ny * g.nh + TB - 1) / TB;
}

// Elements of a stage-private spectrum (pre, Na, Nb of k_zstage): group-major,
// [z workgroup][kz][TB], which may exceed the padded spectrum for unaligned rows.
__host__ __device__ inline long long zs_private(const Geo& g, int TB) {
  const long long n = zgroups(g, TB) * g.nz * TB;
  return n > g.ncp ? n : g.ncp;
}

// |k|^2 of the gradient with each component's own Nyquist plane zeroed
// (gradient_kernel, Calc.cu:310-339), times the half-spectrum weight.
__device__ __forceinline__ double grad_w(const Geo& g, int kx, int ky, int kz) {
  const double gx = (kx == g.nx / 2) ? 0.0 : double(kx);
  const double gy = (ky == g.ny / 2) ? 0.0 : double(sgnk(ky, g.ny));
  const double gz = (kz == g.nz / 2) ? 0.0 : double(sgnk(kz, g.nz));
  const double w = (kx == 0 || kx == g.nx / 2) ? 1.0 : 2.0;
  return w * (gx * gx + gy * gy + gz * gz) * g.kfac2;
}

__device__ __forceinline__ int mindex(const Geo& g, int kx, int ky, int kz) {
  const int fy = fold(ky, g.ny), fz = fold(kz, g.nz);
  return kx * kx + fy * fy + fz * fz;
}

template <typename T, int N, int S, int TB = tbw<T, N>()>
__global__ void __launch_bounds__((TB * TFof<N>())) k_zpass(cpx<T>* buf, Geo g, const cpx<T>* __restrict__ tw) {
  constexpr int E = Plan<N>::E, TF = N / E;
  __shared__ cpx<T> sm[N * TB];
  const int j = threadIdx.x / TB;
  ColLds<T, TB> lds{sm + threadIdx.x % TB};
  const ZIdx id = zidx<TB>(g);
  const long long plane = g.pp;
  cpx<T>* col = buf + id.q;
  cpx<T> x[E];
#pragma unroll
  for (int e = 0; e < E; ++e) x[e] = col[(long long)(j + TF * e) * plane];
  fft<T, N, S>(x, j, tw, lds);
  if (id.valid) {
#pragma unroll
    for (int e = 0; e < E; ++e) col[(long long)(j + TF * e) * plane] = x[e];
  }
}

template <typename T>
struct ZlArgs {
  cpx<T>* buf;
  const cpx<T>* tw;
  Geo g;
  FinArgs fin;   // workgroups [0, fin_on) evaluate these integrals
  int fin_on;
};

// forward along z, multiply by -|k|^2 (laplace_kernel, Calc.cu:114-122), inverse
template <typename T, int N, int TB = tbw<T, N>(), bool TAB = false, int V = 0>
__global__ void __launch_bounds__((TB * TFof<N, V>()), 4) k_zlap(ZlArgs<T> a0, const ZlArgs<T>* tab) {
  constexpr int E = Plan<N, V>::E, TF = N / E;
  __shared__ cpx<T> sm[N * TB];
  const ZlArgs<T> a = TAB ? tab[blockIdx.z] : a0;
  if ((int)blockIdx.x < a.fin_on) {
    // the integrals only read partials of kernels that ran before this one;
    // their workgroups are dispatched first and overlap the column work
    __shared__ double red[6 * ((TB * TF) / 64)];
    __shared__ double tot[7];
    finalize_body<TB * TF>(a.fin, red, tot, blockIdx.x, a.fin_on);
    return;
  }
  const Geo& g = a.g;
  const cpx<T>* __restrict__ tw = a.tw;
  cpx<T>* buf = a.buf;
  const int j = threadIdx.x / TB;
  ColLds<T, TB> lds{sm + threadIdx.x % TB};
  const ZIdx id = zidx<TB>(g, (long long)blockIdx.x - a.fin_on);
  const long long plane = g.pp;
  cpx<T>* col = buf + id.q;
  cpx<T> x[E];
#pragma unroll
  for (int e = 0; e < E; ++e) x[e] = col[(long long)(j + TF * e) * plane];
  fft<T, N, -1, V>(x, j, tw, lds);
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int kz = j + TF * e;
    const T mult = T(-double(mindex(g, id.kx, id.ky, kz)) * g.kfac2);
    x[e] = scale(x[e], mult);
  }
  fft<T, N, +1, V>(x, j, tw, lds);
  if (id.valid) {
#pragma unroll
    for (int e = 0; e < E; ++e) col[(long long)(j + TF * e) * plane] = x[e];
  }
}

// Shared tail of zprep / zstage, applied to the next stage field nf in FP64
// (the state and the combine stay FP64 in both precisions), one element at a
// time as the combine produces it: its Parseval partial of |grad|^2 and its
// mean (the k = 0 coefficient, written to dc[0]) are taken in FP64; then
// P = z-inverse(nf), Q = z-inverse(-|k|^2 nf) in T.
template <typename T>
__device__ __forceinline__ void next_field_elem(double re, double im, int kz, const ZIdx& id, const Geo& g,
                                                double& acc, cpx<T>& xo, cpx<T>& yo) {
  if (id.valid) acc += grad_w(g, id.kx, id.ky, kz) * (re * re + im * im);
  const double m = -double(mindex(g, id.kx, id.ky, kz)) * g.kfac2;
  xo = mk<T>(T(re), T(im));
  yo = mk<T>(T(re * m), T(im * m));
}

// PONLY: the z-inverse of the field alone into P (Q, the Laplacian, is not
// needed); the Parseval partials are stored either way.
template <typename T, int N, int TB, int V, bool PONLY = false, class Lds>
__device__ __forceinline__ void next_field_tail(cpx<T>* x, cpx<T>* y, double acc, int j, const ZIdx& id,
                                                const Geo& g, Lds& lds, Lds& lds2, const cpx<T>* __restrict__ tw,
                                                cpx<T>* P, cpx<T>* Q, double* part, double* red, int tile) {
  constexpr int E = Plan<N, V>::E, TF = N / E;
  const long long plane = g.pp;
  if constexpr (PONLY)
    fft<T, N, +1, V>(x, j, tw, lds);
  else
    fft2<T, N, +1, V>(x, y, j, tw, lds, lds2);
  if (id.valid) {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      P[id.q + (long long)(j + TF * e) * plane] = x[e];
      if constexpr (!PONLY) Q[id.q + (long long)(j + TF * e) * plane] = y[e];
    }
  }
  double s[1] = {acc};
  block_sum_store<1, TB * TF>(s, red, part + tile);
}

template <typename T>
struct ZprArgs {
  const cpx<double>* vh;   // state spectrum, group-major
  cpx<T>* P;
  cpx<T>* Q;
  double* part;
  double* dc;
  const cpx<T>* tw;
  Geo g;
};

template <typename T, int N, int TB = tbw<T, N>(), bool TAB = false>
__global__ void __launch_bounds__((TB * TFof<N>())) k_zprep(ZprArgs<T> a0, const ZprArgs<T>* tab) {
  constexpr int E = Plan<N>::E, TF = N / E;
  const ZprArgs<T> a = TAB ? tab[blockIdx.z] : a0;
  const Geo& g = a.g;
  const cpx<T>* __restrict__ tw = a.tw;
  const cpx<double>* __restrict__ vh = a.vh;
  __shared__ cpx<T> sm[2 * N * TB];
  __shared__ double red[(TB * TF) / 64];
  const int j = threadIdx.x / TB;
  ColLds<T, TB> lds{sm + threadIdx.x % TB};
  ColLds<T, TB> lds2{sm + N * TB + threadIdx.x % TB};
  const ZIdx id = zidx<TB>(g);
  // the state spectrum is group-major (see k_zstage)
  const long long ib0 = (long long)blockIdx.x * N * TB + threadIdx.x % TB;
  cpx<double> u[E];
#pragma unroll
  for (int e = 0; e < E; ++e) u[e] = id.valid ? vh[ib0 + (long long)(j + TF * e) * TB] : mk<double>(0.0, 0.0);
  if (id.valid && id.q == 0 && j == 0) a.dc[0] = u[0].x;
  cpx<T> x[E], y[E];
  double acc = 0.0;
#pragma unroll
  for (int e = 0; e < E; ++e) next_field_elem<T>(u[e].x, u[e].y, j + TF * e, id, g, acc, x[e], y[e]);
  next_field_tail<T, N, TB, 0>(x, y, acc, j, id, g, lds, lds2, tw, a.P, a.Q, a.part, red, int(blockIdx.x));
}

template <typename T>
struct ZsArgs {
  const cpx<T>* U;        // x,y-forward of tmp1 (normalised), group-major
  const cpx<T>* U2;       // x,y-forward of s (normalised), group-major
  cpx<T>* Vh;             // spectrum of the stage field (stages after the first)
                          // and of the next one, group-major, in place
  const Scal* sc;         // lambda1, lambda2 of the stage field
  double eps, kappa, eta;
  const cpx<double>* uh;  // state spectrum u^ (group-major)
  cpx<T>* pre;            // ETDRK4: e1 a^ - q N^u (stage 1 -> 3)
  cpx<T>* Na;             // ETDRK4: N^a (stage 2 -> 4); ETDRK2: N^u (stage 1 -> 2)
  cpx<T>* Nb;             // ETDRK4: N^b (stage 3 -> 4)
  cpx<double>* uout;      // new state spectrum (last stage, group-major)
  cpx<T>* P;
  cpx<T>* Q;
  double* part;
  double* dc;
  const double2* c1;      // per m: ETDRK4 {e1, (e1-1)/H}; ETDRK2 {e^z, (e^z-1)/H}
  const double4* c2;      // per m: ETDRK4 {e2, A s, 2 B s, C s}; ETDRK2 {c2, 0, 0, 0}
  const cpx<T>* tw;
  Geo g;
};

// SCH = 4 (ETDRK4, belastic.cpp:519-552 / Calc.cu:488-577) or 2 (ETDRK2,
// belastic.cpp:491-517 / Calc.cu:446-486); STG = stage whose N^ this is.
// The stage's N^ is formed here from the spectra of tmp1 and s (see k_xrow),
// the stage field's spectrum v^ (the state u^ in the first stage, Vh after
// it) and lambda1, lambda2; the stage also stores the next stage field's
// spectrum in Vh.  The N^ of a stage is stored in precision T for a later
// stage; the state and every combine that produces it are FP64.
// All stages load their operands after the forward transforms, and ask for
// no minimum occupancy: a cap of three or four waves per SIMD spills heavily
// and runs slower (the probe form under a three-wave cap ran faster with
// per-element gathers, slower with the staged table below).  Every stream but P, Q (read
// by the next kernel) and the coefficient tables is loaded and stored
// nontemporally (see ld_nt).
// PROBE (ETDRK4 stage 4 only): the result is needed only for its V and A,
// the first evaluation of a penalty-drift step (belastic.cpp:677-683): the
// new state is not stored, and only P (its z-inverse) and the Parseval and
// DC terms are.
// Plan variant of the stage combines (see Plan): the default plan, except
// FP32 at 192, 12 x 4 x 4 (two exchanges instead of three, the same 12
// elements per thread; 218-240 registers against 205-226, two waves per SIMD
// either way; 4 x 192^3 batched 3.61 -> 3.57 ms per step).  At 256 the
// 16 x 16 plan needs 248-256 registers there and ran slower (stage 4 158 ->
// 182 us), and so did 4 x 4 x 4 x 4 over 64 threads (100-128 registers).
template <typename T, int N> constexpr int zsplan() { return (sizeof(T) == 4 && N == 192) ? 3 : 0; }

// LDS bytes of a stage combine: two column tiles for the transforms, or the
// staged coefficient table, whichever is larger
template <typename T, int N, int SCH, int STG, int TB, int V>
constexpr int zs_lds_bytes() {
  return (2 * N * TB * int(sizeof(cpx<T>))) >
                 ((SCH == 4 && STG != 2) ? TB * (N / 2 + 1) * int(STG == 4 ? sizeof(double4) : sizeof(double2)) : 0)
             ? 2 * N * TB * int(sizeof(cpx<T>))
             : TB * (N / 2 + 1) * int(STG == 4 ? sizeof(double4) : sizeof(double2));
}

// The stage combine of column group `tile` of one solve, with the
// workgroup's LDS passed in (smraw: zs_lds_bytes, red: (TB * TF) / 64
// doubles), so that a launch can run the groups of solves of several sizes.
template <typename T, int N, int SCH, int STG, int TB, int V, bool PROBE>
__device__ __forceinline__ void zstage_body(const ZsArgs<T>& a, int tile, unsigned char* smraw, double* red) {
  static_assert(!PROBE || (SCH == 4 && STG == 4), "the probe form is the last ETDRK4 stage");
  constexpr int E = Plan<N, V>::E, TF = N / E;
  // ETDRK4 stages 1, 3 and 4 read their coefficient records from a
  // per-workgroup table in LDS: TB columns x (N/2 + 1) folded kz, gathered once
  // after the forward transforms.  Gathered per element, the 64 lanes of a load
  // touched up to 64 cache lines and the combines issued 4x the L1-to-L2
  // requests of their streams; staged, stages 1, 3, 4 and the probe ran 120,
  // 123, 161, 126 -> 112, 107, 148, 122 us at 256^3 FP32.  Stage 2 keeps the
  // gathers: the table cost it a wave per SIMD and time.
  constexpr bool STAGED = SCH == 4 && STG != 2;
  using Rec = typename std::conditional<STG == 4, double4, double2>::type;
  constexpr int NQ = N / 2 + 1;
  cpx<T>* sm = reinterpret_cast<cpx<T>*>(smraw);
  const int j = threadIdx.x / TB;
  ColLds<T, TB> lds{sm + threadIdx.x % TB};
  ColLds<T, TB> lds2{sm + N * TB + threadIdx.x % TB};
  const Geo g = a.g;
  const ZIdx id = zidx<TB>(g, tile);
  // The state and U, U2, pre, Na, Nb, Vh are kept group-major: this
  // workgroup's TB columns of all N planes are one contiguous run (see
  // zs_private); the strided z traffic of a stage is only P, Q
  const long long ib0 = (long long)tile * N * TB + threadIdx.x % TB;
  // (lanes past the last column of a Nyquist group have no input written;
  // they transform zeros and store nothing)
  cpx<T> x[E], x2[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    x[e] = id.valid ? ld_nt(a.U + ib0 + (long long)(j + TF * e) * TB) : mk<T>(T(0), T(0));
    x2[e] = id.valid ? ld_nt(a.U2 + ib0 + (long long)(j + TF * e) * TB) : mk<T>(T(0), T(0));
  }
  // Calc.cu:428-444 with nld (Calc.cu:265-297) moved to spectral space:
  // N^ = -lambda1/2 delta_k0 - lambda2 3/(2 sqrt 2) F(f) - eta F(s)
  //      - (eta/eps) |k|^2 F(tmp1),  F(f) = eps |k|^2 v^ + (F(tmp1) + kappa v^)/eps.
  // Everything but the v^ terms is formed right after the transforms, so the
  // spectra of tmp1 and s are dead before the combine's operands load:
  // N^ = Np + Av v^ with Av = -lambda2 3/(2 sqrt 2) (eps |k|^2 + kappa/eps).
  // Np stays FP64: near an equilibrium f^ is small and its two parts cancel.
  fft2<T, N, -1, V>(x, x2, j, a.tw, lds, lds2);   // spectra of tmp1 and eps^3 s (stored units)
  const double l1h = 0.5 * a.sc->l1;
  const double l2c = a.sc->l2 * (3.0 / (2.0 * sqrt(2.0)));
  const double ie = 1.0 / a.eps;
  // the x kernel stores 2 N F(tmp1) and 2 N F(eps^3 s) (F normalised): hu and
  // hs turn them into F(tmp1) and F(s)
  const double hu = 0.5 / double(g.nr), hs = hu * (ie * ie * ie);
  const double ieu = ie * hu, etas = a.eta * hs;
  cpx<double> np[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int kz = j + TF * e;
    const double k2 = double(mindex(g, id.kx, id.ky, kz)) * g.kfac2;
    const double at = -(l2c + a.eta * k2) * ieu;   // on the stored spectrum of tmp1
    double tr = double(x[e].x), sr = double(x2[e].x);
    if constexpr (sizeof(T) == 4) {
      // the k = 0 coefficients summed in FP64 by the x kernel (k_xrow), the
      // means of tmp1 and eps^3 s, in stored units
      if (e == 0 && id.q == 0 && j == 0) {
        tr = a.sc->dc_t1 / hu;
        sr = a.sc->dc_s / hu;
      }
    }
    double pr = at * tr - etas * sr;
    const double pi = at * double(x[e].y) - etas * double(x2[e].y);
    if (id.q == 0 && kz == 0) pr -= l1h;
    np[e] = mk<double>(pr, pi);
  }
  const Rec* tbl = reinterpret_cast<const Rec*>(smraw);
  if constexpr (STAGED) {
    // the transforms' last LDS reads are behind their last barrier
    Rec* w = reinterpret_cast<Rec*>(smraw);
    for (int i = threadIdx.x; i < TB * NQ; i += TB * TF) {
      const int c = i / NQ, q = i % NQ;
      const ZIdx ic = zidx<TB>(g, tile, c);
      const int m = mindex(g, ic.kx, ic.ky, q);
      if constexpr (STG == 4) w[i] = a.c2[m]; else w[i] = a.c1[m];
    }
    __syncthreads();
  }
  const int tcol = int(threadIdx.x % TB);
  constexpr bool USES_U = !(SCH == 4 && STG == 3);
  // The last stage loads and combines its operands in two halves in FP64, and
  // in FP32 with 12 elements per thread (192, 384): all at once they need 256
  // registers and more in AGPRs, one wave per SIMD; in halves two (4 x 192^3
  // batched 4.15 -> 3.98 ms per step).  The probe form does too in FP32 at
  // 256 (198 -> 166 registers, two waves per SIMD -> three, 121 -> 118 us);
  // the full FP32 stage 4 stays at two waves either way and ran slower split
  // (149 -> 154 us).
  constexpr int NH = ((sizeof(T) == 8 || E >= 12 || PROBE) && STG == SCH) ? 2 : 1, EH = E / NH;
  cpx<T> y[E];
  double acc = 0.0, dc0 = 0.0;
#pragma unroll
  for (int hh = 0; hh < NH; ++hh) {
    cpx<double> ur[USES_U ? EH : 1];
    cpx<T> pa[(SCH == 4 && STG >= 3) || (SCH == 2 && STG == 2) ? EH : 1];
    cpx<T> pb[(SCH == 4 && STG == 4) ? EH : 1];
    cpx<T> pv[STG > 1 ? EH : 1];
#pragma unroll
    for (int iq = 0; iq < EH; ++iq) {
      const int e = hh * EH + iq;
      const long long ib = ib0 + (long long)(j + TF * e) * TB;
      if constexpr (USES_U) ur[iq] = id.valid ? ld_nt(a.uh + ib) : mk<double>(0.0, 0.0);
      if constexpr (SCH == 4 && STG == 3) pa[iq] = ld_nt(a.pre + ib);
      if constexpr (SCH == 4 && STG == 4) { pa[iq] = ld_nt(a.Na + ib); pb[iq] = ld_nt(a.Nb + ib); }
      if constexpr (SCH == 2 && STG == 2) pa[iq] = ld_nt(a.Na + ib);
      if constexpr (STG > 1) pv[iq] = ld_nt(a.Vh + ib);
    }
#pragma unroll
    for (int iq = 0; iq < EH; ++iq) {
      const int e = hh * EH + iq;
      const int kz = j + TF * e;
      const long long ib = ib0 + (long long)kz * TB;
      const int m = mindex(g, id.kx, id.ky, kz);
      const double av = -l2c * (a.eps * double(m) * g.kfac2 + a.kappa * ie);
      double vr, vi;   // the stage field's spectrum
      if constexpr (STG == 1) {
        vr = ur[iq].x;
        vi = ur[iq].y;
      } else {
        vr = double(pv[iq].x);
        vi = double(pv[iq].y);
      }
      const double nr = np[e].x + av * vr, ni = np[e].y + av * vi;   // N^ of this stage
      x[e] = mk<T>(T(nr), T(ni));
      double fr, fi;   // next field
      if constexpr (SCH == 4 && STG == 1) {
        const double2 c = STAGED ? tbl[tcol * NQ + (kz <= N / 2 ? kz : N - kz)] : a.c1[m];   // e1, q
        const cpx<double> u = ur[iq];
        fr = c.x * u.x + c.y * nr;                      // etd41: a^
        fi = c.x * u.y + c.y * ni;
        if (id.valid) st_nt(a.pre + ib, mk<T>(T(c.x * fr - c.y * nr), T(c.x * fi - c.y * ni)));
      } else if constexpr (SCH == 4 && STG == 2) {
        const double2 c = STAGED ? tbl[tcol * NQ + (kz <= N / 2 ? kz : N - kz)] : a.c1[m];
        const cpx<double> u = ur[iq];
        fr = c.x * u.x + c.y * nr;                      // etd42: b^
        fi = c.x * u.y + c.y * ni;
        if (id.valid) st_nt(a.Na + ib, x[e]);
      } else if constexpr (SCH == 4 && STG == 3) {
        const double q = STAGED ? tbl[tcol * NQ + (kz <= N / 2 ? kz : N - kz)].y : a.c1[m].y;
        const cpx<T> p = pa[iq];
        fr = double(p.x) + 2.0 * q * nr;                // etd43: c^ = e1 F(a) + q (2 N^b - N^u)
        fi = double(p.y) + 2.0 * q * ni;
        if (id.valid) st_nt(a.Nb + ib, x[e]);
      } else if constexpr (SCH == 4 && STG == 4) {
        const double4 c = STAGED ? tbl[tcol * NQ + (kz <= N / 2 ? kz : N - kz)] : a.c2[m];   // e2, A s, 2 B s, C s
        const cpx<double> u = ur[iq];
        const cpx<T> na = pa[iq], nb = pb[iq];
        // etd44: e2 u^ + [N^a A + 2 (N^a + N^b) B + N^c C] s  (first term N^a as in Calc.cu:559)
        fr = c.x * u.x + (double(na.x) * c.y + (double(na.x) + double(nb.x)) * c.z + nr * c.w);
        fi = c.x * u.y + (double(na.y) * c.y + (double(na.y) + double(nb.y)) * c.z + ni * c.w);
        if (id.valid && !PROBE) st_nt(a.uout + ib, mk<double>(fr, fi));
      } else if constexpr (SCH == 2 && STG == 1) {
        const double2 c = a.c1[m];                      // e^z, q1
        const cpx<double> u = ur[iq];
        fr = c.x * u.x + c.y * nr;                      // etd1: a^
        fi = c.x * u.y + c.y * ni;
        if (id.valid) st_nt(a.Na + ib, x[e]);
      } else {                                          // SCH == 2 && STG == 2
        const double2 c = a.c1[m];
        const double c2 = a.c2[m].x;
        const cpx<double> u = ur[iq];
        const cpx<T> nu = pa[iq];
        // etd2: u^+ = F(a) + c2 (N^a - N^u), a^ = e^z u^ + q1 N^u
        const double ar = c.x * u.x + c.y * double(nu.x), ai = c.x * u.y + c.y * double(nu.y);
        fr = ar + c2 * (nr - double(nu.x));
        fi = ai + c2 * (ni - double(nu.y));
        if (id.valid) st_nt(a.uout + ib, mk<double>(fr, fi));
      }
      if constexpr (STG < SCH) {
        if (id.valid) st_nt(a.Vh + ib, mk<T>(T(fr), T(fi)));
      }
      if (e == 0) dc0 = fr;
      next_field_elem<T>(fr, fi, kz, id, g, acc, x[e], y[e]);
    }
  }
  if (id.valid && id.q == 0 && j == 0) a.dc[0] = dc0;
  if constexpr (STAGED) __syncthreads();   // the table's last reads before the transforms reuse LDS
  next_field_tail<T, N, TB, V, PROBE>(x, y, acc, j, id, g, lds, lds2, a.tw, a.P, a.Q, a.part, red, tile);
}

template <typename T, int N, int SCH, int STG, int TB = tbw<T, N>(), bool TAB = false, int V = zsplan<T, N>(),
          bool PROBE = false>
__global__ void __launch_bounds__((TB * TFof<N, V>())) k_zstage(ZsArgs<T> a0, const ZsArgs<T>* tab) {
  const ZsArgs<T> a = TAB ? tab[blockIdx.z] : a0;
  __shared__ __align__(16) unsigned char smraw[zs_lds_bytes<T, N, SCH, STG, TB, V>()];
  __shared__ double red[(TB * TFof<N, V>()) / 64];
  zstage_body<T, N, SCH, STG, TB, V, PROBE>(a, int(blockIdx.x), smraw, red);
}

// Copy of a half spectrum between layouts, widened to FP64: tb_in / tb_out =
// 0 for the row layout [z][ky][pitch], or the column tile of the z kernels'
// group-major layout (gm_index).  One thread per mode.
template <typename TI>
struct RelArgs {
  const cpx<TI>* in;
  cpx<double>* out;
  int tb_in, tb_out;
  Geo g;
};

template <typename TI, bool TAB = false>
__global__ void k_relayout(RelArgs<TI> a0, const RelArgs<TI>* tab) {
  const RelArgs<TI> a = TAB ? tab[blockIdx.z] : a0;
  const Geo& g = a.g;
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  const long long nm = (long long)g.nz * g.ny * (g.nx / 2 + 1);
  if (i >= nm) return;
  const int kx = int(i % (g.nx / 2 + 1));
  const int ky = int((i / (g.nx / 2 + 1)) % g.ny);
  const int z = int(i / ((long long)(g.nx / 2 + 1) * g.ny));
  const long long rw = spec_pos(g, z, ky, kx);
  const cpx<TI> v = a.in[a.tb_in ? gm_index(g, a.tb_in, z, ky, kx) : rw];
  a.out[a.tb_out ? gm_index(g, a.tb_out, z, ky, kx) : rw] = mk<double>(double(v.x), double(v.y));
}

}  // namespace vb
