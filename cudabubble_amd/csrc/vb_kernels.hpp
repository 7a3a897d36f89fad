// vb_kernels.hpp -- kernel templates of the fused pseudo-spectral step.
//
// Pass structure of one nonlinear evaluation N^(v) (reference:
// belastic::get_nld_term, belastic.cpp:473-488), with v^ the stage spectrum:
//
//   z-kernel  : v^ -> z-inverse of v^ and of -|k|^2 v^   (+ Parseval |grad v|^2)
//   y-kernel  : y-inverse of both
//   xa-kernel : x-inverse (two-for-one) -> v, lap v; f, tmp1, sums of v,
//               (v^2-1)^2, f^2; x-forward of tmp1            (P1 fused)
//   y, z(lap), y kernels : lap tmp1 spectrum round trip
//   finalize  : V, A, E_b, lambda1, lambda2, E from the block partials
//   xb-kernel : x-inverse -> lap tmp1; nld, N; x-forward of N  (P2 fused)
//   y-kernel  : y-forward of N
//   zstage    : z-forward -> N^, ETD combine with the state, and the first
//               kernel of the next stage (z-inverse of the next field)
#pragma once
#include "vb_device.hpp"

namespace vb {

constexpr int XWG = 256;  // threads per x-kernel workgroup

template <int N> constexpr int xlanes() { return XWG / TFof<N>(); }

// Column tile width of the y / z kernels (columns that are contiguous in
// memory): enough columns for >= 128-byte segments and >= 256 threads.
template <typename T, int N> constexpr int tbw() {
  constexpr int seg = 128 / int(sizeof(cpx<T>));
  constexpr int occ = 256 / TFof<N>();
  return seg > occ ? seg : occ;
}

// Z_p = A_ext(p) + i * B_ext(p): Hermitian extension along x of two half
// spectra.  The imaginary parts at kx = 0 and kx = N/2 are dropped, which is
// what a real-output inverse does with them.
template <typename T, int N>
__device__ __forceinline__ cpx<T> pack2(const cpx<T>* __restrict__ A, const cpx<T>* __restrict__ B, int p) {
  const bool mir = p > N / 2;
  const int k = mir ? N - p : p;
  cpx<T> a = A[k], b = B[k];
  if (k == 0 || k == N / 2) { a.y = T(0); b.y = T(0); }
  if (mir) { a.y = -a.y; b.y = -b.y; }
  return mk<T>(a.x - b.y, a.y + b.x);
}

// After a forward transform of Z = a + i b (a, b real rows), write the half
// spectra of a and b, times s.  Uses the lane's LDS; ends with a barrier.
template <typename T, int N, class Lds>
__device__ __forceinline__ void unpack_store(const cpx<T>* x, int j, Lds& lds, cpx<T>* __restrict__ outA,
                                             cpx<T>* __restrict__ outB, T s, bool valid) {
  constexpr int E = Plan<N>::E, TF = N / E, NH = N / 2 + 1;
#pragma unroll
  for (int e = 0; e < E; ++e) lds(j + TF * e) = x[e];
  __syncthreads();
#pragma unroll
  for (int e = 0; e < (NH + TF - 1) / TF; ++e) {
    const int k = j + TF * e;
    if (k < NH && valid) {
      const cpx<T> zk = lds(k);
      const cpx<T> zm = lds(k == 0 ? 0 : N - k);
      const T h = T(0.5) * s;
      // a_k = (Z_k + conj Z_{N-k}) / 2,  b_k = (Z_k - conj Z_{N-k}) / (2i)
      outA[k] = mk<T>((zk.x + zm.x) * h, (zk.y - zm.y) * h);
      outB[k] = mk<T>((zk.y + zm.y) * h, (zm.x - zk.x) * h);
    }
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// plain x transforms (pairs of rows)
// ---------------------------------------------------------------------------
template <typename TI, typename T, int N>
__global__ void __launch_bounds__(XWG) k_xr2c(const TI* __restrict__ in, cpx<T>* __restrict__ out,
                                               long long npairs, const cpx<T>* __restrict__ tw, double scl) {
  constexpr int E = Plan<N>::E, TF = N / E, G = xlanes<N>(), LL = row_lds_len<N>(), NH = N / 2 + 1;
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
  fft<T, N, -1>(x, j, tw, lds);
  unpack_store<T, N>(x, j, lds, out + r0 * NH, out + (r0 + 1) * NH, T(scl), valid);
}

template <typename T, typename TO, int N>
__global__ void __launch_bounds__(XWG) k_xc2r(const cpx<T>* __restrict__ in, TO* __restrict__ out,
                                               long long npairs, const cpx<T>* __restrict__ tw) {
  constexpr int E = Plan<N>::E, TF = N / E, G = xlanes<N>(), LL = row_lds_len<N>(), NH = N / 2 + 1;
  __shared__ cpx<T> sm[G * LL];
  const int lane = threadIdx.x / TF, j = threadIdx.x % TF;
  RowLds<T> lds{sm + lane * LL};
  long long pair = (long long)blockIdx.x * G + lane;
  const bool valid = pair < npairs;
  if (!valid) pair = npairs - 1;
  const long long r0 = 2 * pair;
  cpx<T> x[E];
#pragma unroll
  for (int e = 0; e < E; ++e) x[e] = pack2<T, N>(in + r0 * NH, in + (r0 + 1) * NH, j + TF * e);
  fft<T, N, +1>(x, j, tw, lds);
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
  cpx<T>* Tt;        // out: x-forward of tmp1, times inv_n
  cpx<T>* W;         // out (Euler only): x-forward of f, times inv_n
  double* part;      // [gridDim.x][3]
  const cpx<T>* tw;
  long long npairs;
  double eps, kappa, inv_n;
};

template <typename T, int N, bool WITHF>
__global__ void __launch_bounds__(XWG) k_xa(XaArgs<T> a) {
  constexpr int E = Plan<N>::E, TF = N / E, G = xlanes<N>(), LL = row_lds_len<N>(), NH = N / 2 + 1;
  __shared__ cpx<T> sm[G * LL];
  __shared__ double red[3 * (XWG / 64)];
  const int lane = threadIdx.x / TF, j = threadIdx.x % TF;
  RowLds<T> lds{sm + lane * LL};
  long long pair = (long long)blockIdx.x * G + lane;
  const bool valid = pair < a.npairs;
  if (!valid) pair = a.npairs - 1;
  const long long r0 = 2 * pair;
  const double eps = a.eps, kappa = a.kappa;
  double acc[3] = {0.0, 0.0, 0.0};
  T tm[2][E];
  T ff[WITHF ? 2 : 1][E];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const long long r = r0 + h;
    cpx<T> x[E];
#pragma unroll
    for (int e = 0; e < E; ++e) x[e] = pack2<T, N>(a.P + r * NH, a.Q + r * NH, j + TF * e);
    fft<T, N, +1>(x, j, a.tw, lds);
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int p = j + TF * e;
      const double v = double(x[e].x), lv = double(x[e].y);
      const double w = v * v - 1.0;
      const double f = -lv * eps + w * v / eps;          // f_kernel, Calc.cu:160
      const double t1 = v * (v * v - 1.0 - kappa);       // nld_kernel1, Calc.cu:270
      if (valid) {
        acc[0] += v; acc[1] += w * w; acc[2] += f * f;
        a.v[r * N + p] = T(v);
        a.f[r * N + p] = T(f);
      }
      tm[h][e] = T(t1);
      if constexpr (WITHF) ff[h][e] = T(f);
    }
  }
  {
    cpx<T> z[E];
#pragma unroll
    for (int e = 0; e < E; ++e) z[e] = mk<T>(tm[0][e], tm[1][e]);
    fft<T, N, -1>(z, j, a.tw, lds);
    unpack_store<T, N>(z, j, lds, a.Tt + r0 * NH, a.Tt + (r0 + 1) * NH, T(a.inv_n), valid);
  }
  if constexpr (WITHF) {
    cpx<T> z[E];
#pragma unroll
    for (int e = 0; e < E; ++e) z[e] = mk<T>(ff[0][e], ff[1][e]);
    fft<T, N, -1>(z, j, a.tw, lds);
    unpack_store<T, N>(z, j, lds, a.W + r0 * NH, a.W + (r0 + 1) * NH, T(a.inv_n), valid);
  }
  block_sum_store<3, XWG>(acc, red, a.part + 3 * (long long)blockIdx.x);
}

// ---------------------------------------------------------------------------
// xb: stage suffix, P2 fused
// ---------------------------------------------------------------------------
template <typename T>
struct XbArgs {
  cpx<T>* Tt;        // in: x-half-spectra of lap tmp1; out: x-forward of N, times inv_n
  const T* v;
  const T* f;
  const Scal* sc;
  const cpx<T>* tw;
  long long npairs;
  double eps, kappa, D, eta, inv_n;
};

template <typename T, int N>
__global__ void __launch_bounds__(XWG) k_xb(XbArgs<T> a) {
  constexpr int E = Plan<N>::E, TF = N / E, G = xlanes<N>(), LL = row_lds_len<N>(), NH = N / 2 + 1;
  __shared__ cpx<T> sm[G * LL];
  const int lane = threadIdx.x / TF, j = threadIdx.x % TF;
  RowLds<T> lds{sm + lane * LL};
  long long pair = (long long)blockIdx.x * G + lane;
  const bool valid = pair < a.npairs;
  if (!valid) pair = a.npairs - 1;
  const long long r0 = 2 * pair;
  cpx<T> x[E];
#pragma unroll
  for (int e = 0; e < E; ++e) x[e] = pack2<T, N>(a.Tt + r0 * NH, a.Tt + (r0 + 1) * NH, j + TF * e);
  fft<T, N, +1>(x, j, a.tw, lds);
  const double l1 = a.sc->l1, l2 = a.sc->l2;
  const double ie = 1.0 / a.eps, ie2 = ie * ie, ie3 = ie * ie2;
  const double D = a.D, kappa = a.kappa, eta = a.eta;
  const double c32 = 2.0 * sqrt(2.0);
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int p = j + TF * e;
    double nn[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const long long r = r0 + h;
      const double dt1 = double(h ? x[e].y : x[e].x);
      const double v = double(a.v[r * N + p]), f = double(a.f[r * N + p]);
      const double t1 = v * (v * v - 1.0 - kappa);
      const double nl = -dt1 * ie + D * t1 * ie3 + (3 * v * v - D - 1) * ie2 * f;   // Calc.cu:282
      nn[h] = -l1 * 0.5 - l2 * f * 3.0 / c32 - nl * eta;                           // Calc.cu:435
    }
    x[e] = mk<T>(T(nn[0]), T(nn[1]));
  }
  fft<T, N, -1>(x, j, a.tw, lds);
  unpack_store<T, N>(x, j, lds, a.Tt + r0 * NH, a.Tt + (r0 + 1) * NH, T(a.inv_n), valid);
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
};

template <typename T, int N>
__global__ void __launch_bounds__(XWG) k_xe(XeArgs<T> a) {
  constexpr int E = Plan<N>::E, TF = N / E, G = xlanes<N>(), LL = row_lds_len<N>(), NH = N / 2 + 1;
  __shared__ cpx<T> sm[G * LL];
  const int lane = threadIdx.x / TF, j = threadIdx.x % TF;
  RowLds<T> lds{sm + lane * LL};
  long long pair = (long long)blockIdx.x * G + lane;
  const bool valid = pair < a.npairs;
  if (!valid) pair = a.npairs - 1;
  const long long r0 = 2 * pair;
  cpx<T> x[E];
#pragma unroll
  for (int e = 0; e < E; ++e) x[e] = pack2<T, N>(a.W + r0 * NH, a.W + (r0 + 1) * NH, j + TF * e);
  fft<T, N, +1>(x, j, a.tw, lds);
  const double l1 = a.sc->l1, l2 = a.sc->l2, eta = a.eta, eps = a.eps, t = a.t;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int p = j + TF * e;
    double nv[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const long long r = r0 + h;
      const double lf = double(h ? x[e].y : x[e].x);
      const double v = double(a.v[r * N + p]), f = double(a.f[r * N + p]);
      const double g0 = -lf + (3.0 * v * v - 1.0) * f / (eps * eps);    // Calc.cu:178
      const double rhs = a.shape ? g0 : l1 * 0.5 + l2 * f + g0 * eta;    // Calc.cu:417
      nv[h] = v + t * rhs;                                               // Calc.cu:400
    }
    x[e] = mk<T>(T(nv[0]), T(nv[1]));
  }
  fft<T, N, -1>(x, j, a.tw, lds);
  unpack_store<T, N>(x, j, lds, a.out + r0 * NH, a.out + (r0 + 1) * NH, T(a.inv_n), valid);
}

// ---------------------------------------------------------------------------
// y kernels: complex transforms along y, in place, batch over (z, kx).
// blockIdx.y selects one of two arrays.
// ---------------------------------------------------------------------------
template <typename T, int N, int S, int TB = tbw<T, N>()>
__global__ void __launch_bounds__((TB * TFof<N>())) k_ypass(cpx<T>* b0, cpx<T>* b1, Geo g,
                                                          const cpx<T>* __restrict__ tw) {
  constexpr int E = Plan<N>::E, TF = N / E;
  __shared__ cpx<T> sm[N * TB];
  cpx<T>* buf = blockIdx.y ? b1 : b0;
  const int c = threadIdx.x % TB, j = threadIdx.x / TB;
  ColLds<T, TB> lds{sm + c};
  const long long batch = (long long)g.nz * g.nh;
  long long b = (long long)blockIdx.x * TB + c;
  const bool valid = b < batch;
  if (!valid) b = batch - 1;
  const long long z = b / g.nh, kx = b % g.nh;
  cpx<T>* col = buf + z * (long long)N * g.nh + kx;
  cpx<T> x[E];
#pragma unroll
  for (int e = 0; e < E; ++e) x[e] = col[(long long)(j + TF * e) * g.nh];
  fft<T, N, S>(x, j, tw, lds);
  if (valid) {
#pragma unroll
    for (int e = 0; e < E; ++e) col[(long long)(j + TF * e) * g.nh] = x[e];
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

template <int TB>
__device__ __forceinline__ ZIdx zidx(const Geo& g) {
  ZIdx r;
  const long long plane = (long long)g.ny * g.nh;
  long long q = (long long)blockIdx.x * TB + (threadIdx.x % TB);
  r.valid = q < plane;
  if (!r.valid) q = plane - 1;
  r.q = q;
  r.ky = int(q / g.nh);
  r.kx = int(q % g.nh);
  return r;
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
  const long long plane = (long long)g.ny * g.nh;
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

// forward along z, multiply by -|k|^2 (laplace_kernel, Calc.cu:114-122), inverse
template <typename T, int N, int TB = tbw<T, N>()>
__global__ void __launch_bounds__((TB * TFof<N>())) k_zlap(cpx<T>* buf, Geo g, const cpx<T>* __restrict__ tw) {
  constexpr int E = Plan<N>::E, TF = N / E;
  __shared__ cpx<T> sm[N * TB];
  const int j = threadIdx.x / TB;
  ColLds<T, TB> lds{sm + threadIdx.x % TB};
  const ZIdx id = zidx<TB>(g);
  const long long plane = (long long)g.ny * g.nh;
  cpx<T>* col = buf + id.q;
  cpx<T> x[E];
#pragma unroll
  for (int e = 0; e < E; ++e) x[e] = col[(long long)(j + TF * e) * plane];
  fft<T, N, -1>(x, j, tw, lds);
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int kz = j + TF * e;
    const T mult = T(-double(mindex(g, id.kx, id.ky, kz)) * g.kfac2);
    x[e] = scale(x[e], mult);
  }
  fft<T, N, +1>(x, j, tw, lds);
  if (id.valid) {
#pragma unroll
    for (int e = 0; e < E; ++e) col[(long long)(j + TF * e) * plane] = x[e];
  }
}

// Shared tail of zprep / zstage.  nf is the next stage field in FP64 (the
// state and the combine stay FP64 in both precisions): its Parseval partial
// of |grad|^2 and its mean (the k = 0 coefficient, written to dc[0]) are taken
// in FP64; then P = z-inverse(nf), Q = z-inverse(-|k|^2 nf) in T.
template <typename T, int N, int TB, class Lds>
__device__ __forceinline__ void next_field_tail(const cpx<double>* nf, int j, const ZIdx& id, const Geo& g, Lds& lds,
                                                const cpx<T>* __restrict__ tw, cpx<T>* P, cpx<T>* Q,
                                                double* part, double* dc, double* red) {
  constexpr int E = Plan<N>::E, TF = N / E;
  const long long plane = (long long)g.ny * g.nh;
  double acc[1] = {0.0};
  cpx<T> x[E], y[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int kz = j + TF * e;
    const double re = nf[e].x, im = nf[e].y;
    if (id.valid) acc[0] += grad_w(g, id.kx, id.ky, kz) * (re * re + im * im);
    const double m = -double(mindex(g, id.kx, id.ky, kz)) * g.kfac2;
    x[e] = mk<T>(T(re), T(im));
    y[e] = mk<T>(T(re * m), T(im * m));
  }
  if (id.q == 0 && j == 0) dc[0] = nf[0].x;
  fft<T, N, +1>(x, j, tw, lds);
  if (id.valid) {
#pragma unroll
    for (int e = 0; e < E; ++e) P[id.q + (long long)(j + TF * e) * plane] = x[e];
  }
  fft<T, N, +1>(y, j, tw, lds);
  if (id.valid) {
#pragma unroll
    for (int e = 0; e < E; ++e) Q[id.q + (long long)(j + TF * e) * plane] = y[e];
  }
  block_sum_store<1, TB * TF>(acc, red, part + blockIdx.x);
}

template <typename T, int N>
__global__ void __launch_bounds__((tbw<T, N>() * TFof<N>())) k_zprep(const cpx<double>* __restrict__ vh, cpx<T>* P,
                                                                   cpx<T>* Q, double* part, double* dc, Geo g,
                                                                   const cpx<T>* __restrict__ tw) {
  constexpr int E = Plan<N>::E, TF = N / E, TB = tbw<T, N>();
  __shared__ cpx<T> sm[N * TB];
  __shared__ double red[(TB * TF) / 64];
  const int j = threadIdx.x / TB;
  ColLds<T, TB> lds{sm + threadIdx.x % TB};
  const ZIdx id = zidx<TB>(g);
  const long long plane = (long long)g.ny * g.nh;
  cpx<double> x[E];
#pragma unroll
  for (int e = 0; e < E; ++e) x[e] = vh[id.q + (long long)(j + TF * e) * plane];
  next_field_tail<T, N, TB>(x, j, id, g, lds, tw, P, Q, part, dc, red);
}

template <typename T>
struct ZsArgs {
  const cpx<T>* U;        // x,y-forward of N (normalised)
  const cpx<double>* uh;  // state spectrum
  cpx<double>* pre;       // stage precombination
  cpx<double>* acc;       // final-combine accumulator
  cpx<double>* uout;      // new state spectrum (last stage)
  cpx<T>* P;
  cpx<T>* Q;
  double* part;
  double* dc;
  const Coef* lut;
  const cpx<T>* tw;
  Geo g;
};

// SCH = 4 (ETDRK4, belastic.cpp:519-552 / Calc.cu:488-577) or 2 (ETDRK2,
// belastic.cpp:491-517 / Calc.cu:446-486); STG = stage whose N^ this is.
template <typename T, int N, int SCH, int STG>
__global__ void __launch_bounds__((tbw<T, N>() * TFof<N>())) k_zstage(ZsArgs<T> a) {
  constexpr int E = Plan<N>::E, TF = N / E, TB = tbw<T, N>();
  __shared__ cpx<T> sm[N * TB];
  __shared__ double red[(TB * TF) / 64];
  const int j = threadIdx.x / TB;
  ColLds<T, TB> lds{sm + threadIdx.x % TB};
  const Geo g = a.g;
  const ZIdx id = zidx<TB>(g);
  const long long plane = (long long)g.ny * g.nh;
  cpx<T> x[E];
#pragma unroll
  for (int e = 0; e < E; ++e) x[e] = a.U[id.q + (long long)(j + TF * e) * plane];
  fft<T, N, -1>(x, j, a.tw, lds);   // x = N^ of this stage
  cpx<double> nf[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int kz = j + TF * e;
    const long long i = id.q + (long long)kz * plane;
    const Coef& cf = a.lut[mindex(g, id.kx, id.ky, kz)];
    const double nr = double(x[e].x), ni = double(x[e].y);
    double fr, fi;   // next field
    if constexpr (SCH == 4 && STG == 1) {
      const double e1 = cf.c[0], q = cf.c[1];
      const cpx<double> u = a.uh[i];
      fr = e1 * u.x + q * nr;                         // etd41: a^
      fi = e1 * u.y + q * ni;
      if (id.valid) a.pre[i] = mk<double>(e1 * fr - q * nr, e1 * fi - q * ni);
    } else if constexpr (SCH == 4 && STG == 2) {
      const double e1 = cf.c[0], q = cf.c[1], e2 = cf.c[2], fa = cf.c[3], fb = cf.c[4];
      const cpx<double> u = a.uh[i];
      fr = e1 * u.x + q * nr;                         // etd42: b^
      fi = e1 * u.y + q * ni;
      if (id.valid) a.acc[i] = mk<double>(e2 * u.x + fa * nr + fb * nr, e2 * u.y + fa * ni + fb * ni);
    } else if constexpr (SCH == 4 && STG == 3) {
      const double q = cf.c[1], fb = cf.c[4];
      const cpx<double> p = a.pre[i];
      fr = p.x + 2.0 * q * nr;                        // etd43: c^
      fi = p.y + 2.0 * q * ni;
      if (id.valid) {
        const cpx<double> ac = a.acc[i];
        a.acc[i] = mk<double>(ac.x + fb * nr, ac.y + fb * ni);
      }
    } else if constexpr (SCH == 4 && STG == 4) {
      const double fc = cf.c[5];
      const cpx<double> ac = a.acc[i];
      fr = ac.x + fc * nr;                            // etd44: u^+
      fi = ac.y + fc * ni;
      if (id.valid) a.uout[i] = mk<double>(fr, fi);
    } else if constexpr (SCH == 2 && STG == 1) {
      const double ez = cf.c[0], q1 = cf.c[1], c2 = cf.c[2];
      const cpx<double> u = a.uh[i];
      fr = ez * u.x + q1 * nr;                        // etd1: a^
      fi = ez * u.y + q1 * ni;
      if (id.valid) a.pre[i] = mk<double>(fr - c2 * nr, fi - c2 * ni);
    } else {                                          // SCH == 2 && STG == 2
      const double c2 = cf.c[2];
      const cpx<double> p = a.pre[i];
      fr = p.x + c2 * nr;                             // etd2: u^+
      fi = p.y + c2 * ni;
      if (id.valid) a.uout[i] = mk<double>(fr, fi);
    }
    nf[e] = mk<double>(fr, fi);
  }
  next_field_tail<T, N, TB>(nf, j, id, g, lds, a.tw, a.P, a.Q, a.part, a.dc, red);
}

// widening copy of a spectrum into the FP64 state
template <typename T>
__global__ void k_widen(const cpx<T>* __restrict__ in, cpx<double>* __restrict__ out, long long n) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) out[i] = mk<double>(double(in[i].x), double(in[i].y));
}

}  // namespace vb
