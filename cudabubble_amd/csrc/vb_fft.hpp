// vb_fft.hpp -- device-side building blocks of the 3-D real FFT for CDNA4.
//
// One 1-D transform of length N is carried by TF threads ("a lane"); each
// thread holds E = N / TF elements in registers at positions j + TF*e
// (j = thread index inside the lane, e = 0..E-1).  Every Stockham pass reads
// its inputs at exactly those positions, so the first pass takes its data
// straight from global memory, the last pass leaves its results in the same
// positions (ready for coalesced stores), and the passes in between exchange
// data through LDS.  Radices are 2, 3, 4, 8, 12 (3 x 4 in registers), 16
// (4 x 4) and 24 (3 x 8); the radix plan per length is fixed at compile time
// so every loop unrolls.
//
// Twiddles come from a table tw[m] = exp(-2 pi i m / N) computed on the host
// in extended precision and rounded once to the storage type.
#pragma once
#include <hip/hip_runtime.h>

namespace vb {

template <typename T>
struct cpx {
  T x, y;
};

template <typename T>
__device__ __forceinline__ cpx<T> mk(T a, T b) {
  cpx<T> r; r.x = a; r.y = b; return r;
}
template <typename T>
__device__ __forceinline__ cpx<T> operator+(cpx<T> a, cpx<T> b) { return mk(a.x + b.x, a.y + b.y); }
template <typename T>
__device__ __forceinline__ cpx<T> operator-(cpx<T> a, cpx<T> b) { return mk(a.x - b.x, a.y - b.y); }
template <typename T>
__device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) {
  return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T>
__device__ __forceinline__ cpx<T> scale(cpx<T> a, T s) { return mk(a.x * s, a.y * s); }
template <typename T>
__device__ __forceinline__ cpx<T> conj(cpx<T> a) { return mk(a.x, -a.y); }
// multiply by S*i (S = -1 forward, +1 inverse)
template <int S, typename T>
__device__ __forceinline__ cpx<T> rot(cpx<T> a) {
  return S < 0 ? mk(a.y, -a.x) : mk(-a.y, a.x);
}

// ---------------------------------------------------------------------------
// radix plans: E elements per thread, TF threads per transform, radices
// ---------------------------------------------------------------------------
// V = 0 is the plan of every kernel.  V = 1 trades registers for fewer
// Stockham passes (fewer LDS exchanges and workgroup barriers); it differs
// only where measured to pay: 256 = 16 * 16, 128 = 16 * 8 and 192 = 24 * 8
// for the FP32 column kernels and x kernels, 512 = 16 * 8 * 4 for the FP32 x
// kernel.
// V = 3 keeps V = 0's elements per thread with fewer passes (192 = 12 * 4 * 4,
// the FP32 stage combines).  (V = 2, more threads per transform and fewer
// registers, measured slower at 256 and is not defined.)
template <int N, int V = 0> struct Plan;
template <> struct Plan<16, 0>  { static constexpr int E = 4,  NP = 2; static constexpr int R[2] = {4, 4}; };
template <> struct Plan<32, 0>  { static constexpr int E = 8,  NP = 2; static constexpr int R[2] = {8, 4}; };
template <> struct Plan<48, 0>  { static constexpr int E = 12, NP = 3; static constexpr int R[3] = {4, 4, 3}; };
template <> struct Plan<64, 0>  { static constexpr int E = 8,  NP = 2; static constexpr int R[2] = {8, 8}; };
template <> struct Plan<96, 0>  { static constexpr int E = 12, NP = 4; static constexpr int R[4] = {4, 4, 2, 3}; };
template <> struct Plan<128, 0> { static constexpr int E = 8,  NP = 3; static constexpr int R[3] = {8, 4, 4}; };
template <> struct Plan<192, 0> { static constexpr int E = 12, NP = 4; static constexpr int R[4] = {4, 4, 4, 3}; };
template <> struct Plan<256, 0> { static constexpr int E = 8,  NP = 3; static constexpr int R[3] = {8, 8, 4}; };
template <> struct Plan<384, 0> { static constexpr int E = 12, NP = 5; static constexpr int R[5] = {4, 4, 4, 2, 3}; };
template <> struct Plan<512, 0> { static constexpr int E = 8,  NP = 3; static constexpr int R[3] = {8, 8, 8}; };
template <int N> struct Plan<N, 1> : Plan<N, 0> {};
template <> struct Plan<256, 1> { static constexpr int E = 16, NP = 2; static constexpr int R[2] = {16, 16}; };
template <> struct Plan<128, 1> { static constexpr int E = 16, NP = 2; static constexpr int R[2] = {16, 8}; };
template <> struct Plan<512, 1> { static constexpr int E = 16, NP = 3; static constexpr int R[3] = {16, 8, 4}; };
template <> struct Plan<192, 1> { static constexpr int E = 24, NP = 2; static constexpr int R[2] = {24, 8}; };
template <> struct Plan<192, 3> { static constexpr int E = 12, NP = 3; static constexpr int R[3] = {12, 4, 4}; };

template <int N, int V = 0> constexpr int TFof() { return N / Plan<N, V>::E; }

template <int N, int P, int V>
constexpr int ns_before() {
  int ns = 1;
  for (int i = 0; i < P; ++i) ns *= Plan<N, V>::R[i];
  return ns;
}

// ---------------------------------------------------------------------------
// small DFTs in registers
// ---------------------------------------------------------------------------
template <int R, int S, typename T> struct Dft;

template <int S, typename T> struct Dft<2, S, T> {
  static __device__ __forceinline__ void run(cpx<T>* v) {
    cpx<T> a = v[0], b = v[1];
    v[0] = a + b; v[1] = a - b;
  }
};

template <int S, typename T> struct Dft<3, S, T> {
  static __device__ __forceinline__ void run(cpx<T>* v) {
    const T h = T(0.86602540378443864676372317075294);   // sqrt(3)/2
    cpx<T> t1 = v[1] + v[2];
    cpx<T> t2 = mk(v[0].x - T(0.5) * t1.x, v[0].y - T(0.5) * t1.y);
    cpx<T> d = scale(v[1] - v[2], h);
    cpx<T> r = rot<S>(d);
    v[0] = v[0] + t1;
    v[1] = t2 + r;
    v[2] = t2 - r;
  }
};

template <int S, typename T> struct Dft<4, S, T> {
  static __device__ __forceinline__ void run(cpx<T>* v) {
    cpx<T> a = v[0] + v[2], b = v[0] - v[2];
    cpx<T> c = v[1] + v[3], d = rot<S>(v[1] - v[3]);
    v[0] = a + c; v[2] = a - c;
    v[1] = b + d; v[3] = b - d;
  }
};

template <int S, typename T> struct Dft<8, S, T> {
  static __device__ __forceinline__ void run(cpx<T>* v) {
    const T c = T(0.70710678118654752440084436210485);
    cpx<T> e[4] = {v[0], v[2], v[4], v[6]};
    cpx<T> o[4] = {v[1], v[3], v[5], v[7]};
    Dft<4, S, T>::run(e);
    Dft<4, S, T>::run(o);
    // W8^1 = (1 + S i) c, W8^2 = S i, W8^3 = (-1 + S i) c
    cpx<T> o1 = scale(o[1] + rot<S>(o[1]), c);
    cpx<T> o2 = rot<S>(o[2]);
    cpx<T> o3 = scale(rot<S>(o[3]) - o[3], c);
    v[0] = e[0] + o[0]; v[4] = e[0] - o[0];
    v[1] = e[1] + o1;   v[5] = e[1] - o1;
    v[2] = e[2] + o2;   v[6] = e[2] - o2;
    v[3] = e[3] + o3;   v[7] = e[3] - o3;
  }
};

// 16 points as 4 x 4: n = n2 + 4 n1, k = k1 + 4 k2, so
// X[k1 + 4 k2] = sum_n2 W4^(n2 k2) W16^(n2 k1) sum_n1 x[n2 + 4 n1] W4^(n1 k1).
template <int S, typename T> struct Dft<16, S, T> {
  static __device__ __forceinline__ void run(cpx<T>* v) {
    cpx<T> a[4][4];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
      cpx<T> t[4] = {v[n2], v[n2 + 4], v[n2 + 8], v[n2 + 12]};
      Dft<4, S, T>::run(t);
#pragma unroll
      for (int k1 = 0; k1 < 4; ++k1) a[n2][k1] = t[k1];
    }
    // W16^e = cos(2 pi e / 16) + S i sin(2 pi e / 16) for e = n2 k1 in {1, 2, 3, 4, 6, 9}
    const T c1 = T(0.92387953251128675613), s1 = T(0.38268343236508977173), c2 = T(0.70710678118654752440);
#pragma unroll
    for (int n2 = 1; n2 < 4; ++n2) {
#pragma unroll
      for (int k1 = 1; k1 < 4; ++k1) {
        const int e = n2 * k1;
        T c, sn;
        if (e == 1) { c = c1; sn = s1; }
        else if (e == 2) { c = c2; sn = c2; }
        else if (e == 3) { c = s1; sn = c1; }
        else if (e == 4) { c = T(0); sn = T(1); }
        else if (e == 6) { c = -c2; sn = c2; }
        else { c = -c1; sn = -s1; }   // e = 9
        a[n2][k1] = cmul(a[n2][k1], mk<T>(c, T(S) * sn));
      }
    }
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
      cpx<T> t[4] = {a[0][k1], a[1][k1], a[2][k1], a[3][k1]};
      Dft<4, S, T>::run(t);
#pragma unroll
      for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = t[k2];
    }
  }
};

// 12 points as 3 x 4: n = n2 + 3 n1, k = k1 + 4 k2, so
// X[k1 + 4 k2] = sum_n2 W3^(n2 k2) W12^(n2 k1) sum_n1 x[n2 + 3 n1] W4^(n1 k1).
template <int S, typename T> struct Dft<12, S, T> {
  static __device__ __forceinline__ void run(cpx<T>* v) {
    // cos and sin of 2 pi e / 12 (multiples of 30 degrees), e = 0..6
    constexpr double c30 = 0.86602540378443864676;
    constexpr double cs[7][2] = {{1, 0}, {c30, 0.5}, {0.5, c30}, {0, 1}, {-0.5, c30}, {-c30, 0.5}, {-1, 0}};
    cpx<T> a[3][4];
#pragma unroll
    for (int n2 = 0; n2 < 3; ++n2) {
      cpx<T> t[4] = {v[n2], v[n2 + 3], v[n2 + 6], v[n2 + 9]};
      Dft<4, S, T>::run(t);
#pragma unroll
      for (int k1 = 0; k1 < 4; ++k1) a[n2][k1] = t[k1];
    }
#pragma unroll
    for (int n2 = 1; n2 < 3; ++n2) {
#pragma unroll
      for (int k1 = 1; k1 < 4; ++k1) {
        const int e = n2 * k1;
        a[n2][k1] = cmul(a[n2][k1], mk<T>(T(cs[e][0]), T(S * cs[e][1])));
      }
    }
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
      cpx<T> t[3] = {a[0][k1], a[1][k1], a[2][k1]};
      Dft<3, S, T>::run(t);
#pragma unroll
      for (int k2 = 0; k2 < 3; ++k2) v[k1 + 4 * k2] = t[k2];
    }
  }
};

// 24 points as 3 x 8: n = n2 + 3 n1, k = k1 + 8 k2, so
// X[k1 + 8 k2] = sum_n2 W3^(n2 k2) W24^(n2 k1) sum_n1 x[n2 + 3 n1] W8^(n1 k1).
template <int S, typename T> struct Dft<24, S, T> {
  static __device__ __forceinline__ void run(cpx<T>* v) {
    // cos and sin of 2 pi e / 24 (multiples of 15 degrees), e = 0..14
    constexpr double c15 = 0.96592582628906828675, s15 = 0.25881904510252076235;
    constexpr double c30 = 0.86602540378443864676, c45 = 0.70710678118654752440;
    constexpr double cs[15][2] = {{1, 0},        {c15, s15},   {c30, 0.5},  {c45, c45},   {0.5, c30},
                                  {s15, c15},    {0, 1},       {-s15, c15}, {-0.5, c30},  {-c45, c45},
                                  {-c30, 0.5},   {-c15, s15},  {-1, 0},     {-c15, -s15}, {-c30, -0.5}};
    cpx<T> a[3][8];
#pragma unroll
    for (int n2 = 0; n2 < 3; ++n2) {
      cpx<T> t[8];
#pragma unroll
      for (int n1 = 0; n1 < 8; ++n1) t[n1] = v[n2 + 3 * n1];
      Dft<8, S, T>::run(t);
#pragma unroll
      for (int k1 = 0; k1 < 8; ++k1) a[n2][k1] = t[k1];
    }
#pragma unroll
    for (int n2 = 1; n2 < 3; ++n2) {
#pragma unroll
      for (int k1 = 1; k1 < 8; ++k1) {
        const int e = n2 * k1;
        a[n2][k1] = cmul(a[n2][k1], mk<T>(T(cs[e][0]), T(S * cs[e][1])));
      }
    }
#pragma unroll
    for (int k1 = 0; k1 < 8; ++k1) {
      cpx<T> t[3] = {a[0][k1], a[1][k1], a[2][k1]};
      Dft<3, S, T>::run(t);
#pragma unroll
      for (int k2 = 0; k2 < 3; ++k2) v[k1 + 8 * k2] = t[k2];
    }
  }
};

// ---------------------------------------------------------------------------
// Stockham passes.  Lds is a functor: lds(pos) -> cpx<T>& for this lane.
// ---------------------------------------------------------------------------
// Butterflies of pass P for one transform held in x.  Non-last passes write
// their outputs to LDS; the last pass leaves them in y (same positions).
template <typename T, int N, int S, int P, int V, class Lds>
__device__ __forceinline__ void pass_body(const cpx<T>* x, cpx<T>* y, int j, const cpx<T>* __restrict__ tw, Lds& lds) {
  constexpr int E = Plan<N, V>::E;
  constexpr int TF = N / E;
  constexpr int R = Plan<N, V>::R[P];
  constexpr int NS = ns_before<N, P, V>();
  constexpr int ER = E / R;
  constexpr bool last = (P == Plan<N, V>::NP - 1);
  constexpr int TST = N / (NS * R);
#pragma unroll
  for (int t = 0; t < ER; ++t) {
    const int b = j + TF * t;
    const int k = b % NS;
    cpx<T> v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = x[t + r * ER];
    if constexpr (NS > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) {
        cpx<T> w = tw[r * k * TST];
        if constexpr (S > 0) w.y = -w.y;
        v[r] = cmul(v[r], w);
      }
    }
    Dft<R, S, T>::run(v);
    if constexpr (last) {
#pragma unroll
      for (int q = 0; q < R; ++q) y[t + q * ER] = v[q];
    } else {
      const int base = (b / NS) * NS * R + k;
#pragma unroll
      for (int q = 0; q < R; ++q) lds(base + q * NS) = v[q];
    }
  }
}

template <typename T, int N, int S, int P, int V>
struct Pass {
  template <class Lds>
  static __device__ __forceinline__ void run(cpx<T>* x, int j, const cpx<T>* __restrict__ tw, Lds& lds) {
    constexpr int E = Plan<N, V>::E;
    constexpr int TF = N / E;
    constexpr bool last = (P == Plan<N, V>::NP - 1);
    cpx<T> y[last ? E : 1];
    pass_body<T, N, S, P, V>(x, y, j, tw, lds);
    if constexpr (last) {
#pragma unroll
      for (int e = 0; e < E; ++e) x[e] = y[e];
    } else {
      __syncthreads();
#pragma unroll
      for (int e = 0; e < E; ++e) x[e] = lds(j + TF * e);
      __syncthreads();
      Pass<T, N, S, P + 1, V>::run(x, j, tw, lds);
    }
  }
};

// Two independent transforms through the same passes: one set of barriers
// for both, and twice the independent work between barriers.
template <typename T, int N, int S, int P, int V>
struct Pass2 {
  template <class Lds>
  static __device__ __forceinline__ void run(cpx<T>* x0, cpx<T>* x1, int j, const cpx<T>* __restrict__ tw,
                                             Lds& l0, Lds& l1) {
    constexpr int E = Plan<N, V>::E;
    constexpr int TF = N / E;
    constexpr bool last = (P == Plan<N, V>::NP - 1);
    cpx<T> y0[last ? E : 1], y1[last ? E : 1];
    pass_body<T, N, S, P, V>(x0, y0, j, tw, l0);
    pass_body<T, N, S, P, V>(x1, y1, j, tw, l1);
    if constexpr (last) {
#pragma unroll
      for (int e = 0; e < E; ++e) {
        x0[e] = y0[e];
        x1[e] = y1[e];
      }
    } else {
      __syncthreads();
#pragma unroll
      for (int e = 0; e < E; ++e) {
        x0[e] = l0(j + TF * e);
        x1[e] = l1(j + TF * e);
      }
      __syncthreads();
      Pass2<T, N, S, P + 1, V>::run(x0, x1, j, tw, l0, l1);
    }
  }
};

// Full 1-D transform (unnormalised): S = -1 forward, +1 inverse.
// Contains (NP - 1) * 2 workgroup barriers; every thread of the workgroup
// must call it the same number of times.
template <typename T, int N, int S, int V = 0, class Lds>
__device__ __forceinline__ void fft(cpx<T>* x, int j, const cpx<T>* __restrict__ tw, Lds& lds) {
  Pass<T, N, S, 0, V>::run(x, j, tw, lds);
}

// Two transforms of the same length and direction, each with its own LDS view.
template <typename T, int N, int S, int V = 0, class Lds>
__device__ __forceinline__ void fft2(cpx<T>* x0, cpx<T>* x1, int j, const cpx<T>* __restrict__ tw, Lds& l0,
                                     Lds& l1) {
  Pass2<T, N, S, 0, V>::run(x0, x1, j, tw, l0, l1);
}

// LDS view for a lane whose transform is stored contiguously (x kernels):
// one pad slot every 8 entries keeps stride-8 Stockham writes conflict-free.
template <typename T>
struct RowLds {
  cpx<T>* base;
  __device__ __forceinline__ cpx<T>& operator()(int pos) const { return base[pos + (pos >> 3)]; }
};
template <int N> constexpr int row_lds_len() { return N + N / 8; }

// LDS view for column tiles (y and z kernels): TB interleaved transforms,
// column index fastest.
template <typename T, int TB>
struct ColLds {
  cpx<T>* base;  // already offset by the column index
  __device__ __forceinline__ cpx<T>& operator()(int pos) const { return base[pos * TB]; }
};

}  // namespace vb
