// vb_size_impl.hpp -- launch wrappers for one transform length.  Included by
// exactly one translation unit per length with VB_N defined.
#pragma once
#include "vb_kernels.hpp"
#include "vb_ops.hpp"

namespace vb {

template <typename T, int N>
struct Launch {
  static constexpr int TF = TFof<N>();
  static constexpr int G = xlanes<N>();
  static constexpr int TB = tbw<T, N>();
  static constexpr int YT = TB * TF;

  static unsigned xgrid(long long npairs) { return unsigned((npairs + G - 1) / G); }

  static Rows rows(const Geo& g) { return Rows{g.ny, g.pitch, g.pp}; }
  static void xr2c_d(const double* in, cpx<T>* out, long long np, const cpx<T>* tw, double scl, const Geo& g,
                     hipStream_t s) {
    hipLaunchKernelGGL((k_xr2c<double, T, N>), dim3(xgrid(np)), dim3(XWG), 0, s, in, out, np, tw, scl, rows(g));
  }
  static void xr2c(const T* in, cpx<T>* out, long long np, const cpx<T>* tw, double scl, const Geo& g,
                   hipStream_t s) {
    hipLaunchKernelGGL((k_xr2c<T, T, N>), dim3(xgrid(np)), dim3(XWG), 0, s, in, out, np, tw, scl, rows(g));
  }
  static void xc2r_d(const cpx<T>* in, double* out, long long np, const cpx<T>* tw, const Geo& g, hipStream_t s) {
    hipLaunchKernelGGL((k_xc2r<T, double, N>), dim3(xgrid(np)), dim3(XWG), 0, s, in, out, np, tw, rows(g));
  }
  static void xc2r(const cpx<T>* in, T* out, long long np, const cpx<T>* tw, const Geo& g, hipStream_t s) {
    hipLaunchKernelGGL((k_xc2r<T, T, N>), dim3(xgrid(np)), dim3(XWG), 0, s, in, out, np, tw, rows(g));
  }
  static void xa(const XaArgs<T>& a, bool withf, hipStream_t s) {
    if (withf)
      hipLaunchKernelGGL((k_xa<T, N, true>), dim3(xgrid(a.npairs)), dim3(XWG), 0, s, a);
    else
      hipLaunchKernelGGL((k_xa<T, N, false>), dim3(xgrid(a.npairs)), dim3(XWG), 0, s, a);
  }
  static void xb(const XbArgs<T>& a, hipStream_t s) {
    hipLaunchKernelGGL((k_xb<T, N>), dim3(xgrid(a.npairs)), dim3(XWG), 0, s, a);
  }
  static void xe(const XeArgs<T>& a, hipStream_t s) {
    hipLaunchKernelGGL((k_xe<T, N>), dim3(xgrid(a.npairs)), dim3(XWG), 0, s, a);
  }
  static void ypass(cpx<T>* b0, cpx<T>* b1, int narr, int sign, const Geo& g, const cpx<T>* tw, hipStream_t s) {
    dim3 grid(unsigned(ygroups(g, TB)), unsigned(narr));
    if (sign < 0)
      hipLaunchKernelGGL((k_ypass<T, N, -1>), grid, dim3(YT), 0, s, b0, b1, g, tw);
    else
      hipLaunchKernelGGL((k_ypass<T, N, +1>), grid, dim3(YT), 0, s, b0, b1, g, tw);
  }
  static unsigned zgrid(const Geo& g) { return unsigned(zgroups(g, TB)); }
  static void zpass(cpx<T>* b, int sign, const Geo& g, const cpx<T>* tw, hipStream_t s) {
    if (sign < 0)
      hipLaunchKernelGGL((k_zpass<T, N, -1>), dim3(zgrid(g)), dim3(YT), 0, s, b, g, tw);
    else
      hipLaunchKernelGGL((k_zpass<T, N, +1>), dim3(zgrid(g)), dim3(YT), 0, s, b, g, tw);
  }
  static void zlap(cpx<T>* b, const Geo& g, const cpx<T>* tw, hipStream_t s) {
    hipLaunchKernelGGL((k_zlap<T, N>), dim3(zgrid(g)), dim3(YT), 0, s, b, g, tw);
  }
  static void zprep(const cpx<double>* vh, cpx<T>* P, cpx<T>* Q, double* part, double* dc, const Geo& g,
                    const cpx<T>* tw, hipStream_t s) {
    hipLaunchKernelGGL((k_zprep<T, N>), dim3(zgrid(g)), dim3(YT), 0, s, vh, P, Q, part, dc, g, tw);
  }
  static void zstage(const ZsArgs<T>& a, int sch, int stg, hipStream_t s) {
    const dim3 gr(zgrid(a.g)), bl(YT);
    if (sch == 4) {
      switch (stg) {
        case 1: hipLaunchKernelGGL((k_zstage<T, N, 4, 1>), gr, bl, 0, s, a); break;
        case 2: hipLaunchKernelGGL((k_zstage<T, N, 4, 2>), gr, bl, 0, s, a); break;
        case 3: hipLaunchKernelGGL((k_zstage<T, N, 4, 3>), gr, bl, 0, s, a); break;
        default: hipLaunchKernelGGL((k_zstage<T, N, 4, 4>), gr, bl, 0, s, a); break;
      }
    } else {
      if (stg == 1) hipLaunchKernelGGL((k_zstage<T, N, 2, 1>), gr, bl, 0, s, a);
      else hipLaunchKernelGGL((k_zstage<T, N, 2, 2>), gr, bl, 0, s, a);
    }
  }

  static SizeOps<T> table() {
    SizeOps<T> o;
    o.N = N;
    o.xlanes = G;
    o.ytile = TB;
    o.ythreads = YT;
    o.xr2c_d = &xr2c_d;
    o.xr2c = &xr2c;
    o.xc2r_d = &xc2r_d;
    o.xc2r = &xc2r;
    o.xa = &xa;
    o.xb = &xb;
    o.xe = &xe;
    o.ypass = &ypass;
    o.zpass = &zpass;
    o.zlap = &zlap;
    o.zprep = &zprep;
    o.zstage = &zstage;
    return o;
  }
};

}  // namespace vb

#define VB_DEFINE_SIZE(n)                                              \
  namespace vb {                                                       \
  template <> const SizeOps<double>* ops_for<double, n>() {            \
    static const SizeOps<double> t = Launch<double, n>::table();       \
    return &t;                                                         \
  }                                                                    \
  template <> const SizeOps<float>* ops_for<float, n>() {              \
    static const SizeOps<float> t = Launch<float, n>::table();         \
    return &t;                                                         \
  }                                                                    \
  }
