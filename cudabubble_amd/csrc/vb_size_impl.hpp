// vb_size_impl.hpp -- launch wrappers for one transform length.  Included by
// exactly one translation unit per length with VB_N defined.
#pragma once
#include <algorithm>

#include "vb_kernels.hpp"
#include "vb_ops.hpp"

namespace vb {

template <typename T, int N>
struct Launch {
  static constexpr int TF = TFof<N>();
  static constexpr int G = xlanes<N>();
  static constexpr int TB = tbw<T, N>();
  static constexpr int YT = TB * TF;
  static constexpr int CV = colplan<T, N>();          // plan of the y passes and the z Laplacian
  static constexpr int CB = colw<T, N>();             // their columns per workgroup
  static constexpr int CT = CB * TFof<N, CV>();       // and threads per workgroup
  template <int SCH, int STG> static constexpr int zs_threads() { return TB * TFof<N, zsplan<T, N, SCH, STG>()>(); }

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
    const dim3 gr(xgrid(a.npairs)), bl(XWG);
    if (g_recorder) return record_launch(&launch_tab, withf ? OP_XA_F : OP_XA, a, gr, bl);
    if (withf)
      hipLaunchKernelGGL((k_xa<T, N, true>), gr, bl, 0, s, a, nullptr);
    else
      hipLaunchKernelGGL((k_xa<T, N, false>), gr, bl, 0, s, a, nullptr);
  }
  static void xb(const XbArgs<T>& a, hipStream_t s) {
    const dim3 gr(xgrid(a.npairs)), bl(XWG);
    if (g_recorder) return record_launch(&launch_tab, OP_XB, a, gr, bl);
    hipLaunchKernelGGL((k_xb<T, N>), gr, bl, 0, s, a, nullptr);
  }
  static void xe(const XeArgs<T>& a, hipStream_t s) {
    hipLaunchKernelGGL((k_xe<T, N>), dim3(xgrid(a.npairs)), dim3(XWG), 0, s, a);
  }
  static void ypass(cpx<T>* b0, cpx<T>* b1, int narr, int sign, const Geo& g, const cpx<T>* tw, hipStream_t s) {
    const dim3 grid(unsigned(ygroups(g, CB)), unsigned(narr)), bl(CT);
    YArgs<T> a{b0, b1, tw, g};
    if (g_recorder) return record_launch(&launch_tab, sign < 0 ? OP_Y_FWD : OP_Y_INV, a, grid, bl);
    if (sign < 0)
      hipLaunchKernelGGL((k_ypass<T, N, -1, CB, false, CV>), grid, bl, 0, s, a, nullptr);
    else
      hipLaunchKernelGGL((k_ypass<T, N, +1, CB, false, CV>), grid, bl, 0, s, a, nullptr);
  }
  static unsigned zgrid(const Geo& g) { return unsigned(zgroups(g, TB)); }
  static void zpass(cpx<T>* b, int sign, const Geo& g, const cpx<T>* tw, hipStream_t s) {
    if (sign < 0)
      hipLaunchKernelGGL((k_zpass<T, N, -1>), dim3(zgrid(g)), dim3(YT), 0, s, b, g, tw);
    else
      hipLaunchKernelGGL((k_zpass<T, N, +1>), dim3(zgrid(g)), dim3(YT), 0, s, b, g, tw);
  }
  static void zlap(cpx<T>* b, const Geo& g, const cpx<T>* tw, const FinArgs* fin, hipStream_t s) {
    // the integrals take one workgroup per 1024 x partials, up to kFinShares
    const int nf = fin ? std::min(kFinShares, std::max(1, (fin->nxb + 1023) / 1024)) : 0;
    const dim3 gr(unsigned(zgroups(g, CB)) + unsigned(nf)), bl(CT);
    ZlArgs<T> a{b, tw, g, fin ? *fin : FinArgs{}, nf};
    if (g_recorder) return record_launch(&launch_tab, OP_ZLAP, a, gr, bl);
    hipLaunchKernelGGL((k_zlap<T, N, CB, false, CV>), gr, bl, 0, s, a, nullptr);
  }
  static void zprep(const cpx<double>* vh, cpx<T>* P, cpx<T>* Q, double* part, double* dc, const Geo& g,
                    const cpx<T>* tw, hipStream_t s) {
    const dim3 gr(zgrid(g)), bl(YT);
    ZprArgs<T> a{vh, P, Q, part, dc, tw, g};
    if (g_recorder) return record_launch(&launch_tab, OP_ZPREP, a, gr, bl);
    hipLaunchKernelGGL((k_zprep<T, N>), gr, bl, 0, s, a, nullptr);
  }
  static void zstage(const ZsArgs<T>& a, int sch, int stg, hipStream_t s) {
    const dim3 gr(zgrid(a.g));
    int nt = YT;
    if (sch == 4) nt = stg == 1 ? zs_threads<4, 1>() : stg == 2 ? zs_threads<4, 2>() : stg == 3 ? zs_threads<4, 3>()
                                                                                           : zs_threads<4, 4>();
    else nt = stg == 1 ? zs_threads<2, 1>() : zs_threads<2, 2>();
    const dim3 bl = dim3((unsigned)nt);
    if (g_recorder) return record_launch(&launch_tab, sch == 4 ? OP_ZS41 + (stg - 1) : OP_ZS21 + (stg - 1), a, gr, bl);
    if (sch == 4) {
      switch (stg) {
        case 1: hipLaunchKernelGGL((k_zstage<T, N, 4, 1>), gr, bl, 0, s, a, nullptr); break;
        case 2: hipLaunchKernelGGL((k_zstage<T, N, 4, 2>), gr, bl, 0, s, a, nullptr); break;
        case 3: hipLaunchKernelGGL((k_zstage<T, N, 4, 3>), gr, bl, 0, s, a, nullptr); break;
        default: hipLaunchKernelGGL((k_zstage<T, N, 4, 4>), gr, bl, 0, s, a, nullptr); break;
      }
    } else {
      if (stg == 1) hipLaunchKernelGGL((k_zstage<T, N, 2, 1>), gr, bl, 0, s, a, nullptr);
      else hipLaunchKernelGGL((k_zstage<T, N, 2, 2>), gr, bl, 0, s, a, nullptr);
    }
  }

  // the table form of a recorded launch, over `lanes` solves (blockIdx.z)
  static void launch_tab(int op, const void* tab, dim3 gr, dim3 bl, int lanes, hipStream_t s) {
    gr.z = unsigned(lanes);
    switch (op) {
      case OP_XA: hipLaunchKernelGGL((k_xa<T, N, false, true>), gr, bl, 0, s, XaArgs<T>{}, (const XaArgs<T>*)tab); break;
      case OP_XA_F: hipLaunchKernelGGL((k_xa<T, N, true, true>), gr, bl, 0, s, XaArgs<T>{}, (const XaArgs<T>*)tab); break;
      case OP_XB: hipLaunchKernelGGL((k_xb<T, N, true>), gr, bl, 0, s, XbArgs<T>{}, (const XbArgs<T>*)tab); break;
      case OP_Y_FWD:
        hipLaunchKernelGGL((k_ypass<T, N, -1, CB, true, CV>), gr, bl, 0, s, YArgs<T>{}, (const YArgs<T>*)tab);
        break;
      case OP_Y_INV:
        hipLaunchKernelGGL((k_ypass<T, N, +1, CB, true, CV>), gr, bl, 0, s, YArgs<T>{}, (const YArgs<T>*)tab);
        break;
      case OP_ZLAP:
        hipLaunchKernelGGL((k_zlap<T, N, CB, true, CV>), gr, bl, 0, s, ZlArgs<T>{}, (const ZlArgs<T>*)tab);
        break;
      case OP_ZPREP:
        hipLaunchKernelGGL((k_zprep<T, N, TB, true>), gr, bl, 0, s, ZprArgs<T>{}, (const ZprArgs<T>*)tab);
        break;
      case OP_ZS41:
        hipLaunchKernelGGL((k_zstage<T, N, 4, 1, TB, true>), gr, bl, 0, s, ZsArgs<T>{}, (const ZsArgs<T>*)tab);
        break;
      case OP_ZS42:
        hipLaunchKernelGGL((k_zstage<T, N, 4, 2, TB, true>), gr, bl, 0, s, ZsArgs<T>{}, (const ZsArgs<T>*)tab);
        break;
      case OP_ZS43:
        hipLaunchKernelGGL((k_zstage<T, N, 4, 3, TB, true>), gr, bl, 0, s, ZsArgs<T>{}, (const ZsArgs<T>*)tab);
        break;
      case OP_ZS44:
        hipLaunchKernelGGL((k_zstage<T, N, 4, 4, TB, true>), gr, bl, 0, s, ZsArgs<T>{}, (const ZsArgs<T>*)tab);
        break;
      case OP_ZS21:
        hipLaunchKernelGGL((k_zstage<T, N, 2, 1, TB, true>), gr, bl, 0, s, ZsArgs<T>{}, (const ZsArgs<T>*)tab);
        break;
      case OP_ZS22:
        hipLaunchKernelGGL((k_zstage<T, N, 2, 2, TB, true>), gr, bl, 0, s, ZsArgs<T>{}, (const ZsArgs<T>*)tab);
        break;
      default: break;
    }
  }

  static SizeOps<T> table() {
    SizeOps<T> o;
    o.N = N;
    o.xlanes = G;
    o.ytile = TB;
    o.ctile = CB;
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
