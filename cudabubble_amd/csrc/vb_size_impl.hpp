// vb_size_impl.hpp -- launch wrappers for one transform length.  Included by
// exactly one translation unit per length with VB_N defined.
#pragma once
#include <algorithm>

#include "vb_kernels.hpp"
#include "vb_ops.hpp"

namespace vb {

// Kernel schedules of one length: bit 0 selects the other z tile (tbw_alt),
// bit 1 the other plan of the ETD x kernel (xrplan_alt).  Only FP32 has
// alternatives; a schedule is compiled where its choice differs from the
// default (see schedules()).
template <typename T, int N>
constexpr int schedules() {
  return sizeof(T) == 8 ? 1 : (xrplan_alt<T, N>() != xrplan<T, N>() ? 4 : 2);
}

template <typename T, int N, int SC = 0>
struct Launch {
  static constexpr int TF = TFof<N>();
  static constexpr int G = xlanes<T, N>();
  static constexpr int TB = (SC & 1) ? tbw_alt<T, N>() : tbw<T, N>();
  static constexpr int XR = (SC & 2) ? xrplan_alt<T, N>() : xrplan<T, N>();
  static constexpr int YT = TB * TF;
  static constexpr int CV = colplan<T, N>();          // plan of the y passes and the z Laplacian
  static constexpr int CB = colw<T, N>();             // their columns per workgroup
  static constexpr int CT = CB * TFof<N, CV>();       // and threads per workgroup
  // workgroups of the integrals in a launch that evaluates them (nullptr: none)
  static int fin_groups(const FinArgs* fin) {
    return fin ? std::min(kFinShares, std::max(1, (fin->nxb + 1023) / 1024)) : 0;
  }

  static unsigned xgrid(long long npairs) { return unsigned((npairs + G - 1) / G); }

  static Rows rows(const Geo& g) { return Rows{g.ny, g.pitch, g.pp, g.nq}; }
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
  static void xa(const XaArgs<T>& a, int mode, hipStream_t s) {
    // k_xrow: one row per lane, over the same rows per workgroup; k_xsums:
    // kXsLoop row groups per workgroup
    const dim3 gr(mode == XA_SUMS ? (xgrid(a.npairs) + kXsLoop - 1) / kXsLoop : xgrid(a.npairs)),
        bl(mode == XA_WITHS || mode == XA_ROWSUMS ? xrthreads<T, N, XR>() : XWG);
    if (g_recorder) {
      const int op = mode == XA_WITHF ? OP_XA_F : mode == XA_WITHS ? OP_XA_S : mode == XA_SUMS ? OP_XA_SUMS : OP_XA_RS;
      return record_launch(&launch_tab, op, a, gr, bl);
    }
    if (mode == XA_WITHF)
      hipLaunchKernelGGL((k_xa<T, N, true>), gr, bl, 0, s, a, nullptr);
    else if (mode == XA_WITHS)
      hipLaunchKernelGGL((k_xrow<T, N, false, XR>), gr, bl, 0, s, a, nullptr);
    else if (mode == XA_SUMS)
      hipLaunchKernelGGL((k_xsums<T, N>), gr, bl, 0, s, a, nullptr);
    else
      hipLaunchKernelGGL((k_xrow<T, N, false, XR, true>), gr, bl, 0, s, a, nullptr);
  }
  static void xe(const XeArgs<T>& a, hipStream_t s) {
    hipLaunchKernelGGL((k_xe<T, N>), dim3(xgrid(a.npairs)), dim3(XWG), 0, s, a);
  }
  static void ypass(cpx<T>* b0, cpx<T>* b1, int narr, int sign, const Geo& g, const cpx<T>* tw, const FinArgs* fin,
                    hipStream_t s) {
    const int nf = fin_groups(fin);
    const dim3 grid(unsigned(ygroups(g, CB)) + unsigned(nf), unsigned(narr)), bl(CT);
    YArgs<T> a{b0, b1, tw, g, fin ? *fin : FinArgs{}, nf};
    if (g_recorder) return record_launch(&launch_tab, sign < 0 ? OP_Y_FWD : OP_Y_INV, a, grid, bl);
    if (sign < 0)
      hipLaunchKernelGGL((k_ypass<T, N, -1, CB, false, CV>), grid, bl, 0, s, a, nullptr);
    else
      hipLaunchKernelGGL((k_ypass<T, N, +1, CB, false, CV>), grid, bl, 0, s, a, nullptr);
  }
  // forward y pass of two arrays, written out of place in the z kernels'
  // group-major layout (column tile ztb)
  static void ypass_gm(const cpx<T>* b0, const cpx<T>* b1, cpx<T>* o0, cpx<T>* o1, int ztb, const Geo& g,
                       const cpx<T>* tw, const FinArgs* fin, hipStream_t s) {
    const int nf = fin_groups(fin);
    const dim3 grid(unsigned(ygroups(g, CB)) + unsigned(nf), 2u), bl(CT);
    YArgs<T> a{const_cast<cpx<T>*>(b0), const_cast<cpx<T>*>(b1), tw, g, fin ? *fin : FinArgs{}, nf, o0, o1, ztb};
    if (g_recorder) return record_launch(&launch_tab, OP_Y_FWD_GM, a, grid, bl);
    hipLaunchKernelGGL((k_ypass<T, N, -1, CB, false, CV, true>), grid, bl, 0, s, a, nullptr);
  }
  static unsigned zgrid(const Geo& g) { return unsigned(zgroups(g, TB)); }
  static void zpass(cpx<T>* b, int sign, const Geo& g, const cpx<T>* tw, hipStream_t s) {
    if (sign < 0)
      hipLaunchKernelGGL((k_zpass<T, N, -1, TB>), dim3(zgrid(g)), dim3(YT), 0, s, b, g, tw);
    else
      hipLaunchKernelGGL((k_zpass<T, N, +1, TB>), dim3(zgrid(g)), dim3(YT), 0, s, b, g, tw);
  }
  static void zlap(cpx<T>* b, const Geo& g, const cpx<T>* tw, const FinArgs* fin, hipStream_t s) {
    const int nf = fin_groups(fin);
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
    hipLaunchKernelGGL((k_zprep<T, N, TB>), gr, bl, 0, s, a, nullptr);
  }
  template <int SCH, int STG, bool TAB, bool PROBE = false>
  static void zs_launch(dim3 gr, const ZsArgs<T>& a, const ZsArgs<T>* tab, hipStream_t s) {
    hipLaunchKernelGGL((k_zstage<T, N, SCH, STG, TB, TAB, zsplan<T, N>(), PROBE>), gr,
                       dim3(TB * TFof<N, zsplan<T, N>()>()), 0, s, a, tab);
  }
  template <bool TAB>
  static void zs_dispatch(int sch, int stg, dim3 gr, const ZsArgs<T>& a, const ZsArgs<T>* tab, hipStream_t s) {
    if (sch == 4) {
      switch (stg) {
        case 1: zs_launch<4, 1, TAB>(gr, a, tab, s); break;
        case 2: zs_launch<4, 2, TAB>(gr, a, tab, s); break;
        case 3: zs_launch<4, 3, TAB>(gr, a, tab, s); break;
        case 4: zs_launch<4, 4, TAB>(gr, a, tab, s); break;
        default: zs_launch<4, 4, TAB, true>(gr, a, tab, s); break;   // 5: the probe form of stage 4
      }
    } else if (stg == 1) {
      zs_launch<2, 1, TAB>(gr, a, tab, s);
    } else {
      zs_launch<2, 2, TAB>(gr, a, tab, s);
    }
  }
  static void zstage(const ZsArgs<T>& a, int sch, int stg, hipStream_t s) {
    const dim3 gr(zgrid(a.g));
    if (g_recorder)
      return record_launch(&launch_tab, stg == 5 ? OP_ZS44P : (sch == 4 ? OP_ZS41 : OP_ZS21) + (stg - 1), a, gr,
                           dim3(TB * TFof<N, zsplan<T, N>()>()));
    zs_dispatch<false>(sch, stg, gr, a, nullptr, s);
  }

  // the table form of a recorded launch, over `lanes` solves (blockIdx.z)
  static void launch_tab(int op, const void* tab, dim3 gr, dim3 bl, int lanes, hipStream_t s) {
    gr.z = unsigned(lanes);
    switch (op) {
      case OP_XA_RS:
        hipLaunchKernelGGL((k_xrow<T, N, true, XR, true>), gr, bl, 0, s, XaArgs<T>{}, (const XaArgs<T>*)tab);
        break;
      case OP_XA_F: hipLaunchKernelGGL((k_xa<T, N, true, true>), gr, bl, 0, s, XaArgs<T>{}, (const XaArgs<T>*)tab); break;
      case OP_XA_S:
        hipLaunchKernelGGL((k_xrow<T, N, true, XR>), gr, bl, 0, s, XaArgs<T>{}, (const XaArgs<T>*)tab);
        break;
      case OP_XA_SUMS:
        hipLaunchKernelGGL((k_xsums<T, N, true>), gr, bl, 0, s, XaArgs<T>{}, (const XaArgs<T>*)tab);
        break;
      case OP_Y_FWD:
        hipLaunchKernelGGL((k_ypass<T, N, -1, CB, true, CV>), gr, bl, 0, s, YArgs<T>{}, (const YArgs<T>*)tab);
        break;
      case OP_Y_INV:
        hipLaunchKernelGGL((k_ypass<T, N, +1, CB, true, CV>), gr, bl, 0, s, YArgs<T>{}, (const YArgs<T>*)tab);
        break;
      case OP_Y_FWD_GM:
        hipLaunchKernelGGL((k_ypass<T, N, -1, CB, true, CV, true>), gr, bl, 0, s, YArgs<T>{}, (const YArgs<T>*)tab);
        break;
      case OP_ZLAP:
        hipLaunchKernelGGL((k_zlap<T, N, CB, true, CV>), gr, bl, 0, s, ZlArgs<T>{}, (const ZlArgs<T>*)tab);
        break;
      case OP_ZPREP:
        hipLaunchKernelGGL((k_zprep<T, N, TB, true>), gr, bl, 0, s, ZprArgs<T>{}, (const ZprArgs<T>*)tab);
        break;
      case OP_ZS41: case OP_ZS42: case OP_ZS43: case OP_ZS44:
        zs_dispatch<true>(4, op - OP_ZS41 + 1, gr, ZsArgs<T>{}, (const ZsArgs<T>*)tab, s);
        break;
      case OP_ZS44P: zs_dispatch<true>(4, 5, gr, ZsArgs<T>{}, (const ZsArgs<T>*)tab, s); break;
      case OP_ZS21: case OP_ZS22:
        zs_dispatch<true>(2, op - OP_ZS21 + 1, gr, ZsArgs<T>{}, (const ZsArgs<T>*)tab, s);
        break;
      default: break;
    }
  }

  static SizeOps<T> table() {
    SizeOps<T> o;
    o.N = N;
    o.sched = SC;
    o.xr_passes = Plan<N, XR>::NP;
    o.xr_radix = Plan<N, XR>::R[0];
    o.xlanes = G;
    o.ytile = TB;
    o.ctile = CB;
    o.ythreads = YT;
    o.xr2c_d = &xr2c_d;
    o.xr2c = &xr2c;
    o.xc2r_d = &xc2r_d;
    o.xc2r = &xc2r;
    o.xa = &xa;
    o.xe = &xe;
    o.ypass = &ypass;
    o.ypass_gm = &ypass_gm;
    o.zpass = &zpass;
    o.zlap = &zlap;
    o.zprep = &zprep;
    o.zstage = &zstage;
    return o;
  }
};

}  // namespace vb

namespace vb {
// the launch table of schedule sc of length N (nullptr: not compiled)
template <typename T, int N>
const SizeOps<T>* ops_table(int sc) {
  if (sc < 0 || sc >= schedules<T, N>()) return nullptr;
  static const SizeOps<T> t[4] = {
      Launch<T, N, 0>::table(), Launch<T, N, (schedules<T, N>() > 1 ? 1 : 0)>::table(),
      Launch<T, N, (schedules<T, N>() > 2 ? 2 : 0)>::table(), Launch<T, N, (schedules<T, N>() > 2 ? 3 : 0)>::table()};
  return &t[sc];
}
}  // namespace vb

#define VB_DEFINE_SIZE(n)                                                                     \
  namespace vb {                                                                              \
  template <> const SizeOps<double>* ops_for<double, n>(int sc) { return ops_table<double, n>(sc); } \
  template <> const SizeOps<float>* ops_for<float, n>(int sc) { return ops_table<float, n>(sc); }    \
  template <> int sched_count<double, n>() { return schedules<double, n>(); }                 \
  template <> int sched_count<float, n>() { return schedules<float, n>(); }                   \
  }
