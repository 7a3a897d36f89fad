// vb_ops.hpp -- per-transform-length launch table.  Each supported length N
// is compiled in its own translation unit (vb_n<N>.hip) so the build runs in
// parallel; the host pipeline picks the table of nx (x and y kernels) and the
// table of nz (z kernels) at solver creation.
#pragma once
#include <hip/hip_runtime.h>

#include "vb_device.hpp"

namespace vb {

template <typename T> struct XaArgs;
template <typename T> struct XbArgs;
template <typename T> struct XeArgs;
template <typename T> struct ZsArgs;

template <typename T>
struct SizeOps {
  int N;
  int xlanes;      // row pairs per x workgroup
  int ytile;       // columns per y / z workgroup
  int ythreads;    // threads per y / z workgroup
  void (*xr2c_d)(const double* in, cpx<T>* out, long long npairs, const cpx<T>* tw, double scl, const Geo& g,
                 hipStream_t s);
  void (*xr2c)(const T* in, cpx<T>* out, long long npairs, const cpx<T>* tw, double scl, const Geo& g, hipStream_t s);
  void (*xc2r_d)(const cpx<T>* in, double* out, long long npairs, const cpx<T>* tw, const Geo& g, hipStream_t s);
  void (*xc2r)(const cpx<T>* in, T* out, long long npairs, const cpx<T>* tw, const Geo& g, hipStream_t s);
  void (*xa)(const XaArgs<T>& a, bool withf, hipStream_t s);
  void (*xb)(const XbArgs<T>& a, hipStream_t s);
  void (*xe)(const XeArgs<T>& a, hipStream_t s);
  void (*ypass)(cpx<T>* b0, cpx<T>* b1, int narr, int sign, const Geo& g, const cpx<T>* tw, hipStream_t s);
  void (*zpass)(cpx<T>* b, int sign, const Geo& g, const cpx<T>* tw, hipStream_t s);
  void (*zlap)(cpx<T>* b, const Geo& g, const cpx<T>* tw, hipStream_t s);
  void (*zprep)(const cpx<double>* vh, cpx<T>* P, cpx<T>* Q, double* part, double* dc, const Geo& g, const cpx<T>* tw,
                hipStream_t s);
  void (*zstage)(const ZsArgs<T>& a, int sch, int stg, hipStream_t s);
};

template <typename T, int N> const SizeOps<T>* ops_for();

#define VB_SIZE_LIST(X) X(16) X(32) X(48) X(64) X(96) X(128) X(192) X(256) X(384) X(512)

#define VB_DECL(n)                                              \
  template <> const SizeOps<double>* ops_for<double, n>();      \
  template <> const SizeOps<float>* ops_for<float, n>();
VB_SIZE_LIST(VB_DECL)
#undef VB_DECL

template <typename T>
inline const SizeOps<T>* get_ops(int N) {
  switch (N) {
#define VB_CASE(n) case n: return ops_for<T, n>();
    VB_SIZE_LIST(VB_CASE)
#undef VB_CASE
    default: return nullptr;
  }
}

inline bool size_supported(int N) {
  switch (N) {
#define VB_CASE(n) case n: return true;
    VB_SIZE_LIST(VB_CASE)
#undef VB_CASE
    default: return false;
  }
}

}  // namespace vb
