// vb_measure.hip -- the streaming ceiling of the device: a plain copy of a
// large buffer with 16-byte loads and stores, timed with HIP events.  It is
// the practical upper bound of an HBM-bound kernel on this GPU (read + write
// bytes per second), measured beside the solver's kernels.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "vb_host.hpp"

namespace {

typedef unsigned int v4u __attribute__((ext_vector_type(4)));

// grid-stride copy, UN 16-byte elements in flight per thread; NT: nontemporal
template <int UN, bool NT>
__global__ void __launch_bounds__(256) k_copy16(const v4u* __restrict__ src, v4u* __restrict__ dst, long long n) {
  const long long stride = (long long)gridDim.x * 256;
  long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  for (; i + (UN - 1) * stride < n; i += UN * stride) {
    v4u v[UN];
#pragma unroll
    for (int u = 0; u < UN; ++u) v[u] = NT ? __builtin_nontemporal_load(src + i + u * stride) : src[i + u * stride];
#pragma unroll
    for (int u = 0; u < UN; ++u) {
      if (NT) __builtin_nontemporal_store(v[u], dst + i + u * stride);
      else dst[i + u * stride] = v[u];
    }
  }
  for (; i < n; i += stride) dst[i] = NT ? __builtin_nontemporal_load(src + i) : src[i];
}

struct Variant {
  void (*launch)(const v4u*, v4u*, long long, int, hipStream_t);
  const char* name;
};

template <int UN, bool NT>
void launch_copy(const v4u* s, v4u* d, long long n, int blocks, hipStream_t st) {
  hipLaunchKernelGGL((k_copy16<UN, NT>), dim3(blocks), dim3(256), 0, st, s, d, n);
}

}  // namespace

extern "C" {

int vb_copy_bandwidth(int device, size_t bytes, int repeats, double* best_GBps, double* variant_GBps, int cap) {
  if (!best_GBps || bytes < ((size_t)1 << 20) || repeats <= 0) return VB_ERR_PARAM;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return VB_ERR_NO_DEVICE;
  const Variant vars[] = {{&launch_copy<1, false>, "1x16B"}, {&launch_copy<4, false>, "4x16B"},
                          {&launch_copy<1, true>, "1x16B nt"}, {&launch_copy<4, true>, "4x16B nt"}};
  const int nv = int(sizeof(vars) / sizeof(vars[0]));
  v4u* a = nullptr;
  v4u* b = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = VB_OK;
  try {
    VB_HIP(hipSetDevice(device));
    const long long n = (long long)(bytes / 16);
    VB_HIP(hipMalloc(&a, (size_t)n * 16));
    VB_HIP(hipMalloc(&b, (size_t)n * 16));
    VB_HIP(hipMemset(a, 1, (size_t)n * 16));
    VB_HIP(hipEventCreate(&e0));
    VB_HIP(hipEventCreate(&e1));
    int ncu = 0;
    VB_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device));
    double best = 0.0;
    for (int v = 0; v < nv; ++v) {
      // enough workgroups to cover every CU many times over, within the array
      const int blocks = (int)std::min<long long>((long long)ncu * 32, (n + 255) / 256);
      for (int w = 0; w < 2; ++w) vars[v].launch(a, b, n, blocks, nullptr);
      VB_HIP(hipGetLastError());
      VB_HIP(hipEventRecord(e0, nullptr));
      for (int r = 0; r < repeats; ++r) vars[v].launch(a, b, n, blocks, nullptr);
      VB_HIP(hipEventRecord(e1, nullptr));
      VB_HIP(hipEventSynchronize(e1));
      float ms = 0.f;
      VB_HIP(hipEventElapsedTime(&ms, e0, e1));
      const double gbps = 2.0 * double(n) * 16.0 * repeats / (double(ms) * 1e-3) / 1e9;
      if (variant_GBps && v < cap) variant_GBps[v] = gbps;
      best = std::max(best, gbps);
    }
    *best_GBps = best;
  } catch (const vb::HipError& e) {
    vb::set_last_error(std::string(hipGetErrorString(e.err)) + " in " + e.where);
    rc = e.err == hipErrorOutOfMemory ? VB_ERR_OOM : VB_ERR_HIP;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  return rc;
}

}  // extern "C"
