// vb_multi.hip -- the ETD stage combines of solves of several grid sizes in
// one launch (the batched form's shared stage launches).
//
// Workgroup b of a launch runs column group b - first[s] of solve s, the last
// solve with first[s] <= b; the solve's grid length selects the stage body
// compiled for it (the same body the single-size kernels run, so a solve
// computes exactly what it computes alone).  Every size carried here runs its
// default column tile in 256-thread workgroups.  192 is left out: its bodies
// need 205-226 registers, and one launch takes the largest register count of
// the bodies it carries (two waves per SIMD instead of three for the other
// sizes' stages 1-3).
#include "vb_multi.hpp"

#include "vb_kernels.hpp"

namespace vb {

namespace {

constexpr int kMultiThreads = 256;

template <int N>
constexpr bool multi_size_ok() {
  return tbw<float, N>() * TFof<N, zsplan<float, N>()>() == kMultiThreads;
}
static_assert(multi_size_ok<32>() && multi_size_ok<64>() && multi_size_ok<128>() && multi_size_ok<256>(),
              "every size of the shared stage launch runs 256-thread workgroups");

template <int SCH, int STG, int N>
constexpr int lds_of() {
  return zs_lds_bytes<float, N, SCH, STG, tbw<float, N>(), zsplan<float, N>()>();
}
template <int SCH, int STG>
constexpr int multi_lds() {
  constexpr int a = lds_of<SCH, STG, 32>(), b = lds_of<SCH, STG, 64>(), c = lds_of<SCH, STG, 128>(),
                d = lds_of<SCH, STG, 256>();
  constexpr int ab = a > b ? a : b, cd = c > d ? c : d;
  return ab > cd ? ab : cd;
}

template <int SCH, int STG, bool PROBE>
__global__ void __launch_bounds__(kMultiThreads) k_zstage_multi(ZsMultiArgs m) {
  // the solve of this workgroup: first[] is ascending, the search uniform
  const int b = int(blockIdx.x);
  int lo = 0, hi = m.nsolve - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (m.first[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  const ZsArgs<float> a = static_cast<const ZsArgs<float>*>(m.args)[lo];
  const int tile = b - m.first[lo];
  __shared__ __align__(16) unsigned char smraw[multi_lds<SCH, STG>()];
  __shared__ double red[kMultiThreads / 64];
  switch (m.n[lo]) {
    case 32: zstage_body<float, 32, SCH, STG, tbw<float, 32>(), zsplan<float, 32>(), PROBE>(a, tile, smraw, red); break;
    case 64: zstage_body<float, 64, SCH, STG, tbw<float, 64>(), zsplan<float, 64>(), PROBE>(a, tile, smraw, red); break;
    case 128:
      zstage_body<float, 128, SCH, STG, tbw<float, 128>(), zsplan<float, 128>(), PROBE>(a, tile, smraw, red);
      break;
    case 256:
      zstage_body<float, 256, SCH, STG, tbw<float, 256>(), zsplan<float, 256>(), PROBE>(a, tile, smraw, red);
      break;
    default: break;
  }
}

}  // namespace

bool zstage_multi_size(int n, int ztile) {
  switch (n) {
    case 32: return ztile == tbw<float, 32>();
    case 64: return ztile == tbw<float, 64>();
    case 128: return ztile == tbw<float, 128>();
    case 256: return ztile == tbw<float, 256>();
    default: return false;
  }
}

bool zstage_multi_op(int op) { return op >= OP_ZS41 && op <= OP_ZS44 ? true : op == OP_ZS44P; }

void launch_zstage_multi(int op, const ZsMultiArgs& m, unsigned nblocks, hipStream_t s) {
  const dim3 gr(nblocks), bl(kMultiThreads);
  switch (op) {
    case OP_ZS41: hipLaunchKernelGGL((k_zstage_multi<4, 1, false>), gr, bl, 0, s, m); break;
    case OP_ZS42: hipLaunchKernelGGL((k_zstage_multi<4, 2, false>), gr, bl, 0, s, m); break;
    case OP_ZS43: hipLaunchKernelGGL((k_zstage_multi<4, 3, false>), gr, bl, 0, s, m); break;
    case OP_ZS44: hipLaunchKernelGGL((k_zstage_multi<4, 4, false>), gr, bl, 0, s, m); break;
    case OP_ZS44P: hipLaunchKernelGGL((k_zstage_multi<4, 4, true>), gr, bl, 0, s, m); break;
    default: break;
  }
}

}  // namespace vb
