// vb_multi.hpp -- host side of the shared stage launches over solves of
// several grid sizes (vb_multi.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "vb_ops.hpp"

namespace vb {

// One shared stage launch: solve s (of nsolve, FP32-mixed) runs column groups
// [first[s], first[s + 1]) of the launch with its argument struct args[s]
// (a ZsArgs<float>) and its grid length n[s].  Device pointers.
struct ZsMultiArgs {
  const void* args;
  const int* first;
  const int* n;
  int nsolve, pad;
};

// whether solves of length n whose z kernels use column tile ztile can join
// a shared stage launch (FP32-mixed, ETDRK4)
bool zstage_multi_size(int n, int ztile);
// whether a recorded pipeline op is an ETDRK4 stage combine a shared launch carries
bool zstage_multi_op(int op);
void launch_zstage_multi(int op, const ZsMultiArgs& m, unsigned nblocks, hipStream_t s);

}  // namespace vb
