// Launch table for transform length 64 (x/y kernels when nx = 64, z kernels when nz = 64).
#include "vb_size_impl.hpp"
VB_DEFINE_SIZE(64)
