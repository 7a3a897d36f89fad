// Launch table for transform length 32 (x/y kernels when nx = 32, z kernels when nz = 32).
#include "vb_size_impl.hpp"
VB_DEFINE_SIZE(32)
