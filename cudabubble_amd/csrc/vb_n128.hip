// Launch table for transform length 128 (x/y kernels when nx = 128, z kernels when nz = 128).
#include "vb_size_impl.hpp"
VB_DEFINE_SIZE(128)
