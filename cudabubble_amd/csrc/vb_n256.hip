// Launch table for transform length 256 (x/y kernels when nx = 256, z kernels when nz = 256).
#include "vb_size_impl.hpp"
VB_DEFINE_SIZE(256)
