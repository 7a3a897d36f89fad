// Launch table for transform length 512 (x/y kernels when nx = 512, z kernels when nz = 512).
#include "vb_size_impl.hpp"
VB_DEFINE_SIZE(512)
