// Launch table for transform length 192 (x/y kernels when nx = 192, z kernels when nz = 192).
#include "vb_size_impl.hpp"
VB_DEFINE_SIZE(192)
