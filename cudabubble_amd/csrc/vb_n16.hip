// Launch table for transform length 16 (x/y kernels when nx = 16, z kernels when nz = 16).
#include "vb_size_impl.hpp"
VB_DEFINE_SIZE(16)
