// Launch table for transform length 384 (x/y kernels when nx = 384, z kernels when nz = 384).
#include "vb_size_impl.hpp"
VB_DEFINE_SIZE(384)
