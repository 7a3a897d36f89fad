// Launch table for transform length 96 (x/y kernels when nx = 96, z kernels when nz = 96).
#include "vb_size_impl.hpp"
VB_DEFINE_SIZE(96)
