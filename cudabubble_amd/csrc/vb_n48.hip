// Launch table for transform length 48 (x/y kernels when nx = 48, z kernels when nz = 48).
#include "vb_size_impl.hpp"
VB_DEFINE_SIZE(48)
