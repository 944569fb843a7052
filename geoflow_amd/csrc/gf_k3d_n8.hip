// Instantiation of the 3-D element kernels for element orders N = 8.
#include "gf_k3d.hpp"

namespace gfs {
GF_K3D_INSTANTIATE(8)
}  // namespace gfs
