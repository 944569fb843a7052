// Instantiation of the 3-D element kernels for element orders N = 6, 7.
#include "gf_k3d.hpp"

namespace gfs {
GF_K3D_INSTANTIATE(6)
GF_K3D_INSTANTIATE(7)
}  // namespace gfs
