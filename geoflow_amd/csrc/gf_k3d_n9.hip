// Instantiation of the 3-D element kernels for element orders N = 9.
#include "gf_k3d.hpp"

namespace gfs {
GF_K3D_INSTANTIATE(9)
}  // namespace gfs
