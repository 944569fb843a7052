// Instantiation of the 3-D element kernels for element orders N = 2, 3, 4, 5.
#include "gf_k3d.hpp"

namespace gfs {
GF_K3D_INSTANTIATE(2)
GF_K3D_INSTANTIATE(3)
GF_K3D_INSTANTIATE(4)
GF_K3D_INSTANTIATE(5)
}  // namespace gfs
