// Sphere geometry on the reference's world box, shared by the geometry
// kernels and their reverse-mode products.
#pragma once

#include "lbm_physics.hpp"

namespace lbm {

template <typename C>
__device__ __forceinline__ void world_xyz(int x, int y, int z, int nx, int ny, int nz, C& wx, C& wy, C& wz) {
  // node (i,j,k) -> ((i+1/2)/(nx/8) - 4, (j+1/2)/(ny/4) - 2, (k+1/2)/(nz/4) - 2)
  // (primitives.py:26-53): same operation order and precision as the reference
  wx = (C(x) + C(0.5)) / C(nx / 8.0) - C(4.0);
  wy = (C(y) + C(0.5)) / C(ny / 4.0) - C(2.0);
  wz = (C(z) + C(0.5)) / C(nz / 4.0) - C(2.0);
}

__device__ __forceinline__ bool link_from(const int32_t* solid, int x, int y, int z, int i, int nx, int ny,
                                          int nz) {
  // fluid cell whose in-bounds neighbour along e_i is a body cell
  // (primitives.py:95-113, bouzidi.py:46-66, smooth.py:168-185)
  const int bx = x + EX[i], by = y + EY[i], bz = z + EZ[i];
  if (bx < 0 || bx >= nx || by < 0 || by >= ny || bz < 0 || bz >= nz) return false;
  return solid[(size_t(z) * ny + y) * nx + x] == 0 && solid[(size_t(bz) * ny + by) * nx + bx] == 1;
}

}  // namespace lbm
