// LDS layout of a tile's shared J / K sums (eri_tile_impl.hpp), shared by
// the kernels and the host code that sizes their launches.
#pragma once

#include <hip/hip_runtime.h>

#include "device_types.hpp"

namespace cqx {

// Per-wave LDS accumulators of a tile, in doubles: [local][target][plane]
// arrays Jab (per bra), Jdc (per ket), Kac (one), Kda (per ket) and Kbc
// (per bra) for Cartesian shell sizes na, nb, nc, nd; sj / sk are the
// plane strides of the Hermitian J and K kinds (0 when the kind is absent).
struct TileAccLayout {
  int jab, jdc, kac, kda, kbc, total;
};
__host__ __device__ inline TileAccLayout tile_acc_layout(int na, int nb, int nc, int nd, int sj, int sk) {
  TileAccLayout L;
  L.jab = 0;
  L.jdc = L.jab + kTileDim * na * nb * sj;
  L.kac = L.jdc + kTileDim * nd * nc * sj;
  L.kda = L.kac + na * nc * sk;
  L.kbc = L.kda + kTileDim * nd * na * sk;
  L.total = L.kbc + kTileDim * nb * nc * sk;
  return L;
}

}  // namespace cqx
