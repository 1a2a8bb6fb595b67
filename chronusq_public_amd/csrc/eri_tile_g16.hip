// Tile ERI kernels instantiated for G = 16 lanes per shell quartet, FP64 and
// FP32 evaluation.
#include "eri_tile_impl.hpp"

namespace cqx {
template hipError_t launch_tile_g<16, double>(const TileEriLaunch &a, hipStream_t s);
template hipError_t launch_tile_g<16, float>(const TileEriLaunch &a, hipStream_t s);
}  // namespace cqx
