// Wave-per-quartet ERI kernels of the g classes of at most 216 Cartesian
// components instantiated for 4 quartets per wave, FP64 evaluation.
#include "eri_wave_impl.hpp"

namespace cqx {
template hipError_t launch_wave_gq<4>(const EriLaunch &a, hipStream_t s);
template bool wave_gq_built<4>(int la, int lb, int lc, int ld);
}  // namespace cqx
