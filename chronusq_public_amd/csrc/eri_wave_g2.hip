// Wave-per-quartet ERI kernels of the g classes instantiated for 2 quartets
// per wave, FP64 evaluation.
#include "eri_wave_impl.hpp"

namespace cqx {
template hipError_t launch_wave_gq<2>(const EriLaunch &a, hipStream_t s);
template bool wave_gq_built<2>(int la, int lb, int lc, int ld);
}  // namespace cqx
