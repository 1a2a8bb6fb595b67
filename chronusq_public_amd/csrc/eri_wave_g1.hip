// Wave-per-quartet ERI kernels of the g classes instantiated for 1 quartet
// per wave, FP64 evaluation.
#include "eri_wave_impl.hpp"

namespace cqx {
template hipError_t launch_wave_gq<1>(const EriLaunch &a, hipStream_t s);
template bool wave_gq_built<1>(int la, int lb, int lc, int ld);
}  // namespace cqx
