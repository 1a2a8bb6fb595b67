// Wave-per-quartet ERI kernels instantiated for 2 quartets per wave,
// FP64 and FP32 evaluation.
#include "eri_wave_impl.hpp"

namespace cqx {
template hipError_t launch_wave_q<2>(const EriLaunch &a, hipStream_t s);
template bool wave_q_built<2>(int la, int lb, int lc, int ld, bool fp32);
}  // namespace cqx
