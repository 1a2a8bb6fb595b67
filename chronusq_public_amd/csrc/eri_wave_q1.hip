// Wave-per-quartet ERI kernels instantiated for 1 quartet per wave,
// FP64 and FP32 evaluation.
#include "eri_wave_impl.hpp"

namespace cqx {
template hipError_t launch_wave_q<1>(const EriLaunch &a, hipStream_t s);
template bool wave_q_built<1>(int la, int lb, int lc, int ld, bool fp32);
}  // namespace cqx
