// Lane-group ERI kernels instantiated for G = 1 lanes per shell quartet,
// FP64 and FP32 evaluation.
#include "eri_fast_impl.hpp"

namespace cqx {
template hipError_t launch_fast_g<1, double>(const EriLaunch &a, hipStream_t s);
template hipError_t launch_fast_g<1, float>(const EriLaunch &a, hipStream_t s);
}  // namespace cqx
