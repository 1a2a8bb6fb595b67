// Lane-group ERI kernels instantiated for G = 16 lanes per shell quartet.
#include "eri_fast_impl.hpp"

namespace cqx {
template hipError_t launch_fast_g<16>(const EriLaunch &a, hipStream_t s);
}  // namespace cqx
