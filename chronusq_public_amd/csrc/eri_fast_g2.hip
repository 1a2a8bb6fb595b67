// Lane-group ERI kernels instantiated for G = 2 lanes per shell quartet.
#include "eri_fast_impl.hpp"

namespace cqx {
template hipError_t launch_fast_g<2>(const EriLaunch &a, hipStream_t s);
}  // namespace cqx
