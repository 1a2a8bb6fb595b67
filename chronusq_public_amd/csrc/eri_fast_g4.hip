// Lane-group ERI kernels instantiated for G = 4 lanes per shell quartet.
#include "eri_fast_impl.hpp"

namespace cqx {
template hipError_t launch_fast_g<4>(const EriLaunch &a, hipStream_t s);
}  // namespace cqx
