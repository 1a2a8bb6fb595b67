// Dispatch of the lane-group ERI kernels (eri_fast_impl.hpp).
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace cqx {

template <int G, class R>
hipError_t launch_fast_g(const EriLaunch &a, hipStream_t s);

template <class R>
hipError_t launch_fast_r(const EriLaunch &a, int G, hipStream_t s) {
  switch (G) {
    case 1: return launch_fast_g<1, R>(a, s);
    case 2: return launch_fast_g<2, R>(a, s);
    case 4: return launch_fast_g<4, R>(a, s);
    case 8: return launch_fast_g<8, R>(a, s);
    default: return launch_fast_g<16, R>(a, s);
  }
}

bool eri_fast_available(int la, int lb, int lc, int ld) {
  const int q[8][4] = {{0, 0, 0, 0}, {1, 0, 0, 0}, {1, 0, 1, 0}, {1, 1, 0, 0},
                       {1, 1, 1, 0}, {2, 0, 0, 0}, {2, 0, 1, 0}, {2, 1, 0, 0}};
  for (auto &c : q)
    if (c[0] == la && c[1] == lb && c[2] == lc && c[3] == ld) return true;
  return false;
}

hipError_t launch_eri_fast(const EriLaunch &a, hipStream_t s) {
  if (a.nq <= 0) return hipSuccess;
  if (a.store_mode) return hipErrorNotSupported;
  // lanes per quartet: the plane count rounded up to a power of two (<= 16)
  int G = 1;
  while (G < a.planes.n && G < 16) G *= 2;
  return a.fp32 ? launch_fast_r<float>(a, G, s) : launch_fast_r<double>(a, G, s);
}

}  // namespace cqx
