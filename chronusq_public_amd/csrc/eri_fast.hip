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

int fast_class_index(int la, int lb, int lc, int ld) {
  const int q[8][4] = {{0, 0, 0, 0}, {1, 0, 0, 0}, {1, 0, 1, 0}, {1, 1, 0, 0},
                       {1, 1, 1, 0}, {2, 0, 0, 0}, {2, 0, 1, 0}, {2, 1, 0, 0}};
  for (int k = 0; k < 8; ++k)
    if (q[k][0] == la && q[k][1] == lb && q[k][2] == lc && q[k][3] == ld) return k;
  return -1;
}

bool eri_fast_available(int la, int lb, int lc, int ld) { return fast_class_index(la, lb, lc, ld) >= 0; }

bool eri_lane_group_dual(int la, int lb, int lc, int ld) {
  const int q[4][4] = {{2, 0, 2, 0}, {2, 2, 0, 0}, {2, 1, 1, 0}, {2, 0, 1, 1}};
  for (int k = 0; k < 4; ++k)
    if (q[k][0] == la && q[k][1] == lb && q[k][2] == lc && q[k][3] == ld) return true;
  return false;
}

bool eri_lane_group_built(int la, int lb, int lc, int ld) {
  return eri_fast_available(la, lb, lc, ld) || eri_lane_group_dual(la, lb, lc, ld);
}

hipError_t launch_eri_fast(const EriLaunch &a, hipStream_t s) {
  if (a.nq <= 0) return hipSuccess;
  if (a.store_mode) return hipErrorNotSupported;
  // lanes per quartet: about a quarter of the mean primitive quartets per
  // shell quartet (a power of two, 1..16), at least 2 for the larger
  // classes so that a wave's LDS tile stays small
  const int nc = ((a.la + 1) * (a.la + 2) / 2) * ((a.lb + 1) * (a.lb + 2) / 2) *
                 ((a.lc + 1) * (a.lc + 2) / 2) * ((a.ld + 1) * (a.ld + 2) / 2);
  int G = 1;
  while (G < 16 && 4.0 * G < a.mean_prims) G *= 2;
  if (nc > 9 && G < 2) G = 2;
  if (a.force_g == 1 || a.force_g == 2 || a.force_g == 4 || a.force_g == 8 || a.force_g == 16) G = a.force_g;
  return a.fp32 ? launch_fast_r<float>(a, G, s) : launch_fast_r<double>(a, G, s);
}

}  // namespace cqx
