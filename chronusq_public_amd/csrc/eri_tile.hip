// Dispatch of the tile ERI kernels (eri_tile_impl.hpp).
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "tile_layout.hpp"

namespace cqx {

template <int G, class R>
hipError_t launch_tile_g(const TileEriLaunch &a, hipStream_t s);

namespace {

template <class R>
hipError_t launch_tile_r(const TileEriLaunch &a, int G, hipStream_t s) {
  switch (G) {
    case 1: return launch_tile_g<1, R>(a, s);
    case 2: return launch_tile_g<2, R>(a, s);
    case 4: return launch_tile_g<4, R>(a, s);
    case 8: return launch_tile_g<8, R>(a, s);
    default: return launch_tile_g<16, R>(a, s);
  }
}

int ncart_h(int l) { return (l + 1) * (l + 2) / 2; }

}  // namespace

hipError_t launch_eri_tile(const TileEriLaunch &a, hipStream_t s) {
  if (a.ntile <= 0) return hipSuccess;
  if (!eri_fast_available(a.la, a.lb, a.lc, a.ld)) return hipErrorNotSupported;
  // lanes per quartet as for the lane-group kernels; one lane per quartet
  // only where a round's staged quartets stay small
  const int nc = ncart_h(a.la) * ncart_h(a.lb) * ncart_h(a.lc) * ncart_h(a.ld);
  int G = 1;
  while (G < 16 && 4.0 * G < a.mean_prims) G *= 2;
  if (a.force_g == 1 || a.force_g == 2 || a.force_g == 4 || a.force_g == 8 || a.force_g == 16) G = a.force_g;
  if (nc > 9 && G < 2) G = 2;
  return a.fp32 ? launch_tile_r<float>(a, G, s) : launch_tile_r<double>(a, G, s);
}

int tile_acc_doubles(int la, int lb, int lc, int ld, const PlaneSet &ps) {
  if (!eri_fast_available(la, lb, lc, ld)) return 0;
  const int sj = ps.kcount[kJHer] ? ps.kps[kJHer] : 0, sk = ps.kcount[kKHer] ? ps.kps[kKHer] : 0;
  return tile_acc_layout(ncart_h(la), ncart_h(lb), ncart_h(lc), ncart_h(ld), sj, sk).total;
}

}  // namespace cqx
