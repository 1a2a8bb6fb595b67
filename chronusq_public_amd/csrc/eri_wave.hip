// Dispatch of the wave-per-quartet ERI kernels (eri_wave_impl.hpp): the
// class, and the number of quartets a wave evaluates before digesting them
// (forced, e.g. by the autotuner, or the class default).  The f and g
// classes are built for FP64 with 1 and 2 quartets per wave, and 4 up to
// 216 Cartesian components.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kernels.hpp"

namespace cqx {

template <int QT>
hipError_t launch_wave_q(const EriLaunch &a, hipStream_t s);
template <int QT>
bool wave_q_built(int la, int lb, int lc, int ld, bool fp32);
template <int QT>
hipError_t launch_wave_fq(const EriLaunch &a, hipStream_t s);
template <int QT>
bool wave_fq_built(int la, int lb, int lc, int ld);
template <int QT>
hipError_t launch_wave_gq(const EriLaunch &a, hipStream_t s);
template <int QT>
bool wave_gq_built(int la, int lb, int lc, int ld);

static int max_l(int la, int lb, int lc, int ld) { return std::max(std::max(la, lb), std::max(lc, ld)); }

static bool fg_built(int la, int lb, int lc, int ld, int qt) {
  const bool g = max_l(la, lb, lc, ld) == 4;
  if (qt == 1) return g ? wave_gq_built<1>(la, lb, lc, ld) : wave_fq_built<1>(la, lb, lc, ld);
  if (qt == 2) return g ? wave_gq_built<2>(la, lb, lc, ld) : wave_fq_built<2>(la, lb, lc, ld);
  if (qt == 4) return g ? wave_gq_built<4>(la, lb, lc, ld) : wave_fq_built<4>(la, lb, lc, ld);
  return false;
}

bool eri_wave_available(int la, int lb, int lc, int ld) {
  if (max_l(la, lb, lc, ld) >= 3) return fg_built(la, lb, lc, ld, 1);
  return wave_q_built<1>(la, lb, lc, ld, false);
}

bool eri_wave_qt_built(int la, int lb, int lc, int ld, int fp32, int qt) {
  if (max_l(la, lb, lc, ld) >= 3) return !fp32 && fg_built(la, lb, lc, ld, qt);
  switch (qt) {
    case 1: return wave_q_built<1>(la, lb, lc, ld, fp32 != 0);
    case 2: return wave_q_built<2>(la, lb, lc, ld, fp32 != 0);
    case 4: return wave_q_built<4>(la, lb, lc, ld, fp32 != 0);
    case 8: return wave_q_built<8>(la, lb, lc, ld, fp32 != 0);
    default: return false;
  }
}

// 4 quartets per wave for Cartesian quartets of at most 108 components, 2
// up to 216, else 1 (measured on the GHF and C40H82 builds)
int eri_wave_default_qt(int la, int lb, int lc, int ld) {
  auto nc = [](int l) { return (l + 1) * (l + 2) / 2; };
  const int n = nc(la) * nc(lb) * nc(lc) * nc(ld);
  return n <= 108 ? 4 : n <= 216 ? 2 : 1;
}

hipError_t launch_eri_wave(const EriLaunch &a, hipStream_t s) {
  if (a.nq <= 0) return hipSuccess;
  if (a.store_mode) return hipErrorNotSupported;
  int qt = a.force_qt;
  if (!eri_wave_qt_built(a.la, a.lb, a.lc, a.ld, a.fp32, qt)) qt = eri_wave_default_qt(a.la, a.lb, a.lc, a.ld);
  if (max_l(a.la, a.lb, a.lc, a.ld) >= 3) {
    if (!eri_wave_qt_built(a.la, a.lb, a.lc, a.ld, a.fp32, qt)) qt = 1;
    if (max_l(a.la, a.lb, a.lc, a.ld) == 4)
      return qt == 4 ? launch_wave_gq<4>(a, s) : qt == 2 ? launch_wave_gq<2>(a, s) : launch_wave_gq<1>(a, s);
    return qt == 4 ? launch_wave_fq<4>(a, s) : qt == 2 ? launch_wave_fq<2>(a, s) : launch_wave_fq<1>(a, s);
  }
  switch (qt) {
    case 1: return launch_wave_q<1>(a, s);
    case 2: return launch_wave_q<2>(a, s);
    case 4: return launch_wave_q<4>(a, s);
    default: return launch_wave_q<8>(a, s);
  }
}

}  // namespace cqx
