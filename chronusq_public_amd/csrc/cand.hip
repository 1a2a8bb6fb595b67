// Candidate kets per bra pair on the device: the Schwarz pre-cut prefix of
// the bra's own molecule's kets (kets sorted by Q descending), then a scan
// into candidate offsets -- the same arithmetic the screening of the
// reference loop implies before its per-quartet test (direct.hpp:1689-1789).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "kernels.hpp"

namespace cqx {

namespace {

__global__ void cand_count_kernel(CandLaunch a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.nbra) return;
  const PairRec &P = a.bra[i];
  const int m = P.mol;
  const int lo = a.seg[m], hi = a.seg[m + 1];
  long long jmax = hi - lo;
  if (a.screen && jmax > 0) {
    const double bmax = a.molmax[m];
    if (bmax * P.Q * a.ket[lo].Q < a.thresh) {
      jmax = 0;
    } else {
      // largest prefix with Q_i Q_j bmax >= thresh
      const double need = a.thresh / (bmax * P.Q);
      int x = lo, y = hi;
      while (x < y) {
        const int mid = (x + y) / 2;
        if (a.ket[mid].Q >= need) x = mid + 1; else y = mid;
      }
      jmax = x - lo;
    }
  }
  if (a.same_class) jmax = min(jmax, (long long)(a.bra_base + i - lo + 1));
  a.off[i + 1] = jmax > 0 ? jmax : 0;
  a.klo[i] = lo;
  if (i == 0) a.off[0] = 0;
}

}  // namespace

hipError_t launch_candidates(const CandLaunch &a, void *tmp, size_t &tmp_bytes, hipStream_t s) {
  if (a.nbra <= 0) return hipSuccess;
  if (!tmp) {  // size query
    return hipcub::DeviceScan::InclusiveSum(nullptr, tmp_bytes, a.off + 1, a.off + 1, a.nbra, s);
  }
  hipLaunchKernelGGL(cand_count_kernel, dim3((a.nbra + 255) / 256), dim3(256), 0, s, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return hipcub::DeviceScan::InclusiveSum(tmp, tmp_bytes, a.off + 1, a.off + 1, a.nbra, s);
}

}  // namespace cqx
