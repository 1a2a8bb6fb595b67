// Wave-cooperative J/K digestion of one contracted Cartesian shell quartet
// held in LDS, for any class (sizes at run time).
//
// The updates are those of the reference (direct.hpp:1858-2002) written in
// the kernel's canonical labels a, b, c, d (roles 0..3).  Each update
//   A(ar, ac) += coef * sum_{xr, xc} I(a,b,c,d) X(xr, xc)
// is spread over the lanes with the plane index running fastest: planes are
// interleaved in memory, so lanes that share a target element read and add
// consecutive doubles and their loads / atomics coalesce.
#pragma once

#include <hip/hip_runtime.h>

#include "device_types.hpp"

namespace cqx {

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// (ar, ac, xr, xc) of every update, per plane kind; coefficient per kind
// in digest_wave (1 for Hermitian J, sign/2 for Hermitian K, 1/2 otherwise)
struct DigestTerms {
  int n[4];
  signed char t[4][8][4];
};

__device__ __constant__ const DigestTerms kDigestTerms = {
    {2, 4, 8, 8},
    {{{0, 1, 3, 2}, {3, 2, 0, 1}},
     {{0, 2, 3, 1}, {3, 1, 0, 2}, {3, 0, 1, 2}, {1, 2, 3, 0}},
     {{0, 1, 3, 2}, {0, 1, 2, 3}, {1, 0, 3, 2}, {1, 0, 2, 3},
      {2, 3, 1, 0}, {2, 3, 0, 1}, {3, 2, 1, 0}, {3, 2, 0, 1}},
     {{2, 0, 3, 1}, {3, 1, 2, 0}, {3, 0, 2, 1}, {2, 1, 3, 0},
      {0, 2, 1, 3}, {1, 3, 0, 2}, {0, 3, 1, 2}, {1, 2, 0, 3}}}};

__device__ inline void digest_wave(const double *__restrict__ sI, const int n[4], const int bf[4],
                                   const PlaneSet &ps, unsigned mask, int lane, int W) {
  const int N = ps.nbc, S = ps.ps;
  for (int kind = 0; kind < 4; ++kind) {
    const int p0 = ps.kstart[kind], np = ps.kcount[kind];
    if (np == 0) continue;
    const double base_coef = kind == kJHer ? 1.0 : 0.5;
    for (int term = 0; term < kDigestTerms.n[kind]; ++term) {
      const int ar = kDigestTerms.t[kind][term][0], ac = kDigestTerms.t[kind][term][1];
      const int xr = kDigestTerms.t[kind][term][2], xc = kDigestTerms.t[kind][term][3];
      const int nt = n[ar] * n[ac], nsum = n[xr] * n[xc];
      for (int t = lane; t < nt * np; t += W) {
        const int pl = p0 + t % np, tgt = t / np;
        const Plane &P = ps.p[pl];
        if (!((mask >> P.mat) & 1u)) continue;
        int idx[4];
        idx[ar] = tgt / n[ac];
        idx[ac] = tgt % n[ac];
        double sum = 0.0;
        for (int u = 0; u < nsum; ++u) {
          idx[xr] = u / n[xc];
          idx[xc] = u % n[xc];
          const double I = sI[((idx[0] * n[1] + idx[1]) * n[2] + idx[2]) * n[3] + idx[3]];
          sum = fma(I, P.X[((bf[xr] + idx[xr]) + (size_t)(bf[xc] + idx[xc]) * N) * S], sum);
        }
        const double coef = kind == kKHer ? base_coef * P.sign : base_coef;
        unsafeAtomicAdd(&P.A[((bf[ar] + idx[ar]) + (size_t)(bf[ac] + idx[ac]) * N) * S], coef * sum);
      }
    }
  }
}

}  // namespace cqx
