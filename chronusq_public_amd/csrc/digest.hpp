// Wave-cooperative J/K digestion of one contracted Cartesian shell quartet
// held in LDS, for any class (sizes at run time).
//
// The updates are those of the reference (direct.hpp:1858-2002) written in
// the kernel's canonical labels a, b, c, d (roles 0..3).  Each update
//   A(ar, ac) += coef * sum_{xr, xc} I(a,b,c,d) X(xr, xc)
// is spread over the lanes with the plane index running fastest: the planes
// of a kind are interleaved in memory, so lanes that share a target element
// read and add consecutive doubles and their loads / atomics coalesce.
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
                                   const PlaneSet &ps, unsigned mask, int mol, int lane, int W) {
  const MolPlanes &mp = ps.nmol == 1 ? ps.m0 : ps.mols[mol];
  const int N = mp.nbc;
  // stride of each role in the quartet layout I[((a n_b + b) n_c + c) n_d + d]
  const int st[4] = {n[1] * n[2] * n[3], n[2] * n[3], n[3], 1};
  for (int kind = 0; kind < 4; ++kind) {
    const int p0 = ps.kstart[kind], np = ps.kcount[kind];
    if (np == 0) continue;
    const int S = ps.kps[kind];
    const double *xk = mp.kx[kind];
    double *ak = mp.ka[kind];
    const double base_coef = kind == kJHer ? 1.0 : 0.5;
    for (int term = 0; term < kDigestTerms.n[kind]; ++term) {
      const int ar = kDigestTerms.t[kind][term][0], ac = kDigestTerms.t[kind][term][1];
      const int xr = kDigestTerms.t[kind][term][2], xc = kDigestTerms.t[kind][term][3];
      const int n_ac = n[ac], n_xr = n[xr], n_xc = n[xc];
      const int s_ar = st[ar], s_ac = st[ac], s_xr = st[xr], s_xc = st[xc];
      const int bf_ar = bf[ar], bf_ac = bf[ac], bf_xr = bf[xr], bf_xc = bf[xc];
      const int nt = n[ar] * n_ac;
      for (int t = lane; t < nt * np; t += W) {
        const int pl = p0 + t % np, tgt = t / np;
        const Plane &P = ps.p[pl];
        if (!((mask >> P.mat) & 1u)) continue;
        const int t0 = tgt / n_ac, t1 = tgt - (tgt / n_ac) * n_ac;
        const double *__restrict__ Ib = sI + t0 * s_ar + t1 * s_ac;
        const double *__restrict__ Xb = xk + (pl - p0) + ((size_t)bf_xr + (size_t)bf_xc * N) * S;
        double sum0 = 0.0, sum1 = 0.0;
        for (int u0 = 0; u0 < n_xr; ++u0) {
          const double *Iu = Ib + u0 * s_xr;
          const double *Xu = Xb + (size_t)u0 * S;
          int u1 = 0;
          for (; u1 + 1 < n_xc; u1 += 2) {
            sum0 = fma(Iu[u1 * s_xc], Xu[(size_t)u1 * N * S], sum0);
            sum1 = fma(Iu[(u1 + 1) * s_xc], Xu[(size_t)(u1 + 1) * N * S], sum1);
          }
          if (u1 < n_xc) sum0 = fma(Iu[u1 * s_xc], Xu[(size_t)u1 * N * S], sum0);
        }
        const double coef = kind == kKHer ? base_coef * P.sign : base_coef;
        unsafeAtomicAdd(&ak[(pl - p0) + ((bf_ar + t0) + (size_t)(bf_ac + t1) * N) * S], coef * (sum0 + sum1));
      }
    }
  }
}

}  // namespace cqx
