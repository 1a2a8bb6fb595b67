// J/K digestion of a tile of contracted Cartesian shell quartets of one
// class, staged in LDS, by all 64 lanes of a wave.
//
// The updates are the reference's (direct.hpp:1858-2002) in the kernel's
// canonical labels a, b, c, d: each update term
//   A(ar, ac) += coef * sum_{xr, xc} I(a, b, c, d) X(xr, xc)
// becomes work items (quartet, target element, plane slot) with the slot
// fastest, so the lanes that share a target read and add consecutive
// doubles of the kind's interleaved planes and their loads and atomics
// coalesce; target rows come next, so with one plane neighbouring lanes
// hit neighbouring doubles.
// Class sizes are compile-time constants (the index arithmetic folds); the
// number of planes of a kind is padded to a power of two.
#pragma once

#include <hip/hip_runtime.h>

#include "device_types.hpp"

namespace cqx {

// ONE: a one-molecule launch, whose planes sit in the kernel arguments.
// HER: a Hermitian kind, finalised as 1/2 (A + A^H): an update of A(r, c)
// may then land on A(c, r) instead (negated in an imaginary-part plane),
// and the target is stored with its longer dimension down a column so that
// neighbouring lanes add neighbouring doubles.
template <class K, int AR, int AC, int XR, int XC, bool ONE, bool HER>
__device__ __forceinline__ void digest_term_items(const double *__restrict__ sI, const int (*__restrict__ sbf)[4],
                                                  const unsigned *__restrict__ smask, const int *__restrict__ smol,
                                                  int nq, const PlaneSet &ps, int kind, int p0, int np, int lgp,
                                                  double base_coef, bool neg_ok, bool same_bra, int lane) {
  constexpr int NAR = K::n(AR), NT = NAR * K::n(AC);
  constexpr int ST[4] = {K::NCB * K::NCC * K::NCD, K::NCC * K::NCD, K::NCD, 1};
  constexpr bool TR = HER && K::n(AC) > K::n(AR);  // store A(ac, ar) instead
  constexpr int NFAST = TR ? K::n(AC) : NAR;       // target index along the column
  const int S = ps.kps[kind];
  // a bra-only target (J(a, b)) of a tile whose quartets share the bra:
  // one lane per (target, plane) sums over the quartets, one atomic each
  if constexpr (AR < 2 && AC < 2) {
    if (same_bra) {
      const int bitems = NT << lgp;
      const int *bf = sbf[0];
      const MolPlanes &mp = ONE ? ps.m0 : ps.mols[smol[0]];
      const int N = mp.nbc;
      for (int it = lane; it < bitems; it += 64) {
        const int p = it & ((1 << lgp) - 1);
        if (p >= np) continue;
        const int tgt = it >> lgp;
        const int t1 = tgt / NAR, t0 = tgt - t1 * NAR;
        const int slot = p0 + p;
        const int mat = (int)((ps.mats >> (4 * slot)) & 15ull);
        double s0 = 0.0, s1 = 0.0;
        for (int qi = 0; qi < nq; ++qi) {
          if (!((smask[qi] >> mat) & 1u)) continue;
          const int *bq = sbf[qi];
          const double *__restrict__ Iq = sI + qi * K::NC + t0 * ST[AR] + t1 * ST[AC];
          const double *__restrict__ X = mp.kx[kind] + p + ((size_t)bq[XR] + (size_t)bq[XC] * N) * S;
#pragma unroll
          for (int u0 = 0; u0 < K::n(XR); ++u0)
#pragma unroll
            for (int u1 = 0; u1 < K::n(XC); ++u1) {
              const double v = Iq[u0 * ST[XR] + u1 * ST[XC]] * X[((size_t)u0 + (size_t)u1 * N) * S];
              if ((u0 * K::n(XC) + u1) & 1) s1 += v; else s0 += v;
            }
        }
        const double coef = (neg_ok && ((ps.negs >> slot) & 1u)) ? -base_coef : base_coef;
        const double v = coef * (s0 + s1);
        if (v != 0.0)
          unsafeAtomicAdd(mp.ka[kind] + p + ((size_t)(bf[AR] + t0) + (size_t)(bf[AC] + t1) * N) * S, v);
      }
      return;
    }
  }
  const int items = (nq * NT) << lgp;
  for (int it = lane; it < items; it += 64) {
    const int p = it & ((1 << lgp) - 1);
    if (p >= np) continue;
    const int r = it >> lgp;
    const int qi = r / NT, tgt = r - qi * NT;
    // the stored row index fastest: neighbouring lanes hit neighbouring elements
    const int tslow = tgt / NFAST, tfast = tgt - tslow * NFAST;
    const int t0 = TR ? tslow : tfast, t1 = TR ? tfast : tslow;
    const int slot = p0 + p;
    const int mat = (int)((ps.mats >> (4 * slot)) & 15ull);
    if (!((smask[qi] >> mat) & 1u)) continue;
    const int *bf = sbf[qi];
    const MolPlanes &mp = ONE ? ps.m0 : ps.mols[smol[qi]];
    const int N = mp.nbc;
    const double *__restrict__ xk = mp.kx[kind];
    double *ak = mp.ka[kind];
    const double *__restrict__ Iq = sI + qi * K::NC + t0 * ST[AR] + t1 * ST[AC];
    const double *__restrict__ X = xk + p + ((size_t)bf[XR] + (size_t)bf[XC] * N) * S;
    double s0 = 0.0, s1 = 0.0;
#pragma unroll
    for (int u0 = 0; u0 < K::n(XR); ++u0)
#pragma unroll
      for (int u1 = 0; u1 < K::n(XC); ++u1) {
        const double v = Iq[u0 * ST[XR] + u1 * ST[XC]] * X[((size_t)u0 + (size_t)u1 * N) * S];
        if ((u0 * K::n(XC) + u1) & 1) s1 += v; else s0 += v;
      }
    const bool neg = (ps.negs >> slot) & 1u;
    // conj of Im X for Hermitian K; a transposed update of an imaginary plane flips sign again
    const double coef = ((neg_ok && neg) != (TR && neg)) ? -base_coef : base_coef;
    const size_t e = TR ? ((size_t)(bf[AC] + t1) + (size_t)(bf[AR] + t0) * N)
                        : ((size_t)(bf[AR] + t0) + (size_t)(bf[AC] + t1) * N);
    unsafeAtomicAdd(ak + p + e * S, coef * (s0 + s1));
  }
}

template <class K, int AR, int AC, int XR, int XC, bool HER>
__device__ __forceinline__ void digest_term_tile(const double *__restrict__ sI, const int (*__restrict__ sbf)[4],
                                                 const unsigned *__restrict__ smask, const int *__restrict__ smol,
                                                 int nq, const PlaneSet &ps, int kind, int p0, int np, int lgp,
                                                 double base_coef, bool neg_ok, bool same_bra, int lane) {
  if (ps.nmol == 1)
    digest_term_items<K, AR, AC, XR, XC, true, HER>(sI, sbf, smask, smol, nq, ps, kind, p0, np, lgp, base_coef, neg_ok,
                                               same_bra, lane);
  else
    digest_term_items<K, AR, AC, XR, XC, false, HER>(sI, sbf, smask, smol, nq, ps, kind, p0, np, lgp, base_coef,
                                                neg_ok, same_bra, lane);
}

// Every update of every plane kind for nq quartets: sI holds quartet q at
// sI[q * NC] (degeneracy and 1/2 applied), sbf[q] its first Cartesian
// functions in roles a, b, c, d, smask[q] the contractions it passed and
// smol[q] its molecule.
template <class K>
__device__ __forceinline__ void digest_tile(const double *sI, const int (*sbf)[4], const unsigned *smask,
                                            const int *smol, int nq, const PlaneSet &ps, int lane) {
  // do all quartets of the tile share the bra pair (first functions and molecule)?
  const bool differs = lane < nq && (sbf[lane][0] != sbf[0][0] || sbf[lane][1] != sbf[0][1] || smol[lane] != smol[0]);
  const bool same_bra = nq > 1 && __ballot(differs) == 0ull;
  for (int kind = 0; kind < 4; ++kind) {
    const int np = ps.kcount[kind];
    if (np == 0) continue;
    const int p0 = ps.kstart[kind];
    const int lgp = np <= 1 ? 0 : np <= 2 ? 1 : np <= 4 ? 2 : np <= 8 ? 3 : 4;
    switch (kind) {
      case kJHer:  // direct.hpp:1863-1883, real parts
        digest_term_tile<K, 0, 1, 3, 2, true>(sI, sbf, smask, smol, nq, ps, kind, p0, np, lgp, 1.0, false, same_bra, lane);
        digest_term_tile<K, 3, 2, 0, 1, true>(sI, sbf, smask, smol, nq, ps, kind, p0, np, lgp, 1.0, false, same_bra, lane);
        break;
      case kKHer:  // :1889-1917, 1/2 conj(X)
        digest_term_tile<K, 0, 2, 3, 1, true>(sI, sbf, smask, smol, nq, ps, kind, p0, np, lgp, 0.5, true, same_bra, lane);
        digest_term_tile<K, 3, 1, 0, 2, true>(sI, sbf, smask, smol, nq, ps, kind, p0, np, lgp, 0.5, true, same_bra, lane);
        digest_term_tile<K, 3, 0, 1, 2, true>(sI, sbf, smask, smol, nq, ps, kind, p0, np, lgp, 0.5, true, same_bra, lane);
        digest_term_tile<K, 1, 2, 3, 0, true>(sI, sbf, smask, smol, nq, ps, kind, p0, np, lgp, 0.5, true, same_bra, lane);
        break;
      case kJNonHer:  // :1925-1960
        digest_term_tile<K, 0, 1, 3, 2, false>(sI, sbf, smask, smol, nq, ps, kind, p0, np, lgp, 0.5, false, same_bra, lane);
        digest_term_tile<K, 0, 1, 2, 3, false>(sI, sbf, smask, smol, nq, ps, kind, p0, np, lgp, 0.5, false, same_bra, lane);
        digest_term_tile<K, 1, 0, 3, 2, false>(sI, sbf, smask, smol, nq, ps, kind, p0, np, lgp, 0.5, false, same_bra, lane);
        digest_term_tile<K, 1, 0, 2, 3, false>(sI, sbf, smask, smol, nq, ps, kind, p0, np, lgp, 0.5, false, same_bra, lane);
        digest_term_tile<K, 2, 3, 1, 0, false>(sI, sbf, smask, smol, nq, ps, kind, p0, np, lgp, 0.5, false, same_bra, lane);
        digest_term_tile<K, 2, 3, 0, 1, false>(sI, sbf, smask, smol, nq, ps, kind, p0, np, lgp, 0.5, false, same_bra, lane);
        digest_term_tile<K, 3, 2, 1, 0, false>(sI, sbf, smask, smol, nq, ps, kind, p0, np, lgp, 0.5, false, same_bra, lane);
        digest_term_tile<K, 3, 2, 0, 1, false>(sI, sbf, smask, smol, nq, ps, kind, p0, np, lgp, 0.5, false, same_bra, lane);
        break;
      default:  // kKNonHer, :1961-1998
        digest_term_tile<K, 2, 0, 3, 1, false>(sI, sbf, smask, smol, nq, ps, kind, p0, np, lgp, 0.5, false, same_bra, lane);
        digest_term_tile<K, 3, 1, 2, 0, false>(sI, sbf, smask, smol, nq, ps, kind, p0, np, lgp, 0.5, false, same_bra, lane);
        digest_term_tile<K, 3, 0, 2, 1, false>(sI, sbf, smask, smol, nq, ps, kind, p0, np, lgp, 0.5, false, same_bra, lane);
        digest_term_tile<K, 2, 1, 3, 0, false>(sI, sbf, smask, smol, nq, ps, kind, p0, np, lgp, 0.5, false, same_bra, lane);
        digest_term_tile<K, 0, 2, 1, 3, false>(sI, sbf, smask, smol, nq, ps, kind, p0, np, lgp, 0.5, false, same_bra, lane);
        digest_term_tile<K, 1, 3, 0, 2, false>(sI, sbf, smask, smol, nq, ps, kind, p0, np, lgp, 0.5, false, same_bra, lane);
        digest_term_tile<K, 0, 3, 1, 2, false>(sI, sbf, smask, smol, nq, ps, kind, p0, np, lgp, 0.5, false, same_bra, lane);
        digest_term_tile<K, 1, 2, 0, 3, false>(sI, sbf, smask, smol, nq, ps, kind, p0, np, lgp, 0.5, false, same_bra, lane);
        break;
    }
  }
}

}  // namespace cqx
