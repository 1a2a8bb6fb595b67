// Lane-group Rys ERI + J/K digestion kernels for the low angular-momentum
// classes (Cartesian quartet size <= 27, and four classes of 36 and 54
// components that also have wave kernels), templated on the class and on
// the number G of lanes that share one shell quartet.
//
// Every size is a compile-time constant, so the Rys nodes/weights, the 2D
// recursions and the contracted Cartesian quartet stay in registers; the
// quartets of a wave are then staged in LDS and digested by the whole wave
// (digest_tile.hpp) following the reference formulas (direct.hpp:1858-2002).
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "digest.hpp"
#include "digest_tile.hpp"
#include "kernels.hpp"
#include "limits.hpp"
#include "rys.hpp"

namespace cqx {

namespace {

constexpr double kTwoPi52 = 34.98683665524972;  // 2 pi^(5/2)

constexpr int ncart_c(int l) { return (l + 1) * (l + 2) / 2; }

// Cartesian exponents of component c of an l shell (lx descending, then ly)
constexpr int cexp(int l, int c, int dir) {
  int k = c, lx = l;
  while (k >= l - lx + 1) {
    k -= l - lx + 1;
    --lx;
  }
  const int lz = k, ly = l - lx - lz;
  return dir == 0 ? lx : (dir == 1 ? ly : lz);
}

__device__ __forceinline__ void atomic_add(double *p, double v) { unsafeAtomicAdd(p, v); }

template <int LA, int LB, int LC, int LD>
struct Cls {
  static constexpr int NCA = ncart_c(LA), NCB = ncart_c(LB), NCC = ncart_c(LC), NCD = ncart_c(LD);
  static constexpr int NC = NCA * NCB * NCC * NCD;
  static constexpr int NI = LA + LB + 1, NK = LC + LD + 1;
  static constexpr int NR = (LA + LB + LC + LD) / 2 + 1;
  static constexpr int NH = (LA + 1) * (LB + 1) * (LC + 1) * (LD + 1);
  static constexpr int n(int role) { return role == 0 ? NCA : role == 1 ? NCB : role == 2 ? NCC : NCD; }
  static constexpr int oidx(int ia, int ib, int ic, int id) { return ((ia * NCB + ib) * NCC + ic) * NCD + id; }
  static constexpr int hidx(int a, int b, int c, int d) {
    return ((a * (LB + 1) + b) * (LC + 1) + c) * (LD + 1) + d;
  }
};

// For each Cartesian component o of the quartet, the offsets of its x, y
// and z factors in the 2D-integral tables; built at compile time so that an
// unrolled loop over o indexes the tables with constants.
template <int LA, int LB, int LC, int LD>
struct HTab {
  int x[Cls<LA, LB, LC, LD>::NC], y[Cls<LA, LB, LC, LD>::NC], z[Cls<LA, LB, LC, LD>::NC];
};

template <int LA, int LB, int LC, int LD>
constexpr HTab<LA, LB, LC, LD> make_htab() {
  using K = Cls<LA, LB, LC, LD>;
  HTab<LA, LB, LC, LD> t{};
  for (int ia = 0; ia < K::NCA; ++ia)
    for (int ib = 0; ib < K::NCB; ++ib)
      for (int ic = 0; ic < K::NCC; ++ic)
        for (int id = 0; id < K::NCD; ++id) {
          const int o = K::oidx(ia, ib, ic, id);
          t.x[o] = K::hidx(cexp(LA, ia, 0), cexp(LB, ib, 0), cexp(LC, ic, 0), cexp(LD, id, 0));
          t.y[o] = K::hidx(cexp(LA, ia, 1), cexp(LB, ib, 1), cexp(LC, ic, 1), cexp(LD, id, 1));
          t.z[o] = K::hidx(cexp(LA, ia, 2), cexp(LB, ib, 2), cexp(LC, ic, 2), cexp(LD, id, 2));
        }
  return t;
}

// 2D integrals of one root along one direction: VRR on (i, k), then HRR to
// (ia, ib, ic, id); H has Cls::NH entries.
template <int LA, int LB, int LC, int LD, class R>
__device__ __forceinline__ void rys_2d_full(R C00, R D00, R B10, R B01, R B00, R ABd, R CDd, R scale,
                                            R (&H)[Cls<LA, LB, LC, LD>::NH]) {
  using K = Cls<LA, LB, LC, LD>;
  R G[K::NI][K::NK];
  G[0][0] = scale;
  if constexpr (K::NI > 1) G[1][0] = C00 * scale;
#pragma unroll
  for (int i = 1; i + 1 < K::NI; ++i) G[i + 1][0] = C00 * G[i][0] + i * B10 * G[i - 1][0];
  if constexpr (K::NK > 1) {
#pragma unroll
    for (int k = 0; k + 1 < K::NK; ++k) {
      if (k == 0) {
        G[0][1] = D00 * G[0][0];
#pragma unroll
        for (int i = 1; i < K::NI; ++i) G[i][1] = D00 * G[i][0] + i * B00 * G[i - 1][0];
      } else {
        G[0][k + 1] = D00 * G[0][k] + k * B01 * G[0][k - 1];
#pragma unroll
        for (int i = 1; i < K::NI; ++i)
          G[i][k + 1] = D00 * G[i][k] + k * B01 * G[i][k - 1] + i * B00 * G[i - 1][k];
      }
    }
  }
  // bra transfer: E[i][j][k], (i, j+1) = (i+1, j) + AB (i, j)
  R E[K::NI][LB + 1][K::NK];
#pragma unroll
  for (int i = 0; i < K::NI; ++i)
#pragma unroll
    for (int k = 0; k < K::NK; ++k) E[i][0][k] = G[i][k];
#pragma unroll
  for (int j = 1; j <= LB; ++j)
#pragma unroll
    for (int i = 0; i + j < K::NI; ++i)
#pragma unroll
      for (int k = 0; k < K::NK; ++k) E[i][j][k] = E[i + 1][j - 1][k] + ABd * E[i][j - 1][k];
  // ket transfer for each (ia, ib)
#pragma unroll
  for (int ia = 0; ia <= LA; ++ia)
#pragma unroll
    for (int ib = 0; ib <= LB; ++ib) {
      R F[K::NK][LD + 1];
#pragma unroll
      for (int k = 0; k < K::NK; ++k) F[k][0] = E[ia][ib][k];
#pragma unroll
      for (int l = 1; l <= LD; ++l)
#pragma unroll
        for (int k = 0; k + l < K::NK; ++k) F[k][l] = F[k + 1][l - 1] + CDd * F[k][l - 1];
#pragma unroll
      for (int ic = 0; ic <= LC; ++ic)
#pragma unroll
        for (int id = 0; id <= LD; ++id) H[K::hidx(ia, ib, ic, id)] = F[ic][id];
    }
}

// The factors of a primitive quartet (bra exponent sum p, ket q) that need
// a division, from one reciprocal d = 1 / (p q (p + q)):
// 1 / (p + q), 1 / (p q), and the prefactor K_ab K_cd 2 pi^(5/2) / (p q sqrt(p + q)).
template <class R>
struct PrimRecip {
  R rpq, rab, pref;
};
template <class R>
__device__ __forceinline__ PrimRecip<R> prim_recip(R ap, R bq, double kk) {
  const R pq = ap + bq, ab = ap * bq;
  const R d = eval_rcp<R>(ab * pq);
  PrimRecip<R> r;
  r.rpq = ab * d;
  r.rab = pq * d;
  r.pref = (R)(kk * kTwoPi52) * r.rab * eval_sqrt<R>(r.rpq);
  return r;
}

// Rys table of the evaluation precision
template <class R> __device__ __forceinline__ const R *rys_table(const double *r64, const float *r32);
template <> __device__ __forceinline__ const double *rys_table<double>(const double *r64, const float *) { return r64; }
template <> __device__ __forceinline__ const float *rys_table<float>(const double *, const float *r32) { return r32; }

// The Rys table section of the class's root count: staged in LDS by the
// whole block when the launch asks for it (the lane-group classes need at
// most 2 roots: 9 KB / 20 KB of FP64 coefficients), else read from global
// memory.  Every thread of the block must call this before any returns.
template <int NR, class R>
__device__ __forceinline__ const R *rys_section(const double *r64, const float *r32, int lds, R *s_tab) {
  const R *g = rys_table<R>(r64, r32) + rys_tab_off(NR);
  if (!lds) return g;
  for (int e = threadIdx.x; e < rys_tab_len(NR); e += blockDim.x) s_tab[e] = g[e];
  __syncthreads();
  return s_tab;
}

// The contracted Cartesian quartet (bp | kp) by the G lanes of a group:
// lane g takes primitive quartets g, g + G, ... and the group's sum is
// combined by a butterfly of lane shuffles, so every lane of the group ends
// with the whole quartet in I (unscaled).  R is the evaluation precision.
template <int LA, int LB, int LC, int LD, int G, class R>
__device__ __forceinline__ void eval_quartet(const PrimPair *__restrict__ prims, const PrimC32 *__restrict__ pc32,
                                             const R *__restrict__ rysn, const PairRec &bp, const PairRec &kp,
                                             int g, R (&I)[Cls<LA, LB, LC, LD>::NC]) {
  using K = Cls<LA, LB, LC, LD>;
  constexpr bool F32 = std::is_same<R, float>::value;
#pragma unroll
  for (int o = 0; o < K::NC; ++o) I[o] = 0;
  R AB[3], CD[3], AC[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    AB[d] = (R)bp.AB[d];
    CD[d] = (R)kp.AB[d];
    AC[d] = F32 ? (R)(bp.A[d] - kp.A[d]) : (R)0;  // A - C in FP64 before the rounding
  }
  const int nk = kp.npp, ntot = bp.npp * kp.npp;
  int i = 0, j = g;
  while (nk > 0 && j >= nk) {
    j -= nk;
    ++i;
  }
  for (int k = g; k < ntot; k += G) {
    R ap, bq, PA[3], QC[3], PQ[3];
    double kk;
    if constexpr (F32) {
      // 16-byte records: P - A = -(b/p) AB, Q - C = -(d/q) CD
      const PrimC32 A = pc32[bp.pp_off + i];
      const PrimC32 B = pc32[kp.pp_off + j];
      ap = A.p;
      bq = B.p;
      kk = A.K * B.K;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        PA[d] = -A.bp * AB[d];
        QC[d] = -B.bp * CD[d];
        PQ[d] = AC[d] + PA[d] - QC[d];
      }
    } else {
      const PrimPair A = prims[bp.pp_off + i];
      const PrimPair B = prims[kp.pp_off + j];
      ap = A.p;
      bq = B.p;
      kk = A.K * B.K;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        PQ[d] = A.P[d] - B.P[d];
        PA[d] = A.PA[d];
        QC[d] = B.PA[d];
      }
    }
    j += G;
    while (j >= nk) {
      j -= nk;
      ++i;
    }
    const PrimRecip<R> pr = prim_recip<R>(ap, bq, kk);
    const R rpq = pr.rpq, pref = pr.pref;
    const R rho = ap * bq * rpq;
    const R T = rho * (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]);
    R x[K::NR], w[K::NR];
    rys_roots_sec<K::NR, R>(T, rysn, x, w);
    if constexpr (K::NC == 1) {
      R s = 0;
#pragma unroll
      for (int r = 0; r < K::NR; ++r) s += w[r];
      I[0] = fma(pref, s, I[0]);
    } else {
      const R rp = bq * pr.rab, rq = ap * pr.rab;
#pragma unroll
      for (int r = 0; r < K::NR; ++r) {
        const R t2 = x[r];
        const R B00 = (R)0.5 * t2 * rpq;
        const R B10 = (R)0.5 * rp * ((R)1 - rho * rp * t2);
        const R B01 = (R)0.5 * rq * ((R)1 - rho * rq * t2);
        R H[3][K::NH];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const R C00 = PA[d] - bq * rpq * t2 * PQ[d];
          const R D00 = QC[d] + ap * rpq * t2 * PQ[d];
          rys_2d_full<LA, LB, LC, LD, R>(C00, D00, B10, B01, B00, AB[d], CD[d],
                                         d == 2 ? pref * w[r] : (R)1, H[d]);
        }
        constexpr HTab<LA, LB, LC, LD> HT = make_htab<LA, LB, LC, LD>();
#pragma unroll
        for (int o = 0; o < K::NC; ++o) I[o] = fma(H[0][HT.x[o]] * H[1][HT.y[o]], H[2][HT.z[o]], I[o]);
      }
    }
  }
  if constexpr (G > 1) {
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1)
#pragma unroll
      for (int o = 0; o < K::NC; ++o) I[o] += __shfl_xor(I[o], off);
  }
}

// degeneracy of a quartet in the 8-fold loop (direct.hpp:1664-1667,
// 1793-1802) and the 1/2 of :1849-1850
__device__ __forceinline__ double quartet_scale(const PairRec &bp, const PairRec &kp) {
  const bool same_pair = (bp.a == kp.a && bp.b == kp.b);
  return 0.5 * (bp.a == bp.b ? 1.0 : 2.0) * (kp.a == kp.b ? 1.0 : 2.0) * (same_pair ? 1.0 : 2.0);
}

// G lanes per quartet: the lanes split the primitive quartets of the shell
// quartet and combine the contracted Cartesian integrals with a butterfly of
// lane shuffles; the wave's 64 / G quartets then go to LDS and all 64 lanes
// digest them (digest_tile.hpp), so the registers of the evaluation and of
// the digestion are never live together.  A block takes a run of
// consecutive batches of 4 x 64 / G quartets (lane_group_blocks in
// limits.hpp), one after the other, over one copy of the Rys table section.
// R is the evaluation precision (Rys nodes, 2D integrals, the contracted
// quartet); the digestion and the accumulators are FP64 either way.
template <int LA, int LB, int LC, int LD, int G, class R>
__global__ __launch_bounds__(256) void eri_fast_kernel(EriLaunch a) {
  using K = Cls<LA, LB, LC, LD>;
  constexpr int QPW = 64 / G;  // quartets per wave
  __shared__ double s_I[4][QPW * K::NC];
  __shared__ int s_bf[4][QPW][4];
  __shared__ unsigned s_mask[4][QPW];
  __shared__ int s_mol[4][QPW];
  __shared__ R s_tab[rys_tab_len(K::NR)];
  const R *__restrict__ rysn = rys_section<K::NR, R>(a.rys, a.rys32, a.rys_lds, s_tab);
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int g = lane & (G - 1), qw = lane / G;
  // quartet count: the list's device counter for a launch sized to the
  // list's capacity before the list was built
  const long long nq = a.nq_dev ? min((long long)*a.nq_dev - a.nq_base, a.nq) : a.nq;
  // this block's run of batches: the launch's capacity a.nq in batches,
  // split into gridDim.x runs of consecutive batches whose lengths differ
  // by at most one
  const long long nbatch = (a.nq + 4 * QPW - 1) / (4 * QPW);
  const long long b0 = (long long)blockIdx.x * nbatch / gridDim.x;
  const long long b1 = ((long long)blockIdx.x + 1) * nbatch / gridDim.x;
  // each lane's list entry of the next batch, loaded before the current
  // batch's digestion so that its pair-record loads need not wait on it
  Quartet qn{0u, 0u};
  {
    const long long qf = (b0 * 4 + wv) * QPW + qw;
    if (b0 < b1 && qf < nq) qn = a.quartets[qf];
  }
  for (long long bt = b0; bt < b1; ++bt) {
    const long long q0 = (bt * 4 + wv) * QPW;
    if (q0 >= nq) break;  // uniform across the wave; no block barriers below
    if (bt > b0) wave_sync();  // the previous batch's digestion has read the tile
    const long long q = q0 + qw;
    const bool valid = q < nq;
    Quartet qt{0u, 0u};
    PairRec bp{}, kp{};
    if (valid) {
      qt = qn;
      bp = a.bra[qt.bra];
      kp = a.ket[qt.ket_mask & 0xFFFFFFu];
    }
    R I[K::NC];
    eval_quartet<LA, LB, LC, LD, G, R>(a.prims, a.primc32, rysn, bp, kp, g, I);
    const double scale = quartet_scale(bp, kp);
    double *sI = s_I[wv] + qw * K::NC;
#pragma unroll
    for (int o = 0; o < K::NC; ++o)
      if (o % G == g) sI[o] = (double)I[o] * scale;
    if (g == 0) {
      s_bf[wv][qw][0] = bp.bfa;
      s_bf[wv][qw][1] = bp.bfb;
      s_bf[wv][qw][2] = kp.bfa;
      s_bf[wv][qw][3] = kp.bfb;
      s_mask[wv][qw] = valid ? (qt.ket_mask >> 24) : 0u;
      s_mol[wv][qw] = bp.mol;
    }
    {
      const long long qf = ((bt + 1) * 4 + wv) * QPW + qw;
      if (bt + 1 < b1 && qf < nq) qn = a.quartets[qf];
    }
    wave_sync();
    const long long left = nq - q0;
    digest_tile<K>(s_I[wv], s_bf[wv], s_mask[wv], s_mol[wv], left < QPW ? (int)left : QPW, a.planes, lane);
  }
}

template <int LA, int LB, int LC, int LD, int G, class R>
hipError_t launch_cls(const EriLaunch &a, hipStream_t s) {
  constexpr long long per_batch = 4 * (64 / G);
  const long long blocks = lane_group_blocks((a.nq + per_batch - 1) / per_batch);
  hipLaunchKernelGGL((eri_fast_kernel<LA, LB, LC, LD, G, R>), dim3((unsigned)blocks), dim3(256), 0, s, a);
  return hipGetLastError();
}

}  // namespace

// classes (bra class >= ket class in (la,lb) order ss, ps, pp, ds, dp, dd)
// with at most 27 Cartesian components
#define CQX_FAST_CLASSES(X) \
  X(0, 0, 0, 0)             \
  X(1, 0, 0, 0)             \
  X(1, 0, 1, 0)             \
  X(1, 1, 0, 0)             \
  X(1, 1, 1, 0)             \
  X(2, 0, 0, 0)             \
  X(2, 0, 1, 0)             \
  X(2, 1, 0, 0)

// classes of 36 and 54 components built in both families: the autotuner
// picks the lane-group or the wave kernel for each (the lane-group kernels
// win with one density plane per kind, the wave kernels with eight)
#define CQX_FAST_DUAL_CLASSES(X) \
  X(2, 0, 2, 0)                  \
  X(2, 2, 0, 0)                  \
  X(2, 1, 1, 0)                  \
  X(2, 0, 1, 1)

// all classes of the lists for one lane-group size G and precision R
template <int G, class R>
hipError_t launch_fast_g(const EriLaunch &a, hipStream_t s) {
#define CQX_CASE(A, B, C, D) \
  if (a.la == A && a.lb == B && a.lc == C && a.ld == D) return launch_cls<A, B, C, D, G, R>(a, s);
  CQX_FAST_CLASSES(CQX_CASE)
  CQX_FAST_DUAL_CLASSES(CQX_CASE)
#undef CQX_CASE
  return hipErrorNotSupported;
}

}  // namespace cqx
