// Wave kernels for the larger s/p/d classes (Cartesian quartet size > 27)
// and the classes with an f or g shell up to 1080 Cartesian components,
// templated on the class and on QT, the quartets a wave takes together.
//
// The primitive quartets of the wave's QT quartets form one list.  Each
// round takes up to PP of them: lane group g (3 n_r lanes) builds the 2D
// integrals of the round's g-th primitive quartet, one lane per (root,
// direction), each lane evaluating only its own Rys node and weight, and
// stores them to the wave's LDS slice; then every lane adds the round's
// tables to its share of the contracted Cartesian quartet they belong to.
// After the last round the QT quartets go to LDS, in the space the tables
// took, and the lanes digest them together (digest_tile.hpp).  Instantiated
// per QT (eri_wave_q*.hip; the f and g classes, FP64 only, in eri_wave_f*.hip
// and eri_wave_g*.hip, 1, 2 and, up to 216 components, 4 quartets per wave).
#pragma once

#include <hip/hip_runtime.h>

#include "digest.hpp"
#include "eri_fast_impl.hpp"

namespace cqx {

namespace {

constexpr int kWavesPerBlock = 4;

// R is the evaluation precision; the quartet handed to the digestion and
// the accumulators are FP64 either way.
// Primitive quartets whose 2D tables one round builds side by side: as many
// groups of 3 n_r lanes as fit in a wave, at most 8, within 8 KB of LDS.
template <class K, class R>
constexpr int prims_per_round() {
  constexpr int by_lanes = 64 / (3 * K::NR);
  constexpr int by_lds = 8192 / (int)(3 * K::NR * K::NH * sizeof(R));
  constexpr int p = by_lanes < 8 ? by_lanes : 8;
  return p < by_lds ? (p < 1 ? 1 : p) : (by_lds < 1 ? 1 : by_lds);
}

// A wave's LDS buffer in doubles: the 2D tables of one round, and after
// the last round the QT contracted quartets, share it
template <class K, class R, int QT>
constexpr int wave_buf_doubles() {
  constexpr int tables = (prims_per_round<K, R>() * 3 * K::NR * K::NH * (int)sizeof(R) + 7) / 8;
  return tables > QT * K::NC ? tables : QT * K::NC;
}

// LDS of a block of the kernel below with QT quartets per wave
template <class K, class R, int QT>
constexpr int wave_lds_bytes() {
  return kWavesPerBlock * (wave_buf_doubles<K, R, QT>() * 8 + QT * 24 + QT * 2 * (int)sizeof(PairRec));
}

// QT: quartets a wave takes together.  Their primitive quartets form one
// list that the rounds walk through, so the lane groups of a round stay busy
// when a quartet has only a few primitive quartets, and each lane keeps its
// outputs of every quartet.  The QT quartets are then digested together: a
// single quartet gives an update term fewer items than lanes (e.g. 18 for
// (dp|ps) with one plane), so most lanes of a digestion step would idle
// while it waits on memory.
template <int LA, int LB, int LC, int LD, class R, int QT>
__global__ __launch_bounds__(64 * kWavesPerBlock) void eri_wave_kernel(EriLaunch a) {
  using K = Cls<LA, LB, LC, LD>;
  constexpr int NOWN = (K::NC + 63) / 64;
  constexpr int HSZ = 3 * K::NR * K::NH;
  constexpr int PP = prims_per_round<K, R>();
  __shared__ double s_buf[kWavesPerBlock][wave_buf_doubles<K, R, QT>()];
  __shared__ int s_bf[kWavesPerBlock][QT][4];
  __shared__ unsigned s_mask[kWavesPerBlock][QT];
  __shared__ int s_mol[kWavesPerBlock][QT];
  __shared__ PairRec s_pr[kWavesPerBlock][QT][2];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  // quartet count: the list's device counter for a launch sized to the
  // list's capacity before the list was built
  const long long nq = a.nq_dev ? min((long long)*a.nq_dev - a.nq_base, a.nq) : a.nq;
  const long long q0 = ((long long)blockIdx.x * kWavesPerBlock + wv) * QT;
  if (q0 >= nq) return;  // uniform across the wave; no block barriers below
  const int ntile = nq - q0 < QT ? (int)(nq - q0) : QT;
  // lane t < ntile stages quartet t's pair records
  if (lane < ntile) {
    const Quartet qt = a.quartets[q0 + lane];
    const PairRec bp = a.bra[qt.bra];
    const PairRec kp = a.ket[qt.ket_mask & 0xFFFFFFu];
    s_pr[wv][lane][0] = bp;
    s_pr[wv][lane][1] = kp;
    s_bf[wv][lane][0] = bp.bfa;
    s_bf[wv][lane][1] = bp.bfb;
    s_bf[wv][lane][2] = kp.bfa;
    s_bf[wv][lane][3] = kp.bfb;
    s_mask[wv][lane] = qt.ket_mask >> 24;
    s_mol[wv][lane] = bp.mol;
  }
  wave_sync();
  // offsets of the quartets' primitive quartets in the tile's list
  int off[QT + 1];
  off[0] = 0;
#pragma unroll
  for (int t = 0; t < QT; ++t)
    off[t + 1] = off[t] + (t < ntile ? s_pr[wv][t][0].npp * s_pr[wv][t][1].npp : 0);
  const int ntot = off[QT];
  auto quartet_of = [&](int kq) {
    int t = 0;
#pragma unroll
    for (int u = 1; u < QT; ++u)
      if (kq >= off[u]) t = u;
    return t;
  };
  R *H = reinterpret_cast<R *>(s_buf[wv]);  // the round's 2D tables
  const R *__restrict__ rys = rys_table<R>(a.rys, a.rys32);

  // the lane's outputs o = lane + 64 k and their 2D-table offsets
  int ox[NOWN], oy[NOWN], oz[NOWN];
  R acc[QT][NOWN];
#pragma unroll
  for (int k = 0; k < NOWN; ++k) {
#pragma unroll
    for (int t = 0; t < QT; ++t) acc[t][k] = 0;
    int o = lane + 64 * k;
    if (o >= K::NC) o = 0;
    int rem = o;
    const int id = rem % K::NCD; rem /= K::NCD;
    const int ic = rem % K::NCC; rem /= K::NCC;
    const int ib = rem % K::NCB;
    const int ia = rem / K::NCB;
    ox[k] = K::hidx(cexp(LA, ia, 0), cexp(LB, ib, 0), cexp(LC, ic, 0), cexp(LD, id, 0));
    oy[k] = K::NH + K::hidx(cexp(LA, ia, 1), cexp(LB, ib, 1), cexp(LC, ic, 1), cexp(LD, id, 1));
    oz[k] = 2 * K::NH + K::hidx(cexp(LA, ia, 2), cexp(LB, ib, 2), cexp(LC, ic, 2), cexp(LD, id, 2));
  }
  // lane group g = lane / (3 n_r) builds the 2D tables of primitive quartet
  // k0 + g of the tile's list; its lanes (root, direction) one table each
  const int grp = lane / (3 * K::NR), sub = lane - grp * (3 * K::NR);
  const int my_r = sub / 3, my_d = sub - my_r * 3;
  for (int k0 = 0; k0 < ntot; k0 += PP) {
    const int cnt = ntot - k0 < PP ? ntot - k0 : PP;
    if (grp < cnt) {
      const int kq = k0 + grp;
      const int t = quartet_of(kq);
      const PairRec &bp = s_pr[wv][t][0];
      const PairRec &kp = s_pr[wv][t][1];
      const int loc = kq - off[t], nk = kp.npp;
      const int i = loc / nk, j = loc - i * nk;
      R ap, bq, PQ[3], PAv[3], QCv[3];
      double kk;
      if constexpr (std::is_same<R, float>::value) {
        // 16-byte records: P - A = -(b/p) AB, Q - C = -(d/q) CD,
        // P - Q = (A - C) + (P - A) - (Q - C) with A - C in FP64 before the rounding
        const PrimC32 A = a.primc32[bp.pp_off + i];
        const PrimC32 B = a.primc32[kp.pp_off + j];
        ap = A.p;
        bq = B.p;
        kk = A.K * B.K;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          PAv[d] = -A.bp * (R)bp.AB[d];
          QCv[d] = -B.bp * (R)kp.AB[d];
          PQ[d] = (R)(bp.A[d] - kp.A[d]) + PAv[d] - QCv[d];
        }
      } else {
        const PrimPair A = a.prims[bp.pp_off + i];
        const PrimPair B = a.prims[kp.pp_off + j];
        ap = A.p;
        bq = B.p;
        kk = A.K * B.K;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          PQ[d] = A.P[d] - B.P[d];
          PAv[d] = A.PA[d];
          QCv[d] = B.PA[d];
        }
      }
      const PrimRecip<R> pr = prim_recip<R>(ap, bq, kk);
      const R rpq = pr.rpq, pref = pr.pref;
      const R rho = ap * bq * rpq;
      const R T = rho * (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]);
      R t2, wr;
      rys_root<K::NR, R>(T, rys, my_r, t2, wr);
      const R rp = bq * pr.rab, rq = ap * pr.rab;
      const R B00 = (R)0.5 * t2 * rpq;
      const R B10 = (R)0.5 * rp * ((R)1 - rho * rp * t2);
      const R B01 = (R)0.5 * rq * ((R)1 - rho * rq * t2);
      const R PQd = my_d == 0 ? PQ[0] : (my_d == 1 ? PQ[1] : PQ[2]);
      const R PAd = my_d == 0 ? PAv[0] : (my_d == 1 ? PAv[1] : PAv[2]);
      const R QCd = my_d == 0 ? QCv[0] : (my_d == 1 ? QCv[1] : QCv[2]);
      const R ABd = (R)(my_d == 0 ? bp.AB[0] : (my_d == 1 ? bp.AB[1] : bp.AB[2]));
      const R CDd = (R)(my_d == 0 ? kp.AB[0] : (my_d == 1 ? kp.AB[1] : kp.AB[2]));
      const R C00 = PAd - bq * rpq * t2 * PQd;
      const R D00 = QCd + ap * rpq * t2 * PQd;
      R h[K::NH];
      rys_2d_full<LA, LB, LC, LD, R>(C00, D00, B10, B01, B00, ABd, CDd, my_d == 2 ? pref * wr : (R)1, h);
      R *dst = H + grp * HSZ + (my_r * 3 + my_d) * K::NH;
#pragma unroll
      for (int e = 0; e < K::NH; ++e) dst[e] = h[e];
    }
    wave_sync();
    for (int g = 0; g < cnt; ++g) {
      const int t = quartet_of(k0 + g);  // uniform across the wave
#pragma unroll
      for (int k = 0; k < NOWN; ++k) {
        if (lane + 64 * k >= K::NC) continue;
        R s = 0;
#pragma unroll
        for (int r = 0; r < K::NR; ++r) {
          const R *Hr = H + g * HSZ + r * 3 * K::NH;
          s = fma(Hr[ox[k]] * Hr[oy[k]], Hr[oz[k]], s);
        }
#pragma unroll
        for (int u = 0; u < QT; ++u)
          if (u == t) acc[u][k] += s;
      }
    }
    wave_sync();
  }
  // degeneracy (direct.hpp:1664-1667, 1793-1802) and the 1/2 of :1849-1850;
  // the quartets overwrite the tables, which the last round's wave_sync
  // has retired
  double *sI = s_buf[wv];
#pragma unroll
  for (int t = 0; t < QT; ++t) {
    if (t >= ntile) break;
    const PairRec &bp = s_pr[wv][t][0];
    const PairRec &kp = s_pr[wv][t][1];
    const bool same_pair = (bp.a == kp.a && bp.b == kp.b);
    const double scale = 0.5 * (bp.a == bp.b ? 1.0 : 2.0) * (kp.a == kp.b ? 1.0 : 2.0) * (same_pair ? 1.0 : 2.0);
#pragma unroll
    for (int k = 0; k < NOWN; ++k)
      if (lane + 64 * k < K::NC) sI[t * K::NC + lane + 64 * k] = (double)acc[t][k] * scale;
  }
  wave_sync();
  digest_tile<K>(sI, s_bf[wv], s_mask[wv], s_mol[wv], ntile, a.planes, lane);
}

// LDS a block may take on gfx950 (160 KB per CU): variants beyond it are
// not built
constexpr int kWaveLdsMax = 160 * 1024;

template <int LA, int LB, int LC, int LD, int QT>
hipError_t launch_wave_cls(const EriLaunch &a, hipStream_t s) {
  using K = Cls<LA, LB, LC, LD>;
  constexpr long long per_block = (long long)kWavesPerBlock * QT;
  const long long blocks = (a.nq + per_block - 1) / per_block;
  if (a.fp32) {
    if constexpr (wave_lds_bytes<K, float, QT>() <= kWaveLdsMax)
      hipLaunchKernelGGL((eri_wave_kernel<LA, LB, LC, LD, float, QT>), dim3((unsigned)blocks),
                         dim3(64 * kWavesPerBlock), 0, s, a);
    else
      return hipErrorInvalidValue;
  } else {
    if constexpr (wave_lds_bytes<K, double, QT>() <= kWaveLdsMax)
      hipLaunchKernelGGL((eri_wave_kernel<LA, LB, LC, LD, double, QT>), dim3((unsigned)blocks),
                         dim3(64 * kWavesPerBlock), 0, s, a);
    else
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// FP64-only variants (the f and g classes): 1 and 2 quartets per wave for
// every class, 4 for those of at most 216 Cartesian components
template <class K, int QT>
constexpr bool fp64_variant_built() {
  return (QT <= 2 || K::NC <= 216) && wave_lds_bytes<K, double, QT>() <= kWaveLdsMax;
}

template <int LA, int LB, int LC, int LD, int QT>
hipError_t launch_wave_fp64_cls(const EriLaunch &a, hipStream_t s) {
  using K = Cls<LA, LB, LC, LD>;
  constexpr long long per_block = (long long)kWavesPerBlock * QT;
  const long long blocks = (a.nq + per_block - 1) / per_block;
  if (a.fp32) return hipErrorNotSupported;
  if constexpr (fp64_variant_built<K, QT>())
    hipLaunchKernelGGL((eri_wave_kernel<LA, LB, LC, LD, double, QT>), dim3((unsigned)blocks),
                       dim3(64 * kWavesPerBlock), 0, s, a);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

}  // namespace

// s/p/d classes with more than 27 Cartesian components
#define CQX_WAVE_CLASSES(X) \
  X(1, 1, 1, 1)             \
  X(2, 0, 1, 1)             \
  X(2, 0, 2, 0)             \
  X(2, 1, 1, 0)             \
  X(2, 1, 1, 1)             \
  X(2, 1, 2, 0)             \
  X(2, 1, 2, 1)             \
  X(2, 2, 0, 0)             \
  X(2, 2, 1, 0)             \
  X(2, 2, 1, 1)             \
  X(2, 2, 2, 0)             \
  X(2, 2, 2, 1)             \
  X(2, 2, 2, 2)

// every wave class with QT quartets per wave
template <int QT>
hipError_t launch_wave_q(const EriLaunch &a, hipStream_t s) {
#define CQX_CASE(A, B, C, D) \
  if (a.la == A && a.lb == B && a.lc == C && a.ld == D) return launch_wave_cls<A, B, C, D, QT>(a, s);
  CQX_WAVE_CLASSES(CQX_CASE)
#undef CQX_CASE
  return hipErrorNotSupported;
}

// classes with an f shell and at most 1080 Cartesian components (the
// larger ones take the generic kernel); FP64 only, since the FP32 lists
// hold s/p/d classes alone
#define CQX_WAVE_F_CLASSES(X)                                                                   \
  X(3, 0, 0, 0) X(3, 0, 1, 0) X(3, 0, 1, 1) X(3, 0, 2, 0) X(3, 0, 2, 1) X(3, 0, 2, 2) X(3, 0, 3, 0) \
  X(3, 1, 0, 0) X(3, 1, 1, 0) X(3, 1, 1, 1) X(3, 1, 2, 0) X(3, 1, 2, 1) X(3, 1, 2, 2) X(3, 1, 3, 0) \
  X(3, 1, 3, 1) X(3, 2, 0, 0) X(3, 2, 1, 0) X(3, 2, 1, 1) X(3, 2, 2, 0) X(3, 2, 2, 1) X(3, 2, 3, 0) \
  X(3, 3, 0, 0) X(3, 3, 1, 0) X(3, 3, 1, 1) X(3, 3, 2, 0) X(3, 3, 3, 0)

// classes with a g shell and at most 1080 Cartesian components, FP64 only
#define CQX_WAVE_G_CLASSES(X)                                                                   \
  X(4, 0, 0, 0) X(4, 0, 1, 0) X(4, 0, 1, 1) X(4, 0, 2, 0) X(4, 0, 2, 1) X(4, 0, 2, 2) X(4, 0, 3, 0) \
  X(4, 0, 3, 1) X(4, 0, 3, 2) X(4, 0, 4, 0) X(4, 1, 0, 0) X(4, 1, 1, 0) X(4, 1, 1, 1) X(4, 1, 2, 0) \
  X(4, 1, 2, 1) X(4, 1, 3, 0) X(4, 1, 4, 0) X(4, 2, 0, 0) X(4, 2, 1, 0) X(4, 2, 1, 1) X(4, 2, 2, 0) \
  X(4, 2, 3, 0) X(4, 3, 0, 0) X(4, 3, 1, 0) X(4, 3, 2, 0) X(4, 4, 0, 0) X(4, 4, 1, 0)

// every f wave class with QT quartets per wave
template <int QT>
hipError_t launch_wave_fq(const EriLaunch &a, hipStream_t s) {
#define CQX_CASE(A, B, C, D) \
  if (a.la == A && a.lb == B && a.lc == C && a.ld == D) return launch_wave_fp64_cls<A, B, C, D, QT>(a, s);
  CQX_WAVE_F_CLASSES(CQX_CASE)
#undef CQX_CASE
  return hipErrorNotSupported;
}

// whether the QT variant of an f class is built (FP64)
template <int QT>
bool wave_fq_built(int la, int lb, int lc, int ld) {
#define CQX_CASE(A, B, C, D) \
  if (la == A && lb == B && lc == C && ld == D) return fp64_variant_built<Cls<A, B, C, D>, QT>();
  CQX_WAVE_F_CLASSES(CQX_CASE)
#undef CQX_CASE
  return false;
}

// the same for the g classes
template <int QT>
hipError_t launch_wave_gq(const EriLaunch &a, hipStream_t s) {
#define CQX_CASE(A, B, C, D) \
  if (a.la == A && a.lb == B && a.lc == C && a.ld == D) return launch_wave_fp64_cls<A, B, C, D, QT>(a, s);
  CQX_WAVE_G_CLASSES(CQX_CASE)
#undef CQX_CASE
  return hipErrorNotSupported;
}

template <int QT>
bool wave_gq_built(int la, int lb, int lc, int ld) {
#define CQX_CASE(A, B, C, D) \
  if (la == A && lb == B && lc == C && ld == D) return fp64_variant_built<Cls<A, B, C, D>, QT>();
  CQX_WAVE_G_CLASSES(CQX_CASE)
#undef CQX_CASE
  return false;
}

// whether the QT variant of an s/p/d class is built, and its LDS per block
template <int QT>
bool wave_q_built(int la, int lb, int lc, int ld, bool fp32) {
#define CQX_CASE(A, B, C, D)                                                                  \
  if (la == A && lb == B && lc == C && ld == D)                                               \
    return (fp32 ? wave_lds_bytes<Cls<A, B, C, D>, float, QT>()                               \
                 : wave_lds_bytes<Cls<A, B, C, D>, double, QT>()) <= kWaveLdsMax;
  CQX_WAVE_CLASSES(CQX_CASE)
#undef CQX_CASE
  return false;
}

}  // namespace cqx
