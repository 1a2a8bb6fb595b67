// Tile ERI + J/K digestion kernels for the lane-group classes with
// Hermitian J and K planes.
//
// A tile is up to 8 bra pairs (a b_i) sharing their first shell a against
// up to 8 ket pairs (c d_j) sharing c, all of one molecule.  Its quartets are
// evaluated as in the lane-group kernels (eri_fast_impl.hpp: G lanes per
// quartet, 64 / G quartets per round) and digested with the reference's
// Hermitian updates (direct.hpp:1858-1917) in the kernel's labels:
//   J(a, b_i) += X(d, c) I    summed over the tile's kets     (per bra i)
//   J(d_j, c) += X(a, b) I    summed over the tile's bras     (per ket j)
//   K(a, c)   += X(d, b) I    summed over the whole tile
//   K(d_j, a) += X(b, c) I    summed over the tile's bras     (per ket j)
//   K(b_i, c) += X(d, a) I    summed over the tile's kets     (per bra i)
//   K(d_j, b_i) += X(a, c) I  one target per quartet
// The five shared targets are summed in LDS (LDS atomics, per wave) and
// added to the accumulators in HBM once per tile; only K(d, b) takes one
// memory-side atomic per quartet, as every update does in the lane-group
// kernels.  Coefficients, conjugation signs and the transposed storage of
// Hermitian targets are those of digest_tile.hpp, applied when a tile's sums
// are added to HBM.
#pragma once

#include <hip/hip_runtime.h>

#include <mutex>

#include "eri_fast_impl.hpp"
#include "tile_layout.hpp"

namespace cqx {

namespace {

__device__ __forceinline__ int log2_pow2(int s) { return 31 - __clz(s); }

// One shared update term of a round's nq quartets into the tile's LDS sums:
// items (quartet, target element, plane) with the plane fastest; SH = 0 one
// sum for the tile, 1 one per bra (local i), 2 one per ket (local j).
template <class K, int AR, int AC, int XR, int XC, int SH>
__device__ __forceinline__ void digest_shared(const double *__restrict__ sI, const int (*__restrict__ sbf)[4],
                                              const unsigned *__restrict__ smask,
                                              const unsigned char *__restrict__ sloc, int nq, const PlaneSet &ps,
                                              const MolPlanes &mp, int kind, double *acc, int lane) {
  constexpr int NAR = K::n(AR), NT = NAR * K::n(AC);
  constexpr int ST[4] = {K::NCB * K::NCC * K::NCD, K::NCC * K::NCD, K::NCD, 1};
  const int np = ps.kcount[kind];
  if (np == 0) return;
  const int p0 = ps.kstart[kind], S = ps.kps[kind], lgp = log2_pow2(S);
  const int N = mp.nbc;
  const double *__restrict__ xk = mp.kx[kind];
  const int items = (nq * NT) << lgp;
  for (int it = lane; it < items; it += 64) {
    const int p = it & (S - 1);
    if (p >= np) continue;
    const int r = it >> lgp;
    const int qi = r / NT, tgt = r - qi * NT;
    const int t1 = tgt / NAR, t0 = tgt - t1 * NAR;
    const int slot = p0 + p;
    const int mat = (int)((ps.mats >> (4 * slot)) & 15ull);
    if (!((smask[qi] >> mat) & 1u)) continue;
    const int *bf = sbf[qi];
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
    const int loc = SH == 0 ? 0 : SH == 1 ? (sloc[qi] >> 3) : (sloc[qi] & 7);
    atomicAdd(acc + ((loc * NT + tgt) << lgp) + p, s0 + s1);
  }
}

// Add one shared term's LDS sums to the kind's accumulators in HBM, with the
// coefficient, conjugation sign and Hermitian transposition of
// digest_term_items.  tbf[0][i] = (first functions of a, b_i), tbf[1][j] =
// (of c, d_j).
template <class K, int AR, int AC, int SH>
__device__ __forceinline__ void flush_shared(const double *acc, const int (*tbf)[kTileDim][2], const PlaneSet &ps,
                                             const MolPlanes &mp, int kind, double base_coef, bool neg_ok,
                                             int lane) {
  constexpr int NAR = K::n(AR), NAC = K::n(AC), NT = NAR * NAC;
  constexpr bool TR = NAC > NAR;  // Hermitian target: the longer dimension down a column
  const int np = ps.kcount[kind];
  if (np == 0) return;
  const int p0 = ps.kstart[kind], S = ps.kps[kind], lgp = log2_pow2(S);
  const int nloc = SH == 0 ? 1 : kTileDim;
  const int N = mp.nbc;
  double *ak = mp.ka[kind];
  for (int e = lane; e < (nloc * NT) << lgp; e += 64) {
    const int p = e & (S - 1);
    if (p >= np) continue;
    const double v = acc[e];
    if (v == 0.0) continue;  // also every local beyond the tile's bras or kets
    const int rest = e >> lgp;
    const int tgt = rest % NT, loc = rest / NT;
    const int t1 = tgt / NAR, t0 = tgt - t1 * NAR;
    auto role_bf = [&](int role) {
      return role == 0 ? tbf[0][0][0] : role == 1 ? tbf[0][loc][1] : role == 2 ? tbf[1][0][0] : tbf[1][loc][1];
    };
    const int slot = p0 + p;
    const bool neg = (ps.negs >> slot) & 1u;
    const double coef = ((neg_ok && neg) != (TR && neg)) ? -base_coef : base_coef;
    const size_t el = TR ? ((size_t)(role_bf(AC) + t1) + (size_t)(role_bf(AR) + t0) * N)
                         : ((size_t)(role_bf(AR) + t0) + (size_t)(role_bf(AC) + t1) * N);
    unsafeAtomicAdd(ak + p + el * S, coef * v);
  }
}

template <int LA, int LB, int LC, int LD, int G, class R>
__global__ __launch_bounds__(256) void eri_tile_kernel(TileEriLaunch a) {
  using K = Cls<LA, LB, LC, LD>;
  constexpr int QPW = 64 / G;  // quartets per round of a wave
  extern __shared__ double s_acc[];  // 4 waves x the tile's accumulators
  __shared__ double s_I[4][QPW * K::NC];
  __shared__ int s_bf[4][QPW][4];
  __shared__ unsigned s_mask[4][QPW];
  __shared__ int s_mol[4][QPW];
  __shared__ unsigned char s_loc[4][QPW];
  __shared__ unsigned char s_sidx[4][64];
  __shared__ int s_tbf[4][2][kTileDim][2];
  __shared__ R s_tab[rys_tab_len(K::NR)];
  const R *__restrict__ rysn = rys_section<K::NR, R>(a.rys, a.rys32, a.rys_lds, s_tab);
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int g = lane & (G - 1), qw = lane / G;
  const long long t = (long long)blockIdx.x * 4 + wv;
  if (t >= a.ntile) return;  // uniform across the wave; no block barriers below
  const Tile *__restrict__ T = a.tiles + t;
  const unsigned long long bits = T->bits;
  if (!bits) return;  // an empty tile (uniform across the wave)
  const int bm0 = T->bra_m0, km0 = T->ket_m0, nb = T->nb, nk = T->nk;
  const int nsurv = __popcll(bits);
  // survivors in order: s_sidx[k] = local index i * 8 + j of the k-th
  if ((bits >> lane) & 1ull) s_sidx[wv][__popcll(bits & ((1ull << lane) - 1ull))] = (unsigned char)lane;
  if (lane < nb) {
    const PairRec &P = a.bra[a.bmem[bm0 + lane]];
    s_tbf[wv][0][lane][0] = P.bfa;
    s_tbf[wv][0][lane][1] = P.bfb;
  } else if (lane >= 32 && lane - 32 < nk) {
    const PairRec &P = a.ket[a.kmem[km0 + lane - 32]];
    s_tbf[wv][1][lane - 32][0] = P.bfa;
    s_tbf[wv][1][lane - 32][1] = P.bfb;
  }
  const PlaneSet &ps = a.planes;
  const int mol = a.bra[a.bmem[bm0]].mol;
  const MolPlanes &mp = ps.nmol == 1 ? ps.m0 : ps.mols[mol];
  const int sj = ps.kcount[kJHer] ? ps.kps[kJHer] : 0, sk = ps.kcount[kKHer] ? ps.kps[kKHer] : 0;
  const TileAccLayout L = tile_acc_layout(K::NCA, K::NCB, K::NCC, K::NCD, sj, sk);
  double *acc = s_acc + wv * L.total;
  for (int e = lane; e < L.total; e += 64) acc[e] = 0.0;
  const int npK = ps.kcount[kKHer];
  const int lgpK = npK <= 1 ? 0 : npK <= 2 ? 1 : npK <= 4 ? 2 : npK <= 8 ? 3 : 4;
  wave_sync();
  for (int r0 = 0; r0 < nsurv; r0 += QPW) {
    const int q = r0 + qw;
    const bool valid = q < nsurv;
    int loc = 0;
    PairRec bp{}, kp{};
    if (valid) {
      loc = s_sidx[wv][q];
      bp = a.bra[a.bmem[bm0 + (loc >> 3)]];
      kp = a.ket[a.kmem[km0 + (loc & 7)]];
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
      s_mask[wv][qw] = valid ? (unsigned)T->mask[loc] : 0u;
      s_mol[wv][qw] = mol;
      s_loc[wv][qw] = (unsigned char)loc;
    }
    wave_sync();
    const int nq = nsurv - r0 < QPW ? nsurv - r0 : QPW;
    digest_shared<K, 0, 1, 3, 2, 1>(s_I[wv], s_bf[wv], s_mask[wv], s_loc[wv], nq, ps, mp, kJHer, acc + L.jab, lane);
    digest_shared<K, 3, 2, 0, 1, 2>(s_I[wv], s_bf[wv], s_mask[wv], s_loc[wv], nq, ps, mp, kJHer, acc + L.jdc, lane);
    digest_shared<K, 0, 2, 3, 1, 0>(s_I[wv], s_bf[wv], s_mask[wv], s_loc[wv], nq, ps, mp, kKHer, acc + L.kac, lane);
    digest_shared<K, 3, 0, 1, 2, 2>(s_I[wv], s_bf[wv], s_mask[wv], s_loc[wv], nq, ps, mp, kKHer, acc + L.kda, lane);
    digest_shared<K, 1, 2, 3, 0, 1>(s_I[wv], s_bf[wv], s_mask[wv], s_loc[wv], nq, ps, mp, kKHer, acc + L.kbc, lane);
    if (npK)  // K(d_j, b_i): one target per quartet
      digest_term_tile<K, 3, 1, 0, 2, true>(s_I[wv], s_bf[wv], s_mask[wv], s_mol[wv], nq, ps, kKHer,
                                             ps.kstart[kKHer], npK, lgpK, 0.5, true, false, lane);
    wave_sync();  // the next round rewrites the staged quartets
  }
  // the tile's sums to HBM
  flush_shared<K, 0, 1, 1>(acc + L.jab, s_tbf[wv], ps, mp, kJHer, 1.0, false, lane);
  flush_shared<K, 3, 2, 2>(acc + L.jdc, s_tbf[wv], ps, mp, kJHer, 1.0, false, lane);
  flush_shared<K, 0, 2, 0>(acc + L.kac, s_tbf[wv], ps, mp, kKHer, 0.5, true, lane);
  flush_shared<K, 3, 0, 2>(acc + L.kda, s_tbf[wv], ps, mp, kKHer, 0.5, true, lane);
  flush_shared<K, 1, 2, 1>(acc + L.kbc, s_tbf[wv], ps, mp, kKHer, 0.5, true, lane);
}

template <int LA, int LB, int LC, int LD, int G, class R>
hipError_t launch_tile_cls(const TileEriLaunch &a, hipStream_t s) {
  using K = Cls<LA, LB, LC, LD>;
  auto kern = eri_tile_kernel<LA, LB, LC, LD, G, R>;
  const int sj = a.planes.kcount[kJHer] ? a.planes.kps[kJHer] : 0;
  const int sk = a.planes.kcount[kKHer] ? a.planes.kps[kKHer] : 0;
  const size_t dyn = 4 * sizeof(double) * (size_t)tile_acc_layout(K::NCA, K::NCB, K::NCC, K::NCD, sj, sk).total;
  static std::once_flag once;  // allow dynamic LDS beyond the default limit, once per instantiation
  static hipError_t attr = hipSuccess;
  std::call_once(once, [&] {
    attr = hipFuncSetAttribute(reinterpret_cast<const void *>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                               96 * 1024);
  });
  if (attr != hipSuccess) return attr;
  if (dyn > 96 * 1024) return hipErrorInvalidValue;
  hipLaunchKernelGGL(kern, dim3((unsigned)((a.ntile + 3) / 4)), dim3(256), dyn, s, a);
  return hipGetLastError();
}

}  // namespace

// all tile classes for one lane-group size G and precision R
template <int G, class R>
hipError_t launch_tile_g(const TileEriLaunch &a, hipStream_t s) {
#define CQX_CASE(A, B, C, D) \
  if (a.la == A && a.lb == B && a.lc == C && a.ld == D) return launch_tile_cls<A, B, C, D, G, R>(a, s);
  CQX_FAST_CLASSES(CQX_CASE)
#undef CQX_CASE
  return hipErrorNotSupported;
}

}  // namespace cqx
