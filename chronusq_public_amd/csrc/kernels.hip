// HIP kernels of the direct J/K build for gfx950 (MI355X): screening and
// quartet-list construction, the generic Rys ERI + digestion kernel, and the
// matrix transforms around them.
//
// Reference algorithm: GTODirectTPIContraction::directScaffoldNew
// (include/particleintegrals/contract/direct.hpp:1412-2123).  The quartet set,
// its screening and the degeneracy factors follow the reference loop; ERIs
// are evaluated by Rys quadrature in Cartesian form and the density /
// result matrices are carried in the Cartesian AO basis (T X T^T and
// T^T A T around the build), which is the same algebra as transforming every
// integral batch to pure functions.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "digest.hpp"
#include "kernels.hpp"
#include "rys.hpp"

namespace cqx {

namespace {

constexpr double kTwoPi52 = 34.98683665524972;  // 2 pi^(5/2)

__device__ __forceinline__ int ncart_d(int l) { return (l + 1) * (l + 2) / 2; }

// Cartesian exponents of component c of an l shell (lx descending, then ly).
__device__ __forceinline__ void cart_exp(int l, int c, int &lx, int &ly, int &lz) {
  int k = c;
  lx = l;
  while (k >= l - lx + 1) {
    k -= l - lx + 1;
    --lx;
  }
  lz = k;
  ly = l - lx - lz;
}

__device__ __forceinline__ void atomic_add(double *p, double v) { unsafeAtomicAdd(p, v); }

// ---------------------------------------------------------------------------
// Generic Rys ERI kernel: one 64-lane wave per shell quartet.
// Per primitive quartet, lanes r < NR fetch the Rys nodes/weights, lanes
// (r, dir) < 3 NR build the 2D integrals I_dir(ia, ib, ic, id) (vertical
// recursion on the Rys variables, then horizontal transfer to b and d) into
// LDS, and every lane accumulates its share of the Cartesian quartet in LDS.
// Quartets too large for LDS (classes with g shells) are built in passes
// over ranges of the first shell's components (a is the slowest index of
// the quartet layout, so a range is a smaller quartet with a's first
// function shifted) and with the 2D tables of a few roots at a time; each
// pass is digested (or stored) on its own, since every J/K update is a sum
// of terms linear in the integrals.
// ---------------------------------------------------------------------------
struct TeamDims {
  int la, lb, lc, ld, nca, ncb, ncc, ncd, nc, ni, nk, nr, nh;
};

__device__ TeamDims team_dims(int la, int lb, int lc, int ld) {
  TeamDims d;
  d.la = la; d.lb = lb; d.lc = lc; d.ld = ld;
  d.nca = ncart_d(la); d.ncb = ncart_d(lb); d.ncc = ncart_d(lc); d.ncd = ncart_d(ld);
  d.nc = d.nca * d.ncb * d.ncc * d.ncd;
  d.ni = la + lb + 1;
  d.nk = lc + ld + 1;
  d.nr = (la + lb + lc + ld) / 2 + 1;
  d.nh = (la + 1) * (lb + 1) * (lc + 1) * (ld + 1);
  return d;
}

constexpr int kTeamMaxI = 2 * kMaxLDev + 1;  // la + lb + 1

// 2D integrals of one Rys root along one direction, written to H[nh].
__device__ void rys_2d(const TeamDims &D, double t2, double p, double q, double PAd,
                       double QCd, double PQd, double ABd, double CDd, double scale,
                       double *__restrict__ H) {
  double G[kTeamMaxI][kTeamMaxI];                 // [i <= la+lb][k <= lc+ld]
  double E[kTeamMaxI][kMaxLDev + 1][kTeamMaxI];   // bra transfer: [i][j][k]
  const double pq = p + q, rho = p * q / pq;
  const double B00 = 0.5 * t2 / pq;
  const double B10 = 0.5 / p * (1.0 - rho / p * t2);
  const double B01 = 0.5 / q * (1.0 - rho / q * t2);
  const double C00 = PAd - (q / pq) * t2 * PQd;
  const double D00 = QCd + (p / pq) * t2 * PQd;
  G[0][0] = scale;
  if (D.ni > 1) G[1][0] = C00 * scale;
  for (int i = 1; i + 1 < D.ni; ++i) G[i + 1][0] = C00 * G[i][0] + i * B10 * G[i - 1][0];
  for (int k = 0; k + 1 < D.nk; ++k) {
    G[0][k + 1] = D00 * G[0][k] + (k > 0 ? k * B01 * G[0][k - 1] : 0.0);
    for (int i = 1; i < D.ni; ++i)
      G[i][k + 1] = D00 * G[i][k] + (k > 0 ? k * B01 * G[i][k - 1] : 0.0) + i * B00 * G[i - 1][k];
  }
  // horizontal transfer on the bra: (i, j+1) = (i+1, j) + AB (i, j)
  for (int i = 0; i < D.ni; ++i)
    for (int k = 0; k < D.nk; ++k) E[i][0][k] = G[i][k];
  for (int j = 1; j <= D.lb; ++j)
    for (int i = 0; i + j < D.ni; ++i)
      for (int k = 0; k < D.nk; ++k) E[i][j][k] = E[i + 1][j - 1][k] + ABd * E[i][j - 1][k];
  // horizontal transfer on the ket for every (ia, ib)
  for (int ia = 0; ia <= D.la; ++ia)
    for (int ib = 0; ib <= D.lb; ++ib) {
      double F[kTeamMaxI][kMaxLDev + 1];
      for (int k = 0; k < D.nk; ++k) F[k][0] = E[ia][ib][k];
      for (int l = 1; l <= D.ld; ++l)
        for (int k = 0; k + l < D.nk; ++k) F[k][l] = F[k + 1][l - 1] + CDd * F[k][l - 1];
      for (int ic = 0; ic <= D.lc; ++ic)
        for (int id = 0; id <= D.ld; ++id)
          H[((ia * (D.lb + 1) + ib) * (D.lc + 1) + ic) * (D.ld + 1) + id] = F[ic][id];
    }
}

// LDS of the generic kernel, in doubles: Rys nodes and weights of up to 16
// roots, the 2D tables of rpp roots and the quartet components of ca
// components of the first shell
__host__ __device__ constexpr int team_lds_doubles(int rpp, int nh, int ca, int nbcd) {
  return 32 + rpp * 3 * nh + ca * nbcd;
}

__global__ __launch_bounds__(64) void eri_team_kernel(EriLaunch a) {
  extern __shared__ double lds[];
  const int lane = threadIdx.x;
  const TeamDims D = team_dims(a.la, a.lb, a.lc, a.ld);
  const int CA = a.team_ca, RPP = a.team_rpp;  // first-shell components per pass, roots per table pass
  const int nbcd = D.ncb * D.ncc * D.ncd;
  double *s_x = lds;
  double *s_w = lds + 16;
  double *s_H = lds + 32;
  double *s_I = s_H + RPP * 3 * D.nh;
  // quartet count: the list's device counter for a launch sized to the
  // list's capacity before the list was built
  const long long nq = a.nq_dev ? min((long long)*a.nq_dev - a.nq_base, a.nq) : a.nq;
  for (long long q = blockIdx.x; q < nq; q += gridDim.x) {
    const Quartet qt = a.quartets[q];
    const PairRec bp = a.bra[qt.bra];
    const PairRec kp = a.ket[qt.ket_mask & 0xFFFFFFu];
    for (int ia0 = 0; ia0 < D.nca; ia0 += CA) {
      const int na = D.nca - ia0 < CA ? D.nca - ia0 : CA;
      const int nsub = na * nbcd;
      for (int o = lane; o < nsub; o += 64) s_I[o] = 0.0;
      for (int i = 0; i < bp.npp; ++i) {
        const PrimPair A = a.prims[bp.pp_off + i];
        for (int j = 0; j < kp.npp; ++j) {
          const PrimPair B = a.prims[kp.pp_off + j];
          const double pq = A.p + B.p, rho = A.p * B.p / pq;
          const double PQ[3] = {A.P[0] - B.P[0], A.P[1] - B.P[1], A.P[2] - B.P[2]};
          const double T = rho * (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]);
          const double pref = A.K * B.K * kTwoPi52 / (A.p * B.p * sqrt(pq));
          __syncthreads();
          if (lane < D.nr) {
            double x, w;
            rys_root_rt(D.nr, lane, T, a.rys, x, w);
            s_x[lane] = x;
            s_w[lane] = w * pref;
          }
          for (int r0 = 0; r0 < D.nr; r0 += RPP) {
            const int nrp = D.nr - r0 < RPP ? D.nr - r0 : RPP;
            __syncthreads();
            if (lane < 3 * nrp) {
              const int lr = lane / 3, dir = lane % 3, r = r0 + lr;
              rys_2d(D, s_x[r], A.p, B.p, A.PA[dir], B.PA[dir], PQ[dir], bp.AB[dir], kp.AB[dir],
                     dir == 2 ? s_w[r] : 1.0, s_H + (lr * 3 + dir) * D.nh);
            }
            __syncthreads();
            for (int o = lane; o < nsub; o += 64) {
              int rem = o;
              const int id = rem % D.ncd; rem /= D.ncd;
              const int ic = rem % D.ncc; rem /= D.ncc;
              const int ib = rem % D.ncb; rem /= D.ncb;
              const int ia = ia0 + rem;
              int e[4][3];
              cart_exp(D.la, ia, e[0][0], e[0][1], e[0][2]);
              cart_exp(D.lb, ib, e[1][0], e[1][1], e[1][2]);
              cart_exp(D.lc, ic, e[2][0], e[2][1], e[2][2]);
              cart_exp(D.ld, id, e[3][0], e[3][1], e[3][2]);
              int h[3];
              for (int dir = 0; dir < 3; ++dir)
                h[dir] = ((e[0][dir] * (D.lb + 1) + e[1][dir]) * (D.lc + 1) + e[2][dir]) * (D.ld + 1) +
                         e[3][dir];
              double sum = 0.0;
              for (int lr = 0; lr < nrp; ++lr) {
                const double *Hr = s_H + lr * 3 * D.nh;
                sum += Hr[h[0]] * Hr[D.nh + h[1]] * Hr[2 * D.nh + h[2]];
              }
              s_I[o] += sum;
            }
          }
        }
      }
      __syncthreads();
      if (a.store_mode) {
        double *out = a.store + a.store_off[q] + (size_t)ia0 * nbcd;
        for (int o = lane; o < nsub; o += 64) out[o] = s_I[o];
      } else {
        // degeneracy of the quartet in the 8-fold loop (direct.hpp:1664-1667,
        // 1793-1802), times the 1/2 of direct.hpp:1849-1850
        const bool same_pair = (bp.a == kp.a && bp.b == kp.b);
        const double deg = (bp.a == bp.b ? 1.0 : 2.0) * (kp.a == kp.b ? 1.0 : 2.0) * (same_pair ? 1.0 : 2.0);
        for (int o = lane; o < nsub; o += 64) s_I[o] *= 0.5 * deg;
        __syncthreads();
        // the pass is the quartet of a's components ia0 .. ia0 + na - 1
        const int n[4] = {na, D.ncb, D.ncc, D.ncd};
        const int bf[4] = {bp.bfa + ia0, bp.bfb, kp.bfa, kp.bfb};
        digest_wave(s_I, n, bf, a.planes, qt.ket_mask >> 24, bp.mol, lane, 64);
      }
      __syncthreads();
    }
  }
}

// ---------------------------------------------------------------------------
// Screening: quartet candidates (bra i, ket j) after the Schwarz pre-cut,
// tested with the reference's Schwarz x shell-block-norm rule in reference
// labels (direct.hpp:1689-1789); survivors appended to a list with the mask
// of contractions that pass their own separated screen.
// ---------------------------------------------------------------------------
constexpr int kListThreads = 1024;

// The reference's screen of one candidate quartet (bra pair P with its own
// block norms pb, ket pair Q with pk): the Schwarz product against the max
// of the six shell-block norms in reference labels (the pair the reference
// loop visits later is the bra), globally and per matrix (separated
// screen).  Returns whether any contraction survives and their mask.
struct ScreenArgs {
  const double *Bg, *Bm;
  const long long *mol_boff;
  const int *mol_ns;
  int nmat;
  double thresh;
  int screen;
};
__device__ __forceinline__ bool screen_candidate(const ScreenArgs &a, const PairScr &P, const PairScr &Q,
                                                 const double *pb, const double *pk, double qq, unsigned &mask) {
  mask = 0;
  const int ps1 = (int)(P.s12 & 0xFFFFu), ps2 = (int)(P.s12 >> 16);
  const int qs1 = (int)(Q.s12 & 0xFFFFu), qs2 = (int)(Q.s12 >> 16);
  const bool p_first = (ps1 > qs1) || (ps1 == qs1 && ps2 >= qs2);
  const int s1 = p_first ? ps1 : qs1, s2 = p_first ? ps2 : qs2;
  const int s3 = p_first ? qs1 : ps1, s4 = p_first ? qs2 : ps2;
  // lower bounds from the pairs' own block norms: enough to accept
  if (!a.screen) {
    mask = (1u << a.nmat) - 1u;
    return true;
  }
  if (!(fmax(pb[0], pk[0]) * qq < a.thresh)) {
    if (a.nmat == 1) {
      mask = 1u;
      return true;
    }
    unsigned lm = 0;
    for (int m = 0; m < a.nmat; ++m)
      if (!(fmax(pb[1 + m], pk[1 + m]) * qq < a.thresh)) lm |= 1u << m;
    if (lm == (1u << a.nmat) - 1u) {
      mask = lm;
      return true;
    }
  }
  const long long boff = a.mol_boff[P.mol];
  const int ns = a.mol_ns[P.mol];
  auto bmax = [&](const double *B, int ld, int m) {
    auto at = [&](int x, int y) { return B[(boff + x + (size_t)y * ns) * ld + m]; };
    double v = fmax(at(s1, s2), fmax(at(s1, s3), at(s2, s3)));
    v = fmax(v, at(s3, s4));
    v = fmax(v, fmax(at(s1, s4), at(s2, s4)));
    return v;
  };
  if (bmax(a.Bg, 1, 0) * qq < a.thresh) return false;
  if (a.nmat == 1) {
    mask = 1u;
  } else {
    // per-matrix norms interleaved: the nmat values of one block share a line
    for (int m = 0; m < a.nmat; ++m)
      if (!(bmax(a.Bm, a.nmat, m) * qq < a.thresh)) mask |= 1u << m;
  }
  return mask != 0u;
}

// Candidates [base, base + 1024) of one list: the block's bra range is found
// once, each thread tests one candidate, survivors are appended block-wide
// (the waves' counts meet in LDS and one thread reserves the block's slots
// with one atomic per list), and the Ka Kb Kc Kd sums of the chunk are
// reduced in LDS to one atomic per list.
__device__ __forceinline__ void list_chunk(const ListLaunch &a, long long base) {
  constexpr int NW = kListThreads / 64;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  __shared__ int s_lo, s_hi;
  __shared__ unsigned s_cnt[2][NW];
  __shared__ unsigned long long s_start[2][NW];
  __shared__ double s_ps[2][NW];
  // candidates of the list: counted on the device for a list sized in
  // advance (ncand is then its capacity)
  const long long ncand = a.ncand_dev ? min(*a.ncand_dev - a.cand_base, a.ncand) : a.ncand;
  if (base >= ncand) return;  // uniform across the block
  const long long t = base + threadIdx.x;
  bool pass = false;
  unsigned mask = 0;
  int bi = 0, kj = 0;
  double qq = 0.0, kk4 = 0.0;
  // the block's candidates span bra indices [s_lo, s_hi]: two searches per
  // block, then each thread searches that (usually short) range
  if (threadIdx.x == 0 || threadIdx.x == blockDim.x - 1) {
    const long long tt =
        a.cand_base + (threadIdx.x == 0 ? base : min(base + (long long)blockDim.x, ncand) - 1);
    int lo = 0, hi = a.nbra - 1;
    while (lo < hi) {
      int mid = (lo + hi + 1) >> 1;
      if (a.cand_off[mid] <= tt) lo = mid; else hi = mid - 1;
    }
    if (threadIdx.x == 0) s_lo = lo; else s_hi = lo;
  }
  __syncthreads();
  const int blo = s_lo, bhi = s_hi;
  if (t < ncand) {
    // bra index: last i with cand_off[i] <= t (absolute candidate index)
    const long long ta = a.cand_base + t;
    int lo = blo, hi = bhi;
    while (lo < hi) {
      int mid = (lo + hi + 1) >> 1;
      if (a.cand_off[mid] <= ta) lo = mid; else hi = mid - 1;
    }
    bi = lo;
    kj = a.ket_lo[bi] + (int)(ta - a.cand_off[lo]);
    const PairScr P = a.bra[bi], Q = a.ket[kj];
    qq = P.Q * Q.Q;
    kk4 = (double)P.kk * (double)Q.kk;
    const double *pb = a.bra_pn + (size_t)bi * (a.nmat + 1), *pk = a.ket_pn + (size_t)kj * (a.nmat + 1);
    const ScreenArgs sa{a.Bg, a.Bm, a.mol_boff, a.mol_ns, a.nmat, a.thresh, a.screen};
    pass = screen_candidate(sa, P, Q, pb, pk, qq, mask);
  }
  // precision split on the Schwarz product (the FP32 list when below cut32)
  const bool lo32 = pass && a.cut32 > 0.0 && qq < a.cut32;
  const bool hi64 = pass && !lo32;
  double p64 = hi64 ? kk4 : 0.0, p32 = lo32 ? kk4 : 0.0;
  for (int off = 32; off > 0; off >>= 1) {
    p64 += __shfl_xor(p64, off);
    p32 += __shfl_xor(p32, off);
  }
  const unsigned long long bal[2] = {__ballot(hi64), __ballot(lo32)};
  if (lane == 0) {
    s_cnt[0][wv] = (unsigned)__popcll(bal[0]);
    s_cnt[1][wv] = (unsigned)__popcll(bal[1]);
    s_ps[0][wv] = p64;
    s_ps[1][wv] = p32;
  }
  __syncthreads();
  if (threadIdx.x < 2) {  // thread l reserves list l's slots for the whole block
    const int l = threadIdx.x;
    unsigned long long tot = 0;
    double ps = 0.0;
    for (int w = 0; w < NW; ++w) {
      tot += s_cnt[l][w];
      ps += s_ps[l][w];
    }
    unsigned long long st = 0;
    if (tot) {
      st = atomicAdd(l == 0 ? a.count : a.count32, tot);
      atomic_add(l == 0 ? a.prim_sum : a.prim_sum32, ps);
    }
    for (int w = 0; w < NW; ++w) {
      s_start[l][w] = st;
      st += s_cnt[l][w];
    }
  }
  __syncthreads();
  const bool take[2] = {hi64, lo32};
  for (int l = 0; l < 2; ++l) {
    if (!take[l]) continue;
    const unsigned long long slot = s_start[l][wv] + __popcll(bal[l] & ((1ull << lane) - 1ull));
    if ((long long)slot < (l == 0 ? a.cap : a.cap32)) {
      Quartet qv;
      qv.bra = (uint32_t)bi;
      qv.ket_mask = (uint32_t)kj | (mask << 24);
      (l == 0 ? a.out : a.out32)[slot] = qv;
    }
  }
  __syncthreads();  // the LDS above is rewritten by the next chunk
}

// Every list of a batch in one launch: list k covers the 1024-aligned
// chunks [start[k], start[k + 1]) of a common chunk space, so a chunk never
// spans two lists.
__global__ __launch_bounds__(kListThreads) void quartet_lists_kernel(const ListLaunch *lists, const long long *start,
                                                                     int nlist, long long nchunk) {
  __shared__ int s_k;
  for (long long c = blockIdx.x; c < nchunk; c += gridDim.x) {
    if (threadIdx.x == 0) {
      int lo = 0, hi = nlist - 1;  // last k with start[k] <= c
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= c) lo = mid; else hi = mid - 1;
      }
      s_k = lo;
    }
    __syncthreads();
    const int k = s_k;
    __syncthreads();
    list_chunk(lists[k], (c - start[k]) * kListThreads);
  }
}

// Tile lists: one block of kTileListWaves waves per chunk of kTileChunk
// sub-blocks of one list, each wave taking its sub-blocks one after the
// other.  Lane l tests candidate (i, j) = (l / 8, l % 8) of a sub-block
// with the quartet lists' screen; the survivors of each precision make the
// sub-block's tile (a bitmap and the masks) at the sub-block's own slot.
// The block's counts meet in LDS and take one atomic per counter.
constexpr int kTileListWaves = 4;
__global__ __launch_bounds__(64 * kTileListWaves) void tile_lists_kernel(const TileListLaunch *lists,
                                                                         const long long *start, int nlist,
                                                                         long long nchunk) {
  __shared__ unsigned long long s_cnt[kTileListWaves][3];
  __shared__ double s_ps[kTileListWaves][2];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (long long ch = blockIdx.x; ch < nchunk; ch += gridDim.x) {
    const long long t0 = ch * kTileChunk;
    int lo = 0, hi = nlist - 1;  // the chunk's list: last k with start[k] <= t0
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (start[mid] <= t0) lo = mid; else hi = mid - 1;
    }
    const TileListLaunch &a = lists[lo];
    const long long base = t0 - start[lo];
    unsigned long long ncand = 0, nq64 = 0, nq32 = 0;  // lane 0's running sums
    double ps64 = 0.0, ps32 = 0.0;
    for (int u = wv; u < kTileChunk; u += kTileListWaves) {
      const long long slot = base + u;
      const SubBlock sb = a.sub[slot];
      const int b0 = a.bmem[sb.bra_m0];
      const int mol = a.bra[b0].mol;
      // Schwarz pre-cut of the whole sub-block (uniform across the wave);
      // an empty sub-block has no candidates
      const bool live = sb.nb > 0 && !(a.screen && sb.qq * a.molmax[mol] < a.thresh);
      const int i = lane >> 3, j = lane & 7;
      const bool cand = live && i < sb.nb && j < sb.nk && (!sb.diag || i >= j);
      bool pass = false;
      unsigned mask = 0;
      double qq = 0.0, kk4 = 0.0;
      if (cand) {
        const int bi = a.bmem[sb.bra_m0 + i], kj = a.kmem[sb.ket_m0 + j];
        const PairScr P = a.bra[bi], Q = a.ket[kj];
        qq = P.Q * Q.Q;
        kk4 = (double)P.kk * (double)Q.kk;
        const ScreenArgs sa{a.Bg, a.Bm, a.mol_boff, a.mol_ns, a.nmat, a.thresh, a.screen};
        pass = screen_candidate(sa, P, Q, a.bra_pn + (size_t)bi * (a.nmat + 1),
                                a.ket_pn + (size_t)kj * (a.nmat + 1), qq, mask);
      }
      const bool lo32 = pass && a.cut32 > 0.0 && qq < a.cut32;
      const bool hi64 = pass && !lo32;
      const unsigned long long cbits = __ballot(cand);
      for (int l = 0; l < (a.out32 ? 2 : 1); ++l) {
        const bool take = l == 0 ? hi64 : lo32;
        const unsigned long long bits = __ballot(take);
        double ps = take ? kk4 : 0.0;
        for (int off = 32; off > 0; off >>= 1) ps += __shfl_xor(ps, off);
        Tile *T = (l == 0 ? a.out : a.out32) + slot;
        if (bits) T->mask[lane] = (unsigned char)(take ? mask : 0u);
        if (lane == 0) {
          T->bra_m0 = sb.bra_m0;
          T->ket_m0 = sb.ket_m0;
          T->nb = sb.nb;
          T->nk = sb.nk;
          T->pad = 0u;
          T->bits = bits;  // written for every sub-block: an empty tile is skipped
          T->pad2 = 0ull;
          if (l == 0) {
            nq64 += (unsigned long long)__popcll(bits);
            ps64 += ps;
          } else {
            nq32 += (unsigned long long)__popcll(bits);
            ps32 += ps;
          }
        }
      }
      if (lane == 0) ncand += (unsigned long long)__popcll(cbits);
    }
    if (lane == 0) {
      s_cnt[wv][0] = ncand;
      s_cnt[wv][1] = nq64;
      s_cnt[wv][2] = nq32;
      s_ps[wv][0] = ps64;
      s_ps[wv][1] = ps32;
    }
    __syncthreads();
    if (threadIdx.x < 5) {
      const int k = threadIdx.x;
      if (k < 3) {
        unsigned long long v = 0;
        for (int w = 0; w < kTileListWaves; ++w) v += s_cnt[w][k];
        unsigned long long *dst = k == 0 ? a.ncand : k == 1 ? a.nquart : a.nquart32;
        if (v && dst) atomicAdd(dst, v);
      } else {
        double v = 0.0;
        for (int w = 0; w < kTileListWaves; ++w) v += s_ps[w][k - 3];
        double *dst = k == 3 ? a.prim_sum : a.prim_sum32;
        if (v != 0.0 && dst) atomic_add(dst, v);
      }
    }
    __syncthreads();  // the LDS above is rewritten by the next chunk
  }
}

// ---------------------------------------------------------------------------
// ShellBlockNorm (direct.hpp:73-105): B(s1, s2) = max_i sum_j |X(i, j)| over
// the (s1, s2) block.
// ---------------------------------------------------------------------------
__global__ void block_norm_kernel(const double *X, int nmat, int cplx, int nb, int ns,
                                  const int *sh_bf, int bf0, const int *sh_nf, double *Bm) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long tot = (long long)nmat * ns * ns;
  if (t >= tot) return;
  // Bm[(s1 + s2 NS) nmat + m]
  const int m = (int)(t % nmat);
  const int rem = (int)(t / nmat);
  const int s1 = rem % ns, s2 = rem / ns;
  const size_t nb2 = (size_t)nb * nb;
  const double *Xm = X + (size_t)m * nb2 * (cplx ? 2 : 1);
  double mx = 0.0;
  for (int i = 0; i < sh_nf[s1]; ++i) {
    double row = 0.0;
    for (int j = 0; j < sh_nf[s2]; ++j) {
      const size_t idx = (sh_bf[s1] - bf0 + i) + (size_t)(sh_bf[s2] - bf0 + j) * nb;
      row += cplx ? hypot(Xm[2 * idx], Xm[2 * idx + 1]) : fabs(Xm[idx]);
    }
    mx = fmax(mx, row);
  }
  Bm[t] = mx;
}

__global__ void pair_norm_kernel(const PairScr *pairs, int npair, const double *Bg, const double *Bm,
                                 const long long *mol_boff, const int *mol_ns, int nmat, double *pn) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npair) return;
  const PairScr P = pairs[i];
  const int s1 = (int)(P.s12 & 0xFFFFu), s2 = (int)(P.s12 >> 16);
  const size_t b = (size_t)mol_boff[P.mol] + s1 + (size_t)s2 * mol_ns[P.mol];
  pn[(size_t)i * (nmat + 1)] = Bg[b];
  for (int m = 0; m < nmat; ++m) pn[(size_t)i * (nmat + 1) + 1 + m] = Bm[b * nmat + m];
}

__global__ void block_norm_finish_kernel(int nmat, int ns, const double *Bm, double *Bg) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)ns * ns) return;
  const int k = (int)(t % ns), l = (int)(t / ns);
  // per-matrix norms are already symmetrised (block_norm_symm_kernel)
  double g = 0.0;
  for (int m = 0; m < nmat; ++m) g = fmax(g, Bm[(k + (size_t)l * ns) * nmat + m]);
  Bg[k + (size_t)l * ns] = g;
}

__global__ void block_norm_symm_kernel(int nmat, int ns, const int *her, double *Bm) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)nmat * ns * ns) return;
  const int m = (int)(t % nmat);
  if (her[m]) return;
  const int rem = (int)(t / nmat);
  const int k = rem % ns, l = rem / ns;
  if (k <= l) return;
  // symmetrised for non-Hermitian X, direct.hpp:1537-1545
  double *B1 = Bm + (k + (size_t)l * ns) * nmat + m, *B2 = Bm + (l + (size_t)k * ns) * nmat + m;
  const double mx = fmax(*B1, *B2);
  *B1 = mx;
  *B2 = mx;
}

__device__ __forceinline__ int sph_off(int l) {
  // sum_{k<l} (2k+1) ncart(k): 0, 1, 10, 40, 110
  const int off[5] = {0, 1, 10, 40, 110};
  return off[l];
}

// coefficient T(cart comp c, pure function m) of a shell
__device__ __forceinline__ double tcoef(const CartMap &M, int s, int c, int m) {
  if (!M.sh_pure[s]) return c == m ? 1.0 : 0.0;
  const int l = M.sh_l[s];
  return M.sph[sph_off(l) + m * ncart_d(l) + c];
}

__device__ __forceinline__ int nfunc(const CartMap &M, int s) {
  return M.sh_pure[s] ? 2 * M.sh_l[s] + 1 : ncart_d(M.sh_l[s]);
}

__global__ void to_cart_kernel(CartMap M, const double *X, int cplx, double *Pre, double *Pim,
                               int ps) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)M.nbc * M.nbc) return;
  const int i = (int)(t % M.nbc), j = (int)(t / M.nbc);
  const int si = M.cart_shell[i], ci = M.cart_comp[i];
  const int sj = M.cart_shell[j], cj = M.cart_comp[j];
  double re = 0.0, im = 0.0;
  for (int mu = 0; mu < nfunc(M, si); ++mu) {
    const double ti = tcoef(M, si, ci, mu);
    if (ti == 0.0) continue;
    for (int nu = 0; nu < nfunc(M, sj); ++nu) {
      const double tj = tcoef(M, sj, cj, nu);
      if (tj == 0.0) continue;
      const size_t idx = (M.sh_bf[si] - M.bf0 + mu) + (size_t)(M.sh_bf[sj] - M.bf0 + nu) * M.nb;
      if (cplx) {
        re += ti * tj * X[2 * idx];
        im += ti * tj * X[2 * idx + 1];
      } else {
        re += ti * tj * X[idx];
      }
    }
  }
  Pre[t * ps] = re;
  if (Pim) Pim[t * ps] = im;
}

__global__ void from_cart_kernel(CartMap M, const double *Are, const double *Aim, int ps,
                                 int her, int cplx, double *AX) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)M.nb * M.nb) return;
  const int mu = (int)(t % M.nb), nu = (int)(t / M.nb);
  const int sm = M.bf_shell[mu], cm = M.bf_comp[mu];
  const int sn = M.bf_shell[nu], cn = M.bf_comp[nu];
  const int ncm = ncart_d(M.sh_l[sm]), ncn = ncart_d(M.sh_l[sn]);
  double a_re = 0.0, a_im = 0.0, b_re = 0.0, b_im = 0.0;
  for (int i = 0; i < ncm; ++i) {
    const double ti = tcoef(M, sm, i, cm);
    if (ti == 0.0) continue;
    for (int j = 0; j < ncn; ++j) {
      const double tj = tcoef(M, sn, j, cn);
      if (tj == 0.0) continue;
      const size_t ij = (M.sh_bfc[sm] - M.bfc0 + i) + (size_t)(M.sh_bfc[sn] - M.bfc0 + j) * M.nbc;
      const size_t ji = (M.sh_bfc[sn] - M.bfc0 + j) + (size_t)(M.sh_bfc[sm] - M.bfc0 + i) * M.nbc;
      a_re += ti * tj * Are[ij * ps];
      b_re += ti * tj * Are[ji * ps];
      if (Aim) {
        a_im += ti * tj * Aim[ij * ps];
        b_im += ti * tj * Aim[ji * ps];
      }
    }
  }
  double re, im;
  if (her) {  // 1/2 (A + A^H), direct.hpp:2040-2049
    re = 0.5 * (a_re + b_re);
    im = 0.5 * (a_im - b_im);
  } else {    // 1/2 A, direct.hpp:2051-2062
    re = 0.5 * a_re;
    im = 0.5 * a_im;
  }
  if (cplx) {
    AX[2 * t] += re;
    AX[2 * t + 1] += im;
  } else {
    AX[t] += re;
  }
}

// max of each molecule's block of global block norms (one block per molecule)
__global__ void mol_max_kernel(const double *Bg, const long long *mol_boff, const int *mol_ns, double *out) {
  const int m = blockIdx.x;
  const long long n = (long long)mol_ns[m] * mol_ns[m];
  const double *B = Bg + mol_boff[m];
  double v = 0.0;
  for (long long i = threadIdx.x; i < n; i += blockDim.x) v = fmax(v, B[i]);
  __shared__ double red[256];
  red[threadIdx.x] = v;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[m] = red[0];
}

__global__ void gather_kernel(const long long *src, const long long *idx, int n, long long *dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[idx[i]];
}

__global__ void zero_kernel(double *p, size_t n) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = 0.0;
}

inline unsigned nblocks(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

bool team_shape(int la, int lb, int lc, int ld, int lds_max_bytes, int &ca, int &rpp) {
  const int nr = (la + lb + lc + ld) / 2 + 1;
  const int nh = (la + 1) * (lb + 1) * (lc + 1) * (ld + 1);
  auto nc = [](int l) { return (l + 1) * (l + 2) / 2; };
  const int nca = nc(la), nbcd = nc(lb) * nc(lc) * nc(ld);
  const int budget = lds_max_bytes / (int)sizeof(double);
  // the most first-shell components per pass (fewest passes over the
  // primitives), then the most roots per table pass that still fit
  for (ca = nca; ca >= 1; --ca) {
    if (team_lds_doubles(1, nh, ca, nbcd) > budget) continue;
    rpp = nr;
    while (rpp > 1 && team_lds_doubles(rpp, nh, ca, nbcd) > budget) --rpp;
    return true;
  }
  return false;
}

hipError_t launch_eri_team(const EriLaunch &a0, hipStream_t s) {
  if (a0.nq <= 0) return hipSuccess;
  if (a0.la > kMaxLDev || a0.lb > kMaxLDev || a0.lc > kMaxLDev || a0.ld > kMaxLDev) return hipErrorInvalidValue;
  EriLaunch a = a0;
  if (!team_shape(a.la, a.lb, a.lc, a.ld, 160 * 1024, a.team_ca, a.team_rpp)) return hipErrorInvalidValue;
  const int nh = (a.la + 1) * (a.lb + 1) * (a.lc + 1) * (a.ld + 1);
  auto nc = [](int l) { return (l + 1) * (l + 2) / 2; };
  const size_t lds = sizeof(double) * (size_t)team_lds_doubles(a.team_rpp, nh, a.team_ca, nc(a.lb) * nc(a.lc) * nc(a.ld));
  // once per process (contexts on several host threads launch concurrently)
  static std::once_flag attr_once;
  static hipError_t attr_err = hipSuccess;
  std::call_once(attr_once, [] {
    attr_err = hipFuncSetAttribute((const void *)eri_team_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   160 * 1024);
  });
  if (attr_err != hipSuccess) return attr_err;
  const long long grid = a.nq < 65536 ? a.nq : 65536;
  hipLaunchKernelGGL(eri_team_kernel, dim3((unsigned)grid), dim3(64), lds, s, a);
  return hipGetLastError();
}

hipError_t launch_tile_lists(const TileListLaunch *lists, const long long *start, int nlist, long long nsub,
                             hipStream_t s) {
  if (nsub <= 0 || nlist <= 0) return hipSuccess;
  if (nsub % kTileChunk) return hipErrorInvalidValue;  // lists are padded to whole chunks
  const long long nchunk = nsub / kTileChunk;
  const long long blocks = std::min<long long>(nchunk, 1ll << 20);
  hipLaunchKernelGGL(tile_lists_kernel, dim3((unsigned)blocks), dim3(64 * kTileListWaves), 0, s, lists, start,
                     nlist, nchunk);
  return hipGetLastError();
}

hipError_t launch_quartet_lists(const ListLaunch *lists, const long long *start, int nlist, long long nchunk,
                                hipStream_t s) {
  if (nlist <= 0 || nchunk <= 0) return hipSuccess;
  const long long blocks = nchunk < 65536 ? nchunk : 65536;
  hipLaunchKernelGGL(quartet_lists_kernel, dim3((unsigned)blocks), dim3(kListThreads), 0, s, lists, start, nlist,
                     nchunk);
  return hipGetLastError();
}

hipError_t launch_pair_norms(const PairScr *pairs, int npair, const double *Bg, const double *Bm,
                             const long long *mol_boff, const int *mol_ns, int nmat, double *pn,
                             hipStream_t s) {
  if (npair <= 0) return hipSuccess;
  hipLaunchKernelGGL(pair_norm_kernel, dim3(nblocks(npair, 256)), dim3(256), 0, s, pairs, npair, Bg, Bm,
                     mol_boff, mol_ns, nmat, pn);
  return hipGetLastError();
}

hipError_t launch_block_norms(const double *X, int nmat, int is_complex, const int *her, int nb,
                              int ns, const int *sh_bf, int bf0, const int *sh_nf, double *Bm, double *Bg,
                              hipStream_t s) {
  const long long n1 = (long long)nmat * ns * ns;
  hipLaunchKernelGGL(block_norm_kernel, dim3(nblocks(n1, 256)), dim3(256), 0, s, X, nmat,
                     is_complex, nb, ns, sh_bf, bf0, sh_nf, Bm);
  hipLaunchKernelGGL(block_norm_symm_kernel, dim3(nblocks(n1, 256)), dim3(256), 0, s, nmat, ns,
                     her, Bm);
  hipLaunchKernelGGL(block_norm_finish_kernel, dim3(nblocks((long long)ns * ns, 256)), dim3(256),
                     0, s, nmat, ns, Bm, Bg);
  return hipGetLastError();
}

hipError_t launch_to_cart(const CartMap &m, const double *X, int is_complex, double *pre,
                          double *pim, int ps, hipStream_t s) {
  hipLaunchKernelGGL(to_cart_kernel, dim3(nblocks((long long)m.nbc * m.nbc, 256)), dim3(256), 0,
                     s, m, X, is_complex, pre, pim, ps);
  return hipGetLastError();
}

hipError_t launch_from_cart(const CartMap &m, const double *are, const double *aim, int ps, int her,
                            int is_complex, double *AX, hipStream_t s) {
  hipLaunchKernelGGL(from_cart_kernel, dim3(nblocks((long long)m.nb * m.nb, 256)), dim3(256), 0,
                     s, m, are, aim, ps, her, is_complex, AX);
  return hipGetLastError();
}

hipError_t launch_mol_max(const double *Bg, const long long *mol_boff, const int *mol_ns, int nmol,
                          double *out, hipStream_t s) {
  if (nmol <= 0) return hipSuccess;
  hipLaunchKernelGGL(mol_max_kernel, dim3(nmol), dim3(256), 0, s, Bg, mol_boff, mol_ns, out);
  return hipGetLastError();
}

hipError_t launch_gather(const long long *src, const long long *idx, int n, long long *dst, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(gather_kernel, dim3(nblocks(n, 256)), dim3(256), 0, s, src, idx, n, dst);
  return hipGetLastError();
}

hipError_t launch_zero(double *p, size_t n, hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(zero_kernel, dim3(nblocks((long long)n, 256)), dim3(256), 0, s, p, n);
  return hipGetLastError();
}

}  // namespace cqx
