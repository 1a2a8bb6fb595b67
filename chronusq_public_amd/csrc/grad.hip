// Nuclear-gradient J/K contraction: the derivative integrals of every
// shell quartet digested into one J or K matrix per nuclear coordinate,
// DirectGradContraction::directScaffoldGrad (direct.hpp:2198-2580).
//
// One 64-lane wave per shell quartet, any class through (gg|gg).  Per primitive
// quartet the Rys 2D integrals are built with one more unit of angular
// momentum on centres A, B and C; the derivative of a Cartesian component
// along direction x is 2 alpha I(a + 1_x) - a_x I(a - 1_x) on the x factor,
// and the centre-D derivative follows from translational invariance,
// d/dD = -(d/dA + d/dB + d/dC).  The twelve derivative quartets, scaled by
// the 8-fold degeneracy, are digested with the Hermitian J / K updates of
// the reference (direct.hpp:2450-2497) into the Cartesian accumulator of
// the coordinate (atom of the centre, direction).  The derivative quartets
// are built in LDS for a chunk of the Cartesian components at a time (the
// whole quartet when it fits beside the 2D tables; the primitive loop runs
// once per chunk), with the 2D tables of rpp roots at a time (all of them
// unless the tables of a g class would not fit).
#include <hip/hip_runtime.h>

#include <mutex>

#include "kernels.hpp"
#include "rys.hpp"

namespace cqx {

namespace {

constexpr double kTwoPi52g = 34.98683665524972;  // 2 pi^(5/2)

__device__ __forceinline__ int ncart_g(int l) { return (l + 1) * (l + 2) / 2; }

__device__ __forceinline__ void cart_exp_g(int l, int c, int e[3]) {
  int k = c, lx = l;
  while (k >= l - lx + 1) {
    k -= l - lx + 1;
    --lx;
  }
  e[0] = lx;
  e[2] = k;
  e[1] = l - lx - k;
}

struct GDims {
  int la, lb, lc, ld;      // the quartet's class
  int ea, eb, ec;          // table extents on A, B, C: l + 2 entries (0..l+1)
  int ni, nk, nr, nh, nc;  // VRR extents, Rys roots, 2D table size, Cartesian quartet size
};

__device__ GDims gdims(int la, int lb, int lc, int ld) {
  GDims d;
  d.la = la; d.lb = lb; d.lc = lc; d.ld = ld;
  d.ea = la + 2; d.eb = lb + 2; d.ec = lc + 2;
  d.ni = la + lb + 2;
  d.nk = lc + ld + 2;
  d.nr = (la + lb + lc + ld + 1) / 2 + 1;
  d.nh = d.ea * d.eb * d.ec * (ld + 1);
  d.nc = ncart_g(la) * ncart_g(lb) * ncart_g(lc) * ncart_g(ld);
  return d;
}

__device__ __forceinline__ int hix(const GDims &D, int ia, int ib, int ic, int id) {
  return ((ia * D.eb + ib) * D.ec + ic) * (D.ld + 1) + id;
}

// 2D integrals of one root along one direction with A, B, C raised by one
// (entries with both ia = la + 1 and ib = lb + 1 are not needed and not set)
__device__ void rys_2d_grad(const GDims &D, double t2, double p, double q, double PAd, double QCd, double PQd,
                            double ABd, double CDd, double scale, double *__restrict__ H) {
  constexpr int NI = 2 * kMaxLDev + 2;
  double G[NI][NI];                 // [i <= la + lb + 1][k <= lc + ld + 1]
  double E[NI][kMaxLDev + 2][NI];   // bra transfer: [i][j <= lb + 1][k]
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
  for (int i = 0; i < D.ni; ++i)
    for (int k = 0; k < D.nk; ++k) E[i][0][k] = G[i][k];
  for (int j = 1; j <= D.lb + 1; ++j)
    for (int i = 0; i + j < D.ni; ++i)
      for (int k = 0; k < D.nk; ++k) E[i][j][k] = E[i + 1][j - 1][k] + ABd * E[i][j - 1][k];
  for (int ia = 0; ia <= D.la + 1; ++ia)
    for (int ib = 0; ib <= D.lb + 1; ++ib) {
      if (ia + ib >= D.ni) continue;
      double F[NI][kMaxLDev + 1];  // ket transfer: [k][l <= ld]
      for (int k = 0; k < D.nk; ++k) F[k][0] = E[ia][ib][k];
      for (int l = 1; l <= D.ld; ++l)
        for (int k = 0; k + l < D.nk; ++k) F[k][l] = F[k + 1][l - 1] + CDd * F[k][l - 1];
      for (int ic = 0; ic <= D.lc + 1; ++ic)
        for (int id = 0; id <= D.ld; ++id) H[hix(D, ia, ib, ic, id)] = F[ic][id];
    }
}

__device__ __forceinline__ void add(double *p, double v) { unsafeAtomicAdd(p, v); }

__global__ __launch_bounds__(64) void eri_grad_kernel(GradLaunch a) {
  extern __shared__ double lds[];
  const int lane = threadIdx.x;
  const GDims D = gdims(a.la, a.lb, a.lc, a.ld);
  double *s_x = lds;
  double *s_w = lds + 16;
  double *s_H = lds + 32;
  const int chunk = a.chunk;                      // Cartesian components per pass
  const int RPP = a.rpp;                          // roots per 2D-table pass
  double *s_dI = s_H + (size_t)RPP * 3 * D.nh;    // [9][chunk]: centre A, B, C x direction
  const int na = ncart_g(D.la), nbb = ncart_g(D.lb), ncc = ncart_g(D.lc), ndd = ncart_g(D.ld);
  const size_t nbc2 = (size_t)a.nbc * a.nbc;
  const long long npass = (D.nc + chunk - 1) / chunk;
  for (long long qp = blockIdx.x; qp < a.nq * npass; qp += gridDim.x) {
    const long long q = qp / npass;
    const int o0 = (int)(qp - q * npass) * chunk, o1 = min(D.nc, o0 + chunk);
    int i, j;
    if (a.same_class) {  // j <= i, row-major lower triangle
      long long r = (long long)((sqrt(8.0 * (double)q + 1.0) - 1.0) * 0.5);
      while (r * (r + 1) / 2 > q) --r;
      while ((r + 1) * (r + 2) / 2 <= q) ++r;
      i = (int)r;
      j = (int)(q - r * (r + 1) / 2);
    } else {
      i = (int)(q / a.nket);
      j = (int)(q - (long long)i * a.nket);
    }
    const PairRec bp = a.bra[i];
    const PairRec kp = a.ket[j];
    for (int o = lane; o < 9 * chunk; o += 64) s_dI[o] = 0.0;
    for (int pi = 0; pi < bp.npp; ++pi) {
      const PrimPair A = a.prims[bp.pp_off + pi];
      const double al_a = a.prim_alpha[bp.pp_off + pi], al_b = A.p - al_a;
      for (int pj = 0; pj < kp.npp; ++pj) {
        const PrimPair B = a.prims[kp.pp_off + pj];
        const double al_c = a.prim_alpha[kp.pp_off + pj];
        const double pq = A.p + B.p, rho = A.p * B.p / pq;
        const double PQ[3] = {A.P[0] - B.P[0], A.P[1] - B.P[1], A.P[2] - B.P[2]};
        const double T = rho * (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]);
        const double pref = A.K * B.K * kTwoPi52g / (A.p * B.p * sqrt(pq));
        __syncthreads();
        if (lane < D.nr) {
          double x, w;
          rys_root_rt(D.nr, lane, T, a.rys, x, w);
          s_x[lane] = x;
          s_w[lane] = w * pref;
        }
        const double twoal[3] = {2.0 * al_a, 2.0 * al_b, 2.0 * al_c};
        for (int r0 = 0; r0 < D.nr; r0 += RPP) {
        const int nrp = D.nr - r0 < RPP ? D.nr - r0 : RPP;
        __syncthreads();
        if (lane < 3 * nrp) {
          const int lr = lane / 3, dir = lane % 3, r = r0 + lr;
          rys_2d_grad(D, s_x[r], A.p, B.p, A.PA[dir], B.PA[dir], PQ[dir], bp.AB[dir], kp.AB[dir],
                      dir == 2 ? s_w[r] : 1.0, s_H + (size_t)(lr * 3 + dir) * D.nh);
        }
        __syncthreads();
        for (int o = o0 + lane; o < o1; o += 64) {
          int rem = o;
          const int id = rem % ndd; rem /= ndd;
          const int ic = rem % ncc; rem /= ncc;
          const int ib = rem % nbb; rem /= nbb;
          const int ia = rem;
          int e[4][3];
          cart_exp_g(D.la, ia, e[0]);
          cart_exp_g(D.lb, ib, e[1]);
          cart_exp_g(D.lc, ic, e[2]);
          cart_exp_g(D.ld, id, e[3]);
          double acc[9];
          for (int t = 0; t < 9; ++t) acc[t] = 0.0;
          for (int lr = 0; lr < nrp; ++lr) {
            const double *Hr = s_H + (size_t)lr * 3 * D.nh;
            double f[3], df[3][3];  // plain factor per direction; derivative per centre and direction
            for (int dir = 0; dir < 3; ++dir) {
              const double *Hd = Hr + (size_t)dir * D.nh;
              const int x0 = e[0][dir], x1 = e[1][dir], x2 = e[2][dir], x3 = e[3][dir];
              f[dir] = Hd[hix(D, x0, x1, x2, x3)];
              df[0][dir] = twoal[0] * Hd[hix(D, x0 + 1, x1, x2, x3)] - (x0 ? x0 * Hd[hix(D, x0 - 1, x1, x2, x3)] : 0.0);
              df[1][dir] = twoal[1] * Hd[hix(D, x0, x1 + 1, x2, x3)] - (x1 ? x1 * Hd[hix(D, x0, x1 - 1, x2, x3)] : 0.0);
              df[2][dir] = twoal[2] * Hd[hix(D, x0, x1, x2 + 1, x3)] - (x2 ? x2 * Hd[hix(D, x0, x1, x2 - 1, x3)] : 0.0);
            }
            for (int k = 0; k < 3; ++k) {
              acc[k * 3 + 0] += df[k][0] * f[1] * f[2];
              acc[k * 3 + 1] += f[0] * df[k][1] * f[2];
              acc[k * 3 + 2] += f[0] * f[1] * df[k][2];
            }
          }
          for (int t = 0; t < 9; ++t) s_dI[(size_t)t * chunk + (o - o0)] += acc[t];
        }
        }
      }
    }
    __syncthreads();
    // degeneracy of the quartet in the 8-fold loop and the 1/2 (:2382-2388)
    const bool same_pair = (bp.a == kp.a && bp.b == kp.b);
    const double deg = 0.5 * (bp.a == bp.b ? 1.0 : 2.0) * (kp.a == kp.b ? 1.0 : 2.0) * (same_pair ? 1.0 : 2.0);
    const int sh[4] = {bp.a, bp.b, kp.a, kp.b};
    const int bf[4] = {bp.bfa, bp.bfb, kp.bfa, kp.bfb};
    int gc[4];
    for (int k = 0; k < 4; ++k) gc[k] = a.shell_atom[sh[k]] * 3;
    const int N = a.nbc;
    for (int o = o0 + lane; o < o1; o += 64) {
      int rem = o;
      const int id = rem % ndd; rem /= ndd;
      const int ic = rem % ncc; rem /= ncc;
      const int ib = rem % nbb; rem /= nbb;
      const int ia = rem;
      const int f1 = bf[0] + ia, f2 = bf[1] + ib, f3 = bf[2] + ic, f4 = bf[3] + id;
      for (int dir = 0; dir < 3; ++dir) {
        double v[4];
        v[0] = deg * s_dI[(size_t)(0 + dir) * chunk + (o - o0)];
        v[1] = deg * s_dI[(size_t)(3 + dir) * chunk + (o - o0)];
        v[2] = deg * s_dI[(size_t)(6 + dir) * chunk + (o - o0)];
        v[3] = -(v[0] + v[1] + v[2]);  // translational invariance
        for (int k = 0; k < 4; ++k) {
          const double I = v[k];
          if (I == 0.0) continue;
          const int g = gc[k] + dir;
          for (int m = 0; m < a.nmat; ++m) {
            const double *X = a.X + (size_t)m * nbc2;
            double *A_ = a.acc + ((size_t)g * a.nmat + m) * nbc2;
            if (!((a.kmask >> m) & 1u)) {
              // J(1,2) += I X(4,3); J(4,3) += I X(1,2)
              add(A_ + f1 + (size_t)f2 * N, I * X[f4 + (size_t)f3 * N]);
              add(A_ + f4 + (size_t)f3 * N, I * X[f1 + (size_t)f2 * N]);
            } else {
              // K(1,3) += I X(4,2)/2; K(4,2) += I X(1,3)/2; K(4,1) += I X(2,3)/2; K(2,3) += I X(4,1)/2
              const double h = 0.5 * I;
              add(A_ + f1 + (size_t)f3 * N, h * X[f4 + (size_t)f2 * N]);
              add(A_ + f4 + (size_t)f2 * N, h * X[f1 + (size_t)f3 * N]);
              add(A_ + f4 + (size_t)f1 * N, h * X[f2 + (size_t)f3 * N]);
              add(A_ + f2 + (size_t)f3 * N, h * X[f4 + (size_t)f1 * N]);
            }
          }
        }
      }
    }
    __syncthreads();
  }
}

}  // namespace

size_t grad_lds_bytes(int la, int lb, int lc, int ld, int chunk, int rpp) {
  const size_t nh = (size_t)(la + 2) * (lb + 2) * (lc + 2) * (ld + 1);
  return sizeof(double) * (32 + (size_t)rpp * 3 * nh + 9 * (size_t)chunk);
}

int grad_chunk(int la, int lb, int lc, int ld, size_t lds_max, int *rpp_out) {
  auto nca = [](int l) { return (l + 1) * (l + 2) / 2; };
  const int nc = nca(la) * nca(lb) * nca(lc) * nca(ld);
  const int nr = (la + lb + lc + ld + 1) / 2 + 1;
  // the fewest passes over the components, then the most roots per table pass
  int best = 0, best_rpp = 0;
  long long best_passes = 0;
  for (int rpp = nr; rpp >= 1; --rpp) {
    const size_t fixed = grad_lds_bytes(la, lb, lc, ld, 0, rpp);
    if (fixed >= lds_max) continue;
    const size_t fit = (lds_max - fixed) / (9 * sizeof(double));
    const int chunk = fit >= (size_t)nc ? nc : (fit >= 64 ? (int)(fit / 64 * 64) : 0);
    if (chunk <= 0) continue;
    const long long passes = (nc + chunk - 1) / chunk;
    if (best == 0 || passes < best_passes) {
      best = chunk;
      best_rpp = rpp;
      best_passes = passes;
    }
  }
  if (rpp_out) *rpp_out = best_rpp;
  return best;
}

hipError_t launch_eri_grad(const GradLaunch &a, hipStream_t s) {
  if (a.nq <= 0) return hipSuccess;
  if (a.la > kMaxLDev || a.lb > kMaxLDev || a.lc > kMaxLDev || a.ld > kMaxLDev || a.chunk <= 0 || a.rpp <= 0)
    return hipErrorNotSupported;
  const int nc = ((a.la + 1) * (a.la + 2) / 2) * ((a.lb + 1) * (a.lb + 2) / 2) * ((a.lc + 1) * (a.lc + 2) / 2) *
                 ((a.ld + 1) * (a.ld + 2) / 2);
  const long long work = a.nq * ((nc + a.chunk - 1) / a.chunk);
  const long long blocks = work < 262144 ? work : 262144;
  const size_t lds = grad_lds_bytes(a.la, a.lb, a.lc, a.ld, a.chunk, a.rpp);
  static std::once_flag once;  // dynamic LDS beyond the default limit for the f classes
  static hipError_t attr = hipSuccess;
  std::call_once(once, [&] {
    attr = hipFuncSetAttribute(reinterpret_cast<const void *>(eri_grad_kernel),
                               hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
  });
  if (attr != hipSuccess) return attr;
  hipLaunchKernelGGL(eri_grad_kernel, dim3((unsigned)blocks), dim3(64), lds, s, a);
  return hipGetLastError();
}

}  // namespace cqx
