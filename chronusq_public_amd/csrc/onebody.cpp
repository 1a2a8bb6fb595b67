// One-electron integrals over the engine's shells: overlap S, kinetic
// energy T and the attraction V of point nuclei -- the three matrices that
// ChronusQ's OnePInts::computeAOInts builds with libint2 for a
// non-relativistic reference (src/particleintegrals/aointegrals/
// aointegrals_onee_drivers.cxx:762-784; point nuclei unless finiteWidthNuc).
//
// Method (McMurchie-Davidson): the product of two Cartesian Gaussians is
// expanded in Hermite Gaussians with coefficients E^{ij}_t from the
// two-term recursions; overlap and kinetic energy follow from E^{ij}_0,
// the Coulomb integral of a Hermite Gaussian from the Hermite integrals
// R_{tuv} over the Boys function.  Host code, threads over shell pairs.
#include "onebody.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <stdexcept>
#include <thread>

namespace cqx {

namespace {

constexpr int kMaxLL = 2 * kMaxL;  // highest la + lb
constexpr int kTaylor = 7;         // Taylor terms of the tabulated Boys function
constexpr double kBoysStep = 0.05;
constexpr double kBoysTmax = 36.0;  // above: F_0 = sqrt(pi / T) / 2 to double precision

// F_m(T) on the grid T_i = i h, m = 0 .. kMaxLL + kTaylor, by the series
// F_m(T) = exp(-T) sum_k (2T)^k / ((2m+1)(2m+3)...(2m+2k+1)).
struct BoysGrid {
  static constexpr int kM = kMaxLL + kTaylor + 1;
  int n = 0;
  std::vector<double> v;  // v[i * kM + m]
  BoysGrid() {
    n = (int)(kBoysTmax / kBoysStep) + 2;
    v.resize((size_t)n * kM);
    for (int i = 0; i < n; ++i) {
      const double T = i * kBoysStep, e = std::exp(-T);
      for (int m = 0; m < kM; ++m) {
        double term = 1.0 / (2 * m + 1), sum = term;
        for (int k = 1; k < 2000; ++k) {
          term *= 2.0 * T / (2 * m + 2 * k + 1);
          sum += term;
          if (term < 1e-18 * sum) break;
        }
        v[(size_t)i * kM + m] = e * sum;
      }
    }
  }
};

const BoysGrid &boys_grid() {
  static const BoysGrid g;
  return g;
}

}  // namespace

void boys_host(int mmax, double T, double *F) {
  if (mmax < 0 || mmax > kMaxLL) throw std::runtime_error("boys_host: order out of range");
  if (T < kBoysTmax) {
    // Taylor expansion about the nearest grid point: dF_m/dT = -F_{m+1}
    const BoysGrid &g = boys_grid();
    const int i = (int)std::lround(T / kBoysStep);
    const double d = i * kBoysStep - T;
    const double *row = g.v.data() + (size_t)i * BoysGrid::kM;
    for (int m = 0; m <= mmax; ++m) {
      double s = 0.0, c = 1.0;
      for (int k = 0; k < kTaylor; ++k) {
        s += row[m + k] * c;
        c *= d / (k + 1);
      }
      F[m] = s;
    }
  } else {
    const double e = std::exp(-T);
    F[0] = 0.5 * std::sqrt(M_PI / T);
    for (int m = 0; m < mmax; ++m) F[m + 1] = ((2 * m + 1) * F[m] - e) / (2.0 * T);
  }
}

namespace {

// Hermite coefficients of one direction: E[(i * (lbx + 1) + j) * ntd + t],
// i <= la, j <= lbx, t <= i + j, without the Gaussian-product exponential.
struct Herm {
  int la, lbx, ntd;
  std::vector<double> E;
  Herm(int la_, int lbx_) : la(la_), lbx(lbx_), ntd(la_ + lbx_ + 1), E((size_t)(la_ + 1) * (lbx_ + 1) * (la_ + lbx_ + 1)) {}
  double &at(int i, int j, int t) { return E[((size_t)i * (lbx + 1) + j) * ntd + t]; }
  double get(int i, int j, int t) const {
    if (t < 0 || t > i + j) return 0.0;
    return E[((size_t)i * (lbx + 1) + j) * ntd + t];
  }
  void build(double p, double PA, double PB) {
    std::fill(E.begin(), E.end(), 0.0);
    const double h = 0.5 / p;
    at(0, 0, 0) = 1.0;
    for (int i = 0; i < la; ++i)
      for (int t = 0; t <= i + 1; ++t)
        at(i + 1, 0, t) = h * get(i, 0, t - 1) + PA * get(i, 0, t) + (t + 1) * get(i, 0, t + 1);
    for (int i = 0; i <= la; ++i)
      for (int j = 0; j < lbx; ++j)
        for (int t = 0; t <= i + j + 1; ++t)
          at(i, j + 1, t) = h * get(i, j, t - 1) + PB * get(i, j, t) + (t + 1) * get(i, j, t + 1);
  }
};

// Hermite Coulomb integrals R^0_{tuv}, t + u + v <= L, of exponent p about
// PC; R is indexed [(t * (L+1) + u) * (L+1) + v].
void hermite_R(int L, double p, const double PC[3], std::vector<double> &R, std::vector<double> &W) {
  const int D = L + 1;
  const size_t S = (size_t)D * D * D;
  W.assign(S * D, 0.0);  // W[n * S + idx]
  double F[kMaxLL + 1];
  const double T = p * (PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2]);
  boys_host(L, T, F);
  double f = 1.0;
  for (int n = 0; n <= L; ++n) {
    W[n * S] = f * F[n];
    f *= -2.0 * p;
  }
  auto idx = [D](int t, int u, int v) { return ((size_t)t * D + u) * D + v; };
  for (int n = L - 1; n >= 0; --n) {
    const double *Rn1 = W.data() + (size_t)(n + 1) * S;
    double *Rn = W.data() + (size_t)n * S;
    for (int t = 0; t <= L - n; ++t)
      for (int u = 0; u + t <= L - n; ++u)
        for (int v = 0; v + u + t <= L - n; ++v) {
          if (t + u + v == 0) continue;
          double r;
          if (t > 0)
            r = (t > 1 ? (t - 1) * Rn1[idx(t - 2, u, v)] : 0.0) + PC[0] * Rn1[idx(t - 1, u, v)];
          else if (u > 0)
            r = (u > 1 ? (u - 1) * Rn1[idx(t, u - 2, v)] : 0.0) + PC[1] * Rn1[idx(t, u - 1, v)];
          else
            r = (v > 1 ? (v - 1) * Rn1[idx(t, u, v - 2)] : 0.0) + PC[2] * Rn1[idx(t, u, v - 1)];
          Rn[idx(t, u, v)] = r;
        }
  }
  R.assign(W.begin(), W.begin() + S);
}

void cart_exps(int l, std::vector<int> &e) {
  e.clear();
  for (int lx = l; lx >= 0; --lx)
    for (int ly = l - lx; ly >= 0; --ly) {
      e.push_back(lx);
      e.push_back(ly);
      e.push_back(l - lx - ly);
    }
}

// Cartesian block (row-major nca x ncb) of one operator for shells A, B.
void cart_block(const Shell &A, const Shell &B, int kind, const std::vector<double> &Zc,
                const std::vector<double> &Cxyz, std::vector<double> &out) {
  const int la = A.l, lb = B.l, nca = ncart(la), ncb = ncart(lb);
  out.assign((size_t)nca * ncb, 0.0);
  std::vector<int> ea, eb;
  cart_exps(la, ea);
  cart_exps(lb, eb);
  const int lbx = kind == kKinetic ? lb + 2 : lb;
  Herm H[3] = {Herm(la, lbx), Herm(la, lbx), Herm(la, lbx)};
  double AB2 = 0.0;
  for (int d = 0; d < 3; ++d) AB2 += (A.origin[d] - B.origin[d]) * (A.origin[d] - B.origin[d]);
  std::vector<double> R, W;
  const int L = la + lb, D = L + 1;
  for (size_t pa = 0; pa < A.alpha.size(); ++pa)
    for (size_t pb = 0; pb < B.alpha.size(); ++pb) {
      const double a = A.alpha[pa], b = B.alpha[pb], p = a + b;
      const double K = A.coef[pa] * B.coef[pb] * std::exp(-a * b / p * AB2);
      if (std::fabs(K) < 1e-24) continue;
      double P[3];
      for (int d = 0; d < 3; ++d) {
        P[d] = (a * A.origin[d] + b * B.origin[d]) / p;
        H[d].build(p, P[d] - A.origin[d], P[d] - B.origin[d]);
      }
      if (kind == kOverlap || kind == kKinetic) {
        const double pre = K * std::pow(M_PI / p, 1.5);
        for (int x = 0; x < nca; ++x)
          for (int y = 0; y < ncb; ++y) {
            const int *ia = &ea[3 * x], *jb = &eb[3 * y];
            double S1[3], T1[3];
            for (int d = 0; d < 3; ++d) {
              const int i = ia[d], j = jb[d];
              S1[d] = H[d].get(i, j, 0);
              if (kind == kKinetic) {
                // -1/2 d^2/dx^2 on the ket: j(j-1) x^{j-2} - 2b(2j+1) x^j + 4b^2 x^{j+2}
                const double lo = j >= 2 ? j * (j - 1) * H[d].get(i, j - 2, 0) : 0.0;
                T1[d] = -0.5 * (lo - 2.0 * b * (2 * j + 1) * S1[d] + 4.0 * b * b * H[d].get(i, j + 2, 0));
              }
            }
            double v = kind == kOverlap
                           ? S1[0] * S1[1] * S1[2]
                           : T1[0] * S1[1] * S1[2] + S1[0] * T1[1] * S1[2] + S1[0] * S1[1] * T1[2];
            out[(size_t)x * ncb + y] += pre * v;
          }
      } else {
        const double pre = K * 2.0 * M_PI / p;
        for (size_t c = 0; c < Zc.size(); ++c) {
          const double PC[3] = {P[0] - Cxyz[3 * c], P[1] - Cxyz[3 * c + 1], P[2] - Cxyz[3 * c + 2]};
          hermite_R(L, p, PC, R, W);
          const double zc = -Zc[c] * pre;
          for (int x = 0; x < nca; ++x)
            for (int y = 0; y < ncb; ++y) {
              const int *ia = &ea[3 * x], *jb = &eb[3 * y];
              double s = 0.0;
              for (int t = 0; t <= ia[0] + jb[0]; ++t) {
                const double et = H[0].get(ia[0], jb[0], t);
                if (et == 0.0) continue;
                for (int u = 0; u <= ia[1] + jb[1]; ++u) {
                  const double eu = et * H[1].get(ia[1], jb[1], u);
                  if (eu == 0.0) continue;
                  for (int v = 0; v <= ia[2] + jb[2]; ++v)
                    s += eu * H[2].get(ia[2], jb[2], v) * R[((size_t)t * D + u) * D + v];
                }
              }
              out[(size_t)x * ncb + y] += zc * s;
            }
        }
      }
    }
}

}  // namespace

std::vector<double> one_body_matrix(const Basis &B, int kind, const std::vector<double> &Zc,
                                    const std::vector<double> &Cxyz, int nthreads) {
  if (kind < kOverlap || kind > kPotential) throw std::runtime_error("one_body_matrix: unknown kind");
  if (Cxyz.size() != 3 * Zc.size()) throw std::runtime_error("one_body_matrix: charge arrays disagree");
  const int ns = (int)B.shells.size(), NB = B.nbasis;
  std::vector<std::vector<double>> sph;
  for (int l = 0; l <= kMaxL; ++l) sph.push_back(solid_harmonics(l));
  std::vector<double> M((size_t)NB * NB, 0.0);
  const long long npair = (long long)ns * (ns + 1) / 2;
  std::atomic<long long> next{0};
  auto work = [&]() {
    std::vector<double> cart;
    for (;;) {
      const long long k = next.fetch_add(1);
      if (k >= npair) break;
      // k -> (a >= b)
      int a = (int)((std::sqrt(8.0 * (double)k + 1.0) - 1.0) / 2.0);
      while ((long long)a * (a + 1) / 2 > k) --a;
      while ((long long)(a + 1) * (a + 2) / 2 <= k) ++a;
      const int b = (int)(k - (long long)a * (a + 1) / 2);
      const Shell &A = B.shells[a], &Bs = B.shells[b];
      cart_block(A, Bs, kind, Zc, Cxyz, cart);
      const int nca = ncart(A.l), ncb = ncart(Bs.l), na = A.nfunc(), nb = Bs.nfunc();
      for (int i = 0; i < na; ++i)
        for (int j = 0; j < nb; ++j) {
          double s = 0.0;
          for (int x = 0; x < nca; ++x) {
            const double ta = A.pure ? sph[A.l][(size_t)i * nca + x] : (x == i ? 1.0 : 0.0);
            if (ta == 0.0) continue;
            for (int y = 0; y < ncb; ++y) {
              const double tb = Bs.pure ? sph[Bs.l][(size_t)j * ncb + y] : (y == j ? 1.0 : 0.0);
              if (tb != 0.0) s += ta * tb * cart[(size_t)x * ncb + y];
            }
          }
          const int r = B.bf[a] + i, c = B.bf[b] + j;
          M[r + (size_t)c * NB] = s;
          M[c + (size_t)r * NB] = s;
        }
    }
  };
  if (nthreads <= 0) nthreads = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<std::thread> pool;
  for (int t = 1; t < nthreads; ++t) pool.emplace_back(work);
  work();
  for (auto &t : pool) t.join();
  return M;
}

}  // namespace cqx
