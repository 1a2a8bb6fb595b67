// Shell sets for the direct J/K engine: G94 parsing, libint2-convention
// normalisation, real-solid-harmonic transforms and the significant
// shell-pair list.
//
// Reference behaviour restated here:
//   ReferenceBasisSet::parseBasisFile   src/basisset/reference.cxx:127-289
//   ReferenceBasisSet::generateShellSet src/basisset/reference.cxx:299-357
//   BasisSet::update (bf maps)          src/basisset/basisset.cxx:98-196
//   ShellPairData::computeShellPairs    src/basisset/basisset.cxx:652-705
#include "basis.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <sstream>
#include <stdexcept>

namespace cqx {

namespace {

const char *kElements[] = {
    "X",  "H",  "He", "Li", "Be", "B",  "C",  "N",  "O",  "F",  "Ne", "Na",
    "Mg", "Al", "Si", "P",  "S",  "Cl", "Ar", "K",  "Ca", "Sc", "Ti", "V",
    "Cr", "Mn", "Fe", "Co", "Ni", "Cu", "Zn", "Ga", "Ge", "As", "Se", "Br",
    "Kr", "Rb", "Sr", "Y",  "Zr", "Nb", "Mo", "Tc", "Ru", "Rh", "Pd", "Ag",
    "Cd", "In", "Sn", "Sb", "Te", "I",  "Xe"};

int element_number(std::string s) {
  for (auto &c : s) c = (char)std::tolower((unsigned char)c);
  for (int z = 1; z < (int)(sizeof kElements / sizeof kElements[0]); ++z) {
    std::string e = kElements[z];
    for (auto &c : e) c = (char)std::tolower((unsigned char)c);
    if (e == s) return z;
  }
  return -1;
}

// ReferenceBasisSet::str2doub: Fortran 'D' exponents
double to_double(std::string s) {
  for (auto &c : s)
    if (c == 'D' || c == 'd') c = 'E';
  return std::stod(s);
}

int l_of_symbol(const std::string &s) {
  static const std::string sym = "SPDFGHIJKLMNOQRTUVWXYZ";
  if (s.size() != 1) return -1;
  auto p = sym.find((char)std::toupper((unsigned char)s[0]));
  return p == std::string::npos ? -1 : (int)p;
}

std::string rtrim(std::string s) {
  while (!s.empty() && std::isspace((unsigned char)s.back())) s.pop_back();
  return s;
}

std::vector<std::string> tokens(const std::string &s) {
  std::istringstream is(s);
  std::vector<std::string> t;
  std::string w;
  while (is >> w) t.push_back(w);
  return t;
}

double double_factorial(int n) {
  double r = 1.0;
  for (int k = n; k > 1; k -= 2) r *= k;
  return r;
}

double choose(int n, int k) {
  if (k < 0 || k > n) return 0.0;
  double r = 1.0;
  for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return r;
}

}  // namespace

int cart_index(int lx, int ly, int lz) {
  int l = lx + ly + lz;
  return (l - lx) * (l - lx + 1) / 2 + lz;
}

// Coefficients of the (2l+1) real solid harmonics in the x^l-normalised
// Cartesian monomials of degree l.  Built from the associated Legendre form
//   r^l P_l^|m|(z/r) {cos m phi, sin |m| phi}
//     = (d^|m|/dt^|m| P_l)(z, r^2 homogenised) * {Re, Im} (x + i y)^|m|
// (no Condon-Shortley phase), expanded into monomials and scaled to unit
// norm.  Rows m = -l..l, columns in cart_index order.
std::vector<double> solid_harmonics(int l) {
  int nc = ncart(l), ns = 2 * l + 1;
  std::vector<double> T((size_t)ns * nc, 0.0);
  // Legendre P_l(t) = 2^-l sum_k (-1)^k C(l,k) C(2l-2k,l) t^(l-2k)
  for (int m = -l; m <= l; ++m) {
    int am = std::abs(m);
    // polynomial in (x, y, z): map (px,py,pz) -> coefficient
    std::vector<double> poly((size_t)(l + 1) * (l + 1) * (l + 1), 0.0);
    auto at = [&](int px, int py, int pz) -> double & {
      return poly[((size_t)px * (l + 1) + py) * (l + 1) + pz];
    };
    // (x + i y)^am -> real / imaginary parts: sum_j C(am,j) x^(am-j) (iy)^j
    std::vector<std::pair<double, int>> xy;  // (coef, power of y), x power = am - j
    for (int j = 0; j <= am; ++j) {
      // i^j: real for even j (sign (-1)^(j/2)), imaginary for odd j
      bool real_part = (j % 2 == 0);
      if ((m >= 0) != real_part) continue;
      double sgn = ((j / 2) % 2) ? -1.0 : 1.0;
      xy.push_back({sgn * choose(am, j), j});
    }
    for (int k = 0; 2 * k <= l; ++k) {
      int deg = l - 2 * k;  // power of t in P_l
      if (deg < am) continue;
      double c = std::pow(0.5, l) * ((k % 2) ? -1.0 : 1.0) * choose(l, k) *
                 choose(2 * l - 2 * k, l);
      for (int f = 0; f < am; ++f) c *= (deg - f);  // d^am/dt^am
      int zp = deg - am;                             // z power
      int rk = k;                                    // (r^2)^k to homogenise
      // expand (x^2 + y^2 + z^2)^rk
      for (int i = 0; i <= rk; ++i)
        for (int j = 0; i + j <= rk; ++j) {
          int kk = rk - i - j;
          double mult = choose(rk, i) * choose(rk - i, j);
          for (auto &t : xy) {
            int px = (am - t.second) + 2 * i, py = t.second + 2 * j,
                pz = zp + 2 * kk;
            at(px, py, pz) += c * mult * t.first;
          }
        }
    }
    double *row = &T[(size_t)(m + l) * nc];
    for (int px = 0; px <= l; ++px)
      for (int py = 0; px + py <= l; ++py) {
        int pz = l - px - py;
        row[cart_index(px, py, pz)] = at(px, py, pz);
      }
    // unit norm w.r.t. Cartesian Gaussians normalised like x^l:
    // <x^a y^b z^c | x^a' y^b' z^c'> / <x^l|x^l> = prod (s-1)!! / (2l-1)!!
    double nrm = 0.0;
    for (int i = 0; i < nc; ++i)
      for (int j = 0; j < nc; ++j) {
        if (row[i] == 0.0 || row[j] == 0.0) continue;
        int e[2][3];
        for (int q = 0, idx = 0; q < 2; ++q) {
          int want = q ? j : i;
          idx = 0;
          for (int px = l; px >= 0; --px)
            for (int py = l - px; py >= 0; --py, ++idx)
              if (idx == want) { e[q][0] = px; e[q][1] = py; e[q][2] = l - px - py; }
        }
        int sx = e[0][0] + e[1][0], sy = e[0][1] + e[1][1], sz = e[0][2] + e[1][2];
        if ((sx | sy | sz) & 1) continue;
        nrm += row[i] * row[j] * double_factorial(sx - 1) * double_factorial(sy - 1) *
               double_factorial(sz - 1);
      }
    nrm /= double_factorial(2 * l - 1);
    double s = 1.0 / std::sqrt(nrm);
    for (int i = 0; i < nc; ++i) row[i] *= s;
  }
  return T;
}

// Primitive normalisation as x^l, then unit self-overlap of the contraction
// (libint2::Shell::renorm, the normalisation ChronusQ's shells carry).
static void renormalise(Shell &sh) {
  const double pi32 = std::pow(M_PI, 1.5);
  const int l = sh.l;
  const double dfl = double_factorial(2 * l - 1);
  for (size_t p = 0; p < sh.alpha.size(); ++p) {
    double ta = 2.0 * sh.alpha[p];
    sh.coef[p] *= std::sqrt(std::pow(2.0, l) * std::pow(ta, l + 1.5) / (pi32 * dfl));
  }
  double s = 0.0;
  for (size_t p = 0; p < sh.alpha.size(); ++p)
    for (size_t q = 0; q < sh.alpha.size(); ++q) {
      double g = sh.alpha[p] + sh.alpha[q];
      s += dfl * pi32 * sh.coef[p] * sh.coef[q] / (std::pow(2.0, l) * std::pow(g, l + 1.5));
    }
  double f = 1.0 / std::sqrt(s);
  for (auto &c : sh.coef) c *= f;
}

Basis Basis::from_g94(const std::string &text, const std::vector<int> &Z,
                      const std::vector<double> &xyz, bool force_cart) {
  // bad: the first shell of the record beyond kMaxL (reported only when an
  // atom of the element is asked for; basis files carry heavier elements
  // with higher shells)
  struct Rec { int z; std::vector<Shell> sh; std::string bad; };
  std::vector<Rec> recs;
  std::istringstream in(text);
  std::string line;
  bool in_record = false, new_record = false;
  while (std::getline(in, line)) {
    std::string s = rtrim(line);
    size_t b = s.find_first_not_of(" \t");
    if (b == std::string::npos) continue;
    s = s.substr(b);
    if (s[0] == '!') continue;
    if (s == "****") {
      std::string next;
      if (!std::getline(in, next)) { in_record = false; break; }
      next = rtrim(next);
      size_t nb = next.find_first_not_of(" \t");
      if (nb == std::string::npos) { in_record = false; continue; }
      s = next.substr(nb);
      in_record = true;
      new_record = true;
    }
    if (!in_record) continue;
    auto t = tokens(s);
    if (t.empty()) continue;
    if (new_record) {
      recs.push_back({element_number(t[0]), {}, {}});
      new_record = false;
      continue;
    }
    if (t.size() < 2) throw std::runtime_error("G94: malformed shell line: " + s);
    bool sp = (t[0] == "SP" || t[0] == "sp");
    int depth = std::stoi(t[1]);
    int L = sp ? 0 : l_of_symbol(t[0]);
    if (L < 0 || depth <= 0) throw std::runtime_error("G94: unsupported shell " + t[0]);
    Shell a, p;
    for (int i = 0; i < depth; ++i) {
      std::string pl;
      if (!std::getline(in, pl)) throw std::runtime_error("G94: truncated shell");
      auto pt = tokens(pl);
      if (pt.size() < (sp ? 3u : 2u)) throw std::runtime_error("G94: bad primitive line");
      a.alpha.push_back(to_double(pt[0]));
      a.coef.push_back(to_double(pt[1]));
      if (sp) { p.alpha.push_back(to_double(pt[0])); p.coef.push_back(to_double(pt[2])); }
    }
    a.l = L;
    a.pure = (L > 1) && !force_cart;
    if (L > kMaxL) {
      if (recs.back().bad.empty()) recs.back().bad = t[0];
      continue;
    }
    recs.back().sh.push_back(a);
    if (sp) { p.l = 1; p.pure = false; recs.back().sh.push_back(p); }
  }
  Basis bs;
  for (size_t at = 0; at < Z.size(); ++at) {
    const Rec *r = nullptr;
    for (auto &x : recs)
      if (x.z == Z[at]) r = &x;  // a later record for the same element wins
    if (!r) throw std::runtime_error("Cannot find Z=" + std::to_string(Z[at]) + " in Basis Definition");
    if (!r->bad.empty()) throw std::runtime_error("G94: unsupported shell " + r->bad);
    for (auto sh : r->sh) {
      for (int d = 0; d < 3; ++d) sh.origin[d] = xyz[3 * at + d];
      sh.atom = (int)at;
      renormalise(sh);
      bs.shells.push_back(sh);
    }
  }
  bs.finalise();
  return bs;
}

void Basis::finalise() {
  nbasis = 0;
  ncart_total = 0;
  max_prim = 0;
  max_l = 0;
  bf.clear();
  bf_cart.clear();
  nmol = 0;
  mol_sh0.clear();
  mol_bf0.clear();
  mol_bfc0.clear();
  for (size_t k = 0; k < shells.size(); ++k) {
    const auto &s = shells[k];
    if (k == 0 || s.mol != shells[k - 1].mol) {
      if (s.mol != nmol) throw std::runtime_error("batched shells must be grouped by molecule in order");
      mol_sh0.push_back((int)k);
      mol_bf0.push_back(nbasis);
      mol_bfc0.push_back(ncart_total);
      ++nmol;
    }
    bf.push_back(nbasis);
    bf_cart.push_back(ncart_total);
    nbasis += s.nfunc();
    ncart_total += ncart(s.l);
    max_prim = std::max(max_prim, (int)s.alpha.size());
    max_l = std::max(max_l, s.l);
  }
  if (nmol == 0) nmol = 1, mol_sh0.push_back(0), mol_bf0.push_back(0), mol_bfc0.push_back(0);
  mol_sh0.push_back((int)shells.size());
  mol_bf0.push_back(nbasis);
  mol_bfc0.push_back(ncart_total);
}

Basis Basis::batch(const std::vector<const Basis *> &parts) {
  Basis out;
  int atom0 = 0;
  for (size_t m = 0; m < parts.size(); ++m) {
    int amax = -1;
    for (auto sh : parts[m]->shells) {
      amax = std::max(amax, sh.atom);
      sh.atom += atom0;  // distinct centre labels across molecules
      sh.mol = (int)m;
      out.shells.push_back(sh);
    }
    atom0 += amax + 1;
  }
  out.finalise();
  return out;
}

// Contracted overlap block (pure where the shell is pure), row-major n1 x n2.
static std::vector<double> overlap_block(const Shell &A, const Shell &B,
                                         const std::vector<std::vector<double>> &sph) {
  int la = A.l, lb = B.l, nca = ncart(la), ncb = ncart(lb);
  std::vector<double> cart((size_t)nca * ncb, 0.0);
  double AB2 = 0;
  for (int d = 0; d < 3; ++d) AB2 += (A.origin[d] - B.origin[d]) * (A.origin[d] - B.origin[d]);
  std::vector<int> ea(3 * nca), eb(3 * ncb);
  for (int k = 0, lx = la; lx >= 0; --lx)
    for (int ly = la - lx; ly >= 0; --ly, ++k) { ea[3 * k] = lx; ea[3 * k + 1] = ly; ea[3 * k + 2] = la - lx - ly; }
  for (int k = 0, lx = lb; lx >= 0; --lx)
    for (int ly = lb - lx; ly >= 0; --ly, ++k) { eb[3 * k] = lx; eb[3 * k + 1] = ly; eb[3 * k + 2] = lb - lx - ly; }
  for (size_t p = 0; p < A.alpha.size(); ++p)
    for (size_t q = 0; q < B.alpha.size(); ++q) {
      double a = A.alpha[p], b = B.alpha[q], g = a + b;
      double pre = A.coef[p] * B.coef[q] * std::exp(-a * b / g * AB2) * std::pow(M_PI / g, 1.5);
      // 1D overlaps by the Obara-Saika two-centre recursion
      double S[3][kMaxL + 1][kMaxL + 1];
      for (int d = 0; d < 3; ++d) {
        double P = (a * A.origin[d] + b * B.origin[d]) / g;
        double PA = P - A.origin[d], PB = P - B.origin[d];
        for (int i = 0; i <= la; ++i)
          for (int j = 0; j <= lb; ++j) {
            double v;
            if (i == 0 && j == 0) v = 1.0;
            else if (j == 0) v = PA * S[d][i - 1][0] + (i > 1 ? (i - 1) / (2 * g) * S[d][i - 2][0] : 0.0);
            else v = PB * S[d][i][j - 1] + (j > 1 ? (j - 1) / (2 * g) * S[d][i][j - 2] : 0.0) +
                     (i > 0 ? i / (2 * g) * S[d][i - 1][j - 1] : 0.0);
            S[d][i][j] = v;
          }
      }
      for (int x = 0; x < nca; ++x)
        for (int y = 0; y < ncb; ++y)
          cart[(size_t)x * ncb + y] += pre * S[0][ea[3 * x]][eb[3 * y]] *
                                       S[1][ea[3 * x + 1]][eb[3 * y + 1]] *
                                       S[2][ea[3 * x + 2]][eb[3 * y + 2]];
    }
  int na = A.nfunc(), nb = B.nfunc();
  std::vector<double> out((size_t)na * nb, 0.0);
  for (int i = 0; i < na; ++i)
    for (int j = 0; j < nb; ++j) {
      double s = 0.0;
      for (int x = 0; x < nca; ++x) {
        double ta = A.pure ? sph[la][(size_t)i * nca + x] : (x == i ? 1.0 : 0.0);
        if (ta == 0.0) continue;
        for (int y = 0; y < ncb; ++y) {
          double tb = B.pure ? sph[lb][(size_t)j * ncb + y] : (y == j ? 1.0 : 0.0);
          if (tb != 0.0) s += ta * tb * cart[(size_t)x * ncb + y];
        }
      }
      out[(size_t)i * nb + j] = s;
    }
  return out;
}

std::vector<double> Basis::overlap_matrix() const {
  std::vector<std::vector<double>> sph;
  for (int l = 0; l <= kMaxL; ++l) sph.push_back(solid_harmonics(l));
  std::vector<double> S((size_t)nbasis * nbasis);
  for (size_t a = 0; a < shells.size(); ++a)
    for (size_t b = 0; b < shells.size(); ++b) {
      auto blk = overlap_block(shells[a], shells[b], sph);
      int na = shells[a].nfunc(), nb = shells[b].nfunc();
      for (int i = 0; i < na; ++i)
        for (int j = 0; j < nb; ++j) S[(bf[a] + i) + (size_t)(bf[b] + j) * nbasis] = blk[(size_t)i * nb + j];
    }
  return S;
}

// Significant pairs s2 <= s1: same centre (coordinates equal to 1e-5, the
// centre match of BasisSet::update) or 2-norm of the overlap block >= 1e-12;
// in a batched set only pairs within one molecule.
std::vector<std::pair<int, int>> Basis::significant_pairs() const {
  std::vector<std::vector<double>> sph;
  for (int l = 0; l <= kMaxL; ++l) sph.push_back(solid_harmonics(l));
  std::vector<std::pair<int, int>> out;
  int ns = (int)shells.size();
  for (int s1 = 0; s1 < ns; ++s1)
    for (int s2 = mol_sh0.empty() ? 0 : mol_sh0[shells[s1].mol]; s2 <= s1; ++s2) {
      // pairs never cross molecules of a batch (their shells are contiguous)
      const Shell &A = shells[s1], &B = shells[s2];
      bool same = std::fabs(A.origin[0] - B.origin[0]) < 1e-5 &&
                  std::fabs(A.origin[1] - B.origin[1]) < 1e-5 &&
                  std::fabs(A.origin[2] - B.origin[2]) < 1e-5;
      bool sig = same;
      if (!same) {
        auto blk = overlap_block(A, B, sph);
        double n2 = 0.0;
        for (double v : blk) n2 += v * v;
        sig = std::sqrt(n2) >= 1e-12;
      }
      if (sig) out.push_back({s1, s2});
    }
  return out;
}

}  // namespace cqx
