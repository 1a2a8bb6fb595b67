// Host-side shell sets of the direct J/K engine (see basis.cpp).
#pragma once

#include <string>
#include <utility>
#include <vector>

namespace cqx {

constexpr int kMaxL = 4;  // highest shell angular momentum accepted

inline int ncart(int l) { return (l + 1) * (l + 2) / 2; }
int cart_index(int lx, int ly, int lz);
std::vector<double> solid_harmonics(int l);  // (2l+1) x ncart(l), row-major

struct Shell {
  int l = 0;
  bool pure = false;
  int atom = 0;
  int mol = 0;  // molecule of a batched shell set (0 for a single molecule)
  double origin[3] = {0, 0, 0};
  std::vector<double> alpha, coef;  // coef normalised (libint2 convention)
  int nfunc() const { return pure ? 2 * l + 1 : ncart(l); }
  int nprim() const { return (int)alpha.size(); }
};

struct Basis {
  std::vector<Shell> shells;
  std::vector<int> bf;       // first (pure-or-Cartesian) function of each shell
  std::vector<int> bf_cart;  // first Cartesian function of each shell
  int nbasis = 0, ncart_total = 0, max_prim = 0, max_l = 0;
  // molecules of a batched set: shells of molecule m are the contiguous
  // range [mol_sh0[m], mol_sh0[m + 1]) (nmol = 1 for a single molecule)
  int nmol = 1;
  std::vector<int> mol_sh0, mol_bf0, mol_bfc0;  // nmol + 1 prefix offsets

  static Basis from_g94(const std::string &text, const std::vector<int> &Z,
                        const std::vector<double> &xyz, bool force_cart);
  void finalise();
  // the molecules' shell sets side by side (molecule m's shells get mol = m)
  static Basis batch(const std::vector<const Basis *> &parts);
  std::vector<double> overlap_matrix() const;  // column-major NB x NB
  std::vector<std::pair<int, int>> significant_pairs() const;
};

}  // namespace cqx
