// One-electron integrals (overlap, kinetic energy, point-nucleus
// attraction) over a host shell set; see onebody.cpp.
#pragma once

#include <vector>

#include "basis.hpp"

namespace cqx {

enum OneBodyKind { kOverlap = 0, kKinetic = 1, kPotential = 2 };

// Boys function F_m(T), m = 0..mmax (mmax <= 2 kMaxL), to ~1e-15 relative.
void boys_host(int mmax, double T, double *F);

// Column-major NB x NB matrix (pure functions where the shell is pure).
// kPotential: sum over point charges Zc[c] at Cxyz[3c..3c+2] (bohr) of
// -Zc <a| 1/|r - C| |b>.  nthreads <= 0 picks the host's (capped at 16).
std::vector<double> one_body_matrix(const Basis &B, int kind, const std::vector<double> &Zc,
                                    const std::vector<double> &Cxyz, int nthreads = 0);

}  // namespace cqx
