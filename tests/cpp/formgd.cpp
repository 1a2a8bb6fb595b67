// Drives the C++ adapter (include/cqx/cqx.hpp) the way ChronusQ's SCF
// drives FockBuilder: formGD / formFock on one RHF or two-component density,
// through the reference's signatures (twoBodyContract(MPI_Comm, screen,
// list, pert), formGD(ss, pert, increment, xHFX, HerDen)).
//
// usage: formgd <basis.g94> <geom.txt: Z x y z (bohr) per line> <ncomp>
//               <density.bin> <hcore.bin> <out.bin>
// density.bin / hcore.bin: ncomp NB x NB column-major matrices, double
// (ncomp = 1) or complex<double> (ncomp = 4, Pauli S, Z, Y, X).
// out.bin receives G (ncomp matrices) then F (ncomp matrices); stdout the
// line "energies E1 E2" of computeEnergy on D, H and G.
// Checks, each of which fails the run:
//  * the context records its build into a graph: the second formGD
//    replays it and gives the same G;
//  * formGD with increment on deltaOnePDM = D - D/2 after a plain build of
//    D/2 gives K[D] and G[D] - K[D/2]: K accumulates, and G accumulates
//    2 J[delta] - K with the accumulated K (impl.hpp:223-237);
//  * SpinGather of the components followed by SpinScatter returns them.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "cqx/cqx.hpp"

using namespace cqx_cpp;

template <typename T>
double maxdiff(const T *a, const T *b, size_t n) {
  double d = 0.0;
  for (size_t k = 0; k < n; ++k) d = std::max(d, std::abs(a[k] - b[k]));
  return d;
}

template <typename T>
int run(const std::string &g94, const std::vector<int> &Z, const std::vector<double> &xyz, int ncomp,
        const char *din, const char *hin, const char *dout) {
  BasisSet bs(g94, Z, xyz);
  auto tpi = std::make_shared<DirectTPI>(bs, 1e-12);
  tpi->setOption("graph", 1.0);
  auto gto = std::make_shared<GpuDirectTPIContraction<T, double>>(tpi);
  const size_t NB = bs.nBasis, nb2 = NB * NB;
  SingleSlater<T> ss(gto, NB, ncomp == 1 ? 1 : 2);
  std::ifstream f(din, std::ios::binary), h(hin, std::ios::binary);
  for (int c = 0; c < ncomp; ++c) {
    f.read(reinterpret_cast<char *>(ss.onePDM->component(c)), sizeof(T) * nb2);
    h.read(reinterpret_cast<char *>(ss.coreH->component(c)), sizeof(T) * nb2);
  }
  if (!f || !h) CErr("short density or core-Hamiltonian file");
  FockBuilder<T> fb;
  EMPerturbation pert;
  // 1. eager + recorded, then replayed
  fb.formGD(ss, pert, false, 1.0, true);
  const int first = tpi->lastStats().graph;
  std::vector<std::vector<T>> G1(ncomp), K1(ncomp);
  for (int c = 0; c < ncomp; ++c) {
    G1[c].assign(ss.twoeH->component(c), ss.twoeH->component(c) + nb2);
    K1[c].assign(ss.exchangeMatrix->component(c), ss.exchangeMatrix->component(c) + nb2);
  }
  fb.formFock(ss, pert);
  const int second = tpi->lastStats().graph;
  double diff = 0.0;
  for (int c = 0; c < ncomp; ++c) diff = std::max(diff, maxdiff(G1[c].data(), ss.twoeH->component(c), nb2));
  if (first != 1 || second != 2 || !(diff < 1e-11)) {
    std::fprintf(stderr, "recorded build: graph modes %d, %d; max |G1 - G2| %.3e\n", first, second, diff);
    return 1;
  }
  ss.computeEnergy(*tpi, 0.0);
  // 2. incremental build: G[D/2], then the increment D - D/2 gives K[D] and
  //    G[D] - K[D/2] = G[D] - K[D] / 2
  SingleSlater<T> inc(gto, NB, ncomp == 1 ? 1 : 2);
  double scale = 0.0;
  for (int c = 0; c < ncomp; ++c)
    for (size_t k = 0; k < nb2; ++k) {
      const T d = ss.onePDM->component(c)[k];
      inc.onePDM->component(c)[k] = T(0.5) * d;
      inc.deltaOnePDM->component(c)[k] = d - T(0.5) * d;
      scale = std::max(scale, std::abs(G1[c][k]));
    }
  fb.formGD(inc, pert, false);
  fb.formGD(inc, pert, true);
  double dinc = 0.0;
  for (int c = 0; c < ncomp; ++c) {
    std::vector<T> Gexp(nb2);
    for (size_t k = 0; k < nb2; ++k) Gexp[k] = G1[c][k] - T(0.5) * K1[c][k];
    dinc = std::max(dinc, maxdiff(Gexp.data(), inc.twoeH->component(c), nb2));
    dinc = std::max(dinc, maxdiff(K1[c].data(), inc.exchangeMatrix->component(c), nb2));
  }
  if (!(dinc <= 1e-10 * std::max(1.0, scale))) {
    std::fprintf(stderr, "incremental build: max deviation from K[D], G[D] - K[D/2]: %.3e (scale %.3e)\n", dinc, scale);
    return 1;
  }
  // 3. SpinGather then SpinScatter returns the components
  if (ncomp == 4) {
    const auto A = ss.onePDM->spinGather(*tpi);
    auto P = PauliSpinorMatrices<T>::spinScatter(*tpi, A.data(), NB);
    double drt = 0.0;
    for (int c = 0; c < 4; ++c) drt = std::max(drt, maxdiff(P.component(c), ss.onePDM->component(c), nb2));
    if (!(drt < 1e-14)) {
      std::fprintf(stderr, "SpinGather / SpinScatter round trip: %.3e\n", drt);
      return 1;
    }
  }
  if (ss.fockMatrix->hasNaN(*tpi)) CErr("NaN in the Fock matrix");
  std::ofstream o(dout, std::ios::binary);
  for (int c = 0; c < ncomp; ++c) o.write(reinterpret_cast<const char *>(ss.twoeH->component(c)), sizeof(T) * nb2);
  for (int c = 0; c < ncomp; ++c)
    o.write(reinterpret_cast<const char *>(ss.fockMatrix->component(c)), sizeof(T) * nb2);
  std::printf("formGD NB=%zu ncomp=%d ok\nenergies %.15e %.15e\n", NB, ncomp, ss.OBEnergy, ss.MBEnergy);
  std::printf("scratch per batch (MatsT elements) %zu\n", fb.formRawGDSCRSizePerBatch(ss, true, true));
  return 0;
}

int main(int argc, char **argv) {
  if (argc != 7) {
    std::fprintf(stderr, "usage: %s basis.g94 geom.txt ncomp density.bin hcore.bin out.bin\n", argv[0]);
    return 2;
  }
  std::ifstream gb(argv[1]);
  std::stringstream ss;
  ss << gb.rdbuf();
  std::vector<int> Z;
  std::vector<double> xyz;
  std::ifstream geom(argv[2]);
  int z;
  double x, y, w;
  while (geom >> z >> x >> y >> w) {
    Z.push_back(z);
    xyz.insert(xyz.end(), {x, y, w});
  }
  const int ncomp = std::atoi(argv[3]);
  try {
    return ncomp == 1 ? run<double>(ss.str(), Z, xyz, 1, argv[4], argv[5], argv[6])
                      : run<std::complex<double>>(ss.str(), Z, xyz, ncomp, argv[4], argv[5], argv[6]);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
}
