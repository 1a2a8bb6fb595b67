// Drives the C++ adapter (include/cqx/cqx.hpp) the way ChronusQ's SCF
// drives FockBuilder::formGD: one RHF or two-component G[D] build.
//
// usage: formgd <basis.g94> <geom.txt: Z x y z (bohr) per line> <ncomp>
//               <density.bin> <out.bin>
// density.bin: ncomp NB x NB column-major matrices, double (ncomp = 1) or
// complex<double> (ncomp = 4, Pauli S, Z, Y, X); out.bin receives G likewise.
// The context records its build into a graph (option "graph"): the first
// formGD runs eagerly and records, the second replays, and the two G must
// agree before the second is written.
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

template <typename T>
int run(const std::string &g94, const std::vector<int> &Z, const std::vector<double> &xyz, int ncomp,
        const char *din, const char *dout) {
  cqx_cpp::BasisSet bs(g94, Z, xyz);
  auto tpi = std::make_shared<cqx_cpp::DirectTPI>(bs, 1e-12);
  tpi->setOption("graph", 1.0);
  cqx_cpp::GpuDirectTPIContraction<T> gto(tpi);
  const size_t nb2 = (size_t)bs.nBasis * bs.nBasis;
  cqx_cpp::FockState<T> ss;
  ss.onePDM.assign(ncomp, std::vector<T>(nb2));
  std::ifstream f(din, std::ios::binary);
  for (auto &d : ss.onePDM) f.read(reinterpret_cast<char *>(d.data()), sizeof(T) * nb2);
  if (!f) throw std::runtime_error("FATAL: short density file");
  cqx_cpp::FockBuilder<T>::formGD(ss, gto);
  const int first = tpi->lastStats().graph;
  const auto G1 = ss.twoeH;
  cqx_cpp::FockBuilder<T>::formGD(ss, gto);
  const int second = tpi->lastStats().graph;
  double diff = 0.0;
  for (size_t c = 0; c < G1.size(); ++c)
    for (size_t k = 0; k < nb2; ++k) diff = std::max(diff, std::abs(G1[c][k] - ss.twoeH[c][k]));
  if (first != 1 || second != 2 || !(diff < 1e-11)) {
    std::fprintf(stderr, "recorded build: graph modes %d, %d; max |G1 - G2| %.3e\n", first, second, diff);
    return 1;
  }
  std::ofstream o(dout, std::ios::binary);
  for (auto &g : ss.twoeH) o.write(reinterpret_cast<const char *>(g.data()), sizeof(T) * nb2);
  std::printf("formGD NB=%d ncomp=%d ok\n", bs.nBasis, ncomp);
  return 0;
}

int main(int argc, char **argv) {
  if (argc != 6) {
    std::fprintf(stderr, "usage: %s basis.g94 geom.txt ncomp density.bin out.bin\n", argv[0]);
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
    return ncomp == 1 ? run<double>(ss.str(), Z, xyz, 1, argv[4], argv[5])
                      : run<std::complex<double>>(ss.str(), Z, xyz, ncomp, argv[4], argv[5]);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
}
