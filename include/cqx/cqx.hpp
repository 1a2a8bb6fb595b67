// C++ adapter over the cqx C ABI, shaped like the ChronusQ classes it
// replaces on the DIRECT path:
//
//   TwoBodyContraction<T>           include/particleintegrals/twopints.hpp:59-75
//   TPIContractions<MatsT,IntsT>    include/particleintegrals/twopints.hpp:172-260
//   GTODirectTPIContraction         include/particleintegrals/twopints/gtodirecttpi.hpp:143-228
//   FockBuilder::formGD             include/fockbuilder/impl.hpp:75-244
//
// Header-only; link against libcqx.so.  A non-zero status from the C ABI is
// raised as std::runtime_error("FATAL: ..."), the exception ChronusQ's CErr
// throws (include/cerr.hpp:33-51).
#pragma once

#include <complex>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../cqx.h"

namespace cqx_cpp {

using dcomplex = std::complex<double>;

enum TWOBODY_CONTRACTION_TYPE { COULOMB = CQX_COULOMB, EXCHANGE = CQX_EXCHANGE };

inline void check(int status, const cqx_ctx *ctx = nullptr) {
  if (status != CQX_OK) {
    const char *m = cqx_last_error(ctx);
    throw std::runtime_error(std::string("FATAL: ") + (m ? m : "cqx error"));
  }
}

// {X, AX, HER, contType}: X and AX are NB x NB column-major buffers
template <typename T>
struct TwoBodyContraction {
  T *X;
  T *AX;
  bool HER;
  TWOBODY_CONTRACTION_TYPE contType;
};

// Shell set built by the engine's G94 reader (or from arrays that a caller's
// own BasisSet already holds, normalised the libint2 way).
class BasisSet {
 public:
  BasisSet(const std::string &g94, const std::vector<int> &Z, const std::vector<double> &xyz,
           bool forceCart = false) {
    check(cqx_basis_from_g94(g94.c_str(), (int)Z.size(), Z.data(), xyz.data(), forceCart, &b_));
    check(cqx_basis_dims(b_, &nShell, &nBasis, &maxPrim, &maxL));
  }
  ~BasisSet() { cqx_basis_destroy(b_); }
  BasisSet(const BasisSet &) = delete;
  BasisSet &operator=(const BasisSet &) = delete;
  const cqx_basis *handle() const { return b_; }
  int nShell = 0, nBasis = 0, maxPrim = 0, maxL = 0;

 private:
  cqx_basis *b_ = nullptr;
};

// DirectTPI<double>: basis + Schwarz threshold, device-resident pair data
class DirectTPI {
 public:
  DirectTPI(const BasisSet &bs, double threshSchwarz = 1e-12, int device = -1)
      : nBasis_(bs.nBasis), thresh_(threshSchwarz) {
    check(cqx_ctx_create(bs.handle(), threshSchwarz, device, &ctx_));
  }
  ~DirectTPI() { cqx_ctx_destroy(ctx_); }
  DirectTPI(const DirectTPI &) = delete;
  DirectTPI &operator=(const DirectTPI &) = delete;
  cqx_ctx *handle() const { return ctx_; }
  int nBasis() const { return nBasis_; }
  double threshSchwarz() const { return thresh_; }
  // engine options (cqx_set_option: "graph", "fp32_cut", "list_cap", ...)
  void setOption(const std::string &name, double value) { check(cqx_set_option(ctx_, name.c_str(), value), ctx_); }
  // statistics of the last contraction (quartets, timings, graph replay)
  cqx_stats lastStats() const {
    cqx_stats st{};
    check(cqx_last_stats(ctx_, &st), ctx_);
    return st;
  }

 private:
  cqx_ctx *ctx_ = nullptr;
  int nBasis_;
  double thresh_;
};

// The virtual interface the reference's callers see
template <typename MatsT>
class TPIContractions {
 public:
  virtual ~TPIContractions() = default;
  virtual void twoBodyContract(bool screen, std::vector<TwoBodyContraction<MatsT>> &list) const = 0;
  void twoBodyContract(std::vector<TwoBodyContraction<MatsT>> &list) const { twoBodyContract(true, list); }
  bool contractSecond = false;
  bool printContractionTiming = false;
};

// GPU replacement of GTODirectTPIContraction<MatsT, double> for host
// matrices (MatsT = double or std::complex<double>)
template <typename MatsT>
class GpuDirectTPIContraction : public TPIContractions<MatsT> {
  static_assert(std::is_same<MatsT, double>::value || std::is_same<MatsT, dcomplex>::value,
                "MatsT must be double or std::complex<double>");

 public:
  explicit GpuDirectTPIContraction(std::shared_ptr<DirectTPI> tpi) : tpi_(std::move(tpi)) {}
  void twoBodyContract(bool screen, std::vector<TwoBodyContraction<MatsT>> &list) const override {
    std::vector<cqx_contraction> c(list.size());
    for (size_t i = 0; i < list.size(); ++i)
      c[i] = {list[i].X, list[i].AX, list[i].HER ? 1 : 0, (int)list[i].contType};
    check(cqx_two_body_contract(tpi_->handle(), screen, (int)c.size(), c.data(),
                                std::is_same<MatsT, dcomplex>::value ? 1 : 0, CQX_HOST),
          tpi_->handle());
  }
  const std::shared_ptr<DirectTPI> &ints() const { return tpi_; }

 private:
  std::shared_ptr<DirectTPI> tpi_;
};

// GradContractions<double, double> for the DIRECT algorithm
// (particleintegrals/gradints/direct.hpp:31-76): cList[g][m], g = 3 atom +
// xyz, one density per matrix m shared by every g (FockBuilder::getGDGrad).
class GpuDirectGradContraction {
 public:
  explicit GpuDirectGradContraction(std::shared_ptr<DirectTPI> tpi) : tpi_(std::move(tpi)) {}
  void gradTwoBodyContract(std::vector<std::vector<TwoBodyContraction<double>>> &cList) const {
    if (cList.empty() || cList[0].empty()) throw std::runtime_error("FATAL: Nothing specified to compute");
    const size_t nmat = cList[0].size();
    std::vector<cqx_contraction> c;
    c.reserve(cList.size() * nmat);
    for (auto &row : cList) {
      if (row.size() != nmat) throw std::runtime_error("FATAL: ragged gradient contraction list");
      for (auto &x : row) c.push_back({x.X, x.AX, x.HER ? 1 : 0, (int)x.contType});
    }
    check(cqx_grad_two_body_contract(tpi_->handle(), (int)cList.size(), (int)nmat, c.data(), CQX_HOST),
          tpi_->handle());
  }

 private:
  std::shared_ptr<DirectTPI> tpi_;
};

// The parts of SingleSlater that formGD reads and writes: Pauli components
// of the density (S; or S, Z, Y, X) and the J / K / G outputs.
template <typename MatsT>
struct FockState {
  std::vector<std::vector<MatsT>> onePDM;         // nComp x (NB*NB)
  std::vector<MatsT> coulombMatrix;                // NB*NB
  std::vector<std::vector<MatsT>> exchangeMatrix;  // nComp x (NB*NB)
  std::vector<std::vector<MatsT>> twoeH;           // nComp x (NB*NB)
};

template <typename MatsT>
struct FockBuilder {
  // FockBuilder::formGD for an electronic, non-incremental build:
  // J on S, K on each component, G = 2J - xHFX K (impl.hpp:109-237)
  static void formGD(FockState<MatsT> &ss, const TPIContractions<MatsT> &tpi, double xHFX = 1.0,
                     bool HerDen = true) {
    const size_t nc = ss.onePDM.size();
    if (nc == 0) throw std::runtime_error("FATAL: Nothing specified to compute");
    const size_t nb2 = ss.onePDM[0].size();
    ss.coulombMatrix.assign(nb2, MatsT(0.));
    ss.exchangeMatrix.assign(nc, std::vector<MatsT>(nb2, MatsT(0.)));
    ss.twoeH.assign(nc, std::vector<MatsT>(nb2, MatsT(0.)));
    std::vector<TwoBodyContraction<MatsT>> cont;
    cont.push_back({ss.onePDM[0].data(), ss.coulombMatrix.data(), HerDen, COULOMB});
    const bool doK = std::abs(xHFX) > 1e-12;
    if (doK)
      for (size_t c = 0; c < nc; ++c)
        cont.push_back({ss.onePDM[c].data(), ss.exchangeMatrix[c].data(), HerDen, EXCHANGE});
    tpi.twoBodyContract(true, cont);
    for (size_t c = 0; c < nc; ++c)
      for (size_t k = 0; k < nb2; ++k) {
        MatsT g = c == 0 ? MatsT(2.0) * ss.coulombMatrix[k] : MatsT(0.);
        if (doK) g -= MatsT(xHFX) * ss.exchangeMatrix[c][k];
        ss.twoeH[c][k] = g;
      }
  }
};

}  // namespace cqx_cpp
